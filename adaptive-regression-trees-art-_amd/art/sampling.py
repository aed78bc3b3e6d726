"""Hit-and-run polytope samplers on the GPU (art/sampling.py of the reference).

``HitAndRunSampler``, ``BatchedHitAndRunSampler`` and the two Numba sampler
classes keep the reference's dataclass fields, validation and
``sample(region, n, random_state=None, x0=None) -> (n, d)`` contract
(art/sampling.py:53-64,432-1025).  Every transition runs in
``art_har_sample``: one lane per chain, FP64 chord endpoints with slacks
recomputed from the point every move, FP32 Philox4x32-10 Box-Muller
directions.  ``PolytopeBatch`` samples many polytopes in one launch, the shape
of a tree level's leaves.

Random streams differ from the reference's NumPy streams (the reference's
own tests pin equality only for injected streams); ``random_state`` seeds the
device stream, so a seed reproduces its samples exactly.  Host-side helpers
for start points and direction spectra are the reference's CPU code paths
restated (find_feasible_point, infer_box_bounds, covariance spectra).
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field
from typing import ClassVar, Literal, Protocol, runtime_checkable

import numpy as np

from . import _hip
from .domain import BoxDomain, PolytopeRegion

RandomState = int | np.random.Generator | None
ThinningCandidateMode = Literal["linear", "powers_of_two"]


def as_rng(random_state: RandomState = None) -> np.random.Generator:
    if isinstance(random_state, np.random.Generator):
        return random_state
    return np.random.default_rng(random_state)


@runtime_checkable
class Sampler(Protocol):
    def sample(self, region, n, random_state=None, x0=None):
        ...


def sample_uniform_box(bounds, n, random_state: RandomState = None):
    """art/sampling.py:67-83."""
    bounds = np.asarray(bounds, dtype=float)
    if bounds.ndim != 2 or bounds.shape[1] != 2:
        raise ValueError(f"bounds must have shape (d, 2), got {bounds.shape}.")
    if np.any(bounds[:, 0] >= bounds[:, 1]):
        raise ValueError("Each bound must satisfy low < high.")
    if n < 0:
        raise ValueError("n must be nonnegative.")
    return as_rng(random_state).uniform(bounds[:, 0], bounds[:, 1], size=(int(n), bounds.shape[0]))


def sample_covariance_eigendecomposition(samples, *, assume_centered=False):
    """art/sampling.py:86-104."""
    values = np.asarray(samples, dtype=float)
    if values.ndim != 2:
        raise ValueError(f"samples must have shape (n_samples, d), got {values.shape}.")
    if values.shape[0] < 2:
        raise ValueError("samples must contain at least two points.")
    if values.shape[1] < 1:
        raise ValueError("samples must contain at least one coordinate.")
    if not np.all(np.isfinite(values)):
        raise ValueError("samples must contain only finite values.")
    centered = values if assume_centered else values - values.mean(axis=0, keepdims=True)
    eigenvalues, eigenvectors = np.linalg.eigh(centered.T @ centered / (centered.shape[0] - 1))
    return eigenvectors, eigenvalues


def floor_covariance_eigenvalues(eigenvalues, floor_ratio):
    """Trace-scaled positive floor (art/sampling.py:107-127)."""
    values = np.asarray(eigenvalues, dtype=float).reshape(-1)
    if values.size == 0:
        raise ValueError("eigenvalues must contain at least one value.")
    if not np.all(np.isfinite(values)):
        raise ValueError("eigenvalues must be finite.")
    if not np.isfinite(floor_ratio) or floor_ratio <= 0.0:
        raise ValueError("floor_ratio must be positive.")
    mean = float(np.mean(values))
    if mean <= 0.0:
        raise ValueError("eigenvalues must have a positive mean before flooring.")
    floor = float(floor_ratio * mean)
    return np.maximum(values, floor), floor


def infer_box_bounds(region: PolytopeRegion, tol: float = 1e-12):
    """Box bounds from single-coordinate rows, or None (art/sampling.py:145-168)."""
    lows = np.full(region.dimension, -np.inf)
    highs = np.full(region.dimension, np.inf)
    for row, bound in zip(region.A, region.b):
        nz = np.flatnonzero(np.abs(row) > tol)
        if nz.size != 1:
            continue
        i = int(nz[0])
        if row[i] > 0:
            highs[i] = min(highs[i], bound / row[i])
        else:
            lows[i] = max(lows[i], bound / row[i])
    if np.any(~np.isfinite(lows)) or np.any(~np.isfinite(highs)) or np.any(lows >= highs):
        return None
    return np.column_stack([lows, highs])


def find_feasible_point(region, bounds=None, max_tries=20000, random_state=None, tol=1e-10):
    """Rejection sampling from a bounding box (art/sampling.py:171-190)."""
    if bounds is None:
        bounds = infer_box_bounds(region)
    if bounds is None:
        raise ValueError("bounds are required when they cannot be inferred from the region.")
    rng = as_rng(random_state)
    for _ in range(max_tries):
        x = sample_uniform_box(bounds, 1, random_state=rng)[0]
        if bool(region.contains(x, tol=tol)[0]):
            return x
    raise RuntimeError("Could not find a feasible point inside the region.")


def split_axis_rows(A, b, offsets):
    """Separate rows x_j <= c and -x_j <= c from the general rows.

    Returns (A_general, b_general, offsets_general, lo (P, d), hi (P, d)); lo
    and hi hold the tightest axis bound per side (-inf / +inf where none), or
    are None when no polytope has an axis row.
    """
    P, d = offsets.size - 1, A.shape[1]
    owner = np.repeat(np.arange(P), np.diff(offsets))
    nz = A != 0.0
    col = np.argmax(nz, axis=1)
    val = A[np.arange(A.shape[0]), col]
    axis = (nz.sum(axis=1) == 1) & ((val == 1.0) | (val == -1.0))
    if not axis.any():
        return A, b, offsets, None, None
    lo = np.full((P, d), -np.inf)
    hi = np.full((P, d), np.inf)
    up, dn = axis & (val == 1.0), axis & (val == -1.0)
    np.minimum.at(hi, (owner[up], col[up]), b[up])
    np.maximum.at(lo, (owner[dn], col[dn]), -b[dn])
    keep = ~axis
    counts = np.bincount(owner[keep], minlength=P)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    return np.ascontiguousarray(A[keep]), np.ascontiguousarray(b[keep]), off, lo, hi


class PolytopeBatch:
    """P polytopes of one dimension d, resident on the device.

    ``regions``: a sequence of PolytopeRegion, or a pair (A_list, b_list).
    Rows are stored back to back (CSR by polytope) as float64.  With
    ``split_box`` (default) rows +-e_j are held as per-coordinate bounds,
    which gives the same chords for less arithmetic.
    """

    def __init__(self, regions, *, stream=None, split_box=True):
        if isinstance(regions, tuple) and len(regions) == 2:
            As, bs = regions
        else:
            As = [r.A for r in regions]
            bs = [r.b for r in regions]
        As = [np.asarray(a, dtype=float) for a in As]
        bs = [np.asarray(b, dtype=float).reshape(-1) for b in bs]
        if not As:
            raise ValueError("At least one polytope is required.")
        d = As[0].shape[1] if As[0].ndim == 2 else -1
        for a, b in zip(As, bs):
            if a.ndim != 2 or a.shape[1] != d or a.shape[0] != b.shape[0] or a.shape[0] < 1:
                raise ValueError("Every polytope needs A (m, d) and b (m,) with one d.")
        rows = np.array([a.shape[0] for a in As], dtype=np.int64)
        offsets = np.concatenate([[0], np.cumsum(rows)]).astype(np.int64)
        self._setup(np.concatenate(As), np.concatenate(bs), offsets, stream, split_box)

    @classmethod
    def from_csr(cls, A, b, offsets, *, stream=None, split_box=True):
        """Polytope q owns rows [offsets[q], offsets[q+1]) of A (rows, d), b (rows,)."""
        A = np.ascontiguousarray(A, dtype=float)
        b = np.ascontiguousarray(b, dtype=float).reshape(-1)
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        if (A.ndim != 2 or offsets.ndim != 1 or offsets.size < 2 or offsets[0] != 0
                or offsets[-1] != A.shape[0] or b.shape[0] != A.shape[0]
                or np.any(np.diff(offsets) < 1)):
            raise ValueError("Need A (rows, d), b (rows,) and offsets (P+1,) from 0 to rows "
                             "with at least one row per polytope.")
        self = cls.__new__(cls)
        self._setup(A, b, offsets, stream, split_box)
        return self

    def _setup(self, A, b, offsets, stream, split_box):
        if not (np.isfinite(A).all() and np.isfinite(b).all()):
            raise ValueError("A and b must be finite.")
        d = A.shape[1]
        if not 1 <= d <= 32:
            raise ValueError("The device sampler supports 1 <= d <= 32.")
        self.P, self.d = offsets.size - 1, d
        self.rows = np.diff(offsets)
        lo = hi = None
        if split_box:
            A, b, offsets, lo, hi = split_axis_rows(A, b, offsets)
        torch = _hip.torch_device()
        self.torch, self.lib = torch, _hip.load()
        self.device = torch.device("cuda", torch.cuda.current_device())
        self.stream = torch.cuda.current_stream() if stream is None else stream
        with torch.cuda.stream(self.stream):
            self.A = torch.from_numpy(np.ascontiguousarray(A)).to(self.device)
            self.b = torch.from_numpy(np.ascontiguousarray(b)).to(self.device)
            self.offsets = torch.from_numpy(offsets).to(self.device)
            self.box_lo = None if lo is None else torch.from_numpy(lo).to(self.device)
            self.box_hi = None if hi is None else torch.from_numpy(hi).to(self.device)
        self.host_offsets = offsets
        self.max_rows = int(np.max(np.diff(offsets))) if self.P else 0
        self.last_kernel_ms = None

    def sample(self, x0, n, *, first, stride, seed=0, shape=None, directions=None,
               uniforms=None, out=None, timed=False, seeds=None):
        """Samples (P, n, d) on the device for starts x0 (P, C, d) (host or device).

        Chain c records after moves first, first + stride, ...; record k is
        sample k*C + c.  ``directions (P, S, C, d)`` / ``uniforms (P, S, C)``
        replace the Philox stream when given.  ``shape`` is one (d, d)
        shaping matrix or one per polytope (P, d, d); ``seeds`` (P,) gives
        each polytope the stream a one-polytope call with that seed draws.
        """
        torch, P, d = self.torch, self.P, self.d
        dev = dict(dtype=torch.float64, device=self.device)
        x0 = torch.as_tensor(x0, **dev)
        if x0.dim() != 3 or x0.shape[0] != P or x0.shape[2] != d:
            raise ValueError(f"x0 must have shape ({P}, chains, {d}).")
        C = int(x0.shape[1])
        n, first, stride = int(n), int(first), int(stride)
        if not 1 <= C <= n:
            raise ValueError("Need 1 <= chains <= n.")
        if first < 1 or stride < 1:
            raise ValueError("first and stride must be at least 1.")
        x0 = x0.contiguous()
        a = _hip.HarArgs()
        a.A, a.b, a.row_offsets = _hip.ptr(self.A), _hip.ptr(self.b), _hip.ptr(self.offsets)
        a.x0 = _hip.ptr(x0)
        if self.box_lo is not None:
            a.box_lo, a.box_hi = _hip.ptr(self.box_lo), _hip.ptr(self.box_hi)
        keep = [x0]
        if shape is not None:
            sh = torch.as_tensor(np.asarray(shape, dtype=float), **dev).contiguous()
            if tuple(sh.shape) == (P, d, d):
                a.shape_stride = d * d
            elif tuple(sh.shape) != (d, d):
                raise ValueError(f"shape must be ({d}, {d}) or ({P}, {d}, {d}).")
            a.shape = _hip.ptr(sh)
            keep.append(sh)
        if seeds is not None:
            sd = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint64).reshape(-1))
            if sd.shape[0] != P:
                raise ValueError(f"seeds must have shape ({P},).")
            st = torch.from_numpy(sd.view(np.int64)).to(self.device)
            a.seeds = _hip.ptr(st)
            keep.append(st)
        a.steps = 0
        if directions is not None:
            dr = torch.as_tensor(directions, **dev).contiguous()
            un = torch.as_tensor(uniforms, **dev).contiguous()
            if dr.dim() != 4 or dr.shape[0] != P or dr.shape[2] != C or dr.shape[3] != d:
                raise ValueError(f"directions must have shape ({P}, steps, {C}, {d}).")
            if tuple(un.shape) != tuple(dr.shape[:3]):
                raise ValueError("uniforms must have shape (P, steps, chains).")
            a.directions, a.uniforms, a.steps = _hip.ptr(dr), _hip.ptr(un), int(dr.shape[1])
            keep += [dr, un]
        if out is None:
            out = torch.empty((P, n, d), **dev)
        elif tuple(out.shape) != (P, n, d) or out.dtype != torch.float64 or not out.is_contiguous():
            raise ValueError(f"out must be a contiguous float64 tensor of shape ({P}, {n}, {d}).")
        a.samples = _hip.ptr(out)
        a.row_bound = self.max_rows + 1
        a.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        a.P, a.d, a.chains, a.n, a.first, a.stride = P, d, C, n, first, stride
        with torch.cuda.stream(self.stream):
            if timed:
                ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
                ev[0].record(self.stream)
            _hip.check(self.lib.art_har_sample(ctypes.byref(a), _hip.stream_handle(self.stream)))
            if timed:
                ev[1].record(self.stream)
        if timed:
            self.stream.synchronize()
            self.last_kernel_ms = ev[0].elapsed_time(ev[1])
        self._keep = keep  # inputs stay alive until the next call
        return out


@dataclass
class HitAndRunSampler:
    """Hit-and-run sampler for convex polytopes A x <= b (art/sampling.py:432-564).

    One chain; a record after moves burn_in + 1, burn_in + 1 + thinning, ...
    """

    burn_in: int = 500
    thinning: int = 10
    feasibility_tol: float = 1e-10
    max_feasible_tries: int = 20000
    bounds: np.ndarray | None = None
    direction_eigenvectors: np.ndarray | None = None
    direction_eigenvalues: np.ndarray | None = None
    direction_eigenvalue_floor: float = 1e-4

    def sample(self, region, n, random_state=None, x0=None):
        if n < 0:
            raise ValueError("n must be nonnegative.")
        if self.burn_in < 0:
            raise ValueError("burn_in must be nonnegative.")
        if self.thinning < 1:
            raise ValueError("thinning must be at least 1.")
        rng = as_rng(random_state)
        shape = self._shaping(region.dimension)
        if x0 is None:
            x = find_feasible_point(region, bounds=self.bounds, max_tries=self.max_feasible_tries,
                                    random_state=rng, tol=self.feasibility_tol)
        else:
            x = np.asarray(x0, dtype=float).reshape(-1)
            if x.shape[0] != region.dimension:
                raise ValueError(f"x0 must have shape ({region.dimension},), got {x.shape}.")
            if not bool(region.contains(x, tol=self.feasibility_tol)[0]):
                raise ValueError("x0 is not feasible for the region.")
        n = int(n)
        if n == 0:
            return np.empty((0, region.dimension))
        seed = int(rng.integers(0, np.iinfo(np.int64).max))
        batch = PolytopeBatch([region])
        out = batch.sample(x[None, None], n, first=self.burn_in + 1, stride=self.thinning,
                           seed=seed, shape=shape)
        return out[0].cpu().numpy()

    def _direction_spectrum(self, dimension):
        """(eigenvectors, sqrt eigenvalues scaled to max 1) or (None, None)
        (art/sampling.py:496-537)."""
        V, lam = self.direction_eigenvectors, self.direction_eigenvalues
        if V is None and lam is None:
            return None, None
        if V is None or lam is None:
            raise ValueError(
                "direction_eigenvectors and direction_eigenvalues must be provided together.")
        V = np.asarray(V, dtype=float)
        lam = np.asarray(lam, dtype=float).reshape(-1)
        if V.shape != (dimension, dimension):
            raise ValueError("direction_eigenvectors must have shape "
                             f"{(dimension, dimension)}, got {V.shape}.")
        if lam.shape != (dimension,):
            raise ValueError(f"direction_eigenvalues must have shape ({dimension},), "
                             f"got {lam.shape}.")
        if not np.all(np.isfinite(V)) or not np.all(np.isfinite(lam)):
            raise ValueError("Direction eigenvectors and eigenvalues must be finite.")
        if not np.allclose(V.T @ V, np.eye(dimension), rtol=1e-8, atol=1e-10):
            raise ValueError("direction_eigenvectors must be orthonormal.")
        safe, _ = floor_covariance_eigenvalues(lam, self.direction_eigenvalue_floor)
        s = np.sqrt(safe)
        s /= np.max(s)
        return V, s

    def _shaping(self, dimension):
        V, s = self._direction_spectrum(dimension)
        return None if V is None else V * s[None, :]


@dataclass
class BatchedHitAndRunSampler(HitAndRunSampler):
    """Several chains per region; records after every `thinning` moves in order
    round * chains + chain, the last partial round on the first chains
    (art/sampling.py:605-803)."""

    n_chains: int = 20
    precompute_ad: bool = False
    profile_timing: bool = False
    last_phase_timings: dict = field(default_factory=dict, init=False, repr=False,
                                     compare=False)
    _chain_major: ClassVar[bool] = False

    def sample(self, region, n, random_state=None, x0=None):
        if n < 0:
            raise ValueError("n must be nonnegative.")
        if self.burn_in < 0:
            raise ValueError("burn_in must be nonnegative.")
        if self.thinning < 0:
            raise ValueError("thinning must be nonnegative.")
        if self.n_chains < 1:
            raise ValueError("n_chains must be at least 1.")
        self._check_random_settings()
        self.last_phase_timings = {}
        n = int(n)
        if n == 0:
            return np.empty((0, region.dimension), dtype=float)
        rng = as_rng(random_state)
        chains = min(self.n_chains, n)
        points = self._initial_points(region, chains, rng, x0)
        seed = int(rng.integers(0, np.iinfo(np.int64).max))
        stride = max(1, self.thinning)
        batch = PolytopeBatch([region])
        t0 = time.perf_counter()
        out = batch.sample(points[None], n, first=self.burn_in + stride, stride=stride,
                           seed=seed, shape=self._shaping(region.dimension),
                           timed=self.profile_timing)
        result = out[0].cpu().numpy()
        if self.profile_timing:
            # Slacks, direction products and random draws are fused into the
            # transition kernel; its device time is reported as transitions.
            self.last_phase_timings = {
                "slack_matmul_seconds": 0.0, "constraint_product_matmul_seconds": 0.0,
                "random_generation_seconds": 0.0,
                "transitions_seconds": batch.last_kernel_ms * 1e-3,
                "wall_seconds": time.perf_counter() - t0,
            }
        return result

    def _check_random_settings(self):
        pass

    def _initial_points(self, region, n_chains, rng, x0):
        """art/sampling.py:817-852."""
        if x0 is None:
            seed = find_feasible_point(region, bounds=self.bounds,
                                       max_tries=self.max_feasible_tries, random_state=rng,
                                       tol=self.feasibility_tol)
            return np.repeat(seed[None, :], n_chains, axis=0)
        candidates = np.asarray(x0, dtype=float)
        if candidates.ndim == 1:
            candidates = candidates[None, :]
        if candidates.ndim != 2 or candidates.shape[1] != region.dimension:
            raise ValueError(f"x0 must have shape ({region.dimension},) or "
                             f"(n_points, {region.dimension}), got {candidates.shape}.")
        if candidates.shape[0] == 0:
            raise ValueError("x0 must contain at least one starting point.")
        if not np.all(region.contains(candidates, tol=self.feasibility_tol)):
            raise ValueError("All x0 candidates must be feasible for the region.")
        idx = rng.choice(candidates.shape[0], size=n_chains,
                         replace=candidates.shape[0] < n_chains)
        return candidates[idx].copy()


@dataclass
class NumbaBatchedHitAndRunSampler(BatchedHitAndRunSampler):
    """Reference name of the compiled batched sampler (art/sampling.py:965-970)."""

    precompute_ad: bool = field(default=True, init=False)


@dataclass
class ParallelNumbaBatchedHitAndRunSampler(NumbaBatchedHitAndRunSampler):
    """Reference name of the chain-parallel sampler (art/sampling.py:973-1025)."""

    n_chains: int = 50
    parallel_random: bool = True
    random_threads: int = 10
    _chain_major: ClassVar[bool] = True

    def _check_random_settings(self):
        if self.parallel_random and self.random_threads < 1:
            raise ValueError("random_threads must be at least 1.")


def sample_polytopes(regions, n, *, n_chains=64, burn_in=0, thinning=10, seed=0, x0=None,
                     shape=None, stream=None):
    """Batched-sampler semantics for many polytopes at once: (P, n, d) device samples.

    x0: (P, d) one start per polytope (every chain starts there) or
    (P, chains, d).
    """
    batch = PolytopeBatch(regions, stream=stream)
    x0 = np.asarray(x0, dtype=float)
    chains = min(int(n_chains), int(n))
    if x0.ndim == 2:
        x0 = np.repeat(x0[:, None, :], chains, axis=1)
    stride = max(1, int(thinning))
    return batch.sample(x0, n, first=int(burn_in) + stride, stride=stride, seed=seed, shape=shape)


def sample_hit_and_run_many(samplers, regions, ns, random_states, x0s):
    """``samplers[k].sample(regions[k], ns[k], random_states[k], x0s[k])`` for
    every k, with one GPU launch per (burn_in, thinning) group.

    Every request must give its start x0.  Request k draws exactly what its
    own ``sample`` call draws (one Philox seed from random_states[k]) and gets
    the same samples: per-polytope keys and shaping matrices make the chains
    of one launch independent, and a chain recording fewer samples than the
    group's longest is the prefix of the longer run.  Returns a list of (n_k,
    d) host arrays, or the exception a request raised, per request.
    """
    out = [None] * len(regions)
    groups = {}
    for k, (s, region, n, x0) in enumerate(zip(samplers, regions, ns, x0s)):
        try:
            if n < 0:
                raise ValueError("n must be nonnegative.")
            if s.burn_in < 0:
                raise ValueError("burn_in must be nonnegative.")
            if s.thinning < 1:
                raise ValueError("thinning must be at least 1.")
            rng = as_rng(random_states[k])
            shape = s._shaping(region.dimension)
            x = np.asarray(x0, dtype=float).reshape(-1)
            if x.shape[0] != region.dimension:
                raise ValueError(f"x0 must have shape ({region.dimension},), got {x.shape}.")
            if not bool(region.contains(x, tol=s.feasibility_tol)[0]):
                raise ValueError("x0 is not feasible for the region.")
            if int(n) == 0:
                out[k] = np.empty((0, region.dimension))
                continue
            seed = int(rng.integers(0, np.iinfo(np.int64).max))
        except (ValueError, RuntimeError) as exc:
            out[k] = exc
            continue
        groups.setdefault((s.burn_in, s.thinning, region.dimension), []).append(
            (k, region, int(n), x, seed, shape))
    for (burn_in, thinning, d), items in groups.items():
        n_max = max(n for _, _, n, _, _, _ in items)
        shapes = None
        if any(sh is not None for *_, sh in items):
            shapes = np.stack([np.eye(d) if sh is None else sh for *_, sh in items])
        batch = PolytopeBatch([region for _, region, *_ in items])
        x0 = np.stack([x for _, _, _, x, _, _ in items])[:, None, :]
        res = batch.sample(x0, n_max, first=burn_in + 1, stride=thinning,
                           seeds=[seed for *_, seed, _ in items], shape=shapes)
        host = res.cpu().numpy()
        for i, (k, _, n, *_rest) in enumerate(items):
            out[k] = host[i, :n].copy()
    return out
