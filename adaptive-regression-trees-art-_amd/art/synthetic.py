"""Seeded synthetic workloads for the leaf-ridge, split and sampler paths.

These generate inputs of the shapes the hot path is benchmarked on; they are
plain NumPy and hold no reference arithmetic.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def sigmoid(t):
    """Logistic gate with the +-50 clip of art/objectives.py:49-51."""
    return 1.0 / (1.0 + np.exp(-np.clip(t, -50.0, 50.0)))


@dataclass
class BumpMixture:
    """y(x) = sum_k A_k exp(-||x - c_k||^2 / (2 s_k^2))."""

    centers: np.ndarray
    widths: np.ndarray
    amplitudes: np.ndarray

    @classmethod
    def random(cls, rng, n_bumps, d):
        return cls(rng.uniform(-1.0, 1.0, size=(n_bumps, d)),
                   rng.uniform(0.1, 0.5, size=n_bumps),
                   rng.uniform(-1.0, 1.0, size=n_bumps))

    def __call__(self, X):
        X = np.asarray(X, dtype=np.float64)
        out = np.zeros(X.shape[0])
        for c, s, a in zip(self.centers, self.widths, self.amplitudes):
            out += a * np.exp(-np.sum((X - c) ** 2, axis=1) / (2.0 * s * s))
        return out


@dataclass
class LeafRidgeWorkload:
    """Variable-length weighted leaves in CSR form.

    X (N,d) float64 row-major, y (N,), w (N,) in (0,1], offsets (K+1,) int64.
    """

    X: np.ndarray
    y: np.ndarray
    w: np.ndarray
    offsets: np.ndarray
    ridge: float
    degree: int = 2

    @property
    def n_leaves(self):
        return len(self.offsets) - 1

    @property
    def n_rows(self):
        return int(self.offsets[-1])

    def leaf(self, j):
        s, e = int(self.offsets[j]), int(self.offsets[j + 1])
        return self.X[s:e], self.y[s:e], self.w[s:e]

    def subset(self, leaves):
        """A new workload holding only the given leaves, in the given order."""
        leaves = np.asarray(leaves, dtype=np.int64)
        parts = [self.leaf(j) for j in leaves]
        sizes = np.array([len(p[1]) for p in parts], dtype=np.int64)
        offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        return LeafRidgeWorkload(np.concatenate([p[0] for p in parts]),
                                 np.concatenate([p[1] for p in parts]),
                                 np.concatenate([p[2] for p in parts]),
                                 offsets, self.ridge, self.degree)


def leaf_ridge_workload(n_leaves=4096, d=8, min_rows=64, max_rows=16384,
                        n_bumps=6, ridge=1e-6, seed=2, degree=2):
    """Batched weighted-ridge leaves (default: 4096 leaves, d=8, ~12M rows).

    Leaf sizes are log-uniform in [min_rows, max_rows]; x ~ U[-1,1]^d; each
    leaf's weights are the product of 1..6 sigmoid gates
    sigma((a.x - c)/tau), a a unit Gaussian direction, c ~ U[-.5,.5],
    tau ~ U[.05,.5]; y is a sum of ``n_bumps`` Gaussian bumps with centers in
    [-1,1]^d, widths U[.1,.5] and amplitudes U[-1,1].
    """
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(np.log(min_rows), np.log(max_rows),
                                        size=n_leaves))).astype(np.int64)
    sizes = np.clip(sizes, min_rows, max_rows)
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    N = int(offsets[-1])
    X = rng.uniform(-1.0, 1.0, size=(N, d))
    target = BumpMixture.random(rng, n_bumps, d)
    y = target(X)
    w = np.empty(N)
    n_gates = rng.integers(1, 7, size=n_leaves)
    for j in range(n_leaves):
        s, e = offsets[j], offsets[j + 1]
        g = n_gates[j]
        a = rng.normal(size=(g, d))
        a /= np.linalg.norm(a, axis=1, keepdims=True)
        c = rng.uniform(-0.5, 0.5, size=g)
        tau = rng.uniform(0.05, 0.5, size=g)
        w[s:e] = np.prod(sigmoid((X[s:e] @ a.T - c) / tau), axis=1)
    return LeafRidgeWorkload(X, y, w, offsets, float(ridge), degree)


@dataclass
class PolytopeWorkload:
    """P polytopes: the box [-1, 1]^d (2d rows) plus k_j random cuts a.x <= b
    with unit Gaussian a and b = a.x*_j + 0.05 + U[0, 0.5], so the seeded
    interior point x*_j keeps slack >= 0.05 on every cut (and >= 0.05 on the
    box, x* ~ U[-0.95, 0.95]^d).  CSR rows: A (rows, d), b (rows,),
    offsets (P+1,); starts (P, d) = x*."""

    A: np.ndarray
    b: np.ndarray
    offsets: np.ndarray
    starts: np.ndarray

    @property
    def n_polytopes(self):
        return len(self.offsets) - 1

    def polytope(self, j):
        s, e = int(self.offsets[j]), int(self.offsets[j + 1])
        return self.A[s:e], self.b[s:e]


def polytope_workload(n_polytopes=65536, d=12, min_cuts=4, max_cuts=40, seed=4):
    rng = np.random.default_rng(seed)
    P = int(n_polytopes)
    starts = rng.uniform(-0.95, 0.95, size=(P, d))
    cuts = rng.integers(min_cuts, max_cuts + 1, size=P)
    rows = 2 * d + cuts
    offsets = np.concatenate([[0], np.cumsum(rows)]).astype(np.int64)
    A = np.empty((int(offsets[-1]), d))
    b = np.empty(int(offsets[-1]))
    eye = np.eye(d)
    a_all = rng.normal(size=(int(cuts.sum()), d))
    a_all /= np.linalg.norm(a_all, axis=1, keepdims=True)
    slack = 0.05 + rng.uniform(0.0, 0.5, size=int(cuts.sum()))
    owner = np.repeat(np.arange(P), cuts)
    b_cut = np.einsum("ij,ij->i", a_all, starts[owner]) + slack
    cut_start = np.concatenate([[0], np.cumsum(cuts)])
    for j in range(P):
        s = int(offsets[j])
        A[s:s + d] = eye
        A[s + d:s + 2 * d] = -eye
        b[s:s + 2 * d] = 1.0
        c0, c1 = int(cut_start[j]), int(cut_start[j + 1])
        A[s + 2 * d:s + 2 * d + (c1 - c0)] = a_all[c0:c1]
        b[s + 2 * d:s + 2 * d + (c1 - c0)] = b_cut[c0:c1]
    return PolytopeWorkload(A, b, offsets, starts)


def split_workload(n_nodes=2048, n=8192, d=16, n_bumps=8, seed=3):
    """Equal-size split nodes: X (K, n, d) ~ U[-1, 1], y (K, n) a seeded
    n_bumps-bump Gaussian mixture per node."""
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1.0, 1.0, size=(n_nodes, n, d))
    y = np.empty((n_nodes, n))
    for k in range(n_nodes):
        y[k] = BumpMixture.random(rng, n_bumps, d)(X[k])
    return X, y


def split_workload_device(n_nodes, device, n=8192, d=16, n_bumps=8, seed=3, block=64):
    """The split workload's shape generated on the device: node k's rows
    X_k ~ U[-1, 1]^d from a torch generator seeded (seed, k) and its target
    a BumpMixture drawn from numpy's default_rng(seed) in node order (the
    bumps of split_workload's distribution; the rows are another seeded
    stream).  Returns float64 device tensors X (K, n, d), y (K, n) and the
    per-node mixtures."""
    import torch

    rng = np.random.default_rng(seed)
    mixtures = [BumpMixture.random(rng, n_bumps, d) for _ in range(n_nodes)]
    X = torch.empty((n_nodes, n, d), dtype=torch.float64, device=device)
    y = torch.zeros((n_nodes, n), dtype=torch.float64, device=device)
    g = torch.Generator(device=device)
    for k in range(n_nodes):
        g.manual_seed(int(seed) * 1_000_003 + k)
        X[k].uniform_(-1.0, 1.0, generator=g)
    for k0 in range(0, n_nodes, block):
        k1 = min(k0 + block, n_nodes)
        c = torch.as_tensor(np.stack([m.centers for m in mixtures[k0:k1]]), device=device)
        w = torch.as_tensor(np.stack([m.widths for m in mixtures[k0:k1]]), device=device)
        a = torch.as_tensor(np.stack([m.amplitudes for m in mixtures[k0:k1]]), device=device)
        for b in range(n_bumps):
            r2 = ((X[k0:k1] - c[:, b][:, None, :]) ** 2).sum(dim=2)
            y[k0:k1] += a[:, b][:, None] * torch.exp(-r2 / (2.0 * w[:, b][:, None] ** 2))
    return X, y, mixtures


class PlainFitOracle:
    """The 2-D plain-fit target: f(x) = sin(3 x_1) cos(2 x_2)
    + 0.5 exp(-8 |x - (0.3, -0.2)|^2), plus N(0, noise^2) from a seeded
    generator (every call draws fresh noise, like a real experiment)."""

    def __init__(self, noise=1e-3, seed=0):
        self.noise = float(noise)
        self.rng = np.random.default_rng(seed)

    @staticmethod
    def clean(X):
        X = np.asarray(X, dtype=float)
        r2 = (X[..., 0] - 0.3) ** 2 + (X[..., 1] + 0.2) ** 2
        return np.sin(3.0 * X[..., 0]) * np.cos(2.0 * X[..., 1]) + 0.5 * np.exp(-8.0 * r2)

    def __call__(self, X):
        f = self.clean(X)
        return f + self.noise * self.rng.normal(size=np.shape(f))


def plain_fit_builder(seed=0, node_batch_size=32, error_threshold=1e-2, **overrides):
    """Builder of the plain ART fit: [-1, 1]^2, 4000 samples per node, soft
    oblique splits at T = 0.05 on the batched device optimizer (BFS cohorts),
    quadratic ridge leaves (lambda = 1e-6), max depth 6, at least 64 points
    per side, relative L2 stopping error 1e-2."""
    from .builder import RegressionTreeBuilder
    from .domain import BoxDomain
    from .models import PolynomialRidgeModel
    from .splitters import SoftObliqueSplitter
    from .temperature import TemperatureConfig

    opts = dict(
        domain=BoxDomain.hypercube(2, -1.0, 1.0), oracle=PlainFitOracle(seed=seed),
        oracle_vectorized=True, model_template=PolynomialRidgeModel(2, ridge=1e-6),
        splitter=SoftObliqueSplitter(PolynomialRidgeModel(2, ridge=1e-6), 0.05,
                                     min_side_points=64, optimizer="cupy_batched_cholesky"),
        temperature_config=TemperatureConfig(strategy="splitter"),
        error_threshold=error_threshold, sample_count=4000, max_depth=6, traversal="bfs",
        node_batch_size=node_batch_size, random_state=seed)
    opts.update(overrides)
    return RegressionTreeBuilder(**opts)


@dataclass
class ServingEvent:
    """One request of the serving stream: a fit of a new tree or a refit
    adding rows to an earlier fit's tree.  ``rows`` and ``seed`` fix its data."""

    arrival: float
    kind: str            # "fit" or "refit"
    d: int
    rows: int
    seed: int
    target: BumpMixture
    tree: int            # index of the fit this event creates or refits


def serving_stream(n_requests, *, rate=200.0, dims=(2, 4, 8, 12, 16, 20), min_rows=1000,
                   max_rows=2_000_000, refit_fraction=0.3, refit_rows=(0.01, 0.10),
                   bumps=(4, 16), seed=5):
    """Poisson arrivals at ``rate``/s; fits draw d uniformly from ``dims`` and
    N log-uniformly from [min_rows, max_rows]; with probability
    ``refit_fraction`` (once a fit exists) an event instead adds a uniform
    1-10% of an earlier fit's N as new rows of the same target."""
    rng = np.random.default_rng(seed)
    t, events, fits = 0.0, [], []
    for _ in range(n_requests):
        t += rng.exponential(1.0 / rate)
        if fits and rng.uniform() < refit_fraction:
            f = events[fits[rng.integers(len(fits))]]
            k = max(1, int(np.ceil(rng.uniform(*refit_rows) * f.rows)))
            events.append(ServingEvent(t, "refit", f.d, k, int(rng.integers(2**31)), f.target,
                                       f.tree))
            continue
        d = int(rng.choice(dims))
        n = int(np.exp(rng.uniform(np.log(min_rows), np.log(max_rows))))
        target = BumpMixture.random(rng, int(rng.integers(bumps[0], bumps[1] + 1)), d)
        events.append(ServingEvent(t, "fit", d, n, int(rng.integers(2**31)), target, len(fits)))
        fits.append(len(events) - 1)
    return events


def serving_rows(event, device):
    """The event's rows on the device: X ~ U[-1,1]^d from its seed, y its
    bump mixture (float64 tensors)."""
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(event.seed)
    X = torch.rand((event.rows, event.d), generator=g, dtype=torch.float64, device=device)
    X.mul_(2.0).sub_(1.0)
    y = torch.zeros(event.rows, dtype=torch.float64, device=device)
    m = event.target
    for c, s, a in zip(m.centers, m.widths, m.amplitudes):
        cd = torch.as_tensor(c, dtype=torch.float64, device=device)
        y.add_(torch.exp(-((X - cd) ** 2).sum(dim=1) / (2.0 * s * s)), alpha=float(a))
    return X, y
