"""Batched weighted ridge fits over variable-length leaves (CSR row ranges).

The reference fits every leaf separately on the CPU with
``solve_weighted_ridge(solver="normal")`` (art/models.py:218-343) after
``PolynomialRidgeModel.prepare_design`` (art/models.py:602-620); its GPU
batch (art/_cuda_batched_cholesky.py) needs equal-size slots.  Here one
resident plan holds the chunking of K uneven leaves and every fit builds the
features from X on the fly, contracts the weighted Gram by MFMA and solves
each leaf's system on the device.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _hip
from .features import _columns

PRECISIONS = {"fp64": _hip.PRECISION_FP64, "mixed": _hip.PRECISION_MIXED,
              "fp64_refined": _hip.PRECISION_FP64_REFINED}
DEFAULT_CHUNK_ROWS = 4096  # the library's default for leaf plans (chunk_rows=0)


class LeafRidgeBatch:
    """Resident plan for K leaves given by CSR ``offsets`` (K+1,).

    ``fit(X, y, w, ridge)`` takes device tensors X (N,d), y (N,), w (N,)
    float64 and returns device tensors ``coefficients (K,p)``, ``status (K,)``
    and ``ridge_effective (K,)`` (overwritten by the next fit).  Status codes
    are those of the batched Cholesky workspace; a negative or non-finite
    weight flags its leaf with -1.
    """

    def __init__(self, offsets, d, *, degree=2, include_bias=True, bias_last=False,
                 chunk_rows=0, stream=None):
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        if offsets.ndim != 1 or offsets.size < 2 or offsets[0] != 0:
            raise ValueError("offsets must be a one-dimensional CSR array starting at 0.")
        if np.any(np.diff(offsets) < 0):
            raise ValueError("offsets must be nondecreasing.")
        if not isinstance(d, (int, np.integer)) or d < 1:
            raise ValueError("d must be a positive integer.")
        columns = _columns(int(d), degree, include_bias, bias_last)
        self.K, self.d, self.p = offsets.size - 1, int(d), columns.shape[1]
        self.n_rows = int(offsets[-1])
        self.offsets = offsets
        self.degree, self.include_bias, self.bias_last = degree, bool(include_bias), bool(bias_last)
        self.plan = None
        torch = _hip.torch_device()
        self.torch, self.lib = torch, _hip.load()
        self.device_id = torch.cuda.current_device()
        self.device = torch.device("cuda", self.device_id)
        self.stream = torch.cuda.current_stream() if stream is None else stream
        handle = ctypes.c_void_p()
        _hip.check(self.lib.art_leaf_plan_create(
            offsets.ctypes.data, self.K, self.d, int(degree), int(include_bias), int(bias_last),
            int(chunk_rows), ctypes.byref(handle)))
        self.plan = handle
        p, nchunks, tiles = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _hip.check(self.lib.art_plan_info(self.plan, ctypes.byref(p), ctypes.byref(nchunks),
                                          ctypes.byref(tiles)))
        self.n_chunks, self.tiles = nchunks.value, tiles.value
        f64 = dict(dtype=torch.float64, device=self.device)
        self.coefficients = torch.full((self.K, self.p), float("nan"), **f64)
        self.ridge_effective = torch.zeros(self.K, **f64)
        self.status = torch.zeros(self.K, dtype=torch.int32, device=self.device)

    @classmethod
    def autotuned(cls, offsets, d, X, y, w, ridge, *, degree=2, include_bias=True,
                  bias_last=False, precision="fp64", stream=None, candidates=None):
        """A batch whose chunk size the autotuner picked for these leaves
        (``art.autotune.tune_chunk_rows``: timed once per shape key, then
        cached); ``batch.tuning`` holds (chunk_rows, timings or None)."""
        from . import autotune

        kwargs = {} if candidates is None else {"candidates": candidates}
        chunk, timings = autotune.tune_chunk_rows(
            offsets, d, X, y, w, ridge, degree=degree, include_bias=include_bias,
            bias_last=bias_last, precision=precision, stream=stream, **kwargs)
        batch = cls(offsets, d, degree=degree, include_bias=include_bias, bias_last=bias_last,
                    chunk_rows=chunk, stream=stream)
        batch.tuning = (chunk, timings)
        return batch

    def _check_inputs(self, X, y, w):
        torch = self.torch
        for name, t, shape in (("X", X, (self.n_rows, self.d)), ("y", y, (self.n_rows,)),
                               ("w", w, (self.n_rows,))):
            if (not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float64
                    or not t.is_contiguous() or t.device.index != self.device_id):
                raise ValueError(f"{name} must be a C-contiguous float64 tensor on the device.")
            if tuple(t.shape) != shape:
                raise ValueError(f"{name} must have shape {shape}.")

    def fit(self, X, y, w, ridge, weight_floor=1e-12, precision="fp64"):
        """Enqueue one fit of every leaf.

        precision="fp64": FP64 Gram (MFMA) and FP64 Cholesky.
        precision="mixed" (p <= 64): split-bf16 MFMA Gram with an exact FP64
        diagonal and right-hand side, one FP64 residual refinement, and the
        FP64 path for every leaf whose correction exceeds the tolerance.
        precision="fp64_refined" (p <= 64): the FP64 fit followed by FP64
        data-space refinement to the least-squares solution (the "auto"
        solver's answer for ill-conditioned leaves).
        """
        if self.plan is None:
            raise RuntimeError("The plan is closed.")
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {tuple(PRECISIONS)}.")
        ridge, weight_floor = float(ridge), float(weight_floor)
        if not np.isfinite(ridge) or ridge < 0:
            raise ValueError("ridge must be finite and nonnegative.")
        if not np.isfinite(weight_floor) or weight_floor < 0:
            raise ValueError("weight_floor must be finite and nonnegative.")
        self._check_inputs(X, y, w)
        _hip.check(self.lib.art_leaf_ridge_fit(
            self.plan, _hip.ptr(X), _hip.ptr(y), _hip.ptr(w), ridge, weight_floor,
            PRECISIONS[precision], _hip.ptr(self.coefficients), _hip.ptr(self.status),
            _hip.ptr(self.ridge_effective), _hip.stream_handle(self.stream)))
        return self.coefficients, self.status, self.ridge_effective

    def capture(self, X, y, w, ridge, weight_floor=1e-12, precision="fp64"):
        """Record one fit of these inputs as a HIP graph; returns a
        ``LeafFitGraph`` whose ``replay()`` reruns every kernel of the fit
        (Gram, solve, refinement rounds, fallback) as one graph launch on
        this batch's stream, writing the same outputs as ``fit``.

        The fit is run once eagerly first, so that every buffer it needs
        exists before capture (a captured fit allocates nothing and issues
        no memset or host synchronization).  X, y and w stay bound to the
        graph: refill them in place between replays to fit new values.
        Capturing switches the plan's kernel timing off (its event records
        would otherwise enter the graph)."""
        if self.plan is None:
            raise RuntimeError("The plan is closed.")
        torch = self.torch
        _hip.check(self.lib.art_plan_timing_enable(self.plan, 0))
        self.fit(X, y, w, ridge, weight_floor, precision)
        self.synchronize()
        # Capture on a stream of the graph's own (the default stream cannot be
        # captured); the batch's fits go there for the duration.
        side = torch.cuda.Stream(device=self.device)
        graph = torch.cuda.CUDAGraph()
        home = self.stream
        try:
            self.stream = side
            with torch.cuda.graph(graph, stream=side):
                self.fit(X, y, w, ridge, weight_floor, precision)
        finally:
            self.stream = home
        side.synchronize()
        return LeafFitGraph(self, graph, side, (X, y, w))

    def grams(self, X, y, w, weight_floor=1e-12, *, precision="fp64", mask=None, out=None):
        """Each leaf's augmented system, without ridge: a device tensor
        (K, gram_size) in the kernels' tile layout, and its invalid-weight
        flags (K,).  Systems of disjoint row sets add (rank-k refits).

        precision "fp64" (default) or "mixed" (the split-bf16 MFMA system of
        the mixed fit, p <= 64, d <= 16: entry (i, j) within 4e-7 of
        sqrt(G_ii G_jj) of the FP64 system, ~1e-8 on large leaves).
        mask, a (K,) device int32 tensor, selects the leaves to form; the
        others keep what ``out`` = (grams, bad) held (fresh tensors when out
        is None, so pass out with a mask)."""
        torch = self.torch
        if self.plan is None:
            raise RuntimeError("The plan is closed.")
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'.")
        self._check_inputs(X, y, w)
        size = self.lib.art_gram_size(self.p)
        if out is None:
            grams = torch.empty((self.K, size), dtype=torch.float64, device=self.device)
            bad = torch.empty(self.K, dtype=torch.int32, device=self.device)
        else:
            grams, bad = out
            if grams.shape != (self.K, size) or grams.dtype != torch.float64 \
                    or not grams.is_contiguous() or bad.shape != (self.K,) \
                    or bad.dtype != torch.int32 or not bad.is_contiguous():
                raise ValueError("out must be contiguous (K, gram_size) float64 and (K,) int32 "
                                 "device tensors.")
        if mask is not None and (mask.shape != (self.K,) or mask.dtype != torch.int32
                                 or not mask.is_contiguous() or mask.device != grams.device):
            raise ValueError("mask must be a contiguous (K,) int32 tensor on the batch's device.")
        code = _hip.PRECISION_MIXED if precision == "mixed" else _hip.PRECISION_FP64
        _hip.check(self.lib.art_leaf_gram_ex(
            self.plan, _hip.ptr(X), _hip.ptr(y), _hip.ptr(w), float(weight_floor), code,
            None if mask is None else _hip.ptr(mask), _hip.ptr(grams), _hip.ptr(bad),
            _hip.stream_handle(self.stream)))
        return grams, bad

    def rcond(self):
        """Device tensor (K,): 1 / (|A|_1 |A^-1|_1) of each leaf's regularized
        system from the last "mixed" or "fp64_refined" fit (0 where the fit
        failed); the quantity LAPACK dpocon estimates (art/models.py:288-296)."""
        if self.plan is None:
            raise RuntimeError("The plan is closed.")
        out = self.torch.empty(self.K, dtype=self.torch.float64, device=self.device)
        _hip.check(self.lib.art_leaf_rcond(self.plan, _hip.ptr(self.status), _hip.ptr(out),
                                           _hip.stream_handle(self.stream)))
        return out

    def fallback_count(self):
        """Leaves that took the exact FP64 path in the last mixed fit (synchronizes)."""
        count = ctypes.c_int()
        _hip.check(self.lib.art_leaf_fallback_count(self.plan, ctypes.byref(count)))
        return count.value

    def set_refine_steps(self, steps):
        """Refinement steps after the first, on leaves still flagged (default 1)."""
        _hip.check(self.lib.art_leaf_set_refine_steps(self.plan, int(steps)))

    def set_refine_tol(self, tol):
        """Acceptance tolerance |delta| <= tol |beta| of the mixed refinement."""
        _hip.check(self.lib.art_leaf_set_refine_tol(self.plan, float(tol)))

    def synchronize(self):
        self.stream.synchronize()

    def close(self):
        if getattr(self, "plan", None) is not None:
            self.synchronize()
            self.lib.art_plan_destroy(self.plan)
            self.plan = None

    def __enter__(self):
        return self

    def __exit__(self, *args):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LeafFitGraph:
    """A captured leaf fit (``LeafRidgeBatch.capture``): ``replay()`` enqueues
    the whole fit as one graph launch, ordered after the work already on the
    batch's stream and before the work enqueued there next, and returns the
    batch's (coefficients, status, ridge_effective) tensors."""

    def __init__(self, batch, graph, stream, inputs):
        self.batch, self.graph, self.stream, self.inputs = batch, graph, stream, inputs

    def replay(self):
        batch = self.batch
        if batch.plan is None:
            raise RuntimeError("The plan is closed.")
        self.stream.wait_stream(batch.stream)
        with batch.torch.cuda.stream(self.stream):
            self.graph.replay()
        batch.stream.wait_stream(self.stream)
        return batch.coefficients, batch.status, batch.ridge_effective


def solve_grams(grams, bad, d, ridge, *, degree=2, include_bias=True, bias_last=False,
                stream=None):
    """Fit every resident system of ``LeafRidgeBatch.grams`` (summed over any
    number of row sets): returns device tensors (coefficients (K, p), status
    (K,), ridge_effective (K,)), as ``LeafRidgeBatch.fit`` in FP64."""
    torch = _hip.torch_device()
    lib = _hip.load()
    p = _columns(int(d), degree, include_bias, bias_last).shape[1]
    K = int(grams.shape[0])
    if grams.dtype != torch.float64 or not grams.is_contiguous() or grams.dim() != 2 \
            or grams.shape[1] != lib.art_gram_size(p):
        raise ValueError("grams must be a contiguous float64 (K, gram_size) device tensor.")
    stream = torch.cuda.current_stream() if stream is None else stream
    f64 = dict(dtype=torch.float64, device=grams.device)
    coef = torch.empty((K, p), **f64)
    status = torch.empty(K, dtype=torch.int32, device=grams.device)
    lam = torch.empty(K, **f64)
    _hip.check(lib.art_gram_solve(_hip.ptr(grams), _hip.ptr(bad.contiguous()), K,
                                  _hip.Basis(int(d), p, int(degree), int(include_bias),
                                             int(bias_last)), float(ridge), _hip.ptr(coef),
                                  _hip.ptr(status), _hip.ptr(lam), _hip.stream_handle(stream)))
    return coef, status, lam
