"""Batched soft-oblique split optimization on the GPU (device-controlled).

Drop-in for ``BatchedOptimizerWorkspace`` / ``optimize_batch`` of
art/_cuda_batched_optimizer.py:88-660: same constructor, ``allocate``,
``load``, ``optimize`` and ``close``, the same validation, result types
(``SplitResult`` or ``SplitNotFoundError`` per requested slot), stop reasons
and metadata keys.  Where the reference runs a host loop with one
device->host transfer per iteration and per backtrack batch, this workspace
makes one launch of ``art_split_optimize`` for every (node, restart) problem
of the call and downloads the results once; when a launch has too few
problems to fill the device, each problem's rows are spread over several
workgroups of one cooperative launch (``art_split_optimize_parts``, same
results up to the order of the partial sums).  Random starts (Gaussian unit w,
z = median(X w), art/splitters.py:692-700) are drawn on the device
(``art_split_init``: Philox4x32-10 normals seeded from the splitter's RNG,
radix-select median).
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _hip
from .batched_ridge import _dimensions, _node_ids
from .features import feature_count
from .models import AffineRidgeModel, PolynomialRidgeModel, RidgeSolveResult, model_basis
from .optimizers import AdaptiveAlpha
from .splitters import (SplitNotFoundError, SplitResult, _partition_result, project_unit_w,
                        project_unit_w_rows)

STOP_REASONS = ("max_iters", "gradient_tolerance", "non_descent_direction", "line_search_failed",
                "max_line_search_failures", "alpha_min_reached", "cholesky_failure",
                "min_side_points", "nonfinite_split")
ERROR_MESSAGES = {
    "cholesky_failure": "GPU fit or gradient failed.",
    "min_side_points": "Hard split has too few samples on one side.",
    "nonfinite_split": "Final partition is nonfinite.",
}
MAX_P, MAX_D = 32, 31
SCALARS = 6
# A node's rows may be spread over up to MAX_PARTS workgroups of one
# cooperative launch (art_split_optimize_parts): one part per PART_ROWS rows
# (rounded up) of the launch's largest node, the size of one staged chunk of
# the narrow kernel.  On 4000-row nodes (d = 2, 32 nodes a launch) 8 parts
# take the split searches of a tree build from 73.8 to 31.9 ms.
MAX_PARTS = 8
PART_ROWS = 512


def _batch_data(X, y, temperature):
    """art/_cuda_batched_optimizer.py:46-56."""
    X, y = np.asarray(X, dtype=float), np.asarray(y, dtype=float)
    if X.ndim != 3 or min(X.shape) == 0 or y.shape != X.shape[:2]:
        raise ValueError("X/y must have nonempty shapes (K,n,d)/(K,n).")
    if not np.isfinite(X).all() or not np.isfinite(y).all():
        raise ValueError("X and y must be finite.")
    temperature = np.broadcast_to(np.asarray(temperature, dtype=float), (X.shape[0],))
    if not np.isfinite(temperature).all() or np.any(temperature <= 0):
        raise ValueError("Temperatures must be finite and positive.")
    _dimensions(*X.shape)
    return X, y, temperature


def _validate_settings(s, d, reorder):
    """art/_cuda_batched_optimizer.py:59-85."""
    if type(s.model_template) not in (AffineRidgeModel, PolynomialRidgeModel):
        raise TypeError("Batched Cholesky supports affine and polynomial ridge models.")
    if type(s.model_template) is PolynomialRidgeModel and (
            not isinstance(s.model_template.degree, (int, np.integer))
            or s.model_template.degree not in (1, 2)):
        raise ValueError("Batched GPU features support only polynomial degree 1 or 2.")
    if s.optimizer == "hybrid_gram" or s.refit_during_line_search:
        raise ValueError("Batched optimization requires refit_during_line_search=False and a "
                         "resident optimizer.")
    if not isinstance(reorder, (bool, np.bool_)):
        raise ValueError("reorder must be a boolean.")
    s._validate_inputs(np.empty((0, d)), np.empty(0))
    for name in ("alpha0", "alpha_min", "alpha_max", "grad_atol", "grad_rtol", "weight_floor"):
        value = getattr(s, name)
        if not np.isfinite(value) or value < 0 or (name.startswith("alpha") and value == 0):
            raise ValueError(f"Invalid {name}.")
    for name, minimum in (("max_iters", 0), ("max_backtracks", 0), ("Xw_refresh", 0),
                          ("backtrack_batch_size", 1), ("max_line_search_failures", 1),
                          ("n_restarts", 0)):
        if not isinstance(getattr(s, name), (int, np.integer)) or getattr(s, name) < minimum:
            raise ValueError(f"Invalid {name}.")
    if not 0 < s.rho < 1 or not 0 < s.armijo_c < 1:
        raise ValueError("rho and armijo_c must lie strictly between 0 and 1.")
    if not np.isfinite(s.model_template.ridge) or s.model_template.ridge < 0:
        raise ValueError("ridge must be finite and nonnegative.")
    AdaptiveAlpha(s.alpha0, s.alpha_min, s.alpha_max, s.alpha_grow, s.alpha_recovery,
                  s.heavy_backtrack_threshold)


class BatchedOptimizerWorkspace:
    """Allocate capacity K once; load and optimize cohorts of k <= K nodes.

    X/y are (k,n,d)/(k,n); temperatures may vary by node.  optimize() returns
    a list in requested slot order, each entry a SplitResult or a
    SplitNotFoundError; exports hold host data only.  ``reorder`` is accepted
    for API compatibility: every problem runs to completion in one launch,
    so there is no active-prefix compaction to choose.
    """

    def __init__(self, splitter, X, y, *, temperature=None, reorder=True):
        temp = splitter.temperature if temperature is None else temperature
        X, y, temp = _batch_data(X, y, temp)
        self._allocate(splitter, *X.shape, reorder=reorder)
        self._load(X, y, temp)

    @classmethod
    def allocate(cls, splitter, capacity, n, d, *, reorder=True, stream=None):
        ws = cls.__new__(cls)
        ws._allocate(splitter, capacity, n, d, reorder=reorder, stream=stream)
        return ws

    def _allocate(self, splitter, capacity, n, d, *, reorder, stream=None):
        self._setup(splitter, capacity, n, d, reorder=reorder, stream=stream)
        torch = self.torch
        f64 = dict(dtype=torch.float64, device=self.device)
        self.x = torch.empty((self.K, self.n, self.d), **f64)
        self.yd = torch.empty((self.K, self.n), **f64)
        self.host_X = np.empty((self.K, self.n, self.d))
        self.host_y = np.empty((self.K, self.n))
        self.X, self.y = self.host_X[:0], self.host_y[:0]

    def _setup(self, splitter, capacity, n, d, *, reorder, stream=None):
        self.lib = None
        self.K, self.n, self.d = _dimensions(capacity, n, d)
        _validate_settings(splitter, self.d, reorder)
        self.s, self.capacity, self.k = splitter, self.K, 0
        self.reorder = bool(reorder)
        self.seed = int(splitter._rng().integers(0, np.iinfo(np.int64).max))
        self._init_calls = 0
        self.required = splitter._required_side_points(self.n)
        self.degree, self.bias, self.bias_last = model_basis(splitter.model_template)
        self.p = feature_count(self.d, self.degree, self.bias)
        if self.p > MAX_P or self.d > MAX_D:
            raise ValueError(f"The device split optimizer supports p <= {MAX_P} and d <= {MAX_D}.")
        self.R = max(1, int(splitter.n_restarts))
        self.S = int(splitter.max_backtracks) + 1
        self.H = min(int(splitter.backtrack_batch_size), self.S)
        torch = _hip.torch_device()
        self.torch, self.lib = torch, _hip.load()
        self.device = torch.device("cuda", torch.cuda.current_device())
        self.stream = torch.cuda.current_stream() if stream is None else stream
        f64 = dict(dtype=torch.float64, device=self.device)
        self.temp = torch.ones(self.K, **f64)
        self.temperature = np.ones(self.K)
        self.all_ids = np.arange(self.K, dtype=np.int32)
        self.fit_sizes = []
        self.last_launch_problems = 0
        # Workgroups per problem: None picks by node size and device room;
        # an int fixes it (1 is one workgroup per problem).
        self.split_parts = None
        self.last_launch_parts = 1
        self.part_retries = 0  # multi-part launches rerun on one workgroup per problem
        self._part_bufs = None

    def _check_open(self):
        if self.lib is None:
            raise RuntimeError("The workspace is closed.")

    def load(self, X, y, *, temperature=None):
        """Load one packed cohort into the fixed device buffers."""
        self._check_open()
        temp = self.s.temperature if temperature is None else temperature
        X, y, temp = _batch_data(X, y, temp)
        if X.shape[0] > self.K or X.shape[1:] != (self.n, self.d):
            raise ValueError("X/y must have shapes (k,n,d)/(k,n), 0<k<=capacity.")
        return self._load(X, y, temp)

    def load_device(self, X, y, *, temperature=None):
        """Load a cohort already resident on the device: X (k, n, d) and
        y (k, n) contiguous float64 tensors on the workspace's device, copied
        into the fixed buffers on the workspace stream (no host staging; the
        data are finite by the caller's contract).  The host views X and y
        stay empty until a host load."""
        self._check_open()
        torch = self.torch
        if (not isinstance(X, torch.Tensor) or not isinstance(y, torch.Tensor) or not X.is_cuda
                or X.dtype != torch.float64 or y.dtype != torch.float64 or X.dim() != 3
                or tuple(y.shape) != tuple(X.shape[:2])):
            raise ValueError("Device X/y must be float64 tensors of shapes (k,n,d)/(k,n).")
        k = int(X.shape[0])
        if not 0 < k <= self.K or tuple(X.shape[1:]) != (self.n, self.d):
            raise ValueError("X/y must have shapes (k,n,d)/(k,n), 0<k<=capacity.")
        temp = np.broadcast_to(np.asarray(self.s.temperature if temperature is None
                                          else temperature, dtype=float), (k,))
        if not np.isfinite(temp).all() or np.any(temp <= 0):
            raise ValueError("Temperatures must be finite and positive.")
        self.temperature[:k] = temp
        with torch.cuda.stream(self.stream):
            self.x[:k].copy_(X)
            self.yd[:k].copy_(y)
            self.temp[:k].copy_(torch.from_numpy(self.temperature[:k].copy()))
        self.k = k
        self.X, self.y = self.host_X[:0], self.host_y[:0]
        self.fit_sizes.clear()
        return self

    def _load(self, X, y, temp):
        k = X.shape[0]
        self.k = 0
        for dst, src in ((self.host_X[:k], X), (self.host_y[:k], y)):
            if dst.ctypes.data != src.ctypes.data or dst.strides != src.strides:
                np.copyto(dst, src)
        self.temperature[:k] = temp
        torch = self.torch
        with torch.cuda.stream(self.stream):
            self.x[:k].copy_(torch.from_numpy(self.host_X[:k]))
            self.yd[:k].copy_(torch.from_numpy(self.host_y[:k]))
            self.temp[:k].copy_(torch.from_numpy(self.temperature[:k].copy()))
        self.k = k
        self.X, self.y = self.host_X[:k], self.host_y[:k]
        self.fit_sizes.clear()
        return self

    def _device_starts(self, nodes, timing=None):
        """Random starts (Gaussian unit w, z = median(X w)) for node slots, drawn
        on the device (art_split_init); each call advances the counter.
        timing, a list, receives the kernel's milliseconds (HIP events on the
        workspace stream)."""
        torch, P = self.torch, len(nodes)
        if P == 0:
            return np.empty((0, self.d + 1))
        node_t = torch.as_tensor(np.ascontiguousarray(nodes, dtype=np.int32), device=self.device)
        out = torch.empty((P, self.d + 1), dtype=torch.float64, device=self.device)
        with torch.cuda.stream(self.stream):
            if timing is not None:
                ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
                ev[0].record(self.stream)
            _hip.check(self.lib.art_split_init(
                _hip.ptr(self.x), self.n, self.d, _hip.ptr(node_t), P,
                self.seed & 0xFFFFFFFFFFFFFFFF, self._init_calls, _hip.ptr(out),
                _hip.stream_handle(self.stream)))
            if timing is not None:
                ev[1].record(self.stream)
        self._init_calls += 1
        self.stream.synchronize()
        if timing is not None:
            timing.append(ev[0].elapsed_time(ev[1]))
        return out.cpu().numpy()

    def _launch(self, nodes, theta0, mode, fallback, hard_refit, anneal=None, timing=None,
                batched=False, child_precision=_hip.PRECISION_FP64, lean=False):
        """One device launch over problems (nodes[i], theta0[i]); host copies out.

        anneal = (iters, t0, t1, step) for mode 1; timing, a list, receives
        the launch's device milliseconds.  lean: only theta, the hard-refit
        results and info come back (beta, histories and scalars are None)."""
        torch, s, P = self.torch, self.s, len(nodes)
        mi = int(s.max_iters) if anneal is None else int(anneal[0])
        dev = dict(device=self.device)
        f64 = dict(dtype=torch.float64, **dev)
        node_t = torch.as_tensor(np.ascontiguousarray(nodes, dtype=np.int32), **dev)
        theta_t = torch.as_tensor(np.ascontiguousarray(theta0, dtype=np.float64), **dev)
        out_theta = torch.empty((P, self.d + 1), **f64)
        out_beta = torch.empty((P, 2, self.p), **f64)
        hard = torch.zeros((P, 8), **f64)
        info = torch.zeros((P, 8), dtype=torch.int32, **dev)
        if lean:  # written by the kernel, never read back
            hist = torch.empty((P, 5, mi + 1), **f64)
        else:
            hist = torch.full((P, 5, mi + 1), float("nan"), **f64)
        scal = torch.zeros((P, SCALARS), **f64)
        m = s.model_template
        a = _hip.SplitArgs()
        a.X, a.y, a.temperature = _hip.ptr(self.x), _hip.ptr(self.yd), _hip.ptr(self.temp)
        a.theta0, a.node = _hip.ptr(theta_t), _hip.ptr(node_t)
        a.n, a.d = self.n, self.d
        a.basis = _hip.Basis(self.d, self.p, self.degree, int(self.bias), int(self.bias_last))
        a.ridge, a.floor = float(m.ridge), float(s.weight_floor)
        a.grad_atol, a.grad_rtol = float(s.grad_atol), float(s.grad_rtol)
        a.alpha0, a.rho, a.armijo_c = float(s.alpha0), float(s.rho), float(s.armijo_c)
        a.alpha_min, a.alpha_max = float(s.alpha_min), float(s.alpha_max)
        a.grow, a.recovery = float(s.alpha_grow), float(s.alpha_recovery)
        a.max_iters, a.max_backtracks = mi, int(s.max_backtracks)
        a.heavy, a.max_failures = int(s.heavy_backtrack_threshold), int(s.max_line_search_failures)
        a.adaptive, a.required = int(bool(s.adaptive_alpha)), int(self.required)
        a.fallback, a.hard_refit, a.H, a.mode = int(bool(fallback)), int(bool(hard_refit)), self.H, mode
        if anneal is None:
            a.anneal_iters, a.t0, a.t1, a.step = 0, 0.0, 0.0, 0.0
        else:
            a.anneal_iters, a.t0, a.t1, a.step = int(anneal[0]), *map(float, anneal[1:])
        a.theta_out, a.beta_out, a.hard = _hip.ptr(out_theta), _hip.ptr(out_beta), _hip.ptr(hard)
        a.info, a.history, a.scalars = _hip.ptr(info), _hip.ptr(hist), _hip.ptr(scal)
        a.child_precision = int(child_precision)
        self._ragged_args(a)
        with torch.cuda.stream(self.stream):
            if timing is not None:
                ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
                ev[0].record(self.stream)
            parts = 1 if batched else self._parts(a, nodes)
            if parts > 1:
                count, buf, failed = self._part_buffers(P, parts)
                rc = self.lib.art_split_optimize_parts(
                    ctypes.byref(a), P, parts, _hip.ptr(count), _hip.ptr(buf),
                    _hip.ptr(failed), _hip.stream_handle(self.stream))
                if rc == _hip.ERR_HIP:  # the runtime refused the launch: nothing ran
                    parts = 1
                else:
                    _hip.check(rc)
            if parts == 1:
                entry = (self.lib.art_split_anneal_batched if batched
                         else self.lib.art_split_optimize)
                _hip.check(entry(ctypes.byref(a), P, _hip.stream_handle(self.stream)))
            if timing is not None:
                ev[1].record(self.stream)
        self.last_launch_problems = P
        self.last_launch_parts = parts
        self.stream.synchronize()
        if parts > 1 and int(failed.item()):
            # A part waited too long for the others (the device was shared with
            # other work): discard the launch and run it on one workgroup per
            # problem, from cleared outputs.
            self.part_retries += 1
            with torch.cuda.stream(self.stream):
                hard.zero_()
                info.zero_()
                scal.zero_()
                if not lean:
                    hist.fill_(float("nan"))
                _hip.check(self.lib.art_split_optimize(ctypes.byref(a), P,
                                                       _hip.stream_handle(self.stream)))
            self.last_launch_parts = 1
            self.stream.synchronize()
        if timing is not None:
            timing.append(ev[0].elapsed_time(ev[1]))
        if lean:
            return out_theta.cpu().numpy(), None, hard.cpu().numpy(), info.cpu().numpy(), None, \
                None
        return (out_theta.cpu().numpy(), out_beta.cpu().numpy(), hard.cpu().numpy(),
                info.cpu().numpy(), hist.cpu().numpy(), scal.cpu().numpy())

    def optimize(self, theta0=None, *, parent_loss=None, nodes=None, min_side_fallback=None,
                 hard_refit=True):
        """Run every restart of every requested node; keep each node's best."""
        self._check_open()
        if not self.k:
            raise RuntimeError("Load data before optimizing.")
        ids = self.all_ids[:self.k] if nodes is None else _node_ids(nodes, self.k)
        count, R, s = len(ids), self.R, self.s
        if not isinstance(hard_refit, (bool, np.bool_)):
            raise ValueError("hard_refit must be a boolean.")
        if not hard_refit and s.n_restarts != 1:
            raise ValueError("Skipping hard refits requires n_restarts == 1.")
        initial = None if theta0 is None else np.asarray(theta0, dtype=float)
        if initial is not None and (initial.shape not in ((count, self.d + 1), (R, count, self.d + 1))
                                    or not np.isfinite(initial).all()):
            raise ValueError("theta0 must be finite with shape (r,d+1) or (R,r,d+1).")
        if parent_loss is not None:
            parent_loss = np.asarray(parent_loss, dtype=float)
            if parent_loss.shape != (count,) or not np.isfinite(parent_loss).all() \
                    or np.any(parent_loss < 0):
                raise ValueError("parent_loss must be finite and nonnegative with shape (r,).")
        if min_side_fallback is not None and not isinstance(min_side_fallback, (bool, np.bool_)):
            raise ValueError("min_side_fallback must be a boolean.")
        if not count:
            return []
        fallback = s.min_side_fallback if min_side_fallback is None else bool(min_side_fallback)
        self.fit_sizes.clear()
        if parent_loss is None:
            zero = np.zeros((count, self.d + 1))
            zero[:, 0] = 1.0
            _, _, hard, info, _, _ = self._launch(ids, zero, 2, False, True)
            parent_loss, parent_ok = hard[:, 0].copy(), info[:, 6] == 0
            self.fit_sizes.append(("parent", count))
        else:
            parent_ok = np.ones(count, dtype=bool)
        # Problems: restart-major, only nodes whose parent fit succeeded.
        prob_node, prob_theta, prob_slot, prob_random = [], [], [], []
        for r in range(R):
            for i, j in enumerate(ids):
                if not parent_ok[i]:
                    continue
                random = initial is None or (r > 0 and initial.ndim == 2)
                start = None if random else project_unit_w(
                    initial[i] if initial.ndim == 2 else initial[r, i])
                prob_random.append(random)
                prob_node.append(j)
                prob_theta.append(start)
                prob_slot.append((r, i))
        draw = [q for q, rnd in enumerate(prob_random) if rnd]
        for q, start in zip(draw, self._device_starts([prob_node[q] for q in draw])):
            prob_theta[q] = start
        results = [None] * count
        for i in range(count):
            if not parent_ok[i]:
                results[i] = SplitNotFoundError("cholesky_failure", "Parent Cholesky fit failed.")
        if not prob_node:
            return results
        theta, beta, hard, info, hist, scal = self._launch(
            np.asarray(prob_node), np.asarray(prob_theta), 0, fallback, hard_refit)
        self._log_fits(prob_slot, info, count, hard_refit)
        best = [None] * count
        for q, (r, i) in enumerate(prob_slot):
            j = int(ids[i])
            meta = self._metadata(j, q, theta, hard, info, hist, scal, hard_refit, prob_random[q])
            meta["initial_theta"] = np.array(prob_theta[q], dtype=float)
            code = int(info[q, 0])
            if info[q, 6]:
                reason = STOP_REASONS[code]
                meta["stop_reason"] = reason
                results[i] = SplitNotFoundError(reason, ERROR_MESSAGES.get(reason, reason),
                                                diagnostics=meta)
                continue
            if not hard_refit:
                results[i] = _partition_result(theta[q], self._rows(j), hard[q, 3],
                                               parent_loss[i], meta["converged"], meta)
                continue
            fb = bool(hard[q, 7])
            b = best[i]
            if b is None or (b[1] and not fb) or (b[1] == fb and hard[q, 0] < b[2]):
                best[i] = (q, fb, hard[q, 0], meta)
        if hard_refit:
            for i in range(count):
                if best[i] is not None:
                    q, _, _, meta = best[i]
                    results[i] = self._export(q, theta[q], beta[q], hard[q], meta, parent_loss[i],
                                              int(ids[i]))
        return results

    def hyperplanes(self, theta0, *, timing=None):
        """Lean ``optimize`` for callers that need only each node's split.

        theta0 (R, k, d+1): starts of every restart of the k loaded nodes.
        Runs the same device launch (hard refit, no min-side fallback unless
        the splitter asks for it) and keeps ``optimize``'s choice of restart,
        but skips the parent-loss launch and the per-node result objects.
        Returns host arrays theta (k, d+1), ok (k,) and reason (k,) -- the
        stop reason code of a node without a split, else -1.
        """
        self._check_open()
        starts = np.asarray(theta0, dtype=float)
        R, count = starts.shape[0], self.k
        if starts.shape != (R, count, self.d + 1) or not np.isfinite(starts).all():
            raise ValueError("theta0 must be finite with shape (R, k, d+1).")
        nodes = np.tile(self.all_ids[:count], R)
        thetas = project_unit_w_rows(starts.reshape(-1, self.d + 1))
        theta, _, hard, info, _, _ = self._launch(nodes, thetas, 0, self.s.min_side_fallback,
                                                  True, timing=timing, lean=True)
        if R == 1:  # one start per node: its split, or its stop reason
            ok = info[:count, 6] == 0
            out = np.where(ok[:, None], theta[:count], 0.0)
            reason = np.where(ok, -1, info[:count, 0]).astype(np.int64)
            return out, ok, reason
        out = np.zeros((count, self.d + 1))
        ok = np.zeros(count, dtype=bool)
        reason = np.full(count, -1, dtype=np.int64)
        best = [None] * count
        for q in range(R * count):
            i = q % count
            if info[q, 6]:
                if best[i] is None:
                    reason[i] = int(info[q, 0])
                continue
            fb = bool(hard[q, 7])
            b = best[i]
            if b is None or (b[1] and not fb) or (b[1] == fb and hard[q, 0] < b[2]):
                best[i] = (q, fb, hard[q, 0])
        for i, b in enumerate(best):
            if b is not None:
                out[i], ok[i], reason[i] = theta[b[0]], True, -1
        return out, ok, reason

    def anneal(self, theta0=None, *, nodes=None, t0=0.2, t1=0.01, iters=200, step=0.05,
               timing=None, batched=None, child_precision="fp64"):
        """Annealed fixed-step projected gradient from every start, then a hard refit.

        theta0: (r, d+1) or (R, r, d+1) starts for the r selected nodes, or
        None for random starts drawn on the device (Gaussian unit w,
        z = median(X w)).  Iteration
        k uses T_k = t0 (t1/t0)^(k/(iters-1)) and
        theta <- retract(theta - step * P(grad)).  No reference counterpart
        (the reference optimizes at a fixed temperature).  Returns a dict of
        host arrays per (restart, node): theta (R, r, d+1), beta (R, r, 2, p),
        hard loss (R, r), n_right (R, r), ok (R, r), soft-loss history
        (R, r, iters), and best (r,): the restart with the lowest hard loss.
        timing, a list, receives the milliseconds of each device launch: the
        random starts (when theta0 is None), then the annealer.

        batched (default: when d <= 16 and p <= 17)
        runs 16 starts of a node per workgroup, the node's rows staged once
        per pass for all of them (art_split_anneal_batched); otherwise one
        workgroup per start.

        child_precision="mixed" (batched only) forms the soft fits' 32 child
        systems per workgroup by split-bf16 MFMA with FP32 accumulation into
        FP64; gates, predictions, gradients, solves and the hard refit stay
        FP64.  "fp64" (default) is FP64 throughout.
        """
        if child_precision not in ("fp64", "mixed"):
            raise ValueError("child_precision must be 'fp64' or 'mixed'.")
        self._check_open()
        if not self.k:
            raise RuntimeError("Load data before optimizing.")
        ids = self.all_ids[:self.k] if nodes is None else _node_ids(nodes, self.k)
        count, R = len(ids), self.R
        if iters < 1 or not (t0 > 0 and t1 > 0 and np.isfinite(step)):
            raise ValueError("Need iters >= 1, positive temperatures and a finite step.")
        if theta0 is None:
            starts = self._device_starts(np.tile(ids, R), timing).reshape(R, count, self.d + 1)
        else:
            starts = np.asarray(theta0, dtype=float)
            if starts.ndim == 2:
                starts = starts[None]
            if starts.shape[1:] != (count, self.d + 1) or not np.isfinite(starts).all():
                raise ValueError("theta0 must be finite with shape (r,d+1) or (R,r,d+1).")
            starts = np.stack([[project_unit_w(t) for t in rs] for rs in starts])
        Rr = starts.shape[0]
        eligible = self.d <= 16 and self.p <= 17
        if batched is None:
            batched = eligible
        if batched and not eligible:
            raise ValueError("The batched annealer needs d <= 16 and at most 17 basis features.")
        if child_precision == "mixed" and not batched:
            raise ValueError("child_precision='mixed' needs the batched annealer.")
        child = _hip.PRECISION_MIXED if child_precision == "mixed" else _hip.PRECISION_FP64
        if batched:
            # Node-major blocks of 16 starts; a node's last block repeats its
            # last start (results of the copies are dropped).
            B = -(-Rr // 16) * 16
            order = np.minimum(np.arange(B), Rr - 1)
            nodes_all = np.repeat(ids, B)
            thetas = starts.transpose(1, 0, 2)[:, order].reshape(-1, self.d + 1)
            outs = self._launch(nodes_all, thetas, 1, False, True,
                                anneal=(iters, t0, t1, step), timing=timing, batched=True,
                                child_precision=child)
            keep = (np.arange(count)[:, None] * B + np.arange(Rr)[None, :]).T.reshape(-1)
            theta, beta, hard, info, hist, _ = (o[keep] for o in outs)
        else:
            nodes_all = np.tile(ids, Rr)
            theta, beta, hard, info, hist, _ = self._launch(
                nodes_all, starts.reshape(-1, self.d + 1), 1, False, True,
                anneal=(iters, t0, t1, step), timing=timing)
        shape = (Rr, count)
        ok = info[:, 6].reshape(shape) == 0
        loss = np.where(ok, hard[:, 0].reshape(shape), np.inf)
        return {"theta": theta.reshape(shape + (self.d + 1,)),
                "beta": beta.reshape(shape + (2, self.p)), "loss": loss,
                "n_right": hard[:, 3].reshape(shape).astype(int), "ok": ok,
                "info": info.reshape(shape + (8,)),
                "soft_loss_history": hist[:, 0, :iters].reshape(shape + (iters,)),
                "best": np.argmin(loss, axis=0), "starts": starts}

    def _log_fits(self, prob_slot, info, count, hard_refit):
        """Record the logical fit batches of the device loop in ``fit_sizes``.

        Per restart: soft round t fits every problem still iterating (it >= t)
        with reorder, or every requested node without; then one hard batch
        (art/_cuda_batched_optimizer.py:277-282,413-430).
        """
        slots = np.array([r for r, _ in prob_slot])
        for r in range(self.R):
            q = np.flatnonzero(slots == r)
            if not q.size:
                continue
            fits = info[q, 2] + 1
            for t in range(int(fits.max())):
                self.fit_sizes.append(("soft", int(np.count_nonzero(fits > t))
                                       if self.reorder else count))
            finished = int(np.count_nonzero(info[q, 7]))
            if hard_refit and finished:
                self.fit_sizes.append(("hard", finished if self.reorder else count))

    def _metadata(self, j, q, theta, hard, info, hist, scal, hard_refit, random_init):
        s = self.s
        it = int(info[q, 2])
        mi = int(s.max_iters)
        loss = [float(v) for v in hist[q, 0, :it + 1]]
        gnorm = [float(v) for v in hist[q, 1, :it + 1]]
        meta = dict(
            soft_loss_history=loss, projected_grad_norm_history=gnorm,
            step_size_history=[float(v) for v in hist[q, 2, :it]],
            backtrack_history=[int(v) for v in hist[q, 3, :it]],
            line_search_loss_history=[float(v) for v in hist[q, 4, :it]],
            optimizer="cupy_batched_cholesky", device_optimizer="hip_persistent",
            reorder=self.reorder, initialization="gpu" if random_init else "provided",
            initialization_seed=self.seed if random_init else None, gpu_solver="cholesky",
            solver_fallback=False, condition_estimation=False, complementary_grams=False,
            backtrack_batch_size=self.H, Xw_refresh=s.Xw_refresh,
            temperature=float(self.temperature[j]), n_iters=it,
            stop_reason=STOP_REASONS[int(info[q, 0])] if not info[q, 6] else None,
            converged=bool(info[q, 1]), grad_tolerance=float(scal[q, 1]),
            initial_grad_norm=gnorm[0] if gnorm else None,
            final_grad_norm=gnorm[-1] if gnorm else None,
            adaptive_alpha=s.adaptive_alpha, refit_during_line_search=False,
            physical_line_search_candidates=int(info[q, 4]),
            logical_line_search_candidates=int(info[q, 5]),
            speculative_line_search_candidates=int(info[q, 4] - info[q, 5]),
            final_alpha=float(scal[q, 2]), hard_refit=bool(hard_refit),
            n_left=self._rows(j) - int(hard[q, 3]), n_right=int(hard[q, 3]),
            left_loss=float(hard[q, 1]) if hard_refit else None,
            right_loss=float(hard[q, 2]) if hard_refit else None,
            required_side_points=self._required(j),
        )
        meta["final_soft_metadata"] = {
            "temperature": float(self.temperature[j]),
            "model_type": type(s.model_template).__name__, "models_refit": True,
            "prepared_design": not self.bias_last,
            **{f"{side}_solve": {"solver_requested": "normal", "solver_used": "cholesky",
                                 "condition_estimator": None, "cond_estimate": None, "rank": None,
                                 "fallback_reason": None, "ridge_effective": float(scal[q, 4 + k])}
               for k, side in enumerate(("left", "right"))},
        }
        if hard[q, 7]:
            meta.update(min_side_fallback=True, fallback_iteration=int(info[q, 3]))
        return meta

    def _export(self, q, theta, beta, hard, meta, parent_loss, j):
        models = []
        for side in range(2):
            model = self.s.model_template.clone()
            model.coef_ = beta[side].copy()
            model.solve_result_ = RidgeSolveResult(
                coefficients=model.coef_, solver_requested="normal", solver_used="cholesky",
                condition_estimator=None, cond_estimate=None, ridge_effective=float(hard[4 + side]),
                rank=None)
            models.append(model)
        right = int(hard[3])
        gain = float(parent_loss - hard[0])
        return SplitResult(w=theta[:-1].copy(), z=float(theta[-1]), left_model=models[0],
                           right_model=models[1], loss=float(hard[0]),
                           parent_loss=float(parent_loss), split_gain=gain,
                           relative_split_gain=gain / max(float(parent_loss), 1e-12),
                           n_left=self._rows(j) - right, n_right=right, converged=meta["converged"],
                           n_iters=meta["n_iters"], stop_reason=meta["stop_reason"], metadata=meta)

    def _parts(self, a, nodes):
        """Workgroups per problem for one launch: the fixed split_parts, else
        one per PART_ROWS rows of the largest node (at most MAX_PARTS),
        capped by what the device holds at once for this many problems."""
        if self.split_parts is not None:
            return max(1, min(MAX_PARTS, int(self.split_parts)))
        want = min(MAX_PARTS, -(-self._max_rows(nodes) // PART_ROWS))
        if want < 2:
            return 1
        return max(1, min(want, int(self.lib.art_split_max_parts(ctypes.byref(a), len(nodes)))))

    def _part_buffers(self, problems, parts):
        """Exchange buffers of a multi-part launch (kept and grown as needed)."""
        torch = self.torch
        need = int(self.lib.art_split_part_doubles(self.p, parts)) * problems
        bufs = self._part_bufs
        if bufs is None or bufs[0].numel() < problems or bufs[1].numel() < need:
            dev = dict(device=self.device)
            bufs = (torch.empty(max(problems, 1), dtype=torch.int32, **dev),
                    torch.empty(need, dtype=torch.float64, **dev),
                    torch.empty(1, dtype=torch.int32, **dev))
            self._part_bufs = bufs
        return bufs

    def _max_rows(self, nodes):
        return self.n

    # Node sizes: every slot holds n rows here; the ragged workspace overrides.
    def _rows(self, j):
        return self.n

    def _required(self, j):
        return self.required

    def _ragged_args(self, a):
        pass

    def close(self):
        if getattr(self, "lib", None) is not None:
            self.stream.synchronize()
            self.lib = None
            self.x = self.yd = self.temp = None
            self._part_bufs = None
            self.k = 0

    def __enter__(self):
        return self

    def __exit__(self, *args):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RaggedOptimizerWorkspace(BatchedOptimizerWorkspace):
    """Nodes of uneven size in one device launch.

    Node k owns rows [offsets[k], offsets[k+1]) of X (N, d) and y (N,), kept
    concatenated on the device; each problem reads its own range and needs
    max(min_side_points, ceil(min_side_fraction * n_k)) points per side.
    ``optimize`` keeps the signature, results and metadata of
    ``BatchedOptimizerWorkspace.optimize``.  This is the uneven-cohort form
    the reference's workspace lacks (it requires (K, n, d) blocks,
    art/_cuda_batched_optimizer.py:88-103); the batched annealer needs equal
    sizes, so ``anneal`` runs one workgroup per start here.
    """

    def __init__(self, splitter, X, y, offsets, *, temperature=None):
        torch = _hip.torch_device()
        on_device = isinstance(X, torch.Tensor)
        if on_device:
            # Device rows (already resident): float64, C-contiguous, finite by
            # the caller's contract; they are used in place.
            if (not isinstance(y, torch.Tensor) or not X.is_cuda or not y.is_cuda
                    or X.dtype != torch.float64 or y.dtype != torch.float64
                    or not X.is_contiguous() or not y.is_contiguous()):
                raise ValueError("Device X and y must be contiguous float64 tensors.")
            y = y.reshape(-1)
        else:
            X = np.ascontiguousarray(X, dtype=float)
            y = np.ascontiguousarray(y, dtype=float).reshape(-1)
        offsets = np.ascontiguousarray(offsets, dtype=np.int64).reshape(-1)
        if X.ndim != 2 or X.shape[0] != y.shape[0] or offsets.size < 2:
            raise ValueError("X/y must have shapes (N,d)/(N,) and offsets (K+1,).")
        sizes = np.diff(offsets)
        if offsets[0] != 0 or offsets[-1] != X.shape[0] or np.any(sizes < 2):
            raise ValueError("offsets must run from 0 to N with at least two rows per node.")
        if not on_device and (not np.isfinite(X).all() or not np.isfinite(y).all()):
            raise ValueError("X and y must be finite.")
        K = sizes.size
        temp = np.broadcast_to(np.asarray(splitter.temperature if temperature is None
                                          else temperature, dtype=float), (K,))
        if not np.isfinite(temp).all() or np.any(temp <= 0):
            raise ValueError("Temperatures must be finite and positive.")
        self._setup(splitter, K, int(sizes.max()), X.shape[1], reorder=True)
        self.sizes = sizes
        self.required_k = np.maximum(int(splitter.min_side_points),
                                     np.ceil(splitter.min_side_fraction * sizes).astype(np.int64))
        torch = self.torch
        with torch.cuda.stream(self.stream):
            self.x = X if on_device else torch.from_numpy(X).to(self.device)
            self.yd = y if on_device else torch.from_numpy(y).to(self.device)
            self.offsets_dev = torch.from_numpy(offsets).to(self.device)
            self.temperature[:] = temp
            self.temp.copy_(torch.from_numpy(self.temperature.copy()))
        self.X, self.y = X, y
        self.offsets = offsets
        self.k = K

    def load(self, *args, **kwargs):
        raise RuntimeError("A ragged workspace takes its data at construction.")

    def _rows(self, j):
        return int(self.sizes[j])

    def _max_rows(self, nodes):
        return int(self.sizes[np.asarray(nodes, dtype=np.int64)].max()) if len(nodes) else 0

    def _required(self, j):
        return int(self.required_k[j])

    def _ragged_args(self, a):
        a.n = 0
        a.row_offsets = _hip.ptr(self.offsets_dev)
        a.min_side_points = int(self.s.min_side_points)
        a.min_side_fraction = float(self.s.min_side_fraction)

    def _device_starts(self, nodes, timing=None, keys=None):
        """Random starts of node slots; keys (one uint64 per start) make each
        start depend on its key alone, not on its position in the call."""
        torch, P = self.torch, len(nodes)
        if P == 0:
            return np.empty((0, self.d + 1))
        node_t = torch.as_tensor(np.ascontiguousarray(nodes, dtype=np.int32), device=self.device)
        out = torch.empty((P, self.d + 1), dtype=torch.float64, device=self.device)
        key_t = None
        if keys is not None:
            kv = np.ascontiguousarray(np.asarray(keys, dtype=np.uint64).reshape(-1))
            if kv.shape[0] != P:
                raise ValueError("keys must give one key per start.")
            key_t = torch.from_numpy(kv.view(np.int64)).to(self.device)
        with torch.cuda.stream(self.stream):
            _hip.check(self.lib.art_split_init_ranges(
                _hip.ptr(self.x), _hip.ptr(self.offsets_dev), self.d, _hip.ptr(node_t), P,
                self.seed & 0xFFFFFFFFFFFFFFFF, self._init_calls,
                0 if key_t is None else _hip.ptr(key_t), _hip.ptr(out),
                _hip.stream_handle(self.stream)))
        self._init_calls += 1
        self.stream.synchronize()
        return out.cpu().numpy()

    def anneal(self, theta0=None, **kwargs):
        kwargs["batched"] = False
        return super().anneal(theta0, **kwargs)

    def close(self):
        super().close()
        self.offsets_dev = None


def optimize_batch(splitter, X, y, *, theta0=None, parent_loss=None, temperature=None,
                   reorder=True, hard_refit=True):
    """Convenience entry point (art/_cuda_batched_optimizer.py:656-660)."""
    with BatchedOptimizerWorkspace(splitter, X, y, temperature=temperature, reorder=reorder) as ws:
        return ws.optimize(theta0, parent_loss=parent_loss, hard_refit=hard_refit)
