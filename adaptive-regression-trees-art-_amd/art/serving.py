"""Serving loop: continuously batched tree fits and rank-k refits on one GPU.

The reference builds one tree per process call (builder.py) and has no
request loop.  ``ServingEngine`` accepts fit and refit requests from any
thread and serves them from one scheduler thread:

* **Fit** (``FitRequest``): a data-driven oblique regression tree on the
  request's rows.  Every level, each frontier node's weighted ridge model
  comes from its augmented system ``[Phi | y]^T W [Phi | y]`` (FP64 MFMA,
  ``art_leaf_gram``) solved on the device (``art_gram_solve``); its weighted
  SSE is read from the same system (``art_gram_sse``).  A node stays a leaf
  at ``max_depth``, below ``2 * min_leaf`` rows or at a weighted MSE of at
  most ``error_threshold``; otherwise the soft-oblique split optimizer
  (``RaggedOptimizerWorkspace``) searches a strided sample of at most
  ``split_rows`` of its rows from keyed random starts, and
  ``art_partition_rows`` moves the rows to its children (right when
  ``w.x - z >= 0``) in the next level buffer.
* **Refit** (``RefitRequest``): new rows of a resident tree are routed to
  their leaves (``art_tree_leaf``), their systems are added to the leaves'
  resident systems (a rank-k update ``G += Phi_new^T W Phi_new``) and every
  leaf is re-solved; the tree's structure is kept.

Continuous batching: all requests of one basis (d, degree) share one level
buffer, so one launch of each kernel serves every admitted request's
frontier; requests that arrive while a level runs are admitted at the next
level, as slots and HBM budget allow.  Every node's result depends only on
its own rows and key, so a request's result is the same alone and in any
mix (bit for bit).

Leaf systems are exact FP64: they stay resident for refits, and a reduced
precision system would carry its error into every later rank-k update.

``ServingConfig(precision="mixed")`` (bases with p <= 64 and d <= 16) forms
the statistics of the nodes that may still split -- whose fit is read only
for the stop test -- as split-bf16 MFMA systems (``art_leaf_gram_ex``: entry (i, j) within 4e-7
of sqrt(G_ii G_jj) of the FP64 system), and FP64 systems only for the nodes
that end as leaves: those the depth and size rules stop are formed in FP64
directly, the others once the mixed statistics (or a failed split search, or
a partition that leaves one side empty) make them leaves, from the rows still
in their level buffer.  Every kept system and leaf model is the FP64 one, so
refits are exact and a leaf's result is that of the FP64 engine whenever the
stop tests agree; the FP64 Gram pass over a request's rows drops from one per
level to one in all.
"""

from __future__ import annotations

import collections
import itertools
import threading
import time
from concurrent.futures import Future, ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import Any, Optional

import numpy as np

from . import _hip

_MASK64 = (1 << 64) - 1


class AdmissionError(RuntimeError):
    """A request that can never fit in the engine's HBM budget."""


@dataclass
class FitRequest:
    """Fit a tree to rows X (N, d), y (N,), optional weights w (N,) >= 0.

    X, y and w may be host arrays or float64 device tensors (used in place).
    ``degree`` defaults to 2 for d <= 12 and 1 above (leaf basis
    [1, x, x_i x_j] or [1, x]).  ``seed`` keys the split search's random
    starts; ``tree_id`` names the resident tree (default: a fresh name).
    """

    X: Any
    y: Any
    w: Any = None
    max_depth: int = 10
    min_leaf: int = 64
    degree: Optional[int] = None
    error_threshold: float = 0.0
    temperature: float = 0.05
    seed: int = 0
    tree_id: Optional[str] = None


@dataclass
class RefitRequest:
    """Add rows X (k, d), y (k,), w (k,) to resident tree ``tree_id``."""

    tree_id: str
    X: Any
    y: Any
    w: Any = None


@dataclass
class RequestTiming:
    submitted: float
    admitted: float = float("nan")
    finished: float = float("nan")

    @property
    def latency(self):
        return self.finished - self.submitted

    @property
    def queued(self):
        return self.admitted - self.submitted


@dataclass
class FitResult:
    tree_id: str
    coefficients: np.ndarray   # (leaves, p) in the flattened tree's leaf order
    status: np.ndarray         # (leaves,) solve status per leaf (0 = ok)
    n_rows: int
    timing: RequestTiming
    resident: Any = field(default=None, repr=False)

    @property
    def tree(self):
        """The fitted ``RegressionTree`` (built on first access; a refit
        updates its leaf models in place)."""
        return self.resident.tree


@dataclass
class RefitResult:
    tree_id: str
    coefficients: np.ndarray
    status: np.ndarray
    n_new_rows: int
    timing: RequestTiming
    resident: Any = field(default=None, repr=False)

    @property
    def tree(self):
        return self.resident.tree


@dataclass
class ServingConfig:
    """Engine-wide settings (shared by every request of a launch)."""

    hbm_budget_bytes: float = 200e9
    max_active: int = 512          # fit requests on the device at once
    ridge: float = 1e-6            # leaf ridge (trace-scaled, models.py:209-215)
    weight_floor: float = 1e-12
    split_rows: int = 4096         # split-search sample per node
    split_ridge: float = 1e-4
    n_restarts: int = 1
    max_iters: int = 30
    min_side_fraction: float = 0.05
    quadratic_max_d: int = 12
    precision: str = "fp64"        # "fp64", or "mixed": split-bf16 statistics of open nodes


def _splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & _MASK64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & _MASK64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & _MASK64
    return x ^ (x >> 31)


def start_key(seed, heap, restart):
    """Random-start key of node ``heap`` (root 1, children 2h and 2h+1)."""
    return _splitmix64((int(seed) & _MASK64) ^ _splitmix64((int(heap) << 8) | int(restart)))


# Mixed-precision stop test: a node goes to the split search on its mixed
# statistics only when its mixed MSE exceeds the error threshold by
# _MIXED_SSE_MARGIN * S^2 / sum(w), S = sum_i |b_i| sqrt(G_ii) + sqrt(y'Wy);
# the others are decided on their FP64 fit.  The split-bf16 system's entries
# are within 4e-7 of sqrt(G_ii G_jj), so the SSE y'Wy - 2 b'h + b'G b,
# i.e. [b; -1]' G [b; -1], errs by at most 4e-7 S^2 at a given b; the margin
# leaves a factor of 25 for the fitted coefficients' own error.  Ill-
# conditioned fits (large cancelling b) widen S and go to FP64 more often.
_MIXED_SSE_MARGIN = 1e-5


def _diag_offset(col, p):
    """Offset of diagonal entry (col, col) in a resident system of the
    (p+1)-column augmented matrix (tile pairs (t1 <= t2), t1 outermost)."""
    T = (p + 1 + 15) // 16
    t, a = divmod(int(col), 16)
    q = sum(T - u for u in range(t))
    return 256 * q + 16 * a + a


def _heap_path(heap):
    """Node id of the reference's scheme ("root", "root/L", ...) for a heap index."""
    bits = bin(heap)[3:]
    return "/".join(["root"] + ["L" if b == "0" else "R" for b in bits])


# --------------------------------------------------------------------------- device helpers

def gram_sse(grams, beta, p, *, bias_col=0, stream=None):
    """Weighted SSE and weight sum of each system's fit (device tensors (K,))."""
    torch = _hip.torch_device()
    lib = _hip.load()
    K = int(grams.shape[0])
    sse = torch.empty(K, dtype=torch.float64, device=grams.device)
    wsum = torch.empty(K, dtype=torch.float64, device=grams.device)
    stream = torch.cuda.current_stream() if stream is None else stream
    _hip.check(lib.art_gram_sse(_hip.ptr(grams), K, int(p), int(bias_col), _hip.ptr(beta),
                                _hip.ptr(sse), _hip.ptr(wsum), _hip.stream_handle(stream)))
    return sse, wsum


def gather_strided(X, y, w, begin, end, counts, *, stream=None, staging=None):
    """Strided row samples of ranges [begin[k], end[k]) with counts[k] rows
    each (staging: a PinnedRing on the stream for the range table)."""
    torch = _hip.torch_device()
    lib = _hip.load()
    dev, d = X.device, int(X.shape[1])
    counts = np.asarray(counts, dtype=np.int64)
    dst = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    total = int(dst[-1])
    Xo = torch.empty((total, d), dtype=torch.float64, device=dev)
    yo = torch.empty(total, dtype=torch.float64, device=dev)
    wo = None if w is None else torch.empty(total, dtype=torch.float64, device=dev)
    put = _hip.upload if staging is None else staging.upload
    table = put(np.concatenate([np.asarray(begin, np.int64), np.asarray(end, np.int64), dst]),
                dev)
    K = len(counts)
    b, e, dd = table[:K], table[K:2 * K], table[2 * K:]
    stream = torch.cuda.current_stream() if stream is None else stream
    _hip.check(lib.art_gather_strided(
        _hip.ptr(X), _hip.ptr(y), 0 if w is None else _hip.ptr(w), d, len(counts), _hip.ptr(b),
        _hip.ptr(e), _hip.ptr(dd), total, _hip.ptr(Xo), _hip.ptr(yo),
        0 if wo is None else _hip.ptr(wo), _hip.stream_handle(stream)))
    return Xo, yo, wo, dst


def partition_rows(X, y, w, begin, end, node_w, node_z, out, out_base, *, stream=None,
                   staging=None):
    """Stable partition of ranges [begin[k], end[k]) by hyperplanes (node_w, node_z)
    into ``out`` = (Xo, yo, wo) at rows out_base[k] + ...; returns the right
    counts (host int64 array, synchronizes).  staging: a PinnedRing on the
    stream for the tables."""
    torch = _hip.torch_device()
    lib = _hip.load()
    dev, d = X.device, int(X.shape[1])
    begin = np.asarray(begin, dtype=np.int64)
    end = np.asarray(end, dtype=np.int64)
    K = begin.shape[0]
    if K == 0:
        return np.zeros(0, dtype=np.int64)
    rows = lib.art_partition_block_rows()
    nblk = (end - begin + rows - 1) // rows
    blk = np.concatenate([[0], np.cumsum(nblk)]).astype(np.int64)
    blocks = int(blk[-1])
    put = _hip.upload if staging is None else staging.upload
    table = put(np.concatenate([begin, end, blk, np.asarray(out_base, np.int64)]), dev)
    tb, te = table[:K], table[K:2 * K]
    tblk, tdst = table[2 * K:3 * K + 1], table[3 * K + 1:]
    planes = put(np.concatenate([np.asarray(node_w, np.float64).reshape(-1),
                                 np.asarray(node_z, np.float64).reshape(-1)]), dev)
    nw, nz = planes[:K * d], planes[K * d:]
    flags = torch.empty(int(X.shape[0]), dtype=torch.uint8, device=dev)
    blk_right = torch.empty(max(blocks, 1), dtype=torch.int32, device=dev)
    node_right = torch.empty(K, dtype=torch.int64, device=dev)
    Xo, yo, wo = out
    stream = torch.cuda.current_stream() if stream is None else stream
    _hip.check(lib.art_partition_rows(
        _hip.ptr(X), _hip.ptr(y), 0 if w is None else _hip.ptr(w), d, K, _hip.ptr(tb),
        _hip.ptr(te), _hip.ptr(tblk), blocks, _hip.ptr(tdst), _hip.ptr(nw), _hip.ptr(nz),
        _hip.ptr(flags), _hip.ptr(blk_right), _hip.ptr(node_right), _hip.ptr(Xo), _hip.ptr(yo),
        0 if wo is None else _hip.ptr(wo), _hip.stream_handle(stream)))
    return node_right.cpu().numpy()


# --------------------------------------------------------------------------- engine state

class _Fits:
    """Host results of one node-fit launch, row i for the launch's node i:
    coefficients (K, p), solve status and weighted MSE.  A node's fit is the
    pair (record, i), so a level costs three arrays, not a record per node."""

    __slots__ = ("coef", "status", "mse", "bad")

    def __init__(self, coef, status, sse, wsum, bad=None):
        self.coef, self.status, self.bad = coef, status, bad
        with np.errstate(divide="ignore", invalid="ignore"):
            self.mse = np.where(wsum > 0, sse / wsum, np.inf)


@dataclass(slots=True)
class _Node:
    heap: int
    depth: int
    n_rows: int
    split: Optional[tuple] = None      # (w, z, sample rows)
    leaf: Optional[tuple] = None       # (_Fits, row) of a leaf's model
    gram: Any = None                   # (level systems tensor, row) of a leaf
    bad: int = 0
    leaf_candidate: Optional[tuple] = None   # the node's own fit while its split is open
    gram_candidate: Any = None
    index: int = -1                    # row of the node in its level's tables
    err: Optional[str] = None          # why the split search failed, for a leaf


@dataclass
class _FitJob:
    req: FitRequest
    future: Future
    timing: RequestTiming
    tree_id: str
    d: int
    degree: int
    n_rows: int
    est_bytes: float
    X: Any = None
    y: Any = None
    w: Any = None
    nodes: dict = field(default_factory=dict)
    open: int = 1                       # nodes not yet decided


@dataclass
class _Refit:
    req: RefitRequest
    future: Future
    timing: RequestTiming
    X: Any = None
    y: Any = None
    w: Any = None


class _Resident:
    """A fitted tree kept on the device: flattened split and leaf tables
    (breadth-first: heap order), the leaves' FP64 systems for refits, and the
    host records from which the ``RegressionTree`` is built on demand."""

    def __init__(self, tree_id, flat, grams, bad, d, degree, p, ridge, splits, leaves, n_rows):
        self.tree_id, self.flat, self.grams, self.bad = tree_id, flat, grams, bad
        self.d, self.degree, self.p, self.ridge = d, degree, p, ridge
        self.splits = splits      # [(heap, depth, w, z, sample rows)] in heap order
        # Leaves in heap order (= flat leaf order): heaps and depths (lists),
        # status, MSE and row counts (arrays), {leaf: split-search error}.
        self.leaf_heap, self.leaf_depth, self.status, self.mse, self.leaf_rows, \
            self.errors = leaves
        self._meta = None         # leaf metadata dicts, made with the tree
        self.coef = None          # host (L, p), set by the owner
        self.n_rows = n_rows
        self.bytes = 0.0
        self.last_used = time.perf_counter()
        self._tree = None
        self._lock = threading.Lock()

    @property
    def tree(self):
        with self._lock:
            if self._tree is None:
                self._tree = self._build()
            return self._tree

    def _build(self):
        from .models import PolynomialRidgeModel
        from .tree import LeafNode, RegressionTree, SplitNode

        split_of = {h: (depth, w, z, {"sample_rows": n}) for h, depth, w, z, n in self.splits}
        metas = []
        for j in range(len(self.leaf_heap)):
            meta = {"status": int(self.status[j]), "mse": float(self.mse[j]),
                    "n_rows": int(self.leaf_rows[j])}
            if j in self.errors:
                meta["split_error"] = self.errors[j]
            metas.append(meta)
        self._meta = metas
        leaf_of = {h: (j, depth, metas[j])
                   for j, (h, depth) in enumerate(zip(self.leaf_heap, self.leaf_depth))}
        self._models = [None] * len(self.leaf_heap)

        def build(h):
            if h in split_of:
                depth, w, z, meta = split_of[h]
                return SplitNode(w, z, build(2 * h), build(2 * h + 1), depth=depth,
                                 node_id=_heap_path(h), metadata=dict(meta))
            j, depth, meta = leaf_of[h]
            m = PolynomialRidgeModel(self.degree, ridge=self.ridge)
            m.coef_ = self.coef[j].copy()
            self._models[j] = m
            # The leaf shares its metadata dict with the resident record:
            # refit statuses show through.
            return LeafNode(m, depth=depth, node_id=_heap_path(h), metadata=meta)

        return RegressionTree(build(1), metadata={"n_rows": self.n_rows, "d": self.d,
                                                  "degree": self.degree})

    def set_coefficients(self, coef, status):
        """Leaf coefficients after a refit (host copies; the device table is
        updated by the caller)."""
        with self._lock:
            self.coef = coef
            self.status = np.asarray(status, dtype=np.int64).copy()
            if self._meta is not None:
                for j, meta in enumerate(self._meta):
                    meta["status"] = int(self.status[j])
            if self._tree is not None:
                for j, m in enumerate(self._models):
                    m.coef_ = coef[j].copy()


class _Group:
    """Level buffer of every fit request with basis (d, degree)."""

    def __init__(self, d, degree):
        self.d, self.degree = d, degree
        self.X = self.y = self.w = None
        self.entries = []        # (job, node) in buffer order; node None = unused range
        self.sizes = []
        self.pending = None      # splits to apply: [(job, node, begin, end, w, z)]
        self.new_jobs = []
        self.offsets = None      # row offsets of entries in the level buffer
        self.stream = None       # the group's split-search stream
        self.staging = None      # pinned upload ring on that stream
        self.plan = None         # mixed precision: the level's leaf plan, kept for late leaves
        self.grams = None        # mixed precision: the level's FP64 systems (leaves only)
        self.bad = None


class ServingEngine:
    """Request loop over one GPU; ``submit`` returns a ``Future``."""

    def __init__(self, config: Optional[ServingConfig] = None, *, start=True):
        self.config = ServingConfig() if config is None else config
        if self.config.precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'.")
        torch = _hip.torch_device()
        self.torch, self.lib = torch, _hip.load()
        self.device = torch.device("cuda", torch.cuda.current_device())
        self._cv = threading.Condition()
        self._queue = collections.deque()
        self._waiting = collections.deque()   # submitted, not yet admitted
        self._groups = {}
        self._refits = []
        self._resident = collections.OrderedDict()
        self._fitting = set()                 # tree ids of fits in flight
        self._inflight_bytes = 0.0
        self._resident_bytes = 0.0
        self._active = 0
        self._names = itertools.count()
        self._closing = False
        self._error = None
        self.steps = 0
        self.completed = 0
        self.phase_seconds = collections.defaultdict(float)
        self._retired_plans = []       # leaf plans to close after the level's searches
        self._diag_cache = {}
        self._thread = None
        self.stream = None
        self._pool = ThreadPoolExecutor(max_workers=8, thread_name_prefix="art-search")
        if start:
            self.start()

    # ----------------------------------------------------------------- public API

    def start(self):
        if self._thread is not None:
            return
        self.stream = self.torch.cuda.Stream(device=self.device)
        self.staging = _hip.PinnedRing(self.stream)
        self._thread = threading.Thread(target=self._run, name="art-serving", daemon=True)
        self._thread.start()

    def submit(self, request):
        """Queue a FitRequest or RefitRequest; returns a Future of its result."""
        if not isinstance(request, (FitRequest, RefitRequest)):
            raise TypeError("submit takes a FitRequest or a RefitRequest.")
        fut = Future()
        with self._cv:
            if self._closing:
                raise RuntimeError("The engine is closed.")
            if self._error is not None:
                raise RuntimeError("The engine stopped after an error.") from self._error
            self._queue.append((request, fut, RequestTiming(time.perf_counter())))
            self._cv.notify()
        return fut

    def fit(self, request):
        return self.submit(request).result()

    def tree(self, tree_id):
        with self._cv:
            res = self._resident[tree_id]
        return res.tree

    def evict(self, tree_id):
        with self._cv:
            r = self._resident.pop(tree_id)
            self._resident_bytes -= r.bytes

    def resident_ids(self):
        with self._cv:
            return list(self._resident)

    def close(self, wait=True):
        with self._cv:
            self._closing = True
            self._cv.notify()
        if wait and self._thread is not None:
            self._thread.join()
        self._thread = None
        if wait:
            self._pool.shutdown(wait=True)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    # ----------------------------------------------------------------- scheduler

    def _busy(self):
        return self._active > 0 or bool(self._refits) or bool(self._waiting)

    def _run(self):
        torch = self.torch
        torch.cuda.set_device(self.device)
        try:
            with torch.cuda.stream(self.stream):
                while True:
                    with self._cv:
                        while not self._queue and not self._closing and not self._busy():
                            self._cv.wait()
                        incoming = list(self._queue)
                        self._queue.clear()
                        if self._closing and not incoming and not self._busy():
                            break
                    self._waiting.extend(incoming)
                    self._admit()
                    self._step()
        except BaseException as exc:  # fail every open request loudly
            self._error = exc
            self._fail_all(exc)
        finally:
            for g in self._groups.values():
                try:
                    self._close_plan(g)
                except Exception:  # the failure is already reported to every request
                    pass
            try:
                self._close_retired_plans()
            except Exception:
                pass

    def _fail_all(self, exc):
        with self._cv:
            pending = list(self._queue) + list(self._waiting)
            self._queue.clear()
            self._waiting.clear()
        for _, fut, _ in pending:
            if not fut.done():
                fut.set_exception(exc)
        for g in self._groups.values():
            for job in {id(j): j for j, _ in g.entries if j is not None}.values():
                if not job.future.done():
                    job.future.set_exception(exc)
            for job in g.new_jobs:
                if not job.future.done():
                    job.future.set_exception(exc)
        for r in self._refits:
            if not r.future.done():
                r.future.set_exception(exc)

    # ----------------------------------------------------------------- admission

    def _basis(self, d, degree):
        if degree is None:
            degree = 2 if d <= self.config.quadratic_max_d else 1
        if degree not in (1, 2):
            raise ValueError("degree must be 1 or 2.")
        p = (1 + d + d * (d + 1) // 2) if degree == 2 else d + 1
        return degree, p

    def _fit_bytes(self, n, d, p, max_depth, min_leaf):
        gsize = self.lib.art_gram_size(p)
        leaves = min(2 ** max_depth, max(1, n // max(min_leaf, 1)))
        return 8.0 * (3 * n * (d + 2) + 2 * leaves * gsize + 2 * leaves * p)

    def _to_device(self, a, shape, name):
        torch = self.torch
        if a is None:
            return None
        if isinstance(a, torch.Tensor):
            if a.dtype != torch.float64 or tuple(a.shape) != shape or not a.is_cuda:
                raise ValueError(f"{name} must be a float64 device tensor of shape {shape}.")
            return a.contiguous()
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != shape:
            raise ValueError(f"{name} must have shape {shape}.")
        if not np.isfinite(a).all():
            raise ValueError(f"{name} must be finite.")
        return torch.from_numpy(a).to(self.device)

    def _validate_fit(self, req):
        X = req.X
        shape = tuple(X.shape)
        if len(shape) != 2 or shape[0] < 1 or shape[1] < 1 or shape[1] > 31:
            raise ValueError("X must have shape (N, d) with N >= 1 and 1 <= d <= 31.")
        if not isinstance(req.max_depth, (int, np.integer)) or req.max_depth < 0 \
                or req.max_depth > 60:
            raise ValueError("max_depth must be an integer in [0, 60].")
        if not isinstance(req.min_leaf, (int, np.integer)) or req.min_leaf < 1:
            raise ValueError("min_leaf must be a positive integer.")
        if not np.isfinite(req.temperature) or req.temperature <= 0:
            raise ValueError("temperature must be finite and positive.")
        if not np.isfinite(req.error_threshold) or req.error_threshold < 0:
            raise ValueError("error_threshold must be finite and nonnegative.")
        return shape

    def _admit(self):
        """Admit waiting requests in arrival order while slots and HBM allow
        (a later request may pass one that does not fit yet)."""
        cfg = self.config
        keep = collections.deque()
        refit_trees = {r.req.tree_id for r in self._refits}
        named = set()     # trees of fits still waiting ahead of a refit
        while self._waiting:
            req, fut, timing = self._waiting.popleft()
            try:
                if isinstance(req, RefitRequest):
                    if req.tree_id in refit_trees or req.tree_id in self._fitting \
                            or req.tree_id in named:
                        # One refit per tree per level, after the tree's fit.
                        keep.append((req, fut, timing))
                        refit_trees.add(req.tree_id)
                        continue
                    res = self._resident.get(req.tree_id)
                    if res is None:
                        raise KeyError(f"No resident tree {req.tree_id!r}.")
                    k = int(req.X.shape[0]) if hasattr(req.X, "shape") else -1
                    if len(tuple(req.X.shape)) != 2 or req.X.shape[1] != res.d or k < 1:
                        raise ValueError(f"Refit rows must have shape (k, {res.d}), k >= 1.")
                    refit = _Refit(req, fut, timing)
                    refit.X = self._to_device(req.X, (k, res.d), "X")
                    refit.y = self._to_device(req.y, (k,), "y")
                    refit.w = self._to_device(req.w, (k,), "w")
                    refit.timing.admitted = time.perf_counter()
                    self._refits.append(refit)
                    refit_trees.add(req.tree_id)
                    continue
                n, d = self._validate_fit(req)
                degree, p = self._basis(d, req.degree)
                est = self._fit_bytes(n, d, p, req.max_depth, req.min_leaf)
                if est > cfg.hbm_budget_bytes:
                    raise AdmissionError(f"Request needs ~{est / 1e9:.1f} GB, over the "
                                         f"{cfg.hbm_budget_bytes / 1e9:.1f} GB budget.")
                if self._active >= cfg.max_active or not self._reserve(est, keep):
                    keep.append((req, fut, timing))
                    if req.tree_id is not None:
                        named.add(req.tree_id)
                    continue
                tree_id = req.tree_id if req.tree_id is not None else f"tree-{next(self._names)}"
                job = _FitJob(req, fut, timing, tree_id, d, degree, n, est)
                job.X = self._to_device(req.X, (n, d), "X")
                job.y = self._to_device(req.y, (n,), "y")
                job.w = self._to_device(req.w, (n,), "w")
                job.nodes[1] = _Node(1, 0, n)
                job.timing.admitted = time.perf_counter()
                self._inflight_bytes += est
                self._active += 1
                self._fitting.add(tree_id)
                self._groups.setdefault((d, degree), _Group(d, degree)).new_jobs.append(job)
            except Exception as exc:  # a bad request fails alone
                fut.set_exception(exc)
        self._waiting = keep

    def _reserve(self, est, keep=()):
        """True when est more bytes fit the budget, evicting least recently
        used resident trees that no admitted or waiting refit needs."""
        budget = self.config.hbm_budget_bytes
        if self._inflight_bytes + self._resident_bytes + est <= budget:
            return True
        busy = {r.req.tree_id for r in self._refits}
        busy |= {req.tree_id for req, _, _ in itertools.chain(self._waiting, keep)
                 if isinstance(req, RefitRequest)}
        with self._cv:
            for tid in list(self._resident):
                if self._inflight_bytes + self._resident_bytes + est <= budget:
                    break
                if tid in busy:
                    continue
                r = self._resident.pop(tid)
                self._resident_bytes -= r.bytes
        return self._inflight_bytes + self._resident_bytes + est <= budget

    # ----------------------------------------------------------------- one level

    def _clock(self, phase, t0):
        t = time.perf_counter()
        self.phase_seconds[phase] += t - t0
        return t

    def _step(self):
        """One level of every basis group; refits first (they finish in one level).

        Node fits run group by group on the engine stream.  A group's split
        search starts, on the group's own stream from a pool thread, as soon
        as its node fits are decided, so it runs on the device while the
        engine thread partitions and fits the next groups.  Leaf plans freed
        during the level (a free synchronizes the device) are closed after
        the searches, so that no close waits on a search.
        """
        self.steps += 1
        if self._refits:
            t = time.perf_counter()
            self._run_refits()
            self._clock("refit", t)
        searches = []
        for key in list(self._groups):
            g = self._groups[key]
            t = time.perf_counter()
            self._build_level(g)
            t = self._clock("partition", t)
            if g.entries:
                idx = self._level(g)
                if idx:
                    searches.append((g, idx, self._pool.submit(self._search, g, idx)))
            if not g.entries and not g.new_jobs and g.pending is None:
                del self._groups[key]
        if searches:
            t = time.perf_counter()
            found = [(g, idx, fut.result()) for g, idx, fut in searches]
            t = self._clock("split", t)
            for g, idx, f in found:
                self._apply_splits(g, idx, *f)
            self._clock("decide", t)
        self._close_retired_plans()

    def _close_retired_plans(self):
        plans, self._retired_plans = self._retired_plans, []
        for plan in plans:
            plan.close()

    def _build_level(self, g):
        """Next level buffer: children of last level's splits, then new roots."""
        torch, d = self.torch, g.d
        pend = g.pending or []
        child_rows = sum(e - b for _, _, b, e, _, _ in pend)
        root_rows = sum(job.n_rows for job in g.new_jobs)
        total = child_rows + root_rows
        if total == 0:
            self._close_plan(g)
            g.entries, g.sizes, g.pending = [], [], None
            g.X = g.y = g.w = None
            return
        f64 = dict(dtype=torch.float64, device=self.device)
        X = torch.empty((total, d), **f64)
        y = torch.empty(total, **f64)
        w = torch.empty(total, **f64)
        entries, sizes = [], []
        if pend:
            begin = [b for _, _, b, _, _, _ in pend]
            end = [e for _, _, _, e, _, _ in pend]
            base = np.concatenate([[0], np.cumsum(np.subtract(end, begin))[:-1]])
            right = partition_rows(g.X, g.y, g.w, begin, end, np.stack([p[4] for p in pend]),
                                   [p[5] for p in pend], (X, y, w), base, stream=self.stream,
                                   staging=self.staging)
            right = right.tolist()
            if g.plan is not None:
                # A partition that leaves one side empty makes its node a leaf
                # (below); its FP64 system comes from the previous level buffer.
                self._late_leaves(g, [node for (_, node, b, e, _, _), nr in zip(pend, right)
                                      if nr == 0 or nr == e - b])
            self._close_plan(g)
            for (job, node, b, e, _, _), nr in zip(pend, right):
                nl = e - b - nr
                if nl == 0 or nr == 0:
                    # Degenerate partition (sample and rows disagree at the
                    # boundary): the node becomes a leaf, its rows unused.
                    entries += ((None, None), (None, None))
                    sizes += (nl, nr)
                    node.split = None
                    node.leaf = node.leaf_candidate
                    node.gram = node.gram_candidate
                    job.open -= 1
                    self._maybe_finish(job)
                    continue
                h, depth, nodes = 2 * node.heap, node.depth + 1, job.nodes
                left = nodes[h] = _Node(h, depth, nl)
                rgt = nodes[h + 1] = _Node(h + 1, depth, nr)
                entries += ((job, left), (job, rgt))
                sizes += (nl, nr)
                node.leaf_candidate = node.gram_candidate = None
                job.open += 1   # two children open, the node itself decided
        self._close_plan(g)
        off = child_rows
        for job in g.new_jobs:
            n = job.n_rows
            X[off:off + n].copy_(job.X)
            y[off:off + n].copy_(job.y)
            if job.w is None:
                w[off:off + n].fill_(1.0)
            else:
                w[off:off + n].copy_(job.w)
            job.X = job.y = job.w = None   # the level buffer holds the rows now
            entries.append((job, job.nodes[1]))
            sizes.append(n)
            off += n
        g.new_jobs = []
        g.X, g.y, g.w = X, y, w
        g.entries, g.sizes, g.pending = entries, sizes, None

    def _mixed(self, d, p):
        return self.config.precision == "mixed" and p <= 64 and d <= 16

    def _diag_index(self, p):
        """Device offsets of the diagonal entries of a packed (p+1)-system."""
        t = self._diag_cache.get(p)
        if t is None:
            t = self.staging.upload([_diag_offset(c, p) for c in range(p + 1)], self.device,
                                    np.int64)
            self._diag_cache[p] = t
        return t

    def _level(self, g):
        from .leaf_ridge import LeafRidgeBatch, solve_grams

        torch, cfg, d = self.torch, self.config, g.d
        degree, p = self._basis(d, g.degree)
        t = time.perf_counter()
        offsets = np.concatenate([[0], np.cumsum(g.sizes)]).astype(np.int64)
        g.offsets = offsets
        if self._mixed(d, p):
            return self._level_mixed(g, degree, p, t)
        lb = LeafRidgeBatch(offsets, d, degree=degree, stream=self.stream)
        self._retired_plans.append(lb)  # closed after the level's searches
        grams, bad = lb.grams(g.X, g.y, g.w, cfg.weight_floor)
        coef, status, _ = solve_grams(grams, bad, d, cfg.ridge, degree=degree,
                                      stream=self.stream)
        sse, wsum = gram_sse(grams, coef, p, bias_col=0, stream=self.stream)
        self.stream.synchronize()
        rec = _Fits(coef.cpu().numpy(), status.cpu().numpy(), sse.cpu().numpy(),
                    wsum.cpu().numpy(), bad.cpu().numpy())
        mse_h, bad_h = rec.mse, rec.bad
        t = self._clock("node_fits", t)
        split_idx = []
        for i, (job, node) in enumerate(g.entries):
            if job is None:
                continue
            req = job.req
            stop = (node.depth >= req.max_depth or node.n_rows < 2 * req.min_leaf
                    or node.n_rows < 2 * (d + 2) or mse_h[i] <= req.error_threshold)
            if stop:
                node.leaf, node.gram, node.bad = (rec, i), (grams, i), int(bad_h[i])
                job.open -= 1
                self._maybe_finish(job)
            else:
                node.leaf_candidate, node.gram_candidate = (rec, i), (grams, i)
                node.bad = int(bad_h[i])
                node.index = i
                split_idx.append(i)
        self._clock("decide", t)
        return split_idx

    def _final_by_size(self, job, node, d):
        req = job.req
        return (node.depth >= req.max_depth or node.n_rows < 2 * req.min_leaf
                or node.n_rows < 2 * (d + 2))

    def _level_mixed(self, g, degree, p, t):
        """One level at mixed precision.  Nodes the depth and size rules leave
        open get split-bf16 statistics; a node whose mixed MSE clears its
        error threshold by the margin goes to the split search on them.
        Every other live node -- stopped by depth or size, or with a mixed MSE
        within the margin of its threshold -- gets its FP64 system and fit,
        and the stop test of the FP64 engine decides it.  So the decisions are
        the FP64 engine's whenever the mixed MSE errs by less than the margin.
        The level's plan and FP64 system tensor stay with the group until the
        next level, for nodes that become leaves after their split search."""
        from .leaf_ridge import LeafRidgeBatch, solve_grams

        torch, cfg, d = self.torch, self.config, g.d
        K = len(g.entries)
        live = np.array([job is not None for job, _ in g.entries], dtype=bool)
        final = np.array([job is not None and self._final_by_size(job, node, d)
                          for job, node in g.entries], dtype=bool)
        cand = live & ~final
        thr = np.array([job.req.error_threshold if job is not None else 0.0
                        for job, _ in g.entries])
        g.plan = LeafRidgeBatch(g.offsets, d, degree=degree, stream=self.stream)
        size = self.lib.art_gram_size(p)
        f64 = dict(dtype=torch.float64, device=self.device)
        g.grams = torch.zeros((K, size), **f64)
        g.bad = torch.zeros(K, dtype=torch.int32, device=self.device)
        clear = np.zeros(K, dtype=bool)
        # FP64 systems of the nodes the depth and size rules stop go first, on
        # the same stream, so that both passes complete under one synchronization.
        early = self._fp64_launch(g, degree, p, final) if final.any() else None
        if cand.any():
            mg = torch.zeros((K, size), **f64)
            mb = torch.zeros(K, dtype=torch.int32, device=self.device)
            mask = self.staging.upload(cand, self.device, np.int32)
            g.plan.grams(g.X, g.y, g.w, cfg.weight_floor, precision="mixed", mask=mask,
                         out=(mg, mb))
            coef, status, _ = solve_grams(mg, mb, d, cfg.ridge, degree=degree, stream=self.stream)
            sse, wsum = gram_sse(mg, coef, p, bias_col=0, stream=self.stream)
            diag = self._diag_index(p)
            root = mg.index_select(1, diag).clamp_min(0.0).sqrt()
            scale = (coef.abs() * root[:, :p]).sum(dim=1) + root[:, p]
            scale = scale * scale
            self.stream.synchronize()
            wsum_h = wsum.cpu().numpy()
            mrec = _Fits(coef.cpu().numpy(), status.cpu().numpy(), sse.cpu().numpy(), wsum_h)
            s2_h = scale.cpu().numpy()
            with np.errstate(divide="ignore", invalid="ignore"):
                margin = np.where(wsum_h > 0, _MIXED_SSE_MARGIN * s2_h / wsum_h, np.inf)
            mse = mrec.mse
            clear = cand & (mrec.status == 0) & np.isfinite(mse) & (mse > thr + margin)
        frec = self._fp64_collect(g, *early) if early is not None else None
        self._clock("node_fits", t)
        near = cand & ~clear
        nrec = self._fp64_fits(g, degree, p, near) if near.any() else None
        t = time.perf_counter()
        split_idx = []
        grams = g.grams
        for i in np.flatnonzero(live).tolist():
            job, node = g.entries[i]
            node.index = i
            if clear[i]:
                node.leaf_candidate = (mrec, i)
                node.gram_candidate = None   # formed in FP64 only if the node ends as a leaf
                split_idx.append(i)
                continue
            rec = frec if final[i] else nrec
            if final[i] or rec.mse[i] <= job.req.error_threshold:
                node.leaf, node.gram, node.bad = (rec, i), (grams, i), int(rec.bad[i])
                job.open -= 1
                self._maybe_finish(job)
            else:
                node.leaf_candidate, node.gram_candidate = (rec, i), (grams, i)
                node.bad = int(rec.bad[i])
                split_idx.append(i)
        self._clock("decide", t)
        return split_idx

    def _fp64_launch(self, g, degree, p, sel):
        """Enqueue the FP64 systems and fits of the level's nodes sel (bool
        (K,)), from the rows in the level buffer, into the group's FP64
        tensors; returns the device results (coefficients, status, sse, wsum)."""
        from .leaf_ridge import solve_grams

        torch, cfg, d = self.torch, self.config, g.d
        mask = self.staging.upload(sel, self.device, np.int32)
        g.plan.grams(g.X, g.y, g.w, cfg.weight_floor, mask=mask, out=(g.grams, g.bad))
        coef, status, _ = solve_grams(g.grams, g.bad, d, cfg.ridge, degree=degree,
                                      stream=self.stream)
        sse, wsum = gram_sse(g.grams, coef, p, bias_col=0, stream=self.stream)
        return coef, status, sse, wsum

    def _fp64_fits(self, g, degree, p, sel):
        """_fp64_launch, then its host record (see _fp64_collect)."""
        t = time.perf_counter()
        out = self._fp64_collect(g, *self._fp64_launch(g, degree, p, sel))
        self._clock("node_fits", t)
        return out

    def _fp64_collect(self, g, coef, status, sse, wsum):
        """The host record (_Fits, with the group's weight flags) of an FP64
        launch; rows of the launch's nodes hold their fits (synchronizes the
        engine stream)."""
        self.stream.synchronize()
        return _Fits(coef.cpu().numpy(), status.cpu().numpy(), sse.cpu().numpy(),
                     wsum.cpu().numpy(), g.bad.cpu().numpy())

    def _late_leaves(self, g, nodes):
        """Mixed precision: nodes split on mixed statistics that end as leaves
        (failed search, one-sided partition) take their FP64 system and fit
        from the rows still in their level buffer."""
        todo = [n for n in nodes if n.gram_candidate is None]
        if not todo:
            return
        sel = np.zeros(len(g.entries), dtype=bool)
        sel[[n.index for n in todo]] = True
        degree, p = self._basis(g.d, g.degree)
        rec = self._fp64_fits(g, degree, p, sel)
        for n in todo:
            i = n.index
            n.leaf_candidate, n.gram_candidate, n.bad = (rec, i), (g.grams, i), int(rec.bad[i])

    def _close_plan(self, g):
        if g.plan is not None:
            self._retired_plans.append(g.plan)
            g.plan = None

    def _search(self, g, idx):
        """Split search of nodes idx of group g on the group's own stream:
        strided row samples, keyed random starts, the optimizer launch."""
        from .batched_optimizer import RaggedOptimizerWorkspace
        from .models import AffineRidgeModel
        from .splitters import SoftObliqueSplitter

        torch, cfg, d = self.torch, self.config, g.d
        torch.cuda.set_device(self.device)
        if g.stream is None:
            g.stream = torch.cuda.Stream(device=self.device)
        g.stream.wait_stream(self.stream)
        offsets = g.offsets
        begin = offsets[idx]
        end = offsets[np.asarray(idx) + 1]
        counts = np.minimum(end - begin, cfg.split_rows)
        temps = np.array([g.entries[i][0].req.temperature for i in idx])
        splitter = SoftObliqueSplitter(AffineRidgeModel(cfg.split_ridge), float(temps[0]),
                                       max_iters=cfg.max_iters, min_side_points=d + 2,
                                       min_side_fraction=cfg.min_side_fraction,
                                       n_restarts=cfg.n_restarts, random_state=0)
        R, count = cfg.n_restarts, len(idx)
        keys = [start_key(g.entries[i][0].req.seed, g.entries[i][1].heap, r)
                for r in range(R) for i in idx]
        if g.staging is None:
            g.staging = _hip.PinnedRing(g.stream, 1 << 20)
        with torch.cuda.stream(g.stream):
            Xs, ys, _, dst = gather_strided(g.X, g.y, None, begin, end, counts, stream=g.stream,
                                            staging=g.staging)
            with RaggedOptimizerWorkspace(splitter, Xs, ys, dst, temperature=temps) as ws:
                slots = [k for _ in range(R) for k in range(count)]
                theta0 = ws._device_starts(slots, keys=keys).reshape(R, count, d + 1)
                theta, ok, reason = ws.hyperplanes(theta0)
        return begin, end, counts, theta, ok, reason

    def _apply_splits(self, g, idx, begin, end, counts, theta, ok, reason):
        from .batched_optimizer import STOP_REASONS

        pending = []
        ok = np.asarray(ok, dtype=bool).tolist()
        if g.plan is not None:
            self._late_leaves(g, [g.entries[i][1] for k, i in enumerate(idx) if not ok[k]])
        # Each split's weights are a row of one copy of the batch's planes.
        W = np.array(theta[:, :-1], dtype=np.float64)
        Z, C = theta[:, -1].tolist(), np.asarray(counts).tolist()
        B, E = np.asarray(begin).tolist(), np.asarray(end).tolist()
        entries = g.entries
        for k, i in enumerate(idx):
            job, node = entries[i]
            if not ok[k]:
                node.leaf, node.gram = node.leaf_candidate, node.gram_candidate
                node.err = STOP_REASONS[int(reason[k])]
                node.leaf_candidate = node.gram_candidate = None
                job.open -= 1
                self._maybe_finish(job)
                continue
            wk, zk = W[k], Z[k]
            node.split = (wk, zk, C[k])
            pending.append((job, node, B[k], E[k], wk, zk))
        g.pending = pending or None

    # ----------------------------------------------------------------- completion

    def _maybe_finish(self, job):
        if job.open > 0:
            return
        t0 = time.perf_counter()
        self._finish(job)
        self.phase_seconds["assemble"] += time.perf_counter() - t0

    def _finish(self, job):
        """Flatten a finished fit straight from its node records: splits and
        leaves in heap order (the breadth-first order of ``FlatTree``), leaf
        systems gathered from the level tensors that hold them."""
        from .tree import FlatTree

        torch, d, degree = self.torch, job.d, job.degree
        _, p = self._basis(d, degree)
        nodes = job.nodes
        split_h, leaf_h = [], []
        for h in sorted(nodes):
            (split_h if nodes[h].split is not None else leaf_h).append(h)
        # Child link of every node: its split row, or -1 - its leaf index.
        link = {h: i for i, h in enumerate(split_h)}
        link.update((h, -1 - j) for j, h in enumerate(leaf_h))

        S = max(len(split_h), 1)
        w = np.zeros((S, d))
        z = np.zeros(S)
        child = np.zeros((S, 2), dtype=np.int32)
        splits = []
        for i, h in enumerate(split_h):
            node = nodes[h]
            sw, sz, n = node.split
            w[i], z[i] = sw, sz
            child[i, 0], child[i, 1] = link[2 * h], link[2 * h + 1]
            splits.append((h, node.depth, sw, sz, n))
        # Leaf fits and systems are rows of their launches' records and level
        # tensors: gather them launch by launch.
        L = len(leaf_h)
        depth, rows, bads, errors = [0] * L, [0] * L, [0] * L, {}
        by_rec, by_gram = {}, {}
        for j, h in enumerate(leaf_h):
            node = nodes[h]
            rec, i = node.leaf
            e = by_rec.get(id(rec))
            if e is None:
                e = by_rec[id(rec)] = (rec, [], [])
            e[1].append(j)
            e[2].append(i)
            t, i = node.gram
            e = by_gram.get(id(t))
            if e is None:
                e = by_gram[id(t)] = (t, [], [])
            e[1].append(j)
            e[2].append(i)
            depth[j], rows[j], bads[j] = node.depth, node.n_rows, node.bad
            if node.err is not None:
                errors[j] = node.err
        coef = np.empty((L, p))
        status = np.empty(L, dtype=np.int64)
        mse = np.empty(L)
        for rec, js, rs in by_rec.values():
            coef[js], status[js], mse[js] = rec.coef[rs], rec.status[rs], rec.mse[rs]
        flat = FlatTree.from_tables(w, z, child, link[1], coef, d, (degree, True, False),
                                    upload=self.staging.upload)
        lists = list(by_gram.values())
        grams = torch.empty((L, lists[0][0].shape[1]), dtype=torch.float64, device=self.device)
        idx = self.staging.upload(np.array([j for lst in lists for j in lst[1]]
                                   + [i for lst in lists for i in lst[2]] + bads,
                                   dtype=np.int64), self.device)
        at = 0
        for t, js, _ in lists:
            n = len(js)
            grams.index_copy_(0, idx[at:at + n], t.index_select(0, idx[L + at:L + at + n]))
            at += n
        bad = idx[2 * L:].to(torch.int32)
        leaves = (leaf_h, depth, status.copy(), mse, np.asarray(rows, dtype=np.int64), errors)
        res = _Resident(job.tree_id, flat, grams, bad, d, degree, p, self.config.ridge, splits,
                        leaves, job.n_rows)
        res.coef = coef
        res.bytes = 8.0 * (grams.numel() + flat.coef.numel() + flat.w.numel() * 2 + len(status))
        with self._cv:
            old = self._resident.pop(job.tree_id, None)
            if old is not None:
                self._resident_bytes -= old.bytes
            self._resident[job.tree_id] = res
            self._resident_bytes += res.bytes
            self._inflight_bytes -= job.est_bytes
            self._active -= 1
            self._fitting.discard(job.tree_id)
        job.nodes.clear()
        job.timing.finished = time.perf_counter()
        self.completed += 1
        job.future.set_result(FitResult(job.tree_id, coef.copy(), status, job.n_rows, job.timing,
                                        res))

    # ----------------------------------------------------------------- refits

    def _run_refits(self):
        """Every admitted refit in one grams launch per basis and one solve per
        basis: route, sort by leaf, add the new systems, re-solve."""
        from .leaf_ridge import LeafRidgeBatch, solve_grams

        torch, cfg = self.torch, self.config
        by_basis = collections.defaultdict(list)
        for r in self._refits:
            res = self._resident.get(r.req.tree_id)
            if res is None:
                r.future.set_exception(KeyError(f"No resident tree {r.req.tree_id!r}."))
                continue
            by_basis[(res.d, res.degree)].append((r, res))
        self._refits = []
        for (d, degree), items in by_basis.items():
            Xs, ys, ws, offs = [], [], [], [0]
            for r, res in items:
                L = res.flat.n_leaves
                ids = res.flat.leaves(r.X).to(torch.int64)
                order = torch.sort(ids, stable=True).indices
                counts = torch.bincount(ids, minlength=L).cpu().numpy()
                Xs.append(r.X.index_select(0, order))
                ys.append(r.y.index_select(0, order))
                ws.append(torch.ones_like(r.y) if r.w is None else r.w.index_select(0, order))
                offs.extend((offs[-1] + np.cumsum(counts)).tolist())
            offsets = np.asarray(offs, dtype=np.int64)
            X, y, w = torch.cat(Xs), torch.cat(ys), torch.cat(ws)
            with LeafRidgeBatch(offsets, d, degree=degree, stream=self.stream) as lb:
                new, bad = lb.grams(X, y, w, cfg.weight_floor)
                start = 0
                for r, res in items:
                    L = res.flat.n_leaves
                    res.grams += new[start:start + L]
                    res.bad |= bad[start:start + L]
                    start += L
                allg = torch.cat([res.grams for _, res in items])
                allb = torch.cat([res.bad for _, res in items])
                coef, status, _ = solve_grams(allg, allb, d, cfg.ridge, degree=degree,
                                              stream=self.stream)
                self.stream.synchronize()
            coef_h, status_h = coef.cpu().numpy(), status.cpu().numpy()
            start = 0
            for r, res in items:
                L = res.flat.n_leaves
                c, s = coef_h[start:start + L], status_h[start:start + L]
                res.flat.coef.copy_(coef[start:start + L])
                res.set_coefficients(c.copy(), s)
                start += L
                res.last_used = time.perf_counter()
                with self._cv:
                    self._resident.move_to_end(res.tree_id)
                r.timing.finished = time.perf_counter()
                self.completed += 1
                r.future.set_result(RefitResult(res.tree_id, c.copy(), s.copy(),
                                                int(r.X.shape[0]), r.timing, res))
