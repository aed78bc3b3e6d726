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


def _splitmix64_u64(x):
    """_splitmix64 of a uint64 array (numpy's uint64 arithmetic wraps mod 2^64)."""
    u = np.uint64
    x = x + u(0x9E3779B97F4A7C15)
    x = (x ^ (x >> u(30))) * u(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> u(27))) * u(0x94D049BB133111EB)
    return x ^ (x >> u(31))


def start_keys(seeds, heaps, restarts):
    """start_key of every (seed, heap, restart) triple of equal-length
    integer arrays, as a uint64 array."""
    u = np.uint64
    seeds = np.asarray(seeds, dtype=np.uint64)   # nonnegative seeds below 2^64
    heaps = np.asarray(heaps, dtype=np.uint64)
    restarts = np.asarray(restarts, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return _splitmix64_u64(seeds ^ _splitmix64_u64((heaps << u(8)) | restarts))


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
        # Rows of nodes a launch did not form hold unread values: no warnings.
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            self.mse = np.where(wsum > 0, sse / wsum, np.inf)


@dataclass
class _FitJob:
    """An admitted fit.  Its decided nodes accumulate as blocks of arrays, one
    per decision batch: leaf blocks (heap, depth, rows, record, record rows,
    FP64 systems tensor, tensor rows, weight flags, split-search errors or
    None) and split blocks (heap, depth, w (k, d), z, sample rows)."""

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
    open: int = 1                       # nodes not yet decided
    slot: int = -1                      # the job's slot in its group
    leaves: list = field(default_factory=list)
    splits: list = field(default_factory=list)


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
        self.splits = splits      # (heap, depth, w (S, d), z, sample rows) arrays, heap order
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

        H, D, W, Z, C = self.splits
        split_of = {h: (depth, w, z, {"sample_rows": n})
                    for h, depth, w, z, n in zip(H.tolist(), D.tolist(), W, Z.tolist(),
                                                 C.tolist())}
        metas = []
        for j in range(len(self.leaf_heap)):
            meta = {"status": int(self.status[j]), "mse": float(self.mse[j]),
                    "n_rows": int(self.leaf_rows[j])}
            if j in self.errors:
                meta["split_error"] = self.errors[j]
            metas.append(meta)
        self._meta = metas
        leaf_of = {h: (j, depth, metas[j])
                   for j, (h, depth) in enumerate(zip(self.leaf_heap.tolist(),
                                                      self.leaf_depth.tolist()))}
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
    """Level buffer of every fit request with basis (d, degree).

    The level's entries are row ranges of the buffer, as arrays: the job
    slot (into ``jobs``; -1 for a range no node uses), the node's heap index
    (root 1, children 2h and 2h+1), depth and row count.  Slots keep their
    request's stop rules and search settings as arrays."""

    _PARAMS = ("max_depth", "min_leaf", "thr", "temp", "seed")

    def __init__(self, d, degree):
        self.d, self.degree = d, degree
        self.X = self.y = self.w = None
        empty = np.zeros(0, dtype=np.int64)
        self.jid, self.heap, self.depth, self.sizes = empty, empty, empty, empty
        self.jobs = []           # job per slot (None once finished)
        self._params = {k: [] for k in self._PARAMS}
        self._param_arrays = None
        self.pending = None      # splits to apply: dict of arrays (idx, begin, end, W, Z, counts)
        self.new_jobs = []
        self.offsets = None      # row offsets of entries in the level buffer
        self.stream = None       # the group's split-search stream
        self.staging = None      # pinned upload ring on that stream
        self.plan = None         # mixed precision: the level's leaf plan, kept for late leaves
        self.grams = None        # mixed precision: the level's FP64 systems (leaves only)
        self.bad = None
        # The searched entries' own fits, used if they end as leaves: record
        # c_recs[c_sel[i]] row i; FP64 system row i of c_gram_t when c_gram[i]
        # (mixed: formed late, from the rows still in the level buffer).
        self.c_recs, self.c_sel, self.c_gram, self.c_gram_t, self.c_bad = [], None, None, None, None

    def add_job(self, job):
        req = job.req
        job.slot = len(self.jobs)
        self.jobs.append(job)
        for k, v in zip(self._PARAMS, (req.max_depth, req.min_leaf, req.error_threshold,
                                       req.temperature, int(req.seed) & _MASK64)):
            self._params[k].append(v)
        self._param_arrays = None

    def params(self):
        """Per-slot arrays of the stop rules and search settings."""
        if self._param_arrays is None:
            P = {k: np.asarray(v) for k, v in self._params.items()}
            P["seed"] = np.asarray(self._params["seed"], dtype=np.uint64)
            self._param_arrays = P
        return self._param_arrays

    def live_jobs(self):
        return [j for j in self.jobs if j is not None]


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
        self.idle_periods = collections.deque(maxlen=100000)  # (start, end) with no work
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
                        if not self._queue and not self._closing and not self._busy():
                            t0 = time.perf_counter()
                            while not self._queue and not self._closing and not self._busy():
                                self._cv.wait()
                            self.idle_periods.append((t0, time.perf_counter()))
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
            for job in g.live_jobs():
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
                job.timing.admitted = time.perf_counter()
                self._inflight_bytes += est
                self._active += 1
                self._fitting.add(tree_id)
                g = self._groups.setdefault((d, degree), _Group(d, degree))
                g.add_job(job)
                g.new_jobs.append(job)
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
            if g.sizes.size:
                idx = self._level(g)
                if idx.size:
                    searches.append((g, idx, self._pool.submit(self._search, g, idx)))
            if not g.sizes.size and not g.new_jobs and g.pending is None:
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
        pend = g.pending
        child_rows = int((pend["end"] - pend["begin"]).sum()) if pend is not None else 0
        root_rows = sum(job.n_rows for job in g.new_jobs)
        total = child_rows + root_rows
        if total == 0:
            self._close_plan(g)
            empty = np.zeros(0, dtype=np.int64)
            g.jid, g.heap, g.depth, g.sizes = empty, empty, empty, empty
            g.pending = None
            g.X = g.y = g.w = None
            return
        f64 = dict(dtype=torch.float64, device=self.device)
        X = torch.empty((total, d), **f64)
        y = torch.empty(total, **f64)
        w = torch.empty(total, **f64)
        parts = []
        if pend is not None:
            idx, begin, end = pend["idx"], pend["begin"], pend["end"]
            n = end - begin
            base = np.concatenate([[0], np.cumsum(n)[:-1]])
            right = partition_rows(g.X, g.y, g.w, begin, end, pend["W"], pend["Z"],
                                   (X, y, w), base, stream=self.stream, staging=self.staging)
            left = n - right
            # A partition that leaves one side empty (sample and rows disagree
            # at the boundary) makes its node a leaf; its rows go unused, and
            # under mixed precision its FP64 system comes from the previous
            # level buffer.
            degen = (left == 0) | (right == 0)
            if degen.any():
                if g.plan is not None:
                    self._late_leaves(g, idx[degen])
                self._close_plan(g)
                self._candidate_leaves(g, idx[degen], None)
            self._close_plan(g)
            keep = ~degen
            self._add_splits(g, idx[keep], pend["W"][keep], pend["Z"][keep],
                             pend["counts"][keep])
            h = g.heap[idx]
            parts.append((np.repeat(np.where(degen, -1, g.jid[idx]), 2),
                          np.stack([2 * h, 2 * h + 1], axis=1).reshape(-1),
                          np.repeat(g.depth[idx] + 1, 2),
                          np.stack([left, right], axis=1).reshape(-1)))
        self._close_plan(g)
        if g.new_jobs:
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
                off += n
            k = len(g.new_jobs)
            parts.append((np.array([job.slot for job in g.new_jobs], dtype=np.int64),
                          np.ones(k, dtype=np.int64), np.zeros(k, dtype=np.int64),
                          np.array([job.n_rows for job in g.new_jobs], dtype=np.int64)))
        g.new_jobs = []
        g.X, g.y, g.w = X, y, w
        g.jid, g.heap, g.depth, g.sizes = (np.concatenate(c).astype(np.int64)
                                           for c in zip(*parts))
        g.pending = None

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
        t = self._clock("node_fits", t)
        live, final, thr = self._stop_rules(g)
        stop = live & (final | (rec.mse <= thr))
        K = g.sizes.size
        g.c_recs, g.c_sel = [rec], np.zeros(K, dtype=np.int64)
        g.c_gram, g.c_gram_t, g.c_bad = np.ones(K, dtype=bool), grams, rec.bad
        self._add_leaves(g, np.flatnonzero(stop), rec, grams, rec.bad)
        self._clock("decide", t)
        return np.flatnonzero(live & ~stop)

    def _stop_rules(self, g):
        """Per entry: live (a node uses it), final (its depth or size stops
        it) and its request's error threshold."""
        P, jid, d = g.params(), g.jid, g.d
        live = jid >= 0
        j = np.where(live, jid, 0)
        n = g.sizes
        final = live & ((g.depth >= P["max_depth"][j]) | (n < 2 * P["min_leaf"][j])
                        | (n < 2 * (d + 2)))
        return live, final, np.where(live, P["thr"][j], 0.0)

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
        K = g.sizes.size
        live, final, thr = self._stop_rules(g)
        cand = live & ~final
        g.plan = LeafRidgeBatch(g.offsets, d, degree=degree, stream=self.stream)
        size = self.lib.art_gram_size(p)
        f64 = dict(dtype=torch.float64, device=self.device)
        # Rows are formed for the nodes that need them (masked launches); the
        # others are never read.
        g.grams = torch.empty((K, size), **f64)
        g.bad = torch.zeros(K, dtype=torch.int32, device=self.device)
        clear = np.zeros(K, dtype=bool)
        mrec = None
        # FP64 systems of the nodes the depth and size rules stop go first, on
        # the same stream, so that both passes complete under one synchronization.
        early = self._fp64_launch(g, degree, p, final) if final.any() else None
        if cand.any():
            mg = torch.empty((K, size), **f64)
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
        # Searched on their mixed statistics: their fit is the mixed one, and
        # their FP64 system is formed only if they end as leaves.  Near the
        # threshold: the FP64 fit decides (leaf, or searched with its FP64
        # system in hand).  Stopped by depth or size: FP64 leaves.
        g.c_recs = [mrec, nrec]
        g.c_sel = np.where(clear, 0, 1)
        g.c_gram, g.c_gram_t = near.copy(), g.grams
        g.c_bad = np.zeros(K, dtype=np.int64) if nrec is None else np.where(near, nrec.bad, 0)
        if frec is not None:
            self._add_leaves(g, np.flatnonzero(final), frec, g.grams, frec.bad)
        split = clear.copy()
        if nrec is not None:
            stop = near & (nrec.mse <= thr)
            self._add_leaves(g, np.flatnonzero(stop), nrec, g.grams, nrec.bad)
            split |= near & ~stop
        self._clock("decide", t)
        return np.flatnonzero(split)

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

    def _late_leaves(self, g, idx):
        """Mixed precision: entries idx split on mixed statistics that end as
        leaves (failed search, one-sided partition) take their FP64 system
        and fit from the rows still in their level buffer."""
        todo = idx[~g.c_gram[idx]]
        if not todo.size:
            return
        sel = np.zeros(g.sizes.size, dtype=bool)
        sel[todo] = True
        degree, p = self._basis(g.d, g.degree)
        rec = self._fp64_fits(g, degree, p, sel)
        g.c_recs.append(rec)
        g.c_sel[todo] = len(g.c_recs) - 1
        g.c_gram[todo] = True
        g.c_bad = np.where(sel, rec.bad, g.c_bad)

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
        offsets, P, jid = g.offsets, g.params(), g.jid[idx]
        begin = offsets[idx]
        end = offsets[idx + 1]
        counts = np.minimum(end - begin, cfg.split_rows)
        temps = P["temp"][jid].astype(float)
        splitter = SoftObliqueSplitter(AffineRidgeModel(cfg.split_ridge), float(temps[0]),
                                       max_iters=cfg.max_iters, min_side_points=d + 2,
                                       min_side_fraction=cfg.min_side_fraction,
                                       n_restarts=cfg.n_restarts, random_state=0)
        R, count = cfg.n_restarts, len(idx)
        seeds, heaps = P["seed"][jid], g.heap[idx].astype(np.uint64)
        keys = start_keys(np.tile(seeds, R), np.tile(heaps, R),
                          np.repeat(np.arange(R, dtype=np.uint64), count))
        if g.staging is None:
            g.staging = _hip.PinnedRing(g.stream, 1 << 20)
        with torch.cuda.stream(g.stream):
            Xs, ys, _, dst = gather_strided(g.X, g.y, None, begin, end, counts, stream=g.stream,
                                            staging=g.staging)
            with RaggedOptimizerWorkspace(splitter, Xs, ys, dst, temperature=temps) as ws:
                # One workgroup per node: with several, the number of parts
                # (and so the order in which a pass's partial sums are added)
                # would depend on how many nodes share the launch, and a
                # request served in a mix must get bitwise the tree it gets
                # alone.
                ws.split_parts = 1
                slots = [k for _ in range(R) for k in range(count)]
                theta0 = ws._device_starts(slots, keys=keys).reshape(R, count, d + 1)
                theta, ok, reason = ws.hyperplanes(theta0)
        return begin, end, counts, theta, ok, reason

    def _apply_splits(self, g, idx, begin, end, counts, theta, ok, reason):
        from .batched_optimizer import STOP_REASONS

        ok = np.asarray(ok, dtype=bool)
        fail = idx[~ok]
        if fail.size:
            if g.plan is not None:
                self._late_leaves(g, fail)
            self._candidate_leaves(g, fail, [STOP_REASONS[int(r)] for r in reason[~ok]])
        if not ok.any():
            g.pending = None
            return
        # Each split's weights are a row of one copy of the batch's planes.
        theta = np.asarray(theta, dtype=np.float64)
        g.pending = dict(idx=idx[ok], begin=np.asarray(begin)[ok], end=np.asarray(end)[ok],
                         W=np.array(theta[ok, :-1]), Z=theta[ok, -1].copy(),
                         counts=np.asarray(counts)[ok])

    # ----------------------------------------------------------------- decisions

    def _job_runs(self, g, idx):
        """Entries idx grouped by job: (job, positions in idx) per job."""
        jid = g.jid[idx]
        order = np.argsort(jid, kind="stable")
        js = jid[order]
        cuts = np.flatnonzero(js[1:] != js[:-1]) + 1
        starts = np.concatenate([[0], cuts]).tolist()
        ends = np.concatenate([cuts, [js.size]]).tolist()
        return [(g.jobs[int(js[a])], order[a:b]) for a, b in zip(starts, ends)]

    def _add_leaves(self, g, idx, rec, gram_t, bad, errs=None):
        """Entries idx end as leaves with fits rec rows idx, FP64 systems
        gram_t rows idx and weight flags bad[idx] (errs: per entry, the split
        search's stop reason or None); finishes every job left with no open
        node."""
        if not idx.size:
            return
        heap, depth, rows, flags = g.heap[idx], g.depth[idx], g.sizes[idx], bad[idx]
        done = []
        for job, pos in self._job_runs(g, idx):
            e = None if errs is None else [errs[k] for k in pos.tolist()]
            job.leaves.append((heap[pos], depth[pos], rows[pos], rec, idx[pos], gram_t, idx[pos],
                               flags[pos], e))
            job.open -= pos.size
            if job.open == 0:
                done.append(job)
        for job in done:
            g.jobs[job.slot] = None
            t0 = time.perf_counter()
            self._finish(job)
            self.phase_seconds["assemble"] += time.perf_counter() - t0

    def _candidate_leaves(self, g, idx, errs):
        """Searched entries idx end as leaves on their own fits (the
        candidate state of their level)."""
        sel = g.c_sel[idx]
        for s in np.unique(sel).tolist():
            m = sel == s
            self._add_leaves(g, idx[m], g.c_recs[s], g.c_gram_t, g.c_bad,
                             None if errs is None else [e for e, k in zip(errs, m) if k])

    def _add_splits(self, g, idx, W, Z, counts):
        """Entries idx split (planes W, Z; search samples of counts rows):
        each adds two open children and closes itself."""
        if not idx.size:
            return
        heap, depth = g.heap[idx], g.depth[idx]
        for job, pos in self._job_runs(g, idx):
            job.splits.append((heap[pos], depth[pos], W[pos], Z[pos], counts[pos]))
            job.open += pos.size

    # ----------------------------------------------------------------- completion

    def _finish(self, job):
        """Flatten a finished fit from its leaf and split blocks: splits and
        leaves in heap order (the breadth-first order of ``FlatTree``), leaf
        fits gathered from their records and leaf systems from the level
        tensors that hold them."""
        from .tree import FlatTree

        torch, d, degree = self.torch, job.d, job.degree
        _, p = self._basis(d, degree)
        if job.splits:
            sh, sd, W, Z, C = (np.concatenate(c) for c in zip(*job.splits))
            o = np.argsort(sh, kind="stable")
            sh, sd, W, Z, C = sh[o], sd[o], W[o], Z[o], C[o]
        else:
            sh = sd = C = np.zeros(0, dtype=np.int64)
            W, Z = np.zeros((0, d)), np.zeros(0)
        blocks = job.leaves
        lh = np.concatenate([b[0] for b in blocks])
        order = np.argsort(lh, kind="stable")   # heap order of the leaves
        L = lh.size
        at = np.empty(L, dtype=np.int64)        # leaf j of the blocks -> its heap-order slot
        at[order] = np.arange(L)
        coef = np.empty((L, p))
        status = np.empty(L, dtype=np.int64)
        mse = np.empty(L)
        errors = {}
        k = 0
        for h, _, _, rec, rr, _, _, _, e in blocks:
            dst = at[k:k + h.size]
            coef[dst], status[dst], mse[dst] = rec.coef[rr], rec.status[rr], rec.mse[rr]
            if e is not None:
                errors.update((int(dst[i]), v) for i, v in enumerate(e) if v is not None)
            k += h.size
        lh = lh[order]
        ldepth = np.concatenate([b[1] for b in blocks])[order]
        lrows = np.concatenate([b[2] for b in blocks])[order]
        bads = np.concatenate([b[7] for b in blocks])[order]
        # Child link of every split: its children's split rows, or -1 - leaf index.
        S = sh.size
        ch = np.stack([2 * sh, 2 * sh + 1], axis=1)
        ks = np.minimum(np.searchsorted(sh, ch), max(S - 1, 0))
        is_split = (S > 0) & (sh[ks] == ch) if S else np.zeros(ch.shape, dtype=bool)
        kl = np.searchsorted(lh, ch)
        child = np.where(is_split, ks, -1 - kl).astype(np.int32)
        root = 0 if S and sh[0] == 1 else -1
        if not S:
            W, Z, child = np.zeros((1, d)), np.zeros(1), np.zeros((1, 2), dtype=np.int32)
        flat = FlatTree.from_tables(W, Z, child, root, coef, d, (degree, True, False),
                                    upload=self.staging.upload)
        # Leaf systems: gathered block by block into their heap-order rows.
        grams = torch.empty((L, blocks[0][5].shape[1]), dtype=torch.float64, device=self.device)
        gidx = self.staging.upload(np.concatenate([at, np.concatenate([b[6] for b in blocks]),
                                                   bads.astype(np.int64)]), self.device,
                                   np.int64)
        k = 0
        for b in blocks:
            n = b[0].size
            grams.index_copy_(0, gidx[k:k + n], b[5].index_select(0, gidx[L + k:L + k + n]))
            k += n
        bad = gidx[2 * L:].to(torch.int32)
        splits = (sh, sd, W[:S], Z[:S], C)
        leaves = (lh, ldepth, status.copy(), mse, lrows, errors)
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
        job.leaves, job.splits = [], []
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
