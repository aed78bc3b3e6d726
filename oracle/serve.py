"""FP64 oracle for the serving loop's data-driven tree fit and rank-k refit.

Test infrastructure only (see ``oracle/__init__.py``).

The reference has no request loop and no data-driven fit (its builder queries
an oracle function, art/builder.py:502-551); this module restates the serving
engine's own rule from the reference's pieces, one node at a time:

* node model: weighted ridge on the leaf basis with the trace-scaled ridge
  (art/models.py:218-343 via ``ridge.solve_weighted_ridge``) and its weighted
  SSE on the node's rows;
* stop rule: depth >= max_depth, n < 2 min_leaf, n < 2 (d + 2) or weighted
  MSE <= error_threshold;
* split: the soft-oblique optimizer (``split.optimize``, art/splitters.py:365-601)
  on the strided row sample i -> floor(i n / m), m = min(n, split_rows), from
  the keyed device start (``split.device_start``), then the hard refit's side
  check (``split.hard_result``, art/splitters.py:603-690);
* routing: right when w.x - z >= 0 (art/tree.py:45-50), children keep the
  parent's row order;
* refit: a leaf's coefficients are the ridge fit on all of its rows, old and
  new (what the rank-k system update must reproduce).

Pinned by (tests/test_oracle_serve.py): depth 0 reduces to one
``solve_weighted_ridge`` call; a noiseless piecewise-affine target on two
half-planes is split once, near the true boundary, with one child fitted
exactly; the stable partition keeps
every row exactly once in input order; brute force over tiny inputs for the
strided sample.  The optimizer, ridge and start pieces carry the pins of
their own modules.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import ridge as R
from . import split as O

_MASK64 = (1 << 64) - 1


def _splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & _MASK64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & _MASK64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & _MASK64
    return x ^ (x >> 31)


def start_key(seed, heap, restart=0):
    """SplitMix64 (Steele, Lea and Flood 2014) of the node's heap index and
    restart, mixed with the request seed: the key of the node's random start."""
    return _splitmix64((int(seed) & _MASK64) ^ _splitmix64((int(heap) << 8) | int(restart)))


@dataclass
class FitSettings:
    max_depth: int = 10
    min_leaf: int = 64
    degree: int = 2
    error_threshold: float = 0.0
    temperature: float = 0.05
    seed: int = 0
    ridge: float = 1e-6
    weight_floor: float = 1e-12
    split_rows: int = 4096
    split_ridge: float = 1e-4
    max_iters: int = 30
    min_side_fraction: float = 0.05


def strided_sample(n, m):
    """Row positions of the split search's sample of m of n rows."""
    return (np.arange(m, dtype=np.int64) * n) // m


def stable_partition(X, w, z):
    """Boolean right mask (w.x - z >= 0) and the stable left-then-right order."""
    right = (X @ w - z) >= 0.0
    return right, np.concatenate([np.flatnonzero(~right), np.flatnonzero(right)])


def node_fit(X, y, w, s: FitSettings):
    """(coefficients, weighted SSE, weight sum) of one node's ridge model."""
    phi = R.features(X, s.degree, True, False)
    beta, _ = R.solve_weighted_ridge(phi, y, w, s.ridge, s.weight_floor)
    wf = np.maximum(w, s.weight_floor)
    r = y - phi @ beta
    return beta, float(np.sum(wf * r * r)), float(np.sum(wf))


def split_settings(s: FitSettings, d):
    return O.Settings(temperature=s.temperature, ridge=s.split_ridge, max_iters=s.max_iters,
                      min_side_points=d + 2, min_side_fraction=s.min_side_fraction,
                      degree=1, bias_last=True)


def fit_tree(X, y, w=None, s: FitSettings = FitSettings()):
    """Breadth-first data-driven fit; returns {heap: node} with node either
    {"split": (w, z), "rows": idx} or {"coef": beta, "rows": idx, "mse": m}."""
    X = np.asarray(X, dtype=float)
    y = np.asarray(y, dtype=float)
    w = np.ones(len(y)) if w is None else np.asarray(w, dtype=float)
    d = X.shape[1]
    so = split_settings(s, d)
    nodes = {}
    level = [(1, 0, np.arange(len(y)))]
    while level:
        nxt = []
        for heap, depth, idx in level:
            Xn, yn, wn = X[idx], y[idx], w[idx]
            beta, sse, wsum = node_fit(Xn, yn, wn, s)
            mse = sse / wsum if wsum > 0 else np.inf
            n = len(idx)
            leaf = {"coef": beta, "rows": idx, "mse": mse}
            if depth >= s.max_depth or n < 2 * s.min_leaf or n < 2 * (d + 2) \
                    or mse <= s.error_threshold:
                nodes[heap] = leaf
                continue
            m = min(n, s.split_rows)
            pos = strided_sample(n, m)
            Xs, ys = Xn[pos], yn[pos]
            theta0 = O.device_start(Xs, 0, 0, 0, key=start_key(s.seed, heap))
            theta = O.project_unit_w(O.optimize(Xs, ys, theta0, so)["theta"])
            try:
                O.hard_result(Xs, ys, theta, so)
            except (ValueError, np.linalg.LinAlgError):
                nodes[heap] = leaf
                continue
            wv, z = theta[:-1], float(theta[-1])
            right, _ = stable_partition(Xn, wv, z)
            if right.all() or not right.any():
                nodes[heap] = leaf
                continue
            nodes[heap] = {"split": (wv, z), "rows": idx}
            nxt.append((2 * heap, depth + 1, idx[~right]))
            nxt.append((2 * heap + 1, depth + 1, idx[right]))
        level = nxt
    return nodes


def route(nodes, X):
    """Heap index of the leaf each row of X reaches."""
    out = np.empty(len(X), dtype=np.int64)
    for r, x in enumerate(np.asarray(X, dtype=float)):
        h = 1
        while "split" in nodes[h]:
            wv, z = nodes[h]["split"]
            h = 2 * h + (1 if float(x @ wv) - z >= 0.0 else 0)
        out[r] = h
    return out


def refit_leaves(nodes, X_old, y_old, w_old, X_new, y_new, w_new, s: FitSettings):
    """Leaf coefficients after adding rows: the fit on each leaf's old and new rows."""
    X = np.vstack([X_old, X_new])
    y = np.concatenate([y_old, y_new])
    w = np.concatenate([np.ones(len(y_old)) if w_old is None else w_old,
                        np.ones(len(y_new)) if w_new is None else w_new])
    leaf_of = route(nodes, X)
    out = {}
    for h, node in nodes.items():
        if "coef" in node:
            idx = np.flatnonzero(leaf_of == h)
            out[h] = node_fit(X[idx], y[idx], w[idx], s)[0]
    return out
