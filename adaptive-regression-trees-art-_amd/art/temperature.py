"""Sigmoid temperature scales of soft oblique splits (art/temperature.py).

A node's temperature is c times a distance scale of its (subsampled) points:
the median nearest-neighbour distance ("median_nn") or the median pairwise
distance times n^(-1/d) ("median_pairwise_scaled").  The scales run on the
GPU (``art_knn_median``: brute-force distances of up to 4096 points per set in
one workgroup, median by radix select), one launch for any number of nodes;
``estimate_temperatures`` serves a whole cohort at once.  The subsample is
drawn on the host from the caller's generator, in the reference's order.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Literal

import numpy as np

TemperatureMode = Literal["median_nn", "median_pairwise_scaled"]
NearestNeighborMethod = Literal["kdtree", "bruteforce"]
TemperatureStrategy = Literal["splitter", "fixed", "tune_root", "tune_node"]

DEFAULT_TEMPERATURE_GRID = (1e-4, 5e-4, 0.001, 0.005, 0.01, 0.05, 0.1, 0.5, 1.0)
MAX_SET_POINTS = 4096


@dataclass(frozen=True)
class TemperatureConfig:
    """Node-level temperature scaling and tuning (art/temperature.py:19-44)."""

    strategy: TemperatureStrategy = "fixed"
    scale_mode: TemperatureMode = "median_nn"
    c: float = 0.1
    c_values: tuple[float, ...] = DEFAULT_TEMPERATURE_GRID
    validation_fraction: float = 0.2
    max_points: int | None = 512
    nn_method: NearestNeighborMethod | None = None
    bruteforce_dimension_threshold: int = 20

    def __post_init__(self):
        if self.strategy not in ("splitter", "fixed", "tune_root", "tune_node"):
            raise ValueError("Unknown temperature strategy.")
        if self.c <= 0.0:
            raise ValueError("Temperature c must be positive.")
        if not self.c_values or any(v <= 0.0 for v in self.c_values):
            raise ValueError("Temperature c_values must be nonempty and positive.")
        if not (0.0 < self.validation_fraction < 0.5):
            raise ValueError("validation_fraction must satisfy 0 < fraction < 0.5.")
        if self.max_points is not None and self.max_points < 2:
            raise ValueError("max_points must be at least 2 or None.")
        if self.bruteforce_dimension_threshold < 1:
            raise ValueError("bruteforce_dimension_threshold must be at least 1.")


def subsample_points(X, max_points=512, random_state=None):
    """At most max_points rows drawn without replacement (art/temperature.py:47-61)."""
    X = np.asarray(X, dtype=float)
    if X.ndim != 2:
        raise ValueError(f"X must have shape (n, d), got {X.shape}.")
    if max_points is None or X.shape[0] <= max_points:
        return X
    if max_points < 1:
        raise ValueError("max_points must be positive or None.")
    rng = (random_state if isinstance(random_state, np.random.Generator)
           else np.random.default_rng(random_state))
    return X[rng.choice(X.shape[0], size=int(max_points), replace=False)]


def pairwise_distances(X):
    """Dense Euclidean distance matrix of a small host array."""
    X = np.asarray(X, dtype=float)
    if X.ndim != 2:
        raise ValueError(f"X must have shape (n, d), got {X.shape}.")
    return np.linalg.norm(X[:, None, :] - X[None, :, :], axis=2)


def _check_points(X):
    X = np.asarray(X, dtype=float)
    if X.ndim != 2:
        raise ValueError(f"X must have shape (n, d), got {X.shape}.")
    if X.shape[0] < 2:
        raise ValueError("At least two points are required.")
    if X.shape[0] > MAX_SET_POINTS:
        raise ValueError(f"At most {MAX_SET_POINTS} points per set; subsample first.")
    return X


def set_distance_medians(point_sets, mode):
    """Median distance scale of each point set in one GPU launch.

    mode 0: median nearest-neighbour distance; 1: median pairwise distance.
    """
    import torch

    from . import _hip

    sets = [_check_points(X) for X in point_sets]
    if not sets:
        return np.empty(0)
    d = sets[0].shape[1]
    if any(X.shape[1] != d for X in sets):
        raise ValueError("Every point set must have the same dimension.")
    offsets = np.concatenate([[0], np.cumsum([X.shape[0] for X in sets])]).astype(np.int64)
    lib = _hip.load()
    torch = _hip.torch_device()
    dev = torch.device("cuda", torch.cuda.current_device())
    Xd = torch.from_numpy(np.ascontiguousarray(np.concatenate(sets))).to(dev)
    off = torch.from_numpy(offsets).to(dev)
    out = torch.empty(len(sets), dtype=torch.float64, device=dev)
    _hip.check(lib.art_knn_median(_hip.ptr(Xd), _hip.ptr(off), offsets.ctypes.data, len(sets), d,
                                  int(mode), _hip.ptr(out),
                                  _hip.stream_handle(torch.cuda.current_stream())))
    return out.cpu().numpy()


def median_nearest_neighbor_distance(X, method: NearestNeighborMethod = "kdtree"):
    """Median over points of the distance to the nearest other point.

    Both methods give the exact value; the GPU kernel is brute force.
    """
    if method not in ("kdtree", "bruteforce"):
        raise ValueError("method must be 'kdtree' or 'bruteforce'.")
    return float(set_distance_medians([X], 0)[0])


def median_pairwise_distance(X):
    return float(set_distance_medians([X], 1)[0])


def _scale(base, X_used, mode, c):
    if mode == "median_pairwise_scaled":
        base = base * (X_used.shape[0] ** (-1.0 / X_used.shape[1]))
    t = float(c) * float(base)
    if not np.isfinite(t) or t <= 0.0:
        raise RuntimeError("Estimated temperature is not positive and finite.")
    return t


def estimate_temperature(X, mode: TemperatureMode, c=1.0, max_points=512, random_state=None,
                         nn_method: NearestNeighborMethod = "kdtree"):
    """c times the distance scale of a subsample (art/temperature.py:121-142)."""
    return estimate_temperatures([X], mode, c, max_points, [random_state], nn_method)[0]


def estimate_temperatures(point_sets, mode: TemperatureMode, c=1.0, max_points=512,
                          random_states=None, nn_method: NearestNeighborMethod = "kdtree"):
    """``estimate_temperature`` of several nodes with one GPU launch; node k
    subsamples with random_states[k]."""
    if random_states is None:
        random_states = [None] * len(point_sets)
    used = [subsample_points(X, max_points=max_points, random_state=r)
            for X, r in zip(point_sets, random_states)]
    if c <= 0.0:
        raise ValueError("c must be positive.")
    if mode not in ("median_nn", "median_pairwise_scaled"):
        raise ValueError("mode must be 'median_nn' or 'median_pairwise_scaled'.")
    if nn_method not in ("kdtree", "bruteforce"):
        raise ValueError("method must be 'kdtree' or 'bruteforce'.")
    bases = set_distance_medians(used, 0 if mode == "median_nn" else 1)
    return [_scale(b, X, mode, c) for b, X in zip(bases, used)]


def temperature_grid(X, mode: TemperatureMode, c_values, max_points=512, random_state=None,
                     nn_method: NearestNeighborMethod = "kdtree"):
    """Distinct temperatures c * scale over c_values (art/temperature.py:145-164).

    Each candidate subsamples with the same random_state argument, as the
    reference does.
    """
    temps = [estimate_temperature(X, mode=mode, c=float(c), max_points=max_points,
                                  random_state=random_state, nn_method=nn_method)
             for c in c_values]
    unique = []
    for t in temps:
        if not any(np.isclose(t, u) for u in unique):
            unique.append(t)
    return unique
