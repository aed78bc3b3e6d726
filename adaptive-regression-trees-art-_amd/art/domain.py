"""Polytope regions and box domains (art/domain.py of the reference).

Same classes, fields, validation and split convention: a split by w.x - z
sends w.x - z < 0 to the left child (row w, bound z) and w.x - z >= 0 to the
right child (row -w, bound -z) (art/domain.py:11-119).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np


@dataclass(frozen=True)
class PolytopeRegion:
    """A convex polytope {x : A x <= b}."""

    A: np.ndarray
    b: np.ndarray
    depth: int = 0
    tag: Optional[str] = None

    def __post_init__(self):
        A = np.asarray(self.A, dtype=float)
        b = np.asarray(self.b, dtype=float).reshape(-1)
        if A.ndim != 2:
            raise ValueError(f"A must have shape (m, d), got {A.shape}.")
        if b.ndim != 1:
            raise ValueError(f"b must have shape (m,), got {b.shape}.")
        if A.shape[0] != b.shape[0]:
            raise ValueError(
                f"A and b must have compatible rows, got {A.shape[0]} and {b.shape[0]}.")
        object.__setattr__(self, "A", A)
        object.__setattr__(self, "b", b)

    @property
    def dimension(self) -> int:
        return int(self.A.shape[1])

    def contains(self, X, tol: float = 1e-10):
        X = np.asarray(X, dtype=float)
        if X.ndim == 1:
            X = X.reshape(1, -1)
        if X.ndim != 2 or X.shape[1] != self.dimension:
            raise ValueError(f"X must have shape (n, {self.dimension}), got {X.shape}.")
        return np.all(X @ self.A.T <= self.b[None, :] + tol, axis=1)


@dataclass(frozen=True)
class BoxDomain:
    """Axis-aligned box with bounds shaped (d, 2)."""

    bounds: np.ndarray

    def __post_init__(self):
        bounds = np.asarray(self.bounds, dtype=float)
        if bounds.ndim != 2 or bounds.shape[1] != 2:
            raise ValueError(f"bounds must have shape (d, 2), got {bounds.shape}.")
        if np.any(bounds[:, 0] >= bounds[:, 1]):
            raise ValueError("Each bound must satisfy low < high.")
        object.__setattr__(self, "bounds", bounds)

    @property
    def dimension(self) -> int:
        return int(self.bounds.shape[0])

    @classmethod
    def hypercube(cls, dimension: int, low: float, high: float) -> "BoxDomain":
        dimension = int(dimension)
        if dimension < 1:
            raise ValueError("dimension must be at least 1.")
        return cls(np.tile(np.array([[float(low), float(high)]]), (dimension, 1)))

    def as_region(self, depth: int = 0, tag: Optional[str] = "root") -> PolytopeRegion:
        eye = np.eye(self.dimension)
        return PolytopeRegion(A=np.vstack([eye, -eye]),
                              b=np.concatenate([self.bounds[:, 1], -self.bounds[:, 0]]),
                              depth=depth, tag=tag)


def split_region(region: PolytopeRegion, w, z: float, depth: Optional[int] = None,
                 left_tag: Optional[str] = None, right_tag: Optional[str] = None):
    """Children of `region` split by w.x - z: left (< 0) and right (>= 0)."""
    w = np.asarray(w, dtype=float).reshape(-1)
    if w.shape[0] != region.dimension:
        raise ValueError(f"w must have shape ({region.dimension},), got {w.shape}.")
    child_depth = region.depth + 1 if depth is None else int(depth)
    left = PolytopeRegion(np.vstack([region.A, w[None]]), np.append(region.b, float(z)),
                          depth=child_depth, tag=left_tag)
    right = PolytopeRegion(np.vstack([region.A, -w[None]]), np.append(region.b, -float(z)),
                           depth=child_depth, tag=right_tag)
    return left, right
