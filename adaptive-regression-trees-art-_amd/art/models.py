"""Leaf model containers with the reference's interface (art/models.py).

``AffineRidgeModel`` and ``PolynomialRidgeModel`` keep the reference's fields
and column orders (affine: ``[x, 1]``; polynomial: sklearn order with the
bias first).  Fits run on the GPU through the leaf-ridge kernels
(``LeafRidgeBatch`` with one leaf, FP64 normal equations with the
trace-scaled ridge of art/models.py:209-215); prediction is a host-side
matrix-vector product on the stored coefficients.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Literal

import numpy as np

from .features import _columns, feature_count

RidgeSolver = Literal["auto", "normal", "qr", "svd"]
RIDGE_SOLVERS = ("auto", "normal", "qr", "svd")
DEFAULT_AUTO_RCOND_THRESHOLD = 1e-10


@dataclass(frozen=True)
class RidgeSolveResult:
    """Coefficients and diagnostics of one ridge solve (art/models.py:24-35)."""

    coefficients: np.ndarray
    solver_requested: RidgeSolver
    solver_used: str
    condition_estimator: str | None
    cond_estimate: float | None
    ridge_effective: float
    rank: int | None = None
    fallback_reason: str | None = None


def polynomial_feature_count(d, degree=2, include_bias=True):
    return feature_count(int(d), int(degree), bool(include_bias))


def _design(X, degree, include_bias, bias_last):
    X = np.asarray(X, dtype=float)
    if X.ndim == 1:
        X = X.reshape(1, -1)
    a, b = _columns(X.shape[1], degree, include_bias, bias_last)
    ext = np.concatenate([X, np.ones((X.shape[0], 1))], axis=1)
    return ext[:, a] * ext[:, b]


def device_fit(X, y, weights, *, degree, include_bias, bias_last, ridge, weight_floor):
    """One weighted ridge fit on the GPU; returns (coefficients, ridge_effective).

    Raises ``np.linalg.LinAlgError`` when the Cholesky factorization fails, as
    the reference's normal-equation solver does (art/models.py:280-286).
    """
    import torch

    from .leaf_ridge import LeafRidgeBatch

    X = np.ascontiguousarray(X, dtype=np.float64)
    if X.ndim == 1:
        X = X.reshape(1, -1)
    y = np.ascontiguousarray(y, dtype=np.float64).reshape(-1)
    w = np.ascontiguousarray(weights, dtype=np.float64).reshape(-1)
    if X.shape[0] != y.shape[0] or y.shape[0] != w.shape[0]:
        raise ValueError("design, y, and weights must have the same number of rows.")
    if np.any(w < 0.0):
        raise ValueError("weights must be nonnegative.")
    if ridge < 0.0:
        raise ValueError("ridge must be nonnegative.")
    dev = torch.device("cuda", torch.cuda.current_device())
    with LeafRidgeBatch(np.array([0, X.shape[0]]), X.shape[1], degree=degree,
                        include_bias=include_bias, bias_last=bias_last) as batch:
        coef, status, lam = batch.fit(torch.from_numpy(X).to(dev), torch.from_numpy(y).to(dev),
                                      torch.from_numpy(w).to(dev), ridge, weight_floor)
        st = int(status.cpu()[0])
        if st != 0:
            raise np.linalg.LinAlgError(
                "Normal-equation Cholesky factorization failed; increase ridge regularization.")
        return coef.cpu().numpy()[0].copy(), float(lam.cpu()[0])


class _RidgeBase:
    def _fit(self, X, y, weights, weight_floor):
        coef, lam = device_fit(X, y, weights, degree=self._degree, include_bias=self._bias,
                               bias_last=self._bias_last, ridge=self.ridge,
                               weight_floor=weight_floor)
        self.coef_ = coef
        self.solve_result_ = RidgeSolveResult(coef, self.solver, "cholesky", None, None, lam)
        return self

    def fit(self, X, y):
        y = np.asarray(y, dtype=float).reshape(-1)
        return self._fit(X, y, np.ones_like(y), 1e-12)

    def fit_weighted(self, X, y, weights, weight_floor=1e-12):
        return self._fit(X, y, weights, weight_floor)

    def design(self, X):
        return _design(X, self._degree, self._bias, self._bias_last)

    def predict(self, X):
        if self.coef_ is None:
            raise ValueError("Model must be fit before calling predict.")
        return self.design(X) @ self.coef_


@dataclass
class AffineRidgeModel(_RidgeBase):
    """Affine model [x, 1] . coef (art/models.py:472-554)."""

    ridge: float = 1e-8
    solver: RidgeSolver = "normal"
    auto_rcond_threshold: float = DEFAULT_AUTO_RCOND_THRESHOLD
    coef_: np.ndarray | None = field(default=None, init=False, repr=False)
    solve_result_: RidgeSolveResult | None = field(default=None, init=False, repr=False)

    _degree = 1
    _bias = True
    _bias_last = True

    def effective_dimension(self, input_dimension):
        return int(input_dimension) + 1

    def clone(self):
        return AffineRidgeModel(self.ridge, self.solver, self.auto_rcond_threshold)


@dataclass
class PolynomialRidgeModel(_RidgeBase):
    """Polynomial model in sklearn column order (art/models.py:557-694)."""

    degree: int = 2
    ridge: float = 1e-8
    include_bias: bool = True
    solver: RidgeSolver = "normal"
    auto_rcond_threshold: float = DEFAULT_AUTO_RCOND_THRESHOLD
    coef_: np.ndarray | None = field(default=None, init=False, repr=False)
    solve_result_: RidgeSolveResult | None = field(default=None, init=False, repr=False)

    _bias_last = False

    @property
    def _degree(self):
        return int(self.degree)

    @property
    def _bias(self):
        return bool(self.include_bias)

    def effective_dimension(self, input_dimension):
        return polynomial_feature_count(input_dimension, self.degree, self.include_bias)

    def clone(self):
        return PolynomialRidgeModel(self.degree, self.ridge, self.include_bias, self.solver,
                                    self.auto_rcond_threshold)


def model_basis(model):
    """(degree, include_bias, bias_last) of a leaf model template."""
    if isinstance(model, AffineRidgeModel):
        return 1, True, True
    if isinstance(model, PolynomialRidgeModel):
        return int(model.degree), bool(model.include_bias), False
    raise TypeError("GPU fits support AffineRidgeModel and PolynomialRidgeModel.")
