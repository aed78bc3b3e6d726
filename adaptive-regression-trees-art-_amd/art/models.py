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
from typing import Literal, Protocol, runtime_checkable

import numpy as np

from .features import _columns, feature_count

RidgeSolver = Literal["auto", "normal", "qr", "svd"]
RIDGE_SOLVERS = ("auto", "normal", "qr", "svd")
DEFAULT_AUTO_RCOND_THRESHOLD = 1e-10
MODEL_CONDITION_WARNING_THRESHOLD = 1e10


class RegressionModel(Protocol):
    """What the tree builder needs of a node model (art/models.py:58-68)."""

    def fit(self, X, y): ...

    def predict(self, X): ...

    def clone(self): ...


@runtime_checkable
class WeightedRegressionModel(RegressionModel, Protocol):
    """A node model with weighted fits (art/models.py:71-82)."""

    def fit_weighted(self, X, y, weights, weight_floor=1e-12): ...


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


def _check_solver(solver, auto_rcond_threshold=DEFAULT_AUTO_RCOND_THRESHOLD):
    """The reference's solver validation (art/models.py:460-470) and the GPU
    path's support: "normal" and "auto"; "qr" and "svd" (host LAPACK
    least squares in the reference) are not provided on the device."""
    if solver not in RIDGE_SOLVERS:
        raise ValueError(f"solver must be one of {RIDGE_SOLVERS}, got {solver!r}.")
    if solver in ("qr", "svd"):
        raise ValueError(f"solver={solver!r} is not available on the GPU ridge path, which "
                         "solves the normal equations; use solver='normal' or solver='auto' "
                         "(Cholesky with FP64 iterative refinement of ill-conditioned systems).")
    threshold = float(auto_rcond_threshold)
    if not np.isfinite(threshold) or not 0.0 <= threshold <= 1.0:
        raise ValueError("auto_rcond_threshold must be finite and in [0, 1].")


def _solve_results(solver, coef, lam, rcond, threshold):
    """RidgeSolveResult per fit.  "normal": Cholesky, no estimate.  "auto":
    the condition estimate sqrt(1 / min(rcond, 1)) (art/models.py:288-296)
    and, below the threshold, the FP64-refined solution that replaces the
    reference's QR fallback (solver_used "cholesky_refined",
    fallback_reason "rcond_below_threshold")."""
    out = []
    for j in range(coef.shape[0]):
        if solver == "normal":
            out.append(RidgeSolveResult(coef[j], solver, "cholesky", None, None, float(lam[j])))
            continue
        rc = float(rcond[j])
        ok = rc > 0.0
        cond = float(np.sqrt(1.0 / min(rc, 1.0))) if ok else float("inf")
        below = not ok or rc < threshold
        out.append(RidgeSolveResult(
            coef[j], solver, "cholesky_refined" if below else "cholesky",
            "sqrt_gram_1norm" if ok else "condition_estimate_failure", cond, float(lam[j]),
            fallback_reason="rcond_below_threshold" if below else None))
    return out


def device_fit(X, y, weights, *, degree, include_bias, bias_last, ridge, weight_floor,
               solver="normal", auto_rcond_threshold=DEFAULT_AUTO_RCOND_THRESHOLD):
    """One weighted ridge fit on the GPU; returns (coefficients,
    RidgeSolveResult).

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
    _check_solver(solver, auto_rcond_threshold)
    dev = torch.device("cuda", torch.cuda.current_device())
    with LeafRidgeBatch(np.array([0, X.shape[0]]), X.shape[1], degree=degree,
                        include_bias=include_bias, bias_last=bias_last) as batch:
        coef, status, lam, rcond = _fit_batch(
            batch, solver, auto_rcond_threshold, torch.from_numpy(X).to(dev),
            torch.from_numpy(y).to(dev), torch.from_numpy(w).to(dev), ridge, weight_floor)
        if status[0] != 0:
            raise np.linalg.LinAlgError(
                "Normal-equation Cholesky factorization failed; increase ridge regularization.")
        result = _solve_results(solver, coef, lam, rcond, float(auto_rcond_threshold))[0]
        return coef[0].copy(), result


def _fit_batch(batch, solver, threshold, X, y, w, ridge, weight_floor):
    """Host arrays (coef, status, lambda, rcond or None) of one batch fit.
    "auto": the FP64 Cholesky solve and its condition estimate; leaves below
    the threshold are refitted with FP64 data-space refinement (p <= 64)."""
    coef, status, lam = batch.fit(X, y, w, ridge, weight_floor)
    if solver != "auto":
        return coef.cpu().numpy(), status.cpu().numpy(), lam.cpu().numpy(), None
    if batch.p > 64 or batch.d > 16:
        raise ValueError("solver='auto' on the GPU needs at most 64 basis features and d <= 16 "
                         "(its condition estimate uses the refined fit's stored factors).")
    plain = coef.cpu().numpy()
    rc, st, lam_h = batch.fit(X, y, w, ridge, weight_floor, precision="fp64_refined")
    rcond = batch.rcond().cpu().numpy()
    refined = rc.cpu().numpy()
    st, lam_h = st.cpu().numpy(), lam_h.cpu().numpy()
    below = (rcond < float(threshold)) | (rcond <= 0.0)
    out = np.where(below[:, None], refined, plain)
    return out, st, lam_h, rcond


class _RidgeBase:
    def _fit(self, X, y, weights, weight_floor):
        coef, result = device_fit(X, y, weights, degree=self._degree, include_bias=self._bias,
                                  bias_last=self._bias_last, ridge=self.ridge,
                                  weight_floor=weight_floor, solver=self.solver,
                                  auto_rcond_threshold=self.auto_rcond_threshold)
        self.coef_ = coef
        self.solve_result_ = result
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


def model_effective_dimension(model, input_dimension):
    """Fitting dimension of a model for sample budgets (art/models.py:166-178)."""
    method = getattr(model, "effective_dimension", None)
    if method is None:
        raise TypeError(f"{type(model).__name__} does not define effective_dimension; "
                        "provide an explicit sample_count policy to the tree builder.")
    dimension = int(method(input_dimension))
    if dimension < 1:
        raise ValueError("Model effective_dimension must be at least 1.")
    return dimension


def ridge_solve_diagnostics(model):
    """Scalar diagnostics of a model's last ridge solve (art/models.py:181-199)."""
    result = model if isinstance(model, RidgeSolveResult) else getattr(model, "solve_result_",
                                                                       None)
    if not isinstance(result, RidgeSolveResult):
        return None
    return {"solver_requested": result.solver_requested, "solver_used": result.solver_used,
            "condition_estimator": result.condition_estimator,
            "cond_estimate": result.cond_estimate, "ridge_effective": result.ridge_effective,
            "rank": result.rank, "fallback_reason": result.fallback_reason}


def fit_models(template, samples, stream=None):
    """Fit one clone of a GPU ridge template per (X, y) pair in a single batched
    launch (uniform weights, as the builder's node fits); returns the fitted
    models and their predictions on their own rows.

    This is the cohort form of the builder's per-node ``_fit_node_model``
    (art/builder.py:812-829): every node of a cohort becomes one leaf of a
    ``LeafRidgeBatch`` (FP64 Gram and Cholesky, trace-scaled ridge).  A node
    whose factorization fails raises ``numpy.linalg.LinAlgError`` as the
    reference's normal-equation solve does.
    """
    import torch

    from .leaf_ridge import LeafRidgeBatch

    degree, bias, bias_last = model_basis(template)
    _check_solver(template.solver, template.auto_rcond_threshold)
    if not samples:
        return [], []
    d = int(np.asarray(samples[0][0]).shape[1])
    sizes = np.array([len(y) for _, y in samples], dtype=np.int64)
    offsets = np.concatenate([[0], np.cumsum(sizes)])
    X = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=float) for x, _ in samples]))
    y = np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=float).reshape(-1)
                                             for _, v in samples]))
    dev = torch.device("cuda", torch.cuda.current_device())
    with LeafRidgeBatch(offsets, d, degree=degree, include_bias=bias, bias_last=bias_last,
                        stream=stream) as batch:
        Xt = torch.from_numpy(X).to(dev)
        coef, status, lam, rcond = _fit_batch(
            batch, template.solver, template.auto_rcond_threshold, Xt,
            torch.from_numpy(y).to(dev), torch.ones(len(y), dtype=torch.float64, device=dev),
            template.ridge, 1e-12)
    results = _solve_results(template.solver, coef, lam, rcond,
                             float(template.auto_rcond_threshold))
    models, predictions = [], []
    for j, (x, _) in enumerate(samples):
        if status[j] != 0:
            raise np.linalg.LinAlgError(
                "Normal-equation Cholesky factorization failed; increase ridge regularization.")
        m = template.clone()
        m.coef_ = coef[j].copy()
        m.solve_result_ = results[j]
        models.append(m)
        predictions.append(m.predict(x))
    return models, predictions
