"""Soft-oblique splitter API (art/splitters.py) backed by the device optimizer.

``SoftObliqueSplitter`` keeps the reference's fields and ``split`` signature;
every optimizer literal runs the device-controlled HIP optimizer
(``BatchedOptimizerWorkspace`` with one node), which restates the reference
loop (art/splitters.py:252-690) on the GPU.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Literal, Protocol, get_args, runtime_checkable

import numpy as np

Optimizer = Literal["cpu", "hybrid_gram", "cupy_qr", "cupy_cholesky", "cupy_batched_cholesky"]
OPTIMIZERS = get_args(Optimizer)


class SplitNotFoundError(RuntimeError):
    """Raised when every attempt yields an invalid hard partition (art/splitters.py:36-53)."""

    def __init__(self, reason, message, diagnostics=None, restarts_on_failure=0,
                 failure_reasons=(), context=None):
        super().__init__(message)
        self.reason = reason
        self.diagnostics = diagnostics
        self.restarts_on_failure = int(restarts_on_failure)
        self.failure_reasons = tuple(failure_reasons)
        self.context = {} if context is None else dict(context)


@runtime_checkable
class Splitter(Protocol):
    def split(self, X, y, parent_model=None, parent_loss=None, prepared_design=None):
        ...


def project_unit_w(theta):
    """Project theta = [w, z] onto |w| = 1 (art/splitters.py:71-83)."""
    theta = np.asarray(theta, dtype=float).reshape(-1)
    if theta.size < 2:
        raise ValueError("theta must contain at least one w component and one z component.")
    w = theta[:-1]
    n = np.linalg.norm(w)
    if n <= 1e-12:
        raise ValueError("Cannot project theta with near-zero w.")
    out = theta.copy()
    out[:-1] = w / n
    return out


def project_unit_w_rows(thetas):
    """project_unit_w of every row of thetas (n, d+1) at once."""
    thetas = np.asarray(thetas, dtype=float)
    if thetas.ndim != 2 or thetas.shape[1] < 2:
        raise ValueError("theta must contain at least one w component and one z component.")
    w = thetas[:, :-1]
    n = np.sqrt(np.einsum("ij,ij->i", w, w))
    if not np.all(n > 1e-12):
        raise ValueError("Cannot project theta with near-zero w.")
    out = thetas.copy()
    out[:, :-1] = w / n[:, None]
    return out


def project_unit_w_gradient(theta, grad):
    """Tangent projection of the w-gradient (art/splitters.py:86-98)."""
    theta = project_unit_w(theta)
    grad = np.asarray(grad, dtype=float).reshape(-1)
    if grad.shape != theta.shape:
        raise ValueError("grad and theta must have the same shape.")
    w = theta[:-1]
    out = grad.copy()
    out[:-1] = grad[:-1] - float(grad[:-1] @ w) * w
    return out


@dataclass
class SplitResult:
    """A split with optional hard models (art/splitters.py:114-196)."""

    w: np.ndarray
    z: float
    left_model: Any
    right_model: Any
    loss: float | None
    parent_loss: float
    split_gain: float | None
    relative_split_gain: float | None
    n_left: int
    n_right: int
    converged: bool
    n_iters: int
    stop_reason: str
    metadata: dict = field(default_factory=dict)

    @property
    def has_hard_fit(self):
        return self.loss is not None

    def predict(self, X):
        if not self.has_hard_fit:
            raise RuntimeError("This partition has no hard models; top up and fit its children first.")
        X = np.asarray(X, dtype=float)
        if X.ndim == 1:
            X = X.reshape(1, -1)
        if X.ndim != 2 or X.shape[1] != self.w.shape[0]:
            raise ValueError(f"X must have shape (n, {self.w.shape[0]}), got {X.shape}.")
        right = (X @ self.w - self.z) >= 0.0
        out = np.empty(X.shape[0])
        if np.any(~right):
            out[~right] = self.left_model.predict(X[~right])
        if np.any(right):
            out[right] = self.right_model.predict(X[right])
        return out


def _partition_result(theta, n, n_right, parent_loss, converged, metadata):
    """Geometry and counts without hard models (art/splitters.py:199-207)."""
    return SplitResult(
        w=np.asarray(theta[:-1], dtype=float).copy(), z=float(theta[-1]), left_model=None,
        right_model=None, loss=None, parent_loss=float(parent_loss), split_gain=None,
        relative_split_gain=None, n_left=n - int(n_right), n_right=int(n_right),
        converged=converged, n_iters=int(metadata["n_iters"]),
        stop_reason=str(metadata["stop_reason"]), metadata={**metadata, "hard_refit": False})


@dataclass
class SoftObliqueSplitter:
    """Soft sigmoid oblique splitter optimized by projected gradient descent."""

    model_template: Any
    temperature: float
    max_iters: int = 100
    grad_atol: float = 1e-8
    grad_rtol: float = 1e-5
    min_side_points: int = 1
    min_side_fraction: float = 0.05
    min_side_fallback: bool = False
    n_restarts: int = 1
    alpha0: float = 1.0
    rho: float = 0.5
    armijo_c: float = 1e-4
    max_backtracks: int = 25
    adaptive_alpha: bool = True
    alpha_min: float = 1e-12
    alpha_max: float = 1e3
    alpha_grow: float = 10.0
    alpha_recovery: float = 10.0
    heavy_backtrack_threshold: int = 8
    max_line_search_failures: int = 3
    weight_floor: float = 1e-12
    refit_during_line_search: bool = False
    random_state: Any = None
    split_timing_profile: Any = field(default=None, repr=False, compare=False)
    optimizer: Optimizer = "cpu"
    backtrack_batch_size: int = 8
    qr_parallel: bool = True
    Xw_refresh: int = 32

    def split(self, X, y, parent_model=None, parent_loss=None, prepared_design=None, *,
              gram_workspace=None, qr_workspace=None, cholesky_workspace=None, hard_refit=True,
              theta0=None):
        """Best split over n_restarts starts for one node, run on the GPU."""
        from .batched_optimizer import BatchedOptimizerWorkspace

        X = np.asarray(X, dtype=float)
        y = np.asarray(y, dtype=float).reshape(-1)
        self._validate_inputs(X, y)
        if not isinstance(hard_refit, (bool, np.bool_)):
            raise ValueError("hard_refit must be a boolean.")
        if not hard_refit and self.n_restarts != 1:
            raise ValueError("Skipping hard refits requires n_restarts == 1.")
        if self.optimizer not in OPTIMIZERS:
            raise ValueError(f"optimizer must be one of {OPTIMIZERS}.")
        if theta0 is not None:
            theta0 = np.asarray(theta0, dtype=float)
            theta0 = theta0[None] if theta0.ndim == 1 else theta0[:, None]
        pl = None if parent_loss is None else np.array([float(parent_loss)])
        with BatchedOptimizerWorkspace(self, X[None], y[None]) as ws:
            result = ws.optimize(theta0, parent_loss=pl, hard_refit=hard_refit)[0]
        if isinstance(result, SplitNotFoundError):
            raise result
        return result

    def _validate_inputs(self, X, y):
        if X.ndim != 2:
            raise ValueError(f"X must have shape (n, d), got {X.shape}.")
        if X.shape[0] != y.shape[0]:
            raise ValueError(f"X and y length mismatch: {X.shape[0]} != {y.shape[0]}.")
        if self.temperature <= 0.0:
            raise ValueError("temperature must be positive.")
        if self.max_iters < 0:
            raise ValueError("max_iters must be nonnegative.")
        if self.grad_atol < 0.0:
            raise ValueError("grad_atol must be nonnegative.")
        if self.grad_rtol < 0.0:
            raise ValueError("grad_rtol must be nonnegative.")
        if self.min_side_points < 1:
            raise ValueError("min_side_points must be at least 1.")
        if not (0.0 <= self.min_side_fraction < 0.5):
            raise ValueError("min_side_fraction must satisfy 0 <= fraction < 0.5.")
        if self.max_line_search_failures < 1:
            raise ValueError("max_line_search_failures must be at least 1.")

    def _required_side_points(self, n_total):
        return max(int(self.min_side_points), int(np.ceil(self.min_side_fraction * n_total)))

    def _rng(self):
        if isinstance(self.random_state, np.random.Generator):
            return self.random_state
        return np.random.default_rng(self.random_state)
