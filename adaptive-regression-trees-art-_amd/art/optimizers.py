"""Line-search result and the adaptive initial-step controller (art/optimizers.py).

The GPU split optimizer runs the same AdaptiveAlpha rules on the device; these
host classes keep the reference's API for callers and metadata.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class LineSearchResult:
    theta: np.ndarray
    value: float
    step_size: float
    success: bool
    n_backtracks: int


@dataclass
class AdaptiveAlpha:
    """Next initial backtracking step (art/optimizers.py:20-69).

    Failure: alpha = clip(step * rho).  Success with no backtrack:
    clip(step * grow); with at least heavy_backtrack_threshold backtracks:
    clip(step * recovery); otherwise clip(step / rho).
    """

    alpha: float = 1.0
    alpha_min: float = 1e-12
    alpha_max: float = 1e3
    grow: float = 10.0
    recovery: float = 10.0
    heavy_backtrack_threshold: int = 8

    def __post_init__(self):
        if self.alpha <= 0.0:
            raise ValueError("alpha must be positive.")
        if self.alpha_min <= 0.0:
            raise ValueError("alpha_min must be positive.")
        if self.alpha_max < self.alpha_min:
            raise ValueError("alpha_max must be at least alpha_min.")
        if self.grow < 1.0:
            raise ValueError("grow must be at least 1.")
        if self.recovery < 1.0:
            raise ValueError("recovery must be at least 1.")
        if self.heavy_backtrack_threshold < 0:
            raise ValueError("heavy_backtrack_threshold must be nonnegative.")
        self.alpha = self._clip(self.alpha)

    def update(self, result: LineSearchResult, rho: float) -> float:
        if not (0.0 < rho < 1.0):
            raise ValueError("rho must satisfy 0 < rho < 1.")
        if not result.success:
            self.alpha = self._clip(result.step_size * rho)
        elif result.n_backtracks == 0:
            self.alpha = self._clip(result.step_size * self.grow)
        elif result.n_backtracks >= self.heavy_backtrack_threshold:
            self.alpha = self._clip(result.step_size * self.recovery)
        else:
            self.alpha = self._clip(result.step_size / rho)
        return self.alpha

    def _clip(self, alpha):
        return min(max(float(alpha), self.alpha_min), self.alpha_max)
