"""Coarse build-time accounting (art/timing.py:18-62).

Wall-clock seconds per category, as the reference reports them in
``tree.metadata["build_timing"]``; device work is attributed to the category
whose host call waited for it.
"""

from __future__ import annotations

from contextlib import contextmanager
from dataclasses import dataclass
from time import perf_counter
from typing import Literal

BuildTimingCategory = Literal["sampling", "splitter", "model_refit"]


@dataclass
class BuildTimingProfile:
    sampling_seconds: float = 0.0
    splitter_seconds: float = 0.0
    optimizer_model_refit_seconds: float = 0.0

    @contextmanager
    def measure(self, category: BuildTimingCategory):
        start = perf_counter()
        try:
            yield
        finally:
            spent = perf_counter() - start
            name = {"sampling": "sampling_seconds", "splitter": "splitter_seconds"}.get(
                category, "optimizer_model_refit_seconds")
            setattr(self, name, getattr(self, name) + spent)

    def summary(self, total_build_seconds):
        total = max(float(total_build_seconds), 0.0)
        other = max(total - self.sampling_seconds - self.splitter_seconds, 0.0)

        def pct(x):
            return 0.0 if total == 0.0 else 100.0 * x / total

        return {
            "total_build_seconds": total,
            "sampling_seconds": self.sampling_seconds,
            "sampling_percent": pct(self.sampling_seconds),
            "splitter_seconds": self.splitter_seconds,
            "splitter_percent": pct(self.splitter_seconds),
            "optimizer_model_refit_seconds": self.optimizer_model_refit_seconds,
            "optimizer_model_refit_percent": pct(self.optimizer_model_refit_seconds),
            "other_seconds": other,
            "other_percent": pct(other),
        }
