"""MI355X-native hot path of Adaptive Regression Trees (ART).

Drop-in replacements for the reference's GPU tree-building entry points, backed
by hand-written gfx950 HIP kernels in libart_hip.so (C ABI: include/art_hip.h).
Importing the package loads neither torch nor the HIP library; both are
loaded when a device object is first constructed.
"""

from .features import FeaturePlan, build_features
from .batched_ridge import BatchedCholeskyWorkspace
from .leaf_ridge import LeafRidgeBatch
from .models import AffineRidgeModel, PolynomialRidgeModel, RidgeSolveResult
from .optimizers import AdaptiveAlpha, LineSearchResult
from .splitters import (SoftObliqueSplitter, SplitNotFoundError, SplitResult, project_unit_w,
                        project_unit_w_gradient)
from .batched_optimizer import BatchedOptimizerWorkspace, optimize_batch

__all__ = [
    "AdaptiveAlpha",
    "AffineRidgeModel",
    "BatchedCholeskyWorkspace",
    "BatchedOptimizerWorkspace",
    "FeaturePlan",
    "LeafRidgeBatch",
    "LineSearchResult",
    "PolynomialRidgeModel",
    "RidgeSolveResult",
    "SoftObliqueSplitter",
    "SplitNotFoundError",
    "SplitResult",
    "build_features",
    "optimize_batch",
    "project_unit_w",
    "project_unit_w_gradient",
]
