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
from .domain import BoxDomain, PolytopeRegion, split_region
from .sampling import HitAndRunSampler
from .temperature import TemperatureConfig, estimate_temperature
from .tree import LeafNode, RegressionTree, SplitNode
from .builder import RegressionTreeBuilder, TreeBuildResult

__all__ = [
    "AdaptiveAlpha",
    "AffineRidgeModel",
    "BatchedCholeskyWorkspace",
    "BatchedOptimizerWorkspace",
    "BoxDomain",
    "FeaturePlan",
    "HitAndRunSampler",
    "LeafNode",
    "LeafRidgeBatch",
    "LineSearchResult",
    "PolynomialRidgeModel",
    "PolytopeRegion",
    "RegressionTree",
    "RegressionTreeBuilder",
    "RidgeSolveResult",
    "SoftObliqueSplitter",
    "SplitNotFoundError",
    "SplitNode",
    "SplitResult",
    "TemperatureConfig",
    "TreeBuildResult",
    "build_features",
    "estimate_temperature",
    "optimize_batch",
    "project_unit_w",
    "project_unit_w_gradient",
    "split_region",
]
