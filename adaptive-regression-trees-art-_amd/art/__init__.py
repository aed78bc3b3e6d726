"""MI355X-native hot path of Adaptive Regression Trees (ART).

Drop-in replacements for the reference's GPU tree-building entry points, backed
by hand-written gfx950 HIP kernels in libart_hip.so (C ABI: include/art_hip.h).
Importing the package loads neither torch nor the HIP library; both are
loaded when a device object is first constructed.
"""

from .features import FeaturePlan, build_features
from .batched_ridge import BatchedCholeskyWorkspace
from .leaf_ridge import LeafRidgeBatch

__all__ = [
    "BatchedCholeskyWorkspace",
    "FeaturePlan",
    "LeafRidgeBatch",
    "build_features",
]
