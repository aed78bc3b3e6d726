"""Reference import path art._cuda_features (art/_cuda_features.py), served by HIP."""

from .features import FeaturePlan, _columns, build_features

__all__ = ["FeaturePlan", "build_features", "_columns"]
