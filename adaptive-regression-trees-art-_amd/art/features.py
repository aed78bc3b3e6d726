"""Leaf-feature construction on the GPU (hot-path piece 4).

Drop-in for ``FeaturePlan`` / ``build_features`` of art/_cuda_features.py:42-95.
Device arrays are torch tensors on the HIP device (the reference uses CuPy
arrays); X stays C-contiguous float64 (k,n,d) and Phi is C-contiguous float64
(k,n,p).  The basis is decoded inside the kernel from (d, degree, bias), so no
column table is uploaded.
"""

from __future__ import annotations

import numpy as np

from . import _hip

_INT32_MAX = np.iinfo(np.int32).max


def _columns(d, degree, include_bias, bias_last):
    """Coordinate pairs in coefficient order; -1 means no factor.

    Same order as art/_cuda_features.py:13-29: bias (first, or last for the
    affine order), linear terms, then x_i*x_j for i <= j with i outermost.
    """
    if not isinstance(degree, (int, np.integer)) or degree not in (1, 2):
        raise ValueError("GPU features support only degree 1 or 2.")
    if bias_last and degree != 1:
        raise ValueError("bias_last is reserved for affine features (degree 1).")
    pairs = [(i, -1) for i in range(d)]
    if degree == 2:
        pairs.extend((i, j) for i in range(d) for j in range(i, d))
    if include_bias:
        pairs.insert(len(pairs) if bias_last else 0, (-1, -1))
    return np.asarray(pairs, dtype=np.int32).T.copy()


def feature_count(d, degree=2, include_bias=True):
    return (d + d * (d + 1) // 2 if degree == 2 else d) + (1 if include_bias else 0)


def _is_device_f64_c(t, torch, device_index=None):
    return (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64
            and t.is_contiguous()
            and (device_index is None or t.device.index == device_index))


class FeaturePlan:
    """Fixed linear/quadratic basis; fills supplied Phi buffers on the device.

    Create/use on the same stream (the current torch stream), or establish
    ordering before switching.  Inputs and output must be C-contiguous
    float64 on the plan's device; no sample-sized scratch is allocated.
    """

    def __init__(self, d, *, degree=2, include_bias=True, bias_last=False):
        if not isinstance(d, (int, np.integer)) or not 0 < d <= _INT32_MAX:
            raise ValueError("d must be a positive CUDA integer.")
        columns = _columns(d, degree, include_bias, bias_last)
        self.d, self.p = int(d), columns.shape[1]
        if self.p > _INT32_MAX:
            raise ValueError("Feature dimensions exceed the CUDA integer limit.")
        self.degree, self.include_bias, self.bias_last = int(degree), bool(include_bias), bool(bias_last)
        self.a, self.b = columns  # host copies of the column table (metadata only)
        torch = _hip.torch_device()
        self.lib = _hip.load()
        self.torch = torch
        self.device_id = torch.cuda.current_device()

    def build(self, X, out=None):
        torch = self.torch
        if X.ndim != 3 or min(X.shape) < 1 or X.shape[2] != self.d:
            raise ValueError("X must have nonempty shape (k,n,d) matching the basis.")
        if (not _is_device_f64_c(X, torch, self.device_id)
                or torch.cuda.current_device() != self.device_id):
            raise ValueError("X must be C-contiguous float64 on the current CUDA device.")
        k, n, d = X.shape
        if max(k, n) > _INT32_MAX:
            raise ValueError("Feature dimensions exceed the CUDA integer limit.")
        if out is None:
            out = torch.empty((k, n, self.p), dtype=torch.float64, device=X.device)
        if (not _is_device_f64_c(out, torch, self.device_id)
                or tuple(out.shape) != (k, n, self.p)):
            raise ValueError("out must be C-contiguous float64 with shape (k,n,p).")
        stream = torch.cuda.current_stream()
        _hip.check(self.lib.art_features_build(
            _hip.ptr(X), k * n, d, self.degree, int(self.include_bias), int(self.bias_last),
            _hip.ptr(out), _hip.stream_handle(stream)))
        return out


def build_features(X, *, degree=2, include_bias=True, bias_last=False):
    """Allocate a design for one call; use FeaturePlan to reuse the output.

    Polynomial order matches sklearn degree 1/2; bias_last=True selects the
    affine model's order (art/_cuda_features.py:88-95).
    """
    if X.ndim != 3 or min(X.shape) < 1:
        raise ValueError("X must have nonempty shape (k,n,d).")
    return FeaturePlan(X.shape[2], degree=degree, include_bias=include_bias,
                       bias_last=bias_last).build(X)
