"""ctypes binding of libart_hip.so (declared in include/art_hip.h).

The library is built in-tree by ``art._build`` (hipcc, gfx950).  There is no
fallback: when the library is missing every device entry point raises.
Validation helpers here run before the library is loaded, mirroring the
reference's rule that argument errors surface before any GPU import
(tests/test_cuda_batched_cholesky.py:142-152).
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

LIB_DIR = Path(__file__).resolve().with_name("_lib")
LIB_PATH = LIB_DIR / "libart_hip.so"

ART_OK = 0
ERR_ARG, ERR_HIP = -1, -2
PRECISION_FP64 = 0
PRECISION_MIXED = 1
PRECISION_FP64_REFINED = 2

_lib = None


class HipLibraryError(RuntimeError):
    """Raised when libart_hip.so is missing or a call fails."""


_p = ctypes.c_void_p
_i = ctypes.c_int
_i64 = ctypes.c_int64
_d = ctypes.c_double

class Basis(ctypes.Structure):
    """art_basis of include/art_hip.h."""

    _fields_ = [("d", ctypes.c_int), ("p", ctypes.c_int), ("degree", ctypes.c_int),
                ("bias", ctypes.c_int), ("bias_last", ctypes.c_int)]


class SplitArgs(ctypes.Structure):
    """art_split_args of include/art_hip.h."""

    _fields_ = [
        ("X", ctypes.c_void_p), ("y", ctypes.c_void_p), ("temperature", ctypes.c_void_p),
        ("theta0", ctypes.c_void_p), ("node", ctypes.c_void_p),
        ("n", ctypes.c_int), ("d", ctypes.c_int), ("basis", Basis),
        *[(name, ctypes.c_double) for name in (
            "ridge", "floor", "grad_atol", "grad_rtol", "alpha0", "rho", "armijo_c",
            "alpha_min", "alpha_max", "grow", "recovery")],
        *[(name, ctypes.c_int) for name in (
            "max_iters", "max_backtracks", "heavy", "max_failures", "adaptive", "required",
            "fallback", "hard_refit", "H", "mode", "anneal_iters")],
        ("t0", ctypes.c_double), ("t1", ctypes.c_double), ("step", ctypes.c_double),
        ("theta_out", ctypes.c_void_p), ("beta_out", ctypes.c_void_p), ("hard", ctypes.c_void_p),
        ("info", ctypes.c_void_p), ("history", ctypes.c_void_p), ("scalars", ctypes.c_void_p),
        ("row_offsets", ctypes.c_void_p), ("min_side_points", ctypes.c_int),
        ("min_side_fraction", ctypes.c_double), ("child_precision", ctypes.c_int),
    ]


class HarArgs(ctypes.Structure):
    """art_har_args of include/art_hip.h."""

    _fields_ = [
        *[(name, ctypes.c_void_p) for name in ("A", "b", "row_offsets", "x0", "shape",
                                               "box_lo", "box_hi", "directions", "uniforms",
                                               "samples")],
        ("seed", ctypes.c_uint64),
        *[(name, ctypes.c_int) for name in ("P", "d", "chains", "n", "first", "stride",
                                            "steps")],
        ("seeds", ctypes.c_void_p),
        ("shape_stride", ctypes.c_int),
        ("row_bound", ctypes.c_int),
    ]


_PROTOTYPES = {
    "art_last_error": (ctypes.c_char_p, []),
    "art_version": (_i, []),
    "art_feature_count": (_i, [_i, _i, _i, _i]),
    "art_features_build": (_i, [_p, _i64, _i, _i, _i, _i, _p, _p]),
    "art_check_finite": (_i, [_p, _p, _i64, _i64, _i, _p, _p]),
    "art_leaf_plan_create": (_i, [_p, _i, _i, _i, _i, _i, _i, ctypes.POINTER(_p)]),
    "art_leaf_ridge_fit": (_i, [_p, _p, _p, _p, _d, _d, _i, _p, _p, _p, _p]),
    "art_leaf_ridge_fit_static": (_i, [_p, _i, Basis, _p, _p, _p, _d, _d, _p, _p, _p, _p]),
    "art_gated_plan_create": (_i, [_i, _i64, _i, _i, ctypes.POINTER(_p)]),
    "art_gated_ridge_fit": (_i, [_p, _p, _p, _p, _p, _p, _i, _d, _d, _p, _p, _p, _p]),
    "art_plan_info": (_i, [_p, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "art_leaf_fallback_count": (_i, [_p, ctypes.POINTER(_i)]),
    "art_leaf_rcond": (_i, [_p, _p, _p, _p]),
    "art_leaf_set_refine_tol": (_i, [_p, _d]),
    "art_leaf_set_refine_steps": (_i, [_p, _i]),
    "art_plan_timing_enable": (_i, [_p, _i]),
    "art_plan_timing_read": (_i, [_p, _p, _p, _i]),
    "art_plan_destroy": (None, [_p]),
    "art_split_optimize": (_i, [ctypes.POINTER(SplitArgs), _i, _p]),
    "art_split_optimize_parts": (_i, [ctypes.POINTER(SplitArgs), _i, _i, _p, _p, _p, _p]),
    "art_split_part_doubles": (_i, [_i, _i]),
    "art_split_max_parts": (_i, [ctypes.POINTER(SplitArgs), _i]),
    "art_split_anneal_batched": (_i, [ctypes.POINTER(SplitArgs), _i, _p]),
    "art_split_init": (_i, [_p, _i, _i, _p, _i, ctypes.c_uint64, ctypes.c_uint64, _p, _p]),
    "art_split_init_ranges": (_i, [_p, _p, _i, _p, _i, ctypes.c_uint64, ctypes.c_uint64, _p,
                                   _p, _p]),
    "art_har_sample": (_i, [ctypes.POINTER(HarArgs), _p]),
    "art_philox4x32_10": (_i, [_p, _p, _i64, _p, _p]),
    "art_tree_predict": (_i, [_p, _i64, _p, _p, _p, _i, _p, Basis, _p, _p]),
    "art_leaf_gram": (_i, [_p, _p, _p, _p, ctypes.c_double, _p, _p, _p]),
    "art_leaf_gram_ex": (_i, [_p, _p, _p, _p, ctypes.c_double, _i, _p, _p, _p, _p]),
    "art_gram_size": (_i, [_i]),
    "art_gram_solve": (_i, [_p, _p, _i, Basis, ctypes.c_double, _p, _p, _p, _p]),
    "art_knn_median": (_i, [_p, _p, _p, _i, _i, _i, _p, _p]),
    "art_tree_leaf": (_i, [_p, _i64, _i, _p, _p, _p, _i, _p, _p]),
    "art_gram_sse": (_i, [_p, _i, _i, _i, _p, _p, _p, _p]),
    "art_gather_strided": (_i, [_p, _p, _p, _i, _i, _p, _p, _p, _i64, _p, _p, _p, _p]),
    "art_partition_rows": (_i, [_p, _p, _p, _i, _i, _p, _p, _p, _i64, _p, _p, _p, _p, _p, _p,
                                _p, _p, _p, _p]),
    "art_partition_block_rows": (_i, []),
}


def exported_symbols():
    """Names every build of libart_hip.so must export."""
    return tuple(_PROTOTYPES)


def load(path: os.PathLike | None = None):
    """Load (once) and return the ctypes handle of libart_hip.so."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    target = Path(path) if path is not None else LIB_PATH
    if not target.exists():
        raise HipLibraryError(
            f"{target} is missing: build it with `python -c 'import __graft_entry__ as g; "
            f"g.build()'` (hipcc --offload-arch=gfx950).")
    lib = ctypes.CDLL(str(target))
    for name, (res, args) in _PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.art_version() != 1:
        raise HipLibraryError("libart_hip.so ABI version mismatch; rebuild it.")
    if path is None:
        _lib = lib
    return lib


def check(rc: int) -> None:
    """Translate a negative ART_ERR_* return into an exception."""
    if rc != ART_OK:
        msg = load().art_last_error().decode(errors="replace")
        if rc == -1:
            raise ValueError(msg)
        raise HipLibraryError(f"libart_hip error {rc}: {msg}")


def torch_device():
    """torch and a HIP device check; raises when no GPU is visible."""
    import torch

    if not torch.cuda.is_available():
        raise HipLibraryError("No HIP device is visible; the ART hot path runs only on a GPU.")
    return torch


def stream_handle(stream) -> int:
    return int(stream.cuda_stream)


def ptr(t) -> int:
    return int(t.data_ptr())


def upload(a, device, dtype=None):
    """Small host array -> device tensor, enqueued on the current stream with
    no host wait: the values are staged in pinned host memory (torch's
    caching host allocator keeps the staging block until the copy has run),
    so a later change to ``a`` does not reach the copy.  A pageable copy
    would first wait for every kernel already queued on the stream."""
    import numpy as np
    import torch

    t = torch.from_numpy(np.ascontiguousarray(a, dtype=dtype))
    return t.pin_memory().to(device, non_blocking=True)


class PinnedRing:
    """Pinned host staging for small uploads that one thread enqueues on one
    stream: a bump allocator over one pinned buffer (no allocation per
    upload).  When it wraps, the stream is synchronized first, so every copy
    of the previous lap has run before its bytes are overwritten.  Uploads
    larger than a quarter of the ring go through ``upload``."""

    ALIGN = 256

    def __init__(self, stream, nbytes=8 << 20):
        import torch

        self.torch, self.stream = torch, stream
        self.buf = torch.empty(nbytes, dtype=torch.uint8, pin_memory=True)
        self.host = self.buf.numpy()
        self.size, self.at = nbytes, 0

    def upload(self, a, device, dtype=None):
        import numpy as np

        torch = self.torch
        a = np.ascontiguousarray(a, dtype=dtype)
        n = a.nbytes
        if 4 * n > self.size:
            return upload(a, device)
        if self.at + n > self.size:
            self.stream.synchronize()
            self.at = 0
        at = self.at
        self.host[at:at + n] = a.reshape(-1).view(np.uint8)
        self.at = (at + n + self.ALIGN - 1) // self.ALIGN * self.ALIGN
        src = self.buf[at:at + n].view(torch.from_numpy(a[:0]).dtype).view(a.shape)
        with torch.cuda.stream(self.stream):
            return src.to(device, non_blocking=True)
