"""Resident, batched left/right weighted ridge fits for equally sized nodes.

Drop-in for ``BatchedCholeskyWorkspace`` (art/_cuda_batched_cholesky.py:45-363):
same constructor, ``allocate``/``load``/``fit``/``synchronize``/``close``,
the same device views ``coefficients (K,2,p)``, ``status (K,2)``,
``ridge_effective (K,2)``, the same status codes and the same validation
messages.  The arithmetic runs in libart_hip.so: one MFMA kernel builds both
sides' weighted Grams and right-hand sides per chunk of rows, one wave per
system reduces them, adds the trace-scaled ridge, factors and solves in LDS.
Device arrays are torch tensors; torch is imported only after validation.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _hip

_INT32_MAX = np.iinfo(np.int32).max


def _dimensions(*shape):
    if any(not isinstance(v, (int, np.integer)) or not 0 < v <= _INT32_MAX for v in shape):
        raise ValueError("Workspace dimensions must be positive CUDA integers.")
    return tuple(int(v) for v in shape)


def _node_ids(nodes, limit):
    ids = np.asarray(nodes)
    if ids.ndim != 1 or (ids.size and not np.issubdtype(ids.dtype, np.integer)):
        raise ValueError("nodes must be a one-dimensional integer list.")
    if np.any(ids < 0) or np.any(ids >= limit) or len(np.unique(ids)) != len(ids):
        raise ValueError("nodes must contain distinct loaded slot indices.")
    return np.ascontiguousarray(ids, dtype=np.int32)


def _is_complex(a):
    if hasattr(a, "is_complex") and callable(a.is_complex):
        return bool(a.is_complex())
    return np.iscomplexobj(a)


def _shape(a):
    return tuple(a.shape) if hasattr(a, "shape") else np.shape(a)


class BatchedCholeskyWorkspace:
    """Cache Phi (K,n,p), y (K,n), and buffers for 2K float64 solves.

    ``fit(pi)`` accepts gates of shape (K,n) and returns device views
    ``(coefficients, status)`` of shapes (K,2,p) and (K,2).  Side 0 is left.
    These views are overwritten by the next fit; copy them to retain a result.

    Status is 0 on success, positive for a non-positive Cholesky pivot (its
    one-based index), -1 for a nonfinite/out-of-range gate, -2 for nonfinite
    assembly/factors, and -3 for a failed/nonfinite solve.  Failed
    coefficients are NaN.  There is no condition estimate or fallback.

    Calls enqueue work on ``stream`` (the current torch stream at
    construction by default) without synchronizing.  ``copy=False`` borrows
    C-contiguous float64 device tensors.
    """

    def __init__(self, phi, y, ridge: float = 0.0, weight_floor: float = 1e-12, *,
                 stream=None, copy=True, chunk_rows=0):
        shape = _shape(phi)
        if len(shape) != 3 or any(size < 1 for size in shape):
            raise ValueError("Phi must have nonempty shape (K, n, p).")
        if _shape(y) != shape[:2]:
            raise ValueError("y must have shape (K, n).")
        if _is_complex(phi) or _is_complex(y):
            raise ValueError("Phi and y must be real.")
        if not isinstance(copy, (bool, np.bool_)):
            raise ValueError("copy must be a boolean.")
        self._allocate(shape, ridge, weight_floor, stream, phi, y, copy, chunk_rows)

    @classmethod
    def allocate(cls, capacity, n, p, ridge=0., weight_floor=1e-12, *, stream=None, chunk_rows=0):
        """Reserve all numerical buffers before loading data."""
        ws = cls.__new__(cls)
        ws._allocate(_dimensions(capacity, n, p), ridge, weight_floor, stream,
                     chunk_rows=chunk_rows)
        return ws

    def _allocate(self, shape, ridge, weight_floor, stream, phi=None, y=None, copy=True,
                  chunk_rows=0):
        self.plan = None
        self.K, self.n, self.p = _dimensions(*shape)
        self.capacity, self.k = self.K, 0
        self.ridge, self.weight_floor = float(ridge), float(weight_floor)
        if not np.isfinite(self.ridge) or self.ridge < 0:
            raise ValueError("ridge must be finite and nonnegative.")
        if not np.isfinite(self.weight_floor) or self.weight_floor < 0:
            raise ValueError("weight_floor must be finite and nonnegative.")
        torch = _hip.torch_device()
        self.torch, self.lib = torch, _hip.load()
        self._nodes = None
        self.last_batch_size = 0
        self.fit_count = 0
        self.device_id = torch.cuda.current_device()
        self.device = torch.device("cuda", self.device_id)
        self.stream = torch.cuda.current_stream() if stream is None else stream
        f64 = dict(dtype=torch.float64, device=self.device)
        with torch.cuda.stream(self.stream):
            if phi is None:
                self.phi = torch.empty(shape, **f64)
                self.y = torch.empty(shape[:2], **f64)
            elif copy:
                self.phi = torch.as_tensor(phi).to(**f64).clone().contiguous()
                self.y = torch.as_tensor(y).to(**f64).clone().contiguous()
            else:
                if any(not isinstance(a, torch.Tensor) or not a.is_cuda
                       or a.dtype != torch.float64 or not a.is_contiguous()
                       or a.device.index != self.device_id for a in (phi, y)):
                    raise ValueError("copy=False requires C-contiguous float64 device Phi/y.")
                self.phi, self.y = phi, y
            # Side-major storage; the public views are transposes, as in the reference.
            self._rhs = torch.full((2, self.K, self.p), float("nan"), **f64)
            self._ridge = torch.zeros((2, self.K), **f64)
            self._status = torch.zeros((2, self.K), dtype=torch.int32, device=self.device)
            self._invalid = torch.zeros(self.K, dtype=torch.int32, device=self.device)
            self._pi = torch.empty((self.K, self.n), **f64)
            self._ids = torch.arange(self.K, dtype=torch.int32, device=self.device)
            self.coefficients = self._rhs.transpose(0, 1)
            self.status = self._status.T
            self.ridge_effective = self._ridge.T
            self._all_ids = np.arange(self.K, dtype=np.int32)
            handle = ctypes.c_void_p()
            _hip.check(self.lib.art_gated_plan_create(self.K, self.n, self.p, int(chunk_rows),
                                                      ctypes.byref(handle)))
            self.plan = handle
            if phi is not None:
                self._check_data(self.K)
                self.k = self.K

    def load(self, phi, y):
        """Replace k equally sized datasets in place, preserving all buffers.

        Inputs can be host arrays or device tensors.  A view of self.phi[:k]
        already filled by a feature kernel is accepted without copying.
        """
        self._check_open()
        torch = self.torch
        shape = _shape(phi)
        if (len(shape) != 3 or not 0 < shape[0] <= self.K
                or shape[1:] != (self.n, self.p) or _shape(y) != shape[:2]):
            raise ValueError("Phi/y must have shapes (k,n,p)/(k,n), 0<k<=capacity.")
        if _is_complex(phi) or _is_complex(y):
            raise ValueError("Phi and y must be real.")
        self.k = 0  # A failed upload/validation must not leave a usable stale batch.
        k = shape[0]
        with torch.cuda.stream(self.stream):
            for dst, src in ((self.phi[:k], phi), (self.y[:k], y)):
                if isinstance(src, torch.Tensor) and src.is_cuda:
                    if src.device.index != self.device_id:
                        raise ValueError("Inputs must be on the workspace's device.")
                    if (src.data_ptr() != dst.data_ptr() or src.stride() != dst.stride()
                            or src.dtype != dst.dtype):
                        dst.copy_(src)
                else:
                    dst.copy_(torch.as_tensor(np.ascontiguousarray(src, dtype=np.float64)))
            self._check_data(k)
        self.k, self.last_batch_size, self.fit_count = k, 0, 0
        self._nodes = None
        return self

    def _check_data(self, k):
        _hip.check(self.lib.art_check_finite(
            _hip.ptr(self.phi), _hip.ptr(self.y), k, self.n, self.p, _hip.ptr(self._invalid),
            _hip.stream_handle(self.stream)))
        invalid = self._invalid[:k].cpu()  # synchronizes the stream, as the reference does
        if bool(invalid.any()):
            raise ValueError("Phi and y must contain only finite values.")

    def _check_open(self):
        if self.plan is None:
            raise RuntimeError("The workspace is closed.")
        if self.torch.cuda.current_device() != self.device_id:
            raise ValueError("Use the workspace on the device where it was created.")

    def fit(self, pi, *, nodes=None, weight_floor=None):
        """Enqueue Gram assembly, Cholesky and solves for both sides.

        A contiguous float64 device tensor of gates needs no conversion;
        NumPy gates are uploaded for this call.  ``nodes`` is an optional CPU
        list of distinct slot indices; other slots retain their results.
        ``weight_floor=0`` enables hard refits.
        """
        self._check_open()
        if not self.k:
            raise RuntimeError("Load data before fitting.")
        torch = self.torch
        if _shape(pi) not in ((self.K, self.n), (self.k, self.n)) or _is_complex(pi):
            raise ValueError("pi must be real with shape (k,n) or (capacity,n).")
        on_device = isinstance(pi, torch.Tensor) and pi.is_cuda
        if on_device and pi.device.index != self.device_id:
            raise ValueError("pi must be on the workspace's device.")
        floor = self.weight_floor if weight_floor is None else float(weight_floor)
        if not np.isfinite(floor) or floor < 0:
            raise ValueError("weight_floor must be finite and nonnegative.")
        if nodes is not None:
            nodes = _node_ids(nodes, self.k)
        else:
            nodes = self._all_ids[:self.k]
        count = len(nodes)
        self.last_batch_size = 2 * count
        if not count:
            return self.coefficients, self.status
        self.fit_count += count
        with torch.cuda.stream(self.stream):
            if on_device:
                pi = pi.to(torch.float64).contiguous()
            else:
                rows = _shape(pi)[0]
                self._pi[:rows].copy_(torch.as_tensor(np.ascontiguousarray(pi, dtype=np.float64)))
                pi = self._pi
            if self._nodes is None or not np.array_equal(nodes, self._nodes):
                self._ids[:count].copy_(torch.from_numpy(nodes))
                self._nodes = nodes.copy()
            _hip.check(self.lib.art_gated_ridge_fit(
                self.plan, _hip.ptr(self.phi), _hip.ptr(self.y), _hip.ptr(pi),
                _hip.ptr(self._ids), self._nodes.ctypes.data, count, self.ridge, floor,
                _hip.ptr(self._rhs), _hip.ptr(self._status), _hip.ptr(self._ridge),
                _hip.stream_handle(self.stream)))
        return self.coefficients, self.status

    def synchronize(self) -> None:
        """Wait until this workspace's queued operations have completed."""
        self.stream.synchronize()

    def close(self) -> None:
        if getattr(self, "plan", None) is not None:
            self.synchronize()
            self.lib.art_plan_destroy(self.plan)
            self.plan = None
            for name in ("phi", "y", "_rhs", "_ridge", "_status", "_invalid", "_pi", "_ids",
                         "coefficients", "status", "ridge_effective"):
                setattr(self, name, None)
            self.k = 0

    def __enter__(self):
        return self

    def __exit__(self, *args):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
