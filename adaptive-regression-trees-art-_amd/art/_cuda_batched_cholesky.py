"""Reference import path art._cuda_batched_cholesky (art/_cuda_batched_cholesky.py), served by HIP."""

from .batched_ridge import BatchedCholeskyWorkspace

__all__ = ["BatchedCholeskyWorkspace"]
