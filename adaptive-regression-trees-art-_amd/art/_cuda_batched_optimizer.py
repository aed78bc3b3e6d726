"""Reference import path art._cuda_batched_optimizer (art/_cuda_batched_optimizer.py), served by HIP."""

from .batched_optimizer import BatchedOptimizerWorkspace, optimize_batch

__all__ = ["BatchedOptimizerWorkspace", "optimize_batch"]
