"""Chunk-size autotuning for batched leaf fits.

A leaf fit's unit of Gram work is a chunk of at most ``chunk_rows`` rows of
one leaf (one workgroup each, largest first).  Smaller chunks spread big
leaves over more workgroups but write and reduce more partial systems;
larger ones cut the reduction and lengthen the tail.  The best size depends
on the basis (d, degree, bias), the precision and the leaf-size mix, so it
is measured: ``tune_chunk_rows`` times a fit per candidate on the caller's
own data and keeps the fastest, once per key

    (d, p, degree, bias, bias_last, precision, K bucket, N bucket)

with K (leaves) and N (rows) bucketed by powers of two.  Choices are cached
in memory and in a JSON file (``ART_TUNE_CACHE``, default
``~/.cache/art_hip/tune.json``); a cached key runs no timing.
"""

from __future__ import annotations

import json
import math
import os
from pathlib import Path

import numpy as np

DEFAULT_CANDIDATES = (1024, 2048, 4096, 8192)

_memory: dict[str, int] = {}


def cache_path() -> Path:
    return Path(os.environ.get("ART_TUNE_CACHE",
                               Path.home() / ".cache" / "art_hip" / "tune.json"))


def tune_key(d, p, degree, include_bias, bias_last, precision, n_leaves, n_rows):
    """The cache key of a leaf-fit shape (leaf and row counts bucketed by
    powers of two)."""
    kb = int(math.floor(math.log2(max(int(n_leaves), 1))))
    nb = int(math.floor(math.log2(max(int(n_rows), 1))))
    return (f"leaf_fit:d{int(d)}:p{int(p)}:deg{int(degree)}:b{int(bool(include_bias))}"
            f":last{int(bool(bias_last))}:{precision}:K2^{kb}:N2^{nb}")


def _load_file():
    path = cache_path()
    try:
        with open(path) as fh:
            data = json.load(fh)
        return {str(k): int(v) for k, v in data.items()}
    except (OSError, ValueError, TypeError):
        return {}


def lookup(key):
    """The cached chunk size of a key, or None."""
    if key in _memory:
        return _memory[key]
    found = _load_file().get(key)
    if found is not None:
        _memory[key] = found
    return found


def store(key, chunk_rows):
    """Record a choice in memory and in the cache file (best effort)."""
    _memory[key] = int(chunk_rows)
    path = cache_path()
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        data = _load_file()
        data[key] = int(chunk_rows)
        tmp = path.with_suffix(".tmp")
        with open(tmp, "w") as fh:
            json.dump(data, fh, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def tune_chunk_rows(offsets, d, X, y, w, ridge, *, degree=2, include_bias=True,
                    bias_last=False, precision="fp64", candidates=DEFAULT_CANDIDATES, reps=3,
                    weight_floor=1e-12, stream=None, force=False):
    """Fastest chunk size for fitting these leaves, and its timings.

    Returns (chunk_rows, timings) where timings maps each candidate to its
    median fit time in milliseconds (HIP events on the fit stream, after
    one warm-up fit), or is None when the key was cached (unless force)."""
    from .features import _columns
    from .leaf_ridge import LeafRidgeBatch

    offsets = np.asarray(offsets, dtype=np.int64)
    p = _columns(int(d), degree, include_bias, bias_last).shape[1]
    key = tune_key(d, p, degree, include_bias, bias_last, precision, offsets.size - 1,
                   offsets[-1])
    if not force:
        hit = lookup(key)
        if hit is not None:
            return hit, None
    cands = sorted({int(c) for c in candidates})
    if not cands or any(c < 16 or c % 16 for c in cands):
        raise ValueError("Chunk-size candidates must be positive multiples of 16.")
    timings = {}
    for c in cands:
        with LeafRidgeBatch(offsets, d, degree=degree, include_bias=include_bias,
                            bias_last=bias_last, chunk_rows=c, stream=stream) as batch:
            torch = batch.torch
            batch.fit(X, y, w, ridge, weight_floor, precision)
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
            with torch.cuda.stream(batch.stream):
                ev[0].record(batch.stream)
                for i in range(reps):
                    batch.fit(X, y, w, ridge, weight_floor, precision)
                    ev[i + 1].record(batch.stream)
            batch.synchronize()
            timings[c] = float(np.median([ev[i].elapsed_time(ev[i + 1]) for i in range(reps)]))
    best = min(cands, key=lambda c: (timings[c], c))
    store(key, best)
    return best, timings
