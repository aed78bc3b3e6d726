"""Argument validation happens before torch or the HIP library is touched.

Mirrors tests/test_cuda_batched_cholesky.py:132-152 of the reference, where
validation must precede any CuPy import; here it must precede torch.
"""

import builtins

import numpy as np
import pytest


def _guard(monkeypatch):
    original = builtins.__import__

    def guarded(name, *args, **kw):
        if name == "torch" or name.startswith("torch."):
            raise AssertionError("Validation must not require torch")
        return original(name, *args, **kw)

    monkeypatch.setattr(builtins, "__import__", guarded)


@pytest.mark.parametrize("phi,y,kwargs,message", [
    (np.ones((2, 3)), np.ones(3), {}, "Phi"),
    (np.ones((0, 3, 2)), np.ones((0, 3)), {}, "Phi"),
    (np.ones((2, 3, 2)), np.ones((2, 3, 1)), {}, "y"),
    (np.ones((2, 3, 2)), np.ones((2, 3)), {"ridge": -1}, "ridge"),
    (np.ones((2, 3, 2)), np.ones((2, 3)), {"ridge": np.nan}, "ridge"),
    (np.ones((2, 3, 2)), np.ones((2, 3)), {"weight_floor": np.inf}, "weight_floor"),
    (np.ones((2, 3, 2)), np.ones((2, 3)), {"weight_floor": -1}, "weight_floor"),
    (np.ones((2, 3, 2), dtype=complex), np.ones((2, 3)), {}, "real"),
])
def test_batched_workspace_validation_precedes_torch(monkeypatch, phi, y, kwargs, message):
    from art.batched_ridge import BatchedCholeskyWorkspace

    _guard(monkeypatch)
    with pytest.raises(ValueError, match=message):
        BatchedCholeskyWorkspace(phi, y, **kwargs)


@pytest.mark.parametrize("kwargs,message", [
    ({"degree": 3}, "degree"), ({"degree": 2, "bias_last": True}, "bias_last"),
])
def test_feature_plan_validation_precedes_torch(monkeypatch, kwargs, message):
    from art.features import FeaturePlan

    _guard(monkeypatch)
    with pytest.raises(ValueError, match=message):
        FeaturePlan(3, **kwargs)
    with pytest.raises(ValueError, match="positive"):
        FeaturePlan(0)


@pytest.mark.parametrize("offsets,message", [
    ([1, 5], "CSR"), ([0, 5, 3], "nondecreasing"), ([[0, 1]], "CSR"),
])
def test_leaf_batch_validation_precedes_torch(monkeypatch, offsets, message):
    from art.leaf_ridge import LeafRidgeBatch

    _guard(monkeypatch)
    with pytest.raises(ValueError, match=message):
        LeafRidgeBatch(offsets, 2)


@pytest.mark.parametrize("solver,match", [
    ("lu", "solver must be one of"), ("qr", "not available on the GPU ridge path"),
    ("svd", "not available on the GPU ridge path")])
def test_ridge_solver_literals_are_checked_before_any_device_work(solver, match):
    # art/models.py:460-470 validates the literal; "qr" and "svd" (host
    # LAPACK least squares in the reference) are refused by the GPU path
    # instead of silently running Cholesky.
    from art.models import AffineRidgeModel, PolynomialRidgeModel

    X = np.random.default_rng(0).uniform(-1, 1, (20, 2))
    for model in (AffineRidgeModel(1e-6, solver=solver), PolynomialRidgeModel(solver=solver)):
        with pytest.raises(ValueError, match=match):
            model.fit(X, X[:, 0])


def test_autotune_keys_and_cache_round_trip(tmp_path, monkeypatch):
    from art import autotune

    monkeypatch.setenv("ART_TUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(autotune, "_memory", {})
    k1 = autotune.tune_key(8, 45, 2, True, False, "mixed", 4096, 12_217_562)
    # Leaf and row counts are bucketed by powers of two.
    assert k1 == autotune.tune_key(8, 45, 2, True, False, "mixed", 4100, 9_000_000)
    assert k1 != autotune.tune_key(8, 45, 2, True, False, "mixed", 8192, 12_217_562)
    assert k1 != autotune.tune_key(8, 45, 2, True, False, "fp64", 4096, 12_217_562)
    assert autotune.lookup(k1) is None
    autotune.store(k1, 2048)
    monkeypatch.setattr(autotune, "_memory", {})  # a fresh process reads the file
    assert autotune.lookup(k1) == 2048
    # A cached key runs no timing (the device is never touched here).
    offsets = np.array([0, 3000, 12_217_562 - 7], dtype=np.int64)
    chunk, timings = autotune.tune_chunk_rows(
        np.concatenate([offsets[:1], np.linspace(0, 12_217_562, 4097).astype(np.int64)[1:]]),
        8, None, None, None, 1e-6, precision="mixed")
    assert chunk == 2048 and timings is None
