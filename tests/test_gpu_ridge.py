"""GPU parity of features, gated batched fits and varlen leaf fits vs the oracle.

Parameter grids and known answers follow tests/test_cuda_features.py and
tests/test_cuda_batched_cholesky.py of the reference.  Coefficient tolerance
for the FP64 path: rtol 3e-11 / atol 1e-12 on well-conditioned systems, as the
reference requires of its GPU path (test_cuda_batched_cholesky.py:196-197).
"""

import numpy as np
import pytest
from sklearn.preprocessing import PolynomialFeatures

from oracle import ridge as R

pytestmark = pytest.mark.gpu


def dev(torch, a, dtype=None):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype or torch.float64, device="cuda")


# --------------------------------------------------------------------------- features

@pytest.mark.parametrize("K,n,d,degree,bias,last", [
    (2, 7, 3, 1, True, True), (3, 5, 2, 1, True, False), (2, 5, 3, 1, False, False),
    (2, 7, 1, 2, False, False), (20, 3, 5, 2, True, False), (3, 3, 10, 2, False, False),
    (2, 2, 20, 2, True, False), (1, 2, 40, 2, True, False), (2, 259, 2, 2, True, False),
])
def test_features_match_sklearn_exactly(torch_gpu, K, n, d, degree, bias, last):
    from art import build_features

    torch = torch_gpu
    X = np.random.default_rng(901).normal(size=(K, n, d))
    X[:, 0] = 0.
    x = dev(torch, X)
    ptr = x.data_ptr()
    phi = build_features(x, degree=degree, include_bias=bias, bias_last=last)
    torch.cuda.synchronize()
    expected = PolynomialFeatures(degree=degree, include_bias=bias).fit_transform(X.reshape(-1, d))
    if last:
        expected = np.concatenate((expected[:, 1:], expected[:, :1]), axis=1)
    assert tuple(phi.shape) == (K, n, expected.shape[1]) and phi.is_contiguous()
    np.testing.assert_array_equal(phi.cpu().numpy(), expected.reshape(phi.shape))
    np.testing.assert_array_equal(x.cpu().numpy(), X)
    assert x.data_ptr() == ptr


def test_quadratic_feature_known_answer(torch_gpu):
    from art import build_features

    phi = build_features(dev(torch_gpu, [[[2., 3.], [5., 7.]], [[11., 13.], [17., 19.]]]))
    np.testing.assert_array_equal(phi.cpu().numpy(), [
        [[1., 2., 3., 4., 6., 9.], [1., 5., 7., 25., 35., 49.]],
        [[1., 11., 13., 121., 143., 169.], [1., 17., 19., 289., 323., 361.]],
    ])


def test_feature_plan_reuses_supplied_output(torch_gpu):
    from art import FeaturePlan

    torch, rng = torch_gpu, np.random.default_rng(903)
    plan = FeaturePlan(3, degree=2, include_bias=False)
    x = torch.empty((5, 7, 3), dtype=torch.float64, device="cuda")
    phi = torch.empty((5, 7, plan.p), dtype=torch.float64, device="cuda")
    address = phi.data_ptr()
    for k in (5, 2, 1):
        X = rng.normal(size=(k, 7, 3))
        x[:k].copy_(torch.from_numpy(X))
        phi.fill_(-91.)
        out = phi[:k]
        assert plan.build(x[:k], out=out) is out
        torch.cuda.synchronize()
        expected = PolynomialFeatures(degree=2, include_bias=False).fit_transform(X.reshape(-1, 3))
        np.testing.assert_array_equal(phi.cpu().numpy()[:k], expected.reshape(k, 7, -1))
        np.testing.assert_array_equal(phi.cpu().numpy()[k:], -91.)
        assert phi.data_ptr() == address
    with pytest.raises(ValueError, match="out"):
        plan.build(x, out=phi[:, :, :-1])


@pytest.mark.parametrize("bad", ["dtype", "strides", "degree", "bias_last", "shape"])
def test_feature_builder_rejects_unsupported_inputs(torch_gpu, bad):
    from art import build_features

    torch = torch_gpu
    x, kwargs = dev(torch, np.ones((2, 5, 3))), {}
    if bad == "dtype": x = x.float()
    if bad == "strides": x = x.transpose(1, 2).contiguous().transpose(1, 2)
    if bad == "degree": kwargs["degree"] = 3
    if bad == "bias_last": kwargs["bias_last"] = True
    if bad == "shape": x = x[0]
    with pytest.raises(ValueError):
        build_features(x, **kwargs)


# --------------------------------------------------------------------------- gated fits

@pytest.mark.parametrize("K,p,ridge,floor", [
    (1, 1, 0., 0.), (1, 6, 0., 1e-12), (3, 6, .03, .07), (3, 6, 0., .75),
    (20, 21, 1e-5, 1e-12), (3, 231, 1e-5, .07), (1, 259, .03, 0.),
])
def test_gated_fits_match_oracle_and_reuse_storage(torch_gpu, K, p, ridge, floor):
    from art import BatchedCholeskyWorkspace

    torch, rng = torch_gpu, np.random.default_rng(603)
    phi, y = rng.normal(size=(K, 5 * p + 4, p)), rng.normal(size=(K, 5 * p + 4))
    pi = rng.uniform(size=y.shape)
    pi[:, :4] = [0., 1., 1e-20, 1. - 1e-15]
    stream = torch.cuda.Stream()
    with BatchedCholeskyWorkspace(phi, y, ridge, floor, stream=stream) as ws:
        addresses = [t.data_ptr() for t in (ws.phi, ws.y, ws._rhs, ws._status)]
        for iteration, gates in enumerate((pi, 1. - pi)):
            with torch.cuda.stream(stream):
                device_pi = dev(torch, gates)
            coefficients, status = ws.fit(device_pi)
            assert coefficients is ws.coefficients and status is ws.status
            ws.synchronize()
            np.testing.assert_array_equal(status.cpu().numpy(), 0)
            actual = coefficients.cpu().numpy()
            if iteration == 0:
                first = actual.copy()
            for j in range(K):
                for side in range(2):
                    q = 1. - gates[j] if side == 0 else gates[j]
                    ref, lam = R.solve_weighted_ridge(phi[j], y[j], q, ridge, floor)
                    np.testing.assert_allclose(actual[j, side], ref, rtol=3e-11, atol=1e-12)
                    assert float(ws.ridge_effective[j, side]) == pytest.approx(lam, rel=1e-12)
            assert addresses == [t.data_ptr() for t in (ws.phi, ws.y, ws._rhs, ws._status)]
        phi.fill(0.); y.fill(0.)
        ws.fit(pi)  # host gates; device data were copied once
        ws.synchronize()
        np.testing.assert_array_equal(ws.status.cpu().numpy(), 0)
        np.testing.assert_allclose(ws.coefficients.cpu().numpy(), first, rtol=2e-12, atol=1e-13)
    with pytest.raises(RuntimeError, match="closed"):
        ws.fit(pi)


def test_gated_failure_isolation_and_recovery(torch_gpu):
    from art import BatchedCholeskyWorkspace

    rng = np.random.default_rng(604)
    phi, y = rng.normal(size=(3, 20, 3)), rng.normal(size=(3, 20))
    phi[1, :, 2] = 0.  # Both fits at node 1 fail at pivot 3.
    pi = np.full((3, 20), .4)
    pi[2] = 0.  # Only the right fit at node 2 is empty.
    with BatchedCholeskyWorkspace(phi, y, weight_floor=0.) as ws:
        coefficients, status = ws.fit(pi)
        ws.synchronize()
        np.testing.assert_array_equal(status.cpu().numpy(), [[0, 0], [3, 3], [0, 1]])
        result = coefficients.cpu().numpy()
        assert np.isfinite(result[0]).all() and np.isfinite(result[2, 0]).all()
        assert np.isnan(result[1]).all() and np.isnan(result[2, 1]).all()
        for side, q in enumerate((1. - pi[0], pi[0])):
            ref, _ = R.solve_weighted_ridge(phi[0], y[0], q, 0., 0.)
            np.testing.assert_allclose(result[0, side], ref, rtol=2e-12, atol=1e-13)
        pi[0, 0] = np.nan
        ws.fit(pi); ws.synchronize()
        np.testing.assert_array_equal(ws.status.cpu().numpy()[0], [-1, -1])
        pi[0, 0] = .4
        pi[2] = .4
        ws.fit(pi); ws.synchronize()
        np.testing.assert_array_equal(ws.status.cpu().numpy(), [[0, 0], [3, 3], [0, 0]])


@pytest.mark.parametrize("value", [np.nan, np.inf, -.1, 1.1])
def test_gated_bad_gate_marks_only_its_node(torch_gpu, value):
    from art import BatchedCholeskyWorkspace

    pi = np.full((3, 8), .4)
    pi[1, 2] = value
    with BatchedCholeskyWorkspace(np.ones((3, 8, 1)), np.ones((3, 8))) as ws:
        ws.fit(pi); ws.synchronize()
        np.testing.assert_array_equal(ws.status.cpu().numpy().T, [[0, -1, 0], [0, -1, 0]])
        pi[1, 2] = .4
        ws.fit(pi); ws.synchronize()
        np.testing.assert_array_equal(ws.status.cpu().numpy(), 0)


def test_gated_solve_overflow_is_status_minus_three(torch_gpu):
    from art import BatchedCholeskyWorkspace

    # One-feature system 1e-300 * b = 1e10 -> inf: factor succeeds, solve overflows.
    phi = np.full((1, 4, 1), 1e-150)
    y = np.full((1, 4), 1e160)
    with BatchedCholeskyWorkspace(phi, y, weight_floor=0.) as ws:
        ws.fit(np.full((1, 4), .5)); ws.synchronize()
        np.testing.assert_array_equal(ws.status.cpu().numpy(), [[-3, -3]])
        assert np.isnan(ws.coefficients.cpu().numpy()).all()
    _, st, _ = R.batched_gate_fit(phi, y, np.full((1, 4), .5), floor=0.)
    np.testing.assert_array_equal(st, [[-3, -3]])


def test_gated_subset_nodes_keep_other_slots(torch_gpu):
    from art import BatchedCholeskyWorkspace

    rng = np.random.default_rng(902)
    phi, y, gates = rng.normal(size=(4, 31, 5)), rng.normal(size=(4, 31)), rng.uniform(size=(4, 31))
    with BatchedCholeskyWorkspace.allocate(6, 31, 5, ridge=.01) as ws:
        ws.load(phi, y)
        ws.fit(gates)
        ws.synchronize()
        before = ws.coefficients.cpu().numpy()
        ws.fit(1. - gates, nodes=[2, 0])
        ws.synchronize()
        after = ws.coefficients.cpu().numpy()
        np.testing.assert_array_equal(after[[1, 3]], before[[1, 3]])
        assert ws.last_batch_size == 4
        for j in (2, 0):
            for side, q in enumerate((gates[j], 1. - gates[j])):
                ref, _ = R.solve_weighted_ridge(phi[j], y[j], q, .01)
                np.testing.assert_allclose(after[j, side], ref, rtol=3e-11, atol=1e-12)
        with pytest.raises(ValueError, match="distinct"):
            ws.fit(gates, nodes=[0, 0])
        with pytest.raises(ValueError, match="distinct"):
            ws.fit(gates, nodes=[4])


def test_gated_input_validation(torch_gpu):
    from art import BatchedCholeskyWorkspace

    phi, y = np.ones((2, 8, 2)), np.ones((2, 8))
    phi[0, 0, 0] = np.nan
    with pytest.raises(ValueError, match="finite"):
        BatchedCholeskyWorkspace(phi, y)
    phi[0, 0, 0], y[0, 0] = 1., np.inf
    with pytest.raises(ValueError, match="finite"):
        BatchedCholeskyWorkspace(phi, y)
    y[0, 0] = 1.
    with BatchedCholeskyWorkspace(phi, y, ridge=.1) as ws:
        with pytest.raises(ValueError, match="shape"):
            ws.fit(np.ones(8))
        with pytest.raises(ValueError, match="real"):
            ws.fit(np.ones((2, 8), dtype=complex))
    with pytest.raises(RuntimeError, match="Load data"):
        BatchedCholeskyWorkspace.allocate(2, 8, 2).fit(np.ones((2, 8)))


def test_borrowed_device_design(torch_gpu):
    from art import BatchedCholeskyWorkspace, build_features

    torch, rng = torch_gpu, np.random.default_rng(902)
    X, y = rng.normal(size=(3, 31, 3)), rng.normal(size=(3, 31))
    gates = rng.uniform(size=y.shape)
    x, target = dev(torch, X), dev(torch, y)
    phi = build_features(x)
    with BatchedCholeskyWorkspace(phi, target, ridge=.01, copy=False) as ws:
        assert ws.phi is phi and ws.y is target
        ws.fit(dev(torch, gates), nodes=[2, 0])
        ws.synchronize()
        actual = ws.coefficients.cpu().numpy()
        expected_phi = R.features(X)
        for j in (2, 0):
            for side, q in enumerate((1. - gates[j], gates[j])):
                ref, _ = R.solve_weighted_ridge(expected_phi[j], y[j], q, .01)
                np.testing.assert_allclose(actual[j, side], ref, rtol=3e-11, atol=1e-12)
    with BatchedCholeskyWorkspace(phi, target) as ws:
        assert ws.phi.data_ptr() != phi.data_ptr()
    with pytest.raises(ValueError, match="copy=False"):
        BatchedCholeskyWorkspace(phi.cpu(), target, copy=False)


# --------------------------------------------------------------------------- varlen leaves

def _rel_err(a, b):
    return np.linalg.norm(a - b, axis=-1) / np.maximum(np.linalg.norm(b, axis=-1), 1e-300)


@pytest.mark.parametrize("d,degree,bias,last,chunk", [
    (2, 2, True, False, 0), (8, 2, True, False, 0), (8, 2, True, False, 64),
    (12, 2, True, False, 0), (16, 1, True, True, 0), (20, 1, True, False, 0), (3, 2, False, False, 16),
])
def test_leaf_fits_match_oracle(torch_gpu, d, degree, bias, last, chunk):
    from art import LeafRidgeBatch

    torch, rng = torch_gpu, np.random.default_rng(1000 + d)
    sizes = np.array([300, 1, 17, 64, 129, 2048 + 7, 0, 500])
    offsets = np.concatenate([[0], np.cumsum(sizes)])
    N = offsets[-1]
    X = rng.uniform(-1, 1, size=(N, d))
    y, w = np.sin(3 * X[:, 0]) + X[:, -1] ** 2, rng.uniform(0.05, 1.0, size=N)
    ref, st_ref, lam_ref = R.leaf_fit(X, y, w, offsets, 1e-3, 1e-12, degree, bias, last)
    with LeafRidgeBatch(offsets, d, degree=degree, include_bias=bias, bias_last=last,
                        chunk_rows=chunk) as batch:
        coef, status, lam = batch.fit(dev(torch, X), dev(torch, y), dev(torch, w), 1e-3)
        batch.synchronize()
        st = status.cpu().numpy()
        np.testing.assert_array_equal(st, st_ref)
        ok = st == 0
        got = coef.cpu().numpy()
        assert np.isnan(got[~ok]).all()
        assert _rel_err(got[ok], ref[ok]).max() < 1e-9
        np.testing.assert_allclose(lam.cpu().numpy(), lam_ref, rtol=1e-12)


def test_leaf_bad_weights_flag_only_their_leaf(torch_gpu):
    from art import LeafRidgeBatch

    torch, rng = torch_gpu, np.random.default_rng(11)
    offsets = np.array([0, 40, 80, 120, 160])
    X, y = rng.uniform(-1, 1, size=(160, 2)), rng.normal(size=160)
    w = rng.uniform(size=160)
    w[45], w[130] = np.nan, -1.0
    with LeafRidgeBatch(offsets, 2) as batch:
        _, status, _ = batch.fit(dev(torch, X), dev(torch, y), dev(torch, w), 1e-6)
        np.testing.assert_array_equal(status.cpu().numpy(), [0, -1, 0, -1])


def test_leaf_workload_subset_matches_oracle(torch_gpu):
    """Synthetic leaf-ridge workload (d=8 quadratic): FP64 GPU vs FP64 oracle."""
    from art import LeafRidgeBatch
    from art.synthetic import leaf_ridge_workload

    torch = torch_gpu
    wl = leaf_ridge_workload(n_leaves=64, seed=2)
    ref, st_ref, _ = R.leaf_fit(wl.X, wl.y, wl.w, wl.offsets, wl.ridge)
    with LeafRidgeBatch(wl.offsets, 8) as batch:
        coef, status, _ = batch.fit(dev(torch, wl.X), dev(torch, wl.y), dev(torch, wl.w), wl.ridge)
        np.testing.assert_array_equal(status.cpu().numpy(), st_ref)
        # Both sides are FP64; they differ only by rounding amplified by the
        # leaf condition number (up to ~1e7 here).
        assert _rel_err(coef.cpu().numpy(), ref).max() < 1e-7


# --------------------------------------------------------------------------- mixed precision

@pytest.mark.parametrize("n_leaves,seed", [(256, 2), (64, 5)])
def test_mixed_leaf_fits_within_1e6_of_fp64(torch_gpu, n_leaves, seed):
    """Mixed precision (split-bf16 Gram + FP64 refinement + FP64 fallback) stays
    within 1e-6 relative coefficient error of the FP64 path on every leaf."""
    from art import LeafRidgeBatch
    from art.synthetic import leaf_ridge_workload

    torch = torch_gpu
    wl = leaf_ridge_workload(n_leaves=n_leaves, seed=seed)
    X, y, w = dev(torch, wl.X), dev(torch, wl.y), dev(torch, wl.w)
    with LeafRidgeBatch(wl.offsets, 8) as batch:
        c64, s64, l64 = (t.clone() for t in batch.fit(X, y, w, wl.ridge, precision="fp64"))
        cm, sm, lm = batch.fit(X, y, w, wl.ridge, precision="mixed")
        np.testing.assert_array_equal(sm.cpu().numpy(), s64.cpu().numpy())
        np.testing.assert_allclose(lm.cpu().numpy(), l64.cpu().numpy(), rtol=1e-12)
        err = _rel_err(cm.cpu().numpy(), c64.cpu().numpy())
        assert err.max() <= 1e-6, err.max()
        # The split-bf16 Gram is accurate to ~1e-8, so only the worst-conditioned
        # leaves (~1 %) need the FP64 path; a broken split sends most of them there.
        assert 0 <= batch.fallback_count() <= n_leaves // 8, batch.fallback_count()
    ref, st_ref, _ = R.leaf_fit(wl.X, wl.y, wl.w, wl.offsets, wl.ridge)
    np.testing.assert_array_equal(sm.cpu().numpy(), st_ref)
    assert _rel_err(cm.cpu().numpy(), ref).max() <= 1e-6


@pytest.mark.parametrize("d", [5, 9])
def test_mixed_fits_other_basis_sizes(torch_gpu, d):
    """Quadratic bases of p = 21 and p = 55 (the 32- and 64-column solve and
    correction kernels): mixed within 1e-6 of FP64 on every leaf."""
    from art import LeafRidgeBatch

    torch, rng = torch_gpu, np.random.default_rng(30 + d)
    sizes = rng.integers(300, 3000, size=48)
    offsets = np.concatenate([[0], np.cumsum(sizes)])
    N = offsets[-1]
    X = rng.uniform(-1, 1, size=(N, d))
    y = np.sin(X @ rng.normal(size=d)) + 0.01 * rng.normal(size=N)
    w = rng.uniform(0.2, 1.0, size=N)
    Xd, yd, wd = dev(torch, X), dev(torch, y), dev(torch, w)
    with LeafRidgeBatch(offsets, d) as batch:
        c64, s64, _ = (t.clone() for t in batch.fit(Xd, yd, wd, 1e-6, precision="fp64"))
        cm, sm, _ = batch.fit(Xd, yd, wd, 1e-6, precision="mixed")
        np.testing.assert_array_equal(sm.cpu().numpy(), s64.cpu().numpy())
        assert (s64.cpu().numpy() == 0).all()
        err = _rel_err(cm.cpu().numpy(), c64.cpu().numpy())
        assert err.max() <= 1e-6, err.max()


@pytest.mark.parametrize("d,degree,bias,last,chunk", [
    (3, 2, False, False, 0), (8, 2, False, False, 0), (8, 2, False, False, 16384),
    (3, 2, True, False, 16384),
    (3, 1, True, False, 0), (12, 1, True, False, 0), (16, 1, True, False, 0),
    (3, 1, True, True, 0), (12, 1, True, True, 0), (16, 1, True, True, 0),
    (3, 1, False, False, 0), (12, 1, False, False, 0), (16, 1, False, False, 0),
])
def test_mixed_fits_every_basis_layout(torch_gpu, d, degree, bias, last, chunk):
    """Mixed within 1e-6 of FP64 for quadratic bases without a bias (the
    register residual kernel, whose partials have stride p) and every linear
    layout (bias first, bias last, none: the feature-layout residual kernel),
    with one leaf above 8192 rows (several residual chunks) and, in two
    cases, 16384-row Gram chunks."""
    from art import LeafRidgeBatch

    torch, rng = torch_gpu, np.random.default_rng(400 + 7 * d + degree + 2 * bias + last)
    sizes = np.concatenate([rng.integers(200, 3000, size=20), [9000, 17000, 64]])
    offsets = np.concatenate([[0], np.cumsum(sizes)])
    N = offsets[-1]
    X = rng.uniform(-1, 1, size=(N, d))
    y = np.sin(X @ rng.normal(size=d)) + 0.3 * X[:, 0] ** 2 + 0.01 * rng.normal(size=N)
    w = rng.uniform(0.2, 1.0, size=N)
    Xd, yd, wd = dev(torch, X), dev(torch, y), dev(torch, w)
    with LeafRidgeBatch(offsets, d, degree=degree, include_bias=bias, bias_last=last,
                        chunk_rows=chunk) as batch:
        c64, s64, l64 = (t.clone() for t in batch.fit(Xd, yd, wd, 1e-6, precision="fp64"))
        cm, sm, lm = batch.fit(Xd, yd, wd, 1e-6, precision="mixed")
        np.testing.assert_array_equal(sm.cpu().numpy(), s64.cpu().numpy())
        assert (s64.cpu().numpy() == 0).all()
        np.testing.assert_allclose(lm.cpu().numpy(), l64.cpu().numpy(), rtol=1e-12)
        err = _rel_err(cm.cpu().numpy(), c64.cpu().numpy())
        assert err.max() <= 1e-6, err.max()
        # Well-conditioned leaves: a wrong residual sends most of them to the
        # FP64 path (or, unflagged, past the bound above).
        assert batch.fallback_count() <= len(sizes) // 4, batch.fallback_count()
    ref, st_ref, _ = R.leaf_fit(X, y, w, offsets, 1e-6, 1e-12, degree, bias, last)
    np.testing.assert_array_equal(sm.cpu().numpy(), st_ref)
    assert _rel_err(cm.cpu().numpy(), ref).max() <= 1e-6


def test_mixed_edge_leaves_match_fp64(torch_gpu):
    """Empty, one-row, flagged and tiny-weight leaves take the exact path."""
    from art import LeafRidgeBatch

    torch, rng = torch_gpu, np.random.default_rng(17)
    sizes = np.array([0, 1, 50, 3000, 64, 700])
    offsets = np.concatenate([[0], np.cumsum(sizes)])
    N = offsets[-1]
    X = rng.uniform(-1, 1, size=(N, 2))
    y = np.cos(2 * X[:, 0]) * X[:, 1]
    w = rng.uniform(size=N)
    w[offsets[2] + 3] = np.nan            # leaf 2 flagged (-1)
    w[offsets[4]:offsets[5]] = 1e-13      # leaf 4: every weight below the floor
    Xd, yd, wd = dev(torch, X), dev(torch, y), dev(torch, w)
    with LeafRidgeBatch(offsets, 2) as batch:
        c64, s64, _ = (t.clone() for t in batch.fit(Xd, yd, wd, 1e-6, precision="fp64"))
        cm, sm, _ = batch.fit(Xd, yd, wd, 1e-6, precision="mixed")
        np.testing.assert_array_equal(sm.cpu().numpy(), s64.cpu().numpy())
        ok = s64.cpu().numpy() == 0
        assert not ok[0] and not ok[2]
        assert _rel_err(cm.cpu().numpy()[ok], c64.cpu().numpy()[ok]).max() <= 1e-6
        assert np.isnan(cm.cpu().numpy()[~ok]).all()


def test_mixed_rejects_wide_bases(torch_gpu):
    from art import LeafRidgeBatch

    torch = torch_gpu
    offsets = np.array([0, 200])
    X = np.random.default_rng(3).uniform(-1, 1, size=(200, 12))
    with LeafRidgeBatch(offsets, 12) as batch:  # p = 91 > 64
        with pytest.raises(ValueError, match="mixed"):
            batch.fit(dev(torch, X), dev(torch, X[:, 0]), dev(torch, np.ones(200)), 1e-6,
                      precision="mixed")


# --------------------------------------------------------------------------- resident systems

@pytest.mark.parametrize("d,degree", [(2, 2), (5, 2), (8, 2), (12, 1)])
def test_resident_grams_solve_like_the_fp64_fit(torch_gpu, d, degree):
    """grams() followed by solve_grams() gives LeafRidgeBatch.fit's FP64 answer."""
    from art import LeafRidgeBatch
    from art.leaf_ridge import solve_grams

    torch, rng = torch_gpu, np.random.default_rng(70 + d)
    sizes = np.array([300, 1, 17, 0, 129, 2055, 64])
    offsets = np.concatenate([[0], np.cumsum(sizes)])
    N = offsets[-1]
    X = rng.uniform(-1, 1, size=(N, d))
    y, w = np.sin(2 * X[:, 0]) + X[:, -1] ** 2, rng.uniform(0.05, 1.0, size=N)
    w[offsets[2] + 4] = np.nan
    Xd, yd, wd = dev(torch, X), dev(torch, y), dev(torch, w)
    with LeafRidgeBatch(offsets, d, degree=degree) as batch:
        c64, s64, l64 = (t.clone() for t in batch.fit(Xd, yd, wd, 1e-4))
        grams, bad = batch.grams(Xd, yd, wd)
        coef, status, lam = solve_grams(grams, bad, d, 1e-4, degree=degree)
        torch.cuda.synchronize()
    np.testing.assert_array_equal(status.cpu().numpy(), s64.cpu().numpy())
    assert status.cpu().numpy()[2] == -1
    ok = s64.cpu().numpy() == 0
    np.testing.assert_allclose(coef.cpu().numpy()[ok], c64.cpu().numpy()[ok], rtol=1e-11,
                               atol=1e-13)
    np.testing.assert_allclose(lam.cpu().numpy(), l64.cpu().numpy(), rtol=1e-13)


def _unpack_tiles(row, T):
    """Full (16T, 16T) symmetric matrix of one resident system in tile layout
    (tile pair (t1 <= t2, t1 outermost), element (a, b) = G[16 t1 + a][16 t2 + b])."""
    G = np.zeros((16 * T, 16 * T))
    q = 0
    for t1 in range(T):
        for t2 in range(t1, T):
            blk = row[256 * q:256 * (q + 1)].reshape(16, 16)
            G[16 * t1:16 * t1 + 16, 16 * t2:16 * t2 + 16] = blk
            G[16 * t2:16 * t2 + 16, 16 * t1:16 * t1 + 16] = blk.T
            q += 1
    return G


@pytest.mark.parametrize("d,degree", [(4, 2), (8, 2), (16, 1)])
def test_masked_and_mixed_resident_grams(torch_gpu, d, degree):
    """art_leaf_gram_ex: a mask forms only the selected leaves' systems (the
    others keep what the output held) and bitwise what the unmasked call
    forms; mixed-precision systems are within 4e-7 of the FP64 ones on the
    Cauchy-Schwarz scale sqrt(G_ii G_jj), padding exactly zero.  The bound:
    each operand v = sqrt(w) x_a x_b is formed in FP32 (up to five roundings,
    <= 2.5 * 2^-24 relative), so a product v_i v_j is off by <= 3e-7 of
    |v_i v_j|, and the sum over rows by <= 3e-7 of sum |v_i v_j| <= the scale;
    a one-row leaf reaches it, large leaves average down to ~1e-8."""
    from art import LeafRidgeBatch

    torch, rng = torch_gpu, np.random.default_rng(90 + d)
    sizes = np.array([5000, 1, 64, 0, 333, 9000, 17, 2048])
    offsets = np.concatenate([[0], np.cumsum(sizes)])
    N = offsets[-1]
    X = rng.uniform(-1, 1, size=(N, d))
    y, w = np.cos(3 * X[:, 0]) * X[:, -1], rng.uniform(0.05, 1.0, size=N)
    Xd, yd, wd = dev(torch, X), dev(torch, y), dev(torch, w)
    sel = np.array([1, 0, 1, 1, 0, 1, 0, 1], dtype=np.int32)
    mask = torch.as_tensor(sel, device="cuda")
    with LeafRidgeBatch(offsets, d, degree=degree) as batch:
        T = (batch.p + 1 + 15) // 16
        full64 = [t.clone() for t in batch.grams(Xd, yd, wd)]
        fullmx = [t.clone() for t in batch.grams(Xd, yd, wd, precision="mixed")]
        outs = {}
        for prec in ("fp64", "mixed"):
            g = torch.full_like(full64[0], 7.0)
            b = torch.full_like(full64[1], 5)
            batch.grams(Xd, yd, wd, precision=prec, mask=mask, out=(g, b))
            outs[prec] = (g, b)
        torch.cuda.synchronize()
        with pytest.raises(ValueError):
            batch.grams(Xd, yd, wd, precision="bf16")
    for prec, full in (("fp64", full64), ("mixed", fullmx)):
        g, b = (t.cpu().numpy() for t in outs[prec])
        fg, fb = (t.cpu().numpy() for t in full)
        np.testing.assert_array_equal(g[sel == 1], fg[sel == 1])
        np.testing.assert_array_equal(b[sel == 1], fb[sel == 1])
        assert (g[sel == 0] == 7.0).all() and (b[sel == 0] == 5).all()
    g64, gmx = full64[0].cpu().numpy(), fullmx[0].cpu().numpy()
    worst = 0.0
    for j in range(len(sizes)):
        A, M = _unpack_tiles(g64[j], T), _unpack_tiles(gmx[j], T)
        scale = np.sqrt(np.outer(np.diag(A), np.diag(A)))
        pad = scale == 0
        assert (M[pad] == 0).all() and (A[pad] == 0).all()
        if (~pad).any():
            worst = max(worst, float(np.max(np.abs(M - A)[~pad] / scale[~pad])))
    assert worst < 4e-7, worst


def test_resident_grams_of_disjoint_rows_add(torch_gpu):
    """Rank-k refits: the systems of two disjoint row sets of each leaf sum to
    the system of their union, and its solve matches the oracle on all rows."""
    from art import LeafRidgeBatch
    from art.leaf_ridge import solve_grams

    torch, rng = torch_gpu, np.random.default_rng(71)
    d, sizes_a, sizes_b = 3, np.array([200, 40, 1000, 7]), np.array([30, 500, 2, 90])
    Xa = [rng.uniform(-1, 1, (n, d)) for n in sizes_a]
    Xb = [rng.uniform(-1, 1, (n, d)) for n in sizes_b]
    f = lambda X: np.exp(X[:, 0]) - X[:, 1] * X[:, 2]
    wa = [rng.uniform(0.1, 1, n) for n in sizes_a]
    wb = [rng.uniform(0.1, 1, n) for n in sizes_b]

    def grams_of(Xs, ws):
        offsets = np.concatenate([[0], np.cumsum([len(x) for x in Xs])])
        X = np.vstack(Xs)
        with LeafRidgeBatch(offsets, d) as batch:
            g, b = batch.grams(dev(torch, X), dev(torch, f(X)), dev(torch, np.concatenate(ws)))
            torch.cuda.synchronize()
        return g, b

    ga, ba = grams_of(Xa, wa)
    gb, bb = grams_of(Xb, wb)
    gu, bu = grams_of([np.vstack(p) for p in zip(Xa, Xb)],
                      [np.concatenate(p) for p in zip(wa, wb)])
    np.testing.assert_allclose((ga + gb).cpu().numpy(), gu.cpu().numpy(), rtol=1e-12,
                               atol=1e-12)
    coef, status, _ = solve_grams(ga + gb, ba | bb, d, 1e-6)
    Xu = np.vstack([np.vstack(p) for p in zip(Xa, Xb)])
    wu = np.concatenate([np.concatenate(p) for p in zip(wa, wb)])
    offsets = np.concatenate([[0], np.cumsum(sizes_a + sizes_b)])
    ref, st_ref, _ = R.leaf_fit(Xu, f(Xu), wu, offsets, 1e-6, 1e-12, 2, True, False)
    np.testing.assert_array_equal(status.cpu().numpy(), st_ref)
    assert _rel_err(coef.cpu().numpy(), ref).max() < 1e-9


# --------------------------------------------------------------------------- solvers

def _exact_rcond(G):
    return 1.0 / (np.linalg.norm(G, 1) * np.linalg.norm(np.linalg.inv(G), 1))


def test_auto_solver_on_a_well_conditioned_fit_is_the_cholesky_fit(torch_gpu):
    from art.models import PolynomialRidgeModel

    rng = np.random.default_rng(5)
    X = rng.uniform(-1, 1, (400, 3))
    y = np.sin(X @ np.array([1.0, -2.0, .5]))
    normal = PolynomialRidgeModel(2, ridge=1e-6).fit(X, y)
    auto = PolynomialRidgeModel(2, ridge=1e-6, solver="auto").fit(X, y)
    r = auto.solve_result_
    assert r.solver_requested == "auto" and r.solver_used == "cholesky"
    assert r.condition_estimator == "sqrt_gram_1norm" and r.fallback_reason is None
    phi = R.features(X, 2, True, False)
    G = phi.T @ phi
    G += 1e-6 * np.trace(G) / G.shape[0] * np.eye(G.shape[0])
    assert r.cond_estimate == pytest.approx(np.sqrt(1.0 / _exact_rcond(G)), rel=1e-8)
    np.testing.assert_allclose(auto.coef_, normal.coef_, rtol=1e-10, atol=1e-12)
    assert normal.solve_result_.cond_estimate is None


def test_auto_solver_refines_an_ill_conditioned_fit_to_the_least_squares_answer(torch_gpu):
    """Nearly collinear quadratic features (x in [1, 1.02]^2, ridge 1e-15):
    the Gram's rcond is ~1e-15, below the 1e-10 threshold, where the
    reference switches to QR (art/models.py:298-357).  The refined answer
    matches SciPy's gelsy least squares on the augmented design (the
    reference's fallback) far better than the plain Cholesky answer does."""
    from scipy import linalg

    from art.models import PolynomialRidgeModel

    rng = np.random.default_rng(8)
    X = rng.uniform(1.0, 1.02, (600, 2))
    y = np.exp(X[:, 0] - X[:, 1]) + 0.1 * X[:, 0] ** 3
    ridge = 1e-15
    normal = PolynomialRidgeModel(2, ridge=ridge).fit(X, y)
    auto = PolynomialRidgeModel(2, ridge=ridge, solver="auto").fit(X, y)
    r = auto.solve_result_
    assert r.solver_used == "cholesky_refined" and r.fallback_reason == "rcond_below_threshold"
    phi = R.features(X, 2, True, False)
    G = phi.T @ phi
    lam = ridge * np.trace(G) / G.shape[0]
    assert r.ridge_effective == pytest.approx(lam, rel=1e-12)
    aug = np.vstack([phi, np.sqrt(lam) * np.eye(phi.shape[1])])
    ref = linalg.lstsq(aug, np.concatenate([y, np.zeros(phi.shape[1])]), lapack_driver="gelsy")[0]
    err = lambda c: np.linalg.norm(phi @ c - phi @ ref) / np.linalg.norm(phi @ ref)
    assert err(auto.coef_) <= 1e-9, err(auto.coef_)
    assert err(auto.coef_) < 1e-2 * err(normal.coef_), (err(auto.coef_), err(normal.coef_))
    # The estimate is the exact 1-norm reciprocal condition of G + lambda I,
    # to the accuracy an inverse of that conditioning allows.
    assert 1.0 / r.cond_estimate ** 2 == pytest.approx(_exact_rcond(G + lam * np.eye(6)), rel=1e-2)


def test_refined_fp64_and_rcond_on_the_leaf_batch(torch_gpu):
    from art import LeafRidgeBatch
    from art.synthetic import leaf_ridge_workload

    torch = torch_gpu
    wl = leaf_ridge_workload(n_leaves=48, seed=4)
    X, y, w = dev(torch, wl.X), dev(torch, wl.y), dev(torch, wl.w)
    with LeafRidgeBatch(wl.offsets, 8) as batch:
        c64, s64, _ = (t.clone() for t in batch.fit(X, y, w, wl.ridge, precision="fp64"))
        cr, sr, _ = batch.fit(X, y, w, wl.ridge, precision="fp64_refined")
        rc = batch.rcond().cpu().numpy()
        np.testing.assert_array_equal(sr.cpu().numpy(), s64.cpu().numpy())
        cr = cr.cpu().numpy()
    from scipy import linalg

    _, _, lam = R.leaf_fit(wl.X, wl.y, wl.w, wl.offsets, wl.ridge)
    # Against a stable least-squares solve of the augmented weighted design
    # (SVD-based gelsd): the refined fit removes the normal equations' kappa^2
    # rounding, so it is far closer than the plain FP64 fit on the worst leaf.
    c64 = c64.cpu().numpy()
    e_plain, e_ref = np.zeros(wl.n_leaves), np.zeros(wl.n_leaves)
    for j in range(wl.n_leaves):
        a, b = wl.offsets[j], wl.offsets[j + 1]
        phi = R.features(wl.X[a:b], 2, True, False)
        sw = np.sqrt(np.maximum(wl.w[a:b], 1e-12))
        aug = np.vstack([phi * sw[:, None], np.sqrt(lam[j]) * np.eye(phi.shape[1])])
        ls = linalg.lstsq(aug, np.concatenate([wl.y[a:b] * sw, np.zeros(phi.shape[1])]),
                          lapack_driver="gelsd")[0]
        e_plain[j] = np.linalg.norm(c64[j] - ls) / np.linalg.norm(ls)
        e_ref[j] = np.linalg.norm(cr[j] - ls) / np.linalg.norm(ls)
    assert e_ref.max() <= max(1e-10, 0.1 * e_plain.max()), (e_ref.max(), e_plain.max())
    for j in range(0, wl.n_leaves, 6):
        a, b = wl.offsets[j], wl.offsets[j + 1]
        phi = R.features(wl.X[a:b], 2, True, False)
        ww = np.maximum(wl.w[a:b], 1e-12)
        G = phi.T @ (phi * ww[:, None]) + lam[j] * np.eye(phi.shape[1])
        assert rc[j] == pytest.approx(_exact_rcond(G), rel=1e-3), j


# --------------------------------------------------------------------------- graphs, tuning

@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_captured_leaf_fit_replays_equal_the_eager_fit(torch_gpu, precision):
    """A fit captured as one HIP graph replays to the eager fit's outputs bit
    for bit, on every replay, with eager work and synchronizations between
    the replays; refilling the bound inputs in place and replaying equals an
    eager fit of the new values."""
    from art import LeafRidgeBatch
    from art.synthetic import leaf_ridge_workload

    torch = torch_gpu
    wl = leaf_ridge_workload(n_leaves=96, seed=6)
    X, y, w = dev(torch, wl.X), dev(torch, wl.y), dev(torch, wl.w)
    other = leaf_ridge_workload(n_leaves=20, seed=7)
    with LeafRidgeBatch(wl.offsets, 8, stream=torch.cuda.Stream()) as batch, \
            LeafRidgeBatch(other.offsets, 8) as side:
        # Outputs live on the batch's stream: synchronize before reading them.
        fitted = batch.fit(X, y, w, wl.ridge, precision=precision)
        batch.synchronize()
        eager = [t.clone() for t in fitted]
        graph = batch.capture(X, y, w, wl.ridge, precision=precision)
        for rep in range(4):
            with torch.cuda.stream(batch.stream):
                batch.coefficients.fill_(float("nan"))
                batch.status.fill_(99)
            side.fit(dev(torch, other.X), dev(torch, other.y), dev(torch, other.w), 1e-3,
                     precision=precision)
            torch.cuda.synchronize()
            out = graph.replay()
            batch.synchronize()
            for got, ref in zip(out, eager):
                np.testing.assert_array_equal(got.cpu().numpy(), ref.cpu().numpy())
        y.mul_(-0.5).add_(0.25)
        torch.cuda.synchronize()
        replayed = graph.replay()
        batch.synchronize()
        out = [t.clone() for t in replayed]
        fitted = batch.fit(X, y, w, wl.ridge, precision=precision)
        batch.synchronize()
        fresh = [t.clone() for t in fitted]
        for got, ref in zip(out, fresh):
            np.testing.assert_array_equal(got.cpu().numpy(), ref.cpu().numpy())
        assert not np.array_equal(out[0].cpu().numpy(), eager[0].cpu().numpy())


def test_autotuned_chunk_rows_are_measured_once_and_fit_like_the_oracle(torch_gpu, tmp_path,
                                                                        monkeypatch):
    from art import LeafRidgeBatch, autotune
    from art.synthetic import leaf_ridge_workload

    torch = torch_gpu
    monkeypatch.setenv("ART_TUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(autotune, "_memory", {})
    wl = leaf_ridge_workload(n_leaves=64, seed=9)
    X, y, w = dev(torch, wl.X), dev(torch, wl.y), dev(torch, wl.w)
    with LeafRidgeBatch.autotuned(wl.offsets, 8, X, y, w, wl.ridge,
                                  candidates=(512, 4096)) as batch:
        chunk, timings = batch.tuning
        assert set(timings) == {512, 4096} and chunk in timings
        assert chunk == min(timings, key=timings.get)
        coef, status, _ = batch.fit(X, y, w, wl.ridge)
        ref, st_ref, _ = R.leaf_fit(wl.X, wl.y, wl.w, wl.offsets, wl.ridge)
        np.testing.assert_array_equal(status.cpu().numpy(), st_ref)
        assert _rel_err(coef.cpu().numpy(), ref).max() < 1e-7
    with LeafRidgeBatch.autotuned(wl.offsets, 8, X, y, w, wl.ridge,
                                  candidates=(512, 4096)) as again:
        assert again.tuning == (chunk, None)
