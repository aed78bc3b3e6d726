"""Pin the FP64 ridge/feature oracle before it is trusted as the GPU checker.

Each check is independent of the oracle's own formula: a library routine
(sklearn, SciPy/LAPACK), a literal known-answer vector transcribed from the
reference's tests, or a closed form.
"""

import numpy as np
import pytest
from scipy import linalg
from sklearn.preprocessing import PolynomialFeatures

from oracle import ridge as R


@pytest.mark.parametrize("K,n,d,degree,bias,last", [
    (2, 7, 3, 1, True, True), (3, 5, 2, 1, True, False), (2, 5, 3, 1, False, False),
    (2, 7, 1, 2, False, False), (20, 3, 5, 2, True, False), (3, 3, 10, 2, False, False),
    (2, 2, 20, 2, True, False), (1, 2, 40, 2, True, False), (2, 259, 2, 2, True, False),
])
def test_features_match_sklearn(K, n, d, degree, bias, last):
    # Parameter grid of tests/test_cuda_features.py:13-23.
    X = np.random.default_rng(901).normal(size=(K, n, d))
    X[:, 0] = 0.0
    expected = PolynomialFeatures(degree=degree, include_bias=bias).fit_transform(X.reshape(-1, d))
    if last:
        expected = np.concatenate((expected[:, 1:], expected[:, :1]), axis=1)
    np.testing.assert_array_equal(R.features(X, degree, bias, last).reshape(-1, expected.shape[1]),
                                  expected)


def test_quadratic_feature_known_answer():
    # Literal vector of tests/test_cuda_features.py:42-50.
    X = np.array([[[2., 3.], [5., 7.]], [[11., 13.], [17., 19.]]])
    np.testing.assert_array_equal(R.features(X), [
        [[1., 2., 3., 4., 6., 9.], [1., 5., 7., 25., 35., 49.]],
        [[1., 11., 13., 121., 143., 169.], [1., 17., 19., 289., 323., 361.]],
    ])


def test_feature_column_errors():
    with pytest.raises(ValueError, match="degree"):
        R.feature_columns(3, degree=3)
    with pytest.raises(ValueError, match="bias_last"):
        R.feature_columns(3, degree=2, bias_last=True)


@pytest.mark.parametrize("p", [1, 2, 6, 45, 91])
def test_cholesky_matches_lapack(p):
    rng = np.random.default_rng(p)
    a = rng.normal(size=(3 * p + 2, p))
    spd = a.T @ a + 1e-3 * np.eye(p)
    L, info = R.cholesky_lower(spd)
    assert info == 0
    np.testing.assert_allclose(L, linalg.cholesky(spd, lower=True), rtol=1e-10, atol=1e-12)
    b = rng.normal(size=p)
    np.testing.assert_allclose(R.cholesky_solve(L, b), linalg.solve(spd, b, assume_a="pos"),
                               rtol=1e-8, atol=1e-10)


def test_cholesky_two_by_two_known_answer():
    # art/_cuda_cholesky.py:102-106: [[4,1],[1,3]] x = [6,7] -> [1,2].
    L, info = R.cholesky_lower(np.array([[4., 1.], [1., 3.]]))
    assert info == 0
    np.testing.assert_allclose(R.cholesky_solve(L, np.array([6., 7.])), [1., 2.], rtol=1e-15)


def test_cholesky_pivot_index_matches_lapack_info():
    # dpotrf reports the order of the first non-positive leading minor.
    rng = np.random.default_rng(5)
    a = rng.normal(size=(20, 4))
    a[:, 2] = 0.0
    g = a.T @ a
    _, info = R.cholesky_lower(g)
    _, lapack_info = linalg.lapack.dpotrf(g, lower=1)
    assert info == lapack_info == 3
    _, info = R.cholesky_lower(np.zeros((3, 3)))
    assert info == 1
    _, info = R.cholesky_lower(np.array([[np.nan]]))
    assert info == 1


@pytest.mark.parametrize("ridge,floor", [(0.0, 0.0), (0.03, 1e-12), (1e-5, 0.07), (0.0, 0.75)])
def test_weighted_ridge_matches_augmented_least_squares(ridge, floor):
    # Independent method: minimize sum q (y - phi b)^2 + lam |b|^2 as an
    # augmented least-squares problem solved by LAPACK gelsd.
    rng = np.random.default_rng(603)
    phi, y, w = rng.normal(size=(34, 6)), rng.normal(size=34), rng.uniform(size=34)
    w[:4] = [0.0, 1.0, 1e-20, 1.0 - 1e-15]
    beta, lam = R.solve_weighted_ridge(phi, y, w, ridge, floor)
    q = np.maximum(w, floor)
    assert lam == pytest.approx(ridge * np.sum(q[:, None] * phi ** 2) / 6, rel=1e-14)
    A = np.vstack([np.sqrt(q)[:, None] * phi, np.sqrt(lam) * np.eye(6)])
    b = np.concatenate([np.sqrt(q) * y, np.zeros(6)])
    expected = linalg.lstsq(A, b, lapack_driver="gelsd")[0]
    np.testing.assert_allclose(beta, expected, rtol=1e-9, atol=1e-11)


def test_trace_reduction_across_all_columns():
    # tests/test_cuda_batched_cholesky.py:71-78: lambda = ridge * sum(q phi^2) / p.
    rng = np.random.default_rng(602)
    for p in (1, 259):
        phi, y = rng.normal(size=(1, 11, p)), rng.normal(size=(1, 11))
        _, _, lam = R.batched_gate_fit(phi, y, np.full((1, 11), .3), ridge=.2)
        np.testing.assert_allclose(lam[0], .2 * np.sum(phi ** 2) / p * np.array([.7, .3]),
                                   rtol=2e-15)


@pytest.mark.parametrize("value", [np.nan, np.inf, -.1, 1.1])
def test_bad_gate_status_known_answer(value):
    # tests/test_cuda_batched_cholesky.py:81-92.
    pi = np.full((3, 8), .4)
    pi[1, 2] = value
    coef, status, _ = R.batched_gate_fit(np.ones((3, 8, 1)), np.ones((3, 8)), pi)
    np.testing.assert_array_equal(status.T, [[0, -1, 0], [0, -1, 0]])
    assert np.isnan(coef[1]).all() and np.isfinite(coef[[0, 2]]).all()


def test_failure_isolation_known_answer():
    # tests/test_cuda_batched_cholesky.py:211-233.
    rng = np.random.default_rng(604)
    phi, y = rng.normal(size=(3, 20, 3)), rng.normal(size=(3, 20))
    phi[1, :, 2] = 0.
    pi = np.full((3, 20), .4)
    pi[2] = 0.
    coef, status, _ = R.batched_gate_fit(phi, y, pi, ridge=0.0, floor=0.0)
    np.testing.assert_array_equal(status, [[0, 0], [3, 3], [0, 1]])
    assert np.isnan(coef[1]).all() and np.isnan(coef[2, 1]).all()
    assert np.isfinite(coef[0]).all() and np.isfinite(coef[2, 0]).all()
    for side, q in enumerate((1. - pi[0], pi[0])):
        np.testing.assert_allclose(coef[0, side], linalg.lstsq(np.sqrt(q)[:, None] * phi[0],
                                                               np.sqrt(q) * y[0])[0], rtol=1e-10)
    pi[0, 0] = np.nan
    _, status, _ = R.batched_gate_fit(phi, y, pi, floor=0.0)
    np.testing.assert_array_equal(status[0], [-1, -1])


def test_nonfinite_and_solve_failure_codes():
    # -2: non-finite assembly (tests/test_cuda_batched_cholesky.py:120-129).
    g = np.eye(2)
    g[0, 0] = np.inf
    beta, _, st = R.solve_system(g, np.ones(2), 0.1)
    assert st == -2 and np.isnan(beta).all()
    # -3: a factor that succeeds but whose solve overflows to inf.
    beta, _, st = R.solve_system(np.eye(2) * 1e-300, np.full(2, 1e10), 0.0)
    assert st == -3 and np.isnan(beta).all()
    # ridge 0 gives lambda exactly 0 even when the trace is huge.
    _, lam, st = R.solve_system(np.eye(2) * 1e300, np.ones(2), 0.0)
    assert lam == 0.0 and st == 0


def test_leaf_fit_matches_per_leaf_solves():
    rng = np.random.default_rng(7)
    offsets = np.array([0, 64, 64 + 200, 64 + 200 + 90])
    X = rng.uniform(-1, 1, size=(offsets[-1], 3))
    y, w = rng.normal(size=offsets[-1]), rng.uniform(size=offsets[-1])
    coef, status, lam = R.leaf_fit(X, y, w, offsets, 1e-6)
    fast, fstatus, flam = R.leaf_fit_fast(X, y, w, offsets, 1e-6)
    np.testing.assert_array_equal(status, 0)
    np.testing.assert_array_equal(fstatus, 0)
    np.testing.assert_allclose(coef, fast, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(lam, flam, rtol=1e-14)
