"""art_leaf_ridge_fit_static (allocation-free FP64 leaf fits) against the
plan-based art_leaf_ridge_fit, through the C ABI.

The static entry point runs the plan's kernels on the plan's default
chunking, so its results must equal the plan fit's bit for bit; two calls in
a row (the second smaller) check that the shared static scratch carries no
state from one call into the next.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(rng, d, sizes, bad_leaf=None):
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    N = int(offsets[-1])
    X = rng.uniform(-1, 1, size=(N, d))
    y = np.sin(2 * X[:, 0]) + X[:, -1] ** 2 + 0.01 * rng.normal(size=N)
    w = rng.uniform(0.05, 1.0, size=N)
    if bad_leaf is not None:
        w[offsets[bad_leaf] + 1] = np.nan
    return offsets, X, y, w


def _static_fit(torch, offsets, d, degree, bias, last, X, y, w, ridge):
    from art import _hip

    lib = _hip.load()
    p = lib.art_feature_count(d, degree, int(bias), int(last))
    K = len(offsets) - 1
    f64 = dict(dtype=torch.float64, device="cuda")
    beta = torch.full((K, p), 7.0, **f64)
    status = torch.full((K,), 99, dtype=torch.int32, device="cuda")
    lam = torch.full((K,), -1.0, **f64)
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    _hip.check(lib.art_leaf_ridge_fit_static(
        off.ctypes.data, K, _hip.Basis(d, p, degree, int(bias), int(last)), _hip.ptr(X),
        _hip.ptr(y), _hip.ptr(w), float(ridge), 1e-12, _hip.ptr(beta), _hip.ptr(status),
        _hip.ptr(lam), _hip.stream_handle(torch.cuda.current_stream())))
    torch.cuda.synchronize()
    return beta.cpu().numpy(), status.cpu().numpy(), lam.cpu().numpy()


def _plan_fit(torch, offsets, d, degree, bias, last, X, y, w, ridge):
    from art import LeafRidgeBatch

    with LeafRidgeBatch(offsets, d, degree=degree, include_bias=bias, bias_last=last) as batch:
        c, s, l = batch.fit(X, y, w, ridge, precision="fp64")
        torch.cuda.synchronize()
        return c.cpu().numpy(), s.cpu().numpy(), l.cpu().numpy()


def test_static_fit_equals_the_plan_fit_on_two_calls_in_a_row(torch_gpu):
    torch = torch_gpu
    rng = np.random.default_rng(2024)
    cases = [
        # d = 8 quadratic (p = 45): leaves of 1..9000 rows (three chunks for
        # the largest), an empty leaf and a leaf with a NaN weight (-1).
        (8, 2, True, False, np.array([300, 1, 0, 9000, 64, 4097, 2500, 17] * 5), 3),
        # smaller second call: d = 3 quadratic without a bias, then affine.
        (3, 2, False, False, np.array([200, 50, 3000, 7]), None),
    ]
    for d, degree, bias, last, sizes, bad in cases:
        offsets, X, y, w = _problem(rng, d, sizes, bad)
        dev = lambda a: torch.as_tensor(a, dtype=torch.float64, device="cuda")
        Xd, yd, wd = dev(X), dev(y), dev(w)
        got = _static_fit(torch, offsets, d, degree, bias, last, Xd, yd, wd, 1e-6)
        ref = _plan_fit(torch, offsets, d, degree, bias, last, Xd, yd, wd, 1e-6)
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b)
        if bad is not None:
            assert got[1][bad] == -1 and np.isnan(got[0][bad]).all()
        solvable = np.asarray(sizes) > 60
        if bad is not None:
            solvable[bad] = False
        assert (got[1][solvable] == 0).all()


def test_static_fit_rejects_problems_beyond_its_scratch(torch_gpu):
    torch = torch_gpu
    # 2000 leaves of 4097 rows: 4000 chunks of 12 KB partials > 16 MiB.  The
    # size check runs on the host before any launch (the data is never read).
    offsets = np.arange(2001, dtype=np.int64) * 4097
    X = torch.zeros((1, 8), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="static scratch"):
        _static_fit(torch, offsets, 8, 2, True, False, X, X[:, 0], X[:, 0], 1e-6)
