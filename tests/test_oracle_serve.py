"""Pins of the serving-fit oracle (oracle/serve.py) on the CPU."""

import numpy as np

from oracle import ridge as R
from oracle import serve as S


def test_depth_zero_is_one_weighted_ridge_fit():
    rng = np.random.default_rng(0)
    X = rng.uniform(-1, 1, (300, 3))
    y = np.sin(X[:, 0]) + X[:, 1] * X[:, 2]
    w = rng.uniform(0.1, 1, 300)
    nodes = S.fit_tree(X, y, w, S.FitSettings(max_depth=0, ridge=1e-6))
    assert list(nodes) == [1]
    ref, _ = R.solve_weighted_ridge(R.features(X, 2, True, False), y, w, 1e-6)
    np.testing.assert_array_equal(nodes[1]["coef"], ref)


def test_two_half_planes_split_once_near_the_true_boundary():
    rng = np.random.default_rng(1)
    X = rng.uniform(-1, 1, (2000, 2))
    normal = np.array([1.0, 0.5]) / np.hypot(1.0, 0.5)
    right = X @ normal >= 0.1
    y = np.where(right, 2.0 + X[:, 0], -1.0 - X[:, 1])
    nodes = S.fit_tree(X, y, None, S.FitSettings(max_depth=1, degree=1, min_leaf=32,
                                                 temperature=0.05, max_iters=60))
    assert sorted(nodes) == [1, 2, 3]
    wv, z = nodes[1]["split"]
    sign = np.sign(wv @ normal)
    assert sign * (wv @ normal) > 0.99 and abs(sign * z - 0.1) < 0.03
    # one child lies inside one half-plane and is fitted exactly; both beat the parent
    root = S.node_fit(X, y, np.ones(len(y)), S.FitSettings(degree=1))
    assert min(nodes[2]["mse"], nodes[3]["mse"]) < 1e-9
    assert max(nodes[2]["mse"], nodes[3]["mse"]) < 0.5 * root[1] / root[2]


def test_partition_keeps_rows_once_in_order_and_routes_back():
    rng = np.random.default_rng(2)
    X = rng.uniform(-1, 1, (3000, 3))
    y = np.exp(X[:, 0]) - X[:, 1] ** 2
    nodes = S.fit_tree(X, y, None, S.FitSettings(max_depth=3, min_leaf=100))
    leaves = [h for h, n in nodes.items() if "coef" in n]
    rows = np.concatenate([nodes[h]["rows"] for h in leaves])
    np.testing.assert_array_equal(np.sort(rows), np.arange(3000))
    for h in leaves:
        r = nodes[h]["rows"]
        assert np.all(np.diff(r) > 0)
        np.testing.assert_array_equal(S.route(nodes, X[r]), np.full(len(r), h))
    right, order = S.stable_partition(X, np.array([1., 0., 0.]), 0.0)
    np.testing.assert_array_equal(order, np.r_[np.flatnonzero(X[:, 0] < 0),
                                               np.flatnonzero(X[:, 0] >= 0)])


def test_strided_sample_known_answers():
    np.testing.assert_array_equal(S.strided_sample(7, 3), [0, 2, 4])
    np.testing.assert_array_equal(S.strided_sample(5, 5), [0, 1, 2, 3, 4])
    np.testing.assert_array_equal(S.strided_sample(10, 4), [0, 2, 5, 7])


def test_refit_of_no_new_rows_keeps_the_fit():
    rng = np.random.default_rng(3)
    X = rng.uniform(-1, 1, (1500, 2))
    y = np.sin(3 * X[:, 0]) * np.cos(2 * X[:, 1])
    s = S.FitSettings(max_depth=2, min_leaf=64)
    nodes = S.fit_tree(X, y, None, s)
    out = S.refit_leaves(nodes, X, y, None, X[:0], y[:0], None, s)
    for h, beta in out.items():
        np.testing.assert_allclose(beta, nodes[h]["coef"], rtol=1e-12, atol=1e-14)


def test_start_keys_are_distinct_per_node_and_restart():
    keys = {S.start_key(7, h, r) for h in range(1, 200) for r in range(4)}
    assert len(keys) == 199 * 4
    assert S.start_key(7, 1, 0) != S.start_key(8, 1, 0)
