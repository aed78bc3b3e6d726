"""GPU parity of the device-controlled split optimizer vs the oracle.

Cases and tolerances follow tests/test_cuda_batched_optimizer.py of the
reference: theta rtol 1e-9 / atol 1e-10, soft-loss history rtol 1e-9,
accepted steps rtol 1e-13, backtrack counts and stop reasons exact, hard-refit
coefficients rtol 1e-10 / atol 1e-12 against an FP64 weighted ridge solve of
the final partition.
"""

import numpy as np
import pytest

from oracle import ridge as R
from oracle import split as O

pytestmark = pytest.mark.gpu


def problems():
    """tests/test_cuda_batched_optimizer.py:19-25 (seed 801)."""
    rng = np.random.default_rng(801)
    X = rng.uniform(-1, 1, (3, 43, 2))
    y = np.where(X[:, :, 0] > .1, 1. + X[:, :, 1], -.5 + .3 * X[:, :, 1])
    y[0] = 0.
    theta = np.tile([1., 0., 0.], (3, 1))
    return X, y, theta


def settings_of(s, T):
    """Oracle settings equal to a SoftObliqueSplitter's."""
    from art import AffineRidgeModel

    m = s.model_template
    affine = isinstance(m, AffineRidgeModel)
    return O.Settings(
        temperature=T, ridge=m.ridge, weight_floor=s.weight_floor, max_iters=s.max_iters,
        grad_atol=s.grad_atol, grad_rtol=s.grad_rtol, min_side_points=s.min_side_points,
        min_side_fraction=s.min_side_fraction, min_side_fallback=s.min_side_fallback,
        alpha0=s.alpha0, rho=s.rho, armijo_c=s.armijo_c, max_backtracks=s.max_backtracks,
        adaptive_alpha=s.adaptive_alpha, alpha_min=s.alpha_min, alpha_max=s.alpha_max,
        alpha_grow=s.alpha_grow, alpha_recovery=s.alpha_recovery,
        heavy_backtrack_threshold=s.heavy_backtrack_threshold,
        max_line_search_failures=s.max_line_search_failures,
        degree=1 if affine else m.degree, bias_last=affine)


def check_against_oracle(result, s, X, y, theta0, T):
    o = O.optimize(X, y, theta0, settings_of(s, T))
    np.testing.assert_allclose(np.r_[result.w, result.z], O.project_unit_w(o["theta"]),
                               rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(result.metadata["soft_loss_history"], o["soft_loss_history"],
                               rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(result.metadata["step_size_history"], o["step_size_history"],
                               rtol=1e-13)
    assert result.metadata["backtrack_history"] == o["backtrack_history"]
    assert result.stop_reason == o["stop_reason"]
    assert result.n_iters == len(o["step_size_history"])
    assert result.metadata["logical_line_search_candidates"] == o["logical_candidates"]
    return o


def check_hard_models(result, s, X, y):
    if result.left_model is None:
        return
    so = settings_of(s, 1.0)
    phi = O.design(X, so)
    left = X @ result.w < result.z
    for side, q in enumerate((left.astype(float), (~left).astype(float))):
        ref, _ = R.solve_weighted_ridge(phi, y, q, s.model_template.ridge, 0.0)
        fitted = result.left_model if side == 0 else result.right_model
        np.testing.assert_allclose(fitted.coef_, ref, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("degree,adaptive,H", [(1, False, 8), (1, True, 1), (2, True, 3),
                                               (2, False, 11)])
def test_nodes_match_independent_oracle_runs(torch_gpu, degree, adaptive, H):
    from art import (AffineRidgeModel, BatchedOptimizerWorkspace, PolynomialRidgeModel,
                     SoftObliqueSplitter, SplitNotFoundError)

    X, y, theta = problems()
    model = (AffineRidgeModel(1e-4, solver="normal") if degree == 1
             else PolynomialRidgeModel(ridge=1e-4, solver="normal"))
    s = SoftObliqueSplitter(model, .7, max_iters=4, alpha0=2., adaptive_alpha=adaptive,
                            min_side_points=2, backtrack_batch_size=H)
    temp = np.array([.3, .7, 1.2])
    for reorder in (False, True):
        with BatchedOptimizerWorkspace(s, X, y, temperature=temp, reorder=reorder) as ws:
            results = ws.optimize(theta)
            assert all(not isinstance(r, SplitNotFoundError) for r in results), results
            for j, result in enumerate(results):
                check_against_oracle(result, s, X[j], y[j], theta[j], temp[j])
                check_hard_models(result, s, X[j], y[j])
                md = result.metadata
                assert md["physical_line_search_candidates"] >= md["logical_line_search_candidates"]
                assert md["backtrack_batch_size"] == min(H, s.max_backtracks + 1)
            assert results[0].n_iters == 0
            assert results[0].stop_reason == "gradient_tolerance"
            parent = [O.parent_loss(X[j], y[j], settings_of(s, 1.)) for j in range(3)]
            np.testing.assert_allclose([r.parent_loss for r in results], parent,
                                       rtol=1e-9, atol=1e-13)


def test_restarts_and_numerical_failures_are_isolated(torch_gpu):
    from art import (AffineRidgeModel, BatchedOptimizerWorkspace, SoftObliqueSplitter,
                     SplitNotFoundError)

    X, y, theta = problems()
    X[1, :, 1] = 0.  # singular design without ridge
    s = SoftObliqueSplitter(AffineRidgeModel(0., solver="normal"), .7, max_iters=2,
                            n_restarts=2, min_side_points=2)
    initial = np.stack((theta, np.tile([0., 1., 0.], (3, 1))))
    with BatchedOptimizerWorkspace(s, X, y, reorder=True) as ws:
        result = ws.optimize(initial)
        assert isinstance(result[1], SplitNotFoundError)
        assert result[1].reason == "cholesky_failure"
        assert all(not isinstance(result[j], SplitNotFoundError) for j in (0, 2))
        # The kept restart is the one with the lower hard loss.
        for j in (0, 2):
            losses = []
            for r in range(2):
                o = O.optimize(X[j], y[j], initial[r, j], settings_of(s, .7))
                losses.append(O.hard_result(X[j], y[j], o["theta"], settings_of(s, .7))["loss"])
            assert result[j].loss == pytest.approx(min(losses), rel=1e-9, abs=1e-13)
        again = ws.optimize(initial)
        np.testing.assert_allclose(result[2].predict(X[2]), again[2].predict(X[2]), atol=1e-12)
    with pytest.raises(RuntimeError, match="closed"):
        ws.optimize(initial)


@pytest.mark.parametrize("adaptive", [False, True])
def test_line_search_exhaustion_preserves_theta_and_counts_retries(torch_gpu, adaptive):
    from art import AffineRidgeModel, BatchedOptimizerWorkspace, SoftObliqueSplitter

    X, y, theta = problems()
    s = SoftObliqueSplitter(AffineRidgeModel(1e-4, solver="normal"), .7, max_iters=2,
                            alpha0=1e12, alpha_max=1e12, max_backtracks=0,
                            max_line_search_failures=2, adaptive_alpha=adaptive)
    with BatchedOptimizerWorkspace(s, X, y) as ws:
        result = ws.optimize(theta)
        for j in (1, 2):
            check_against_oracle(result[j], s, X[j], y[j], theta[j], .7)
            assert result[j].n_iters == 0
            np.testing.assert_allclose(np.r_[result[j].w, result[j].z], theta[j], atol=1e-14)
            assert result[j].metadata["physical_line_search_candidates"] == (2 if adaptive else 1)


@pytest.mark.parametrize("hard_refit", [False, True])
def test_last_feasible_parameters_are_restored_before_hard_refits(torch_gpu, hard_refit):
    from art import AffineRidgeModel, BatchedOptimizerWorkspace, SoftObliqueSplitter

    data = []
    for seed in (2, 3):
        rng = np.random.default_rng(seed)
        x = rng.uniform(-1, 1, (60, 2))
        data.append((x, (x[:, 0] > .55).astype(float)))
    X, y = np.stack([v[0] for v in data]), np.stack([v[1] for v in data])
    theta = np.tile([1., 0., 0.], (2, 1))
    s = SoftObliqueSplitter(AffineRidgeModel(1e-4, solver="normal"), .25, max_iters=6,
                            alpha0=5., min_side_points=28, min_side_fallback=False)
    with BatchedOptimizerWorkspace(s, X, y) as ws:
        results = ws.optimize(theta, min_side_fallback=True, hard_refit=hard_refit)
        assert not s.min_side_fallback
        so = settings_of(s, .25)
        so.min_side_fallback = True
        for j, result in enumerate(results):
            o = O.optimize(X[j], y[j], theta[j], so)
            assert o["min_side_fallback"]
            np.testing.assert_allclose(np.r_[result.w, result.z], O.project_unit_w(o["theta"]),
                                       rtol=1e-8, atol=1e-10)
            assert result.metadata["fallback_iteration"] == o["fallback_iteration"]
            assert result.metadata["min_side_fallback"]
            assert min(result.n_left, result.n_right) >= 28
            assert result.metadata["hard_refit"] is hard_refit
            assert (result.left_model is None) == (not hard_refit)


def test_failed_attempts_can_retry_and_then_accept_independently(torch_gpu):
    from art import (AffineRidgeModel, BatchedOptimizerWorkspace, SoftObliqueSplitter,
                     SplitNotFoundError)

    X, y, theta = problems()
    s = SoftObliqueSplitter(AffineRidgeModel(1e-4, solver="normal"), .7, max_iters=2,
                            alpha0=1000., armijo_c=.8, max_backtracks=0,
                            max_line_search_failures=20)
    with BatchedOptimizerWorkspace(s, X, y) as ws:
        results = ws.optimize(theta)
        for j in (1, 2):
            assert not isinstance(results[j], SplitNotFoundError), results[j]
            check_against_oracle(results[j], s, X[j], y[j], theta[j], .7)
            assert results[j].n_iters == 2
            assert results[j].metadata["physical_line_search_candidates"] > 2


def test_single_slot_and_cached_parent_loss(torch_gpu):
    from art import AffineRidgeModel, BatchedOptimizerWorkspace, SoftObliqueSplitter

    X, y, theta = problems()
    s = SoftObliqueSplitter(AffineRidgeModel(1e-4, solver="normal"), 1., max_iters=0)
    with BatchedOptimizerWorkspace(s, X[2:], y[2:]) as ws:
        result = ws.optimize(theta[2:], parent_loss=np.array([.125]))[0]
        assert ws.fit_sizes == [("soft", 1), ("hard", 1)]
        assert result.parent_loss == .125
        assert result.split_gain == pytest.approx(.125 - result.loss)
        assert result.n_iters == 0 and result.stop_reason == "max_iters"
        np.testing.assert_allclose(np.r_[result.w, result.z], theta[2])
        h = O.hard_result(X[2], y[2], theta[2], settings_of(s, 1.))
        assert result.loss == pytest.approx(h["loss"], rel=1e-10)
        assert (result.n_left, result.n_right) == (h["n_left"], h["n_right"])


@pytest.mark.parametrize("degree", [1, 2])
def test_partition_export_skips_singular_hard_refit(torch_gpu, degree):
    from art import (AffineRidgeModel, BatchedOptimizerWorkspace, PolynomialRidgeModel,
                     SoftObliqueSplitter, SplitNotFoundError)

    rng = np.random.default_rng(912)
    X = rng.uniform(-.5, .5, (2, 20, 2))
    X[:, -1] = [2., 3.]  # a one-point right partition has a singular hard Gram
    y = np.sin(X[:, :, 0]) + X[:, :, 1] ** 2
    model = AffineRidgeModel(ridge=0.) if degree == 1 else PolynomialRidgeModel(ridge=0.)
    s = SoftObliqueSplitter(model, 1., max_iters=0)
    theta = np.tile([1., 0., 1.], (2, 1))
    parent = np.array([O.parent_loss(x, z, settings_of(s, 1.)) for x, z in zip(X, y)])
    with BatchedOptimizerWorkspace(s, X, y) as ws:
        fitted = ws.optimize(theta, parent_loss=parent)
        assert all(isinstance(r, SplitNotFoundError) and r.reason == "cholesky_failure"
                   for r in fitted)
        results = ws.optimize(theta, parent_loss=parent, hard_refit=False)
        assert ws.fit_sizes == [("soft", 2)]
        for result in results:
            assert not isinstance(result, SplitNotFoundError), result
            assert result.n_left == 19 and result.n_right == 1
            assert result.left_model is result.right_model is None
            assert result.loss is result.split_gain is result.relative_split_gain is None
            assert result.metadata["hard_refit"] is False
            np.testing.assert_array_equal(np.r_[result.w, result.z], theta[0])


@pytest.mark.parametrize("reorder", [False, True])
def test_capacity_loads_match_independent_runs(torch_gpu, reorder):
    from art import (BatchedOptimizerWorkspace, PolynomialRidgeModel, SoftObliqueSplitter,
                     SplitNotFoundError)

    X, y, theta = problems()
    s = SoftObliqueSplitter(PolynomialRidgeModel(ridge=1e-4), .7, max_iters=2,
                            min_side_points=2, alpha0=.5, max_backtracks=4)
    with BatchedOptimizerWorkspace.allocate(s, 5, X.shape[1], X.shape[2], reorder=reorder) as ws:
        with pytest.raises(RuntimeError, match="Load"):
            ws.optimize()
        previous = None
        for batch, k in enumerate((3, 1, 2)):
            take = [2, 1, 0][:k]
            bX, by = X[take] + .07 * batch, y[take] + .2 * batch
            ws.host_X[:k], ws.host_y[:k] = bX, by
            temps = np.linspace(.5, 1., k)
            ws.load(ws.host_X[:k], ws.host_y[:k], temperature=temps)
            parent = None
            if batch == 1:
                parent = np.array([O.parent_loss(a, b, settings_of(s, 1.)) for a, b in zip(bX, by)])
            results = ws.optimize(theta[take], parent_loss=parent)
            assert len(results) == k and ws.K == 5 and ws.k == k
            assert all(count <= k for _, count in ws.fit_sizes)
            if not reorder:
                assert all(count == k for _, count in ws.fit_sizes)
            assert any(phase == "parent" for phase, _ in ws.fit_sizes) == (parent is None)
            for j, result in enumerate(results):
                assert not isinstance(result, SplitNotFoundError), result
                check_against_oracle(result, s, bX[j], by[j], theta[take][j], temps[j])
            if previous is not None:
                old, predictions = previous
                np.testing.assert_array_equal(old.predict(X[0]), predictions)
            previous = results[0], results[0].predict(X[0]).copy()
    with pytest.raises(RuntimeError, match="closed"):
        ws.load(X, y)


def test_selected_retry_and_random_starts(torch_gpu):
    from art import (AffineRidgeModel, BatchedOptimizerWorkspace, SoftObliqueSplitter,
                     SplitNotFoundError)

    X, y, theta = problems()
    s = SoftObliqueSplitter(AffineRidgeModel(1e-4), .7, max_iters=1, min_side_points=2,
                            min_side_fallback=True, random_state=28)
    with BatchedOptimizerWorkspace.allocate(s, 5, X.shape[1], X.shape[2]) as ws:
        ws.load(X, y)
        first = ws.optimize(theta)
        retry = ws.optimize(theta[[2, 1]], nodes=[2, 1], min_side_fallback=False)
        assert s.min_side_fallback
        assert all(count <= 2 for _, count in ws.fit_sizes)
        for result, j in zip(retry, (2, 1)):
            assert not isinstance(result, SplitNotFoundError), result
            np.testing.assert_allclose(np.r_[result.w, result.z], np.r_[first[j].w, first[j].z],
                                       rtol=1e-9, atol=1e-10)
        assert ws.optimize(nodes=[]) == []
        for ids in ([3], [1, 1], [-1], [1.2]):
            with pytest.raises(ValueError, match="nodes"):
                ws.optimize(nodes=ids)
        fresh = ws.optimize(nodes=[2], parent_loss=np.array([first[2].parent_loss]))[0]
        assert fresh.metadata["initialization_seed"] is not None
        # A random start still satisfies the oracle's loop from that start.
        w = fresh.metadata["initial_theta"]
        check_against_oracle(fresh, s, X[2], y[2], w, .7)


def test_splitter_split_matches_oracle_for_one_node(torch_gpu):
    from art import PolynomialRidgeModel, SoftObliqueSplitter

    rng = np.random.default_rng(5)
    X = rng.uniform(-1.0, 1.0, size=(100, 2))
    y = np.where(X[:, 0] >= 0.0, 0.8 - 0.4 * X[:, 0] + 1.2 * X[:, 1],
                 -0.5 + 1.1 * X[:, 0] - 0.3 * X[:, 1])
    s = SoftObliqueSplitter(PolynomialRidgeModel(ridge=1e-6), .15, max_iters=30,
                            min_side_points=2, optimizer="cupy_batched_cholesky")
    theta0 = np.array([.9, .3, .05])
    result = s.split(X, y, theta0=theta0)
    check_against_oracle(result, s, X, y, theta0, .15)
    check_hard_models(result, s, X, y)
    assert result.split_gain > 0


@pytest.mark.parametrize("batched", [False, True])
def test_annealed_mode_matches_the_oracle(torch_gpu, batched):
    from art import AffineRidgeModel, BatchedOptimizerWorkspace, SoftObliqueSplitter

    X, y, theta = problems()
    s = SoftObliqueSplitter(AffineRidgeModel(1e-4), .7, n_restarts=2, min_side_points=2)
    starts = np.stack([theta, np.tile([.6, .8, .1], (3, 1))])
    with BatchedOptimizerWorkspace(s, X, y) as ws:
        res = ws.anneal(starts, t0=.5, t1=.05, iters=25, step=.3, batched=batched)
    so = settings_of(s, .7)
    for r in range(2):
        for j in range(3):
            th, losses = O.annealed(X[j], y[j], starts[r, j], so, iters=25, t0=.5, t1=.05,
                                    step=.3)
            np.testing.assert_allclose(res["theta"][r, j], O.project_unit_w(th), rtol=1e-9,
                                       atol=1e-10)
            np.testing.assert_allclose(res["soft_loss_history"][r, j], losses, rtol=1e-9,
                                       atol=1e-12)
            try:
                h = O.hard_result(X[j], y[j], th, so)
            except ValueError:
                assert not res["ok"][r, j]
                continue
            assert res["ok"][r, j]
            assert res["loss"][r, j] == pytest.approx(h["loss"], rel=1e-9, abs=1e-14)
            np.testing.assert_allclose(res["beta"][r, j], h["beta"], rtol=1e-9, atol=1e-12)
    assert res["best"].shape == (3,)


@pytest.mark.parametrize("degree", [1, 2])
def test_batched_annealer_matches_the_per_start_kernel(torch_gpu, degree):
    # 20 starts per node (two blocks of 16, the second padded), d = 5
    # affine or d = 3 quadratic, uneven row count (300).
    from art import (AffineRidgeModel, BatchedOptimizerWorkspace, PolynomialRidgeModel,
                     SoftObliqueSplitter)

    rng = np.random.default_rng(71)
    d = 5 if degree == 1 else 3
    X = rng.uniform(-1, 1, (3, 300, d))
    y = np.sin(2 * X[..., 0]) + X[..., 1] * X[..., -1]
    model = AffineRidgeModel(1e-4) if degree == 1 else PolynomialRidgeModel(ridge=1e-4)
    s = SoftObliqueSplitter(model, .3, n_restarts=20, min_side_points=3, random_state=5)
    with BatchedOptimizerWorkspace(s, X, y) as ws:
        starts = ws._device_starts(np.tile(np.arange(3), 20)).reshape(20, 3, d + 1)
        one = ws.anneal(starts, t0=.3, t1=.03, iters=30, step=.2, batched=False)
        two = ws.anneal(starts, t0=.3, t1=.03, iters=30, step=.2, batched=True)
    np.testing.assert_array_equal(one["ok"], two["ok"])
    np.testing.assert_allclose(two["theta"], one["theta"], rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(two["soft_loss_history"], one["soft_loss_history"], rtol=1e-9,
                               atol=1e-12)
    ok = one["ok"]
    np.testing.assert_allclose(two["loss"][ok], one["loss"][ok], rtol=1e-9, atol=1e-13)
    np.testing.assert_allclose(two["beta"][ok], one["beta"][ok], rtol=1e-8, atol=1e-10)
    np.testing.assert_array_equal(two["n_right"], one["n_right"])
    so = settings_of(s, .3)
    th, losses = O.annealed(X[1], y[1], starts[7, 1], so, iters=30, t0=.3, t1=.03, step=.2)
    np.testing.assert_allclose(two["theta"][7, 1], O.project_unit_w(th), rtol=1e-9, atol=1e-10)


def test_batched_annealer_at_full_width_matches_the_oracle(torch_gpu):
    # d = 16 affine (p = 17: [phi, y] fills all 171 packed products), 16
    # starts of each of 2 nodes in one workgroup each, 700 rows (a partial
    # last chunk).
    from art import AffineRidgeModel, BatchedOptimizerWorkspace, SoftObliqueSplitter

    rng = np.random.default_rng(91)
    d = 16
    X = rng.uniform(-1, 1, (2, 700, d))
    y = np.exp(-2.0 * ((X[..., :3] - .2) ** 2).sum(-1)) - .5 * np.exp(
        -3.0 * ((X[..., 3:6] + .3) ** 2).sum(-1))
    s = SoftObliqueSplitter(AffineRidgeModel(1e-4), .2, n_restarts=16, min_side_points=3,
                            random_state=9)
    with BatchedOptimizerWorkspace(s, X, y) as ws:
        starts = ws._device_starts(np.tile(np.arange(2), 16)).reshape(16, 2, d + 1)
        one = ws.anneal(starts, t0=.2, t1=.02, iters=12, step=.05, batched=False)
        two = ws.anneal(starts, t0=.2, t1=.02, iters=12, step=.05, batched=True)
    np.testing.assert_array_equal(one["ok"], two["ok"])
    np.testing.assert_allclose(two["theta"], one["theta"], rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(two["soft_loss_history"], one["soft_loss_history"], rtol=1e-9,
                               atol=1e-12)
    np.testing.assert_array_equal(two["n_right"], one["n_right"])
    so = settings_of(s, .2)
    for r, j in ((0, 0), (15, 1), (6, 1)):
        th, losses = O.annealed(X[j], y[j], starts[r, j], so, iters=12, t0=.2, t1=.02,
                                step=.05)
        np.testing.assert_allclose(two["theta"][r, j], O.project_unit_w(th), rtol=1e-9,
                                   atol=1e-10)
        np.testing.assert_allclose(two["soft_loss_history"][r, j], losses, rtol=1e-9,
                                   atol=1e-12)
        h = O.hard_result(X[j], y[j], th, so)
        assert two["loss"][r, j] == pytest.approx(h["loss"], rel=1e-9, abs=1e-14)
        np.testing.assert_allclose(two["beta"][r, j], h["beta"], rtol=1e-8, atol=1e-10)


def test_batched_annealer_on_the_split_workload_matches_the_oracle(torch_gpu):
    """The split-optimizer configuration's own shape and schedule: 2 nodes of
    8192 rows in d = 16 with an 8-bump target (seed 3), 16 device starts each,
    200 iterations cooling geometrically from T = 0.2 to 0.01 at step 0.05,
    affine children at ridge 1e-4.  The batched kernel equals the per-start
    kernel, and three starts equal the FP64 oracle run from the same start
    over the whole schedule (saturated gates and the +-50 clip included)."""
    from art import AffineRidgeModel, BatchedOptimizerWorkspace, SoftObliqueSplitter
    from art.synthetic import split_workload

    X, y = split_workload(n_nodes=2, seed=3)
    d = X.shape[2]
    s = SoftObliqueSplitter(AffineRidgeModel(1e-4), .2, n_restarts=16, min_side_points=64,
                            random_state=3)
    with BatchedOptimizerWorkspace(s, X, y) as ws:
        starts = ws._device_starts(np.tile(np.arange(2), 16)).reshape(16, 2, d + 1)
        one = ws.anneal(starts, t0=.2, t1=.01, iters=200, step=.05, batched=False)
        two = ws.anneal(starts, t0=.2, t1=.01, iters=200, step=.05, batched=True)
        mix = ws.anneal(starts, t0=.2, t1=.01, iters=200, step=.05, batched=True,
                        child_precision="mixed")
    np.testing.assert_array_equal(two["ok"], one["ok"])
    assert two["ok"].mean() > .5
    # Split-bf16 child systems (Gram entries to ~1e-9): over the whole
    # schedule theta stays within 1e-10 of the FP64 kernel (measured 2e-14),
    # the soft losses within 1e-9 (measured 3e-12), and the hard partitions,
    # hard losses and refit coefficients (FP64 refit) are the same.
    np.testing.assert_array_equal(mix["ok"], two["ok"])
    np.testing.assert_allclose(mix["theta"], two["theta"], rtol=0, atol=1e-10)
    np.testing.assert_allclose(mix["soft_loss_history"], two["soft_loss_history"], rtol=1e-9)
    np.testing.assert_array_equal(mix["n_right"], two["n_right"])
    ok = two["ok"]
    np.testing.assert_allclose(mix["loss"][ok], two["loss"][ok], rtol=1e-12)
    np.testing.assert_allclose(mix["beta"][ok], two["beta"][ok], rtol=1e-9, atol=1e-12)
    np.testing.assert_array_equal(mix["best"], two["best"])
    np.testing.assert_allclose(two["theta"], one["theta"], rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(two["soft_loss_history"], one["soft_loss_history"], rtol=1e-9,
                               atol=1e-12)
    np.testing.assert_array_equal(two["n_right"], one["n_right"])
    so = settings_of(s, .2)
    for r, j in ((0, 0), (9, 1), (15, 0)):
        th, losses = O.annealed(X[j], y[j], starts[r, j], so, iters=200, t0=.2, t1=.01,
                                step=.05)
        np.testing.assert_allclose(two["theta"][r, j], O.project_unit_w(th), rtol=1e-9,
                                   atol=1e-10)
        np.testing.assert_allclose(two["soft_loss_history"][r, j], losses, rtol=1e-9,
                                   atol=1e-12)
        try:
            h = O.hard_result(X[j], y[j], th, so)
        except ValueError:  # a side below the minimum: the device flags it too
            assert not two["ok"][r, j]
            continue
        assert two["ok"][r, j] and two["n_right"][r, j] == h["n_right"]
        assert two["loss"][r, j] == pytest.approx(h["loss"], rel=1e-9, abs=1e-14)
        np.testing.assert_allclose(mix["theta"][r, j], O.project_unit_w(th), rtol=1e-9,
                                   atol=1e-10)


@pytest.mark.parametrize("n,d,ties", [(8192, 16, False), (301, 3, False), (1, 2, False),
                                       (64, 31, False), (8193, 3, False), (10000, 2, False),
                                       (2, 1, False), (400, 3, True), (9000, 2, True)])
def test_device_random_starts_match_the_oracle(torch_gpu, n, d, ties):
    # Nodes of more than 8192 rows take the kernel's recompute path; tied
    # scores (rows from a 3-point grid) exercise the repeated middle key.
    from art import AffineRidgeModel, BatchedOptimizerWorkspace, SoftObliqueSplitter

    rng = np.random.default_rng(n + d)
    X = (rng.integers(0, 3, (2, n, d)).astype(float) if ties
         else rng.uniform(-1, 1, (2, n, d)))
    y = X[..., 0] ** 2
    s = SoftObliqueSplitter(AffineRidgeModel(1e-4), .5, random_state=11)
    with BatchedOptimizerWorkspace(s, X, y) as ws:
        nodes = np.array([0, 1, 1, 0, 1])
        first = ws._device_starts(nodes)
        second = ws._device_starts(nodes)
        for q, j in enumerate(nodes):
            ref = O.device_start(X[j], ws.seed, q, 0)
            np.testing.assert_allclose(first[q], ref, rtol=1e-12, atol=1e-13)
            np.testing.assert_allclose(second[q], O.device_start(X[j], ws.seed, q, 1),
                                       rtol=1e-12, atol=1e-13)
        assert not np.allclose(first, second)
        np.testing.assert_allclose(np.linalg.norm(first[:, :d], axis=1), 1.0, rtol=1e-14)
        many = ws._device_starts(np.zeros(4000, dtype=int))
    # Isotropy: unit directions have E[w_j^2] = 1/d.
    np.testing.assert_allclose((many[:, :d] ** 2).mean(axis=0), 1.0 / d, atol=6 / np.sqrt(4000) / d * 3)
    # The offset is a median: at most half the rows on either side.
    u = X[0] @ many[0, :d]
    assert np.count_nonzero(u < many[0, d]) <= n // 2 and np.count_nonzero(u > many[0, d]) <= n // 2


def test_random_start_metadata_reports_device_initialization(torch_gpu):
    from art import AffineRidgeModel, BatchedOptimizerWorkspace, SoftObliqueSplitter

    X, y, _ = problems()
    s = SoftObliqueSplitter(AffineRidgeModel(1e-4), .7, max_iters=3, min_side_points=2,
                            random_state=28)
    with BatchedOptimizerWorkspace(s, X, y) as ws:
        out = ws.optimize()
    for j, r in enumerate(out):
        assert r.metadata["initialization"] == "gpu"
        assert r.metadata["initialization_seed"] == ws.seed
        start = r.metadata["initial_theta"]
        np.testing.assert_allclose(start, O.device_start(X[j], ws.seed, j, 0), rtol=1e-12,
                                   atol=1e-13)
        check_against_oracle(r, s, X[j], y[j], start, .7)


def test_ragged_nodes_match_the_oracle_node_by_node(torch_gpu):
    # Three nodes of 150, 301 and 77 rows in one launch; each result equals the
    # oracle's run on that node alone (its own min-side requirement too).
    from art import AffineRidgeModel, SoftObliqueSplitter
    from art.batched_optimizer import RaggedOptimizerWorkspace

    rng = np.random.default_rng(88)
    sizes = [150, 301, 77]
    Xs = [rng.uniform(-1, 1, (n, 2)) for n in sizes]
    ys = [np.where(X[:, 0] + .3 * X[:, 1] > .05, 1. + X[:, 1], -.5 + .4 * X[:, 0]) for X in Xs]
    offsets = np.concatenate([[0], np.cumsum(sizes)])
    s = SoftObliqueSplitter(AffineRidgeModel(1e-6), .3, max_iters=25, min_side_points=5,
                            min_side_fraction=.1)
    theta0 = np.array([[.9, .3, 0.], [.2, 1., .1], [1., -.2, .05]])
    with RaggedOptimizerWorkspace(s, np.vstack(Xs), np.concatenate(ys), offsets) as ws:
        out = ws.optimize(theta0)
        starts = ws._device_starts([2, 0, 1])
    for j in range(3):
        r = out[j]
        check_against_oracle(r, s, Xs[j], ys[j], theta0[j], .3)
        assert r.n_left + r.n_right == sizes[j]
        assert r.metadata["required_side_points"] == max(5, int(np.ceil(.1 * sizes[j])))
    for q, j in enumerate([2, 0, 1]):
        np.testing.assert_allclose(starts[q], O.device_start(Xs[j], ws.seed, q, 0), rtol=1e-12,
                                   atol=1e-13)


def test_ragged_layout_of_equal_nodes_equals_the_block_layout(torch_gpu):
    from art import AffineRidgeModel, BatchedOptimizerWorkspace, SoftObliqueSplitter, SplitResult
    from art.batched_optimizer import RaggedOptimizerWorkspace

    X, y, theta = problems()
    s = SoftObliqueSplitter(AffineRidgeModel(1e-4), .7, max_iters=20, min_side_points=2)
    with BatchedOptimizerWorkspace(s, X, y) as ws:
        a = ws.optimize(theta)
    offsets = np.arange(4) * X.shape[1]
    with RaggedOptimizerWorkspace(s, X.reshape(-1, 2), y.reshape(-1), offsets) as ws:
        b = ws.optimize(theta)
    for u, v in zip(a, b):
        assert type(u) is type(v)
        if isinstance(u, SplitResult):
            np.testing.assert_array_equal(np.r_[u.w, u.z], np.r_[v.w, v.z])
            assert u.loss == v.loss and u.n_right == v.n_right


def test_keyed_starts_do_not_depend_on_position(torch_gpu):
    from art import AffineRidgeModel, SoftObliqueSplitter
    from art.batched_optimizer import RaggedOptimizerWorkspace

    rng = np.random.default_rng(89)
    sizes = [150, 3000, 77, 9000]
    X = rng.uniform(-1, 1, (sum(sizes), 3))
    offsets = np.concatenate([[0], np.cumsum(sizes)])
    s = SoftObliqueSplitter(AffineRidgeModel(1e-6), .3, max_iters=5)
    keys = np.array([11, 2**63 + 5, 7, 12345678901234], dtype=np.uint64)
    with RaggedOptimizerWorkspace(s, X, X[:, 0].copy(), offsets) as ws:
        a = ws._device_starts([0, 1, 2, 3], keys=keys)
        b = ws._device_starts([3, 1, 0], keys=keys[[3, 1, 0]])
        c = ws._device_starts([0, 0], keys=keys[[0, 2]])
    np.testing.assert_array_equal(b, a[[3, 1, 0]])
    assert not np.array_equal(c[0], c[1])
    np.testing.assert_array_equal(c[0], a[0])
    for j in range(4):
        np.testing.assert_allclose(a[j], O.device_start(X[offsets[j]:offsets[j + 1]], 0, 0, 0,
                                                        key=int(keys[j])),
                                   rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("degree,d", [(1, 3), (2, 3), (2, 5)])
def test_node_rows_spread_over_parts_match_one_workgroup(torch_gpu, degree, d):
    # Large nodes searched with each node's rows spread over several
    # workgroups (art_split_optimize_parts) agree with one workgroup per node
    # and with the oracle: the parts exchange every pass's row sums, so only
    # the order of the partial sums differs.  d = 5 quadratic (p = 21) runs
    # the wide kernel and the block Cholesky.
    from art import AffineRidgeModel, PolynomialRidgeModel, SoftObliqueSplitter
    from art.batched_optimizer import RaggedOptimizerWorkspace

    rng = np.random.default_rng(90 + 10 * degree + d)
    sizes = [23000, 9000, 2500, 17001]
    X = rng.uniform(-1, 1, (sum(sizes), d))
    u = X[:, 0] + .4 * X[:, 1] - .2 * X[:, 2]
    y = np.where(u > .1, 1. + X[:, 1], -.5 + .3 * X[:, 0]) + .05 * rng.standard_normal(len(u))
    offsets = np.concatenate([[0], np.cumsum(sizes)])
    model = AffineRidgeModel(1e-6) if degree == 1 else PolynomialRidgeModel(ridge=1e-6)
    s = SoftObliqueSplitter(model, .3, max_iters=12, min_side_points=5, min_side_fraction=.05)
    theta0 = np.tile(np.r_[1., np.zeros(d - 1), 0.], (len(sizes), 1))
    theta0[:, 1] = [.3, -.2, .5, .1]
    runs = {}
    with RaggedOptimizerWorkspace(s, X, y, offsets) as ws:
        for parts in (1, 4, None):
            ws.split_parts = parts
            runs[parts] = (ws.optimize(theta0), ws.last_launch_parts)
        assert ws.part_retries == 0
    assert runs[1][1] == 1 and runs[4][1] == 4
    assert runs[None][1] > 1  # the largest node asks for MAX_PARTS
    for parts in (4, None):
        for a, b in zip(runs[1][0], runs[parts][0]):
            np.testing.assert_allclose(np.r_[b.w, b.z], np.r_[a.w, a.z], rtol=1e-9, atol=1e-10)
            np.testing.assert_allclose(b.metadata["soft_loss_history"],
                                       a.metadata["soft_loss_history"], rtol=1e-10)
            assert b.stop_reason == a.stop_reason and b.n_iters == a.n_iters
            assert b.n_right == a.n_right
            np.testing.assert_allclose(b.loss, a.loss, rtol=1e-10)
    for j in (0, 3):
        lo, hi = offsets[j], offsets[j + 1]
        check_against_oracle(runs[4][0][j], s, X[lo:hi], y[lo:hi], theta0[j], .3)


def test_annealed_mode_over_parts_matches_one_workgroup_and_the_oracle(torch_gpu):
    # The ragged workspace anneals one start per problem on the optimizer
    # kernel (mode 1); with each node's rows over 4 workgroups the annealed
    # trajectories, hard refits and losses agree with one workgroup per node
    # and with the oracle.
    from art import AffineRidgeModel, SoftObliqueSplitter
    from art.batched_optimizer import RaggedOptimizerWorkspace

    rng = np.random.default_rng(131)
    sizes = [6000, 3100]
    X = rng.uniform(-1, 1, (sum(sizes), 3))
    y = np.where(X[:, 0] - .3 * X[:, 2] > 0, 1. + X[:, 1], -.4 * X[:, 0]) \
        + .05 * rng.standard_normal(len(X))
    offsets = np.concatenate([[0], np.cumsum(sizes)])
    s = SoftObliqueSplitter(AffineRidgeModel(1e-4), .5, n_restarts=1, min_side_points=5)
    starts = np.array([[[.8, .6, 0., .05], [.6, 0., .8, -.1]]])
    runs = {}
    with RaggedOptimizerWorkspace(s, X, y, offsets) as ws:
        for parts in (1, 4):
            ws.split_parts = parts
            runs[parts] = ws.anneal(starts, t0=.5, t1=.05, iters=20, step=.2)
            assert ws.last_launch_parts == parts
        assert ws.part_retries == 0
    a, b = runs[1], runs[4]
    np.testing.assert_allclose(b["theta"], a["theta"], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(b["soft_loss_history"], a["soft_loss_history"], rtol=1e-10)
    np.testing.assert_array_equal(b["ok"], a["ok"])
    np.testing.assert_allclose(b["loss"], a["loss"], rtol=1e-10)
    so = settings_of(s, .5)
    for j in range(2):
        lo, hi = offsets[j], offsets[j + 1]
        th, losses = O.annealed(X[lo:hi], y[lo:hi], starts[0, j], so, iters=20, t0=.5, t1=.05,
                                step=.2)
        np.testing.assert_allclose(b["theta"][0, j], O.project_unit_w(th), rtol=1e-9, atol=1e-10)
        np.testing.assert_allclose(b["soft_loss_history"][0, j], losses, rtol=1e-9, atol=1e-12)
