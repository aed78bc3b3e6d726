"""Pins of the split-optimizer oracle (oracle/split.py), CPU only.

Each check is independent of the oracle's own formulas: finite differences,
closed forms (least squares), known answers taken from the reference's tests
(tests/test_cuda_optimizer_kernels.py:208-214,
tests/test_cuda_batched_optimizer_kernels.py:148-170,
tests/test_cuda_batched_optimizer.py:253-300) and brute force.
"""

import numpy as np
import pytest

from oracle import ridge as R
from oracle import split as O


def problems():
    """The reference's three-node fixture (tests/test_cuda_batched_optimizer.py:19-25)."""
    rng = np.random.default_rng(801)
    X = rng.uniform(-1, 1, (3, 43, 2))
    y = np.where(X[:, :, 0] > .1, 1. + X[:, :, 1], -.5 + .3 * X[:, :, 1])
    y[0] = 0.
    theta = np.tile([1., 0., 0.], (3, 1))
    return X, y, theta


def _fd(f, theta, h=1e-6):
    g = np.empty_like(theta)
    for k in range(theta.size):
        e = np.zeros_like(theta)
        e[k] = h
        g[k] = (f(theta + e) - f(theta - e)) / (2 * h)
    return g


@pytest.mark.parametrize("degree", [1, 2])
def test_frozen_gradient_matches_central_differences(degree):
    rng = np.random.default_rng(11)
    X = rng.uniform(-1, 1, (80, 3))
    y = np.sin(2 * X[:, 0]) + X[:, 1] * X[:, 2]
    s = O.Settings(ridge=1e-3, degree=degree, bias_last=degree == 1)
    phi = O.design(X, s)
    theta = np.array([.6, -.3, .5, .1])
    T = .4
    cur = O.value_and_grad(X, y, phi, theta, T, s)
    frozen = lambda t: O.loss_of(O.evaluate(X, y, phi, t, T, s, reference=cur))
    np.testing.assert_allclose(cur.grad, _fd(frozen, theta), rtol=1e-6, atol=1e-9)


def test_refitted_gradient_obeys_the_envelope_theorem_without_ridge():
    # With ridge 0 and no floor, the child fits minimise the weighted losses,
    # so the derivative of the refitted loss equals the frozen-model gradient.
    rng = np.random.default_rng(12)
    X = rng.uniform(-1, 1, (90, 2))
    y = np.where(X[:, 0] > .2, 1. + X[:, 1], -X[:, 1])
    s = O.Settings(ridge=0.0, weight_floor=0.0, degree=1)
    phi = O.design(X, s)
    theta = np.array([.8, .6, .1])
    T = .3
    cur = O.value_and_grad(X, y, phi, theta, T, s)
    full = lambda t: O.value_and_grad(X, y, phi, t, T, s).loss
    np.testing.assert_allclose(cur.grad, _fd(full, theta), rtol=1e-5, atol=1e-8)


def test_gradient_term_signs_fail_a_planted_mistake():
    # A sign error in the offset derivative is caught by the same check.
    rng = np.random.default_rng(13)
    X = rng.uniform(-1, 1, (50, 2))
    y = X[:, 0] ** 2
    s = O.Settings(ridge=1e-4)
    phi = O.design(X, s)
    theta = np.array([.3, .9, .05])
    cur = O.value_and_grad(X, y, phi, theta, .5, s)
    frozen = lambda t: O.loss_of(O.evaluate(X, y, phi, t, .5, s, reference=cur))
    wrong = cur.grad.copy()
    wrong[-1] = -wrong[-1]
    assert not np.allclose(wrong, _fd(frozen, theta), rtol=1e-6, atol=1e-9)


def test_tangent_projection_identities():
    rng = np.random.default_rng(14)
    for _ in range(20):
        theta = rng.normal(size=5)
        g = rng.normal(size=5)
        t = O.project_unit_w(theta)
        assert abs(np.linalg.norm(t[:-1]) - 1) < 1e-15
        assert t[-1] == theta[-1]
        pg = O.project_unit_w_gradient(theta, g)
        assert abs(pg[:-1] @ t[:-1]) < 1e-14
        assert pg[-1] == g[-1]
        np.testing.assert_allclose(O.project_unit_w_gradient(theta, pg), pg, atol=1e-15)
    with pytest.raises(ValueError):
        O.project_unit_w([0., 0., 1.])


def test_sigmoid_clips_scores_at_fifty():
    assert O.sigmoid(np.array([1e6]))[0] == O.sigmoid(np.array([50.]))[0]
    assert O.sigmoid(np.array([-1e6]))[0] == O.sigmoid(np.array([-50.]))[0] > 0
    assert O.sigmoid(np.array([0.]))[0] == .5


@pytest.mark.parametrize("losses,selected", [([9., 8., 1.], 0), ([np.inf, 8., 1.], 1),
                                             ([np.inf, np.nan, 12.], -1)])
def test_armijo_takes_the_first_passing_step_not_the_minimum(losses, selected):
    # tests/test_cuda_optimizer_kernels.py:208-214: f0 = 10, slope = -1, c = .1.
    table = dict(zip([1., .5, .25], losses))
    theta, value, step, ok, k, examined = O.armijo(lambda a: table[a], lambda a: a, 10., -1.,
                                                   1., .5, .1, 2)
    last = selected if selected >= 0 else 2
    assert ok == (selected >= 0)
    assert (value == losses[last]) or (np.isnan(value) and np.isnan(losses[last]))
    assert step == [1., .5, .25][last]
    assert k == (selected if ok else 2)
    assert examined == (selected + 1 if ok else 3)


def test_batched_selection_known_answers():
    # tests/test_cuda_batched_optimizer_kernels.py:148-170: S = 11 candidates
    # alpha_j * .5^k; the earliest passing one wins regardless of batching.
    S = 11
    for j, (loss, sel) in enumerate([({2: 8., 4: 1.}, 2), ({9: 8.}, 9), ({}, -1)]):
        alphas = (j + 1.) * .5 ** np.arange(S)
        table = {a: loss.get(k, 12.) for k, a in enumerate(alphas)}
        _, value, step, ok, k, _ = O.armijo(lambda a: table[a], lambda a: a, 10., -1.,
                                            alphas[0], .5, .1, S - 1)
        if sel >= 0:
            assert ok and k == sel and step == alphas[sel] and value == 8.
        else:
            assert not ok and value == 12. and step == alphas[S - 1]


def test_adaptive_alpha_known_sequence():
    # art/optimizers.py:49-66 with rho = .5, grow = recovery = 10, heavy = 8.
    c = O.AdaptiveAlpha(1.0, 1e-12, 1e3, 10., 10., 8)
    assert c.update(True, 1.0, 0, .5) == 10.0           # no backtrack: grow
    assert c.update(True, 1.25, 3, .5) == 2.5            # light: undo one backtrack
    assert c.update(True, 2.5 / 256, 8, .5) == 2.5 / 25.6  # heavy: recovery
    assert c.update(False, 1e-3, 25, .5) == 5e-4         # failure: last step * rho
    assert c.update(True, 500., 0, .5) == 1e3            # clipped to alpha_max
    assert c.update(False, 1e-12, 25, .5) == 1e-12       # clipped to alpha_min
    assert O.AdaptiveAlpha(1e6).alpha == 1e3


def _settings(**kw):
    return O.Settings(**{"ridge": 1e-4, "temperature": .7, **kw})


def test_exact_fit_stops_on_the_gradient_tolerance():
    X, y, theta = problems()
    r = O.optimize(X[0], y[0], theta[0], _settings(max_iters=4, alpha0=2., min_side_points=2))
    assert r["stop_reason"] == "gradient_tolerance" and r["n_iters"] == 0 and r["converged"]


@pytest.mark.parametrize("adaptive", [False, True])
def test_line_search_exhaustion_keeps_theta(adaptive):
    # tests/test_cuda_batched_optimizer.py:253-270.
    X, y, theta = problems()
    s = _settings(max_iters=2, alpha0=1e12, alpha_max=1e12, max_backtracks=0,
                  max_line_search_failures=2, adaptive_alpha=adaptive)
    for j in (1, 2):
        r = O.optimize(X[j], y[j], theta[j], s)
        assert r["stop_reason"] == ("max_line_search_failures" if adaptive else "line_search_failed")
        assert r["n_iters"] == 0 and r["step_size_history"] == []
        np.testing.assert_array_equal(r["theta"], theta[j])
        assert r["logical_candidates"] == (2 if adaptive else 1)


def test_failed_attempts_retry_and_then_accept():
    # tests/test_cuda_batched_optimizer.py:301-318: alpha0 = 1000 fails at
    # first, the controller shrinks it and the retries are accepted.
    X, y, theta = problems()
    s = _settings(max_iters=2, alpha0=1000., armijo_c=.8, max_backtracks=0,
                  max_line_search_failures=20)
    for j in (1, 2):
        r = O.optimize(X[j], y[j], theta[j], s)
        assert r["n_iters"] == 2 and r["logical_candidates"] > 2


def test_last_feasible_parameters_are_restored():
    # tests/test_cuda_batched_optimizer.py:272-299.
    for seed in (2, 3):
        rng = np.random.default_rng(seed)
        X = rng.uniform(-1, 1, (60, 2))
        y = (X[:, 0] > .55).astype(float)
        s = O.Settings(ridge=1e-4, temperature=.25, max_iters=6, alpha0=5., min_side_points=28,
                       min_side_fallback=True)
        r = O.optimize(X, y, np.array([1., 0., 0.]), s)
        assert r["min_side_fallback"]
        right = np.count_nonzero(X @ r["theta"][:-1] >= r["theta"][-1])
        assert 28 <= right <= 32
        h = O.hard_result(X, y, r["theta"], s)
        assert min(h["n_left"], h["n_right"]) >= 28


def test_parent_and_hard_fits_match_least_squares_without_ridge():
    rng = np.random.default_rng(15)
    X = rng.uniform(-1, 1, (70, 2))
    y = np.where(X[:, 1] > 0, 2 * X[:, 0], 1 - X[:, 0])
    s = O.Settings(ridge=0.0, degree=2, bias_last=False)
    A = np.column_stack([np.ones(70), X, X[:, 0] ** 2, X[:, 0] * X[:, 1], X[:, 1] ** 2])
    b, *_ = np.linalg.lstsq(A, y, rcond=None)
    assert O.parent_loss(X, y, s) == pytest.approx(np.mean((A @ b - y) ** 2), rel=1e-9, abs=1e-15)
    h = O.hard_result(X, y, np.array([0., 1., 0.]), s)
    right = X[:, 1] >= 0
    for side, m in enumerate((~right, right)):
        bs, *_ = np.linalg.lstsq(A[m], y[m], rcond=None)
        np.testing.assert_allclose(h["beta"][side], bs, rtol=1e-8, atol=1e-10)
    assert h["loss"] == pytest.approx(h["n_left"] / 70 * h["left_loss"]
                                      + h["n_right"] / 70 * h["right_loss"])


def test_one_point_hard_side_without_ridge_is_singular():
    # tests/test_cuda_batched_optimizer.py:85-95: a one-point side has a
    # rank-one Gram, so its Cholesky factorisation fails.
    rng = np.random.default_rng(912)
    X = rng.uniform(-.5, .5, (20, 2))
    X[-1] = [2., 3.]
    y = np.sin(X[:, 0]) + X[:, 1] ** 2
    phi = O.design(X, O.Settings(degree=1))
    with pytest.raises(np.linalg.LinAlgError):
        R.solve_weighted_ridge(phi[-1:], y[-1:], np.ones(1), 0.0, 0.0)


def test_converged_split_is_near_the_brute_force_optimum():
    rng = np.random.default_rng(16)
    X = rng.uniform(-1, 1, (120, 2))
    y = np.where(X[:, 0] + .5 * X[:, 1] > .2, 1. + X[:, 1], -1. + .5 * X[:, 0])
    s = O.Settings(ridge=1e-8, temperature=.05, max_iters=200, min_side_points=5)
    start = O.project_unit_w(np.array([1., .3, .1]))
    r = O.optimize(X, y, start, s)
    found = O.hard_result(X, y, r["theta"], s)["loss"]
    best = np.inf
    for ang in np.linspace(0, np.pi, 181):
        w = np.array([np.cos(ang), np.sin(ang)])
        u = X @ w
        for z in np.quantile(u, np.linspace(.05, .95, 91)):
            try:
                best = min(best, O.hard_result(X, y, np.r_[w, z], s)["loss"])
            except (ValueError, np.linalg.LinAlgError):
                pass
    assert found <= best + 1e-3


# Known answers of a geometric schedule from t0 to t1 over K iterations:
# the first temperature is t0, the last t1; with t1 = t0 / 16 over five
# iterations each step halves the temperature.
_T0, _T1 = 0.2, 0.2 / 16
_SCHEDULES = {1: [0.2], 2: [0.2, 0.0125], 3: [0.2, 0.05, 0.0125],
              5: [0.2, 0.1, 0.05, 0.025, 0.0125]}


def _composed(X, y, theta0, s, temps, step):
    """K annealing steps built from the pinned pieces, in the documented
    order: loss and gradient at T_k, tangent projection, step, retraction."""
    phi = O.design(X, s)
    theta, losses = O.project_unit_w(theta0), []
    for T in temps:
        e = O.value_and_grad(X, y, phi, theta, T, s)
        losses.append(e.loss)
        theta = O.project_unit_w(theta - step * O.project_unit_w_gradient(theta, e.grad))
    return theta, np.array(losses)


def anneal_failures(annealer, schedule):
    """Names of the checks an annealer / schedule pair fails (empty: pinned)."""
    failed = []
    for K, temps in _SCHEDULES.items():
        if not np.allclose(schedule(K, _T0, _T1), temps, rtol=1e-15, atol=0.0):
            failed.append(f"schedule K={K}")
    X, y, _ = problems()
    s = O.Settings(ridge=1e-4, min_side_points=2)
    for node, theta0 in ((1, np.array([.6, .8, .1])), (2, np.array([1., -.3, -.2]))):
        for K, temps in _SCHEDULES.items():
            theta, losses = annealer(X[node], y[node], theta0, s, iters=K, t0=_T0, t1=_T1,
                                     step=.05)
            ref_theta, ref_losses = _composed(X[node], y[node], theta0, s, temps, .05)
            if len(losses) != K or not np.allclose(losses, ref_losses, rtol=1e-13, atol=0.0):
                failed.append(f"losses node={node} K={K}")
            if not np.allclose(theta, ref_theta, rtol=1e-13, atol=1e-15):
                failed.append(f"theta node={node} K={K}")
    return failed


def test_annealed_mode_is_pinned_by_its_schedule_and_composition():
    assert anneal_failures(O.annealed, O.anneal_schedule) == []


def _planted(mistake):
    """The annealer with one plausible mistake planted."""
    def schedule(iters, t0, t1):
        k = np.arange(iters, dtype=float)
        if mistake == "exponent_k_over_K":
            return t0 * (t1 / t0) ** (k / iters)
        return O.anneal_schedule(iters, t0, t1)

    def annealer(X, y, theta0, s, iters=200, t0=0.2, t1=0.01, step=0.05):
        phi = O.design(X, s)
        theta, losses = O.project_unit_w(theta0), []
        count = {"dropped_iteration": iters - 1, "extra_iteration": iters + 1}.get(mistake, iters)
        temps = list(schedule(max(count, 1), t0, t1))[:count] if mistake in (
            "dropped_iteration", "extra_iteration") else schedule(iters, t0, t1)
        for T in temps:
            e = O.value_and_grad(X, y, phi, theta, T, s)
            losses.append(e.loss)
            if mistake == "step_before_projection":
                theta = O.project_unit_w(theta - step * e.grad)
            elif mistake == "ascent_step":
                theta = O.project_unit_w(theta + step * O.project_unit_w_gradient(theta, e.grad))
            else:
                theta = O.project_unit_w(theta - step * O.project_unit_w_gradient(theta, e.grad))
        return theta, np.array(losses)

    return annealer, schedule


@pytest.mark.parametrize("mistake", ["exponent_k_over_K", "step_before_projection",
                                     "dropped_iteration", "extra_iteration", "ascent_step"])
def test_annealed_checks_reject_planted_mistakes(mistake):
    assert anneal_failures(*_planted(mistake)), mistake


def test_annealed_mode_decreases_the_loss_on_a_clean_split():
    rng = np.random.default_rng(17)
    X = rng.uniform(-1, 1, (100, 2))
    y = np.where(X[:, 0] > 0, 1., -1.)
    s = O.Settings(ridge=1e-6)
    theta, losses = O.annealed(X, y, np.array([.6, .8, .3]), s, iters=60, t0=.3, t1=.05, step=.2)
    assert losses[-1] < losses[0]
    assert abs(np.linalg.norm(theta[:-1]) - 1) < 1e-14


def test_device_start_is_a_unit_gaussian_direction_and_a_median():
    rng = np.random.default_rng(5)
    X = rng.uniform(-1, 1, (101, 4))
    starts = np.array([O.device_start(X, 77, q, 3) for q in range(3000)])
    np.testing.assert_allclose(np.linalg.norm(starts[:, :4], axis=1), 1.0, rtol=1e-14)
    np.testing.assert_allclose((starts[:, :4] ** 2).mean(axis=0), .25, atol=.02)
    np.testing.assert_allclose(starts[:, :4].mean(axis=0), 0.0, atol=.05)
    for t in starts[:20]:
        u = X @ t[:4]
        assert np.count_nonzero(u < t[4]) == 50 and np.count_nonzero(u > t[4]) == 50
    assert not np.array_equal(O.device_start(X, 77, 0, 3), O.device_start(X, 77, 0, 4))
