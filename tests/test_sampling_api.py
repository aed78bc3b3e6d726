"""Host-side validation of the sampler API (no GPU needed: every check runs
before the device is touched), following tests/test_sampling.py."""

import numpy as np
import pytest

from art.domain import BoxDomain, PolytopeRegion, split_region
from art.sampling import (BatchedHitAndRunSampler, HitAndRunSampler,
                          ParallelNumbaBatchedHitAndRunSampler, find_feasible_point,
                          floor_covariance_eigenvalues, infer_box_bounds,
                          sample_covariance_eigendecomposition, split_axis_rows)


def test_region_and_split_conventions():
    box = BoxDomain.hypercube(2, -1., 1.).as_region()
    left, right = split_region(box, [1., 0.], .25)
    assert left.contains([0., 0.])[0] and not right.contains([0., 0.])[0]
    assert right.contains([.5, 0.])[0] and not left.contains([.5, 0.])[0]
    assert left.depth == right.depth == 1
    with pytest.raises(ValueError):
        PolytopeRegion(np.ones((2, 2)), np.ones(3))
    with pytest.raises(ValueError):
        BoxDomain(np.array([[1., 0.]]))


def test_inferred_bounds_and_feasible_point():
    box = BoxDomain(np.array([[-1., 2.], [-2., 4.]])).as_region()
    np.testing.assert_array_equal(infer_box_bounds(box), [[-1., 2.], [-2., 4.]])
    left, _ = split_region(box, [1., 1.], 0.)
    x = find_feasible_point(left, random_state=3)
    assert left.contains(x)[0]
    assert infer_box_bounds(PolytopeRegion(np.array([[1., 1.]]), np.array([1.]))) is None


def test_covariance_helpers():
    safe, floor = floor_covariance_eigenvalues(np.array([4.0, 0.0]), floor_ratio=1e-2)
    assert floor == 2e-2
    np.testing.assert_allclose(safe, [4.0, 2e-2])
    rng = np.random.default_rng(4)
    samples = rng.normal(size=(500, 3)) @ np.array([[2., .5, 0.], [0., 1., .2], [0., 0., .4]])
    V, lam = sample_covariance_eigendecomposition(samples)
    np.testing.assert_allclose((V * lam) @ V.T, np.cov(samples, rowvar=False), rtol=1e-12,
                               atol=1e-12)
    s = HitAndRunSampler(direction_eigenvectors=np.eye(2),
                         direction_eigenvalues=np.array([1.0, 0.0]),
                         direction_eigenvalue_floor=1e-4)
    _, sq = s._direction_spectrum(2)
    np.testing.assert_allclose(sq, [1.0, np.sqrt(5e-5)])


@pytest.mark.parametrize("kwargs,n,match", [
    ({"n_chains": 0}, 4, "n_chains"), ({"thinning": -1}, 4, "thinning"),
    ({"burn_in": -1}, 4, "burn_in"), ({}, -1, "n must"),
])
def test_batched_sampler_validation(kwargs, n, match):
    box = BoxDomain.hypercube(2, -1., 1.).as_region()
    with pytest.raises(ValueError, match=match):
        BatchedHitAndRunSampler(**kwargs).sample(box, n, x0=np.zeros(2))


def test_parallel_sampler_defaults_and_thread_check():
    s = ParallelNumbaBatchedHitAndRunSampler()
    assert s.n_chains == 50 and s.parallel_random and s.random_threads == 10
    box = BoxDomain.hypercube(2, -1., 1.).as_region()
    with pytest.raises(ValueError, match="random_threads must be at least 1"):
        ParallelNumbaBatchedHitAndRunSampler(n_chains=2, random_threads=0).sample(
            box, 2, random_state=1, x0=np.zeros(2))
    assert BatchedHitAndRunSampler().sample(box, 0).shape == (0, 2)
    with pytest.raises(ValueError, match="thinning"):
        HitAndRunSampler(thinning=0).sample(box, 3, x0=np.zeros(2))
    with pytest.raises(ValueError, match="feasible"):
        HitAndRunSampler().sample(box, 3, x0=np.array([2., 0.]))


def test_axis_rows_become_the_tightest_bounds():
    A = np.array([[1., 0.], [-1., 0.], [0., 1.], [0., -1.], [1., 0.], [2., 0.], [1., 1.],
                  [0., -1.]])
    b = np.array([1., 1., 2., 3., .5, 1., .3, 2.5])
    offsets = np.array([0, 5, 8])
    Ag, bg, og, lo, hi = split_axis_rows(A, b, offsets)
    np.testing.assert_array_equal(hi, [[.5, 2.], [np.inf, np.inf]])
    np.testing.assert_array_equal(lo, [[-1., -3.], [-np.inf, -2.5]])
    np.testing.assert_array_equal(og, [0, 0, 2])
    np.testing.assert_array_equal(Ag, [[2., 0.], [1., 1.]])
    np.testing.assert_array_equal(bg, [1., .3])
    # The chord from bounds equals the chord from the rows.
    from oracle import sampling as S
    x, dvec = np.array([.1, .2]), np.array([.3, -.7])
    lo_r, hi_r = S.step_interval(A[:5], b[:5], x, dvec)
    r = np.concatenate([(lo[0] - x) / dvec, (hi[0] - x) / dvec])
    assert lo_r == max(r[r < 0]) and hi_r == min(r[r > 0])
