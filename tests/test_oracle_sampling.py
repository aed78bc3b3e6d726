"""Pins of the hit-and-run oracle (oracle/sampling.py), CPU only."""

import numpy as np
import pytest

from oracle import sampling as S


def test_philox_known_answer_vectors():
    # Published Philox4x32-10 known-answer vectors (Random123 distribution).
    cases = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, expected in cases:
        np.testing.assert_array_equal(S.philox4x32_10([ctr], [key])[0],
                                      np.array(expected, dtype=np.uint32))


def test_stream_uniforms_are_uniform_and_directions_gaussian():
    d = 5
    dirs, u = S.step_randoms(12345, 3, np.arange(4000)[:, None], np.arange(5)[None, :], d)
    assert dirs.shape == (4000, 5, d) and dirs.dtype == np.float32
    assert u.min() >= 0.0 and u.max() < 1.0
    flat = dirs.reshape(-1, d).astype(float)
    n = flat.shape[0]
    assert abs(flat.mean()) < 5 / np.sqrt(n * d)
    np.testing.assert_allclose(flat.var(axis=0), 1.0, atol=5 * np.sqrt(2 / n))
    assert abs(np.mean(u) - .5) < 5 * np.sqrt(1 / 12 / u.size)
    # distinct streams for distinct chains, steps, polytopes and seeds
    a = S.stream_words(1, 0, 0, 0, d)
    for other in (S.stream_words(1, 0, 1, 0, d), S.stream_words(1, 0, 0, 1, d),
                  S.stream_words(1, 1, 0, 0, d), S.stream_words(2, 0, 0, 0, d)):
        assert not np.array_equal(a, other)


def test_step_interval_matches_box_boundaries():
    # tests/test_sampling.py:446-458.
    A, b = S.box_region([-1., -2.], [2., 4.])
    lo, hi = S.step_interval(A, b, np.array([.5, 1.]), np.array([.6, .8]))
    assert lo == pytest.approx(-2.5) and hi == pytest.approx(2.5)
    with pytest.raises(RuntimeError, match="infeasible"):
        A, b = S.box_region([-1., -1.], [1., 1.])
        S.step_interval(A, b, np.array([0., 1. + 2e-10]), np.array([1., 0.]))


@pytest.mark.parametrize("thinning", [1, 3])
def test_scalar_and_batched_schedules_agree_on_a_fixed_stream(thinning):
    # tests/test_sampling.py:175-228 asserts the scalar and batched samplers
    # agree on one chain with thinning 1; the generic schedule with
    # first = burn_in + 1 reproduces the scalar sampler for any thinning.
    A, b = S.box_region([-1.5, -2.], [2., 1.])
    x0 = np.array([.25, -.4])
    n, burn = 12, 3
    rng = np.random.default_rng(29)
    steps = burn + n * thinning
    dirs = rng.normal(size=(steps, 2))
    us = rng.random(steps)
    scalar = S.scalar_sample(A, b, x0, n, burn, thinning, dirs, us)
    generic = S.run_chains(A, b, x0[None], n, burn + 1, thinning,
                           lambda s, live: (dirs[s][None], us[s:s + 1]))
    np.testing.assert_allclose(generic, scalar, rtol=1e-13, atol=1e-13)
    if thinning == 1:
        batched = S.batched_sample(A, b, x0[None], n, burn, 1, dirs[:, None, :], us[:, None])
        np.testing.assert_allclose(batched, scalar, rtol=1e-13, atol=1e-13)


def test_batched_schedule_tail_round_and_order():
    A, b = S.box_region([-1.5, -2.], [2., 1.])
    starts = np.array([[0., 0.], [.5, -.5], [-.75, .25], [1., .5]])
    n, burn, thin = 11, 2, 3
    C = 4
    full, rem = divmod(n, C)
    rng = np.random.default_rng(37)
    dirs = rng.normal(size=(burn + full * thin, C, 2))
    us = rng.random((burn + full * thin, C))
    tdirs, tus = rng.normal(size=(thin, rem, 2)), rng.random((thin, rem))
    out = S.batched_sample(A, b, starts, n, burn, thin, dirs, us, tdirs, tus)
    # Chain-by-chain restatement: chain c records after burn + k*thin moves.
    for c in range(C):
        x = starts[c].copy()
        slack = b - A @ x
        k_rec = full + (c < rem)
        moves = burn + k_rec * thin
        recs = []
        for s in range(moves):
            d, u = (dirs[s, c], us[s, c]) if s < burn + full * thin else \
                (tdirs[s - burn - full * thin, c], tus[s - burn - full * thin, c])
            p, sl = x[None].copy(), slack[None].copy()
            S.advance(A, p, sl, d[None].copy(), np.array([u]))
            x, slack = p[0], sl[0]
            if s + 1 > burn and (s + 1 - burn) % thin == 0:
                recs.append(x.copy())
        for k, r in enumerate(recs):
            np.testing.assert_array_equal(out[k * C + c], r)


def test_samples_satisfy_every_constraint_and_match_box_moments():
    d, C = 3, 256
    A, b = S.box_region(-np.ones(d), np.ones(d))
    rng = np.random.default_rng(4)
    n = C * 40
    steps = 20 + 40 * 5
    dirs = rng.normal(size=(steps, C, d))
    us = rng.random((steps, C))
    out = S.batched_sample(A, b, np.zeros((C, d)), n, 20, 5, dirs, us)
    assert np.all(out @ A.T <= b + 1e-12)
    se = np.sqrt(1 / 3 / n) * 3  # chains are correlated within; generous bound
    assert np.all(np.abs(out.mean(axis=0)) < 5 * se)
    np.testing.assert_allclose(out.var(axis=0), 1 / 3, atol=.03)


def test_cut_box_moments_match_rejection_sampling():
    # x in [-1,1]^2 with x0 + x1 <= .5: brute-force rejection gives the target.
    A, b = S.box_region([-1., -1.], [1., 1.])
    A, b = np.vstack([A, [1., 1.]]), np.append(b, .5)
    rng = np.random.default_rng(7)
    pool = rng.uniform(-1, 1, (400000, 2))
    exact = pool[pool.sum(axis=1) <= .5]
    C = 512
    n = C * 30
    steps = 30 + 30 * 4
    out = S.batched_sample(A, b, np.full((C, 2), -.2), n, 30, 4, rng.normal(size=(steps, C, 2)),
                           rng.random((steps, C)))
    assert np.all(out @ A.T <= b + 1e-12)
    np.testing.assert_allclose(out.mean(axis=0), exact.mean(axis=0), atol=.03)
    np.testing.assert_allclose(np.cov(out.T), np.cov(exact.T), atol=.03)


def test_shaping_matrix_floor_and_scale():
    # tests/test_sampling.py:429-442: eigenvalues [1, 0] with floor ratio 1e-4
    # give sqrt scales [1, sqrt(5e-5)].
    M = S.shaping_matrix(np.eye(2), [1., 0.], 1e-4)
    np.testing.assert_allclose(np.diag(M), [1., np.sqrt(5e-5)])


def test_device_chord_margin_stays_inside_and_reduces_to_the_reference_rule():
    """chord_point at margin 0 is t_lo + u (t_hi - t_lo) (the reference's
    step, art/sampling.py:877-962); at the device margin every point lies at
    least margin (t_hi - t_lo) inside each end, the ends map to exactly that
    distance, and the move differs from the reference's by at most
    margin (t_hi - t_lo)."""
    rng = np.random.default_rng(3)
    lo, hi = rng.uniform(-2, 0, 1000), rng.uniform(0, 3, 1000)
    u = rng.random(1000)
    np.testing.assert_array_equal(S.chord_point(lo, hi, u), (hi - lo) * u + lo)
    m = S.DEVICE_CHORD_MARGIN
    t = S.chord_point(lo, hi, u, m)
    L = hi - lo
    assert np.all(t - lo >= 0.999 * m * L) and np.all(hi - t >= 0.999 * m * L)
    assert np.all(np.abs(t - S.chord_point(lo, hi, u)) <= 2.001 * m * L)
    np.testing.assert_allclose(S.chord_point(lo, hi, np.zeros(1000), m) - lo, m * L, rtol=1e-3)
    np.testing.assert_allclose(hi - S.chord_point(lo, hi, np.ones(1000), m), m * L, rtol=1e-3)
    # On a thin cut box every sample keeps a strictly positive slack.
    A, b, star = _cut_box_for_margin(rng)
    C, n = 8, 400
    dirs, us = rng.normal(size=(n // C * 5, C, 3)), rng.random((n // C * 5, C))
    out = S.batched_sample(A, b, np.repeat(star[None], C, axis=0), n, 0, 5, dirs, us,
                           margin=m)
    assert np.all(b[None, :] - out @ A.T > 0.0)


def _cut_box_for_margin(rng):
    A, b = S.box_region(np.full(3, -1.0), np.full(3, 1.0))
    a = rng.normal(size=(6, 3))
    a /= np.linalg.norm(a, axis=1, keepdims=True)
    star = rng.uniform(-.5, .5, 3)
    return np.vstack([A, a]), np.concatenate([b, a @ star + 0.01]), star
