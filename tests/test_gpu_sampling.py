"""GPU parity of the hit-and-run sampler vs the oracle.

Fixed-stream runs (directions and uniforms given) match the oracle's
restatement of the reference's batched and scalar samplers to 1e-12
(the reference holds its own implementations to 1e-13 of each other,
tests/test_sampling.py:175-262; the device recomputes slacks each move
instead of updating them, which moves the last bits).  Philox runs match the
oracle's stream for one move to FP32 direction accuracy, satisfy every
constraint, and reproduce the uniform distribution's moments.
"""

import numpy as np
import pytest

from oracle import sampling as S

pytestmark = pytest.mark.gpu

# The device draws from the chord shrunk by this fraction at each end.
M = MARGIN = S.DEVICE_CHORD_MARGIN


def _cut_box(rng, d, k, lo=-1., hi=1.):
    A, b = S.box_region(np.full(d, lo), np.full(d, hi))
    star = rng.uniform(.8 * lo, .8 * hi, d)
    a = rng.normal(size=(k, d))
    a /= np.linalg.norm(a, axis=1, keepdims=True)
    return np.vstack([A, a]), np.concatenate([b, a @ star + .05 + rng.uniform(0, .5, k)]), star


def _pack(directions, uniforms, tail_d, tail_u, C):
    """(steps, C, d) full arrays plus the tail for the first rem chains."""
    if tail_d is None:
        return directions, uniforms
    S_, d = tail_d.shape[0], tail_d.shape[2]
    td = np.zeros((S_, C, d))
    tu = np.zeros((S_, C))
    td[:, :tail_d.shape[1]] = tail_d
    tu[:, :tail_u.shape[1]] = tail_u
    return np.concatenate([directions, td]), np.concatenate([uniforms, tu])


def test_philox_matches_the_oracle_bit_for_bit(torch_gpu):
    from art import _hip

    torch = torch_gpu
    rng = np.random.default_rng(3)
    N = 4099
    ctr = rng.integers(0, 2 ** 32, (N, 4), dtype=np.uint64).astype(np.uint32)
    key = rng.integers(0, 2 ** 32, (N, 2), dtype=np.uint64).astype(np.uint32)
    ctr[0], key[0] = 0, 0
    c = torch.from_numpy(ctr.view(np.int32)).cuda()
    k = torch.from_numpy(key.view(np.int32)).cuda()
    out = torch.empty((N, 4), dtype=torch.int32, device="cuda")
    lib = _hip.load()
    _hip.check(lib.art_philox4x32_10(_hip.ptr(c), _hip.ptr(k), N, _hip.ptr(out), None))
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(np.uint32)
    np.testing.assert_array_equal(got, S.philox4x32_10(ctr, key))
    np.testing.assert_array_equal(got[0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8])


@pytest.mark.parametrize("split_box", [True, False])
@pytest.mark.parametrize("C,n,burn,thin", [(1, 12, 3, 1), (4, 11, 2, 3), (4, 12, 0, 1),
                                           (70, 213, 5, 2), (3, 3, 0, 4)])
def test_fixed_stream_matches_the_batched_oracle(torch_gpu, C, n, burn, thin, split_box):
    from art.sampling import PolytopeBatch

    rng = np.random.default_rng(37 + C)
    A, b, star = _cut_box(rng, 3, 5)
    starts = star + rng.uniform(-.01, .01, (C, 3))
    full, rem = divmod(n, C)
    steps = burn + full * thin
    dirs, us = rng.normal(size=(steps, C, 3)), rng.random((steps, C))
    td = rng.normal(size=(thin, rem, 3)) if rem else None
    tu = rng.random((thin, rem)) if rem else None
    expected = S.batched_sample(A, b, starts, n, burn, thin, dirs, us, td, tu, margin=M)
    D, U = _pack(dirs, us, td, tu, C)
    got = PolytopeBatch(([A], [b]), split_box=split_box).sample(
        starts[None], n, first=burn + thin, stride=thin, directions=D[None], uniforms=U[None])
    np.testing.assert_allclose(got[0].cpu().numpy(), expected, rtol=1e-12, atol=1e-12)


def test_many_polytopes_of_different_sizes_in_one_launch(torch_gpu):
    from art.sampling import PolytopeBatch

    rng = np.random.default_rng(5)
    P, C, n, burn, thin, d = 7, 5, 13, 1, 2, 4
    As, bs, stars = zip(*[_cut_box(rng, d, int(k)) for k in rng.integers(1, 30, P)])
    starts = np.stack([np.repeat(s[None], C, axis=0) for s in stars])
    full, rem = divmod(n, C)
    steps = burn + full * thin + thin
    dirs = rng.normal(size=(P, steps, C, d))
    us = rng.random((P, steps, C))
    got = PolytopeBatch((list(As), list(bs))).sample(starts, n, first=burn + thin, stride=thin,
                                                     directions=dirs, uniforms=us)
    got = got.cpu().numpy()
    for p in range(P):
        nf = burn + full * thin
        exp = S.batched_sample(As[p], bs[p], starts[p], n, burn, thin, dirs[p, :nf], us[p, :nf],
                               dirs[p, nf:, :rem], us[p, nf:, :rem], margin=M)
        np.testing.assert_allclose(got[p], exp, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("thin", [1, 4])
def test_scalar_schedule_matches_the_scalar_oracle(torch_gpu, thin):
    from art.sampling import PolytopeBatch

    A, b = S.box_region([-1.5, -2.], [2., 1.])
    x0 = np.array([.25, -.4])
    n, burn = 12, 3
    rng = np.random.default_rng(29)
    steps = burn + n * thin
    dirs, us = rng.normal(size=(steps, 2)), rng.random(steps)
    expected = S.scalar_sample(A, b, x0, n, burn, thin, dirs, us, margin=M)
    got = PolytopeBatch(([A], [b])).sample(x0[None, None], n, first=burn + 1, stride=thin,
                                           directions=dirs[None, :, None], uniforms=us[None, :, None])
    np.testing.assert_allclose(got[0].cpu().numpy(), expected, rtol=1e-12, atol=1e-12)


def test_shaped_directions_match_the_oracle(torch_gpu):
    from art.sampling import PolytopeBatch

    ang = np.deg2rad(35.)
    V = np.array([[np.cos(ang), -np.sin(ang)], [np.sin(ang), np.cos(ang)]])
    M = S.shaping_matrix(V, [4., .25], 1e-4)
    A, b = S.box_region([-2., -1.], [2., 1.])
    C, n, burn, thin = 4, 12, 2, 3
    rng = np.random.default_rng(41)
    starts = np.array([[0., 0.], [.5, -.5], [-.75, .25], [1., .5]])
    steps = burn + (n // C) * thin
    dirs, us = rng.normal(size=(steps, C, 2)), rng.random((steps, C))
    expected = S.batched_sample(A, b, starts, n, burn, thin, dirs @ M.T, us, margin=MARGIN)
    got = PolytopeBatch(([A], [b])).sample(starts[None], n, first=burn + thin, stride=thin,
                                           shape=M, directions=dirs[None], uniforms=us[None])
    np.testing.assert_allclose(got[0].cpu().numpy(), expected, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("d", [2, 5, 12])
def test_philox_stream_first_move_matches_the_oracle(torch_gpu, d):
    from art.sampling import PolytopeBatch

    rng = np.random.default_rng(d)
    P, C, seed = 3, 70, 0x1234_5678_9ABC
    regs = [_cut_box(rng, d, 6) for _ in range(P)]
    starts = np.stack([np.repeat(r[2][None], C, axis=0) for r in regs])
    got = PolytopeBatch(([r[0] for r in regs], [r[1] for r in regs])).sample(
        starts, C, first=1, stride=1, seed=seed).cpu().numpy()
    for p, (A, b, star) in enumerate(regs):
        dirs, us = S.step_randoms(seed, p, np.arange(C), 0, d)
        pts, slack = starts[p].copy(), b[None] - starts[p] @ A.T
        S.advance(A, pts, slack, dirs.astype(float), us, M)
        np.testing.assert_allclose(got[p], pts, rtol=1e-5, atol=1e-6)


def test_constraints_hold_and_seeds_reproduce(torch_gpu):
    from art.sampling import PolytopeBatch

    rng = np.random.default_rng(4)
    P, C, d = 64, 64, 12
    regs = [_cut_box(rng, d, int(k), -1., 1.) for k in rng.integers(4, 41, P)]
    batch = PolytopeBatch(([r[0] for r in regs], [r[1] for r in regs]))
    starts = np.stack([np.repeat(r[2][None], C, axis=0) for r in regs])
    n = C * 20
    one = batch.sample(starts, n, first=10, stride=10, seed=11).cpu().numpy()
    two = batch.sample(starts, n, first=10, stride=10, seed=11).cpu().numpy()
    three = batch.sample(starts, n, first=10, stride=10, seed=12).cpu().numpy()
    np.testing.assert_array_equal(one, two)
    assert not np.array_equal(one, three)
    worst = max(float(np.max(one[p] @ A.T - b)) for p, (A, b, _) in enumerate(regs))
    assert worst <= 1e-12


def test_box_moments_match_the_uniform_distribution(torch_gpu):
    from art.sampling import PolytopeBatch

    d, P, C = 12, 32, 64
    A, b = S.box_region(-np.ones(d), np.ones(d))
    n = C * 100
    out = PolytopeBatch(([A] * P, [b] * P)).sample(np.zeros((P, C, d)), n, first=200,
                                                   stride=10, seed=5).cpu().numpy()
    flat = out.reshape(-1, d)
    assert np.all(np.abs(flat) <= 1 + 1e-12)
    se = np.sqrt(1 / 3 / flat.shape[0])
    assert np.all(np.abs(flat.mean(axis=0)) < 30 * se)  # autocorrelated chains
    np.testing.assert_allclose(flat.var(axis=0), 1 / 3, atol=.01)


def test_cut_box_moments_match_rejection_sampling(torch_gpu):
    from art.sampling import PolytopeBatch

    A, b = S.box_region([-1., -1.], [1., 1.])
    A, b = np.vstack([A, [1., 1.]]), np.append(b, .5)
    rng = np.random.default_rng(7)
    pool = rng.uniform(-1, 1, (2_000_000, 2))
    exact = pool[pool.sum(axis=1) <= .5]
    P, C = 16, 256
    out = PolytopeBatch(([A] * P, [b] * P)).sample(np.full((P, C, 2), -.2), C * 50, first=50,
                                                   stride=5, seed=9).cpu().numpy().reshape(-1, 2)
    assert np.all(out @ A.T <= b + 1e-12)
    np.testing.assert_allclose(out.mean(axis=0), exact.mean(axis=0), atol=.01)
    np.testing.assert_allclose(np.cov(out.T), np.cov(exact.T), atol=.01)


def test_drop_in_samplers(torch_gpu):
    from art.domain import BoxDomain
    from art.sampling import (BatchedHitAndRunSampler, HitAndRunSampler,
                              NumbaBatchedHitAndRunSampler, ParallelNumbaBatchedHitAndRunSampler)

    region = BoxDomain(np.array([[-1.5, 2.0], [-2.0, 1.0]])).as_region()
    pool = np.array([[0., 0.], [.5, -.5], [-.75, .25], [1., .5], [-1., -1.]])
    for cls in (BatchedHitAndRunSampler, NumbaBatchedHitAndRunSampler,
                ParallelNumbaBatchedHitAndRunSampler):
        s = cls(burn_in=2, thinning=3, n_chains=4)
        first = s.sample(region, 11, random_state=37, x0=pool)
        again = s.sample(region, 11, random_state=37, x0=pool)
        assert first.shape == (11, 2)
        np.testing.assert_array_equal(first, again)
        assert np.all(region.contains(first, tol=1e-12))
    ang = np.deg2rad(35.0)
    V = np.array([[np.cos(ang), -np.sin(ang)], [np.sin(ang), np.cos(ang)]])
    box = BoxDomain(np.array([[-2.0, 2.0], [-1.0, 1.0]])).as_region()
    s = HitAndRunSampler(burn_in=0, thinning=2, direction_eigenvectors=V,
                         direction_eigenvalues=np.array([4.0, 0.25]))
    out = s.sample(box, 100, random_state=3, x0=np.zeros(2))
    assert out.shape == (100, 2) and np.all(box.contains(out))
    # Without x0 the start comes from rejection sampling in the inferred box.
    out = BatchedHitAndRunSampler(burn_in=5, thinning=2, n_chains=3).sample(box, 7,
                                                                            random_state=1)
    assert out.shape == (7, 2) and np.all(box.contains(out, tol=1e-12))
    timed = BatchedHitAndRunSampler(burn_in=0, thinning=2, n_chains=2, profile_timing=True)
    timed.sample(box, 6, random_state=43, x0=np.array([[0., 0.], [.25, -.25]]))
    assert {"slack_matmul_seconds", "constraint_product_matmul_seconds",
            "random_generation_seconds", "transitions_seconds"} <= set(timed.last_phase_timings)


@pytest.mark.parametrize("d,k,shaped", [(2, 6, False), (3, 0, False), (5, 9, True),
                                        (12, 30, False), (20, 4, True)])
def test_single_chains_equal_chain_zero_of_a_two_chain_run(torch_gpu, d, k, shaped):
    # One chain per polytope runs on the wave-cooperative kernel, two chains on
    # the lane-per-chain kernel; chain 0 draws the same Philox stream in both,
    # so its records must agree bit for bit.
    from art.sampling import PolytopeBatch

    rng = np.random.default_rng(100 + d)
    regs = [_cut_box(rng, d, k) for _ in range(3)]
    batch = PolytopeBatch(([r[0] for r in regs], [r[1] for r in regs]))
    shape = None
    if shaped:
        V = np.linalg.qr(rng.normal(size=(d, d)))[0]
        shape = S.shaping_matrix(V, rng.uniform(0.1, 2.0, d), 1e-4)
    n, first, stride, seed = 37, 3, 5, 77
    one = batch.sample(np.stack([r[2] for r in regs])[:, None], n, first=first, stride=stride,
                       seed=seed, shape=shape).cpu().numpy()
    starts = np.stack([np.stack([r[2], r[2]]) for r in regs])
    two = batch.sample(starts, 2 * n, first=first, stride=stride, seed=seed,
                       shape=shape).cpu().numpy()
    np.testing.assert_array_equal(one, two[:, 0::2])
    for (A, b, _), pts in zip(regs, one):
        assert np.all(pts @ A.T <= b + 1e-9)


def test_single_chain_with_cuts_matches_the_scalar_oracle(torch_gpu):
    from art.sampling import PolytopeBatch

    rng = np.random.default_rng(61)
    regs = [_cut_box(rng, 3, 5) for _ in range(2)]
    n, burn, thin = 25, 4, 3
    steps = burn + n * thin
    dirs, us = rng.normal(size=(2, steps, 1, 3)), rng.random((2, steps, 1))
    got = PolytopeBatch(([r[0] for r in regs], [r[1] for r in regs])).sample(
        np.stack([r[2] for r in regs])[:, None], n, first=burn + 1, stride=thin,
        directions=dirs, uniforms=us).cpu().numpy()
    for p, (A, b, star) in enumerate(regs):
        exp = S.scalar_sample(A, b, star, n, burn, thin, dirs[p, :, 0], us[p, :, 0], margin=M)
        np.testing.assert_allclose(got[p], exp, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("d,k", [(2, 10), (12, 30)])
def test_single_chain_understated_row_bound_replays_safely(torch_gpu, d, k):
    """art_har_sample's row_bound sizes the single-chain kernel's LDS entries;
    a polytope with more rows than the bound replays its chain on the
    lane-per-chain code, so the samples equal those of the exact bound."""
    from art.sampling import PolytopeBatch

    rng = np.random.default_rng(500 + d)
    regs = [_cut_box(rng, d, kk) for kk in (1, k, 3)]
    batch = PolytopeBatch(([r[0] for r in regs], [r[1] for r in regs]))
    starts = np.stack([r[2] for r in regs])[:, None]
    exact = batch.sample(starts, 40, first=3, stride=4, seed=9).cpu().numpy()
    true_rows = batch.max_rows
    batch.max_rows = 2  # the kernel is told at most 2 general rows
    try:
        low = batch.sample(starts, 40, first=3, stride=4, seed=9).cpu().numpy()
    finally:
        batch.max_rows = true_rows
    np.testing.assert_array_equal(low, exact)
    for (A, b, _), pts in zip(regs, low):
        assert np.all(pts @ A.T <= b + 1e-9)


def test_multi_move_fixed_stream_on_hit_and_run_workload_polytopes(torch_gpu):
    """The hit-and-run configuration's shapes: d = 12 polytopes of the box and
    4-40 random cuts (synthetic.polytope_workload), 64 chains per polytope,
    200 moves recorded every 10 (20 records per chain), fixed directions and
    uniforms: the device equals the oracle's batched sampler (with the
    device's chord margin) over the whole chains, and every sample keeps a
    strictly positive slack on every row."""
    from art.sampling import PolytopeBatch
    from art.synthetic import polytope_workload

    d, C, thin, records = 12, 64, 10, 20
    wl = polytope_workload(n_polytopes=6, d=d, seed=44)
    n, steps = C * records, thin * records
    rng = np.random.default_rng(45)
    P = wl.n_polytopes
    dirs, us = rng.normal(size=(P, steps, C, d)), rng.random((P, steps, C))
    starts = np.repeat(wl.starts[:, None, :], C, axis=1)
    got = PolytopeBatch.from_csr(wl.A, wl.b, wl.offsets).sample(
        starts, n, first=thin, stride=thin, directions=dirs, uniforms=us).cpu().numpy()
    for q in range(P):
        A, b = wl.polytope(q)
        exp = S.batched_sample(A, b, starts[q], n, 0, thin, dirs[q], us[q], margin=M)
        # Chord ends move with x, so the few-ulp differences of the device's
        # slacks and reciprocal ratios (b - a.x cancels near a face) grow along
        # a chain; the oracle itself, rerun with every uniform moved by one
        # ulp, spreads by up to 2e-11 on these chains.  A wrong chord (sign,
        # row, face) is off by O(1) at once.  So: the first five records of
        # every chain (50 moves) agree to 1e-12, every coordinate of all 20
        # records to 1e-8 (measured 4e-9), and the median chain to 1e-10
        # (measured 1e-12).
        # Record k of chain c is sample k * C + c.
        diff = np.abs(got[q] - exp)
        assert diff.max() <= 1e-8, diff.max()
        assert np.median(diff.reshape(records, C, d).max(axis=(0, 2))) <= 1e-10
        assert np.abs(got[q] - exp)[:5 * C].max() <= 1e-12
        assert np.all(b[None, :] - got[q] @ A.T > 0.0)
