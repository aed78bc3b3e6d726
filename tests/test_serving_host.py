"""Host-side helpers of the serving loop and the split-search batch (no GPU):
vectorized start keys and start projections agree with the per-node forms."""

import numpy as np

from art.serving import start_key, start_keys
from art.splitters import project_unit_w, project_unit_w_rows


def test_start_keys_equal_the_per_node_keys():
    rng = np.random.default_rng(3)
    seeds = rng.integers(0, 2**63, size=500).tolist() + [0, 1, 2**64 - 1, 2**63]
    heaps = rng.integers(1, 2**11, size=len(seeds)).tolist()
    heaps[-4:] = [1, 2**55, 2**60 - 1, 3]       # deep heaps: (heap << 8) wraps mod 2^64
    restarts = rng.integers(0, 64, size=len(seeds)).tolist()
    got = start_keys(seeds, heaps, restarts)
    assert got.dtype == np.uint64
    want = [start_key(s, h, r) for s, h, r in zip(seeds, heaps, restarts)]
    assert [int(v) for v in got] == want


def test_row_projection_matches_the_single_row_projection():
    rng = np.random.default_rng(4)
    thetas = rng.normal(size=(200, 9)) * rng.uniform(1e-3, 1e3, size=(200, 1))
    got = project_unit_w_rows(thetas)
    want = np.stack([project_unit_w(t) for t in thetas])
    np.testing.assert_allclose(got, want, rtol=4e-16, atol=0)
    np.testing.assert_allclose(np.linalg.norm(got[:, :-1], axis=1), 1.0, rtol=1e-15)
    np.testing.assert_array_equal(got[:, -1], thetas[:, -1])
    bad = thetas.copy()
    bad[5, :-1] = 0.0
    try:
        project_unit_w_rows(bad)
    except ValueError as exc:
        assert "near-zero" in str(exc)
    else:
        raise AssertionError("a zero direction must be refused")
