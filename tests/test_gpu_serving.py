"""GPU parity of the serving loop: row kernels, served fits vs the oracle,
results alone vs in the mix, and rank-k refits vs fits on all rows."""

import numpy as np
import pytest

from oracle import ridge as R
from oracle import serve as S

pytestmark = pytest.mark.gpu


def dev(torch, a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def test_partition_rows_is_numpy_stable_partition(torch_gpu):
    from art.serving import partition_rows

    torch, rng = torch_gpu, np.random.default_rng(5)
    d = 3
    sizes = [5000, 0, 1, 1023, 1024, 1025, 3000, 2]
    offsets = np.concatenate([[0], np.cumsum(sizes)])
    N = offsets[-1] + 17                   # rows outside every range stay untouched
    X, y, w = rng.uniform(-1, 1, (N, d)), rng.normal(size=N), rng.uniform(size=N)
    pick = [0, 1, 2, 3, 4, 5, 7]           # range 6 is not partitioned
    begin, end = offsets[pick], offsets[np.asarray(pick) + 1]
    nw = rng.normal(size=(len(pick), d))
    nz = rng.normal(size=len(pick)) * 0.3
    nz[3] = -10.0                           # everything right
    base = np.concatenate([[0], np.cumsum(end - begin)[:-1]])
    M = int((end - begin).sum())
    f64 = dict(dtype=torch.float64, device="cuda")
    out = (torch.full((M, d), np.nan, **f64), torch.full((M,), np.nan, **f64),
           torch.full((M,), np.nan, **f64))
    right = partition_rows(dev(torch, X), dev(torch, y), dev(torch, w), begin, end, nw, nz, out,
                           base)
    Xo, yo, wo = (t.cpu().numpy() for t in out)
    for k, (b, e) in enumerate(zip(begin, end)):
        s = np.zeros(e - b)
        for j in range(d):  # numpy rounds each product; no row lies that close to a plane
            s = nw[k, j] * X[b:e, j] + s
        r = (s - nz[k]) >= 0.0
        order = np.r_[np.flatnonzero(~r), np.flatnonzero(r)] + b
        assert right[k] == r.sum()
        o = base[k]
        np.testing.assert_array_equal(Xo[o:o + e - b], X[order])
        np.testing.assert_array_equal(yo[o:o + e - b], y[order])
        np.testing.assert_array_equal(wo[o:o + e - b], w[order])
    assert right[3] == sizes[3]


def test_gather_strided_and_gram_sse(torch_gpu):
    from art import LeafRidgeBatch
    from art.leaf_ridge import solve_grams
    from art.serving import gather_strided, gram_sse

    torch, rng = torch_gpu, np.random.default_rng(6)
    sizes = np.array([700, 9, 4096, 5000])
    offsets = np.concatenate([[0], np.cumsum(sizes)])
    N, d = offsets[-1], 4
    X, y, w = rng.uniform(-1, 1, (N, d)), rng.normal(size=N), rng.uniform(0.1, 1, N)
    counts = np.minimum(sizes, 512)
    Xs, ys, ws, dst = gather_strided(dev(torch, X), dev(torch, y), dev(torch, w), offsets[:-1],
                                     offsets[1:], counts)
    idx = np.concatenate([b + S.strided_sample(n, m) for b, n, m in zip(offsets, sizes, counts)])
    np.testing.assert_array_equal(Xs.cpu().numpy(), X[idx])
    np.testing.assert_array_equal(ys.cpu().numpy(), y[idx])
    np.testing.assert_array_equal(ws.cpu().numpy(), w[idx])
    with LeafRidgeBatch(offsets, d) as lb:
        grams, bad = lb.grams(dev(torch, X), dev(torch, y), dev(torch, w))
        coef, status, _ = solve_grams(grams, bad, d, 1e-6)
        sse, wsum = gram_sse(grams, coef, lb.p)
        torch.cuda.synchronize()
    beta = coef.cpu().numpy()
    for k in range(len(sizes)):
        sl = slice(offsets[k], offsets[k + 1])
        phi = R.features(X[sl], 2, True, False)
        r = y[sl] - phi @ beta[k]
        ref = float(np.sum(w[sl] * r * r))
        assert abs(sse.cpu().numpy()[k] - ref) <= 1e-9 * float(np.sum(w[sl] * y[sl] ** 2))
        np.testing.assert_allclose(wsum.cpu().numpy()[k], w[sl].sum(), rtol=1e-13)


def _request(rng, n, d, seed, **kw):
    from art.serving import FitRequest

    X = rng.uniform(-1, 1, (n, d))
    c = rng.uniform(-1, 1, (4, d))
    y = sum(np.exp(-np.sum((X - ck) ** 2, axis=1) / 0.5) * a
            for ck, a in zip(c, rng.uniform(-1, 1, 4)))
    return FitRequest(X, y, seed=seed, **kw)


def _tree_arrays(result):
    tree = result.tree
    splits = [(n.node_id, n.w, n.z) for n in tree.iter_nodes() if hasattr(n, "w")]
    return splits, result.coefficients


def test_served_fit_matches_the_oracle(torch_gpu):
    from art.serving import ServingConfig, ServingEngine

    rng = np.random.default_rng(7)
    req = _request(rng, 3000, 2, seed=11, max_depth=3, min_leaf=100)
    cfg = ServingConfig(split_rows=1024, max_iters=30)
    with ServingEngine(cfg) as eng:
        res = eng.fit(req)
    s = S.FitSettings(max_depth=3, min_leaf=100, degree=2, seed=11, split_rows=1024,
                      max_iters=30, ridge=cfg.ridge, split_ridge=cfg.split_ridge,
                      min_side_fraction=cfg.min_side_fraction)
    nodes = S.fit_tree(req.X, req.y, None, s)
    got = {}
    for node in res.tree.iter_nodes():
        got[node.node_id] = node
    assert len(got) == len(nodes)
    for h, ref in nodes.items():
        nid = "/".join(["root"] + ["L" if b == "0" else "R" for b in bin(h)[3:]])
        node = got[nid]
        if "split" in ref:
            np.testing.assert_allclose(np.r_[node.w, node.z], np.r_[ref["split"]], rtol=1e-9,
                                       atol=1e-10)
        else:
            assert node.metadata["n_rows"] == len(ref["rows"])
            np.testing.assert_allclose(node.model.coef_, ref["coef"], rtol=1e-8, atol=1e-10)
    pred = res.tree.predict(req.X)
    assert np.isfinite(pred).all()


def test_results_are_the_same_alone_and_in_the_mix(torch_gpu):
    from art.serving import RefitRequest, ServingConfig, ServingEngine

    rng = np.random.default_rng(8)
    target = _request(rng, 20000, 4, seed=3, max_depth=5, min_leaf=64, tree_id="target")
    others = [_request(rng, n, d, seed=s, max_depth=4, min_leaf=64)
              for n, d, s in ((5000, 4, 1), (30000, 8, 2), (1500, 2, 4), (12000, 4, 5),
                              (8000, 16, 6))]
    # Sixteen more fits of the target's basis, so that its levels share
    # search launches of many nodes (a launch's size must not change a node's
    # arithmetic).
    others += [_request(rng, 9000, 4, seed=20 + i, max_depth=4, min_leaf=64)
               for i in range(16)]
    Xn = rng.uniform(-1, 1, (1500, 4))
    yn = np.cos(Xn.sum(axis=1))
    cfg = ServingConfig(split_rows=2048)
    with ServingEngine(cfg) as eng:
        alone = eng.fit(target)
        alone_refit = eng.submit(RefitRequest("target", Xn, yn)).result()
    with ServingEngine(cfg) as eng:
        futs = [eng.submit(others[0]), eng.submit(others[1])]
        mixed_f = eng.submit(target)
        futs += [eng.submit(o) for o in others[2:]]
        mixed = mixed_f.result()
        refit_f = eng.submit(RefitRequest("target", Xn, yn))
        futs.append(eng.submit(_request(rng, 4000, 4, seed=9, max_depth=3)))
        mixed_refit = refit_f.result()
        for f in futs:
            f.result()
    a, b = _tree_arrays(alone), _tree_arrays(mixed)
    assert [s[0] for s in a[0]] == [s[0] for s in b[0]] and len(a[0]) > 3
    for (_, w1, z1), (_, w2, z2) in zip(a[0], b[0]):
        np.testing.assert_array_equal(w1, w2)
        assert z1 == z2
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_array_equal(alone_refit.coefficients, mixed_refit.coefficients)


def test_mixed_precision_serving_equals_fp64_alone_and_in_the_mix(torch_gpu):
    """precision="mixed": split-bf16 statistics for open nodes, FP64 systems
    for leaves.  On bases it covers (d = 2, 4, 8 quadratic; 16 linear) and one
    it leaves in FP64 (d = 12 quadratic), with error thresholds that stop
    nodes by their MSE, every served tree equals the FP64 engine's (same
    splits, leaf models and kept systems, so equal refits), alone and in a
    mix; the small fit also matches the oracle."""
    from art.serving import RefitRequest, ServingConfig, ServingEngine

    rng = np.random.default_rng(12)
    reqs = [_request(rng, n, d, seed=s, max_depth=dep, min_leaf=64, error_threshold=thr,
                     tree_id=f"t{s}")
            for n, d, s, dep, thr in ((20000, 4, 1, 6, 2e-4), (6000, 2, 2, 5, 1e-4),
                                      (30000, 8, 3, 5, 5e-4), (9000, 16, 4, 4, 1e-3),
                                      (7000, 12, 5, 3, 0.0))]
    Xn = rng.uniform(-1, 1, (900, 4))
    yn = np.sin(Xn.sum(axis=1))
    base = dict(split_rows=2048)
    results = {}
    for prec in ("fp64", "mixed"):
        with ServingEngine(ServingConfig(precision=prec, **base)) as eng:
            futs = [eng.submit(r) for r in reqs]
            fits = [f.result() for f in futs]
            refit = eng.submit(RefitRequest("t1", Xn, yn)).result()
        results[prec] = (fits, refit)
    with ServingEngine(ServingConfig(precision="mixed", **base)) as eng:
        alone = eng.fit(reqs[0])
        alone_refit = eng.submit(RefitRequest("t1", Xn, yn)).result()
    stopped_by_mse = 0
    for req, a, b in zip(reqs, results["fp64"][0], results["mixed"][0]):
        ta, tb = _tree_arrays(a), _tree_arrays(b)
        assert [s[0] for s in ta[0]] == [s[0] for s in tb[0]]
        for (_, w1, z1), (_, w2, z2) in zip(ta[0], tb[0]):
            np.testing.assert_array_equal(w1, w2)
            assert z1 == z2
        np.testing.assert_array_equal(ta[1], tb[1])
        np.testing.assert_array_equal(a.status, b.status)
        for leaf in b.tree.iter_nodes():
            if hasattr(leaf, "w"):
                continue
            meta = leaf.metadata
            if (leaf.depth < req.max_depth and meta["n_rows"] >= 128
                    and meta["mse"] <= req.error_threshold):
                stopped_by_mse += 1
    assert stopped_by_mse > 0
    np.testing.assert_array_equal(results["fp64"][1].coefficients,
                                  results["mixed"][1].coefficients)
    ta, tb = _tree_arrays(alone), _tree_arrays(results["mixed"][0][0])
    assert [s[0] for s in ta[0]] == [s[0] for s in tb[0]] and len(ta[0]) > 3
    np.testing.assert_array_equal(ta[1], tb[1])
    np.testing.assert_array_equal(alone_refit.coefficients, results["mixed"][1].coefficients)

    small = _request(rng, 3000, 2, seed=11, max_depth=3, min_leaf=100)
    cfg = ServingConfig(precision="mixed", split_rows=1024, max_iters=30)
    with ServingEngine(cfg) as eng:
        res = eng.fit(small)
    s = S.FitSettings(max_depth=3, min_leaf=100, degree=2, seed=11, split_rows=1024,
                      max_iters=30, ridge=cfg.ridge, split_ridge=cfg.split_ridge,
                      min_side_fraction=cfg.min_side_fraction)
    nodes = S.fit_tree(small.X, small.y, None, s)
    got = {node.node_id: node for node in res.tree.iter_nodes()}
    assert len(got) == len(nodes)
    for h, ref in nodes.items():
        node = got["/".join(["root"] + ["L" if b == "0" else "R" for b in bin(h)[3:]])]
        if "split" in ref:
            np.testing.assert_allclose(np.r_[node.w, node.z], np.r_[ref["split"]], rtol=1e-9,
                                       atol=1e-10)
        else:
            np.testing.assert_allclose(node.model.coef_, ref["coef"], rtol=1e-8, atol=1e-10)


def test_refit_equals_the_fit_on_all_rows(torch_gpu):
    from art.serving import FitRequest, RefitRequest, ServingConfig, ServingEngine

    rng = np.random.default_rng(9)
    d = 3
    X = rng.uniform(-1, 1, (6000, d))
    y = np.sin(2 * X[:, 0]) + X[:, 1] * X[:, 2]
    w = rng.uniform(0.2, 1, 6000)
    X1, X2 = rng.uniform(-1, 1, (400, d)), rng.uniform(-1, 1, (300, d))
    y1, y2 = np.sin(2 * X1[:, 0]), np.cos(X2[:, 1])
    w2 = rng.uniform(0.2, 1, 300)
    with ServingEngine(ServingConfig(split_rows=2048)) as eng:
        fit = eng.fit(FitRequest(X, y, w, max_depth=3, min_leaf=200, seed=5, tree_id="t"))
        eng.submit(RefitRequest("t", X1, y1)).result()
        res = eng.submit(RefitRequest("t", X2, y2, w2)).result()
    flat = res.tree.flatten()
    Xa = np.vstack([X, X1, X2])
    ya = np.concatenate([y, y1, y2])
    wa = np.concatenate([w, np.ones(400), w2])
    import torch

    leaf = flat.leaves(torch.as_tensor(Xa, device="cuda")).cpu().numpy()
    assert fit.coefficients.shape == res.coefficients.shape
    for j in range(flat.n_leaves):
        idx = np.flatnonzero(leaf == j)
        phi = R.features(Xa[idx], 2, True, False)
        ref, _ = R.solve_weighted_ridge(phi, ya[idx], wa[idx], 1e-6)
        err = np.linalg.norm(res.coefficients[j] - ref) / np.linalg.norm(ref)
        assert err < 1e-9, (j, err)
    host = np.array([res.tree.leaf_for_point(x).predict_one(x) for x in X2[:50]])
    np.testing.assert_allclose(res.tree.predict(X2[:50]), host, rtol=1e-12, atol=1e-12)


def test_admission_rejects_oversized_requests_and_serves_within_budget(torch_gpu):
    from art.serving import AdmissionError, FitRequest, ServingConfig, ServingEngine

    rng = np.random.default_rng(10)
    small = [FitRequest(rng.uniform(-1, 1, (2000, 2)), rng.normal(size=2000), max_depth=2,
                        seed=i) for i in range(4)]
    big = FitRequest(rng.uniform(-1, 1, (200000, 8)), rng.normal(size=200000))
    cfg = ServingConfig(hbm_budget_bytes=2e6, max_active=1)
    with ServingEngine(cfg) as eng:
        futs = [eng.submit(r) for r in small]
        with pytest.raises(AdmissionError):
            eng.submit(big).result()
        res = [f.result() for f in futs]
    assert all(r.tree.num_leaves() >= 1 for r in res)
    # one slot: each request was admitted after the previous one finished
    for a, b in zip(res, res[1:]):
        assert b.timing.admitted >= a.timing.finished
    with ServingEngine() as eng:
        with pytest.raises(KeyError):
            eng.submit(__import__("art.serving", fromlist=["RefitRequest"]).RefitRequest(
                "missing", np.zeros((1, 2)), np.zeros(1))).result()
