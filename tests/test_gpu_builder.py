"""Device parity of the tree-level paths: batched inference, temperature
scales, cohort node fits, and whole builds through the batched optimizer."""

from dataclasses import dataclass, field

import numpy as np
import pytest

from oracle import ridge as R

pytestmark = pytest.mark.gpu


def _random_tree(rng, d, depth, make_model):
    from art.tree import LeafNode, SplitNode

    def node(level, tag):
        if level == depth:
            return LeafNode(make_model(), depth=level, node_id=tag)
        w = rng.normal(size=d)
        w /= np.linalg.norm(w)
        return SplitNode(w, float(rng.uniform(-0.3, 0.3)), node(level + 1, tag + "/L"),
                         node(level + 1, tag + "/R"), depth=level, node_id=tag)

    return node(0, "root")


@pytest.mark.parametrize("kind,d", [("quad", 2), ("quad", 5), ("quad", 12), ("affine", 3),
                                    ("affine", 20), ("linear_bias_first", 7), ("quad", 16)])
def test_tree_predict_matches_host_routing(torch_gpu, kind, d):
    from art.models import AffineRidgeModel, PolynomialRidgeModel
    from art.tree import RegressionTree

    rng = np.random.default_rng(d * 7 + len(kind))

    def make():
        m = {"quad": lambda: PolynomialRidgeModel(2, ridge=1e-6),
             "affine": lambda: AffineRidgeModel(1e-6),
             "linear_bias_first": lambda: PolynomialRidgeModel(1, ridge=1e-6)}[kind]()
        p = len(m.design(np.zeros((1, d)))[0])
        m.coef_ = rng.normal(size=p)
        return m

    tree = RegressionTree(_random_tree(rng, d, 4, make))
    X = rng.uniform(-1, 1, (5000, d))
    gpu = tree.predict(X)
    host = tree._predict_models(X)
    # Points within 1e-12 of a hyperplane may route differently (summation order).
    margin = np.min(np.abs(np.stack([X @ n.w - n.z for n in tree.iter_nodes()
                                     if hasattr(n, "w")])), axis=0)
    ok = margin > 1e-12
    assert ok.mean() > 0.99
    np.testing.assert_allclose(gpu[ok], host[ok], rtol=1e-12, atol=1e-12)


def test_single_leaf_tree_predicts_its_model(torch_gpu):
    from art.domain import BoxDomain
    from art.models import PolynomialRidgeModel
    from art.tree import LeafNode, RegressionTree

    m = PolynomialRidgeModel(2, ridge=0.0)
    m.coef_ = np.arange(6.0)
    tree = RegressionTree(LeafNode(m, region=BoxDomain.hypercube(2, -1., 1.).as_region()))
    X = np.random.default_rng(0).uniform(-1, 1, (100, 2))
    np.testing.assert_allclose(tree.predict(X), m.predict(X), rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("n,d,dupes", [(2, 1, False), (3, 2, False), (100, 2, False),
                                       (511, 7, False), (512, 3, False), (64, 2, True),
                                       (4096, 2, False)])
def test_nearest_neighbour_median_matches_scipy(torch_gpu, n, d, dupes):
    from scipy.spatial import cKDTree

    from art.temperature import set_distance_medians

    rng = np.random.default_rng(n + d)
    X = rng.integers(0, 4, (n, d)).astype(float) if dupes else rng.uniform(-1, 1, (n, d))
    Y = rng.uniform(-2, 2, (max(n // 2, 2), d))
    got = set_distance_medians([X, Y], 0)
    for P, g in zip((X, Y), got):
        ref = float(np.median(cKDTree(P).query(P, k=2)[0][:, 1]))
        assert g == pytest.approx(ref, rel=1e-12, abs=1e-15)


@pytest.mark.parametrize("n,d", [(2, 1), (5, 3), (64, 2), (301, 4)])
def test_pairwise_median_matches_numpy(torch_gpu, n, d):
    from art.temperature import pairwise_distances, set_distance_medians

    X = np.random.default_rng(n).normal(size=(n, d))
    ref = float(np.median(pairwise_distances(X)[np.triu_indices(n, k=1)]))
    assert set_distance_medians([X], 1)[0] == pytest.approx(ref, rel=1e-12)


def test_estimate_temperature_follows_the_reference_formula(torch_gpu):
    from scipy.spatial import cKDTree

    from art.temperature import estimate_temperature, estimate_temperatures, subsample_points

    rng = np.random.default_rng(4)
    X = rng.uniform(-1, 1, (2000, 3))
    used = subsample_points(X, 512, random_state=9)
    ref = 0.3 * float(np.median(cKDTree(used).query(used, k=2)[0][:, 1]))
    assert estimate_temperature(X, "median_nn", c=0.3, random_state=9) == pytest.approx(ref,
                                                                                        rel=1e-12)
    pw = float(np.median(np.linalg.norm(used[:, None] - used[None], axis=2)[
        np.triu_indices(512, k=1)])) * 512 ** (-1 / 3)
    assert estimate_temperature(X, "median_pairwise_scaled", c=2.0, random_state=9) == \
        pytest.approx(2.0 * pw, rel=1e-12)
    many = estimate_temperatures([X, X[:300]], "median_nn", 0.3, 512, [9, 9])
    assert many[0] == pytest.approx(ref, rel=1e-12)


# --------------------------------------------------------------- builds


@dataclass
class HostRidge:
    """Test double: the oracle's FP64 normal-equation ridge on host arrays."""

    ridge: float = 1e-8
    degree: int = 2
    coef_: np.ndarray = field(default=None, repr=False)

    def clone(self):
        return HostRidge(self.ridge, self.degree)

    def effective_dimension(self, d):
        return len(R.feature_columns(d, self.degree, True, False)[0])

    def fit(self, X, y):
        phi = R.features(X, self.degree, True, False)
        self.coef_, _ = R.solve_weighted_ridge(phi, y, np.ones(len(y)), self.ridge)
        return self

    def fit_weighted(self, X, y, weights, weight_floor=1e-12):
        phi = R.features(X, self.degree, True, False)
        self.coef_, _ = R.solve_weighted_ridge(phi, y, weights, self.ridge, weight_floor)
        return self

    def predict(self, X):
        return R.features(np.atleast_2d(X), self.degree, True, False) @ self.coef_


@dataclass
class MedianAxisSplitter:
    """Test double: split at the median of the first coordinate."""

    random_state: object = None

    def split(self, X, y, parent_model=None, parent_loss=None, prepared_design=None, **kw):
        from art.splitters import SplitResult

        w = np.zeros(X.shape[1])
        w[0] = 1.0
        z = float(np.median(X[:, 0]))
        right = X[:, 0] - z >= 0.0
        return SplitResult(w, z, None, None, None, float(parent_loss), None, None,
                           int((~right).sum()), int(right.sum()), True, 1, "converged", {})


def _target(X):
    return np.sin(3 * X[..., 0]) * np.cos(2 * X[..., 1]) + 0.5 * np.exp(
        -8 * ((X[..., 0] - .3) ** 2 + (X[..., 1] + .2) ** 2))


@pytest.mark.parametrize("traversal", ["bfs", "dfs"])
def test_cohort_fits_match_per_node_host_fits(torch_gpu, traversal):
    # The same build with the GPU ridge template (cohort fits in one launch)
    # and with a host FP64 double of the same fit: identical trees, leaf
    # coefficients equal to 1e-9.
    from art.builder import RegressionTreeBuilder
    from art.domain import BoxDomain
    from art.models import PolynomialRidgeModel
    from art.sampling import HitAndRunSampler

    def build(model):
        return RegressionTreeBuilder(
            domain=BoxDomain.hypercube(2, -1.0, 1.0), oracle=_target, oracle_vectorized=True,
            model_template=model, splitter=MedianAxisSplitter(), error_threshold=0.05,
            sample_count=120, max_depth=4, sampler=HitAndRunSampler(burn_in=0, thinning=3),
            isotropic_sampling=False, random_state=5, traversal=traversal, node_batch_size=6,
            store_samples=True).build()

    gpu = build(PolynomialRidgeModel(2, ridge=1e-8))
    host = build(HostRidge(1e-8))
    a, b = list(gpu.tree.iter_nodes()), list(host.tree.iter_nodes())
    assert [n.node_id for n in a] == [n.node_id for n in b]
    assert [n.status for n in a] == [n.status for n in b]
    assert gpu.oracle_queries == host.oracle_queries
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u.metadata["X"], v.metadata["X"])
        assert u.metadata["fit_error"] == pytest.approx(v.metadata["fit_error"], rel=1e-9,
                                                        abs=1e-12)
    for u, v in zip(gpu.tree.iter_leaves(), host.tree.iter_leaves()):
        # Deep leaves cover small boxes, where the quadratic basis is nearly
        # collinear (condition ~1e9 at ridge 1e-8): two correct FP64 solves
        # agree on the fitted values to ~1e-11 (targets are O(1)) and on the
        # coefficients to ~1e-8.
        X = u.metadata["X"]
        np.testing.assert_allclose(u.model.predict(X), v.model.predict(X), rtol=1e-9,
                                   atol=1e-10)
        np.testing.assert_allclose(u.model.coef_, v.model.coef_, rtol=1e-6, atol=1e-9)
    assert gpu.tree.num_leaves() > 4


def test_batched_bfs_build_end_to_end(torch_gpu):
    from art.builder import RegressionTreeBuilder
    from art.domain import BoxDomain
    from art.models import AffineRidgeModel
    from art.sampling import HitAndRunSampler
    from art.splitters import SoftObliqueSplitter
    from art.temperature import TemperatureConfig
    from art.tree import SplitNode

    def oracle(X):
        return np.where(X[..., 0] + 0.2 * X[..., 1] >= 0.1, 1.0 + X[..., 1], -1.0 + 0.5 * X[..., 0])

    # 8 restarts (so every split is hard-refitted and the best restart kept).
    s = SoftObliqueSplitter(AffineRidgeModel(1e-8), 0.05, max_iters=60, min_side_points=8,
                            n_restarts=8, optimizer="cupy_batched_cholesky")
    b = RegressionTreeBuilder(
        domain=BoxDomain.hypercube(2, -1.0, 1.0), oracle=oracle, oracle_vectorized=True,
        model_template=AffineRidgeModel(1e-8), splitter=s, error_threshold=1e-3,
        sample_count=256, max_depth=3, sampler=HitAndRunSampler(burn_in=0, thinning=5),
        temperature_config=TemperatureConfig(strategy="fixed", c=2.0), random_state=3,
        traversal="bfs", node_batch_size=4, store_samples=True)
    result = b.build()
    tree = result.tree
    root = tree.root
    assert isinstance(root, SplitNode)
    # The root split finds the discontinuity x0 + 0.2 x1 = 0.1.
    w = root.w / np.linalg.norm(root.w)
    truth = np.array([1.0, 0.2]) / np.hypot(1.0, 0.2)
    assert abs(w @ truth) > 0.98
    assert root.metadata["temperature_c"] == 2.0 and root.metadata["temperature"] > 0
    assert root.metadata["hard_refit"] and root.metadata["split_gain_mse"] > 0
    total = 0
    for node in tree.iter_nodes():
        X = node.metadata["X"]
        assert len(X) == 256 and node.region.contains(X).all()
        total += node.metadata["n_new"]
    assert total == result.oracle_queries
    X = np.random.default_rng(1).uniform(-1, 1, (4000, 2))
    gpu, host = tree.predict(X), tree._predict_models(X)
    np.testing.assert_allclose(gpu, host, rtol=1e-10, atol=1e-12)
    err = np.sqrt(np.mean((gpu - oracle(X)) ** 2))
    assert err < 0.3
    assert tree.metadata["traversal"] == "bfs" and tree.metadata["node_batch_size"] == 4


def test_many_chain_requests_equal_their_single_calls(torch_gpu):
    # One launch for several regions, lengths, seeds and shapes gives each
    # request exactly the samples of its own HitAndRunSampler.sample call.
    from art.domain import BoxDomain, split_region
    from art.sampling import HitAndRunSampler, sample_hit_and_run_many

    box = BoxDomain.hypercube(3, -1.0, 1.0).as_region()
    left, right = split_region(box, np.array([0.6, 0.8, 0.0]), 0.1)
    lower, _ = split_region(left, np.array([0.0, 0.3, -1.0]) / np.hypot(0.3, 1.0), -0.2)
    regions = [box, left, right, lower]
    rng = np.random.default_rng(3)
    V = np.linalg.qr(rng.normal(size=(3, 3)))[0]
    samplers = [HitAndRunSampler(burn_in=0, thinning=7),
                HitAndRunSampler(burn_in=0, thinning=7, direction_eigenvectors=V,
                                 direction_eigenvalues=np.array([1.0, 0.2, 0.05])),
                HitAndRunSampler(burn_in=0, thinning=7),
                HitAndRunSampler(burn_in=0, thinning=7, direction_eigenvectors=np.eye(3),
                                 direction_eigenvalues=np.array([0.3, 1.0, 0.6]))]
    starts = [np.zeros(3), np.array([-0.5, -0.2, 0.0]), np.array([0.6, 0.5, 0.1]),
              np.array([-0.6, -0.4, 0.5])]
    ns, seeds = [40, 13, 0, 77], [11, 12, 13, 14]
    got = sample_hit_and_run_many(samplers, regions, ns, seeds, starts)
    for s, r, n, seed, x0, g in zip(samplers, regions, ns, seeds, starts, got):
        np.testing.assert_array_equal(g, s.sample(r, n, random_state=seed, x0=x0))
        assert r.contains(g).all()


@pytest.mark.parametrize("isotropic", [True, False])
def test_cohort_top_ups_equal_node_by_node_top_ups(torch_gpu, isotropic):
    from art.builder import RegressionTreeBuilder
    from art.domain import BoxDomain
    from art.models import PolynomialRidgeModel
    from art.sampling import HitAndRunSampler

    class NodeByNode(RegressionTreeBuilder):
        def _top_up_cohort(self, jobs):
            return [self._top_up_one(w) for w in jobs]

    def build(cls):
        return cls(domain=BoxDomain.hypercube(2, -1.0, 1.0), oracle=_target,
                   oracle_vectorized=True, model_template=PolynomialRidgeModel(2, ridge=1e-8),
                   splitter=MedianAxisSplitter(), error_threshold=0.02, sample_count=150,
                   max_depth=4, sampler=HitAndRunSampler(burn_in=0, thinning=4),
                   isotropic_sampling=isotropic, random_state=21, traversal="bfs",
                   node_batch_size=8, store_samples=True).build()

    a, b = build(RegressionTreeBuilder), build(NodeByNode)
    na, nb = list(a.tree.iter_nodes()), list(b.tree.iter_nodes())
    assert [n.node_id for n in na] == [n.node_id for n in nb] and len(na) > 7
    assert a.oracle_queries == b.oracle_queries
    for u, v in zip(na, nb):
        np.testing.assert_array_equal(u.metadata["X"], v.metadata["X"])
        assert u.metadata["sampling_method"] == v.metadata["sampling_method"]


@pytest.mark.parametrize("d", [2, 9, 20])
def test_tree_leaf_ids_match_host_routing(torch_gpu, d):
    from art.models import AffineRidgeModel
    from art.tree import RegressionTree

    torch, rng = torch_gpu, np.random.default_rng(400 + d)

    def make():
        m = AffineRidgeModel(1e-6)
        m.coef_ = rng.normal(size=d + 1)
        return m

    tree = RegressionTree(_random_tree(rng, d, 4, make))
    flat = tree.flatten()
    X = rng.uniform(-1, 1, (3000, d))
    ids = flat.leaves(torch.as_tensor(X, device="cuda")).cpu().numpy()
    host = np.array([next(i for i, leaf in enumerate(flat.leaf_nodes)
                          if leaf is tree.leaf_for_point(x)) for x in X])
    margin = np.min(np.abs(np.stack([X @ n.w - n.z for n in tree.iter_nodes()
                                     if hasattr(n, "w")])), axis=0)
    ok = margin > 1e-12
    np.testing.assert_array_equal(ids[ok], host[ok])


# --------------------------------------------------------------------------- plain ART fit

@pytest.fixture(scope="module")
def plain_fit_builds(torch_gpu):
    """The plain ART fit at its full size (d = 2, 4000 samples per node, depth
    <= 6, soft oblique T = 0.05 on the batched device optimizer, quadratic
    ridge leaves at 1e-6, 256 samples per leaf polytope), built with the GPU
    ridge template and again with the host FP64 ridge double."""
    from art.synthetic import plain_fit_builder

    gpu = plain_fit_builder(seed=0, store_samples=True, store_diagnostics=True).build()
    host = plain_fit_builder(seed=0, store_samples=True, store_diagnostics=True,
                             model_template=HostRidge(1e-6)).build()
    return gpu, host


def test_plain_fit_build_properties(plain_fit_builds):
    from art.sampling import PolytopeBatch  # noqa: F401  (device sampler is loaded)
    from art.synthetic import PlainFitOracle

    gpu, _ = plain_fit_builds
    tree = gpu.tree
    assert tree.max_depth() <= 6 and tree.num_leaves() >= 8
    total_new = 0
    for node in tree.iter_nodes():
        X = node.metadata["X"]
        assert X.shape == (4000, 2)
        assert node.region.contains(X).all(), node.node_id
        total_new += node.metadata["n_new"]
    assert total_new == gpu.oracle_queries
    # 256 hit-and-run samples in every leaf polytope, inside it.
    for leaf in tree.iter_leaves():
        pts = leaf_samples(leaf, 256)
        assert pts.shape == (256, 2) and leaf.region.contains(pts, tol=1e-12).all()
    Xt = np.random.default_rng(9).uniform(-1, 1, (20000, 2))
    pred = tree.predict(Xt)
    np.testing.assert_allclose(pred, tree._predict_models(Xt), rtol=1e-12, atol=1e-12)
    # Every leaf stops at a relative L2 error of 1e-2 on its own samples or
    # at depth 6; over the square the fit stays within 2e-2 (RMS) of the
    # noise-free target, whose RMS is ~0.45.
    rmse = float(np.sqrt(np.mean((pred - PlainFitOracle.clean(Xt)) ** 2)))
    assert rmse < 2e-2, rmse


def leaf_samples(leaf, n):
    from art.sampling import HitAndRunSampler

    x0 = leaf.metadata["X"][0]
    return HitAndRunSampler(burn_in=20, thinning=5).sample(leaf.region, n, random_state=1, x0=x0)


def test_plain_fit_build_structure_matches_the_host_ridge_build(plain_fit_builds):
    gpu, host = plain_fit_builds
    a, b = list(gpu.tree.iter_nodes()), list(host.tree.iter_nodes())
    assert [n.node_id for n in a] == [n.node_id for n in b]
    assert [n.status for n in a] == [n.status for n in b]
    assert gpu.oracle_queries == host.oracle_queries
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u.metadata["X"], v.metadata["X"])
        if hasattr(u, "w"):
            np.testing.assert_array_equal(u.w, v.w)
            assert u.z == v.z
    for u, v in zip(gpu.tree.iter_leaves(), host.tree.iter_leaves()):
        X = u.metadata["X"]
        np.testing.assert_allclose(u.model.predict(X), v.model.predict(X), rtol=1e-9,
                                   atol=1e-10)


def test_plain_fit_build_replays_on_the_oracle(plain_fit_builds):
    """Every decision of the GPU build, replayed on the FP64 oracle from the
    build's own node samples and device starts: each split's final theta (the
    oracle's optimizer from the recorded initial theta, art/splitters.py:365-601)
    and its partition of the node's samples (rows within 1e-9 of the
    hyperplane excepted); each node fit (the oracle's normal-equation ridge,
    art/models.py:218-343) on its fitted values to 1e-9; each node's stop
    decision from the oracle's fit error.

    Deep nodes run 100 iterations at steps up to alpha_max = 1e3 on losses of
    ~1e-5, where the iteration map amplifies rounding: the oracle itself,
    started 1e-12 away, ends up to ~1e-3 away.  A split must therefore agree
    to 1e-7, or, when the oracle's own spread under that perturbation
    exceeds 1e-7, to within 10x that spread (with its histories equal up to
    where the spread appears); such sensitive splits must stay rare."""
    from art.metrics import relative_l2_error
    from oracle import split as O

    gpu, _ = plain_fit_builds
    s = O.Settings(temperature=0.05, ridge=1e-6, min_side_points=64, degree=2,
                   bias_last=False)
    splits, sensitive = 0, []
    for node in gpu.tree.iter_nodes():
        X, y = node.metadata["X"], node.metadata["y"]
        phi = R.features(X, 2, True, False)
        beta, _ = R.solve_weighted_ridge(phi, y, np.ones(len(y)), 1e-6)
        err = relative_l2_error(y, phi @ beta)
        assert node.metadata["fit_error"] == pytest.approx(err, rel=1e-8, abs=1e-12)
        if node.status == "tolerance_met":
            assert err <= 1e-2
        if not hasattr(node, "w"):
            if node.model is not None:
                np.testing.assert_allclose(node.model.predict(X), phi @ beta, rtol=1e-9,
                                           atol=1e-10)
            continue
        assert err > 1e-2 and node.depth < 6
        meta = node.metadata["splitter_metadata"]
        theta0 = meta["initial_theta"]
        ref = O.optimize(X, y, theta0, s)
        theta = np.append(node.w, node.z)
        assert ref["stop_reason"] == node.metadata["split_stop_reason"], node.node_id
        assert node.metadata["split_n_right"] == int(np.count_nonzero(X @ node.w - node.z >= 0))
        splits += 1
        gap = np.max(np.abs(theta - ref["theta"]))
        if gap <= 1e-7:
            u_gpu, u_ref = X @ node.w - node.z, X @ ref["theta"][:-1] - ref["theta"][-1]
            clear = np.abs(u_ref) > 1e-9
            np.testing.assert_array_equal((u_gpu >= 0)[clear], (u_ref >= 0)[clear])
            continue
        # The oracle's own spread from starts 1e-12 away (w rotated both ways).
        spread = 0.0
        for sign in (1.0, -1.0):
            nudged = theta0.copy()
            nudged[:-1] += sign * 1e-12 * np.array([-theta0[1], theta0[0]])
            spread = max(spread, np.max(np.abs(O.optimize(X, y, nudged, s)["theta"]
                                               - ref["theta"])))
        assert spread > 1e-7 and gap <= 10 * spread, (node.node_id, gap, spread)
        # Up to where the spread appears, the histories agree.
        lg, lr = np.array(meta["soft_loss_history"]), np.array(ref["soft_loss_history"])
        np.testing.assert_allclose(lg[:20], lr[:20], rtol=1e-9)
        sensitive.append(node.node_id)
    assert splits >= 4
    assert len(sensitive) <= max(1, splits // 10), sensitive
