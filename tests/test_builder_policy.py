"""Builder policy known answers, on the host (no device work).

Restates the known answers of the reference's tests/test_builder.py:369-777:
oracle query counts, node ids, validation refit sizes, retries with the same
samples and a new seed, the min-side fallback on the last allowed attempt,
gain-check toggles, warnings and storage toggles, through both traversals.

The product's builder policy runs unchanged.  Two test doubles keep the
device out: ``HostAffine`` (an FP64 affine ridge fit in NumPy, the model the
reference's tests fit with AffineRidgeModel) and ``HostHitAndRun`` (a
HitAndRunSampler whose ``sample`` steps chains with the oracle's chord
interval).  A subclass that overrides ``sample`` is called node by node by
the builder, as any sampler other than the device sampler is.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import pytest

from art.builder import RegressionTreeBuilder
from art.domain import BoxDomain
from art.metrics import max_pointwise_relative_error, mean_squared_error
from art.sampling import HitAndRunSampler
from art.splitters import SoftObliqueSplitter, SplitNotFoundError, SplitResult
from art.temperature import TemperatureConfig
from art.tree import LeafNode, SplitNode
from oracle import sampling as S


class HostAffine:
    """Affine ridge double [x, 1] . coef with the trace-scaled ridge; records
    the row count of every unweighted fit in ``fit_sizes``."""

    def __init__(self, ridge=1e-12, fit_sizes=None):
        self.ridge = ridge
        self.fit_sizes = fit_sizes
        self.coef_ = None
        self.solve_result_ = None

    @staticmethod
    def _design(X):
        X = np.asarray(X, dtype=float)
        return np.hstack([X, np.ones((X.shape[0], 1))])

    def fit(self, X, y):
        if self.fit_sizes is not None:
            self.fit_sizes.append(np.asarray(X).shape[0])
        return self.fit_weighted(X, y, np.ones(np.asarray(X).shape[0]))

    def fit_weighted(self, X, y, weights, weight_floor=1e-12):
        phi = self._design(X)
        w = np.maximum(np.asarray(weights, dtype=float), weight_floor)
        G = phi.T @ (phi * w[:, None])
        if self.ridge > 0.0:
            G = G + self.ridge * np.trace(G) / G.shape[0] * np.eye(G.shape[0])
        self.coef_ = np.linalg.lstsq(G, phi.T @ (w * np.asarray(y, dtype=float)), rcond=None)[0]
        return self

    def predict(self, X):
        return self._design(X) @ self.coef_

    def clone(self):
        return HostAffine(self.ridge, self.fit_sizes)

    def effective_dimension(self, input_dimension):
        return int(input_dimension) + 1


@dataclass
class HostHitAndRun(HitAndRunSampler):
    """Hit-and-run on the host: Gaussian directions (shaped by the given
    eigenvectors and floored eigenvalues), the oracle's chord interval and a
    uniform step; records after moves burn_in + 1, burn_in + 1 + thinning..."""

    def sample(self, region, n, random_state=None, x0=None):
        rng = np.random.default_rng(random_state)
        if x0 is None:
            raise RuntimeError("the host double needs a start point")
        x = np.array(x0, dtype=float)
        d = x.size
        shape = np.eye(d)
        if self.direction_eigenvectors is not None:
            lam = np.maximum(np.asarray(self.direction_eigenvalues, dtype=float),
                             self.direction_eigenvalue_floor * np.max(self.direction_eigenvalues))
            shape = np.asarray(self.direction_eigenvectors) * np.sqrt(lam)[None, :]
        out = []
        for move in range(self.burn_in + n * self.thinning):
            direction = shape @ rng.normal(size=d)
            lo, hi = S.step_interval(region.A, region.b, x, direction)
            x = x + direction * (lo + rng.random() * (hi - lo))
            if move >= self.burn_in and (move - self.burn_in) % self.thinning == 0:
                out.append(x.copy())
        return np.array(out).reshape(n, d)


@pytest.fixture(params=["dfs", "bfs"])
def traversal(request):
    return request.param


@dataclass
class FixedAxisSplitter:
    """Splits at x_0 = 0 (or its median) and fits both sides exactly."""

    ridge: float = 1e-12
    converged: bool = True
    n_iters: int = 1
    stop_reason: str = "converged"
    metadata: dict = field(default_factory=dict)
    random_state: object = None
    median: bool = False

    def split(self, X, y, parent_model=None, parent_loss=None, prepared_design=None):
        w = np.zeros(X.shape[1])
        w[0] = 1.0
        z = float(np.median(X[:, 0])) if self.median else 0.0
        right = (X @ w - z) >= 0.0
        fits = [HostAffine(self.ridge).fit(X[m], y[m]) for m in (~right, right)]
        loss = sum(np.sum(m) / X.shape[0] * mean_squared_error(y[m], f.predict(X[m]))
                   for m, f in zip((~right, right), fits))
        gain = float(parent_loss - loss)
        return SplitResult(w=w, z=z, left_model=fits[0], right_model=fits[1], loss=loss,
                           parent_loss=float(parent_loss), split_gain=gain,
                           relative_split_gain=gain / max(float(parent_loss), 1e-12),
                           n_left=int(np.sum(~right)), n_right=int(np.sum(right)),
                           converged=self.converged, n_iters=self.n_iters,
                           stop_reason=self.stop_reason, metadata=dict(self.metadata))


class ScriptSplitter:
    """Scripted outcomes shared across the builder's deep copies: attempt k
    gives outcomes[min(k, last)]: "min_side_points" raises, an
    "insufficient_split_gain" result has zero gain, anything else succeeds."""

    def __init__(self, outcomes, state=None, n_restarts=1):
        self.outcomes, self.n_restarts = tuple(outcomes), n_restarts
        self.state = {"calls": 0, "X": [], "y": [], "seeds": []} if state is None else state
        self.random_state = None

    def __deepcopy__(self, memo):
        twin = type(self)(self.outcomes, self.state, self.n_restarts)
        twin.random_state = self.random_state
        return twin

    def split(self, X, y, parent_model=None, parent_loss=None, prepared_design=None):
        self.state["calls"] += 1
        call = self.state["calls"]
        self.state["X"].append(np.array(X))
        self.state["y"].append(np.array(y))
        self.state["seeds"].append(self.random_state)
        outcome = self.outcomes[min(call, len(self.outcomes)) - 1]
        diagnostics = {"run": call}
        if outcome == "min_side_points":
            raise SplitNotFoundError("min_side_points", "scripted", diagnostics=diagnostics)
        result = FixedAxisSplitter().split(X, y, parent_model, parent_loss)
        result.metadata = diagnostics
        if outcome == "insufficient_split_gain":
            result.loss, result.split_gain, result.relative_split_gain = float(parent_loss), 0., 0.
        return result


class FallbackScriptSplitter(ScriptSplitter):
    """Fails on min-side points unless the builder enabled the fallback."""

    def __init__(self, state=None):
        super().__init__(("min_side_points",), state)
        self.min_side_fallback = True

    def __deepcopy__(self, memo):
        twin = type(self)(self.state)
        twin.random_state, twin.min_side_fallback = self.random_state, self.min_side_fallback
        return twin

    def split(self, X, y, parent_model=None, parent_loss=None, prepared_design=None):
        self.state["calls"] += 1
        self.state["seeds"].append(self.random_state)
        if not self.min_side_fallback:
            raise SplitNotFoundError("min_side_points", "scripted",
                                     diagnostics={"run": self.state["calls"]})
        result = FixedAxisSplitter().split(X, y, parent_model, parent_loss)
        result.metadata = {"run": self.state["calls"], "min_side_fallback": True,
                           "fallback_iteration": 0}
        return result


class RecordingSoftSplitter(SoftObliqueSplitter):
    """Records the fit sizes of the parent model at split time."""

    def __init__(self, model_template, state=None, fail=False):
        super().__init__(model_template=model_template, temperature=0.1, min_side_points=1)
        self.state = {} if state is None else state
        self.fail = fail

    def __deepcopy__(self, memo):
        twin = type(self)(self.model_template.clone(), self.state, self.fail)
        twin.random_state = self.random_state
        return twin

    def split(self, X, y, parent_model=None, parent_loss=None, prepared_design=None, *,
              hard_refit=True, **_):
        self.state["fit_sizes_at_split"] = tuple(parent_model.fit_sizes)
        if self.fail:
            raise SplitNotFoundError("invalid_split", "scripted split failure")
        return FixedAxisSplitter().split(X, y, parent_model, parent_loss)


def piecewise(x):
    return x[0] - 0.5 * x[1] if x[0] < 0.0 else 3.0 - 0.25 * x[0] + x[1]


def square():
    return BoxDomain(np.array([[-1.0, 1.0], [-1.0, 1.0]]))


def retry_builder(splitter, traversal, **kw):
    kw.setdefault("check_split_gain", True)
    kw.setdefault("model_template", HostAffine(1e-12))
    return RegressionTreeBuilder(domain=square(), oracle=piecewise, splitter=splitter,
                                 error_threshold=0.0, sample_count=30, max_depth=1,
                                 sampler=HostHitAndRun(burn_in=0, thinning=1),
                                 store_diagnostics=True, random_state=3, traversal=traversal,
                                 **kw)


def test_root_within_tolerance_counts_only_its_samples(traversal):
    result = RegressionTreeBuilder(
        domain=square(), oracle=lambda x: 1.5 * x[0] - 0.25 * x[1] + 0.75,
        model_template=HostAffine(1e-12), splitter=FixedAxisSplitter(), error_threshold=1e-8,
        sample_count=20, max_depth=3, random_state=2, sampler=HostHitAndRun(burn_in=0),
        traversal=traversal).build()
    root = result.tree.root
    assert isinstance(root, LeafNode) and root.status == "tolerance_met"
    assert result.oracle_queries == 20 and result.tree.oracle_queries == 20
    assert not {"X", "y", "sample_centroid"} & set(root.metadata)
    assert result.build_timing is None and result.tree.metadata["build_timing"] is None


def test_validation_needs_the_soft_splitter_and_the_max_metric():
    common = dict(domain=BoxDomain(np.array([[-1.0, 1.0]])), oracle=lambda x: x[0],
                  model_template=HostAffine(), error_threshold=1e-2, sample_count=10,
                  validation_fraction=0.2, sampler=HostHitAndRun())
    with pytest.raises(ValueError, match="only by SoftObliqueSplitter"):
        RegressionTreeBuilder(splitter=FixedAxisSplitter(), **common)
    soft = SoftObliqueSplitter(model_template=HostAffine(), temperature=0.1, min_side_points=1)
    with pytest.raises(ValueError, match="requires max_pointwise_relative_error"):
        RegressionTreeBuilder(splitter=soft, **common)


def _validating(traversal, fit_sizes, state, fail=False, oracle=None, **kw):
    model = HostAffine(0.0, fit_sizes)
    return RegressionTreeBuilder(
        domain=square(), oracle=oracle or piecewise, model_template=model,
        splitter=RecordingSoftSplitter(model.clone(), state=state, fail=fail),
        error_metric=max_pointwise_relative_error, sample_count=20, validation_fraction=0.2,
        random_state=5, sampler=HostHitAndRun(burn_in=0, thinning=1), traversal=traversal,
        # The splitter's own temperature: the default kNN temperature scale
        # runs on the device, and these answers do not depend on it.
        temperature_config=TemperatureConfig(strategy="splitter"), **kw)


def test_passing_validation_refits_the_leaf_on_all_samples(traversal):
    sizes = []
    b = _validating(traversal, sizes, {}, oracle=lambda x: 2.0 + x[0] - 0.5 * x[1],
                    error_threshold=1e-10)
    root = b.build().tree.root
    assert isinstance(root, LeafNode) and root.status == "tolerance_met"
    assert sizes == [16, 20]
    assert root.metadata["convergence_fit_samples"] == 16
    assert root.metadata["convergence_error"] <= 1e-10 and root.metadata["fit_error"] <= 1e-10


def test_failed_validation_splits_without_a_full_refit(traversal):
    sizes, state = [], {}
    root = _validating(traversal, sizes, state, error_threshold=0.0, max_depth=1).build().tree.root
    assert isinstance(root, SplitNode)
    assert state["fit_sizes_at_split"] == (16,)
    assert root.metadata["convergence_fit_samples"] == 16


def test_split_failure_refits_the_leaf_on_all_samples(traversal):
    sizes, state = [], {}
    root = _validating(traversal, sizes, state, fail=True, error_threshold=0.0,
                       max_depth=1).build().tree.root
    assert isinstance(root, LeafNode) and root.status == "invalid_split"
    assert state["fit_sizes_at_split"] == (16,)
    assert sizes == [16, 20]
    assert root.metadata["convergence_fit_samples"] == 16


def test_build_timing_splits_sampling_and_splitter(traversal):
    result = retry_builder(FixedAxisSplitter(), traversal, profile_build_timing=True).build()
    t = result.build_timing
    assert t is not None and result.tree.metadata["build_timing"] == t
    assert t["total_build_seconds"] > 0 and t["sampling_seconds"] > 0
    assert t["splitter_seconds"] > 0 and t["sampling_percent"] > 0 and t["splitter_percent"] > 0
    # Node fits are not the split optimizer's model refits.
    assert t["optimizer_model_refit_seconds"] == 0.0
    assert np.isclose(t["sampling_percent"] + t["splitter_percent"] + t["other_percent"], 100.0)


def test_children_inherit_samples_and_query_only_the_top_up(traversal):
    result = RegressionTreeBuilder(
        domain=square(), oracle=piecewise, model_template=HostAffine(1e-12),
        splitter=FixedAxisSplitter(), error_threshold=1e-8, sample_count=30, max_depth=1,
        sampler=HostHitAndRun(burn_in=0, thinning=1), random_state=3,
        traversal=traversal).build()
    root, tree = result.tree.root, result.tree
    assert isinstance(root, SplitNode)
    assert root.metadata["split_gain_mse"] > 0 and root.metadata["relative_split_gain_mse"] > 0
    assert result.oracle_queries == 60
    assert [n.node_id for n in tree.iter_nodes()] == ["root", "root/L", "root/R"]
    assert [n.node_id for n in tree.iter_leaves()] == ["root/L", "root/R"]
    assert (tree.num_nodes(), tree.num_split_nodes(), tree.num_leaves()) == (3, 1, 2)
    assert root.left.metadata["n_samples"] == 30 and root.right.metadata["n_samples"] == 30
    assert root.left.metadata["n_new"] + root.right.metadata["n_new"] == 30
    assert root.metadata["sampling_method"] == "exact_uniform_box"
    for child in (root.left, root.right):
        assert child.metadata["sampling_method"] == "isotropic_hit_and_run"
        assert child.metadata["sampling_pilot_length"] == 100
        assert child.metadata["sampling_thinning"] == 1
        assert np.isfinite(child.metadata["sampling_covariance_condition_number"])
    X_test = np.array([[-0.8, 0.2], [0.6, -0.4]])
    y_test = np.array([piecewise(x) for x in X_test])
    assert mean_squared_error(y_test, tree.predict(X_test)) < 1e-12


def test_small_initial_gradient_warns_on_an_accepted_split(traversal):
    root = retry_builder(FixedAxisSplitter(n_iters=0, stop_reason="gradient_tolerance"),
                         traversal).build().tree.root
    assert isinstance(root, SplitNode) and root.status == "split_with_warnings"
    assert root.metadata["split_iterations"] == 0
    assert root.metadata["split_stop_reason"] == "gradient_tolerance"
    assert root.metadata["warnings"] == ("small_init_grad",)


def test_high_model_conditioning_warns_on_an_accepted_split(traversal):
    root = retry_builder(FixedAxisSplitter(metadata={"high_model_conditioning": True}),
                         traversal).build().tree.root
    assert isinstance(root, SplitNode) and root.status == "split_with_warnings"
    assert root.metadata["warnings"] == ("high_model_conditioning",)


def test_isotropic_sampling_can_be_disabled(traversal):
    root = RegressionTreeBuilder(
        domain=square(), oracle=piecewise, model_template=HostAffine(1e-12),
        splitter=FixedAxisSplitter(), error_threshold=0.0, sample_count=30, max_depth=1,
        sampler=HostHitAndRun(burn_in=0, thinning=3), isotropic_sampling=False, random_state=3,
        traversal=traversal).build().tree.root
    assert isinstance(root, SplitNode)
    for child in (root.left, root.right):
        assert child.metadata["sampling_method"] == "HostHitAndRun"
        assert child.metadata["sampling_pilot_length"] == 0
        assert child.metadata["sampling_thinning"] == 3
        assert child.metadata["sampling_covariance_condition_number"] is None


def test_covariance_floor_saturation_is_reported(traversal):
    root = RegressionTreeBuilder(
        domain=square(), oracle=piecewise, model_template=HostAffine(1e-12),
        splitter=FixedAxisSplitter(), error_threshold=0.0, sample_count=30, max_depth=1,
        sampler=HostHitAndRun(burn_in=0, thinning=1, direction_eigenvalue_floor=1.0),
        random_state=3, traversal=traversal).build().tree.root
    assert isinstance(root, SplitNode)
    for child in (root.left, root.right):
        assert child.metadata["sampling_covariance_floor_saturated"] is True
        assert child.metadata["sampling_warnings"] == ("covariance_eigenvalue_floor_saturated",)


def test_storage_toggles_keep_samples_and_the_leaf_centroid(traversal):
    root = RegressionTreeBuilder(
        domain=BoxDomain(np.array([[-1.0, 1.0]])), oracle=lambda x: x[0] ** 2,
        model_template=HostAffine(), splitter=FixedAxisSplitter(), error_threshold=0.0,
        sample_count=12, max_depth=0, store_samples=True, store_leaf_representative=True,
        random_state=4, sampler=HostHitAndRun(), traversal=traversal).build().tree.root
    assert isinstance(root, LeafNode) and root.status == "max_depth"
    assert root.metadata["X"].shape == (12, 1) and root.metadata["y"].shape == (12,)
    np.testing.assert_allclose(root.metadata["sample_centroid"], root.metadata["X"].mean(axis=0))
    assert bool(root.region.contains(root.metadata["sample_centroid"])[0])


def test_min_side_failure_retries_with_the_same_samples_and_a_new_seed(traversal):
    splitter = ScriptSplitter(("min_side_points", "success"))
    result = retry_builder(splitter, traversal, max_retries_on_failure=2).build()
    root = result.tree.root
    assert isinstance(root, SplitNode)
    assert result.restarts_on_failure == 1 and result.tree.metadata["restarts_on_failure"] == 1
    assert root.metadata["restarts_on_failure"] == 1
    assert root.metadata["split_attempt_failure_reasons"] == ("min_side_points",)
    assert root.metadata["splitter_metadata"]["run"] == 2
    np.testing.assert_array_equal(splitter.state["X"][0], splitter.state["X"][1])
    np.testing.assert_array_equal(splitter.state["y"][0], splitter.state["y"][1])
    assert splitter.state["seeds"][0] != splitter.state["seeds"][1]


@pytest.mark.parametrize("check", [False, True])
def test_gain_check_toggle_controls_thresholds_and_retries(traversal, check):
    result = retry_builder(ScriptSplitter(("insufficient_split_gain",)), traversal,
                           check_split_gain=check, max_retries_on_failure=1, min_split_gain=1.,
                           min_relative_split_gain=1.).build()
    assert result.tree.metadata["check_split_gain"] == check
    if check:
        assert isinstance(result.tree.root, LeafNode)
        assert result.tree.root.status == "insufficient_split_gain"
        assert result.restarts_on_failure == 1
    else:
        assert isinstance(result.tree.root, SplitNode)
        assert result.tree.root.metadata["split_gain_mse"] == 0.
        assert result.restarts_on_failure == 0


def test_disabled_gain_checks_keep_the_numerical_guards(traversal):
    builder = retry_builder(FixedAxisSplitter(), traversal, check_split_gain=False)
    result = FixedAxisSplitter().split(np.array([[-1., 0.], [-.5, 1.], [.5, 0.], [1., 1.]]),
                                       np.array([0., 1., 2., 3.]), parent_loss=0.)
    assert builder._split_rejection_reason(result) is None
    result.loss = np.nan
    assert builder._split_rejection_reason(result) == "nonfinite_split"
    result.loss, result.n_iters, result.stop_reason = 1., 0, "line_search_failed"
    assert builder._split_rejection_reason(result) == "optimizer_failed_without_progress"


def test_min_side_fallback_only_on_the_final_retry(traversal):
    splitter = FallbackScriptSplitter()
    result = retry_builder(splitter, traversal, max_retries_on_failure=2).build()
    root = result.tree.root
    assert isinstance(root, SplitNode) and root.status == "split_with_warnings"
    assert result.restarts_on_failure == 2 and splitter.state["calls"] == 3
    assert root.metadata["split_attempt_failure_reasons"] == ("min_side_points",) * 2
    assert root.metadata["warnings"] == ("min_side_fallback",)
    assert root.metadata["splitter_metadata"]["min_side_fallback"] is True
    assert root.metadata["splitter_metadata"]["fallback_iteration"] == 0
    assert len(set(splitter.state["seeds"])) == 3


def test_insufficient_gain_retries_and_accepts_a_later_split(traversal):
    result = retry_builder(ScriptSplitter(("insufficient_split_gain", "success")), traversal,
                           min_split_gain=1e-6, max_retries_on_failure=1).build()
    root = result.tree.root
    assert isinstance(root, SplitNode) and result.restarts_on_failure == 1
    assert root.metadata["split_attempt_failure_reasons"] == ("insufficient_split_gain",)
    assert root.metadata["splitter_metadata"]["run"] == 2


def test_exhausted_min_side_retries_keep_the_last_diagnostics(traversal):
    result = retry_builder(ScriptSplitter(("min_side_points",)), traversal,
                           max_retries_on_failure=2).build()
    root = result.tree.root
    assert isinstance(root, LeafNode) and root.status == "min_side_points"
    assert result.restarts_on_failure == 2 and result.oracle_queries == 30
    assert root.metadata["restarts_on_failure"] == 2
    assert root.metadata["split_attempt_failure_reasons"] == ("min_side_points",) * 3
    assert root.metadata["splitter_metadata"]["run"] == 3


def test_exhausted_gain_retries_keep_the_last_result_diagnostics(traversal):
    result = retry_builder(ScriptSplitter(("insufficient_split_gain",)), traversal,
                           min_split_gain=1e-6, max_retries_on_failure=1).build()
    root = result.tree.root
    assert isinstance(root, LeafNode) and root.status == "insufficient_split_gain"
    assert result.restarts_on_failure == 1 and root.metadata["restarts_on_failure"] == 1
    assert root.metadata["splitter_metadata"]["run"] == 2


def test_retries_need_a_single_restart_splitter():
    with pytest.raises(ValueError, match="requires splitter.n_restarts == 1"):
        retry_builder(ScriptSplitter(("success",), n_restarts=2), "dfs",
                      max_retries_on_failure=1)
