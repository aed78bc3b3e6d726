"""Adaptive regression-tree construction (art/builder.py), GPU-first.

``RegressionTreeBuilder`` keeps the reference's fields, validation messages,
stopping rules, retry policy, random-stream order and node metadata, so a
build with the same seed consumes the same draws and makes the same
decisions.  What runs where:

* node model fits: every node prepared in one scheduling step (a BFS cohort,
  or the single node of a DFS step) is fitted in ONE batched launch of the
  leaf-ridge kernels (``models.fit_models``) when the template is a GPU ridge
  model; any other model type is fitted through its own ``fit``;
* splits: ``SoftObliqueSplitter`` runs on the device optimizer; under BFS
  with optimizer="cupy_batched_cholesky" a whole cohort is one launch
  (``BatchedSplitDispatcher``);
* sampling: ``HitAndRunSampler`` chains run on the GPU sampler;
* temperature scales: ``temperature.estimate_temperature`` (GPU distances).

The host keeps the queue, the policy and the tree.  Oracle queries are the
caller's function.
"""

from __future__ import annotations

import copy
from contextlib import nullcontext
from dataclasses import dataclass, field, replace
from time import perf_counter
from typing import Any, Callable, Literal

import numpy as np

from .domain import BoxDomain, PolytopeRegion, split_region
from .metrics import mean_squared_error, relative_l2_error
from .models import (
    AffineRidgeModel,
    MODEL_CONDITION_WARNING_THRESHOLD,
    PolynomialRidgeModel,
    WeightedRegressionModel,
    fit_models,
    model_effective_dimension,
    ridge_solve_diagnostics,
)
from .sampling import (
    HitAndRunSampler,
    floor_covariance_eigenvalues,
    sample_covariance_eigendecomposition,
    sample_uniform_box,
)
from .splitters import SoftObliqueSplitter, SplitNotFoundError, SplitResult
from .temperature import TemperatureConfig, estimate_temperature
from .timing import BuildTimingProfile
from .tree import LeafNode, RegressionTree, SplitNode, TreeNode

ErrorMetric = Callable[[np.ndarray, np.ndarray], float]

_NUMERICAL_FAILURES = frozenset({"line_search_failed", "max_line_search_failures",
                                 "alpha_min_reached", "non_descent_direction"})


@dataclass(frozen=True)
class TreeBuildResult:
    tree: RegressionTree
    oracle_queries: int
    restarts_on_failure: int
    build_timing: dict[str, float] | None = None


@dataclass
class _NodeSamples:
    X: np.ndarray
    y: np.ndarray
    n_inherited: int
    n_new: int
    sampling_metadata: dict[str, object] = field(default_factory=dict)


@dataclass(slots=True)
class _NodeWork:
    """One node from its inherited samples to a split candidate.

    parent/side name the attachment slot; preparation fills the model, the
    parent MSE and the metadata (art/builder.py:62-83).
    """

    region: PolytopeRegion
    samples: _NodeSamples
    depth: int
    node_id: str
    parent: SplitNode | None = None
    side: Literal["left", "right"] | None = None
    model: Any = None
    parent_mse: float = 0.0
    metadata: dict[str, object] = field(default_factory=dict)
    convergence: dict[str, object] = field(default_factory=dict)
    prepared: Any = None


# Outcomes that earn another attempt while the node's retry budget lasts:
# split failures by reason, accepted results by their rejection reason.
_RETRY_FAILURES = frozenset({"min_side_points"})
_RETRY_REJECTIONS = frozenset({"insufficient_split_gain", "insufficient_relative_split_gain"})


@dataclass
class _SplitLedger:
    """One node's split attempts (art/builder.py:86-113 keeps the same facts).

    ``budget`` counts the retries still allowed; the min-side fallback is
    switched on for the attempt made when none remain.  Every attempt's
    failure or rejection reason is logged in order.
    """

    temperature: float | None
    context: dict[str, object]
    budget: int
    fallback_enabled: bool
    spent: int = 0
    log: list[str] = field(default_factory=list)

    @property
    def fallback(self):
        return self.fallback_enabled and self.budget <= 0

    def settle(self, reason, retryable):
        """Log an attempt's reason; True when it is retried."""
        if reason is not None:
            self.log.append(reason)
        if not retryable or self.budget <= 0:
            return False
        self.budget -= 1
        self.spent += 1
        return True

    def summary(self):
        return {"restarts_on_failure": self.spent,
                "split_attempt_failure_reasons": tuple(self.log)}

    def failure(self, exc):
        """The node's final SplitNotFoundError, carrying the whole ledger."""
        return SplitNotFoundError(exc.reason, str(exc), diagnostics=exc.diagnostics,
                                  restarts_on_failure=self.spent,
                                  failure_reasons=tuple(self.log), context=self.context)


@dataclass
class _FitRequest:
    """A pending node fit: rows to fit and rows to evaluate on."""

    X: np.ndarray
    y: np.ndarray
    eval_X: np.ndarray


def _sampling_info(method, thinning=None, pilot=0, condition=None, warnings=()):
    return {"sampling_method": method, "sampling_thinning": thinning,
            "sampling_pilot_length": pilot, "sampling_covariance_condition_number": condition,
            "sampling_warnings": warnings}


@dataclass
class RegressionTreeBuilder:
    """Build a regression tree depth-first (default) or by FIFO cohorts."""

    domain: BoxDomain | PolytopeRegion
    oracle: Callable[[np.ndarray], float | np.ndarray]
    model_template: Any
    splitter: Any
    error_threshold: float
    max_depth: int = 10
    error_metric: ErrorMetric = relative_l2_error
    sample_multiplier: int = 50
    sample_count: int | Callable | None = None
    sampler: Any = field(default_factory=lambda: HitAndRunSampler(burn_in=0, thinning=20))
    root_sampler: Any = field(default_factory=HitAndRunSampler)
    exact_box_root: bool = True
    isotropic_sampling: bool = True
    isotropic_pilot_multiplier: int = 50
    temperature_config: TemperatureConfig | None = None
    min_split_gain: float = 0.0
    min_relative_split_gain: float = 0.0
    max_retries_on_failure: int = 0
    validation_fraction: float = 0.0
    store_samples: bool = False
    store_leaf_representative: bool = False
    store_diagnostics: bool = False
    profile_build_timing: bool = False
    oracle_vectorized: bool = False
    random_state: int | np.random.Generator | None = None
    traversal: Literal["dfs", "bfs"] = "dfs"
    node_batch_size: int = 20
    batch_reorder: bool = True
    check_split_gain: bool = False

    oracle_queries: int = field(default=0, init=False)
    restarts_on_failure: int = field(default=0, init=False)
    _rng: np.random.Generator = field(init=False, repr=False)
    _root_temperature_c: float | None = field(default=None, init=False, repr=False)
    _target_samples: int = field(init=False, repr=False)
    _effective_dimension: int | None = field(init=False, repr=False)
    _timing_profile: BuildTimingProfile | None = field(default=None, init=False, repr=False)

    # ------------------------------------------------------------------ setup
    def __post_init__(self):
        if self.traversal not in ("dfs", "bfs"):
            raise ValueError("traversal must be 'dfs' or 'bfs'.")
        nbs = self.node_batch_size
        if not isinstance(nbs, (int, np.integer)) or isinstance(nbs, (bool, np.bool_)) or nbs < 1:
            raise ValueError("node_batch_size must be a positive integer.")
        self.node_batch_size = int(nbs)
        if not isinstance(self.batch_reorder, (bool, np.bool_)):
            raise ValueError("batch_reorder must be a boolean.")
        if not isinstance(self.check_split_gain, (bool, np.bool_)):
            raise ValueError("check_split_gain must be a boolean.")
        if (getattr(self.splitter, "optimizer", None) == "cupy_batched_cholesky"
                and self.traversal != "bfs"):
            raise ValueError("optimizer='cupy_batched_cholesky' requires traversal='bfs'.")
        checks = (
            (self.error_threshold < 0.0, "error_threshold must be nonnegative."),
            (self.max_depth < 0, "max_depth must be nonnegative."),
            (self.sample_multiplier < 1, "sample_multiplier must be at least 1."),
            (self.min_split_gain < 0.0 or self.min_relative_split_gain < 0.0,
             "Split-gain thresholds must be nonnegative."),
            (self.max_retries_on_failure < 0, "max_retries_on_failure must be nonnegative."),
            (not (0.0 <= self.validation_fraction < 1.0),
             "validation_fraction must satisfy 0 <= fraction < 1."),
            (self.isotropic_sampling and self.isotropic_pilot_multiplier < 2,
             "isotropic_pilot_multiplier must be at least 2."),
        )
        for bad, message in checks:
            if bad:
                raise ValueError(message)
        if self.isotropic_sampling and not isinstance(self.sampler, HitAndRunSampler):
            raise TypeError("isotropic_sampling requires a HitAndRunSampler.")
        hit_and_run_root = not (isinstance(self.domain, BoxDomain) and self.exact_box_root)
        if (self.isotropic_sampling and hit_and_run_root
                and not isinstance(self.root_sampler, HitAndRunSampler)):
            raise TypeError("isotropic_sampling requires a HitAndRunSampler for non-exact roots.")
        if self.max_retries_on_failure > 0 and int(getattr(self.splitter, "n_restarts", 1)) != 1:
            raise ValueError("max_retries_on_failure requires splitter.n_restarts == 1.")

        self._rng = (self.random_state if isinstance(self.random_state, np.random.Generator)
                     else np.random.default_rng(self.random_state))
        self._effective_dimension = self._resolve_effective_dimension()
        self._target_samples = self._resolve_sample_count()
        if self._target_samples < 2:
            raise ValueError("The node sample count must be at least 2.")
        has_temperature = hasattr(self.splitter, "temperature")
        if isinstance(self.splitter, SoftObliqueSplitter) and not isinstance(
                self.model_template, WeightedRegressionModel):
            raise TypeError("SoftObliqueSplitter requires a model with fit_weighted support.")
        if self.validation_fraction > 0.0:
            if not isinstance(self.splitter, SoftObliqueSplitter):
                raise ValueError("validation_fraction is supported only by SoftObliqueSplitter.")
            if self._metric_name() != "max_pointwise_relative_error":
                raise ValueError("validation_fraction requires max_pointwise_relative_error.")
        if has_temperature and self.temperature_config is None:
            self.temperature_config = TemperatureConfig()
        if not has_temperature and self.temperature_config is not None:
            raise ValueError("temperature_config requires a splitter with a temperature attribute.")
        if (self.temperature_config is not None
                and self.temperature_config.strategy in ("tune_root", "tune_node")
                and self._target_samples < 4):
            raise ValueError("Temperature tuning requires at least four node samples.")

    @property
    def hard_refits(self):
        """Hard losses are needed for gain checks, restart selection or tuning."""
        return (not isinstance(self.splitter, SoftObliqueSplitter) or self.check_split_gain
                or self.splitter.n_restarts != 1
                or self.temperature_config.strategy not in ("fixed", "splitter"))

    @property
    def dimension(self):
        return self.domain.dimension

    @property
    def target_samples(self):
        return self._target_samples

    # ------------------------------------------------------------------ build
    def build(self) -> TreeBuildResult:
        """Build a fresh tree and reset the per-build state."""
        self._timing_profile = BuildTimingProfile() if self.profile_build_timing else None
        start = perf_counter()
        self.oracle_queries = 0
        self.restarts_on_failure = 0
        self._root_temperature_c = None
        region = self.domain.as_region() if isinstance(self.domain, BoxDomain) else self.domain
        X, info = self._sample_root(region)
        y = self._query_oracle(X)
        work = _NodeWork(region, _NodeSamples(X, y, 0, X.shape[0], info), depth=0,
                         node_id="root")
        if self.traversal == "dfs":
            root = self._build_node(work)
        else:
            from ._batched_builder import build_bfs

            if (isinstance(self.splitter, SoftObliqueSplitter)
                    and self.splitter.optimizer == "cupy_batched_cholesky"):
                from ._cuda_batched_builder import BatchedSplitDispatcher

                with BatchedSplitDispatcher(self) as dispatch:
                    root = build_bfs(self, work, dispatch)
            else:
                root = build_bfs(self, work)
        config = self.temperature_config
        tree = RegressionTree(root=root, oracle_queries=self.oracle_queries, metadata={
            "error_metric": self._metric_name(),
            "error_threshold": self.error_threshold,
            "effective_dimension": self._effective_dimension,
            "splitter": type(self.splitter).__name__,
            "optimizer": getattr(self.splitter, "optimizer", None),
            "check_split_gain": self.check_split_gain,
            "hard_refits": self.hard_refits,
            "traversal": self.traversal,
            "node_batch_size": self.node_batch_size if self.traversal == "bfs" else 1,
            "target_samples_per_node": self._target_samples,
            "restarts_on_failure": self.restarts_on_failure,
            "validation_fraction": self.validation_fraction,
            "isotropic_sampling": self.isotropic_sampling,
            "isotropic_pilot_multiplier": self.isotropic_pilot_multiplier,
            "temperature_strategy": None if config is None else config.strategy,
        })
        timing = None
        if self._timing_profile is not None:
            timing = self._timing_profile.summary(perf_counter() - start)
        tree.metadata["build_timing"] = timing
        return TreeBuildResult(tree, self.oracle_queries, self.restarts_on_failure, timing)

    def _build_node(self, work: _NodeWork) -> TreeNode:
        """Depth-first: a node, then its left subtree, then its right."""
        leaf = self._prepare_node(work)
        if leaf is not None:
            return leaf
        node = self._split_node(work)
        if isinstance(node, SplitNode):
            for child in self._children(work, node):
                setattr(node, child.side, self._build_node(child))
        return node

    # ------------------------------------------------------ node preparation
    def _prepare_node(self, work: _NodeWork) -> LeafNode | None:
        """Top up, fit and test one node; a returned leaf is final."""
        return self._prepare_cohort([work])[0]

    def _prepare_cohort(self, jobs: list[_NodeWork]) -> list[LeafNode | None]:
        """Prepare several nodes with one batched model fit.

        Draws are taken node by node in list order (top-up sampling, then the
        validation permutation), exactly as preparing the nodes one at a time
        would take them (art/builder.py:358-396); the chains of the top-ups
        run together (``_top_up_cohort``), and so do the fits, which draw
        nothing.
        """
        plans, requests = [], []
        # A validation split draws its permutation right after its node's
        # top-up, so those builds top up node by node.
        validating = self.validation_fraction > 0.0
        topped = [None] * len(jobs) if validating else self._top_up_cohort(jobs)
        for k, work in enumerate(jobs):
            top = self._top_up_one(work) if validating else topped[k]
            plan = {"work": work, "failed": None, "fit_rows": None}
            if isinstance(top, Exception):
                plan["failed"] = top
            elif top is not None:
                work.samples = top
            samples = work.samples
            if plan["failed"] is None and validating:
                n = samples.X.shape[0]
                n_val = min(max(1, int(round(self.validation_fraction * n))), n - 1)
                plan["fit_rows"] = self._rng.permutation(n)[n_val:]
            rows = plan["fit_rows"]
            fit_X = samples.X if rows is None else samples.X[rows]
            fit_y = samples.y if rows is None else samples.y[rows]
            plan["request"] = len(requests)
            requests.append(_FitRequest(fit_X, fit_y, samples.X))
            plans.append(plan)
        fitted = self._fit_requests(requests)
        out = []
        for plan in plans:
            work = plan["work"]
            model, predictions = fitted[plan["request"]]
            samples = work.samples
            if plan["failed"] is not None:
                metadata = self._node_metadata(samples, float(self.error_metric(samples.y,
                                                                                predictions)),
                                               mean_squared_error(samples.y, predictions), model)
                metadata["sampling_error"] = repr(plan["failed"])
                out.append(self._make_leaf(model, work.region, work.depth, work.node_id,
                                           "sampling_failed", metadata, samples))
                continue
            work.model = model
            fit_error = float(self.error_metric(samples.y, predictions))
            work.parent_mse = mean_squared_error(samples.y, predictions)
            work.metadata = self._node_metadata(samples, fit_error, work.parent_mse, model)
            if plan["fit_rows"] is not None:
                work.convergence = {"convergence_error": fit_error,
                                    "convergence_fit_samples": int(plan["fit_rows"].size)}
                work.metadata.update(work.convergence)
            if fit_error <= self.error_threshold:
                out.append(self._work_leaf(work, "tolerance_met"))
            elif work.depth >= self.max_depth:
                out.append(self._work_leaf(work, "max_depth"))
            else:
                out.append(None)
        return out

    def _gpu_fits(self):
        """The GPU ridge models themselves (not subclasses, whose own fit is
        called) are fitted a cohort at a time."""
        return type(self.model_template) in (AffineRidgeModel, PolynomialRidgeModel)

    def _fit_requests(self, requests: list[_FitRequest]):
        """(model, predictions on eval_X) per request; GPU ridge templates fit
        every request in one launch.  Node fits are not timed as the split
        optimizer's model refits (they count as other build time)."""
        if not requests:
            return []
        if self._gpu_fits():
            models, _ = fit_models(self.model_template, [(r.X, r.y) for r in requests])
            return [(m, m.predict(r.eval_X)) for m, r in zip(models, requests)]
        out = []
        for r in requests:
            m = self.model_template.clone()
            m.fit(r.X, r.y)
            out.append((m, np.asarray(m.predict(r.eval_X), dtype=float).reshape(-1)))
        return out

    def _fit_node_model(self, X, y, prepared_design=None):
        """One node fit: (model, predictions on X) (art/builder.py:812-829)."""
        return self._fit_requests([_FitRequest(X, y, X)])[0]

    def _fit_full_node_model(self, samples, prepared_design=None, extra_metadata=None):
        model, predictions = self._fit_node_model(samples.X, samples.y)
        fit_error = float(self.error_metric(samples.y, predictions))
        fit_mse = mean_squared_error(samples.y, predictions)
        metadata = self._node_metadata(samples, fit_error, fit_mse, model)
        if extra_metadata is not None:
            metadata.update(extra_metadata)
        return model, fit_error, fit_mse, metadata

    # ------------------------------------------------------------ splitting
    def _split_node(self, work: _NodeWork) -> TreeNode:
        """Split one node with its retries, then finish it."""
        X, y = work.samples.X, work.samples.y
        try:
            ledger = self._open_ledger(X, y, work.depth)
        except SplitNotFoundError as exc:
            return self._finish_node(work, exc)
        outcome = self._attempt_split(ledger, X, y, work.model, work.parent_mse)
        while self._settle_attempt(ledger, outcome):
            outcome = self._attempt_split(ledger, X, y, work.model, work.parent_mse)
        if isinstance(outcome, SplitNotFoundError):
            return self._finish_node(work, ledger.failure(outcome))
        return self._finish_node(work, outcome, {**ledger.context, **ledger.summary()})

    def _finish_node(self, work, result, context=None) -> TreeNode:
        """Turn a final split outcome into a leaf or a split with pending children."""
        if isinstance(result, SplitNotFoundError):
            return self._work_leaf(work, result.reason, error=result)
        rejection = self._split_rejection_reason(result)
        if rejection is not None:
            return self._work_leaf(work, rejection, result=result, context=context)
        left_id, right_id = f"{work.node_id}/L", f"{work.node_id}/R"
        regions = split_region(work.region, result.w, result.z, depth=work.depth + 1,
                               left_tag=left_id, right_tag=right_id)
        hard = result.has_hard_fit
        left = LeafNode(result.left_model if hard else work.model, regions[0], work.depth + 1,
                        left_id, status="pending")
        right = LeafNode(result.right_model if hard else work.model, regions[1], work.depth + 1,
                         right_id, status="pending")
        metadata = self._split_metadata(work, result, context)
        self._store_samples(metadata, work.samples)
        work.prepared = None
        return SplitNode(w=result.w, z=result.z, left=left, right=right, region=work.region,
                         depth=work.depth, node_id=work.node_id,
                         status="split_with_warnings" if metadata["warnings"] else "split",
                         metadata=metadata)

    def _split_metadata(self, work, result, context):
        metadata = {**work.metadata, **self._split_summary(result), **(context or {}),
                    "warnings": self._combine_warnings(work.metadata.get("warnings", ()),
                                                       self._split_warnings(result))}
        if self.store_diagnostics:
            metadata["splitter_metadata"] = copy.deepcopy(result.metadata)
        return metadata

    def _work_leaf(self, work, status, *, result=None, error=None, context=None):
        """Finalize a leaf; after a validation fit, refit on all its samples."""
        if self.validation_fraction > 0.0:
            work.model, _, _, work.metadata = self._fit_full_node_model(
                work.samples, work.prepared, extra_metadata=work.convergence)
        if error is not None:
            metadata = {**work.metadata, **error.context,
                        "split_failure_reason": error.reason,
                        "restarts_on_failure": error.restarts_on_failure,
                        "split_attempt_failure_reasons": error.failure_reasons,
                        "warnings": self._combine_warnings(
                            work.metadata.get("warnings", ()),
                            self._diagnostic_warnings(error.diagnostics))}
            if self.store_diagnostics and error.diagnostics is not None:
                metadata["splitter_metadata"] = copy.deepcopy(error.diagnostics)
        elif result is not None:
            metadata = self._split_metadata(work, result, context)
        else:
            metadata = work.metadata
        work.prepared = None
        return self._make_leaf(work.model, work.region, work.depth, work.node_id, status,
                               metadata, work.samples)

    def _children(self, work, node):
        """Inherited-only child jobs, left then right (no sampling here)."""
        samples = work.samples
        right = (samples.X @ node.w - node.z) >= 0.0
        for side, mask in (("left", ~right), ("right", right)):
            child = getattr(node, side)
            X, y = samples.X[mask], samples.y[mask]
            yield _NodeWork(child.region, _NodeSamples(X, y, n_inherited=len(X), n_new=0),
                            child.depth, child.node_id, parent=node, side=side)

    def _top_up_cohort(self, jobs):
        """Top up every child job of a cohort; returns per job its new
        samples, the exception that made sampling fail, or None (root).

        Draws follow the node-by-node order of ``_top_up_child`` (subset,
        start row, pilot seed, main seed; then the oracle), but the chains of
        all nodes run together: one launch for the pilots, one for the main
        chains.  A pilot that fails would have skipped its node's main-seed
        draw, so the cohort then restarts from the saved generator state and
        runs node by node.
        """
        children = [w for w in jobs if w.parent is not None]
        # Only the GPU sampler itself batches; a subclass with its own sample
        # (or any other sampler) is called node by node.
        if len(children) < 2 or type(self.sampler).sample is not HitAndRunSampler.sample:
            return [self._top_up_one(w) for w in jobs]
        saved = copy.deepcopy(self._rng.bit_generator.state)
        plans = []
        for w in jobs:
            if w.parent is None:
                plans.append(None)
                continue
            X = np.asarray(w.samples.X, dtype=float)
            y = np.asarray(w.samples.y, dtype=float).reshape(-1)
            if X.shape[0] == 0:
                plans.append(RuntimeError("Cannot seed child sampling without an inherited point."))
                continue
            if X.shape[0] > self._target_samples:
                keep = self._rng.choice(X.shape[0], size=self._target_samples, replace=False)
                X, y = X[keep], y[keep]
            missing = self._target_samples - X.shape[0]
            plan = {"work": w, "X": X, "y": y, "missing": missing}
            if missing:
                plan["x0"] = X[int(self._rng.integers(0, X.shape[0]))]
                if self.isotropic_sampling:
                    plan["pilot_seed"] = self._next_seed()
                plan["seed"] = self._next_seed()
            plans.append(plan)
        sampling = [p for p in plans if isinstance(p, dict) and p["missing"]]
        with self._timing_context("sampling"):
            ok = self._sample_cohort(sampling)
        if not ok:
            self._rng.bit_generator.state = saved
            return [self._top_up_one(w) for w in jobs]
        out = []
        for p in plans:
            if p is None or isinstance(p, Exception):
                out.append(p)
                continue
            if not p["missing"]:
                out.append(_NodeSamples(p["X"], p["y"], p["X"].shape[0], 0,
                                        _sampling_info("inherited_only")))
                continue
            if isinstance(p["new"], Exception):
                out.append(p["new"])
                continue
            y_new = self._query_oracle(p["new"])
            out.append(_NodeSamples(np.vstack([p["X"], p["new"]]),
                                    np.concatenate([p["y"], y_new]), p["X"].shape[0],
                                    p["missing"], p["info"]))
        return out

    def _sample_cohort(self, plans):
        """Run the sampling of ``_sample_region_untimed`` for every plan in
        batched launches; sets plan["new"] (array or exception) and
        plan["info"].  False when a pilot failure changes the draw order."""
        from .sampling import sample_hit_and_run_many

        d = self.dimension
        if not plans:
            return True
        if not self.isotropic_sampling:
            got = sample_hit_and_run_many([self.sampler] * len(plans),
                                          [p["work"].region for p in plans],
                                          [p["missing"] for p in plans],
                                          [p["seed"] for p in plans], [p["x0"] for p in plans])
            for p, X in zip(plans, got):
                p["new"] = X if isinstance(X, Exception) else self._checked(p["work"].region, X,
                                                                            p["missing"])
                p["info"] = _sampling_info(type(self.sampler).__name__, self.sampler.thinning)
            return True
        pilot_length = self.isotropic_pilot_multiplier * d
        pilot_sampler = replace(self.sampler, burn_in=0, thinning=1,
                                direction_eigenvectors=None, direction_eigenvalues=None)
        pilots = sample_hit_and_run_many([pilot_sampler] * len(plans),
                                         [p["work"].region for p in plans],
                                         [pilot_length] * len(plans),
                                         [p["pilot_seed"] for p in plans],
                                         [p["x0"] for p in plans])
        shaped = []
        for p, pilot in zip(plans, pilots):
            if isinstance(pilot, Exception):
                return False
            try:
                vectors, values = sample_covariance_eigendecomposition(pilot)
                safe, floor = floor_covariance_eigenvalues(
                    values, self.sampler.direction_eigenvalue_floor)
            except (RuntimeError, ValueError):
                return False
            saturated = bool(np.any(values <= floor))
            warning = "covariance_eigenvalue_floor_saturated"
            p["info"] = {**_sampling_info("isotropic_hit_and_run", self.sampler.thinning,
                                          pilot_length, float(np.max(safe) / np.min(safe)),
                                          (warning,) if saturated else ()),
                         "sampling_covariance_eigenvalue_floor": floor,
                         "sampling_covariance_floor_saturated": saturated}
            shaped.append((replace(self.sampler, direction_eigenvectors=vectors,
                                   direction_eigenvalues=values), pilot[-1]))
        got = sample_hit_and_run_many([sm for sm, _ in shaped], [p["work"].region for p in plans],
                                      [p["missing"] for p in plans], [p["seed"] for p in plans],
                                      [x for _, x in shaped])
        for p, X in zip(plans, got):
            p["new"] = X if isinstance(X, Exception) else self._checked(p["work"].region, X,
                                                                        p["missing"])
        return True

    def _checked(self, region, X, n):
        if X.shape != (n, self.dimension):
            return RuntimeError(f"Sampler returned shape {X.shape}; expected "
                                f"{(n, self.dimension)}.")
        if not np.all(region.contains(X)):
            return RuntimeError("Sampler returned points outside the region.")
        return X

    def _top_up_one(self, work):
        if work.parent is None:
            return None
        try:
            return self._top_up_child(work.region, work.samples.X, work.samples.y)
        except (RuntimeError, ValueError) as exc:
            return exc

    def _top_up_child(self, region, inherited_X, inherited_y) -> _NodeSamples:
        """Keep at most target_samples inherited rows and sample the rest."""
        X = np.asarray(inherited_X, dtype=float)
        y = np.asarray(inherited_y, dtype=float).reshape(-1)
        if X.shape[0] == 0:
            raise RuntimeError("Cannot seed child sampling without an inherited point.")
        if X.shape[0] > self._target_samples:
            keep = self._rng.choice(X.shape[0], size=self._target_samples, replace=False)
            X, y = X[keep], y[keep]
        kept = X.shape[0]
        missing = self._target_samples - kept
        if missing == 0:
            return _NodeSamples(X, y, kept, 0, _sampling_info("inherited_only"))
        x0 = X[int(self._rng.integers(0, kept))]
        X_new, info = self._sample_region(region, missing, x0=x0, sampler=self.sampler)
        y_new = self._query_oracle(X_new)
        return _NodeSamples(np.vstack([X, X_new]), np.concatenate([y, y_new]), kept, missing,
                            info)

    def _attempt_split(self, ledger, X, y, parent_model, parent_mse):
        """One run of the splitter under the ledger's temperature and fallback
        setting: a SplitResult, or the SplitNotFoundError it raised."""
        try:
            return self._run_splitter(X, y, parent_model, parent_mse, ledger.temperature, None,
                                      min_side_fallback=ledger.fallback)
        except SplitNotFoundError as exc:
            return exc

    def _open_ledger(self, X, y, depth) -> _SplitLedger:
        """Pick the node's temperature (its random draws come here, before any
        attempt) and open its ledger; a node whose temperature was tuned gets
        no retries."""
        temperature, context, tuned = self._pick_temperature(X, y, depth)
        return _SplitLedger(temperature, context,
                            budget=0 if tuned else self.max_retries_on_failure,
                            fallback_enabled=bool(getattr(self.splitter, "min_side_fallback",
                                                          False)))

    def _settle_attempt(self, ledger, outcome) -> bool:
        """Classify an attempt and log it; True when the node tries again."""
        if isinstance(outcome, SplitNotFoundError):
            reason = outcome.reason
            retryable = reason in _RETRY_FAILURES
        else:
            reason = self._split_rejection_reason(outcome)
            retryable = reason in _RETRY_REJECTIONS
        again = ledger.settle(reason, retryable)
        self.restarts_on_failure += int(again)
        return again

    def _pick_temperature(self, X, y, depth):
        """(temperature or None for the splitter's own, context, tuned).

        The scale constant c comes from the configuration ("fixed"), from the
        root's tuning ("tune_root" after the root), or from tuning at this
        node; the temperature is then c times the kNN scale of X."""
        config = self.temperature_config
        if config is None or config.strategy == "splitter":
            own = getattr(self.splitter, "temperature", None)
            return None, {"temperature": own, "temperature_c": None,
                          "temperature_tuned_at_node": False}, False
        known = {"fixed": config.c, "tune_root": self._root_temperature_c}.get(config.strategy)
        scores = None
        if known is None:
            known, scores = self._tune_temperature(X, y, None)
            if config.strategy == "tune_root" and depth == 0:
                self._root_temperature_c = known
        tuned = scores is not None
        context = {"temperature": self._temperature_for_c(X, known), "temperature_c": known,
                   "temperature_scale_mode": config.scale_mode,
                   "temperature_tuned_at_node": tuned}
        if tuned and self.store_diagnostics:
            context["temperature_candidate_scores"] = scores
        return context["temperature"], context, tuned

    def _tune_temperature(self, X, y, prepared_design):
        """Hold out a validation fraction and keep the c of lowest validation MSE."""
        config = self.temperature_config
        n_val = max(1, int(round(config.validation_fraction * X.shape[0])))
        order = self._rng.permutation(X.shape[0])
        val, fit = order[:n_val], order[n_val:]
        X_fit, y_fit, X_val, y_val = X[fit], y[fit], X[val], y[val]
        parent_model, parent_predictions = self._fit_node_model(X_fit, y_fit)
        parent_mse = mean_squared_error(y_fit, parent_predictions)
        comparison_seed = self._next_seed()
        temperature_seed = self._next_seed()
        candidates, best = [], None
        for c in config.c_values:
            temperature = self._temperature_for_c(X_fit, c, random_seed=temperature_seed)
            try:
                result = self._run_splitter(X_fit, y_fit, parent_model, parent_mse, temperature,
                                            None, random_seed=comparison_seed)
            except SplitNotFoundError as exc:
                candidates.append({"c": c, "temperature": temperature, "validation_mse": np.inf,
                                   "failure_reason": exc.reason})
                continue
            rejection = self._split_rejection_reason(result)
            if rejection is not None:
                candidates.append({"c": c, "temperature": temperature, "validation_mse": np.inf,
                                   "failure_reason": rejection})
                continue
            mse = mean_squared_error(y_val, result.predict(X_val))
            candidates.append({"c": c, "temperature": temperature, "validation_mse": mse,
                               "failure_reason": None})
            if best is None or mse < best[0]:
                best = (mse, float(c))
        if best is None:
            raise SplitNotFoundError("temperature_tuning_failed",
                                     "All temperature candidates failed to produce a valid split.")
        return best[1], candidates

    def _run_splitter(self, X, y, parent_model, parent_mse, temperature, prepared_design,
                      random_seed=None, min_side_fallback=None, **_workspaces):
        splitter = copy.deepcopy(self.splitter)
        if getattr(splitter, "optimizer", None) == "cupy_batched_cholesky":
            splitter.optimizer = "cupy_cholesky"  # single-node candidates
        if temperature is not None:
            splitter.temperature = float(temperature)
        if min_side_fallback is not None and hasattr(splitter, "min_side_fallback"):
            splitter.min_side_fallback = min_side_fallback
        if hasattr(splitter, "random_state"):
            splitter.random_state = self._next_seed() if random_seed is None else random_seed
        kwargs = {"parent_model": parent_model, "parent_loss": parent_mse}
        if isinstance(splitter, SoftObliqueSplitter) and not self.hard_refits:
            kwargs["hard_refit"] = False
        with self._timing_context("splitter"):
            return splitter.split(X, y, **kwargs)

    def _temperature_for_c(self, X, c, random_seed=None):
        config = self.temperature_config
        method = config.nn_method or ("bruteforce"
                                      if self.dimension >= config.bruteforce_dimension_threshold
                                      else "kdtree")
        return estimate_temperature(X, mode=config.scale_mode, c=c, max_points=config.max_points,
                                    random_state=self._next_seed() if random_seed is None
                                    else random_seed, nn_method=method)

    def _split_rejection_reason(self, result: SplitResult):
        values = [result.parent_loss, result.z, *result.w]
        if result.has_hard_fit:
            values.extend((result.loss, result.split_gain, result.relative_split_gain))
        elif self.check_split_gain:
            raise ValueError("Split-gain checks require final hard refits.")
        if not np.all(np.isfinite(values)):
            return "nonfinite_split"
        if self.check_split_gain and result.split_gain <= self.min_split_gain:
            return "insufficient_split_gain"
        if self.check_split_gain and result.relative_split_gain <= self.min_relative_split_gain:
            return "insufficient_relative_split_gain"
        if result.stop_reason in _NUMERICAL_FAILURES and result.n_iters == 0:
            return "optimizer_failed_without_progress"
        return None

    # ---------------------------------------------------- metadata helpers
    def _split_warnings(self, result):
        warnings = []
        if result.n_iters == 0 and result.stop_reason == "gradient_tolerance":
            warnings.append("small_init_grad")
        if not result.converged:
            warnings.append(result.stop_reason)
        for key in ("alpha_min_saturated", "alpha_max_saturated", "high_model_conditioning",
                    "min_side_fallback"):
            if bool(result.metadata.get(key, False)):
                warnings.append(key)
        return tuple(dict.fromkeys(warnings))

    def _diagnostic_warnings(self, diagnostics):
        if diagnostics is None:
            return ()
        warnings = []
        if diagnostics.get("n_iters") == 0 and diagnostics.get("stop_reason") == "gradient_tolerance":
            warnings.append("small_init_grad")
        if bool(diagnostics.get("high_model_conditioning", False)):
            warnings.append("high_model_conditioning")
        return tuple(warnings)

    @staticmethod
    def _combine_warnings(*groups):
        return tuple(dict.fromkeys(w for group in groups for w in group))

    def _split_summary(self, result):
        return {"hard_refit": result.has_hard_fit, "split_mse": result.loss,
                "split_gain_mse": result.split_gain,
                "relative_split_gain_mse": result.relative_split_gain,
                "split_n_left": result.n_left, "split_n_right": result.n_right,
                "split_converged": result.converged, "split_stop_reason": result.stop_reason,
                "split_iterations": result.n_iters}

    def _node_metadata(self, samples, fit_error, fit_mse, model):
        solve = ridge_solve_diagnostics(model) or {}
        cond = solve.get("cond_estimate")
        high = cond is not None and float(cond) >= MODEL_CONDITION_WARNING_THRESHOLD
        return {"fit_error": fit_error, "fit_error_metric": self._metric_name(),
                "fit_mse": fit_mse, "n_samples": samples.X.shape[0],
                "n_inherited": samples.n_inherited, "n_new": samples.n_new,
                "fit_solver_requested": solve.get("solver_requested"),
                "fit_solver_used": solve.get("solver_used"),
                "fit_condition_estimator": solve.get("condition_estimator"),
                "fit_cond_estimate": cond,
                "fit_ridge_effective": solve.get("ridge_effective"),
                "fit_rank": solve.get("rank"),
                "fit_fallback_reason": solve.get("fallback_reason"),
                "warnings": ("high_model_conditioning",) if high else (),
                **samples.sampling_metadata}

    def _make_leaf(self, model, region, depth, node_id, status, metadata, samples):
        metadata = dict(metadata)
        if self.store_leaf_representative:
            metadata["sample_centroid"] = np.mean(samples.X, axis=0)
        self._store_samples(metadata, samples)
        return LeafNode(model=model, region=region, depth=depth, node_id=node_id, status=status,
                        metadata=metadata)

    def _store_samples(self, metadata, samples):
        if self.store_samples:
            metadata["X"] = samples.X.copy()
            metadata["y"] = samples.y.copy()

    # -------------------------------------------------------------- sampling
    def _sample_root(self, region):
        if isinstance(self.domain, BoxDomain) and self.exact_box_root:
            with self._timing_context("sampling"):
                X = sample_uniform_box(self.domain.bounds, self._target_samples,
                                       random_state=self._next_seed())
            return X, _sampling_info("exact_uniform_box")
        return self._sample_region(region, self._target_samples, x0=None,
                                   sampler=self.root_sampler)

    def _sample_region(self, region, n, *, x0, sampler):
        with self._timing_context("sampling"):
            return self._sample_region_untimed(region, n, x0=x0, sampler=sampler)

    def _sample_region_untimed(self, region, n, *, x0, sampler):
        d = self.dimension
        if not self.isotropic_sampling:
            X = np.asarray(sampler.sample(region, n, random_state=self._next_seed(), x0=x0),
                           dtype=float)
            info = _sampling_info(type(sampler).__name__, getattr(sampler, "thinning", None))
        else:
            if not isinstance(sampler, HitAndRunSampler):
                raise TypeError("isotropic_sampling requires a HitAndRunSampler.")
            # A short unthinned pilot chain estimates the region's covariance,
            # whose eigenvectors then shape the directions of the main chain.
            pilot_length = self.isotropic_pilot_multiplier * d
            pilot_sampler = replace(sampler, burn_in=0, thinning=1, direction_eigenvectors=None,
                                    direction_eigenvalues=None)
            pilot = np.asarray(pilot_sampler.sample(region, pilot_length,
                                                    random_state=self._next_seed(), x0=x0),
                               dtype=float)
            if pilot.shape != (pilot_length, d):
                raise RuntimeError(f"Pilot sampler returned shape {pilot.shape}; "
                                   f"expected {(pilot_length, d)}.")
            vectors, values = sample_covariance_eigendecomposition(pilot)
            safe, floor = floor_covariance_eigenvalues(values, sampler.direction_eigenvalue_floor)
            saturated = bool(np.any(values <= floor))
            condition = float(np.max(safe) / np.min(safe))
            shaped = replace(sampler, direction_eigenvectors=vectors, direction_eigenvalues=values)
            X = np.asarray(shaped.sample(region, n, random_state=self._next_seed(),
                                         x0=pilot[-1]), dtype=float)
            warning = "covariance_eigenvalue_floor_saturated"
            info = {**_sampling_info("isotropic_hit_and_run", sampler.thinning, pilot_length,
                                     condition, (warning,) if saturated else ()),
                    "sampling_covariance_eigenvalue_floor": floor,
                    "sampling_covariance_floor_saturated": saturated}
        if X.shape != (n, d):
            raise RuntimeError(f"Sampler returned shape {X.shape}; expected {(n, d)}.")
        if not np.all(region.contains(X)):
            raise RuntimeError("Sampler returned points outside the region.")
        return X, info

    # ------------------------------------------------------------ utilities
    def _timing_context(self, category):
        if self._timing_profile is None:
            return nullcontext()
        return self._timing_profile.measure(category)

    def _query_oracle(self, X):
        X = np.asarray(X, dtype=float)
        if self.oracle_vectorized:
            values = np.asarray(self.oracle(X), dtype=float).reshape(-1)
            if values.shape[0] != X.shape[0]:
                raise ValueError("A vectorized oracle must return one value per input row.")
            self.oracle_queries += X.shape[0]
            return values
        values = np.empty(X.shape[0])
        for i, x in enumerate(X):
            values[i] = float(self.oracle(x))
            self.oracle_queries += 1
        return values

    def _resolve_effective_dimension(self):
        if self.sample_count is not None:
            try:
                return model_effective_dimension(self.model_template, self.dimension)
            except TypeError:
                return None
        return model_effective_dimension(self.model_template, self.dimension)

    def _resolve_sample_count(self):
        if self.sample_count is None:
            return self.sample_multiplier * int(self._effective_dimension)
        if callable(self.sample_count):
            return int(self.sample_count(self.model_template, self.dimension))
        return int(self.sample_count)

    def _metric_name(self):
        return getattr(self.error_metric, "__name__", type(self.error_metric).__name__)

    def _next_seed(self):
        return int(self._rng.integers(0, 2**31 - 1))
