"""Host-side builder logic: FIFO cohorts, validation, metrics and
temperature configuration (no device work)."""

import numpy as np
import pytest

from art._batched_builder import build_bfs
from art.builder import RegressionTreeBuilder, _NodeSamples, _NodeWork
from art.domain import BoxDomain
from art.metrics import (max_pointwise_relative_error, mean_squared_error,
                         median_pointwise_relative_error, relative_l2_error)
from art.models import AffineRidgeModel, PolynomialRidgeModel
from art.splitters import SoftObliqueSplitter
from art.temperature import TemperatureConfig, subsample_points
from art.tree import LeafNode, RegressionTree, SplitNode


class _Model:
    def predict(self, X):
        return np.zeros(len(X))


class CompleteTreeBuilder:
    """Builder double: every node above max_depth splits; records the
    preparation order and the cohorts handed to the split step."""

    def __init__(self, node_batch_size, max_depth=3):
        self.node_batch_size, self.max_depth = node_batch_size, max_depth
        self.prepared, self.cohorts = [], []

    def _prepare_cohort(self, jobs):
        out = []
        for job in jobs:
            self.prepared.append(job.node_id)
            out.append(LeafNode(_Model(), depth=job.depth, node_id=job.node_id)
                       if job.depth >= self.max_depth else None)
        return out

    def _split_node(self, job):
        return SplitNode(np.array([1.0]), 0.0, LeafNode(_Model(), node_id=job.node_id + "/L"),
                         LeafNode(_Model(), node_id=job.node_id + "/R"), depth=job.depth,
                         node_id=job.node_id)

    def _children(self, job, node):
        for side, tag in (("left", "L"), ("right", "R")):
            yield _NodeWork(None, _NodeSamples(np.zeros((0, 1)), np.zeros(0), 0, 0),
                            job.depth + 1, f"{job.node_id}/{tag}", parent=node, side=side)


@pytest.mark.parametrize("size,expected", [(1, [1] * 7), (3, [1, 2, 3, 1]), (20, [1, 2, 4])])
def test_bfs_cohort_sizes_match_the_reference(size, expected):
    # Cohort-size known answers of the reference (tests/test_batched_builder.py:74-102):
    # a complete depth-3 tree under node_batch_size 1, 3 and 20.
    b = CompleteTreeBuilder(size)

    def dispatch(ready):
        b.cohorts.append([job.node_id for job in ready])
        return [b._split_node(job) for job in ready]

    root = build_bfs(b, _NodeWork(None, None, 0, "root"), dispatch)
    assert [len(c) for c in b.cohorts] == expected
    tree = RegressionTree(root)
    assert tree.num_nodes() == 15 and tree.num_leaves() == 8 and tree.max_depth() == 3
    # Breadth-first preparation order: depth, then left before right.
    ids = [n.node_id for n in tree.iter_nodes()]
    assert b.prepared == sorted(ids, key=lambda i: (i.count("/"), i))


def test_bfs_without_dispatch_splits_node_by_node():
    b = CompleteTreeBuilder(4, max_depth=2)
    root = build_bfs(b, _NodeWork(None, None, 0, "root"))
    assert RegressionTree(root).num_leaves() == 4


def _builder(**kw):
    opts = dict(domain=BoxDomain.hypercube(2, -1.0, 1.0), oracle=lambda X: X[..., 0],
                oracle_vectorized=True, model_template=AffineRidgeModel(1e-8),
                splitter=SoftObliqueSplitter(AffineRidgeModel(1e-8), 0.1), error_threshold=0.0,
                sample_count=40)
    opts.update(kw)
    return RegressionTreeBuilder(**opts)


@pytest.mark.parametrize("kw,err,match", [
    (dict(traversal="wide"), ValueError, "traversal must be"),
    (dict(node_batch_size=0), ValueError, "node_batch_size must be a positive integer"),
    (dict(node_batch_size=True), ValueError, "node_batch_size must be a positive integer"),
    (dict(batch_reorder=1), ValueError, "batch_reorder must be a boolean"),
    (dict(error_threshold=-1.0), ValueError, "error_threshold must be nonnegative"),
    (dict(max_depth=-1), ValueError, "max_depth must be nonnegative"),
    (dict(validation_fraction=1.0), ValueError, "validation_fraction must satisfy"),
    (dict(sample_count=1), ValueError, "node sample count must be at least 2"),
    (dict(max_retries_on_failure=1,
          splitter=SoftObliqueSplitter(AffineRidgeModel(), 0.1, n_restarts=4)),
     ValueError, "requires splitter.n_restarts == 1"),
    (dict(splitter=SoftObliqueSplitter(AffineRidgeModel(), 0.1,
                                       optimizer="cupy_batched_cholesky")),
     ValueError, "requires traversal='bfs'"),
    (dict(validation_fraction=0.2), ValueError, "requires max_pointwise_relative_error"),
    (dict(temperature_config=TemperatureConfig(strategy="tune_node"), sample_count=3),
     ValueError, "Temperature tuning requires at least four node samples"),
])
def test_builder_validation_messages(kw, err, match):
    with pytest.raises(err, match=match):
        _builder(**kw)


def test_builder_sample_budget_and_refit_policy():
    b = _builder(sample_count=None, sample_multiplier=7,
                 model_template=PolynomialRidgeModel(ridge=1e-6))
    assert b.target_samples == 7 * 6  # quadratic basis in d = 2: 6 features
    assert not b.hard_refits  # one restart, splitter temperature
    assert _builder(check_split_gain=True).hard_refits
    s4 = SoftObliqueSplitter(AffineRidgeModel(), 0.1, n_restarts=4)
    assert _builder(splitter=s4).hard_refits


def test_metric_known_answers():
    y, yhat = np.array([3.0, -4.0]), np.array([3.0, -3.0])
    assert mean_squared_error(y, yhat) == 0.5
    assert relative_l2_error(y, yhat) == pytest.approx(1.0 / 5.0)
    assert relative_l2_error(np.zeros(2), np.array([1e-13, 0.0])) == pytest.approx(0.1)
    assert max_pointwise_relative_error(y, yhat) == pytest.approx(0.25)
    assert median_pointwise_relative_error(y, yhat) == pytest.approx(0.125)
    with pytest.raises(ValueError):
        mean_squared_error(np.zeros(0), np.zeros(0))


def test_temperature_config_and_subsample():
    with pytest.raises(ValueError, match="Unknown temperature strategy"):
        TemperatureConfig(strategy="auto")
    with pytest.raises(ValueError, match="c must be positive"):
        TemperatureConfig(c=0.0)
    with pytest.raises(ValueError, match="validation_fraction"):
        TemperatureConfig(validation_fraction=0.5)
    X = np.arange(40.0).reshape(20, 2)
    assert subsample_points(X, None) is not None and subsample_points(X, 30).shape == (20, 2)
    sub = subsample_points(X, 5, random_state=1)
    assert sub.shape == (5, 2) and len({tuple(r) for r in sub}) == 5


def test_tree_navigation_without_device():
    leaves = [LeafNode(_Model(), depth=1, node_id=i) for i in ("root/L", "root/R")]
    tree = RegressionTree(SplitNode(np.array([1.0, 0.0]), 0.25, leaves[0], leaves[1],
                                    node_id="root"))
    assert tree.leaf_for_point([0.5, 0.0]).node_id == "root/R"
    assert tree.leaf_for_point([0.0, 9.0]).node_id == "root/L"
    assert tree.get_node("root/R") is leaves[1]
    with pytest.raises(KeyError):
        tree.get_node("root/L/L")
    assert tree.num_split_nodes() == 1 and tree.max_depth() == 1
