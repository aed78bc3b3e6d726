"""Fitted regression trees and their batched GPU inference (art/tree.py).

``LeafNode``, ``SplitNode`` and ``RegressionTree`` keep the reference's
fields, node-id scheme ("root", "root/L", "root/R/L", ...), traversal order
and joblib artifact layout, so trees written by either side read back the
same way.  ``RegressionTree.predict`` flattens the tree into split and leaf
tables and runs ``art_tree_predict`` (one lane per point: route from the
root, then evaluate the leaf's ridge model on the fused feature basis) when
every leaf holds a GPU ridge model of one basis; a tree whose leaves hold
other model types is routed on the host through each model's ``predict``,
as the reference does for any model.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path
from typing import Any, Iterator, Mapping, Optional, Union

import numpy as np

from .domain import PolytopeRegion

TREE_ARTIFACT_TYPE = "art.regression_tree"
TREE_ARTIFACT_VERSION = 1


@dataclass
class LeafNode:
    model: Any
    region: Optional[PolytopeRegion] = None
    depth: int = 0
    node_id: str = "root"
    status: str = "leaf"
    metadata: dict[str, Any] = field(default_factory=dict)

    def predict_one(self, x):
        return float(self.model.predict(np.asarray(x, dtype=float).reshape(1, -1))[0])


TreeNode = Union["SplitNode", LeafNode]


@dataclass
class SplitNode:
    w: np.ndarray
    z: float
    left: TreeNode
    right: TreeNode
    region: Optional[PolytopeRegion] = None
    depth: int = 0
    node_id: str = "root"
    status: str = "split"
    metadata: dict[str, Any] = field(default_factory=dict)

    def __post_init__(self):
        self.w = np.asarray(self.w, dtype=float).reshape(-1)
        self.z = float(self.z)

    def route(self, x):
        """The child a point falls in: right when w.x - z >= 0 (art/tree.py:45-50)."""
        x = np.asarray(x, dtype=float).reshape(-1)
        if x.shape[0] != self.w.shape[0]:
            raise ValueError(f"x must have shape ({self.w.shape[0]},), got {x.shape}.")
        return self.right if float(self.w @ x - self.z) >= 0.0 else self.left


def _gpu_basis(leaves):
    """(degree, bias, bias_last, p) shared by every leaf's GPU ridge model, or None."""
    from .models import AffineRidgeModel, PolynomialRidgeModel, model_basis

    basis = None
    for leaf in leaves:
        m = leaf.model
        if not isinstance(m, (AffineRidgeModel, PolynomialRidgeModel)) or m.coef_ is None:
            return None
        try:
            b = model_basis(m)
        except TypeError:
            return None
        if b[0] not in (1, 2):
            return None
        if basis is None:
            basis = b
        elif b != basis:
            return None
    return basis


class FlatTree:
    """Device tables of a tree for ``art_tree_predict``.

    Split nodes are numbered breadth-first; a child link c >= 0 is a split
    row and c < 0 is leaf -1-c.  ``root`` is the root's link.
    """

    def __init__(self, root, d, basis):
        splits, leaves, links = [], [], {}
        queue = [root]
        while queue:
            node = queue.pop(0)
            if isinstance(node, SplitNode):
                links[id(node)] = len(splits)
                splits.append(node)
                queue.extend((node.left, node.right))
            else:
                links[id(node)] = -1 - len(leaves)
                leaves.append(node)
        self.leaf_nodes = leaves
        degree, bias, bias_last = basis
        p = int(np.asarray(leaves[0].model.coef_).shape[0])
        w = np.zeros((max(len(splits), 1), d))
        z = np.zeros(max(len(splits), 1))
        child = np.zeros((max(len(splits), 1), 2), dtype=np.int32)
        for k, s in enumerate(splits):
            if s.w.shape[0] != d:
                raise ValueError(f"Split {s.node_id} has {s.w.shape[0]} weights, expected {d}.")
            w[k], z[k] = s.w, s.z
            child[k] = links[id(s.left)], links[id(s.right)]
        coef = np.stack([np.asarray(leaf.model.coef_, dtype=float).reshape(-1) for leaf in leaves])
        if coef.shape[1] != p:
            raise ValueError("Leaf models have different coefficient counts.")
        self._upload(w, z, child, links[id(root)], coef, d, basis, len(splits))

    @classmethod
    def from_tables(cls, w, z, child, root, coef, d, basis, *, upload=None):
        """Device tables from host arrays already in this layout: w (S, d),
        z (S,), child links (S, 2) int32, root link, leaf coefficients (L, p)
        and basis (degree, bias, bias_last); S >= 1 rows (a placeholder row
        when the tree is one leaf).  upload(array, device, dtype): one
        non-blocking staged copy (``_hip.upload`` or a ``PinnedRing``'s), for
        callers that use the tables on the stream it enqueues on."""
        self = cls.__new__(cls)
        self.leaf_nodes = None
        n_splits = int(np.asarray(child).shape[0]) if root >= 0 else 0
        self._upload(np.asarray(w, dtype=float), np.asarray(z, dtype=float),
                     np.asarray(child, dtype=np.int32), int(root), np.asarray(coef, dtype=float),
                     int(d), basis, n_splits, upload)
        return self

    def _upload(self, w, z, child, root, coef, d, basis, n_splits, upload=None):
        import torch

        from . import _hip

        degree, bias, bias_last = basis
        self.torch, self.lib = _hip.torch_device(), _hip.load()
        dev = torch.device("cuda", torch.cuda.current_device())
        if upload is not None:
            # One staging copy: the float tables, then the int32 links as
            # float64-sized slots reinterpreted on the device.
            parts = [np.ascontiguousarray(w).reshape(-1), np.ascontiguousarray(z).reshape(-1),
                     np.ascontiguousarray(coef).reshape(-1)]
            links = np.ascontiguousarray(child, dtype=np.int32).reshape(-1)
            pad = np.zeros(links.size + (links.size & 1), dtype=np.int32)
            pad[:links.size] = links
            parts.append(pad.view(np.float64))
            blob = upload(np.concatenate(parts), dev, np.float64)
            nw, nz, nc = w.size, z.size, coef.size
            self.w = blob[:nw].view(w.shape)
            self.z = blob[nw:nw + nz]
            self.coef = blob[nw + nz:nw + nz + nc].view(coef.shape)
            self.child = blob[nw + nz + nc:].view(torch.int32)[:links.size].view(child.shape)
        else:
            self.w = torch.from_numpy(np.ascontiguousarray(w)).to(dev)
            self.z = torch.from_numpy(np.ascontiguousarray(z)).to(dev)
            self.child = torch.from_numpy(np.ascontiguousarray(child)).to(dev)
            self.coef = torch.from_numpy(np.ascontiguousarray(coef)).to(dev)
        self.root = root
        self.d, self.device = d, dev
        self.basis = _hip.Basis(d, coef.shape[1], degree, int(bias), int(bias_last))
        self.n_splits, self.n_leaves = n_splits, coef.shape[0]

    def leaves(self, X, out=None):
        """Leaf index (breadth-first leaf numbering) of device rows X (n, d)."""
        from . import _hip

        torch = self.torch
        n = int(X.shape[0])
        if out is None:
            out = torch.empty(n, dtype=torch.int32, device=self.device)
        _hip.check(self.lib.art_tree_leaf(
            _hip.ptr(X), n, self.d, _hip.ptr(self.w), _hip.ptr(self.z), _hip.ptr(self.child),
            self.root, _hip.ptr(out), _hip.stream_handle(torch.cuda.current_stream())))
        return out

    def predict(self, X, out=None):
        """Predictions at device rows X (n, d) float64; returns a device tensor."""
        from . import _hip

        torch = self.torch
        n = int(X.shape[0])
        if out is None:
            out = torch.empty(n, dtype=torch.float64, device=self.device)
        _hip.check(self.lib.art_tree_predict(
            _hip.ptr(X), n, _hip.ptr(self.w), _hip.ptr(self.z), _hip.ptr(self.child), self.root,
            _hip.ptr(self.coef), self.basis, _hip.ptr(out),
            _hip.stream_handle(torch.cuda.current_stream())))
        return out


class RegressionTree:
    """A fitted tree: prediction, traversal and artifacts (art/tree.py:53-241)."""

    def __init__(self, root=None, oracle_queries=0, metadata=None):
        self.root = root
        self.oracle_queries = int(oracle_queries)
        self.metadata = {} if metadata is None else dict(metadata)

    def save(self, path, run_config: Mapping[str, Any] | None = None, compress=3):
        import joblib

        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        joblib.dump({"artifact_type": TREE_ARTIFACT_TYPE, "format_version": TREE_ARTIFACT_VERSION,
                     "tree": self, "run_config": {} if run_config is None else dict(run_config)},
                    path, compress=compress)
        return path

    @classmethod
    def load(cls, path):
        """Load a trusted artifact written by ``save`` (joblib unpickles it)."""
        import joblib

        path = Path(path)
        artifact = joblib.load(path)
        if not isinstance(artifact, dict):
            raise TypeError(f"{path} is not a regression-tree artifact.")
        if artifact.get("artifact_type") != TREE_ARTIFACT_TYPE:
            raise ValueError(f"{path} has an unrecognized artifact type.")
        version = artifact.get("format_version")
        if version != TREE_ARTIFACT_VERSION:
            raise ValueError(f"Unsupported tree artifact version {version!r}; "
                             f"expected {TREE_ARTIFACT_VERSION}.")
        tree = artifact.get("tree")
        if not isinstance(tree, cls):
            raise TypeError(f"{path} does not contain a RegressionTree.")
        run_config = artifact.get("run_config", {})
        if not isinstance(run_config, dict):
            raise TypeError(f"{path} contains an invalid run configuration.")
        return tree, run_config

    def flatten(self):
        """Device tables for batched prediction (GPU ridge leaves of one basis)."""
        if self.root is None:
            raise ValueError("Tree has no root node.")
        basis = _gpu_basis(self.iter_leaves())
        if basis is None:
            raise TypeError("Batched GPU prediction needs fitted AffineRidgeModel or "
                            "PolynomialRidgeModel leaves of one basis.")
        d = self._dimension()
        return FlatTree(self.root, d, basis)

    def _dimension(self):
        for node in self.iter_nodes():
            if isinstance(node, SplitNode):
                return int(node.w.shape[0])
            if node.region is not None:
                return int(node.region.dimension)
        m = self.root.model
        coef = np.asarray(m.coef_).reshape(-1)
        from .models import model_basis

        degree, bias, _ = model_basis(m)
        p = coef.shape[0] - int(bias)
        return p if degree == 1 else int(round((-3 + np.sqrt(9 + 8 * p)) / 2))

    def predict_one(self, x):
        return self.leaf_for_point(x).predict_one(x)

    def predict(self, X):
        X = np.asarray(X, dtype=float)
        if X.ndim == 1:
            X = X.reshape(1, -1)
        if X.ndim != 2:
            raise ValueError(f"X must have shape (n, d), got {X.shape}.")
        if X.shape[0] == 0:
            return np.empty(0, dtype=float)
        if self.root is None:
            raise ValueError("Tree has no root node.")
        if _gpu_basis(self.iter_leaves()) is not None:
            import torch

            flat = self.flatten()
            if X.shape[1] != flat.d:
                raise ValueError(f"X must have {flat.d} columns, got {X.shape[1]}.")
            Xd = torch.from_numpy(np.ascontiguousarray(X)).to(flat.device)
            return flat.predict(Xd).cpu().numpy()
        return self._predict_models(X)

    def _predict_models(self, X):
        """Host routing through each leaf model's own predict (any model type)."""
        out = np.empty(X.shape[0])
        pending = [(self.root, np.arange(X.shape[0]))]
        while pending:
            node, idx = pending.pop()
            if isinstance(node, LeafNode):
                vals = np.asarray(node.model.predict(X[idx]), dtype=float).reshape(-1)
                if vals.shape[0] != idx.shape[0]:
                    raise ValueError(f"Leaf model returned {vals.shape[0]} predictions "
                                     f"for {idx.shape[0]} samples.")
                out[idx] = vals
                continue
            if X.shape[1] != node.w.shape[0]:
                raise ValueError(f"X must have {node.w.shape[0]} columns, got {X.shape[1]}.")
            right = X[idx] @ node.w - node.z >= 0.0
            if np.any(~right):
                pending.append((node.left, idx[~right]))
            if np.any(right):
                pending.append((node.right, idx[right]))
        return out

    def leaf_for_point(self, x):
        if self.root is None:
            raise ValueError("Tree has no root node.")
        node = self.root
        while isinstance(node, SplitNode):
            node = node.route(x)
        return node

    def iter_nodes(self) -> Iterator[TreeNode]:
        """Depth-first, left before right."""
        if self.root is None:
            return
        stack = [self.root]
        while stack:
            node = stack.pop()
            yield node
            if isinstance(node, SplitNode):
                stack.extend([node.right, node.left])

    def iter_leaves(self) -> Iterator[LeafNode]:
        for node in self.iter_nodes():
            if isinstance(node, LeafNode):
                yield node

    def path_to_node(self, node_id):
        if not isinstance(node_id, str) or not node_id:
            raise ValueError("node_id must be a nonempty string.")
        missing = KeyError(f"Tree does not contain a node with id {node_id!r}.")
        if self.root is None:
            raise missing
        if node_id == self.root.node_id:
            return (self.root,)
        prefix = f"{self.root.node_id}/"
        if not node_id.startswith(prefix):
            raise missing
        path, node = [self.root], self.root
        for step in node_id[len(prefix):].split("/"):
            if step not in ("L", "R") or not isinstance(node, SplitNode):
                raise missing
            node = node.left if step == "L" else node.right
            path.append(node)
        if node.node_id != node_id:
            raise missing
        return tuple(path)

    def get_node(self, node_id):
        return self.path_to_node(node_id)[-1]

    def num_nodes(self):
        return sum(1 for _ in self.iter_nodes())

    def num_split_nodes(self):
        return sum(isinstance(n, SplitNode) for n in self.iter_nodes())

    def num_leaves(self):
        return sum(1 for _ in self.iter_leaves())

    def max_depth(self):
        return max((leaf.depth for leaf in self.iter_leaves()), default=0)
