"""FIFO cohort scheduling of a tree build (art/_batched_builder.py:20-62).

Jobs leave the queue oldest first.  Up to node_batch_size splittable nodes
form a cohort: the queue hands the builder as many jobs at a time as there
are free cohort slots, and the builder prepares them together (sampling in
queue order, then one batched GPU model fit); terminal nodes are attached at
once and free their slot again.  A cohort is split by ``split_batch`` (one
GPU launch for the batched optimizer) or node by node, and its results are
attached in input order, each split's children queued left then right.
"""

from __future__ import annotations

from collections import deque
from typing import TYPE_CHECKING, Callable

from .tree import SplitNode, TreeNode

if TYPE_CHECKING:
    from .builder import RegressionTreeBuilder, _NodeWork


def build_bfs(builder: RegressionTreeBuilder, work: _NodeWork,
              split_batch: Callable[[list[_NodeWork]], list[TreeNode]] | None = None) -> TreeNode:
    """Build breadth-first from the root job and return the root node."""
    pending = deque([work])
    slot = {"root": None}

    def attach(job, node):
        if job.parent is None:
            slot["root"] = node
        else:
            setattr(job.parent, job.side, node)

    while pending:
        ready = []
        while pending and len(ready) < builder.node_batch_size:
            take = min(builder.node_batch_size - len(ready), len(pending))
            jobs = [pending.popleft() for _ in range(take)]
            for job, leaf in zip(jobs, builder._prepare_cohort(jobs)):
                if leaf is None:
                    ready.append(job)
                else:
                    attach(job, leaf)
        if not ready:
            continue
        nodes = ([builder._split_node(job) for job in ready] if split_batch is None
                 else split_batch(ready))
        if len(nodes) != len(ready):
            raise RuntimeError("Split dispatch must return one node per ready job.")
        for job, node in zip(ready, nodes):
            attach(job, node)
            if isinstance(node, SplitNode):
                pending.extend(builder._children(job, node))
    if slot["root"] is None:
        raise RuntimeError("The build queue finished without a root node.")
    return slot["root"]
