"""FIFO cohort scheduling of a tree build (art/_batched_builder.py:20-62).

A ``_Frontier`` holds the queue of node jobs, oldest first, and the place
every finished node goes (the root slot, or its parent's child field).  A
cohort is filled with up to node_batch_size jobs that need a split: the
frontier hands the builder as many jobs at a time as there are free cohort
places, and the builder prepares them together (sampling in queue order,
then one batched GPU model fit); jobs that end as leaves are placed at once
and free their place again.  The cohort is then split by ``split_batch``
(one GPU launch per attempt round) or node by node, and its nodes are placed
in input order, each split's children queued left then right.
"""

from __future__ import annotations

from collections import deque
from typing import TYPE_CHECKING, Callable

from .tree import SplitNode, TreeNode

if TYPE_CHECKING:
    from .builder import RegressionTreeBuilder, _NodeWork


class _Frontier:
    """The build's job queue and the slots its finished nodes fill."""

    def __init__(self, root_job):
        self.queue = deque([root_job])
        self.root = None

    def place(self, job, node):
        if job.parent is None:
            self.root = node
        else:
            setattr(job.parent, job.side, node)

    def next_cohort(self, builder) -> list:
        """Jobs (oldest first) that still need a split, at most
        node_batch_size; prepared in batches of the free places."""
        limit = builder.node_batch_size
        cohort = []
        while self.queue and len(cohort) < limit:
            take = min(limit - len(cohort), len(self.queue))
            batch = [self.queue.popleft() for _ in range(take)]
            for job, leaf in zip(batch, builder._prepare_cohort(batch)):
                if leaf is None:
                    cohort.append(job)
                else:
                    self.place(job, leaf)
        return cohort


def build_bfs(builder: RegressionTreeBuilder, work: _NodeWork,
              split_batch: Callable[[list[_NodeWork]], list[TreeNode]] | None = None) -> TreeNode:
    """Build breadth-first from the root job and return the root node."""
    frontier = _Frontier(work)
    if split_batch is None:
        def split_batch(jobs):
            return [builder._split_node(job) for job in jobs]
    while frontier.queue:
        cohort = frontier.next_cohort(builder)
        if not cohort:
            continue
        nodes = split_batch(cohort)
        if len(nodes) != len(cohort):
            raise RuntimeError("Split dispatch must return one node per ready job.")
        for job, node in zip(cohort, nodes):
            frontier.place(job, node)
            if isinstance(node, SplitNode):
                frontier.queue.extend(builder._children(job, node))
    if frontier.root is None:
        raise RuntimeError("The build queue finished without a root node.")
    return frontier.root
