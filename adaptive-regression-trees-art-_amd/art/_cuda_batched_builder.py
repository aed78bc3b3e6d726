"""One resident GPU optimizer serving every cohort of a BFS build
(art/_cuda_batched_builder.py:18-101).

The dispatcher owns a capacity workspace of node_batch_size nodes.  A cohort
is loaded once; each optimize call runs every pending node of the cohort in
one device launch (random starts drawn on the device), and min-side retries
re-run only the selected resident slots, with the fallback enabled on a
node's last allowed attempt.  Temperature choice, acceptance and retry
policy stay with the builder.
"""

from __future__ import annotations

import copy

import numpy as np

from .batched_optimizer import BatchedOptimizerWorkspace
from .splitters import SplitNotFoundError


class BatchedSplitDispatcher:
    """``split_batch`` of ``build_bfs``: ready jobs in, finished nodes out."""

    def __init__(self, builder):
        self.builder = builder
        splitter = copy.deepcopy(builder.splitter)
        splitter.random_state = builder._next_seed()  # one draw from the build stream
        with builder._timing_context("splitter"):
            self.ws = BatchedOptimizerWorkspace.allocate(
                splitter, builder.node_batch_size, builder.target_samples, builder.dimension,
                reorder=builder.batch_reorder)

    def __call__(self, ready):
        b, ws = self.builder, self.ws
        nodes = [None] * len(ready)
        jobs, states, positions = [], [], []
        for pos, job in enumerate(ready):
            try:
                state = b._split_state(job.samples.X, job.samples.y, job.depth, job.prepared)
            except SplitNotFoundError as exc:
                nodes[pos] = b._finish_node(job, exc)
                continue
            jobs.append(job)
            states.append(state)
            positions.append(pos)
        if not jobs:
            return nodes
        k = len(jobs)
        parent_loss = np.array([job.parent_mse for job in jobs])
        temperature = np.array([ws.s.temperature if s.temperature is None else s.temperature
                                for s in states])
        outcomes = [None] * k
        with b._timing_context("splitter"):
            for j, job in enumerate(jobs):
                ws.host_X[j] = job.samples.X
                ws.host_y[j] = job.samples.y
            ws.load(ws.host_X[:k], ws.host_y[:k], temperature=temperature)
            pending = list(range(k))
            while pending:
                again = []
                for fallback in (False, True):
                    ids = [j for j in pending if states[j].fallback == fallback]
                    if not ids:
                        continue
                    results = ws.optimize(nodes=ids, parent_loss=parent_loss[ids],
                                          min_side_fallback=fallback, hard_refit=b.hard_refits)
                    for j, result in zip(ids, results):
                        if b._retry_split(states[j], result):
                            again.append(j)
                        else:
                            outcomes[j] = result
                pending = sorted(again)
        for pos, job, state, result in zip(positions, jobs, states, outcomes):
            if isinstance(result, SplitNotFoundError):
                nodes[pos] = b._finish_node(job, state.error(result))
            else:
                nodes[pos] = b._finish_node(job, result, {**state.context, **state.metadata})
        return nodes

    def close(self):
        self.ws.close()

    def __enter__(self):
        return self

    def __exit__(self, *args):
        self.close()
