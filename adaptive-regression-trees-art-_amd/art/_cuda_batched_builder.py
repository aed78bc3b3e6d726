"""The split step of a breadth-first build on the device.

``BatchedSplitDispatcher(builder)`` is the ``split_batch`` callable of
``build_bfs`` and keeps the reference's entry point
(art/_cuda_batched_builder.py:18-101): ``__call__(ready)`` takes a cohort of
prepared node jobs and returns one finished node per job, in input order.

How a cohort runs here:

* the builder opens a split ledger per node (``_open_ledger``: temperature
  pick and retry budget; every random draw of the build stream happens
  here, in job order, before any attempt);
* the cohort's rows -- nodes of any sizes -- are packed into ONE ragged
  device workspace (rows concatenated, CSR offsets, per-node side
  minimums), loaded once;
* each round of attempts is one optimizer launch per min-side fallback
  setting over the nodes still attempting; the builder settles every
  outcome (``_settle_attempt``) and a retried node reruns on its resident
  rows with fresh device starts;
* finished outcomes become nodes through ``builder._finish_node``.

The build stream gives the dispatcher one seed; each cohort's device
starts are drawn from a generator of the dispatcher's own.
"""

from __future__ import annotations

import copy

import numpy as np

from .batched_optimizer import RaggedOptimizerWorkspace
from .splitters import SplitNotFoundError


class BatchedSplitDispatcher:
    """``split_batch`` of ``build_bfs``: ready jobs in, finished nodes out."""

    def __init__(self, builder):
        self.builder = builder
        self.splitter = copy.deepcopy(builder.splitter)
        self.rng = np.random.default_rng(builder._next_seed())
        self.launches = 0
        self.cohorts = 0

    def __call__(self, ready):
        b = self.builder
        finished = [None] * len(ready)
        pending = []  # (position in ready, job, ledger)
        for pos, job in enumerate(ready):
            try:
                ledger = b._open_ledger(job.samples.X, job.samples.y, job.depth)
            except SplitNotFoundError as exc:
                finished[pos] = b._finish_node(job, exc)
            else:
                pending.append((pos, job, ledger))
        if pending:
            outcomes = self._run_cohort([job for _, job, _ in pending],
                                        [ledger for _, _, ledger in pending])
            for (pos, job, ledger), outcome in zip(pending, outcomes):
                if isinstance(outcome, SplitNotFoundError):
                    finished[pos] = b._finish_node(job, ledger.failure(outcome))
                else:
                    finished[pos] = b._finish_node(job, outcome,
                                                   {**ledger.context, **ledger.summary()})
        return finished

    def _cohort_splitter(self):
        """The configured splitter with this cohort's start seed."""
        splitter = copy.copy(self.splitter)
        splitter.random_state = int(self.rng.integers(0, 2**63 - 1))
        return splitter

    def _run_cohort(self, jobs, ledgers):
        """Attempt rounds until every node has a final outcome; returns the
        outcomes in job order."""
        b = self.builder
        offsets = np.zeros(len(jobs) + 1, dtype=np.int64)
        offsets[1:] = np.cumsum([job.samples.X.shape[0] for job in jobs])
        X = np.concatenate([np.asarray(job.samples.X, dtype=float) for job in jobs])
        y = np.concatenate([np.asarray(job.samples.y, dtype=float).reshape(-1) for job in jobs])
        default_t = self.splitter.temperature
        temperature = np.array([default_t if led.temperature is None else led.temperature
                                for led in ledgers], dtype=float)
        parent_loss = np.array([job.parent_mse for job in jobs], dtype=float)
        outcomes = [None] * len(jobs)
        with b._timing_context("splitter"):
            with RaggedOptimizerWorkspace(self._cohort_splitter(), X, y, offsets,
                                          temperature=temperature) as ws:
                self.cohorts += 1
                attempting = np.arange(len(jobs))
                while attempting.size:
                    flags = np.array([ledgers[j].fallback for j in attempting], dtype=bool)
                    again = []
                    for flag in (False, True):
                        group = attempting[flags == flag]
                        if not group.size:
                            continue
                        results = ws.optimize(nodes=group, parent_loss=parent_loss[group],
                                              min_side_fallback=flag, hard_refit=b.hard_refits)
                        self.launches += 1
                        for j, result in zip(group.tolist(), results):
                            outcomes[j] = result
                            if b._settle_attempt(ledgers[j], result):
                                again.append(j)
                    attempting = np.array(sorted(again), dtype=np.int64)
        return outcomes

    def close(self):
        self.builder = None

    def __enter__(self):
        return self

    def __exit__(self, *args):
        self.close()
