"""FP64 CPU restatement of the reference ART hot path -- TEST INFRASTRUCTURE ONLY.

This package is the checker that the HIP kernels are compared against.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  The product path (the ``art`` package and
``libart_hip.so``) never imports, links or calls anything under ``oracle/``
and fails loudly when the HIP library is missing.

Each function cites the reference file:line whose arithmetic it restates.
How each function is pinned (golden vectors, known-answer tests, closed
forms) is listed in the module headers; a function without such a check says
"parity unpinned".
"""
