"""CPU oracle for the montepath Monte Carlo path engine (test infrastructure).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this package, and only as the checker or the
timed CPU baseline -- never as the thing measured or shipped.  The product
library (``montepath_amd``) does not import it.
"""
