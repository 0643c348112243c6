"""CPU oracle for the CHNS-2D ETD2RK step (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
path (``ternary-chns_amd``) never imports, links or calls it.
"""
from .chns_oracle import *  # noqa: F401,F403
