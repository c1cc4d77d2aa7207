"""CPU oracle for the wavelet-tree hot path (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg import this package.  The product (``ecl_wavelet_amd``)
never does.
"""
from .oracle import *  # noqa: F401,F403
