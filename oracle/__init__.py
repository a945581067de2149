"""Test infrastructure: FP64 CPU restatement of PyCuAmpcor's dense offset-tracking pipeline.

This package is the checker. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product path
(``isce2_amd``) never does, and fails loudly when its HIP library is missing.
"""
from .ampcor_oracle import *  # noqa: F401,F403
