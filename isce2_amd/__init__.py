"""MI355X-native dense offset tracking (the PyCuAmpcor path of isce2).

``from isce2_amd import PyCuAmpcor`` gives the drop-in class; the compute runs
in hand-written HIP kernels of ``libisce2_amd.so`` (built by
``python -m isce2_amd.build``).
"""
from .PyCuAmpcor import PyCuAmpcor  # noqa: F401

__version__ = "0.1.0"
