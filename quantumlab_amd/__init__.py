"""quantumlab_amd: MI355X-native QuEST state-vector core.

The product is the C-ABI shared library ``quantumlab_amd/lib/libquest_mi355x.so``
(QuEST 3.4's C API on hand-written HIP kernels for gfx950, declared in
include/QuEST.h and include/quest_mi355x.h). This package holds its sources
(csrc/), its ctypes binding (quest.py) and the host-side workload builders.
"""
from .quest import QuEST, QuESTError, load, library_path  # noqa: F401
