"""MI355X-native NDDO (MNDO/AM1/PM3) ground-state SCF engine.

The compute path is libnddo.so (hand-written gfx950 HIP kernels behind the C ABI
of include/nddo.h); this package holds its ctypes binding, the build script and
the synthetic workload generators.
"""
from .nddo import Engine, Job, NddoError, load_library  # noqa: F401
