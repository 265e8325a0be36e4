"""MI355X (gfx950) implementation of stab's Chebyshev-collocation stability solve.

The arithmetic runs in libstab_hip.so (hand-written HIP kernels behind the C ABI
of include/stab_hip.h); this package holds the host-side mirror of the
reference's interface (solver.py), the mean-flow similarity solver that writes
profile files (meanflow.py) and the build step (build.py).
"""
from . import _lib  # noqa: F401
