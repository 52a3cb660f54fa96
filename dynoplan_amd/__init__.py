"""dynoplan_amd: MI355X (gfx950) kernels for Dynoplan's db-A* node expansion.

The product is the C-ABI library ``libdynoplan_amd.so`` (include/dynoplan_amd.h);
this package holds its build recipe, a ctypes view, the seed-deterministic
synthetic workloads and the host-side expansion context.
"""
from . import _native  # noqa: F401
from ._native import DynoplanError, load  # noqa: F401
