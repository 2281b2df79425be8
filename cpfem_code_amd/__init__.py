"""MI355X-native FASOLT3 per-increment hot path (crystal-plasticity FEM).

The compute path is the HIP library ``lib/libcpfem.so`` (gfx950), reached
through the C ABI declared in ``include/cpfem.h``; :mod:`cpfem_code_amd.capi`
binds it with ctypes.  There is no CPU fallback: every entry point raises when
the library or a GPU is missing.
"""
from .problem import Control, Material, Problem  # noqa: F401
