"""mbx_amd -- MI355X-native MBX many-body short-range (2B/3B) engine.

Layers:
  * ``libmbxg.so`` (mbx_amd/csrc): HIP kernels for gfx950 behind the C ABI in include/mbxg.h,
    and the C++ ``bblock::System`` drop-in plus the reference-compatible extern "C" entry points
    (include/mbx/).
  * :mod:`mbx_amd.engine`: ctypes binding of the C ABI.
"""
from .engine import Engine, MbxgError, water_engine  # noqa: F401
