"""lgx: MI355X-native screen-space indirect lighting (LegitEngine SSVGI path).

The passes run as hand-written HIP kernels for gfx950 behind the C ABI in
include/lgx.h (liblgx.so).  `lgx` is the ctypes binding; `device` holds
torch-backed device buffers for tests and benchmarks.
"""
__version__ = "0.1.0"
