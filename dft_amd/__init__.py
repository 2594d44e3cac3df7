"""MI355X-native Gaussian-integral and direct Fock-build engine.

Host code is Python; compute runs in hand-written gfx950 HIP kernels of
libdft_hip.so, reached through a ctypes C ABI (include/dft_hip.h).
"""
__version__ = "0.1.0"
