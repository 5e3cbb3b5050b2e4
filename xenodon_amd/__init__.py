"""xenodon_amd: MI355X-native volume ray traversal (Xenodon's per-pixel path).

Native pieces live in libxenodon_amd.so (HIP kernels for gfx950 behind the
C-ABI of include/xenodon.h); `abi` is its ctypes binding.
"""
__all__ = ["abi"]
