"""MI355X spiking-network engine: HIP kernels behind the C ABI of include/bghip.h.

Python modules here are plumbing around libbghip.so (braingrid_amd/lib/): ctypes bindings
(`bghip`), host-side parameter tables (`model`), the batched driver (`batch`) and the serving
loop (`serve`). All simulation arithmetic runs in the HIP kernels.
"""
__all__ = ["bghip", "model"]
