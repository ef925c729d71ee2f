"""MI355X-native spinsim propagator: HIP kernels behind a C ABI (libspx.so).

`spinsim_amd.spx` is the ctypes view of that ABI; the C++ drop-in for the
reference's spinsim classes lives in spinsim_amd/host.
"""
__all__ = ["spx"]
