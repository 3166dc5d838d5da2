"""libuipc_amd: MI355X-native per-Newton linear solve of libuipc.

Native code:
  libmi_linsys.so  -- HIP/gfx950 kernels and the C ABI of include/mi_linsys.h
  libmi_synth.so   -- seeded synthetic systems in libuipc's triplet format

The Python modules are ctypes plumbing for tests and benchmarks.
"""
from . import abi  # noqa: F401

__all__ = ["abi"]
