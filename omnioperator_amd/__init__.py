"""MI355X-native hash aggregation and inner hash join behind OmniOperator's operator surface.

Native parts (built in-tree by `python -m omnioperator_amd.build`):
  lib/libomni_gpu.so  HIP kernels for gfx950 + the C ABI of include/omni_gpu.h + the C++ GPU operators
Python parts: `capi` (ctypes face of the C ABI) and `workloads` (seeded benchmark inputs).
"""
__all__ = ["capi", "workloads"]
