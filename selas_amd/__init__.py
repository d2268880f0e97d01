"""selas_amd: an MI355X-native (gfx950) implementation of Selas's rendering hot path.

The compute path is libselas_gpu.so (hand-written HIP kernels behind the C ABI in
include/selas_gpu.h). `selas_amd.native` is its ctypes binding; `selas_amd.workloads` defines
the synthetic workloads.
"""
__all__ = ["native", "workloads"]
