"""GPUropolis for AMD MI355X: Metropolis-Hastings over expression programs on gfx950.

The sampler is libgpuropolis.so (HIP kernels behind the C ABI in include/gpuropolis.h)
and the drop-in CLI bin/gpuropolis.  This package exposes the ctypes binding
(`gpuropolis_amd.lib`), the build script (`gpuropolis_amd.build`) and the workload
generators (`gpuropolis_amd.workloads`).
"""
from . import lib  # noqa: F401

__all__ = ["lib"]
