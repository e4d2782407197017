"""pic3d_amd: MI355X-native hot loop of the PIC3D particle-in-cell code.

The compute path is libpic3d_amd.so (hand-written HIP kernels for gfx950
behind a C ABI, include/pic3d_amd.h); this package is the thin Python host
mirror used by tests, the benchmark and the differentiable form.
"""
from ._abi import (FIELD_NAMES, PIC_BX, PIC_BY, PIC_BZ, PIC_EX, PIC_EY, PIC_EZ, PIC_FP32,
                   PIC_FP64, PIC_JX, PIC_JY, PIC_JZ, PIC_RHO0, PIC_RHO1, PicError, PicLoader,
                   PicParams, lib)
from .simulation import Simulation, field_shape, setup
from . import workloads

__all__ = ["Simulation", "PicParams", "PicLoader", "PicError", "lib", "workloads", "setup",
           "field_shape", "FIELD_NAMES", "PIC_FP32", "PIC_FP64"]
