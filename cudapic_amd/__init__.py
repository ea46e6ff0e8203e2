"""cudapic_amd -- MI355X (gfx950) PIC/MCC engine for the 1d3v argon RF
discharge of sekulas/cudaPIC, behind the C ABI of include/cudapic_amd.h.

Python is a thin host here (tests, benchmarks, scripting); the engine is
the HIP library ``libcudapic_amd.so`` and its C++ host (``cudapic`` CLI).

    from cudapic_amd import kernels, Sim, DeviceBuffer, params
"""
from ._abi import (PIC_ELECTRONS, PIC_IONS, PIC_PATH_FUSED, PIC_PATH_REFERENCE, PIC_PREC_COMP,
                   PIC_PREC_FP32, PIC_PREC_FP64, PicError, lib)
from .device import DeviceBuffer, device_info, synchronize
from .params import params
from .sim import Sim
from .ensemble import Ensemble
from .adjoint import Adjoint
from .step import StepState
from . import kernels

__all__ = [
    "DeviceBuffer", "Sim", "StepState", "Ensemble", "Adjoint", "kernels", "params", "lib", "synchronize", "device_info", "PicError",
    "PIC_ELECTRONS", "PIC_IONS", "PIC_PATH_FUSED", "PIC_PATH_REFERENCE",
    "PIC_PREC_FP32", "PIC_PREC_COMP", "PIC_PREC_FP64",
]
