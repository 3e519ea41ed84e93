"""hoe: MI355X (gfx950) Explicit higher-order MRI encoding operator.

Host-side mirror of HighOrderMRI.jl's Explicit operator API
(HighOrderExplicitOp / normalOperator / explicit_gpu_normal / recon_HOOp) and
its delay estimation (FindDelay / FindDelay_multishot / InterpTrajTime), field-model
inputs (apply_girf / grad2traj / traj2grad) and the 7 T demo data's field-camera
inputs (load_demo_mat / prepare_demo / samplingDensity) over the C ABI of libhoe
(include/hoe.h).
"""
from ._lib import (ArgumentError, AssertionFailure, DimensionMismatch, HIPError, LibraryNotBuilt,
                   NoDeviceError, LIB_PATH, load)
from .grid import Grid, basisfunc_spha
from .operator import (CGNR, AdjointOp, ComposedOp, HighOrderExplicitOp, L2Regularization, NormalOp,
                       WeightingOp, compose, explicit_gpu_normal, normalOperator, recon_HOOp)
from .girf import GIRFModel, apply_girf, grad2traj, traj2grad, unpack_girf_channels
from .sync import (AkimaMonotonicInterpolation, FindDelay, FindDelay_multishot, InterpTrajTime,
                   LinearMonotonicInterpolation, kspha_derivative)
from .fieldcam import demodulate, load_demo_mat, prepare_demo, samplingDensity

__all__ = [
    "ArgumentError", "AssertionFailure", "DimensionMismatch", "HIPError", "LibraryNotBuilt",
    "NoDeviceError", "LIB_PATH", "load", "Grid", "basisfunc_spha", "CGNR", "AdjointOp", "ComposedOp",
    "HighOrderExplicitOp", "L2Regularization", "NormalOp", "WeightingOp", "compose",
    "explicit_gpu_normal", "normalOperator", "recon_HOOp", "AkimaMonotonicInterpolation", "FindDelay",
    "FindDelay_multishot", "InterpTrajTime", "LinearMonotonicInterpolation", "kspha_derivative",
    "GIRFModel", "apply_girf", "grad2traj", "traj2grad", "unpack_girf_channels",
    "demodulate", "load_demo_mat", "prepare_demo", "samplingDensity",
]


def device_count() -> int:
    """Number of visible GPUs according to libhoe (0 without a GPU)."""
    return int(load().hoe_device_count())
