"""mri_recon_amd: MI355X-native NUFFT/SENSE encoding operator.

The compute path is libmrn.so (hand-written HIP kernels for gfx950, C ABI in
include/mrn.h).  ``GridFFT`` mirrors the reference's gridFFT call surface for
torch tensors; ``SenseEncoding`` (autograd) differentiates it with respect to
the image and the trajectory.
"""
from ._ffi import MrnError, lib  # noqa: F401
from .gridfft import GridFFT  # noqa: F401
from .autograd import SenseEncoding  # noqa: F401
from .dcf import dcf_calc, dcf_params  # noqa: F401
from .llr import llr_params, llr_threshold  # noqa: F401
from .gate import gate_params, gate_scale_times, pregate  # noqa: F401

__all__ = ["GridFFT", "MrnError", "SenseEncoding", "dcf_calc", "dcf_params", "llr_params", "llr_threshold", "gate_params", "gate_scale_times",
           "pregate", "lib"]
