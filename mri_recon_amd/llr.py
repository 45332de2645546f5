"""Locally low-rank threshold over frames on the HIP library.

``llr_threshold`` is one LOWRANKCOIL::update_threshold + ::thresh call
(clear.cpp:233-556) on an image of Nt x Ne frames, as recon_lib applies it
after each gradient step (recon_lib.cpp:2117-2126); ``GridFFT.set_fista_llr``
puts the time variant inside the library's FISTA iteration.  Settings are the
reference's -clear_* flags (clear.cpp:40-96).
"""
import ctypes

import torch

from ._ffi import LlrParams, check, lib
from .gridfft import _need, _ptr, _stream


def llr_params(alpha_time=0.0, alpha_encode=0.0, block=(4, 4, 4), block_iter=4, amp=4.0, t2=6.0, seed=0,
               args=None):
    """mrn_llr_params; ``args`` (list of strings) is parsed like the reference's
    command line after the keyword values are set."""
    p = LlrParams()
    lib().mrn_llr_params_default(ctypes.byref(p))
    for a in range(3):
        p.block[a] = int(block[a])
    p.block_iter, p.alpha_time, p.alpha_encode = int(block_iter), float(alpha_time), float(alpha_encode)
    p.amp, p.t2, p.seed = float(amp), float(t2), int(seed)
    if args:
        arr = (ctypes.c_char_p * len(args))(*[a.encode() for a in args])
        check(lib().mrn_llr_parse_args(ctypes.byref(p), len(args), arr))
    return p


def llr_threshold(X, dim=0, iteration=0, call=0, params=None, **kw):
    """In place on X (Nt, Ne, Nz, Ny, Nx) complex64 on the device; returns smax.
    ``call`` numbers the calls for the block-shift stream."""
    if params is None:
        params = llr_params(**kw)
    if X.dim() != 5:
        raise ValueError("X must be (Nt, Ne, Nz, Ny, Nx)")
    _need(X, torch.complex64, None, "X")
    Nt, Ne, Nz, Ny, Nx = X.shape
    N = (ctypes.c_int32 * 3)(Nx, Ny, Nz)
    smax = ctypes.c_float()
    check(lib().mrn_llr(ctypes.byref(params), _ptr(X), Nt, Ne, N, int(dim), int(iteration), int(call),
                        ctypes.byref(smax), _stream()))
    return smax.value
