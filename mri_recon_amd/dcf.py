"""Pipe density compensation on the HIP library.

``dcf_calc`` is the body of RECON::dcf_calc for one encoding
(recon_lib.cpp:879-913): iterative Pipe weights from DCFgridFFT's
polynomial-kernel gridding (DCFgridFFT.cpp:56-440), computed on the device in
one call.  The reference stores the result in data.kw(e); here it is returned
(or written into ``out``).  Settings and their defaults are the reference's
-dcf_* flags (recon_lib.cpp:124-130, 415-421).
"""
import ctypes

import torch

from ._ffi import DcfParams, check, lib
from .gridfft import _need, _ptr, _stream


def dcf_params(N, iters=20, dwin=2.1, scale=1.0, overgrid=2.1, acc=1.0, kernel="poly", args=None):
    """mrn_dcf_params for an image of size N (int or (Nx, Ny, Nz)); ``args``
    (a list of strings) is parsed like the reference's command line."""
    p = DcfParams()
    lib().mrn_dcf_params_default(ctypes.byref(p))
    n = (N, N, N) if isinstance(N, int) else tuple(N)
    for a in range(3):
        p.N[a] = int(n[a])
    p.iters, p.dwin, p.scale, p.overgrid, p.acc = int(iters), float(dwin), float(scale), float(overgrid), float(acc)
    p.kernel = {"poly": 0, "poly2": 1}[kernel]
    if args:
        arr = (ctypes.c_char_p * len(args))(*[a.encode() for a in args])
        check(lib().mrn_dcf_parse_args(ctypes.byref(p), len(args), arr))
    return p


def dcf_grid_size(p):
    """(Sx, Sy, Sz) of the DCF grid: trunc(os * (Nz, Ny, Nx) + 16)."""
    S = (ctypes.c_int32 * 3)()
    check(lib().mrn_dcf_grid_size(ctypes.byref(p), S))
    return tuple(S)


def dcf_calc(kx, ky, kz, N=None, params=None, out=None, return_sums=False, **kw):
    """Pipe DCF weights of one encoding's samples (float32 (M,) tensors on the
    device, the trajectory's units).  Returns kw (M,) float32 on the device,
    and with ``return_sums`` the per-iteration sum of |kw| the reference
    prints.  Keyword arguments go to ``dcf_params``."""
    if params is None:
        if N is None:
            raise ValueError("give N or params")
        params = dcf_params(N, **kw)
    M = kx.numel()
    for t, name in ((kx, "kx"), (ky, "ky"), (kz, "kz")):
        _need(t, torch.float32, (M,), name)
    if out is None:
        out = torch.empty(M, dtype=torch.float32, device=kx.device)
    _need(out, torch.float32, (M,), "out")
    sums = (ctypes.c_double * max(params.iters, 1))()
    check(lib().mrn_dcf_calc(ctypes.byref(params), M, _ptr(kx), _ptr(ky), _ptr(kz), _ptr(out),
                             sums if return_sums else None, _stream()))
    if return_sums:
        return out, [sums[i] for i in range(params.iters)]
    return out


def dcf_last_atomics():
    """Float atomics the grid scatters of the last dcf_calc issued (all iterations)."""
    n = ctypes.c_int64()
    check(lib().mrn_dcf_last_atomics(ctypes.byref(n)))
    return n.value
