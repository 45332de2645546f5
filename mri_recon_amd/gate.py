"""Gated frame selection on the HIP library.

``gate_scale_times`` is GATING::init_time_resolved's scaling of per-readout
gate times to frame coordinates (gating.cpp:880-1078, host); ``pregate`` is
RECON::pregate_data (recon_lib.cpp:675-776) with GATING::weight_data
(gating.cpp:1082-1200) on the device.  Its per-frame outputs feed
``GridFFT.set_traj`` / ``set_kdata`` frame by frame.
"""
import ctypes

import numpy as np
import torch

from ._ffi import GateParams, check, lib
from .gridfft import _need, _ptr, _stream

GATE_TYPES = {"none": 0, "retro_ecg": 1, "ecg": 2, "resp": 3, "time": 4, "prep": 5}
VS_TYPES = {"tornado": 0, "none": 1, "hist": 2}
TORNADO = {"flat": 0, "radial": 1, "vipr": 2}


def gate_params(frames, gate="time", vs="none", shape="vipr", wdth_low=1, wdth_high=4, kmax=128.0,
                min_quantile=0.0, max_quantile=1.0, ecg_outlier_reject=True, reset_dens=False, args=None):
    p = GateParams()
    lib().mrn_gate_params_default(ctypes.byref(p))
    p.gate_type, p.vs_type, p.frames, p.tornado_shape = GATE_TYPES[gate], VS_TYPES[vs], int(frames), TORNADO[shape]
    p.wdth_low, p.wdth_high, p.kmax = int(wdth_low), int(wdth_high), float(kmax)
    p.min_quantile, p.max_quantile = float(min_quantile), float(max_quantile)
    p.ecg_outlier_reject, p.reset_dens = int(bool(ecg_outlier_reject)), int(bool(reset_dens))
    if args:
        arr = (ctypes.c_char_p * len(args))(*[a.encode() for a in args])
        check(lib().mrn_gate_parse_args(ctypes.byref(p), len(args), arr))
    return p


def gate_scale_times(params, times):
    """(scaled times as float64 numpy, info = [offset, scale, temporal resolution])."""
    t = np.array(times, np.float64)
    info = np.zeros(3)
    check(lib().mrn_gate_scale_times(ctypes.byref(params), t.size, t.ctypes.data, info.ctypes.data))
    return t, info


def pregate(params, gt, kx, ky, kz, kw, kdata, resp=None):
    """Frame selection on the device.  gt (R,) float64, resp (R,) float32 or
    None, kx..kw (R, xres) float32, kdata (C, R, xres) complex64, all on the
    device.  Returns (counts, frames) with frames[t] = (kx, ky, kz, kw, kdata
    (C, n_t)) views into one frame-major allocation."""
    R, xres = kx.shape
    C = kdata.shape[0]
    _need(gt, torch.float64, (R,), "gt")
    for t, name in ((kx, "kx"), (ky, "ky"), (kz, "kz"), (kw, "kw")):
        _need(t, torch.float32, (R, xres), name)
    _need(kdata, torch.complex64, (C, R, xres), "kdata")
    if resp is not None:
        _need(resp, torch.float32, (R,), "resp")
    T = params.frames
    counts = (ctypes.c_int64 * T)()
    args = [ctypes.byref(params), xres, R, _ptr(gt), _ptr(resp), _ptr(kx), _ptr(ky), _ptr(kz), _ptr(kw), C,
            _ptr(kdata), counts]
    check(lib().mrn_pregate(*(args + [None] * 5 + [_stream()])))
    n = [int(c) for c in counts]
    tot = sum(n)
    dev = kx.device
    out = [torch.empty(max(tot, 1), dtype=torch.float32, device=dev) for _ in range(4)]
    okd = torch.empty(max(tot * C, 1), dtype=torch.complex64, device=dev)
    check(lib().mrn_pregate(*(args + [_ptr(o) for o in out] + [_ptr(okd), _stream()])))
    frames, base = [], 0
    for t in range(T):
        sl = slice(base, base + n[t])
        frames.append(tuple(o[sl] for o in out) + (okd[C * base:C * (base + n[t])].view(C, n[t]),))
        base += n[t]
    return n, frames
