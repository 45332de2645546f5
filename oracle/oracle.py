"""ctypes front end of the CPU oracle (test infrastructure only).

The oracle is the C restatement of the reference's gridFFT operator and its
CG/FISTA callers in ``gridfft_oracle.c`` (built by ``oracle/Makefile`` into
``liboracle_f32.so`` and ``liboracle_f64.so``).  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import this
module, as the checker; the product path (``mri_recon_amd``) never does.

Array conventions follow the reference (Blitz ColumnMajorArray, x fastest):
images are numpy arrays of shape (Nz, Ny, Nx); coil maps (C, Nz, Ny, Nx);
k-space (C, M); the trajectory kx/ky/kz/kw are float32 (M,) in the
reference's units (cycles per FOV, |k| <= N/2; src/mri_data.h:50-57).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
KERNELS = {"triangle": 0, "kaiser": 1, "sinc": 2, "poly": 3}


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("N", ctypes.c_int32 * 3),
        ("overgrid", ctypes.c_float),
        ("grid", ctypes.c_float * 3),
        ("dwin", ctypes.c_float * 3),
        ("kernel", ctypes.c_int32),
        ("grid_in", ctypes.c_int32 * 3),
        ("fft_in", ctypes.c_int32 * 3),
        ("grid_scale", ctypes.c_float * 3),
        ("k_rad", ctypes.c_float),
    ]


class DcfParams(ctypes.Structure):
    """Pipe DCF settings (recon_lib.cpp:124-130 defaults)."""
    _fields_ = [
        ("N", ctypes.c_int32 * 3),
        ("iters", ctypes.c_int32),
        ("dwin", ctypes.c_float),
        ("scale", ctypes.c_float),
        ("overgrid", ctypes.c_float),
        ("acc", ctypes.c_float),
        ("kernel", ctypes.c_int32),
    ]


def dcf_params(N, iters=20, dwin=2.1, scale=1.0, overgrid=2.1, acc=1.0, kernel=0):
    p = DcfParams()
    n = (N, N, N) if np.isscalar(N) else tuple(N)
    for a in range(3):
        p.N[a] = int(n[a])
    p.iters, p.dwin, p.scale, p.overgrid, p.acc, p.kernel = int(iters), dwin, scale, overgrid, acc, int(kernel)
    return p


class LlrParams(ctypes.Structure):
    """Locally low-rank settings (clear.cpp:40-96 defaults)."""
    _fields_ = [
        ("block", ctypes.c_int32 * 3),
        ("block_iter", ctypes.c_int32),
        ("alpha_time", ctypes.c_float),
        ("alpha_encode", ctypes.c_float),
        ("amp", ctypes.c_float),
        ("t2", ctypes.c_float),
        ("seed", ctypes.c_uint64),
    ]


def llr_params(alpha_time=0.0, alpha_encode=0.0, block=(4, 4, 4), block_iter=4, amp=4.0, t2=6.0, seed=0):
    p = LlrParams()
    for a in range(3):
        p.block[a] = int(block[a])
    p.block_iter, p.alpha_time, p.alpha_encode, p.amp, p.t2 = int(block_iter), alpha_time, alpha_encode, amp, t2
    p.seed = int(seed)
    return p


class GateParams(ctypes.Structure):
    """Gated frame selection settings (gating.cpp:84-112 defaults)."""
    _fields_ = [
        ("gate_type", ctypes.c_int32),
        ("vs_type", ctypes.c_int32),
        ("frames", ctypes.c_int32),
        ("tornado_shape", ctypes.c_int32),
        ("wdth_low", ctypes.c_int32),
        ("wdth_high", ctypes.c_int32),
        ("kmax", ctypes.c_float),
        ("min_quantile", ctypes.c_float),
        ("max_quantile", ctypes.c_float),
        ("ecg_outlier_reject", ctypes.c_int32),
        ("reset_dens", ctypes.c_int32),
    ]


GATE_TYPES = {"none": 0, "retro_ecg": 1, "ecg": 2, "resp": 3, "time": 4, "prep": 5}
VS_TYPES = {"tornado": 0, "none": 1, "hist": 2}
TORNADO = {"flat": 0, "radial": 1, "vipr": 2}


def gate_params(frames, gate="time", vs="none", shape="vipr", wdth_low=1, wdth_high=4, kmax=128.0,
                min_quantile=0.0, max_quantile=1.0, ecg_outlier_reject=True, reset_dens=False):
    p = GateParams()
    p.gate_type, p.vs_type, p.frames, p.tornado_shape = GATE_TYPES[gate], VS_TYPES[vs], int(frames), TORNADO[shape]
    p.wdth_low, p.wdth_high, p.kmax = int(wdth_low), int(wdth_high), float(kmax)
    p.min_quantile, p.max_quantile = float(min_quantile), float(max_quantile)
    p.ecg_outlier_reject, p.reset_dens = int(bool(ecg_outlier_reject)), int(bool(reset_dens))
    return p


_LIBS = {}


def build():
    """Compile both oracle libraries (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _lib(precision):
    if precision not in _LIBS:
        path = os.path.join(_HERE, "liboracle_%s.so" % precision)
        if not os.path.exists(path):
            build()
        lib = ctypes.CDLL(path)
        vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
        lib.orc_create.restype = vp
        lib.orc_create.argtypes = [ctypes.POINTER(OrcParams)]
        lib.orc_destroy.argtypes = [vp]
        lib.orc_geometry.argtypes = [vp] + [vp] * 6
        lib.orc_table.argtypes = [vp, ctypes.c_int, vp]
        lib.orc_table.restype = ctypes.c_int
        lib.orc_window.argtypes = [vp, ctypes.c_int, vp]
        lib.orc_window.restype = ctypes.c_int
        lib.orc_bessi0_eval.argtypes = [dbl]
        lib.orc_bessi0_eval.restype = dbl
        lib.orc_fft1d.argtypes = [ctypes.c_int, ctypes.c_int, vp, ctypes.c_int]
        lib.orc_backward.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, i64, vp]
        lib.orc_forward.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, i64]
        lib.orc_forward_sos.argtypes = [vp, vp, vp, vp, vp, vp, vp, i64]
        lib.orc_wavelet_levels.argtypes = [vp, vp, vp]
        lib.orc_wavelet3d.argtypes = [vp, vp, vp, vp, ctypes.c_int]
        lib.orc_fraction_threshold.argtypes = [vp, i64, dbl]
        lib.orc_fraction_threshold.restype = dbl
        lib.orc_soft_threshold.argtypes = [vp, i64, dbl]
        lib.orc_l1_wavelet.argtypes = [vp, ctypes.c_int, vp, vp, dbl, ctypes.c_int, vp]
        lib.orc_l1_wavelet.restype = dbl
        lib.orc_spin_shifts.argtypes = [ctypes.c_uint64, i64, vp]
        lib.orc_fista_l1.argtypes = [vp, vp, vp, ctypes.c_int, vp, vp, vp, vp, vp, i64, ctypes.c_int,
                                     dbl, dbl, vp, ctypes.c_int, vp, vp, vp]
        lib.orc_stage_pre_fft.argtypes = [vp, vp, vp, vp]
        lib.orc_stage_gather.argtypes = [vp, vp, vp, vp, vp, vp, vp, i64, vp]
        lib.orc_stage_scatter.argtypes = [vp, vp, vp, vp, vp, vp, vp, i64]
        lib.orc_stage_ifft_post.argtypes = [vp, vp, vp, vp]
        lib.orc_fft3.argtypes = [vp, vp, ctypes.c_int]
        lib.orc_image_to_kspace.argtypes = [vp, vp, vp, ctypes.c_int, vp, vp, vp, vp, vp, i64, vp]
        lib.orc_kspace_to_image.argtypes = [vp, vp, vp, ctypes.c_int, vp, vp, vp, vp, vp, i64]
        lib.orc_normal.argtypes = [vp, vp, vp, ctypes.c_int, vp, vp, vp, vp, vp, i64]
        lib.orc_cg.argtypes = [vp, vp, vp, ctypes.c_int, vp, vp, vp, vp, vp, i64, ctypes.c_int, vp]
        lib.orc_power_method.argtypes = [vp, vp, ctypes.c_int, vp, vp, vp, vp, i64, ctypes.c_int, vp]
        lib.orc_power_method.restype = dbl
        lib.orc_fista.argtypes = [vp, vp, vp, ctypes.c_int, vp, vp, vp, vp, vp, i64, ctypes.c_int,
                                  dbl, vp]
        lib.orc_fista_cauchy.argtypes = [vp, vp, vp, ctypes.c_int, vp, vp, vp, vp, vp, i64, ctypes.c_int,
                                         dbl, ctypes.c_int, vp, vp]
        lib.orc_threads.restype = ctypes.c_int
        lib.orc_use_mkl_fft.argtypes = [ctypes.c_int]
        lib.orc_use_mkl_fft.restype = ctypes.c_int
        lib.orc_dcf_table_len.argtypes = [ctypes.c_float]
        lib.orc_dcf_table_len.restype = ctypes.c_int
        lib.orc_dcf_table.argtypes = [ctypes.c_float, ctypes.c_int, vp]
        lib.orc_dcf_table.restype = ctypes.c_int
        lib.orc_dcf_grid_size.argtypes = [ctypes.POINTER(DcfParams), vp]
        lib.orc_dcf_forward.argtypes = [ctypes.POINTER(DcfParams), vp, vp, vp, i64, vp, vp, vp]
        lib.orc_dcf_backward.argtypes = [ctypes.POINTER(DcfParams), vp, vp, vp, i64, vp, vp, vp]
        lib.orc_dcf_calc.argtypes = [ctypes.POINTER(DcfParams), i64, vp, vp, vp, vp, vp]
        lib.orc_dcf_calc.restype = ctypes.c_int
        LP = ctypes.POINTER(LlrParams)
        lib.orc_llr_shift.argtypes = [LP, vp, i64, ctypes.c_int, vp]
        lib.orc_svd_values.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
        lib.orc_llr_block_smax.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp, ctypes.c_int, vp]
        lib.orc_llr_block_smax.restype = i64
        lib.orc_median_nonzero.argtypes = [vp, i64]
        lib.orc_median_nonzero.restype = dbl
        lib.orc_llr_update_threshold.argtypes = [LP, vp, ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int]
        lib.orc_llr_update_threshold.restype = ctypes.c_float
        lib.orc_llr_pass.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp, ctypes.c_int, vp, dbl]
        lib.orc_llr.argtypes = [LP, vp, ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int, i64]
        lib.orc_llr.restype = ctypes.c_float
        GP = ctypes.POINTER(GateParams)
        lib.orc_quantile5.argtypes = [vp, i64, dbl]
        lib.orc_quantile5.restype = dbl
        lib.orc_gate_scale.argtypes = [GP, i64, vp, vp]
        lib.orc_gate_scale.restype = ctypes.c_int
        lib.orc_pregate.argtypes = [GP, i64, i64, vp, vp, vp, vp, vp, vp, ctypes.c_int, vp, vp, vp, vp, vp, vp, vp]
        _LIBS[precision] = lib
    return _LIBS[precision]


def use_mkl_fft(on=True, precision="f32"):
    """Route the oracle's 3-D FFT line transforms through MKL's FFTW3 interface
    (the CPU timing baseline only); returns whether MKL was found."""
    return bool(_lib(precision).orc_use_mkl_fft(int(bool(on))))


def threads():
    return _lib("f32").orc_threads()


def _p(a):
    return None if a is None else a.ctypes.data


def bessi0(x, precision="f64"):
    return _lib(precision).orc_bessi0_eval(float(x))


def fft1d(x, sign=-1, precision="f64"):
    """Unitary 1-D DFT along the last axis (sign -1 forward, +1 backward)."""
    ct = np.complex64 if precision == "f32" else np.complex128
    a = np.ascontiguousarray(x, dtype=ct).copy()
    n = a.shape[-1]
    _lib(precision).orc_fft1d(n, sign, a.ctypes.data, a.size // n)
    return a


def _i3(v):
    return np.ascontiguousarray(v, dtype=np.int32)


def wavelet_levels(shape, levels=(4, 4, 4)):
    """Levels per axis (x, y, z) the transform uses for an image of `shape`
    (Nz, Ny, Nx)."""
    N = _i3([shape[2], shape[1], shape[0]])
    L = np.zeros(3, np.int32)
    req = _i3(levels)  # arrays held in locals while C reads them
    _lib("f64").orc_wavelet_levels(N.ctypes.data, req.ctypes.data, L.ctypes.data)
    return tuple(int(v) for v in L)


def wavelet3d(X, levels=(4, 4, 4), shift=(0, 0, 0), inverse=False, precision="f64"):
    """WAVELET3D forward / backward of one image (Nz, Ny, Nx); returns a copy."""
    ct = np.complex64 if precision == "f32" else np.complex128
    Y = np.array(X, dtype=ct, order="C", copy=True)
    N = _i3([Y.shape[2], Y.shape[1], Y.shape[0]])
    lv, sh = _i3(levels), _i3(shift)
    _lib(precision).orc_wavelet3d(Y.ctypes.data, N.ctypes.data, lv.ctypes.data, sh.ctypes.data,
                                  int(bool(inverse)))
    return Y


def fraction_threshold(C, fraction, precision="f64"):
    ct = np.complex64 if precision == "f32" else np.complex128
    C = np.ascontiguousarray(C, dtype=ct)
    return float(_lib(precision).orc_fraction_threshold(C.ctypes.data, C.size, float(fraction)))


def soft_threshold(C, value, precision="f64"):
    ct = np.complex64 if precision == "f32" else np.complex128
    Y = np.array(C, dtype=ct, order="C", copy=True)
    _lib(precision).orc_soft_threshold(Y.ctypes.data, Y.size, float(value))
    return Y


def l1_wavelet(X, fraction, levels=(4, 4, 4), shifts=((0, 0, 0),), precision="f64"):
    """L1_threshold of frames X (F, Nz, Ny, Nx) with the given per-spin shifts;
    returns (X', threshold)."""
    ct = np.complex64 if precision == "f32" else np.complex128
    Y = np.array(X, dtype=ct, order="C", copy=True)
    if Y.ndim == 3:
        Y = Y[None]
    N = _i3([Y.shape[3], Y.shape[2], Y.shape[1]])
    sh = _i3(np.asarray(shifts).reshape(-1, 3))
    lv = _i3(levels)
    t = _lib(precision).orc_l1_wavelet(Y.ctypes.data, Y.shape[0], N.ctypes.data, lv.ctypes.data,
                                       float(fraction), sh.shape[0], sh.ctypes.data)
    return Y, float(t)


def spin_shifts(seed, count):
    out = np.zeros(count, np.int32)
    _lib("f64").orc_spin_shifts(int(seed), int(count), out.ctypes.data)
    return out


def _rtype(precision):
    return np.float32 if precision == "f32" else np.float64


def dcf_table(dwin=2.1, kernel=0, precision="f32"):
    """DCFgridFFT's normalised polynomial table (DCFgridFFT.cpp:56-165)."""
    lib = _lib(precision)
    tab = np.zeros(lib.orc_dcf_table_len(dwin), _rtype(precision))
    lib.orc_dcf_table(dwin, int(kernel), tab.ctypes.data)
    return tab


def dcf_grid_size(p):
    """(Sx, Sy, Sz) of dcf_calc's grid (recon_lib.cpp:873-875)."""
    S = np.zeros(3, np.int32)
    _lib("f64").orc_dcf_grid_size(ctypes.byref(p), S.ctypes.data)
    return tuple(int(v) for v in S)


def _traj3(kx, ky, kz):
    return [np.ascontiguousarray(np.asarray(k, np.float32).ravel()) for k in (kx, ky, kz)]


def dcf_forward(p, kw, kx, ky, kz, precision="f32"):
    """DCFgridFFT::forward onto a zero grid: returns X of shape (Sz, Sy, Sx)."""
    rt = _rtype(precision)
    tab = dcf_table(p.dwin, p.kernel, precision)
    Sx, Sy, Sz = dcf_grid_size(p)
    X = np.zeros((Sz, Sy, Sx), rt)
    kw = np.ascontiguousarray(kw, rt)
    t = _traj3(kx, ky, kz)
    _lib(precision).orc_dcf_forward(ctypes.byref(p), tab.ctypes.data, X.ctypes.data, kw.ctypes.data, kw.size,
                                    *[a.ctypes.data for a in t])
    return X


def dcf_backward(p, X, kx, ky, kz, precision="f32"):
    """DCFgridFFT::backward: the grid X interpolated at every sample."""
    rt = _rtype(precision)
    tab = dcf_table(p.dwin, p.kernel, precision)
    X = np.ascontiguousarray(X, rt)
    t = _traj3(kx, ky, kz)
    out = np.zeros(t[0].size, rt)
    _lib(precision).orc_dcf_backward(ctypes.byref(p), tab.ctypes.data, X.ctypes.data, out.ctypes.data,
                                     out.size, *[a.ctypes.data for a in t])
    return out


def dcf_calc(p, kx, ky, kz, precision="f32"):
    """RECON::dcf_calc for one encoding: returns (kw, per-iteration sums)."""
    t = _traj3(kx, ky, kz)
    kw = np.zeros(t[0].size, _rtype(precision))
    sums = np.zeros(max(p.iters, 1), np.float64)
    rc = _lib(precision).orc_dcf_calc(ctypes.byref(p), kw.size, *[a.ctypes.data for a in t], kw.ctypes.data,
                                      sums.ctypes.data)
    if rc != 0:
        raise MemoryError("oracle dcf_calc: allocation failed")
    return kw, sums[: p.iters]


def _img5(X, precision):
    ct = np.complex64 if precision == "f32" else np.complex128
    X = np.ascontiguousarray(X, ct)
    if X.ndim != 5:
        raise ValueError("X must be (Nt, Ne, Nz, Ny, Nx)")
    Nt, Ne, Nz, Ny, Nx = X.shape
    return X, Nt, Ne, np.array([Nx, Ny, Nz], np.int32)


def svd_values(A):
    """Singular values (descending) of a complex matrix by the oracle's
    one-sided Jacobi SVD."""
    A = np.asarray(A, np.complex128)
    Np, N = A.shape
    buf = np.ascontiguousarray(A.T).view(np.float64)  # column-major, re/im interleaved
    s = np.zeros(N)
    _lib("f64").orc_svd_values(buf.ctypes.data, Np, N, s.ctypes.data)
    return s


def median_nonzero(v):
    v = np.ascontiguousarray(v, np.float64)
    return _lib("f64").orc_median_nonzero(v.ctypes.data, v.size)


def llr_shift(p, bs, call, pass_):
    b = np.ascontiguousarray(bs, np.int32)
    out = np.zeros(3, np.int32)
    _lib("f64").orc_llr_shift(ctypes.byref(p), b.ctypes.data, int(call), int(pass_), out.ctypes.data)
    return out


def llr_block_smax(X, dim, block=(4, 4, 4), precision="f64"):
    X, Nt, Ne, Nv = _img5(X, precision)
    bs = np.minimum(np.array(block, np.int32), Nv)
    nb = int(np.prod(Nv // bs))
    out = np.zeros(nb)
    b = np.array(block, np.int32)
    _lib(precision).orc_llr_block_smax(X.ctypes.data, Nt, Ne, Nv.ctypes.data, b.ctypes.data, int(dim),
                                       out.ctypes.data)
    return out


def llr_update_threshold(p, X, dim, it, precision="f64"):
    X, Nt, Ne, Nv = _img5(X, precision)
    return _lib(precision).orc_llr_update_threshold(ctypes.byref(p), X.ctypes.data, Nt, Ne, Nv.ctypes.data,
                                                    int(dim), int(it))


def llr_pass(X, dim, shift, tau, block=(4, 4, 4), precision="f64"):
    """One thresh pass (given shift and tau) on a copy of X."""
    X, Nt, Ne, Nv = _img5(np.array(X), precision)
    b = np.array(block, np.int32)
    sh = np.array(shift, np.int32)
    _lib(precision).orc_llr_pass(X.ctypes.data, Nt, Ne, Nv.ctypes.data, b.ctypes.data, int(dim), sh.ctypes.data,
                                 float(tau))
    return X


def llr(p, X, dim, it, call, precision="f64"):
    """update_threshold + thresh on a copy of X: returns (X, smax)."""
    X, Nt, Ne, Nv = _img5(np.array(X), precision)
    smax = _lib(precision).orc_llr(ctypes.byref(p), X.ctypes.data, Nt, Ne, Nv.ctypes.data, int(dim), int(it),
                                   int(call))
    return X, float(smax)


def quantile5(v, q):
    s = np.sort(np.asarray(v, np.float64))
    return _lib("f64").orc_quantile5(s.ctypes.data, s.size, float(q))


def gate_scale(p, times):
    """init_time_resolved's scaling of per-readout gate times: (times, info)."""
    t = np.array(times, np.float64)
    info = np.zeros(3)
    if _lib("f64").orc_gate_scale(ctypes.byref(p), t.size, t.ctypes.data, info.ctypes.data) != 0:
        raise ValueError("unknown gating setting")
    return t, info


def pregate(p, gt, kx, ky, kz, kw, kdata, resp=None):
    """pregate_data for one encoding.  kx.. are (R, xres) float32, kdata
    (C, R, xres) complex64, gt (R,) scaled gate times; returns (counts, kx, ky,
    kz, kw, kdata) with every frame's samples in sample order, frame after
    frame, and kdata frame-major as [frame][C][n_frame]."""
    R, xres = np.shape(kx)
    gt = np.ascontiguousarray(gt, np.float64)
    tr = [np.ascontiguousarray(a, np.float32) for a in (kx, ky, kz, kw)]
    kd = np.ascontiguousarray(kdata, np.complex64)
    C = kd.shape[0]
    rs = None if resp is None else np.ascontiguousarray(resp, np.float32)
    counts = np.zeros(p.frames, np.int64)
    lib = _lib("f64")
    args = [ctypes.byref(p), xres, R, gt.ctypes.data, None if rs is None else rs.ctypes.data] + \
        [a.ctypes.data for a in tr] + [C, kd.ctypes.data, counts.ctypes.data]
    lib.orc_pregate(*(args + [None] * 5))
    n = int(counts.sum())
    outs = [np.zeros(max(n, 1), np.float32) for _ in range(4)]
    kdo = np.zeros(max(n * C, 1), np.complex64)
    lib.orc_pregate(*(args + [o.ctypes.data for o in outs] + [kdo.ctypes.data]))
    return (counts,) + tuple(o[:n] for o in outs) + (kdo[:n * C],)


class OracleOp:
    """One gridFFT instance (geometry, tables, windows) at one precision."""

    def __init__(self, N, overgrid=1.5, dwin=-1.0, kernel="kaiser", grid_in=(1, 1, 1),
                 fft_in=(1, 1, 1), grid=-1.0, grid_scale=1.0, k_rad=9999.0, precision="f64"):
        if np.isscalar(N):
            N = (N, N, N)
        self.precision = precision
        self.ctype = np.complex64 if precision == "f32" else np.complex128
        self.lib = _lib(precision)
        p = OrcParams()
        trip = lambda v: v if not np.isscalar(v) else (v, v, v)
        for a in range(3):
            p.N[a] = int(N[a])
            p.grid[a] = float(trip(grid)[a])
            p.dwin[a] = float(trip(dwin)[a])
            p.grid_in[a] = int(grid_in[a])
            p.fft_in[a] = int(fft_in[a])
            p.grid_scale[a] = float(trip(grid_scale)[a])
        p.overgrid = float(overgrid)
        p.kernel = KERNELS[kernel] if isinstance(kernel, str) else int(kernel)
        p.k_rad = float(k_rad)
        self.params = p
        self.h = self.lib.orc_create(ctypes.byref(p))
        self.N = tuple(int(n) for n in N)
        g = self.geometry()
        self.S = tuple(g["S"])
        self.og = tuple(g["og"])

    def __del__(self):
        try:
            self.lib.orc_destroy(self.h)
        except Exception:
            pass

    def geometry(self):
        S = np.zeros(3, np.int32)
        og = np.zeros(3, np.int32)
        L = np.zeros(3, np.int32)
        grid = np.zeros(3)
        dwin = np.zeros(3)
        beta = np.zeros(3)
        self.lib.orc_geometry(self.h, S.ctypes.data, og.ctypes.data, L.ctypes.data,
                              grid.ctypes.data, dwin.ctypes.data, beta.ctypes.data)
        return dict(S=S, og=og, L=L, grid=grid, dwin=dwin, beta=beta)

    def table(self, axis=0):
        out = np.zeros(4096)
        n = self.lib.orc_table(self.h, axis, out.ctypes.data)
        return out[:n].copy()

    def window(self, axis=0):
        out = np.zeros(8192)
        n = self.lib.orc_window(self.h, axis, out.ctypes.data)
        return out[:n].copy()

    # --- helpers -------------------------------------------------------
    def _c(self, a):
        return None if a is None else np.ascontiguousarray(a, dtype=self.ctype)

    @staticmethod
    def _traj(traj):
        return [np.ascontiguousarray(t, dtype=np.float32) for t in traj]

    def grid_shape(self):
        return (self.S[2], self.S[1], self.S[0])

    # --- per-coil operator (reference names) ---------------------------
    def backward(self, X, smap, traj, sub=None):
        """A_c X (gridFFT::backward); minus sub when given (backward_residual)."""
        kx, ky, kz, kw = self._traj(traj)
        X, smap, sub = self._c(X), self._c(smap), self._c(sub)
        out = np.zeros(kx.shape[0], self.ctype)
        self.lib.orc_backward(self.h, X.ctypes.data, _p(smap), out.ctypes.data, kx.ctypes.data,
                              ky.ctypes.data, kz.ctypes.data, kw.ctypes.data, kx.shape[0], _p(sub))
        return out

    def forward(self, X, smap, data, traj):
        """X += A_c^H (kw taper data) (gridFFT::forward); X updated in place."""
        kx, ky, kz, kw = self._traj(traj)
        assert X.dtype == self.ctype and X.flags.c_contiguous
        smap, data = self._c(smap), self._c(data)
        self.lib.orc_forward(self.h, X.ctypes.data, _p(smap), data.ctypes.data, kx.ctypes.data,
                             ky.ctypes.data, kz.ctypes.data, kw.ctypes.data, kx.shape[0])
        return X

    def forward_sos(self, X, data, traj):
        """X += |deapodised A^H (kw taper data)|^2 (gridFFT::forward_sos); X updated in place."""
        kx, ky, kz, kw = self._traj(traj)
        assert X.dtype == self.ctype and X.flags.c_contiguous
        data = self._c(data)
        self.lib.orc_forward_sos(self.h, X.ctypes.data, data.ctypes.data, kx.ctypes.data,
                                 ky.ctypes.data, kz.ctypes.data, kw.ctypes.data, kx.shape[0])
        return X

    # --- stages ---------------------------------------------------------
    def stage_pre_fft(self, X, smap):
        g = np.zeros(self.grid_shape(), self.ctype)
        X, smap = self._c(X), self._c(smap)
        self.lib.orc_stage_pre_fft(self.h, X.ctypes.data, _p(smap), g.ctypes.data)
        return g

    def stage_gather(self, grid, traj, sub=None):
        kx, ky, kz, kw = self._traj(traj)
        grid, sub = self._c(grid), self._c(sub)
        out = np.zeros(kx.shape[0], self.ctype)
        self.lib.orc_stage_gather(self.h, grid.ctypes.data, out.ctypes.data, kx.ctypes.data,
                                  ky.ctypes.data, kz.ctypes.data, kw.ctypes.data, kx.shape[0],
                                  _p(sub))
        return out

    def stage_scatter(self, data, traj):
        kx, ky, kz, kw = self._traj(traj)
        g = np.zeros(self.grid_shape(), self.ctype)
        data = self._c(data)
        self.lib.orc_stage_scatter(self.h, g.ctypes.data, data.ctypes.data, kx.ctypes.data,
                                   ky.ctypes.data, kz.ctypes.data, kw.ctypes.data, kx.shape[0])
        return g

    def stage_ifft_post(self, grid, smap, X=None):
        g = self._c(grid).copy()
        if X is None:
            X = np.zeros((self.N[2], self.N[1], self.N[0]), self.ctype)
        smap = self._c(smap)
        self.lib.orc_stage_ifft_post(self.h, g.ctypes.data, X.ctypes.data, _p(smap))
        return X

    def fft3(self, grid, sign):
        g = self._c(grid).copy()
        self.lib.orc_fft3(self.h, g.ctypes.data, int(sign))
        return g

    # --- multi-coil -----------------------------------------------------
    def image_to_kspace(self, X, smaps, traj, sub=None):
        kx, ky, kz, kw = self._traj(traj)
        X, smaps, sub = self._c(X), self._c(smaps), self._c(sub)
        C = 1 if smaps is None else smaps.shape[0]
        out = np.zeros((C, kx.shape[0]), self.ctype)
        self.lib.orc_image_to_kspace(self.h, X.ctypes.data, _p(smaps), C, out.ctypes.data,
                                     kx.ctypes.data, ky.ctypes.data, kz.ctypes.data,
                                     kw.ctypes.data, kx.shape[0], _p(sub))
        return out

    def kspace_to_image(self, data, smaps, traj, X=None):
        kx, ky, kz, kw = self._traj(traj)
        data, smaps = self._c(data), self._c(smaps)
        C = data.shape[0]
        if X is None:
            X = np.zeros((self.N[2], self.N[1], self.N[0]), self.ctype)
        self.lib.orc_kspace_to_image(self.h, X.ctypes.data, _p(smaps), C, data.ctypes.data,
                                     kx.ctypes.data, ky.ctypes.data, kz.ctypes.data,
                                     kw.ctypes.data, kx.shape[0])
        return X

    def normal(self, P, smaps, traj):
        kx, ky, kz, kw = self._traj(traj)
        P, smaps = self._c(P), self._c(smaps)
        C = 1 if smaps is None else smaps.shape[0]
        out = np.zeros_like(P)
        self.lib.orc_normal(self.h, P.ctypes.data, _p(smaps), C, out.ctypes.data, kx.ctypes.data,
                            ky.ctypes.data, kz.ctypes.data, kw.ctypes.data, kx.shape[0])
        return out

    def cg(self, smaps, data, traj, iters):
        """SENSE-CG (recon_lib.cpp:1523-1759); returns (X, trace[iters,4])."""
        kx, ky, kz, kw = self._traj(traj)
        data, smaps = self._c(data), self._c(smaps)
        X = np.zeros((self.N[2], self.N[1], self.N[0]), self.ctype)
        trace = np.zeros((iters, 4))
        self.lib.orc_cg(self.h, X.ctypes.data, _p(smaps), data.shape[0], data.ctypes.data,
                        kx.ctypes.data, ky.ctypes.data, kz.ctypes.data, kw.ctypes.data,
                        kx.shape[0], iters, trace.ctypes.data)
        return X, trace

    def power_method(self, smaps, traj, iters=30):
        kx, ky, kz, kw = self._traj(traj)
        smaps = self._c(smaps)
        C = 1 if smaps is None else smaps.shape[0]
        eig = np.zeros(iters)
        step = self.lib.orc_power_method(self.h, _p(smaps), C, kx.ctypes.data, ky.ctypes.data,
                                         kz.ctypes.data, kw.ctypes.data, kx.shape[0], iters,
                                         eig.ctypes.data)
        return step, eig

    def fista_l1(self, smaps, data, traj, iters, step, fraction, levels=(4, 4, 4), spins=4, seed=0):
        """FISTA with the L1 wavelet proximal step; cycle-spin shifts from
        spin_shifts(seed, 3 * spins * iters).  Returns (X, resid, thresholds)."""
        kx, ky, kz, kw = self._traj(traj)
        data, smaps = self._c(data), self._c(smaps)
        X = np.zeros((self.N[2], self.N[1], self.N[0]), self.ctype)
        resid = np.zeros(iters)
        thr = np.zeros(iters)
        sh = spin_shifts(seed, 3 * spins * iters)
        lv = _i3(levels)
        self.lib.orc_fista_l1(self.h, X.ctypes.data, _p(smaps), data.shape[0], data.ctypes.data,
                              kx.ctypes.data, ky.ctypes.data, kz.ctypes.data, kw.ctypes.data,
                              kx.shape[0], iters, float(step), float(fraction), lv.ctypes.data,
                              int(spins), sh.ctypes.data, resid.ctypes.data, thr.ctypes.data)
        return X, resid, thr

    def fista_cauchy(self, smaps, data, traj, iters, step=0.0, update_number=30):
        """FISTA with the Cauchy step; returns (X, resid, steps (complex))."""
        kx, ky, kz, kw = self._traj(traj)
        data, smaps = self._c(data), self._c(smaps)
        X = np.zeros((self.N[2], self.N[1], self.N[0]), self.ctype)
        resid = np.zeros(iters)
        steps = np.zeros(2 * iters)
        self.lib.orc_fista_cauchy(self.h, X.ctypes.data, _p(smaps), data.shape[0], data.ctypes.data,
                                  kx.ctypes.data, ky.ctypes.data, kz.ctypes.data, kw.ctypes.data,
                                  kx.shape[0], iters, float(step), int(update_number), resid.ctypes.data,
                                  steps.ctypes.data)
        return X, resid, steps[0::2] + 1j * steps[1::2]

    def fista(self, smaps, data, traj, iters, step):
        kx, ky, kz, kw = self._traj(traj)
        data, smaps = self._c(data), self._c(smaps)
        X = np.zeros((self.N[2], self.N[1], self.N[0]), self.ctype)
        resid = np.zeros(iters)
        self.lib.orc_fista(self.h, X.ctypes.data, _p(smaps), data.shape[0], data.ctypes.data,
                           kx.ctypes.data, ky.ctypes.data, kz.ctypes.data, kw.ctypes.data,
                           kx.shape[0], iters, float(step), resid.ctypes.data)
        return X, resid
