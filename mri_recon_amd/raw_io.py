"""MRI_Raw.h5 reader and writer (libmrn_io.so, host code).

``read_raw`` / ``write_raw`` follow MRI_DATA::read_external_data /
write_external_data (src/mri_data.cpp:571-902): the /Kdata attributes
(Num_Encodings, Num_Coils, trajectory_type*, dft_needed*, FOV, rotation),
per encoding the trajectory KX/KY/KZ/KW/KT_E# and the k-space KData_E#_C#,
the /Gating arrays and the noise samples.  Arrays are numpy arrays in the
reference's memory order: a Blitz array of lengths (n0, n1, n2), n0 fastest,
is returned with shape (n2, n1, n0) in C order (the HDF5 dataset's shape).

``RawData.gridfft_kwargs(recon_N)`` gives GridFFT's geometry arguments as
gridFFT::precalc_gridding takes them from MRI_DATA (gridFFT.cpp:484-498), with
an explicit reconstruction size (read_external_data never sets recon_res).
"""
import ctypes
import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from ._ffi import MrnParams

_PKG = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class RawInfo(ctypes.Structure):
    _fields_ = [
        ("num_encodings", ctypes.c_int32),
        ("num_coils", ctypes.c_int32),
        ("num_frames", ctypes.c_int32),
        ("trajectory_type", ctypes.c_int32 * 3),
        ("dft_needed", ctypes.c_int32 * 3),
        ("fov", ctypes.c_float * 3),
        ("rot", ctypes.c_float * 12),
        ("reserved", ctypes.c_int32 * 8),
    ]


class RawError(RuntimeError):
    pass


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_PKG, "libmrn_io.so")
        if not os.path.exists(path):
            raise RawError("libmrn_io.so is not built (python -m mri_recon_amd.build)")
        L = ctypes.CDLL(path)
        vp, i32, i64p = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)
        L.mrn_raw_open.argtypes = [ctypes.c_char_p, i32, ctypes.POINTER(vp)]
        L.mrn_raw_close.argtypes = [vp]
        L.mrn_raw_get_info.argtypes = [vp, ctypes.POINTER(RawInfo)]
        L.mrn_raw_put_info.argtypes = [vp, ctypes.POINTER(RawInfo)]
        L.mrn_raw_exists.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p]
        L.mrn_raw_encoding_dims.argtypes = [vp, i32, i64p]
        L.mrn_raw_read_encoding.argtypes = [vp, i32] + [vp] * 6
        L.mrn_raw_write_encoding.argtypes = [vp, i32, i64p] + [vp] * 6
        L.mrn_raw_gating_dims.argtypes = [vp, i32, i64p]
        L.mrn_raw_read_gating.argtypes = [vp, i32] + [vp] * 6
        L.mrn_raw_write_gating.argtypes = [vp, i32, i64p] + [vp] * 6
        L.mrn_raw_noise_dims.argtypes = [vp, i64p]
        L.mrn_raw_read_noise.argtypes = [vp, vp]
        L.mrn_raw_write_noise.argtypes = [vp, i64p, vp]
        L.mrn_raw_params.argtypes = [ctypes.POINTER(RawInfo), ctypes.POINTER(ctypes.c_int32),
                                     ctypes.POINTER(MrnParams)]
        L.mrn_raw_error.restype = ctypes.c_char_p
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        raise RawError("%s (status %d)" % (lib().mrn_raw_error().decode(), rc))


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _dims(shape):
    """Blitz lengths (fastest first) of a C-order numpy shape."""
    return (ctypes.c_int64 * len(shape))(*reversed([int(s) for s in shape]))


@dataclass
class RawEncoding:
    kx: np.ndarray          # float32 (n2, n1, n0)
    ky: np.ndarray
    kz: np.ndarray
    kw: np.ndarray
    kt: Optional[np.ndarray]
    kdata: np.ndarray       # complex64 (C, n2, n1, n0)


@dataclass
class RawGating:
    ecg: np.ndarray         # float64 (m1, m0)
    resp: Optional[np.ndarray] = None
    prep: Optional[np.ndarray] = None
    time: Optional[np.ndarray] = None
    clock: Optional[np.ndarray] = None   # uint64
    k0: Optional[np.ndarray] = None      # complex64 (C, m1, m0)


@dataclass
class RawData:
    num_coils: int
    trajectory_type: tuple = (1, 1, 1)   # 0 Cartesian, 1 non-Cartesian per axis (x, y, z)
    dft_needed: tuple = (1, 1, 1)
    num_frames: int = 1
    fov: tuple = (0.0, 0.0, 0.0)
    rot: tuple = (0.0,) * 12
    encodings: List[RawEncoding] = field(default_factory=list)
    gating: List[Optional[RawGating]] = field(default_factory=list)
    noise: Optional[np.ndarray] = None   # complex64 (m1, m0)

    def info(self):
        r = RawInfo()
        r.num_encodings = len(self.encodings)
        r.num_coils = int(self.num_coils)
        r.num_frames = int(self.num_frames)
        for a in range(3):
            r.trajectory_type[a] = int(self.trajectory_type[a])
            r.dft_needed[a] = int(bool(self.dft_needed[a]))
            r.fov[a] = float(self.fov[a])
        for q in range(12):
            r.rot[q] = float(self.rot[q])
        return r

    def gridfft_kwargs(self, recon_N):
        """N, grid_in and fft_in for GridFFT (mrn_raw_params)."""
        p = MrnParams()
        n = (ctypes.c_int32 * 3)(*[int(v) for v in (recon_N if np.ndim(recon_N) else (recon_N,) * 3)])
        _check(lib().mrn_raw_params(ctypes.byref(self.info()), n, ctypes.byref(p)))
        return dict(N=tuple(p.N), grid_in=tuple(p.grid_in), fft_in=tuple(p.fft_in))


class RawFile:
    """An open MRI_Raw.h5 file (context manager)."""

    def __init__(self, path, write=False):
        h = ctypes.c_void_p()
        _check(lib().mrn_raw_open(os.fsencode(path), int(bool(write)), ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            _check(lib().mrn_raw_close(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def exists(self, group, name):
        r = lib().mrn_raw_exists(self.h, group.encode(), name.encode())
        if r < 0:
            _check(-r)
        return bool(r)

    def info(self):
        r = RawInfo()
        _check(lib().mrn_raw_get_info(self.h, ctypes.byref(r)))
        return r

    def read_encoding(self, e, C, with_kt=True):
        d = (ctypes.c_int64 * 3)()
        _check(lib().mrn_raw_encoding_dims(self.h, e, d))
        shape = (d[2], d[1], d[0])
        arrs = [np.empty(shape, np.float32) for _ in range(4)]
        kt = np.empty(shape, np.float32) if with_kt else None
        kd = np.empty((C,) + shape, np.complex64)
        _check(lib().mrn_raw_read_encoding(self.h, e, *[_p(a) for a in arrs], _p(kt), _p(kd)))
        return RawEncoding(*arrs, kt, kd)

    def read_gating(self, e, C):
        d = (ctypes.c_int64 * 2)()
        if lib().mrn_raw_gating_dims(self.h, e, d) != 0:
            return None
        shape = (d[1], d[0])
        g = RawGating(ecg=np.empty(shape), resp=np.empty(shape), prep=np.empty(shape), time=np.empty(shape),
                      clock=np.empty(shape, np.uint64))
        if self.exists("Gating", "K0_E%d_C0" % e) and C > 0:
            g.k0 = np.empty((C,) + shape, np.complex64)
        _check(lib().mrn_raw_read_gating(self.h, e, _p(g.ecg), _p(g.resp), _p(g.prep), _p(g.time), _p(g.clock),
                                         _p(g.k0)))
        return g

    def read_noise(self):
        d = (ctypes.c_int64 * 2)()
        if lib().mrn_raw_noise_dims(self.h, d) != 0:
            return None
        out = np.empty((d[1], d[0]), np.complex64)
        _check(lib().mrn_raw_read_noise(self.h, _p(out)))
        return out


def read_raw(path, with_kt=True):
    """Read a whole MRI_Raw.h5 file into a RawData."""
    with RawFile(path) as f:
        r = f.info()
        C = r.num_coils
        out = RawData(num_coils=C, trajectory_type=tuple(r.trajectory_type), dft_needed=tuple(r.dft_needed),
                      num_frames=r.num_frames, fov=tuple(r.fov), rot=tuple(r.rot))
        for e in range(r.num_encodings):
            out.encodings.append(f.read_encoding(e, C, with_kt=with_kt and f.exists("Kdata", "KT_E%d" % e)))
            out.gating.append(f.read_gating(e, C))
        out.noise = f.read_noise()
    return out


def write_raw(path, data):
    """Write a RawData as MRI_Raw.h5 (write_external_data's layout)."""
    C = int(data.num_coils)
    with RawFile(path, write=True) as f:
        _check(lib().mrn_raw_put_info(f.h, ctypes.byref(data.info())))
        for e, enc in enumerate(data.encodings):
            arrs = [np.ascontiguousarray(a, np.float32) for a in (enc.kx, enc.ky, enc.kz, enc.kw)]
            shape = arrs[0].shape
            if len(shape) != 3 or any(a.shape != shape for a in arrs):
                raise ValueError("encoding %d: kx, ky, kz, kw must share one 3-D shape" % e)
            kt = None if enc.kt is None else np.ascontiguousarray(enc.kt, np.float32)
            kd = np.ascontiguousarray(enc.kdata, np.complex64)
            if kd.shape != (C,) + shape or (kt is not None and kt.shape != shape):
                raise ValueError("encoding %d: kdata must be (C,) + %s" % (e, shape))
            _check(lib().mrn_raw_write_encoding(f.h, e, _dims(shape), *[_p(a) for a in arrs], _p(kt), _p(kd)))
            g = data.gating[e] if e < len(data.gating) else None
            if g is not None:
                ecg = np.ascontiguousarray(g.ecg, np.float64)
                cv = lambda a, t: None if a is None else np.ascontiguousarray(a, t)
                rs, pr, tm = cv(g.resp, np.float64), cv(g.prep, np.float64), cv(g.time, np.float64)
                ck, k0 = cv(g.clock, np.uint64), cv(g.k0, np.complex64)
                _check(lib().mrn_raw_write_gating(f.h, e, _dims(ecg.shape), _p(ecg), _p(rs), _p(pr), _p(tm), _p(ck),
                                                  _p(k0)))
        if data.noise is not None:
            nz = np.ascontiguousarray(data.noise, np.complex64)
            _check(lib().mrn_raw_write_noise(f.h, _dims(nz.shape), _p(nz)))
