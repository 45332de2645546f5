"""MRI_Raw.h5 reader and writer (libmrn_io) on the CPU.

The reference ships no .h5 files, so no fixture of its own pins this row
(parity unpinned against the reference's bytes).  What pins the layout is the
reference's writer code restated here element by element through the HDF5 C
API (tests/cpp/h5_layout.c, which shares nothing with the library):
write_external_data (mri_data.cpp:571-766) writes /Kdata attributes as
STD_I32LE ints and NATIVE_FLOAT floats (hdf5_interface.cpp:82-135) and arrays
with their Blitz lengths reversed as HDF5 dimensions, complex<float> as the
compound {real: float at 0, imag: float at 4} (hdf5_interface.h:249-368).
  * a file written that way reads back through libmrn_io bit for bit;
  * a file libmrn_io writes has those stored types, names and dimensions, and
    its raw bytes are the arrays';
  * the reader's errors: missing attributes and datasets, mis-sized arrays.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from mri_recon_amd import raw_io

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpp", "h5_layout.c")
SO = os.path.join(HERE, "cpp", "libh5_layout.so")
HDF5 = "/opt/conda"

ROT = ("ix", "iy", "iz", "jx", "jy", "jz", "kx", "ky", "kz", "sx", "sy", "sz")


@pytest.fixture(scope="module")
def h5l():
    if not os.path.exists(os.path.join(HDF5, "include", "hdf5.h")):
        pytest.skip("no HDF5 headers")
    subprocess.run(["gcc", "-O1", "-fPIC", "-shared", "-idirafter", os.path.join(HDF5, "include"), SRC,
                    "-L", os.path.join(HDF5, "lib"), "-lhdf5", "-Wl,-rpath," + os.path.join(HDF5, "lib"),
                    "-o", SO], check=True)
    L = ctypes.CDLL(SO)
    i64, cp, vp, i32 = ctypes.c_int64, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int
    L.h5l_create.restype = L.h5l_open.restype = i64
    L.h5l_create.argtypes = L.h5l_open.argtypes = [cp]
    L.h5l_close.argtypes = [i64]
    L.h5l_attr_i32.argtypes = [i64, cp, cp, ctypes.c_int32]
    L.h5l_attr_f32.argtypes = [i64, cp, cp, ctypes.c_float]
    L.h5l_dataset.argtypes = [i64, cp, cp, i32, i32, ctypes.POINTER(ctypes.c_uint64), vp]
    IP = ctypes.POINTER(ctypes.c_int)
    L.h5l_attr_info.argtypes = [i64, cp, cp, IP, IP, IP, vp]
    L.h5l_dataset_info.argtypes = [i64, cp, cp, IP, ctypes.POINTER(ctypes.c_uint64), IP, IP, IP, IP, IP, cp, i32]
    L.h5l_dataset_raw.argtypes = [i64, cp, cp, vp]
    return L


KIND = {np.dtype(np.float32): 0, np.dtype(np.float64): 1, np.dtype(np.uint64): 2, np.dtype(np.complex64): 3}


def _put(L, fid, g, name, a):
    a = np.ascontiguousarray(a)
    hd = (ctypes.c_uint64 * a.ndim)(*a.shape)  # C-order shape = reversed Blitz lengths
    assert L.h5l_dataset(fid, g.encode(), name.encode(), KIND[a.dtype], a.ndim, hd, a.ctypes.data) == 0


def _problem(E=2, C=3, seed=0, gating=True, noise=True):
    rng = np.random.default_rng(seed)
    d = raw_io.RawData(num_coils=C, trajectory_type=(1, 1, 0), dft_needed=(1, 1, 1), num_frames=4,
                       fov=(220.0, 220.0, 180.5), rot=tuple(float(v) for v in rng.standard_normal(12)))
    for e in range(E):
        shape = (2 + e, 5, 7)  # (zres, yres, xres): Blitz lengths (7, 5, 2 + e)
        f = lambda: rng.standard_normal(shape).astype(np.float32)
        kd = (rng.standard_normal((C,) + shape) + 1j * rng.standard_normal((C,) + shape)).astype(np.complex64)
        d.encodings.append(raw_io.RawEncoding(f(), f(), f(), np.abs(f()), f(), kd))
        if gating:
            gs = (3, 4 + e)
            d.gating.append(raw_io.RawGating(
                ecg=rng.standard_normal(gs), resp=rng.standard_normal(gs), prep=rng.standard_normal(gs),
                time=rng.standard_normal(gs), clock=rng.integers(0, 2 ** 62, gs, dtype=np.uint64),
                k0=(rng.standard_normal((C,) + gs) + 1j * rng.standard_normal((C,) + gs)).astype(np.complex64)))
        else:
            d.gating.append(None)
    if noise:
        d.noise = (rng.standard_normal((C, 50)) + 1j * rng.standard_normal((C, 50))).astype(np.complex64)
    return d


def _write_like_reference(L, path, d):
    """write_external_data (mri_data.cpp:571-766), element by element."""
    fid = L.h5l_create(path.encode())
    assert fid >= 0
    K = b"Kdata"
    for nm, v in (("Num_Encodings", len(d.encodings)), ("Num_Coils", d.num_coils), ("Num_Frames", d.num_frames)):
        assert L.h5l_attr_i32(fid, K, nm.encode(), v) == 0
    for nm, v in zip(("fovx", "fovy", "fovz"), d.fov):
        assert L.h5l_attr_f32(fid, K, nm.encode(), v) == 0
    for nm, v in zip(ROT, d.rot):
        assert L.h5l_attr_f32(fid, K, nm.encode(), v) == 0
    for a, ax in enumerate("XYZ"):
        assert L.h5l_attr_i32(fid, K, ("trajectory_type" + ax).encode(), d.trajectory_type[a]) == 0
    for a, ax in enumerate("XYZ"):
        assert L.h5l_attr_i32(fid, K, ("dft_needed" + ax).encode(), int(d.dft_needed[a])) == 0
    for e, enc in enumerate(d.encodings):
        for nm, a in (("KT", enc.kt), ("KX", enc.kx), ("KY", enc.ky), ("KZ", enc.kz), ("KW", enc.kw)):
            _put(L, fid, "Kdata", "%s_E%d" % (nm, e), a)
        for c in range(d.num_coils):
            _put(L, fid, "Kdata", "KData_E%d_C%d" % (e, c), enc.kdata[c])
        g = d.gating[e]
        if g is not None:
            _put(L, fid, "Gating", "ECG_E%d" % e, g.ecg)
            _put(L, fid, "Gating", "CLOCK_E%d" % e, g.clock)
            _put(L, fid, "Gating", "RESP_E%d" % e, g.resp)
            _put(L, fid, "Gating", "PREP_E%d" % e, g.prep)
            _put(L, fid, "Gating", "TIME_E%d" % e, g.time)
            for c in range(d.num_coils):
                _put(L, fid, "Gating", "K0_E%d_C%d" % (e, c), g.k0[c])
    if d.noise is not None:
        _put(L, fid, "Kdata", "Noise", d.noise)
    assert L.h5l_close(fid) >= 0


def _assert_same(a, b):
    assert a.num_coils == b.num_coils and a.num_frames == b.num_frames
    assert tuple(a.trajectory_type) == tuple(b.trajectory_type)
    assert tuple(int(v) for v in a.dft_needed) == tuple(int(v) for v in b.dft_needed)
    assert np.array_equal(np.float32(a.fov), np.float32(b.fov)) and np.array_equal(np.float32(a.rot), np.float32(b.rot))
    assert len(a.encodings) == len(b.encodings)
    for x, y in zip(a.encodings, b.encodings):
        for f in ("kx", "ky", "kz", "kw", "kt", "kdata"):
            u, v = getattr(x, f), getattr(y, f)
            assert u.shape == v.shape and u.dtype == v.dtype and np.array_equal(u.view(np.uint8), v.view(np.uint8)), f
    for x, y in zip(a.gating, b.gating):
        assert (x is None) == (y is None)
        if x is not None:
            for f in ("ecg", "resp", "prep", "time", "clock", "k0"):
                assert np.array_equal(getattr(x, f), getattr(y, f)), f
    assert (a.noise is None) == (b.noise is None)
    if a.noise is not None:
        assert np.array_equal(a.noise, b.noise)


def test_reads_a_file_written_in_the_reference_layout(h5l, tmp_path):
    d = _problem()
    p = str(tmp_path / "ref_layout.h5")
    _write_like_reference(h5l, p, d)
    _assert_same(raw_io.read_raw(p), d)


def test_written_file_has_the_reference_types_names_and_bytes(h5l, tmp_path):
    d = _problem(E=1, C=2, seed=3)
    p = str(tmp_path / "ours.h5")
    raw_io.write_raw(p, d)
    L = h5l
    fid = L.h5l_open(p.encode())
    cls, size, order = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    v = (ctypes.c_uint8 * 8)()
    ints = {"Num_Encodings": 1, "Num_Coils": 2, "Num_Frames": 4, "trajectory_typeX": 1, "trajectory_typeY": 1,
            "trajectory_typeZ": 0, "dft_neededX": 1, "dft_neededY": 1, "dft_neededZ": 1}
    for nm, want in ints.items():
        assert L.h5l_attr_info(fid, b"Kdata", nm.encode(), ctypes.byref(cls), ctypes.byref(size), ctypes.byref(order), v) == 0
        assert (cls.value, size.value, order.value) == (L.h5l_class_integer(), 4, L.h5l_order_le()), nm
        assert int(np.frombuffer(bytes(v)[:4], "<i4")[0]) == want, nm
    for nm, want in list(zip(("fovx", "fovy", "fovz"), d.fov)) + list(zip(ROT, d.rot)):
        assert L.h5l_attr_info(fid, b"Kdata", nm.encode(), ctypes.byref(cls), ctypes.byref(size), ctypes.byref(order), v) == 0
        assert (cls.value, size.value) == (L.h5l_class_float(), 4), nm
        assert np.frombuffer(bytes(v)[:4], np.float32)[0] == np.float32(want), nm
    rank, csz, nm_, = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    hd = (ctypes.c_uint64 * 8)()
    mc, mo = (ctypes.c_int * 8)(), (ctypes.c_int * 8)()
    names = ctypes.create_string_buffer(256)
    enc = d.encodings[0]

    def info(g, name):
        assert L.h5l_dataset_info(fid, g.encode(), name.encode(), ctypes.byref(rank), hd, ctypes.byref(cls),
                                  ctypes.byref(csz), ctypes.byref(nm_), mc, mo, names, 256) == 0
        return tuple(hd[:rank.value])

    for nm, a in (("KX", enc.kx), ("KY", enc.ky), ("KZ", enc.kz), ("KW", enc.kw), ("KT", enc.kt)):
        assert info("Kdata", nm + "_E0") == a.shape
        assert (cls.value, csz.value) == (L.h5l_class_float(), 4)
        buf = np.empty_like(a)
        assert L.h5l_dataset_raw(fid, b"Kdata", (nm + "_E0").encode(), buf.ctypes.data) == 0
        assert np.array_equal(buf, a)
    for c in range(2):
        assert info("Kdata", "KData_E0_C%d" % c) == enc.kdata[c].shape
        assert (cls.value, csz.value, nm_.value) == (L.h5l_class_compound(), 8, 2)
        assert names.value == b"real,imag" and (mo[0], mo[1]) == (0, 4)
        assert (mc[0], mc[1]) == (L.h5l_class_float(), L.h5l_class_float())
        buf = np.empty_like(enc.kdata[c])
        assert L.h5l_dataset_raw(fid, b"Kdata", ("KData_E0_C%d" % c).encode(), buf.ctypes.data) == 0
        assert np.array_equal(buf, enc.kdata[c])
    g = d.gating[0]
    for nm, a, want_cls, want_size in (("ECG", g.ecg, L.h5l_class_float(), 8), ("CLOCK", g.clock, L.h5l_class_integer(), 8),
                                       ("TIME", g.time, L.h5l_class_float(), 8)):
        assert info("Gating", nm + "_E0") == a.shape and (cls.value, csz.value) == (want_cls, want_size)
    assert info("Kdata", "Noise") == d.noise.shape
    L.h5l_close(fid)


def test_round_trip_and_optional_parts(tmp_path):
    d = _problem(E=3, C=2, seed=5, gating=False, noise=False)
    d.encodings[1].kt = None
    p = str(tmp_path / "rt.h5")
    raw_io.write_raw(p, d)
    r = raw_io.read_raw(p)
    assert r.gating == [None, None, None] and r.noise is None
    assert r.encodings[1].kt is None
    for x, y in zip(r.encodings, d.encodings):
        assert np.array_equal(x.kx, y.kx) and np.array_equal(x.kdata, y.kdata)


def test_reader_errors(h5l, tmp_path):
    L = h5l
    p = str(tmp_path / "bad.h5")
    fid = L.h5l_create(p.encode())
    L.h5l_attr_i32(fid, b"Kdata", b"Num_Encodings", 1)
    L.h5l_close(fid)
    with pytest.raises(raw_io.RawError, match="Num_Coils"):
        raw_io.read_raw(p)
    # a coil dataset of another size
    d = _problem(E=1, C=2, seed=1, gating=False, noise=False)
    _write_like_reference(L, p, d)
    q = str(tmp_path / "bad2.h5")
    d.encodings[0].kdata = d.encodings[0].kdata[:, :, :, :3]
    with pytest.raises(ValueError):
        raw_io.write_raw(q, d)
    fid = L.h5l_create(q.encode())
    for nm in ("Num_Encodings", "Num_Coils", "trajectory_typeX", "trajectory_typeY", "trajectory_typeZ",
               "dft_neededX", "dft_neededY", "dft_neededZ"):
        L.h5l_attr_i32(fid, b"Kdata", nm.encode(), 1 if nm != "Num_Coils" else 1)
    enc = d.encodings[0]
    for nm, a in (("KX", enc.kx), ("KY", enc.ky), ("KZ", enc.kz), ("KW", enc.kw)):
        _put(L, fid, "Kdata", nm + "_E0", a)
    _put(L, fid, "Kdata", "KData_E0_C0", enc.kdata[0])  # 3 of 7 x-points
    L.h5l_close(fid)
    with raw_io.RawFile(q) as f:
        with pytest.raises(raw_io.RawError, match="points"):
            f.read_encoding(0, 1, with_kt=False)
        with pytest.raises(raw_io.RawError, match="KT_E0"):
            f.read_encoding(0, 1, with_kt=True)
    with pytest.raises(raw_io.RawError):
        raw_io.read_raw(str(tmp_path / "missing.h5"))


def test_gridfft_geometry_from_file_flags(tmp_path):
    d = _problem(E=1, C=1, gating=False, noise=False)
    d.trajectory_type, d.dft_needed = (1, 1, 0), (1, 1, 0)
    p = str(tmp_path / "flags.h5")
    raw_io.write_raw(p, d)
    kw = raw_io.read_raw(p).gridfft_kwargs((64, 64, 32))
    # precalc_gridding: grid where NONCARTESIAN, transform where dft_needed (gridFFT.cpp:484-498)
    assert kw == dict(N=(64, 64, 32), grid_in=(1, 1, 0), fft_in=(1, 1, 0))
    with pytest.raises(raw_io.RawError):
        raw_io.read_raw(p).gridfft_kwargs((64, 0, 32))
