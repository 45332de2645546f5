"""The C-ABI library loads and exports every symbol include/mrn.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from mri_recon_amd import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "mrn.h")).read()
    return sorted(set(re.findall(r"\b(mrn_[a-z_0-9]+)\s*\(", txt)))


def test_header_matches_binding_list():
    assert header_symbols() == sorted(_ffi.EXPORTS)


def test_library_loads_and_exports_everything():
    if not os.path.exists(_ffi.LIB_PATH):
        pytest.fail("libmrn.so is not built; run python -m mri_recon_amd.build")
    L = ctypes.CDLL(_ffi.LIB_PATH)
    for s in header_symbols():
        assert hasattr(L, s), s


def test_params_default_and_parse_args_host_only():
    L = _ffi.lib()
    p = _ffi.MrnParams()
    L.mrn_params_default(ctypes.byref(p))
    assert abs(p.overgrid - 1.5) < 1e-7 and p.kernel == 1 and abs(p.k_rad - 9999) < 1e-3
    assert list(p.grid_in) == [1, 1, 1] and list(p.dwin) == [-1, -1, -1]
    args = ["-fast_grid"]
    arr = (ctypes.c_char_p * 1)(*[a.encode() for a in args])
    assert L.mrn_parse_args(ctypes.byref(p), 1, arr) == 0
    assert p.overgrid == 1.0 and p.kernel == 0 and list(p.dwin) == [1, 1, 1]
    args = ["-overgrid", "1.25", "-dwinX", "2", "-sinc", "-grid_in_z", "0"]
    arr = (ctypes.c_char_p * len(args))(*[a.encode() for a in args])
    assert L.mrn_parse_args(ctypes.byref(p), len(args), arr) == 0
    assert p.overgrid == 1.25 and p.dwin[0] == 2.0 and p.kernel == 2 and p.grid_in[2] == 0
    bad = (ctypes.c_char_p * 1)(b"-overgrid")
    assert L.mrn_parse_args(ctypes.byref(p), 1, bad) == _ffi.MRN_EINVAL


def test_invalid_params_rejected_without_device():
    L = _ffi.lib()
    p = _ffi.MrnParams()
    L.mrn_params_default(ctypes.byref(p))
    h = ctypes.c_void_p()
    # N = 0 is rejected before any device call
    assert L.mrn_op_create(ctypes.byref(p), ctypes.byref(h)) == _ffi.MRN_EINVAL
    p.N[0] = p.N[1] = p.N[2] = 16
    p.ncoils = 0
    assert L.mrn_op_create(ctypes.byref(p), ctypes.byref(h)) == _ffi.MRN_EINVAL
    p.ncoils = 2
    p.coil_chunk = 3
    assert L.mrn_op_create(ctypes.byref(p), ctypes.byref(h)) == _ffi.MRN_EINVAL
    assert L.mrn_op_destroy(None) == _ffi.MRN_EINVAL
    assert L.mrn_status_string(4) == b"unsupported configuration"


def test_cpp_facade_compiles_against_header(tmp_path):
    """include/mrn_gridfft.hpp (the C++ gridFFT facade) compiles as plain C++17
    against include/mrn.h with the host compiler alone."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no host C++ compiler")
    src = tmp_path / "use.cpp"
    src.write_text('#include "mrn_gridfft.hpp"\n'
                   'int main() { mrn::gridFFT g; g.overgrid = 1.25f; mrn_params p; '
                   'mrn_params_default(&p); mrn::LowRankFrames l; mrn::GatedFrames gf; '
                   'mrn_dcf_params d; mrn_dcf_params_default(&d); (void)&mrn::dcf_calc; '
                   'return (int)(g.k_rad != p.k_rad) + l.p.block_iter * 0 + gf.p.frames * 0; }\n')
    r = subprocess.run([gxx, "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_io_library_loads_and_exports_everything():
    """libmrn_io.so (the MRI_Raw.h5 reader/writer) exports every symbol
    include/mrn_io.h declares."""
    from mri_recon_amd import raw_io
    txt = open(os.path.join(ROOT, "include", "mrn_io.h")).read()
    syms = sorted(set(re.findall(r"\b(mrn_raw_[a-z_0-9]+)\s*\(", txt)))
    assert len(syms) == 16
    L = raw_io.lib()
    for s in syms:
        assert hasattr(L, s), s
