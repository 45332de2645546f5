"""Build libmrn.so (the HIP library behind include/mrn.h) in-tree for gfx950.

    python -m mri_recon_amd.build            # compile every source, link
    python -m mri_recon_amd.build --only fft # recompile one source, relink
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libmrn.so")
SOURCES = ["fft.hip", "grid.hip", "ops.hip", "grad_tune.hip", "wavelet.hip", "dcf.hip", "llr.hip", "gate.hip", "setup.cpp", "api.cpp", "tune.cpp"]
ARCH = os.environ.get("MRN_OFFLOAD_ARCH", "gfx950")


def hipcc():
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(p):
        raise RuntimeError("hipcc not found: the HIP library cannot be built")
    return p


def flags():
    return ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", os.path.join(ROOT, "include"), "-Wno-unused-result", "-Wno-unused-value"]


def compile_one(src):
    out = os.path.join(OBJ, src + ".o")
    cmd = [hipcc()] + flags() + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError("hipcc failed on %s:\n%s" % (src, msg))
    return out, msg


def build(only=None, jobs=None, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = SOURCES if not only else [s for s in SOURCES if s.split(".")[0] in only]
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        for out, msg in ex.map(compile_one, srcs):
            if verbose and msg:
                print(msg)
    objs = [os.path.join(OBJ, s + ".o") for s in SOURCES]
    missing = [o for o in objs if not os.path.exists(o)]
    if missing:
        raise RuntimeError("missing objects (build all first): %s" % missing)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=" + ARCH, "-shared", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    build_facade_test()
    build_io()
    return LIB


IO_SRC = os.path.join(CSRC, "raw_io.cpp")
IO_LIB = os.path.join(PKG, "libmrn_io.so")
HDF5_PREFIX = os.environ.get("MRN_HDF5_PREFIX", "/opt/conda")


def build_io():
    """libmrn_io.so: the MRI_Raw.h5 reader/writer (host C++ over the HDF5 C
    library; no device code, so the operator library does not depend on HDF5)."""
    inc = os.path.join(HDF5_PREFIX, "include")
    libdir = os.path.join(HDF5_PREFIX, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        raise RuntimeError("hdf5.h not found under %s (set MRN_HDF5_PREFIX)" % inc)
    tmp = IO_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", os.path.join(ROOT, "include"),
           "-idirafter", inc, IO_SRC, "-L", libdir, "-lhdf5", "-Wl,-rpath," + libdir, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("libmrn_io build failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, IO_LIB)
    return IO_LIB


FACADE_SRC = os.path.join(ROOT, "tests", "cpp", "facade_cg.cpp")
FACADE_BIN = os.path.join(ROOT, "tests", "cpp", "facade_cg")


def build_facade_test():
    """The C++ program that drives the operator through include/mrn_gridfft.hpp
    (host code only, linked against libmrn.so)."""
    if not os.path.exists(FACADE_SRC):
        return None
    cmd = [hipcc(), "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), FACADE_SRC,
           "-L", PKG, "-lmrn", "-Wl,-rpath,$ORIGIN/../../mri_recon_amd", "-o", FACADE_BIN]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("facade test build failed:\n" + r.stdout + r.stderr)
    return FACADE_BIN


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", nargs="*")
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    ap.add_argument("--io", action="store_true", help="build only libmrn_io.so")
    a = ap.parse_args(argv)
    print(build_io() if a.io else build(a.only, a.j, a.v))


if __name__ == "__main__":
    sys.exit(main())
