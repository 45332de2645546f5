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
SOURCES = ["fft.hip", "grid.hip", "ops.hip", "grad_tune.hip", "wavelet.hip", "dcf.hip", "llr.hip", "gate.hip", "setup.cpp", "api.cpp"]
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
    return LIB


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
    a = ap.parse_args(argv)
    print(build(a.only, a.j, a.v))


if __name__ == "__main__":
    sys.exit(main())
