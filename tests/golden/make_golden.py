#!/usr/bin/env python3
"""Writes the golden fixtures of the gridFFT operator (tests/golden/*.npz).

The reference's own tests hold no vectors for this path (its test programs do
not call gridFFT, fft3 or the solvers), so the fixtures are produced by the
pinned CPU oracle (oracle/gridfft_oracle.c): FP32 for the tables and windows
(the reference computes them in float), FP64 for the end-to-end vectors.

    geometry.npz   per configuration: S, og, L, beta, KB table and deapodisation
                   window (x axis; the configurations are isotropic)
    tiny_e2e.npz   N=16 (grid 32^3), 2 coils, 100 golden-means radial spokes x
                   12 centre-out points: inputs X, smaps, kx/ky/kz/kw, d; FP64
                   outputs A X (gridFFT::backward), A^H d (gridFFT::forward,
                   with kw and taper), E'E X, and a 5-iteration SENSE-CG image
                   with its scalar trace
    MANIFEST.json  SHA-256 of each file

Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402

# (name, N, overgrid, dwin): the five benchmark configurations, the
# oversampling-1.5 variant of the 256^3 case and the small rounding cases
CONFIGS = [
    ("c1_128_os1.5_w4", 128, 1.5, 2.0),
    ("c2_256_os1.25_w5", 256, 1.25, 2.5),
    ("c2_256_os1.5_w5", 256, 1.5, 2.5),
    ("c3_160_os1.5_w5", 160, 1.5, 2.5),
    ("c4_192_os1.5_w5", 192, 1.5, 2.5),
    ("c5_320_os1.5_w5", 320, 1.5, 2.5),
    ("n16_os1.5_default", 16, 1.5, -1.0),
    ("n32_os1.5_w3", 32, 1.5, 1.5),
]

TINY = dict(N=16, C=2, nspokes=100, npts=12, overgrid=1.5, dwin=-1.0, cg_iters=5, seed=2024)


def golden_radial(nspokes, npts, N):
    """Golden-means kooshball, centre-out, |k| <= N/2, kw = |k|^2 (float32).
    Same recipe as mri_recon_amd.synth.radial_traj, restated here so that the
    fixture does not depend on the package under test."""
    n = np.arange(nspokes, dtype=np.float64)
    kzd = 2.0 * np.mod(0.465571231876768 * n, 1.0) - 1.0
    phi = 2.0 * math.pi * np.mod(0.682327803828019 * n, 1.0)
    st = np.sqrt(np.clip(1.0 - kzd * kzd, 0.0, None))
    r = 0.5 * np.arange(npts, dtype=np.float64) / (npts - 1)
    kx = (N * r[None, :] * (st * np.cos(phi))[:, None]).ravel()
    ky = (N * r[None, :] * (st * np.sin(phi))[:, None]).ravel()
    kz = (N * r[None, :] * kzd[:, None]).ravel()
    kw = kx * kx + ky * ky + kz * kz
    return [a.astype(np.float32) for a in (kx, ky, kz, kw)]


def crandn(rng, shape):
    return ((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / math.sqrt(2)).astype(np.complex64)


def smooth_maps(rng, C, N):
    """Smooth random-phase coil maps, sum of squares normalised to 1."""
    z, y, x = np.meshgrid(*(np.arange(N) - N / 2,) * 3, indexing="ij")
    maps = []
    for c in range(C):
        a = 2 * math.pi * c / C
        cx, cy, cz = 0.6 * N * math.cos(a), 0.6 * N * math.sin(a), 0.2 * N * (c - (C - 1) / 2)
        d2 = ((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) / (N * N)
        ph = rng.uniform(0, 2 * math.pi) + 2 * math.pi * (x * rng.uniform(-1, 1) + y * rng.uniform(-1, 1)) / (4 * N)
        maps.append(np.exp(-d2) * np.exp(1j * ph))
    maps = np.array(maps)
    maps /= np.sqrt(np.sum(np.abs(maps) ** 2, axis=0, keepdims=True))
    return maps.astype(np.complex64)


def make_geometry():
    out = {}
    for name, N, os_, dwin in CONFIGS:
        op = orc.OracleOp(N, overgrid=os_, dwin=dwin, precision="f32")
        g = op.geometry()
        out[name + "/params"] = np.array([N, os_, dwin])
        out[name + "/S"] = g["S"].astype(np.int32)
        out[name + "/og"] = g["og"].astype(np.int32)
        out[name + "/L"] = g["L"].astype(np.int32)
        out[name + "/beta"] = g["beta"].astype(np.float64)
        out[name + "/table"] = op.table(0).astype(np.float32)
        out[name + "/window"] = op.window(0).astype(np.float32)
    return out


def make_tiny():
    t = TINY
    rng = np.random.default_rng(t["seed"])
    N, C = t["N"], t["C"]
    traj = golden_radial(t["nspokes"], t["npts"], N)
    X = crandn(rng, (N, N, N))
    smaps = smooth_maps(rng, C, N)
    d = crandn(rng, (C, traj[0].shape[0]))
    op = orc.OracleOp(N, overgrid=t["overgrid"], dwin=t["dwin"], precision="f64")
    A = op.image_to_kspace(X.astype(np.complex128), smaps, traj)
    AH = op.kspace_to_image(d.astype(np.complex128), smaps, traj)
    EE = op.normal(X.astype(np.complex128), smaps, traj)
    Xcg, trace = op.cg(smaps, d.astype(np.complex128), traj, t["cg_iters"])
    g = op.geometry()
    return {
        "params": np.array([N, C, t["nspokes"], t["npts"], t["overgrid"], t["dwin"], t["cg_iters"], t["seed"]]),
        "S": g["S"].astype(np.int32),
        "kx": traj[0], "ky": traj[1], "kz": traj[2], "kw": traj[3],
        "X": X, "smaps": smaps, "d": d,
        "A": A, "AH": AH, "EE": EE, "cg_X": Xcg, "cg_trace": trace,
    }


def sha256(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        h.update(f.read())
    return h.hexdigest()


def main():
    orc.build()
    files = {"geometry.npz": make_geometry(), "tiny_e2e.npz": make_tiny()}
    manifest = {}
    for fn, arrays in files.items():
        path = os.path.join(HERE, fn)
        np.savez_compressed(path, **arrays)
        manifest[fn] = sha256(path)
        print("wrote %s (%d arrays, %d bytes)" % (fn, len(arrays), os.path.getsize(path)))
    with open(os.path.join(HERE, "MANIFEST.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
