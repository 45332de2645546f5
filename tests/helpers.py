"""Shared helpers for the parity tests (seeded small problems, error norms)."""
import math

import numpy as np
import torch

from mri_recon_amd import synth


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def crandn(shape, seed):
    rng = np.random.default_rng(seed)
    return ((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / math.sqrt(2)).astype(np.complex64)


def radial_problem(N, C, nspokes, npts, seed=0, kw_ones=False):
    """Golden-means radial trajectory (host float32 arrays), random image,
    smooth coil maps, random k-space."""
    kx, ky, kz, kw = [t.numpy() for t in synth.radial_traj(nspokes, npts, N)]
    if kw_ones:
        kw = np.ones_like(kw)
    X = crandn((N, N, N), seed)
    smaps = synth.coil_maps(C, N, seed=seed + 1).numpy()
    M = kx.shape[0]
    d = crandn((C, M), seed + 2)
    return (kx, ky, kz, kw), X, smaps, d


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def traj_dev(traj):
    return [to_dev(np.ascontiguousarray(t, dtype=np.float32)) for t in traj]


def stable_direction(d, k, h, N, overgrid, dwin=-1.0):
    """Zero the rows of d whose tap range would change between k - h d and k + h d.

    The truncated Kaiser-Bessel kernel (and the reference's table) is
    discontinuous at |u| = dwin, so a finite difference across a support edge
    measures the jump, not the derivative.  Tap ranges are computed as the
    operator computes them (dk = k * g + S/2 in float32, taps
    ceil(dk - dwin) .. floor(dk + dwin)) and as the FP64 model does.
    Returns (d with those rows zeroed, number of rows kept)."""
    from oracle import oracle as orc
    geo = orc.OracleOp(N, overgrid=overgrid, dwin=dwin, precision="f64").geometry()
    dw = np.float32(geo["dwin"][0])

    def ranges(kk):
        r = []
        for a in range(3):
            S2 = int(geo["S"][a]) // 2
            dk = (kk[:, a].astype(np.float32) * np.float32(geo["grid"][a])).astype(np.float32) + np.float32(S2)
            dk64 = kk[:, a].astype(np.float64) * float(geo["grid"][a]) + S2
            r += [np.ceil((dk - dw).astype(np.float32)), np.floor((dk + dw).astype(np.float32)),
                  np.ceil(dk64 - float(dw)), np.floor(dk64 + float(dw))]
        return r

    r0, rp, rm = ranges(k), ranges(k + h * d), ranges(k - h * d)
    keep = np.ones(k.shape[0], bool)
    for a, b, c in zip(r0, rp, rm):
        keep &= (a == b) & (a == c)
    return d * keep[:, None], int(keep.sum())
