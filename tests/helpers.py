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
