"""Synthetic inputs for the NUFFT/SENSE operator: trajectories, phantoms, maps.

Everything is generated with torch so that the same code runs on the host for
small test cases and on the GPU for the 256^3 x 32-coil workloads.  Units follow
the reference: k in cycles per FOV (|k| <= N/2, src/mri_data.h:50-57), images
(Nz, Ny, Nx) with x fastest.

Radial directions use the golden-means pair of src/test_voronoi.cpp:44-50; the
density compensation kw = |k|^2 follows the reference simulator
(src/commandline_recon.cxx:60-75).
"""
import math

import numpy as np
import torch

GM1 = 0.465571231876768
GM2 = 0.682327803828019


def radial_traj(nspokes, npts, N, first_spoke=0, device="cpu"):
    """3-D centre-out kooshball: spoke n along a golden-means direction,
    readout p at radius 0.5*p/(npts-1) (normalised), scaled by N.

    Returns float32 tensors kx, ky, kz, kw of shape (nspokes*npts,), readout
    fastest.  kw = |k|^2, so the centre sample has kw = 0 (skipped by the
    operator as in the reference)."""
    n = torch.arange(first_spoke, first_spoke + nspokes, dtype=torch.float64, device=device)
    kzd = 2.0 * torch.remainder(GM1 * n, 1.0) - 1.0
    phi = 2.0 * math.pi * torch.remainder(GM2 * n, 1.0)
    rxy = torch.sqrt(torch.clamp(1.0 - kzd * kzd, min=0.0))
    dx, dy = rxy * torch.cos(phi), rxy * torch.sin(phi)
    r = 0.5 * torch.arange(npts, dtype=torch.float64, device=device) / max(npts - 1, 1)
    kx = (dx[:, None] * r[None, :] * N).reshape(-1)
    ky = (dy[:, None] * r[None, :] * N).reshape(-1)
    kz = (kzd[:, None] * r[None, :] * N).reshape(-1)
    kw = kx * kx + ky * ky + kz * kz
    f = lambda t: t.to(torch.float32).contiguous()
    return f(kx), f(ky), f(kz), f(kw)


def cones_traj(ninterleaves, npts, N, nlevels=64, turns=8.0, device="cpu"):
    """Stack-of-cones-like trajectory: interleave i sits on cone level
    theta_j (j = i mod nlevels, uniform in cos theta), starts at the golden
    angle times i and spirals by 2*pi*turns*sin(theta)*t, t in [0, 1)."""
    i = torch.arange(ninterleaves, dtype=torch.float64, device=device)
    j = torch.remainder(i, nlevels)
    theta = torch.arccos(1.0 - 2.0 * (j + 0.5) / nlevels)
    golden = math.pi * (3.0 - math.sqrt(5.0))
    t = torch.arange(npts, dtype=torch.float64, device=device) / npts
    phi = golden * i[:, None] + 2.0 * math.pi * turns * torch.sin(theta)[:, None] * t[None, :]
    r = 0.5 * t[None, :] * N
    kx = r * torch.sin(theta)[:, None] * torch.cos(phi)
    ky = r * torch.sin(theta)[:, None] * torch.sin(phi)
    kz = r * torch.cos(theta)[:, None].expand_as(phi)
    kw = kx * kx + ky * ky + kz * kz
    f = lambda a: a.reshape(-1).to(torch.float32).contiguous()
    return f(kx), f(ky), f(kz), f(kw)


def gated_radial_frames(nframes, mean_spokes, npts, N, seed, device="cpu"):
    """Frames with Poisson-distributed spoke counts continuing the golden-means
    index; returns a list of (kx, ky, kz, kw) per frame."""
    rng = np.random.default_rng(seed)
    counts = rng.poisson(mean_spokes, size=nframes)
    out, first = [], 0
    for c in counts:
        out.append(radial_traj(int(c), npts, N, first_spoke=first, device=device))
        first += int(c)
    return out


def _gen(seed, device):
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    return g


def ellipsoid_phantom(N, n_ellipsoids=10, seed=0, device="cpu"):
    """Sum of random ellipsoids: centres U(-0.3, 0.3), semi-axes U(0.05, 0.3),
    uniform random rotation, intensity U(0.1, 1); rasterised at voxel centres
    (i - N/2)/N.  Returns complex64 (N, N, N)."""
    rng = np.random.default_rng(seed)
    c = (torch.arange(N, dtype=torch.float32, device=device) - N / 2) / N
    z, y, x = torch.meshgrid(c, c, c, indexing="ij")
    img = torch.zeros((N, N, N), dtype=torch.float32, device=device)
    for _ in range(n_ellipsoids):
        cen = rng.uniform(-0.3, 0.3, 3)
        ax = rng.uniform(0.05, 0.3, 3)
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        a, b, cq, d = q
        R = np.array([[a * a + b * b - cq * cq - d * d, 2 * (b * cq - a * d), 2 * (b * d + a * cq)],
                      [2 * (b * cq + a * d), a * a - b * b + cq * cq - d * d, 2 * (cq * d - a * b)],
                      [2 * (b * d - a * cq), 2 * (cq * d + a * b), a * a - b * b - cq * cq + d * d]])
        val = rng.uniform(0.1, 1.0)
        px, py, pz = x - cen[0], y - cen[1], z - cen[2]
        u = [R[0, 0] * px + R[1, 0] * py + R[2, 0] * pz,
             R[0, 1] * px + R[1, 1] * py + R[2, 1] * pz,
             R[0, 2] * px + R[1, 2] * py + R[2, 2] * pz]
        inside = (u[0] / ax[0]) ** 2 + (u[1] / ax[1]) ** 2 + (u[2] / ax[2]) ** 2 <= 1.0
        img += inside.to(torch.float32) * float(val)
    return img.to(torch.complex64)


def coil_maps(C, N, seed=0, sigma_frac=0.1, device="cpu"):
    """Smooth random-phase coil maps: a CN(0,1) field per coil, periodic
    separable Gaussian blur with sigma = sigma_frac*N, then normalised so that
    sum_c |S_c|^2 = 1.  Returns complex64 (C, N, N, N)."""
    g = _gen(seed, device)
    f = torch.fft.fftfreq(N, device=device, dtype=torch.float32)
    sig = sigma_frac * N
    h = torch.exp(-2.0 * (math.pi * sig * f) ** 2)
    H = h[:, None, None] * h[None, :, None] * h[None, None, :]
    maps = torch.empty((C, N, N, N), dtype=torch.complex64, device=device)
    for c in range(C):
        re = torch.randn((N, N, N), generator=g, device=device)
        im = torch.randn((N, N, N), generator=g, device=device)
        field = torch.complex(re, im)
        maps[c] = torch.fft.ifftn(torch.fft.fftn(field) * H)
        del field, re, im
    sos = torch.zeros((N, N, N), dtype=torch.float32, device=device)
    for c in range(C):
        sos += maps[c].abs() ** 2
    inv = torch.rsqrt(torch.clamp(sos, min=1e-30))
    for c in range(C):
        maps[c] *= inv
    return maps


def cn_image(shape, seed=0, device="cpu"):
    """i.i.d. CN(0,1) complex64 array (unit variance: E|x|^2 = 1)."""
    g = _gen(seed, device)
    re = torch.randn(shape, generator=g, device=device)
    im = torch.randn(shape, generator=g, device=device)
    return (torch.complex(re, im) * math.sqrt(0.5)).to(torch.complex64)


def add_noise(data, rel=0.01, seed=0):
    """data + CN(0, sigma^2) with sigma = rel * rms(data) per coil (rows)."""
    g = _gen(seed, data.device)
    rms = torch.sqrt(torch.mean(data.abs() ** 2, dim=-1, keepdim=True))
    n = torch.complex(torch.randn(data.shape, generator=g, device=data.device),
                      torch.randn(data.shape, generator=g, device=data.device)) * math.sqrt(0.5)
    return (data + rel * rms * n).to(torch.complex64)
