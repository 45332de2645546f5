"""Differentiation of the encoding operator (SenseEncoding / mrn_traj_grad).

Loss L = 0.5 || W^(1/2) (A(x; k) - y) ||^2 with gradients w.r.t. the image x
and the trajectory k (M, 3).  Checks:
* CPU: the FP64 reference gradient (tests/kb_grad_ref.py) is pinned against the
  oracle's gather and against central finite differences (h = 1e-3);
* GPU: the autograd image gradient equals A^H W (A x - y) from the FP64
  oracle and passes the dot-product adjoint test; the trajectory gradient
  matches the FP64 analytic gradient (exact_kb) and central finite
  differences of the loss (h = 1e-3, both kernel modes).
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from mri_recon_amd import synth
from tests.helpers import crandn, rel
from tests.kb_grad_ref import ExactKBGather

N, C, OS = 16, 2, 1.5


def _problem(seed=5, nint=48, npts=40):
    kx, ky, kz, kw = [t.numpy() for t in synth.cones_traj(nint, npts, N, nlevels=8, turns=2.0)]
    k = np.stack([kx, ky, kz], axis=1).astype(np.float32)
    x = crandn((N, N, N), seed)
    smaps = synth.coil_maps(C, N, seed=seed + 1).numpy()
    y = crandn((C, k.shape[0]), seed + 2)
    w = np.maximum(kw, 0.05).astype(np.float32)  # loss weights; all samples kept
    return k, x, smaps, y, w


def _ref_model():
    o = orc.OracleOp(N, overgrid=OS, precision="f64", k_rad=math.inf)
    return o, ExactKBGather(o)


def _grids(o, x, smaps):
    return np.stack([o.stage_pre_fft(x.astype(np.complex128), smaps[c]) for c in range(C)])


def _stable_direction(d, k, h):
    """Zero the rows of d whose tap range would change between k - h d and k + h d.

    The truncated Kaiser-Bessel kernel (and the reference's table) is
    discontinuous at |u| = dwin, so a finite difference across a support edge
    measures the jump, not the derivative.  Tap ranges are computed as the
    operator computes them (dk = k * g + S/2 in float32, taps
    ceil(dk - dwin) .. floor(dk + dwin)) and as the FP64 model does."""
    geo = orc.OracleOp(N, overgrid=OS, precision="f64").geometry()
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


def _loss64(y_model, y, w):
    r = y_model - y[..., None] if y_model.ndim == 3 else y_model - y
    return 0.5 * float(np.sum(w * np.abs(r) ** 2))


# ------------------------------------------------------------------- CPU pins
def test_ref_gather_matches_oracle_table_gather():
    k, x, smaps, _, _ = _problem()
    o, m = _ref_model()
    G = _grids(o, x, smaps)
    y, _ = m.apply(G, k.astype(np.float64))
    traj = [k[:, 0], k[:, 1], k[:, 2], np.ones(k.shape[0], np.float32)]
    yo = np.stack([o.stage_gather(G[c], traj) for c in range(C)])
    # exact KB vs the 600-point linear-interpolation table: interpolation error only
    assert rel(y, yo) < 1e-4


def test_ref_gradient_matches_central_differences():
    k, x, smaps, y, w = _problem()
    o, m = _ref_model()
    G = _grids(o, x, smaps)
    k64 = k.astype(np.float64)
    ym, dy = m.apply(G, k64)
    g_ref = np.real(np.einsum("cm,cma->ma", np.conj(w * (ym - y)), dy))
    h = 1e-3
    rng = np.random.default_rng(0)
    for _ in range(3):
        d, nkeep = _stable_direction(rng.standard_normal(k.shape), k, h)
        assert nkeep > 0.8 * k.shape[0]
        lp = _loss64(m.apply(G, k64 + h * d)[0], y, w)
        lm = _loss64(m.apply(G, k64 - h * d)[0], y, w)
        fd = (lp - lm) / (2 * h)
        an = float(np.sum(g_ref * d))
        assert abs(fd - an) < 1e-5 * abs(an)


# ------------------------------------------------------------------- GPU
def _dev_setup(mode):
    import torch

    from mri_recon_amd import SenseEncoding
    k, x, smaps, y, w = _problem()
    enc = SenseEncoding(N, torch.from_numpy(smaps).cuda(), overgrid=OS, kernel_mode=mode)
    return enc, k, x, smaps, y, w


def _loss_t(enc, xt, kt, yt, wt):
    import torch
    r = enc(xt, kt) - yt
    return 0.5 * torch.sum(wt * (r.real ** 2 + r.imag ** 2))


def _loss_fd(enc, xt, kt, yt, wt):
    """The loss accumulated in float64 from the operator's float32 output (a
    float32 sum of ~1e5 rounds to more than the difference a step of h makes)."""
    import torch
    r = enc(xt, kt).to(torch.complex128) - yt.to(torch.complex128)
    return 0.5 * float(torch.sum(wt.double() * (r.real ** 2 + r.imag ** 2)))


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["table", "exact_kb"])
def test_image_gradient_vs_fp64_oracle(mode):
    import torch
    enc, k, x, smaps, y, w = _dev_setup(mode)
    xt = torch.from_numpy(x).cuda().requires_grad_(True)
    kt = torch.from_numpy(k).cuda()
    yt, wt = torch.from_numpy(y).cuda(), torch.from_numpy(w).cuda()
    L = _loss_t(enc, xt, kt, yt, wt)
    L.backward()
    gx = xt.grad.cpu().numpy()
    o = orc.OracleOp(N, overgrid=OS, precision="f64", k_rad=math.inf)
    ones = np.ones(k.shape[0], np.float32)
    traj = [k[:, 0], k[:, 1], k[:, 2], ones]
    if mode == "table":
        r = o.image_to_kspace(x.astype(np.complex128), smaps, traj) - y
    else:
        _, m = _ref_model()
        r = m.apply(_grids(o, x, smaps), k.astype(np.float64))[0] - y
    ref = o.kspace_to_image(r * 1.0, smaps, [k[:, 0], k[:, 1], k[:, 2], w])
    tol = 1e-5 if mode == "table" else 1e-4  # exact_kb: the adjoint's table vs exact KB differ
    assert rel(gx, ref) < tol
    enc.close()


@pytest.mark.gpu
def test_image_gradient_dot_test():
    import torch
    enc, k, x, smaps, y, _ = _dev_setup("table")
    kt = torch.from_numpy(k).cuda()
    xt = torch.from_numpy(x).cuda().requires_grad_(True)
    v = torch.from_numpy(y).cuda()
    Ax = enc(xt, kt)
    # L = Re <A x, v>  ->  x.grad = A^H v
    L = torch.sum((torch.conj(Ax) * v).real)
    L.backward()
    lhs = torch.sum(torch.conj(Ax.detach().to(torch.complex128)) * v.to(torch.complex128))
    rhs = torch.sum(torch.conj(xt.detach().to(torch.complex128)) * xt.grad.to(torch.complex128))
    assert abs(lhs - rhs) / abs(lhs) < 1e-5
    enc.close()


@pytest.mark.gpu
def test_image_gradient_finite_differences():
    import torch
    enc, k, x, smaps, y, w = _dev_setup("table")
    kt = torch.from_numpy(k).cuda()
    yt, wt = torch.from_numpy(y).cuda(), torch.from_numpy(w).cuda()
    xt = torch.from_numpy(x).cuda().requires_grad_(True)
    _loss_t(enc, xt, kt, yt, wt).backward()
    gx = xt.grad.to(torch.complex128)
    h = 1e-3
    rng = np.random.default_rng(3)
    for _ in range(2):
        d = torch.from_numpy(crandn((N, N, N), int(rng.integers(1 << 30)))).cuda()
        with torch.no_grad():
            lp = _loss_fd(enc, xt + h * d, kt, yt, wt)
            lm = _loss_fd(enc, xt - h * d, kt, yt, wt)
        fd = (lp - lm) / (2 * h)
        an = float(torch.sum(torch.conj(gx) * d.to(torch.complex128)).real)
        # the loss is quadratic in x, so the central difference itself is exact; what
        # remains is the float32 rounding of the two operator outputs, about
        # sqrt(C M) * w |r| * 1e-7 |y| / (2 h), i.e. a few 1e-3 of the derivative
        assert abs(fd - an) < 1e-2 * abs(an), (fd, an)
    enc.close()


@pytest.mark.gpu
def test_traj_gradient_vs_fp64_reference():
    import torch
    enc, k, x, smaps, y, w = _dev_setup("exact_kb")
    xt = torch.from_numpy(x).cuda()
    kt = torch.from_numpy(k).cuda().requires_grad_(True)
    yt, wt = torch.from_numpy(y).cuda(), torch.from_numpy(w).cuda()
    _loss_t(enc, xt, kt, yt, wt).backward()
    gk = kt.grad.cpu().numpy().astype(np.float64)
    o, m = _ref_model()
    ym, dy = m.apply(_grids(o, x, smaps), k.astype(np.float64))
    g_ref = np.real(np.einsum("cm,cma->ma", np.conj(w * (ym - y)), dy))
    assert rel(gk, g_ref) < 1e-4
    enc.close()


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["table", "exact_kb"])
def test_traj_gradient_finite_differences(mode):
    import torch
    enc, k, x, smaps, y, w = _dev_setup(mode)
    xt = torch.from_numpy(x).cuda()
    yt, wt = torch.from_numpy(y).cuda(), torch.from_numpy(w).cuda()
    kt = torch.from_numpy(k).cuda().requires_grad_(True)
    _loss_t(enc, xt, kt, yt, wt).backward()
    gk = kt.grad.double()
    h = 1e-3
    rng = np.random.default_rng(7)
    # the table kernel is piecewise linear at 1/600 spacing: its finite difference over
    # h = 1e-3 averages neighbouring slopes
    tol = 5e-3 if mode == "exact_kb" else 2e-2
    for _ in range(3):
        dn, nkeep = _stable_direction(rng.standard_normal(k.shape).astype(np.float32), k, h)
        assert nkeep > 0.8 * k.shape[0]
        d = torch.from_numpy(dn.astype(np.float32)).cuda()
        with torch.no_grad():
            lp = _loss_fd(enc, xt, (kt + h * d).contiguous(), yt, wt)
            lm = _loss_fd(enc, xt, (kt - h * d).contiguous(), yt, wt)
        fd = (lp - lm) / (2 * h)
        an = float(torch.sum(gk * d.double()))
        assert abs(fd - an) < tol * abs(an), (fd, an)
    enc.close()


@pytest.mark.gpu
def test_rebinding_between_forward_and_backward():
    """A second trajectory bound between forward and backward is replaced by the
    saved one before the gradient runs."""
    import torch
    enc, k, x, smaps, y, w = _dev_setup("exact_kb")
    xt = torch.from_numpy(x).cuda()
    yt, wt = torch.from_numpy(y).cuda(), torch.from_numpy(w).cuda()
    k1 = torch.from_numpy(k).cuda().requires_grad_(True)
    L1 = _loss_t(enc, xt, k1, yt, wt)
    with torch.no_grad():
        enc(xt, (k1 * 0.9).contiguous())
    L1.backward()
    g1 = k1.grad.clone()
    k2 = torch.from_numpy(k).cuda().requires_grad_(True)
    _loss_t(enc, xt, k2, yt, wt).backward()
    assert torch.allclose(g1, k2.grad, rtol=1e-6, atol=0)
    enc.close()


@pytest.mark.gpu
def test_fixed_trajectory_binding():
    """set_traj once, then calls without k: same values and image gradient as
    passing k on every call."""
    import torch
    enc, k, x, smaps, y, w = _dev_setup("table")
    kt = torch.from_numpy(k).cuda()
    xt = torch.from_numpy(x).cuda().requires_grad_(True)
    yt = torch.from_numpy(y).cuda()
    y1 = enc(xt, kt)
    g1 = torch.autograd.grad(torch.sum(torch.abs(y1 - yt) ** 2), xt)[0]
    enc.set_traj(kt)
    y2 = enc(xt)
    g2 = torch.autograd.grad(torch.sum(torch.abs(y2 - yt) ** 2), xt)[0]
    assert torch.equal(y1, y2)
    assert torch.equal(g1, g2)
    enc.close()
