"""Gated frame selection on the GPU (mrn_pregate) against the oracle's
restatement of RECON::pregate_data and GATING::weight_data.  This is copy and
index work, so the bar is bit-exact: counts, trajectory, kw (one float
product, as the reference) and k-space."""
import numpy as np
import pytest
import torch

from oracle import oracle as orc
from mri_recon_amd import gate_params, gate_scale_times, pregate, synth

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(23)


def problem(R, xres, C):
    kx, ky, kz = [RNG.uniform(-30, 30, (R, xres)).astype(np.float32) for _ in range(3)]
    kw = RNG.uniform(0.5, 2, (R, xres)).astype(np.float32)
    kd = (RNG.standard_normal((C, R, xres)) + 1j * RNG.standard_normal((C, R, xres))).astype(np.complex64)
    return kx, ky, kz, kw, kd


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("vs,shape,kw_", [("none", "vipr", {}), ("hist", "vipr", {}), ("tornado", "vipr", {}),
                                          ("tornado", "radial", {"wdth_low": 2, "wdth_high": 5}),
                                          ("tornado", "flat", {}), ("none", "vipr", {"reset_dens": True})])
@pytest.mark.parametrize("with_resp", [False, True])
def test_pregate_bit_exact(vs, shape, kw_, with_resp):
    R, xres, C, T = 300, 37, 3, 7
    kx, ky, kz, kw, kd = problem(R, xres, C)
    times = RNG.uniform(0, 100, R)
    resp = RNG.uniform(0, 1, R).astype(np.float32) * (RNG.uniform(size=R) > 0.2) if with_resp else None
    p_o = orc.gate_params(T, vs=vs, shape=shape, kmax=45.0, **kw_)
    p_g = gate_params(T, vs=vs, shape=shape, kmax=45.0, **kw_)
    gt, _ = orc.gate_scale(p_o, times)
    ref = orc.pregate(p_o, gt, kx, ky, kz, kw, kd, resp=resp)
    n, frames = pregate(p_g, dev(gt), dev(kx), dev(ky), dev(kz), dev(kw), dev(kd),
                        resp=None if resp is None else dev(resp))
    assert np.array_equal(np.array(n), ref[0])
    base = 0
    for t in range(T):
        fx, fy, fz, fw, fd = [f.cpu().numpy() for f in frames[t]]
        sl = slice(base, base + n[t])
        for a, b in ((fx, ref[1][sl]), (fy, ref[2][sl]), (fz, ref[3][sl]), (fw, ref[4][sl])):
            assert np.array_equal(a, b)
        assert np.array_equal(fd.ravel(), ref[5][C * base:C * (base + n[t])])
        base += n[t]


def test_pregate_full_size_partition():
    """The bench trajectory (20000 readouts x 256 samples, 32 coils of k-space):
    without sharing every sample lands in exactly one of 20 frames and each
    frame's samples are its readouts' samples; under tornado sharing frames
    overlap and every count is at least the binned one."""
    kx, ky, kz, kw = [t.view(20000, 256) for t in synth.radial_traj(20000, 256, 256, device="cuda")]
    C = 32
    kd = torch.randn((C, 20000, 256), dtype=torch.complex64, device="cuda")
    times = np.linspace(0, 1, 20000) + RNG.uniform(0, 1e-6, 20000)
    gt, _ = gate_scale_times(gate_params(20, vs="none"), times)
    n, frames = pregate(gate_params(20, vs="none"), dev(gt), kx, ky, kz, kw, kd)
    assert sum(n) == 20000 * 256 and all(c % 256 == 0 and c > 0 for c in n)
    t7 = frames[7]
    rows = np.nonzero(gt == 7)[0]
    assert torch.equal(t7[0].view(-1, 256), kx[rows])
    assert torch.equal(t7[4][5].view(-1, 256), kd[5, rows])
    gtt, _ = gate_scale_times(gate_params(20, vs="tornado"), times)
    n2, _ = pregate(gate_params(20, vs="tornado", kmax=128.0), dev(gtt), kx, ky, kz, kw, kd)
    assert all(a >= b for a, b in zip(n2, n)) and sum(n2) > sum(n)
