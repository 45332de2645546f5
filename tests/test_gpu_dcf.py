"""Pipe DCF on the GPU (mrn_dcf_calc) against the oracle's restatement of
RECON::dcf_calc (recon_lib.cpp:854-917) and DCFgridFFT (DCFgridFFT.cpp).

The GPU scatter adds with float atomics, so the order of the additions into a
grid cell differs from the oracle's sample order: results agree to FP32
rounding of those sums, carried through the ratio iteration.  Bars: relative
L2 <= 1e-5 after one iteration, <= 1e-4 after 20 (the default), against the
FP32 oracle; the FP64 oracle within 1e-4.  The gather, the update, the table
and scale_kw have no ordering freedom.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as orc
from mri_recon_amd import MrnError, dcf_calc, dcf_params, synth
from mri_recon_amd.dcf import dcf_grid_size
from tests.helpers import rel

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def radial(nspokes=300, npts=24, N=32):
    kx, ky, kz, _ = [t.numpy() for t in synth.radial_traj(nspokes, npts, N)]
    return kx, ky, kz


@pytest.mark.parametrize("iters,tol", [(1, 1e-5), (5, 3e-5), (20, 1e-4)])
def test_dcf_matches_oracle(iters, tol):
    kx, ky, kz = radial()
    kw, sums = dcf_calc(dev(kx), dev(ky), dev(kz), N=32, iters=iters, return_sums=True)
    p = orc.dcf_params(32, iters=iters)
    ref, rsums = orc.dcf_calc(p, kx, ky, kz, "f32")
    ref64, _ = orc.dcf_calc(p, kx, ky, kz, "f64")
    got = kw.cpu().numpy()
    e32, e64 = rel(got, ref), rel(got, ref64)
    print("dcf iters=%d: vs f32 %.2e vs f64 %.2e" % (iters, e32, e64))
    assert e32 <= tol and e64 <= 1e-4
    assert np.allclose(sums, rsums, rtol=1e-4)
    assert np.array_equal(got == 0, ref == 0)


@pytest.mark.parametrize("N,acc,kernel,dwin,os_", [((24, 16, 12), 1.0, "poly", 2.1, 2.1),
                                                  (32, 2.0, "poly", 2.1, 2.1),
                                                  (32, 1.0, "poly2", 1.5, 1.5),
                                                  (20, 1.5, "poly", 3.0, 2.5),
                                                  (16, 3.0, "poly", 4.0, 2.1)])
def test_dcf_settings_match_oracle(N, acc, kernel, dwin, os_):
    """Non-cubic sizes (the x extent follows Nz), kernel widening with |k|,
    the second polynomial, other widths and oversampling; the last case has
    rows of up to 10 taps (more than the scatter's 8 lanes per sample)."""
    n = N if isinstance(N, tuple) else (N, N, N)
    kx, ky, kz = radial(200, 16, min(n))
    kw = dcf_calc(dev(kx), dev(ky), dev(kz), N=n, iters=8, acc=acc, kernel=kernel, dwin=dwin, overgrid=os_)
    p = orc.dcf_params(n, iters=8, acc=acc, kernel=1 if kernel == "poly2" else 0, dwin=dwin, overgrid=os_)
    assert dcf_grid_size(dcf_params(n, dwin=dwin, overgrid=os_)) == orc.dcf_grid_size(p)
    ref, _ = orc.dcf_calc(p, kx, ky, kz, "f32")
    assert rel(kw.cpu().numpy(), ref) <= 5e-5


def test_dcf_single_sample_closed_form():
    t = orc.dcf_table(2.1, 0, "f64")
    want = 1.0 / (t[0] ** 2 + 2 * t[600] ** 2 + 2 * t[1200] ** 2) ** 3
    z = torch.zeros(1, dtype=torch.float32, device="cuda")
    kw = dcf_calc(z, z, z, N=32, iters=3).item()
    assert abs(kw / want - 1) < 2e-6


def test_dcf_duplicates_halve_and_edges():
    kx, ky, kz = radial(150, 20)
    # samples off the grid (beyond trunc(os*N+16)/2 cells) get weight 0
    kx = np.r_[kx, [400.0, -400.0]].astype(np.float32)
    ky = np.r_[ky, [0.0, 0.0]].astype(np.float32)
    kz = np.r_[kz, [0.0, 0.0]].astype(np.float32)
    a = dcf_calc(dev(kx), dev(ky), dev(kz), N=32, iters=6).cpu().numpy()
    b = dcf_calc(dev(np.r_[kx, kx]), dev(np.r_[ky, ky]), dev(np.r_[kz, kz]), N=32, iters=6).cpu().numpy()
    M = kx.size
    assert a[-1] == 0 and a[-2] == 0
    assert rel(b[:M], a / 2) < 1e-5 and rel(b[M:], a / 2) < 1e-5
    # no samples: nothing to do
    e = torch.zeros(0, dtype=torch.float32, device="cuda")
    assert dcf_calc(e, e, e, N=32).numel() == 0


def test_dcf_bad_arguments_raise():
    z = torch.zeros(4, dtype=torch.float32, device="cuda")
    with pytest.raises(MrnError, match="dcf_dwin"):
        dcf_calc(z, z, z, N=32, dwin=-1.0)
    with pytest.raises(MrnError, match="dcf_iter"):
        dcf_calc(z, z, z, N=32, iters=-2)
    with pytest.raises(MrnError, match="too large"):
        dcf_calc(z, z, z, N=32, dwin=40.0)


def test_dcf_parse_args():
    p = dcf_params(64, args=["-dcf_iter", "7", "-dcf_dwin", "3.0", "-dcf_overgrid", "1.8", "-dcf_acc", "2",
                             "-dcf_scale", "0.5", "-other"])
    assert (p.iters, p.dwin, p.overgrid, p.acc, p.scale) == (7, 3.0, np.float32(1.8), 2.0, 0.5)


def test_dcf_full_size_duplicate_property():
    """The 256^3 radial workload (5.12 M samples), where the oracle would take
    minutes: duplicating every sample doubles A^T A, so every weight halves."""
    kx, ky, kz, _ = synth.radial_traj(20000, 256, 256, device="cuda")
    a = dcf_calc(kx, ky, kz, N=256, iters=4)
    b = dcf_calc(torch.cat([kx, kx]), torch.cat([ky, ky]), torch.cat([kz, kz]), N=256, iters=4)
    M = kx.numel()
    assert torch.count_nonzero(a).item() > 0.99 * M
    assert rel(b[:M].cpu().numpy(), a.cpu().numpy() / 2) < 1e-4
