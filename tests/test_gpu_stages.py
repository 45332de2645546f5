"""The grid-level stages and the grid's contents between calls.

* mrn_chop_grid_forward / mrn_chop_grid_backward (gridFFT::chop_grid_forward /
  chop_grid_backward, gridFFT.cpp:1009-1322, called directly by
  RECON::kspace_residual, recon_lib.cpp:2244-2252) against the oracle's
  scatter and gather stages: the scatter adds into the grid's contents, the
  gather reads them.
* The grid may hold anything between calls: filled with NaN before A, A
  gives the same bits (every cell a gather or a later pass reads was written
  by the same call, or is masked to the frame's k-space support); filled with
  NaN before A^H, A^H equals the clean result.
"""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from mri_recon_amd import GridFFT
from tests.helpers import crandn, radial_problem, rel, to_dev, traj_dev

pytestmark = pytest.mark.gpu


def _op(N, C, os_=1.5, **kw):
    traj, X, smaps, d = radial_problem(N, C, 300 if N <= 24 else 600, N, seed=4)
    op = GridFFT(N, ncoils=C, overgrid=os_, **kw)
    op.set_traj(*traj_dev(traj))
    op.set_smaps(to_dev(smaps))
    return op, traj, X, smaps, d


@pytest.mark.parametrize("C,kw", [(4, {}), (16, {}), (16, {"scatter_valu": True}), (8, {"gather_tiles": True}),
                                  (16, {"gather_tiles": 2}), (1, {"form": "plain"})])
def test_chop_grid_stages_match_oracle(C, kw):
    N = 24
    op, traj, X, smaps, d = _op(N, C, **kw)
    o32 = orc.OracleOp(N, overgrid=1.5, precision="f32")
    op.grid_zero()
    dd = to_dev(d)
    op.chop_grid_forward(dd)
    g = op.grid().cpu().numpy()
    ref = np.stack([o32.stage_scatter(d[c], traj) for c in range(C)])
    e = rel(g, ref)
    # adds into the contents: a second call doubles the grid
    op.chop_grid_forward(dd)
    e2 = rel(op.grid().cpu().numpy(), 2 * ref)
    # gather of the current grid (2 x the scatter), with and without sub
    y = op.chop_grid_backward().cpu().numpy()
    yref = np.stack([o32.stage_gather(2 * ref[c], traj) for c in range(C)])
    sub = crandn(d.shape, 9)
    ys = op.chop_grid_backward(sub=to_dev(sub)).cpu().numpy()
    ysref = np.stack([o32.stage_gather(2 * ref[c], traj, sub=sub[c]) for c in range(C)])  # skipped samples: 0
    print("chop_grid C=%d %s: forward %.2e, twice %.2e, backward %.2e" % (C, kw, e, e2, rel(y, yref)))
    assert e <= 1e-5 and e2 <= 1e-5
    assert rel(y, yref) <= 1e-5 and rel(ys, ysref) <= 1e-5


@pytest.mark.parametrize("kw", [{}, {"gather_tiles": True}, {"gather_tiles": 2}, {"fft_stockham": True},
                                {"coil_chunk": 8}, {"coil_chunk": 4}])
def test_grid_contents_do_not_reach_results(kw):
    N, C = 32, 16
    op, traj, X, smaps, d = _op(N, C, **kw)
    Xd = to_dev(X)
    y0 = op.backward(Xd).clone()
    a0 = op.forward(to_dev(d), accumulate=False).clone()
    torch.cuda.synchronize()
    op.grid_raw().fill_(complex(math.nan, math.nan))
    torch.cuda.synchronize()
    y1 = op.backward(Xd)
    assert torch.equal(y0, y1), "A read grid cells it did not write"
    op.grid_raw().fill_(complex(math.nan, math.nan))
    torch.cuda.synchronize()
    a1 = op.forward(to_dev(d), accumulate=False)
    assert bool(torch.isfinite(a1).all())
    assert rel(a1.cpu().numpy(), a0.cpu().numpy()) <= 1e-6
    # and a second frame-less round trip after the NaN fill: E'E unchanged
    e0 = op.normal(Xd).clone()
    op.grid_raw().fill_(complex(math.inf, -math.inf))
    torch.cuda.synchronize()
    assert rel(op.normal(Xd).cpu().numpy(), e0.cpu().numpy()) <= 1e-6


@pytest.mark.parametrize("kw", [{}, {"bin_edge": 2}, {"coil_chunk": 8}])
def test_every_tap_lies_in_the_frames_support(kw):
    """The CPU check behind the grid-contents test: with the support ranges the
    passes use (copied from the device), every tap of every sample, computed as
    the gather computes it, lies in its row's support, and the support rows of
    a plane lie in that plane's y range."""
    from tests.test_gpu_configs import _taps_in_support
    N, C = 32, 16
    op, traj, X, smaps, d = _op(N, C, **kw)
    bad, n = _taps_in_support(op, traj, 2.5)
    assert n > 0 and bad == 0
    xr, yr = op.support()
    for z in range(op.S[2]):
        rows = np.nonzero(xr[z, :, 0] <= xr[z, :, 1])[0]
        if rows.size:
            assert yr[z, 0] <= rows.min() and rows.max() <= yr[z, 1]
