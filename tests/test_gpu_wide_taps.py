"""Samples exactly on the kernel window's edge have one more tap per axis
than the rest (taps ceil(dk - dwin) .. floor(dk + dwin) with dk - dwin an
integer, gridFFT.cpp:1061-1094).  The tile scatter sizes each tile's bin
reach for the usual count and scatters those samples separately; the gather
takes them tap by tap.  Trajectories with many such samples (half-integer dk
for dwin 2.5, integer dk for dwin 2.0, at os 2 where dk = 2 k + S/2 is exact)
against the FP32 oracle, for every accumulate form."""
import numpy as np
import pytest

from oracle import oracle as orc
from mri_recon_amd import GridFFT
from tests.helpers import crandn, rel, to_dev, traj_dev
from mri_recon_amd import synth

pytestmark = pytest.mark.gpu


def _edge_traj(N, dwin, n_edge, seed):
    rng = np.random.default_rng(seed)
    kx, ky, kz, kw = [t.numpy() for t in synth.radial_traj(300, 24, N)]
    # k = j / 2 + off: dk = 2 k + S/2 = j + 2 off + S/2, on the window edge for off = dwin % 1 / 2
    off = (dwin % 1.0) / 2.0
    j = rng.integers(-N + 4, N - 4, size=(n_edge, 3))
    e = (j / 2.0 + off).astype(np.float32)
    axes = rng.integers(0, 3, size=n_edge)  # one axis on the edge, the others anywhere
    other = rng.uniform(-N / 2 + 2, N / 2 - 2, size=(n_edge, 3)).astype(np.float32)
    pts = np.where(np.arange(3)[None, :] == axes[:, None], e, other)
    pts[: n_edge // 4] = e[: n_edge // 4]  # all three axes on the edge
    kx = np.concatenate([kx, pts[:, 0]]).astype(np.float32)
    ky = np.concatenate([ky, pts[:, 1]]).astype(np.float32)
    kz = np.concatenate([kz, pts[:, 2]]).astype(np.float32)
    kw = np.concatenate([kw, rng.uniform(0.5, 2.0, n_edge)]).astype(np.float32)
    return kx, ky, kz, kw


@pytest.mark.parametrize("dwin", [2.5, 2.0])
@pytest.mark.parametrize("C,kw", [(16, {}), (8, {}), (4, {}), (16, {"scatter_valu": True}),
                                  (16, {"bin_edge": 2})])
def test_window_edge_samples(dwin, C, kw):
    N = 32
    traj = _edge_traj(N, dwin, 500, seed=int(dwin * 10) + C)
    smaps = synth.coil_maps(C, N, seed=5).numpy()
    X = crandn((N, N, N), 6)
    d = crandn((C, traj[0].size), 7)
    op = GridFFT(N, ncoils=C, overgrid=2.0, dwin=dwin, **kw)
    op.set_traj(*traj_dev(traj))
    op.set_smaps(to_dev(smaps))
    o32 = orc.OracleOp(N, overgrid=2.0, dwin=dwin, precision="f32")
    h = op.forward(to_dev(d), accumulate=False).cpu().numpy()
    a = op.backward(to_dev(X)).cpu().numpy()
    eh = rel(h, o32.kspace_to_image(d, smaps, traj))
    ea = rel(a, o32.image_to_kspace(X, smaps, traj))
    print("dwin %.1f C=%d %s: A^H %.2e A %.2e" % (dwin, C, kw, eh, ea))
    assert eh <= 1e-5 and ea <= 1e-5
