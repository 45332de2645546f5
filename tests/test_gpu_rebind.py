"""One op binding trajectories of different sizes in turn (large, small,
larger again; with and without window-edge samples) reuses its frame arrays
and bind scratch in place and grows them when a trajectory needs more.  Each
bind must give what a fresh op bound once gives: A with a residual, A^H, CG
on stored data, the stored data re-sorted for a new trajectory of the same
size, and A against the FP32 oracle."""
import numpy as np
import pytest

from oracle import oracle as orc
from mri_recon_amd import GridFFT, MrnError, synth
from tests.helpers import crandn, rel, to_dev, traj_dev
from tests.test_gpu_wide_taps import _edge_traj

pytestmark = pytest.mark.gpu


def _radial(nspokes, seed):
    kx, ky, kz, kw = [t.numpy() for t in synth.radial_traj(nspokes, 24, 32, first_spoke=100 * seed)]
    return kx, ky, kz, kw


@pytest.mark.parametrize("dwin", [2.5, 2.0])
@pytest.mark.parametrize("C", [8, 3])
def test_rebind_sizes_match_fresh_op(dwin, C):
    N = 32
    smaps = synth.coil_maps(C, N, seed=3).numpy()
    X = crandn((N, N, N), 4)
    o32 = orc.OracleOp(N, overgrid=2.0, dwin=dwin, precision="f32")
    seq = [_edge_traj(N, dwin, 600, seed=1), _radial(40, 2), _edge_traj(N, dwin, 1500, seed=3), _radial(90, 4)]
    reused = GridFFT(N, ncoils=C, overgrid=2.0, dwin=dwin)
    reused.set_smaps(to_dev(smaps))
    for i, traj in enumerate(seq):
        M = traj[0].size
        d = crandn((C, M), 10 + i)
        fresh = GridFFT(N, ncoils=C, overgrid=2.0, dwin=dwin)
        fresh.set_smaps(to_dev(smaps))
        for op in (reused, fresh):
            op.set_traj(*traj_dev(traj))
            op.set_kdata(to_dev(d))
        a = [op.backward(to_dev(X), sub=to_dev(d)).cpu().numpy() for op in (reused, fresh)]
        h = [op.forward(to_dev(d), accumulate=False).cpu().numpy() for op in (reused, fresh)]
        x = [op.cg(3)[0].cpu().numpy() for op in (reused, fresh)]
        eo = rel(reused.backward(to_dev(X)).cpu().numpy(), o32.image_to_kspace(X, smaps, traj))
        print("dwin %.1f C=%d bind %d (M=%d): A %.1e A^H %.1e CG %.1e, A vs oracle %.1e"
              % (dwin, C, i, M, rel(a[0], a[1]), rel(h[0], h[1]), rel(x[0], x[1]), eo))
        assert rel(a[0], a[1]) <= 1e-6 and rel(h[0], h[1]) <= 1e-6 and rel(x[0], x[1]) <= 1e-6
        assert eo <= 1e-5
        fresh.close()


def test_rebind_resorts_or_drops_stored_data():
    """A new trajectory of the same size re-sorts the stored data (CG equals a
    fresh op given the data); one of another size drops it (CG then asks for
    data instead of reading past the stored array)."""
    N, C = 32, 8
    smaps = synth.coil_maps(C, N, seed=5).numpy()
    t1 = _radial(60, 6)
    rng = np.random.default_rng(7)
    p = rng.permutation(t1[0].size)
    t2 = [np.ascontiguousarray(t[p]) for t in _radial(60, 8)]
    t2[0] = (t2[0] * 0.9).astype(np.float32)  # same M, other coordinates and order
    d = crandn((C, t1[0].size), 9)
    op = GridFFT(N, ncoils=C, overgrid=1.5)
    op.set_smaps(to_dev(smaps))
    op.set_traj(*traj_dev(t1))
    op.set_kdata(to_dev(d))
    op.cg(2)
    op.set_traj(*traj_dev(t2))  # stored d re-sorted for t2
    fresh = GridFFT(N, ncoils=C, overgrid=1.5)
    fresh.set_smaps(to_dev(smaps))
    fresh.set_traj(*traj_dev(t2))
    fresh.set_kdata(to_dev(d))
    e = rel(op.cg(3)[0].cpu().numpy(), fresh.cg(3)[0].cpu().numpy())
    print("re-sorted stored data: CG vs fresh op %.1e" % e)
    assert e <= 1e-6
    op.set_traj(*traj_dev(_radial(80, 10)))
    with pytest.raises(MrnError, match="k-space data"):
        op.cg(1)
