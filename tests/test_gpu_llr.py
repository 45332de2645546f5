"""Locally low-rank threshold on the GPU (mrn_llr, and inside FISTA) against
the oracle's restatement of LOWRANKCOIL (clear.cpp:233-556).

Both sides form the block matrices in double from the float image and store
float results; the GPU diagonalises A^H A by Jacobi where the oracle runs a
one-sided Jacobi SVD, so results agree to FP32 rounding of the stored image:
relative L2 <= 1e-5 per call, smax to 1e-6.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as orc
from mri_recon_amd import GridFFT, MrnError, llr_params, llr_threshold, synth
from tests.helpers import crandn, rel, to_dev, traj_dev

pytestmark = pytest.mark.gpu


def lowrank_frames(Nt, Ne, shape, rank=2, noise=0.1, seed=0):
    r = np.random.default_rng(seed)
    c = lambda *s: (r.standard_normal(s) + 1j * r.standard_normal(s)) / np.sqrt(2)
    base = np.einsum("tr,rv->tv", c(Nt * Ne, rank), c(rank, int(np.prod(shape))))
    X = base + noise * c(Nt * Ne, int(np.prod(shape)))
    return X.reshape((Nt, Ne) + tuple(shape)).astype(np.complex64)


CASES = [
    # Nt, Ne, (Nz, Ny, Nx), dim, block, block_iter, alpha
    (8, 1, (16, 16, 16), 0, (4, 4, 4), 4, 0.05),
    (7, 1, (12, 8, 20), 0, (4, 4, 4), 4, 0.1),      # odd frame count; sizes not multiples of 4 in z, x
    (5, 3, (8, 8, 8), 1, (4, 4, 4), 2, 0.08),       # the encode dimension
    (4, 2, (10, 9, 8), 0, (3, 5, 2), 3, 0.05),      # other block shapes
    (32, 1, (8, 8, 8), 0, (2, 2, 2), 1, 0.02),      # 32 frames, one unshifted pass
    (3, 1, (4, 4, 4), 0, (8, 8, 8), 2, 0.1),        # blocks clamped to the volume
    # block rows stream through LDS in chunks of 32 or 64: more rows than fit at once
    (4, 1, (16, 16, 16), 0, (16, 16, 16), 2, 0.1),  # one block of 4096 rows
    (6, 2, (8, 8, 16), 0, (8, 8, 8), 2, 0.05),      # 2 encodes x 512 voxels per block row set
]


@pytest.mark.parametrize("Nt,Ne,shape,dim,block,block_iter,alpha", CASES)
def test_llr_matches_oracle(Nt, Ne, shape, dim, block, block_iter, alpha):
    X = lowrank_frames(Nt, Ne, shape, seed=Nt + Ne)
    kw = dict(block=block, block_iter=block_iter, seed=11, amp=4.0, t2=6.0)
    if dim == 0:
        kw["alpha_time"] = alpha
    else:
        kw["alpha_encode"] = alpha
    Xd = to_dev(X)
    smax = llr_threshold(Xd, dim=dim, iteration=3, call=5, **kw)
    ref, rsmax = orc.llr(orc.llr_params(**kw), X, dim, 3, 5, "f32")
    e = rel(Xd.cpu().numpy(), ref)
    print("llr %s: rel %.2e smax %.6g / %.6g" % ((Nt, Ne, shape, dim, block), e, smax, rsmax))
    assert abs(smax - rsmax) <= 1e-6 * rsmax
    assert e <= 1e-5
    assert rel(ref, X) > 1e-3  # the step did something


def test_llr_update_only_and_huge_threshold():
    X = lowrank_frames(6, 1, (8, 8, 12), seed=2)
    Xd = to_dev(X)
    smax = llr_threshold(Xd, dim=0, iteration=0, alpha_time=0.0)
    assert np.array_equal(Xd.cpu().numpy(), X)  # alpha 0: update_threshold only
    assert abs(smax - orc.llr_update_threshold(orc.llr_params(), X, 0, 0, "f32")) <= 1e-6 * smax
    Xd = to_dev(X)
    llr_threshold(Xd, dim=0, iteration=0, alpha_time=1e6, block_iter=1)
    Y = Xd.cpu().numpy()
    assert np.all(Y[:, :, :, :, :12] == 0)  # every voxel is in a block here
    # an all-zero image: no nonzero block, smax 0, tau 0, the image stays 0
    Z = torch.zeros((3, 1, 8, 8, 8), dtype=torch.complex64, device="cuda")
    assert llr_threshold(Z, alpha_time=0.1) == 0.0 and torch.count_nonzero(Z).item() == 0


def test_llr_errors_and_flags():
    X = torch.zeros((40, 1, 8, 8, 8), dtype=torch.complex64, device="cuda")
    with pytest.raises(MrnError, match="32 frames"):
        llr_threshold(X, alpha_time=0.1)
    p = llr_params(args=["-clear_nx", "3", "-clear_ny", "5", "-clear_nz", "2", "-clear_block_iter", "6",
                         "-clear_alpha_time", "0.25", "-clear_amp", "2", "-clear_t2", "9", "-x"])
    assert (tuple(p.block), p.block_iter, p.amp, p.t2) == ((3, 5, 2), 6, 2.0, 9.0)
    assert p.alpha_time == pytest.approx(0.25)


def _oracle_fista_llr(o32, smaps, frames, ds, iters, step, p):
    """The FISTA loop of recon_lib.cpp:1925-2212 with the time low-rank step
    after each gradient step, composed from oracle calls frame by frame."""
    F = len(frames)
    N = smaps.shape[-1]
    X = np.zeros((F, N, N, N), np.complex64)
    Xold = np.zeros_like(X)
    st = np.float32(step)
    for it in range(iters):
        a = np.float32(1.0 + (it - 1.0) / (it + 1.0))
        b = np.float32(-(it - 1.0) / (it + 1.0))
        X, Xold = (a * X + b * Xold).astype(np.complex64), X
        for f in range(F):
            diff = o32.image_to_kspace(X[f], smaps, frames[f]) - ds[f]
            R = o32.kspace_to_image(diff.astype(np.complex64), smaps, frames[f])
            X[f] = X[f] - (R * st).astype(np.complex64)
        Y, _ = orc.llr(p, X[:, None], 0, it, it, "f32")
        X = Y[:, 0].copy()
    return X


def test_fista_with_llr_matches_oracle():
    N, C, F = 16, 2, 4
    frames = [[t.numpy() for t in tr] for tr in synth.gated_radial_frames(F, 80, 16, N, seed=7)]
    smaps = synth.coil_maps(C, N, seed=8).numpy()
    o32 = orc.OracleOp(N, overgrid=1.5, precision="f32")
    truth = lowrank_frames(F, 1, (N, N, N), rank=1, noise=0.2, seed=9)[:, 0]
    op = GridFFT(N, ncoils=C, nframes=F, overgrid=1.5)
    op.set_smaps(to_dev(smaps))
    ds = []
    for f in range(F):
        op.set_traj(*traj_dev(frames[f]), frame=f)
        d = o32.image_to_kspace(truth[f], smaps, frames[f])
        op.set_kdata(to_dev(d), frame=f)
        ds.append(d)
    step = 0.5 * op.power_method(10)
    kw = dict(alpha_time=0.05, block=(4, 4, 4), block_iter=2, seed=21)
    ref = _oracle_fista_llr(o32, smaps, frames, ds, 4, step, orc.llr_params(**kw))
    op.set_fista_llr(llr_params(**kw))
    outs = []
    for graph in (True, False):
        Xg, _ = op.fista(4, step, graph=graph)
        outs.append(Xg.cpu().numpy())
        e = rel(outs[-1], ref)
        print("FISTA+LLR graph=%s: %.2e" % (graph, e))
        assert e <= 1e-4
    assert np.array_equal(outs[0], outs[1])
    assert op.fista_llr_smax() > 0
    # off again: plain FISTA
    op.set_fista_llr(None)
    Xp, _ = op.fista(4, step)
    assert rel(Xp.cpu().numpy(), ref) > 1e-3


def test_llr_full_size_time_frames():
    """256^3 x 8 frames, where the oracle would take minutes: alpha 0 leaves X
    unchanged bit for bit, and a threshold removes a block's noise singular
    values (the image repeats a 64^3 rank-2 pattern plus noise)."""
    X = torch.from_numpy(lowrank_frames(8, 1, (64, 64, 64), rank=2, noise=0.05, seed=4)).cuda()
    X = X.repeat(1, 1, 4, 4, 4).contiguous()  # 256^3
    X0 = X.clone()
    llr_threshold(X, alpha_time=0.0)
    assert torch.equal(X, X0)
    llr_threshold(X, alpha_time=0.2, block_iter=1)
    blk = X[:, 0, 100:104, 40:44, 8:12].reshape(8, 64).T.cpu().numpy()
    blk0 = X0[:, 0, 100:104, 40:44, 8:12].reshape(8, 64).T.cpu().numpy()
    s, s0 = np.linalg.svd(blk, compute_uv=False), np.linalg.svd(blk0, compute_uv=False)
    assert s[2] < 1e-3 * s0[0] and s0[2] > 1e-2 * s0[0]
