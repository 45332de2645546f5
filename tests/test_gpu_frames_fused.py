"""Frames fused into one launch per stage (fuse_frames=1: the frames' grids
side by side, the FFT passes over frame x coil block, the gather and scatter
over a per-frame table on the device) against the same op frame by frame
(fuse_frames=0): CG, FISTA with the locally low-rank step, the Cauchy step and
the power method over frames of uneven sample counts, coil blocks of 4, 8 and
16 (24 coils: three blocks of 8), k-space support pruning on and off.  The
per-frame kernels are the same, so the results agree to FP32 summation order
of the split tiles' atomics (<= 1e-6); the fused FISTA+LLR also matches the
composed oracle loop."""
import numpy as np
import pytest

from oracle import oracle as orc
from mri_recon_amd import GridFFT, llr_params, synth
from tests.helpers import rel, to_dev, traj_dev
from tests.test_gpu_llr import _oracle_fista_llr, lowrank_frames

pytestmark = pytest.mark.gpu


def _ops(N, C, frames, data, smaps, **kw):
    ops = []
    for fuse in (1, 0):
        op = GridFFT(N, ncoils=C, nframes=len(frames), fuse_frames=fuse, **kw)
        op.set_smaps(to_dev(smaps))
        for f, tr in enumerate(frames):
            op.set_traj(*traj_dev(tr), frame=f)
            op.set_kdata(to_dev(data[f]), frame=f)
        ops.append(op)
    return ops


def _problem(N, C, F, spokes, seed):
    frames = [[t.numpy() for t in tr] for tr in synth.gated_radial_frames(F, spokes, N, N, seed=seed)]
    # one much shorter frame: the frames' work lists differ in length
    frames[1] = [t[: t.size // 3].copy() for t in frames[1]]
    smaps = synth.coil_maps(C, N, seed=seed + 1).numpy()
    o32 = orc.OracleOp(N, overgrid=1.5, precision="f32")
    truth = lowrank_frames(F, 1, (N, N, N), rank=1, noise=0.2, seed=seed + 2)[:, 0]
    data = [o32.image_to_kspace(truth[f], smaps, frames[f]) for f in range(F)]
    return frames, smaps, data, o32


@pytest.mark.parametrize("C,N,kw", [(16, 24, {}), (24, 24, {}), (4, 24, {}), (8, 32, {"max_visits": 64}),
                                    (16, 24, {"fft_stockham": True}), (16, 24, {"scatter_valu": True})])
def test_fused_frames_match_frame_by_frame(C, N, kw):
    F = 3
    frames, smaps, data, _ = _problem(N, C, F, 120, seed=31 + C)
    fused, single = _ops(N, C, frames, data, smaps, overgrid=1.5, **kw)
    # CG, graph-captured
    Xa, ta = fused.cg(5)
    Xb, tb = single.cg(5)
    e_cg = rel(Xa.cpu().numpy(), Xb.cpu().numpy())
    # power method and FISTA with the Cauchy step
    pa, pb = fused.power_method(6), single.power_method(6)
    for op in (fused, single):
        op.set_fista_step("cauchy", 2)
    Fa, ra = fused.fista(4, 0.0)
    Fb, rb = single.fista(4, 0.0)
    e_fi = rel(Fa.cpu().numpy(), Fb.cpu().numpy())
    print("C=%d N=%d %s: CG %.2e, power %.3e / %.3e, Cauchy FISTA %.2e" % (C, N, kw, e_cg, pa, pb, e_fi))
    # CG trace [sum|R|^2, Re a, Im a, sum|Rnew|^2]: Im(alpha) is rounding noise,
    # so it is compared on the scale of |alpha|
    sc = np.abs(tb.numpy()).max(axis=0)
    sc[2] = max(sc[1], sc[2])
    assert e_cg <= 1e-6 and np.allclose(ta.numpy(), tb.numpy(), rtol=1e-5, atol=1e-6 * sc)
    assert abs(pa - pb) <= 1e-6 * abs(pb)
    assert e_fi <= 1e-6 and np.allclose(ra.numpy(), rb.numpy(), rtol=1e-5)


def test_fused_fista_llr_matches_oracle():
    N, C, F = 16, 8, 4
    frames, smaps, data, o32 = _problem(N, C, F, 80, seed=7)
    fused, single = _ops(N, C, frames, data, smaps, overgrid=1.5)
    step = 0.5 * single.power_method(10)
    kw = dict(alpha_time=0.05, block=(4, 4, 4), block_iter=2, seed=21)
    ref = _oracle_fista_llr(o32, smaps, frames, data, 3, step, orc.llr_params(**kw))
    outs = []
    for op in (fused, single):
        op.set_fista_llr(llr_params(**kw))
        X, _ = op.fista(3, step, graph=True)
        outs.append(X.cpu().numpy())
    print("fused FISTA+LLR: vs frame by frame %.2e, vs oracle %.2e" % (rel(outs[0], outs[1]), rel(outs[0], ref)))
    assert rel(outs[0], outs[1]) <= 1e-6
    assert rel(outs[0], ref) <= 1e-4


def test_new_trajectory_after_a_fused_solve():
    """set_traj between solves: the per-frame table is rebuilt (a stale table
    would gather the old samples)."""
    N, C, F = 16, 16, 2
    frames, smaps, data, o32 = _problem(N, C, F, 80, seed=11)
    fused, single = _ops(N, C, frames, data, smaps, overgrid=1.5)
    fused.cg(2)
    frames2, _, data2, _ = _problem(N, C, F, 60, seed=12)
    for op in (fused, single):
        for f in range(F):
            op.set_traj(*traj_dev(frames2[f]), frame=f)
            op.set_kdata(to_dev(data2[f]), frame=f)
    Xa, _ = fused.cg(3)
    Xb, _ = single.cg(3)
    assert rel(Xa.cpu().numpy(), Xb.cpu().numpy()) <= 1e-6
