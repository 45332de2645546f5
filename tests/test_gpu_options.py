"""The operator's other settings on the GPU against the oracle: the triangle,
sinc and polynomial kernels (gridFFT.cpp:251-383), a non-cubic image, a
Cartesian kz axis (stack of stars: grid_in_z 0, gridFFT.cpp:484-486 and the
single-tap branch of chop_grid_*), the same without a z transform (fft_in_z
0, gridFFT.cpp:488-498; the plain form, which the batched one refers such
geometries to), and every coil block size.  Bar: A and A^H within 1e-4 of
the FP64 oracle, and within 1e-5 of the FP32 one -- or, where the FP32
oracle itself is further than that from FP64 (the sinc kernel: its
deapodisation window falls towards 0 at the edges of the image and
magnifies each FP32 path's own rounding), within 1.5 times its distance.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from mri_recon_amd import GridFFT, MrnError, synth
from tests.helpers import crandn, rel, to_dev, traj_dev

pytestmark = pytest.mark.gpu


def problem(N, C, seed, cartesian_z=False):
    n = (N, N, N) if np.isscalar(N) else N
    kx, ky, kz, kw = [t.numpy() for t in synth.radial_traj(300, 2 * max(n) // 2, min(n))]
    if cartesian_z:  # integer partition positions, one per sample
        rng = np.random.default_rng(seed)
        kz = rng.integers(-n[2] // 2, n[2] // 2, kx.size).astype(np.float32)
    X = crandn((n[2], n[1], n[0]), seed)
    smaps = np.stack([synth.coil_maps(C, max(n), seed=seed + 1).numpy()[c][:n[2], :n[1], :n[0]]
                      for c in range(C)])
    d = crandn((C, kx.size), seed + 2)
    return (kx, ky, kz, kw), X, np.ascontiguousarray(smaps), d


def check(op_kw, N, C, seed, cartesian_z=False, forms=("batched", "plain")):
    traj, X, smaps, d = problem(N, C, seed, cartesian_z)
    okw = {k: v for k, v in op_kw.items() if k in ("overgrid", "dwin", "kernel", "grid_in", "fft_in")}
    o32 = orc.OracleOp(N, precision="f32", **okw)
    o64 = orc.OracleOp(N, precision="f64", **okw)
    A32 = o32.image_to_kspace(X, smaps, traj)
    A64 = o64.image_to_kspace(X.astype(np.complex128), smaps.astype(np.complex128), traj)
    H32 = o32.kspace_to_image(d, smaps, traj)
    H64 = o64.kspace_to_image(d.astype(np.complex128), smaps.astype(np.complex128), traj)
    assert np.linalg.norm(A64) > 0 and np.linalg.norm(H64) > 0
    bar_a = max(1e-5, 1.5 * rel(A32, A64))
    bar_h = max(1e-5, 1.5 * rel(H32, H64))
    for form in forms:
        op = GridFFT(N, ncoils=C, form=form, **op_kw)
        op.set_traj(*traj_dev(traj))
        op.set_smaps(to_dev(smaps))
        a = op.backward(to_dev(X)).cpu().numpy()
        h = op.forward(to_dev(d), accumulate=False).cpu().numpy()
        errs = (rel(a, A32), rel(a, A64), rel(h, H32), rel(h, H64))
        print("%s %s: A %.2e/%.2e  AH %.2e/%.2e" % (form, op_kw, *errs))
        assert errs[0] <= bar_a and errs[1] <= 1e-4 and errs[2] <= bar_h and errs[3] <= 1e-4


@pytest.mark.parametrize("kernel,dwin,os_", [("triangle", 1.0, 1.5), ("triangle", 2.0, 2.0), ("sinc", -1.0, 1.5),
                                             ("poly", -1.0, 1.5), ("kaiser", 3.5, 1.25)])
def test_kernels(kernel, dwin, os_):
    check(dict(kernel=kernel, dwin=dwin, overgrid=os_), 16, 2, 51)


def test_non_cubic_image():
    check(dict(overgrid=1.5, dwin=2.5), (24, 16, 20), 3, 52)


def test_cartesian_kz_stack_of_stars():
    check(dict(overgrid=1.5, dwin=2.5, grid_in=(1, 1, 0)), (16, 16, 12), 2, 53, cartesian_z=True)


def test_cartesian_kz_without_z_transform():
    kw = dict(overgrid=1.5, dwin=2.5, grid_in=(1, 1, 0), fft_in=(1, 1, 0))
    check(kw, (16, 16, 8), 2, 54, cartesian_z=True, forms=("plain",))
    with pytest.raises(MrnError, match="plain form"):
        GridFFT((16, 16, 8), ncoils=2, **kw)


@pytest.mark.parametrize("chunk", [1, 2, 4, 8])
def test_coil_blocks(chunk):
    check(dict(overgrid=1.25, dwin=2.5, coil_chunk=chunk), 16, 8, 55, forms=("batched",))


@pytest.mark.parametrize("N,C,os_", [(48, 16, 1.5), (64, 3, 1.5), (96, 2, 1.5), ((40, 48, 64), 5, 1.5),
                                     (64, 16, 2.5), (128, 2, 2.0), (160, 1, 2.0)])
def test_four_step_fft_vs_oracle_and_stockham(N, C, os_):
    """The four-step register FFT (the default for grid edges S = 16 P,
    64 <= S <= 512; here P = 5, 6, 9, 10, 16, 20, a non-cubic grid 64 x 80 x
    96, coil blocks of 1, 2, 4, 8 and 16) meets the oracle's bar for A and
    A^H (the module's: relative to the FP32 oracle's own distance from FP64
    where that is larger, as at overgrid 2.5), and agrees to FP32 rounding
    with the LDS Stockham passes (fft_stockham=True), which sum in another
    order."""
    traj, X, smaps, d = problem(N, C, 61)
    o32 = orc.OracleOp(N, overgrid=os_, dwin=2.5, precision="f32")
    assert all(s >= 64 and s % 16 == 0 for s in o32.S)
    out = []
    for stockham in (False, True):
        op = GridFFT(N, ncoils=C, overgrid=os_, dwin=2.5, fft_stockham=stockham)
        op.set_traj(*traj_dev(traj))
        op.set_smaps(to_dev(smaps))
        out.append((op.backward(to_dev(X)).cpu().numpy(), op.forward(to_dev(d), accumulate=False).cpu().numpy()))
        del op
    o64 = orc.OracleOp(N, overgrid=os_, dwin=2.5, precision="f64")
    A32, H32 = o32.image_to_kspace(X, smaps, traj), o32.kspace_to_image(d, smaps, traj)
    A64 = o64.image_to_kspace(X.astype(np.complex128), smaps.astype(np.complex128), traj)
    H64 = o64.kspace_to_image(d.astype(np.complex128), smaps.astype(np.complex128), traj)
    bar_a, bar_h = max(1e-5, 1.5 * rel(A32, A64)), max(1e-5, 1.5 * rel(H32, H64))
    a, h = out[0]
    errs = (rel(a, A32), rel(a, A64), rel(h, H32), rel(h, H64), rel(a, out[1][0]), rel(h, out[1][1]))
    print("four-step S=%s C=%d: A %.2e/%.2e AH %.2e/%.2e  vs Stockham %.2e %.2e" % (tuple(o32.S), C, *errs))
    assert errs[0] <= bar_a and errs[1] <= 1e-4 and errs[2] <= bar_h and errs[3] <= 1e-4
    assert errs[4] <= 1e-6 and errs[5] <= 1e-6


def test_four_step_split_transpose_vs_stockham():
    """Grid edges S >= 400 (here 480 and 432; 96 on z) move the four-step
    line passes' transpose through LDS as floats, real parts then imaginary
    parts, so that three workgroups fit a CU.  A and A^H against the LDS
    Stockham passes (pinned to the oracle elsewhere), which sum in another
    order: FP32 rounding, <= 1e-6; 8-coil blocks, where the x and y passes are
    four-step."""
    N, C = (320, 288, 64), 8
    _, X, smaps, _ = problem(N, C, 63)
    # spokes over the whole x-y extent (those leaving the 96-cell z extent are skipped)
    traj = [t.numpy() for t in synth.radial_traj(400, 160, 320)]
    d = crandn((C, traj[0].size), 64)
    out = []
    for stockham in (False, True):
        op = GridFFT(N, ncoils=C, overgrid=1.5, dwin=2.5, fft_stockham=stockham)
        assert tuple(op.S) == (480, 432, 96)
        op.set_traj(*traj_dev(traj))
        op.set_smaps(to_dev(smaps))
        out.append((op.backward(to_dev(X)).cpu().numpy(), op.forward(to_dev(d), accumulate=False).cpu().numpy()))
        op.close()
    ea, eh = rel(out[0][0], out[1][0]), rel(out[0][1], out[1][1])
    print("split transpose S=(480, 432, 96): A %.2e A^H %.2e vs Stockham" % (ea, eh))
    assert ea <= 1e-6 and eh <= 1e-6


@pytest.mark.parametrize("C", [16, 8, 3])
def test_gather_coils_per_lane_same_bits(C):
    """The direct gather with 1, 2, 4 or 8 coils per lane (default 4; capped
    at the coil block) sums every coil's taps in the same order: A and the
    weighted gather inside E'E give the same bits for every setting."""
    traj, X, smaps, d = problem(32, C, 71)
    outs = []
    for nc in (1, 2, 4, 8):
        op = GridFFT(32, ncoils=C, overgrid=1.5, dwin=2.5, gather_coils=nc)
        op.set_traj(*traj_dev(traj))
        op.set_smaps(to_dev(smaps))
        outs.append((op.backward(to_dev(X)).cpu().numpy(), op.normal(to_dev(X)).cpu().numpy()))
    for a, n in outs[1:]:
        assert np.array_equal(a, outs[0][0]) and np.array_equal(n, outs[0][1])
    o32 = orc.OracleOp(32, overgrid=1.5, dwin=2.5, precision="f32")
    assert rel(outs[2][0], o32.image_to_kspace(X, smaps, traj)) <= 1e-5
    with pytest.raises(MrnError, match="gather_coils"):
        GridFFT(32, ncoils=C, gather_coils=3)


def test_xcd_block_order_same_bits():
    """Workgroups go to the 8 XCDs in runs of consecutive blocks: a
    permutation of the block indices for any run length, with the blocks
    that do not fill a whole group of 8 runs left in hardware order.  The
    gather and the FFT passes write each output from one lane, so A gives the
    same bits for every run length; the scatter's atomics add in another
    order, so A^H agrees to FP32 rounding.  Here the gather has 225 blocks per
    coil block: runs of 1, 3 and 7 leave partial groups at the end, runs of
    64 leave every block in hardware order."""
    N, C = 48, 16
    traj, X, smaps, d = problem(N, C, 83)
    outs = []
    for run in (1, 3, 7, 64):
        op = GridFFT(N, ncoils=C, overgrid=1.5, dwin=2.5, xcd_run=run, xcd_run_gather=run,
                     xcd_run_fft=run)
        op.set_traj(*traj_dev(traj))
        op.set_smaps(to_dev(smaps))
        outs.append((op.backward(to_dev(X)).cpu().numpy(), op.forward(to_dev(d)).cpu().numpy()))
    for a, h in outs[1:]:
        assert np.array_equal(a, outs[0][0])
        assert rel(h, outs[0][1]) <= 1e-6
    o32 = orc.OracleOp(N, overgrid=1.5, dwin=2.5, precision="f32")
    assert rel(outs[0][0], o32.image_to_kspace(X, smaps, traj)) <= 1e-5
    assert rel(outs[0][1], o32.kspace_to_image(d, smaps, traj)) <= 1e-5


@pytest.mark.parametrize("C,kw", [(16, {}), (8, {}), (4, {}), (16, {"max_visits": 64}),
                                  (16, {"scatter_valu": True})])
def test_xcd_slab_scatter(C, kw):
    """xcd_run = -1: the scatter's work list is cut into 8 contiguous ranges
    of equal cost and workgroup b takes its items from range b % 8.  Every
    item must run exactly once: A^H as with the hardware order (split tiles'
    atomics aside, the same bits) and as the oracle, for the matrix-core and
    vector-ALU forms, split tiles, accumulation and three fused frames of
    different sizes."""
    N = 32
    traj, X, smaps, d = problem(N, C, 91)
    o32 = orc.OracleOp(N, overgrid=1.5, dwin=2.5, precision="f32")
    outs = []
    for run in (1, -1):
        op = GridFFT(N, ncoils=C, overgrid=1.5, dwin=2.5, xcd_run=run, **kw)
        op.set_traj(*traj_dev(traj))
        op.set_smaps(to_dev(smaps))
        h = op.forward(to_dev(d)).cpu().numpy()
        h2 = op.forward(to_dev(d), X=op.forward(to_dev(d))).cpu().numpy()  # accumulating form
        outs.append((h, h2))
    print("C=%d %s: slabs vs hardware order %.1e / %.1e, vs oracle %.1e"
          % (C, kw, rel(outs[1][0], outs[0][0]), rel(outs[1][1], outs[0][1]),
             rel(outs[1][0], o32.kspace_to_image(d, smaps, traj))))
    for i in range(2):
        if "max_visits" in kw:
            assert rel(outs[1][i], outs[0][i]) <= 1e-6
        else:
            assert np.array_equal(outs[1][i], outs[0][i])
    assert rel(outs[1][0], o32.kspace_to_image(d, smaps, traj)) <= 1e-5
    # fused frames of uneven length: each frame's own ranges
    frames = [[t.numpy() for t in tr] for tr in synth.gated_radial_frames(3, 90, N, N, seed=5)]
    frames[2] = [t[: t.size // 4].copy() for t in frames[2]]
    res = []
    for run in (1, -1):
        op = GridFFT(N, ncoils=C, nframes=3, overgrid=1.5, dwin=2.5, fuse_frames=1, xcd_run=run, **kw)
        op.set_smaps(to_dev(smaps))
        for f, tr in enumerate(frames):
            op.set_traj(*traj_dev(tr), frame=f)
            op.set_kdata(to_dev(o32.image_to_kspace(X, smaps, tr)), frame=f)
        res.append(op.cg(3)[0].cpu().numpy())
    print("fused frames CG: slabs vs hardware order %.1e" % rel(res[1], res[0]))
    assert rel(res[1], res[0]) <= 1e-6


def test_kspace_support_pruning_across_frames():
    """With 16-coil blocks and a four-step grid the forward x pass stores, and
    the inverse x pass loads, only the x cells that some sample of the frame
    reaches, and the scatter skips sample-free tiles outside that support.
    Two frames with different supports (spokes to a quarter of k-max, and to
    k-max), called in turn so that cells left by one frame would show in the
    other: A, A^H and E'E of each frame match the FP32 oracle, and match the
    unpruned LDS Stockham passes (fft_stockham) to FP32 rounding, on every
    call."""
    N, C, os_ = 48, 16, 1.5
    full = [t.numpy() for t in synth.radial_traj(200, N, N)]
    short = [t.numpy().copy() for t in synth.radial_traj(200, N, N)]
    for a in range(3):
        short[a] *= 0.25
    trajs = [tuple(short), tuple(full)]
    X = crandn((N, N, N), 81)
    smaps = synth.coil_maps(C, N, seed=82).numpy()
    ds = [crandn((C, t[0].size), 83 + f) for f, t in enumerate(trajs)]
    o32 = orc.OracleOp(N, overgrid=os_, dwin=2.5, precision="f32")
    assert all(s % 16 == 0 and s >= 64 for s in o32.S)
    ref = [(o32.image_to_kspace(X, smaps, t), o32.kspace_to_image(d, smaps, t), o32.normal(X, smaps, t))
           for t, d in zip(trajs, ds)]
    ops = []
    for stockham in (False, True):
        op = GridFFT(N, ncoils=C, nframes=2, overgrid=os_, dwin=2.5, fft_stockham=stockham)
        for f, t in enumerate(trajs):
            op.set_traj(*traj_dev(t), frame=f)
        op.set_smaps(to_dev(smaps))
        ops.append(op)
    Xd = to_dev(X)
    for f in (1, 0, 1, 0):
        got = []
        for op in ops:
            got.append((op.backward(Xd, frame=f).cpu().numpy(),
                        op.forward(to_dev(ds[f]), accumulate=False, frame=f).cpu().numpy(),
                        op.normal(Xd, frame=f).cpu().numpy()))
        for k in range(3):
            e_ref, e_unpruned = rel(got[0][k], ref[f][k]), rel(got[0][k], got[1][k])
            print("frame %d output %d: vs oracle %.2e, vs unpruned %.2e" % (f, k, e_ref, e_unpruned))
            assert e_ref <= 1e-5 and e_unpruned <= 1e-6


def test_kspace_support_pruning_cartesian_kz_and_empty_frame():
    """Support pruning with one tap in z (stack of stars, grid 80 x 80 x 32,
    16-coil blocks) against the oracle, and a frame whose samples all have
    kw = 0 (no valid sample: empty support, every x / y row skipped, no
    scatter tile) between calls on the first frame: its A is zero, its A^H
    and E'E are zero, and the first frame's results do not change."""
    N, C = (48, 48, 32), 16
    kw_ = dict(overgrid=1.5, dwin=2.5, grid_in=(1, 1, 0))
    traj, X, smaps, d = problem(N, C, 91, cartesian_z=True)
    o32 = orc.OracleOp(N, precision="f32", **kw_)
    o64 = orc.OracleOp(N, precision="f64", **kw_)
    assert tuple(o32.S) == (80, 80, 32)
    A32, H32 = o32.image_to_kspace(X, smaps, traj), o32.kspace_to_image(d, smaps, traj)
    A64 = o64.image_to_kspace(X.astype(np.complex128), smaps.astype(np.complex128), traj)
    H64 = o64.kspace_to_image(d.astype(np.complex128), smaps.astype(np.complex128), traj)
    bar_a, bar_h = max(1e-5, 1.5 * rel(A32, A64)), max(1e-5, 1.5 * rel(H32, H64))
    op = GridFFT(N, ncoils=C, nframes=2, **kw_)
    op.set_traj(*traj_dev(traj), frame=0)
    dead = (traj[0], traj[1], traj[2], np.zeros_like(traj[3]))
    op.set_traj(*traj_dev(dead), frame=1)
    op.set_smaps(to_dev(smaps))
    Xd, dd = to_dev(X), to_dev(d)
    first = None
    for rep in range(2):
        a = op.backward(Xd, frame=0).cpu().numpy()
        h = op.forward(dd, accumulate=False, frame=0).cpu().numpy()
        n = op.normal(Xd, frame=0).cpu().numpy()
        print("cartesian kz pruned: A %.2e AH %.2e" % (rel(a, A32), rel(h, H32)))
        assert rel(a, A32) <= bar_a and rel(a, A64) <= 1e-4
        assert rel(h, H32) <= bar_h and rel(h, H64) <= 1e-4
        if first is None:
            first = (a, h, n)
        else:
            assert all(np.array_equal(u, v) for u, v in zip(first, (a, h, n)))
        assert not np.any(op.backward(Xd, frame=1).cpu().numpy())
        assert not np.any(op.forward(dd, accumulate=False, frame=1).cpu().numpy())
        assert not np.any(op.normal(Xd, frame=1).cpu().numpy())


def test_kspace_support_pruning_bin_edge_2_and_autotune():
    """The support rows depend on the bin edge (a sample binned at b reaches
    [2b, 2b + 1 + W] with 2-cell bins): with bin_edge 2 the pruned passes
    still meet the oracle's bar, and the autotuner, which rebuilds the bins
    (and the support) for each candidate, leaves every result unchanged."""
    N, C = 48, 16
    traj, X, smaps, d = problem(N, C, 101)
    o32 = orc.OracleOp(N, overgrid=1.5, dwin=2.5, precision="f32")
    A32, H32 = o32.image_to_kspace(X, smaps, traj), o32.kspace_to_image(d, smaps, traj)
    for edge in (2, 1):
        op = GridFFT(N, ncoils=C, overgrid=1.5, dwin=2.5, bin_edge=edge)
        op.set_traj(*traj_dev(traj))
        op.set_smaps(to_dev(smaps))
        a = op.backward(to_dev(X)).cpu().numpy()
        h = op.forward(to_dev(d), accumulate=False).cpu().numpy()
        print("bin edge %d: A %.2e AH %.2e" % (edge, rel(a, A32), rel(h, H32)))
        assert rel(a, A32) <= 1e-5 and rel(h, H32) <= 1e-5
    before = op.normal(to_dev(X)).cpu().numpy()
    op.autotune(cache=None)
    after = op.normal(to_dev(X)).cpu().numpy()
    assert rel(after, before) <= 1e-5
    assert rel(op.backward(to_dev(X)).cpu().numpy(), A32) <= 1e-5
