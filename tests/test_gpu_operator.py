"""Parity of the HIP operator with the oracle on seeded small problems.

Bar (FP32 arithmetic, atomics reorder sums): relative L2 error <= 1e-5 against
the FP32 oracle and <= 1e-4 against the FP64 oracle (the same discretisation
evaluated in double).  Every test calls the library through its C ABI.
"""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from mri_recon_amd import GridFFT, MrnError
from tests.helpers import crandn, radial_problem, rel, to_dev, traj_dev

pytestmark = pytest.mark.gpu

TOL32, TOL64 = 1e-5, 1e-4

CASES = [
    # N, C, overgrid, dwin, nspokes, npts
    (16, 2, 1.5, 2.5, 200, 16),
    (32, 4, 1.25, 2.5, 400, 32),
    (24, 3, 1.5, 2.0, 300, 24),
    (32, 8, 1.5, 2.0, 500, 32),
]


def make(case, form="batched", seed=0, **kw):
    N, C, os_, dwin, ns, npts = case
    traj, X, smaps, d = radial_problem(N, C, ns, npts, seed=seed)
    op = GridFFT(N, ncoils=C, overgrid=os_, dwin=dwin, form=form, **kw)
    op.set_traj(*traj_dev(traj))
    op.set_smaps(to_dev(smaps))
    o32 = orc.OracleOp(N, overgrid=os_, dwin=dwin, precision="f32", **{k: v for k, v in kw.items() if k == "k_rad"})
    o64 = orc.OracleOp(N, overgrid=os_, dwin=dwin, precision="f64", **{k: v for k, v in kw.items() if k == "k_rad"})
    return op, o32, o64, traj, X, smaps, d


@pytest.mark.parametrize("case", CASES[:3])
def test_tables_windows_geometry_match_oracle(case):
    N, C, os_, dwin, _, _ = case
    op = GridFFT(N, ncoils=C, overgrid=os_, dwin=dwin)
    o32 = orc.OracleOp(N, overgrid=os_, dwin=dwin, precision="f32")
    g = o32.geometry()
    assert op.S == tuple(g["S"]) and op.og == tuple(g["og"])
    assert abs(op.beta[0] - g["beta"][0]) < 1e-6 * g["beta"][0]
    for a in range(3):
        t = op.table(a).numpy()
        assert np.max(np.abs(t - o32.table(a))) <= 2e-7 * np.max(np.abs(t))
        w = op.window(a).numpy()
        assert np.max(np.abs(w - o32.window(a))) <= 2e-7


@pytest.mark.parametrize("form", ["batched", "plain"])
@pytest.mark.parametrize("case", CASES)
def test_image_to_kspace(case, form):
    op, o32, o64, traj, X, smaps, d = make(case, form)
    y = op.backward(to_dev(X)).cpu().numpy()
    y32 = o32.image_to_kspace(X, smaps, traj)
    y64 = o64.image_to_kspace(X.astype(np.complex128), smaps.astype(np.complex128), traj)
    e32, e64 = rel(y, y32), rel(y, y64)
    print("A %s %s: vs f32 %.2e vs f64 %.2e" % (form, case, e32, e64))
    assert e32 <= TOL32 and e64 <= TOL64


@pytest.mark.parametrize("form", ["batched", "plain"])
@pytest.mark.parametrize("case", CASES)
def test_kspace_to_image_accumulates(case, form):
    op, o32, o64, traj, X, smaps, d = make(case, form)
    X0 = crandn(X.shape, 99)
    Xg = op.forward(to_dev(d), X=to_dev(X0.copy()), accumulate=True).cpu().numpy()
    X32 = o32.kspace_to_image(d, smaps, traj, X=X0.copy())
    X64 = o64.kspace_to_image(d.astype(np.complex128), smaps.astype(np.complex128), traj,
                              X=X0.astype(np.complex128))
    e32, e64 = rel(Xg, X32), rel(Xg, X64)
    print("AH %s %s: vs f32 %.2e vs f64 %.2e" % (form, case, e32, e64))
    assert e32 <= TOL32 and e64 <= TOL64
    # overwrite mode equals accumulate onto zeros
    Xz = op.forward(to_dev(d), X=to_dev(X0.copy()), accumulate=False).cpu().numpy()
    assert rel(Xz, X32 - X0) <= TOL32 * 10


@pytest.mark.parametrize("form", ["batched", "plain"])
def test_residual_and_skipped_samples(form):
    case = CASES[1]
    op, o32, _, traj, X, smaps, d = make(case, form)
    kw = traj[3].copy()
    kw[::7] = 0.0  # skipped samples give 0, not -d
    traj2 = (traj[0], traj[1], traj[2], kw)
    op.set_traj(*traj_dev(traj2))
    y = op.backward_residual(to_dev(X), to_dev(d)).cpu().numpy()
    y32 = o32.image_to_kspace(X, smaps, traj2, sub=d)
    assert rel(y, y32) <= TOL32
    assert np.all(y[:, ::7] == 0)


@pytest.mark.parametrize("form", ["batched", "plain"])
@pytest.mark.parametrize("case", CASES[:2])
def test_normal_operator(case, form):
    op, o32, o64, traj, X, smaps, d = make(case, form)
    L = op.normal(to_dev(X)).cpu().numpy()
    L32 = o32.normal(X, smaps, traj)
    L64 = o64.normal(X.astype(np.complex128), smaps.astype(np.complex128), traj)
    e32, e64 = rel(L, L32), rel(L, L64)
    print("EhE %s %s: vs f32 %.2e vs f64 %.2e" % (form, case, e32, e64))
    assert e32 <= TOL32 and e64 <= TOL64


def test_adjoint_dot_product():
    case = CASES[3]
    N, C = case[0], case[1]
    traj, X, smaps, d = radial_problem(N, C, case[4], case[5], seed=5, kw_ones=True)
    op = GridFFT(N, ncoils=C, overgrid=case[2], dwin=case[3], k_rad=1e30)
    op.set_traj(*traj_dev(traj))
    op.set_smaps(to_dev(smaps))
    Ax = op.backward(to_dev(X)).cpu().numpy().astype(np.complex128)
    AHy = op.forward(to_dev(d), accumulate=False).cpu().numpy().astype(np.complex128)
    lhs, rhs = np.vdot(Ax, d), np.vdot(X, AHy)
    err = abs(lhs - rhs) / abs(lhs)
    print("adjoint dot rel err %.2e" % err)
    assert err < 1e-5


def test_batched_equals_plain_and_coil_entry_points():
    case = CASES[1]
    opb, o32, _, traj, X, smaps, d = make(case, "batched")
    opp, _, _, _, _, _, _ = make(case, "plain")
    yb = opb.backward(to_dev(X)).cpu().numpy()
    yp = opp.backward(to_dev(X)).cpu().numpy()
    assert rel(yb, yp) <= TOL32
    # single-coil reference-style calls (trajectory per call)
    tr = traj_dev(traj)
    sm = to_dev(smaps)
    y0 = opp.coil_backward(to_dev(X), sm[0].contiguous(), tr).cpu().numpy()
    assert rel(y0, yp[0]) <= 1e-6
    Xa = torch.zeros(X.shape, dtype=torch.complex64, device="cuda")
    for c in range(case[1]):
        opp.coil_forward(Xa, sm[c].contiguous(), to_dev(d[c]), tr)
    Xr = o32.kspace_to_image(d, smaps, traj)
    assert rel(Xa.cpu().numpy(), Xr) <= TOL32


def test_edge_samples_empty_and_all_skipped():
    N = 16
    op = GridFFT(N, ncoils=2, overgrid=1.5)
    o32 = orc.OracleOp(N, overgrid=1.5, precision="f32")
    sm = crandn((2, N, N, N), 3)
    op.set_smaps(to_dev(sm))
    X = crandn((N, N, N), 4)
    # samples on the faces and corners of |k| <= N/2 and beyond
    vals = np.array([-N / 2, N / 2, -N / 2 + 0.25, N / 2 - 0.25, 0.0, N / 2 + 3.0], np.float32)
    g = np.stack(np.meshgrid(vals, vals, vals, indexing="ij")).reshape(3, -1)
    traj = (g[0].copy(), g[1].copy(), g[2].copy(), np.ones(g.shape[1], np.float32))
    op.set_traj(*traj_dev(traj))
    y = op.backward(to_dev(X)).cpu().numpy()
    assert rel(y, o32.image_to_kspace(X, sm, traj)) <= TOL32
    d = crandn((2, g.shape[1]), 5)
    Xg = op.forward(to_dev(d), accumulate=False).cpu().numpy()
    assert rel(Xg, o32.kspace_to_image(d, sm, traj)) <= TOL32
    # empty frame
    e = torch.empty(0, dtype=torch.float32, device="cuda")
    op.set_traj(e, e, e, e)
    assert op.backward(to_dev(X)).shape == (2, 0)
    assert torch.count_nonzero(op.forward(torch.empty((2, 0), dtype=torch.complex64, device="cuda"),
                                          accumulate=False)) == 0
    # every sample skipped (kw = 0)
    z = (traj[0], traj[1], traj[2], np.zeros_like(traj[3]))
    op.set_traj(*traj_dev(z))
    assert torch.count_nonzero(op.backward(to_dev(X))) == 0
    assert torch.count_nonzero(op.forward(to_dev(d), accumulate=False)) == 0


def test_bad_arguments_raise():
    op = GridFFT(16, ncoils=2)
    with pytest.raises(MrnError):
        op.backward(torch.zeros((16, 16, 16), dtype=torch.complex64, device="cuda"))  # no trajectory
    with pytest.raises(ValueError):
        op.set_smaps(torch.zeros((3, 16, 16, 16), dtype=torch.complex64, device="cuda"))
    with pytest.raises(MrnError):
        GridFFT(16, ncoils=2, dwin=5.0)


@pytest.mark.parametrize("form,graph", [("batched", True), ("batched", False), ("plain", False)])
def test_cg_matches_oracle(form, graph):
    case = CASES[0]
    N, C = case[0], case[1]
    traj, X, smaps, _ = radial_problem(N, C, case[4], case[5], seed=11)
    o32 = orc.OracleOp(N, overgrid=case[2], dwin=case[3], precision="f32")
    o64 = orc.OracleOp(N, overgrid=case[2], dwin=case[3], precision="f64")
    d = o64.image_to_kspace(X.astype(np.complex128), smaps.astype(np.complex128), traj).astype(np.complex64)
    d = d + 0.01 * crandn(d.shape, 12) * np.sqrt(np.mean(np.abs(d) ** 2))
    op = GridFFT(N, ncoils=C, overgrid=case[2], dwin=case[3], form=form)
    op.set_traj(*traj_dev(traj))
    op.set_smaps(to_dev(smaps))
    op.set_kdata(to_dev(d))
    iters = 8
    Xg, tr = op.cg(iters, graph=graph)
    X32, t32 = o32.cg(smaps, d, traj, iters)
    X64, t64 = o64.cg(smaps.astype(np.complex128), d.astype(np.complex128), traj, iters)
    e32, e64 = rel(Xg[0].cpu().numpy(), X32), rel(Xg[0].cpu().numpy(), X64)
    print("CG %s graph=%s: vs f32 %.2e vs f64 %.2e" % (form, graph, e32, e64))
    assert e32 <= 1e-5 and e64 <= 1e-4
    t = tr.numpy()
    for col in (0, 3):  # sum|R|^2 before / after the step
        assert np.max(np.abs(t[:, col] - t32[:, col]) / t32[:, col]) < 1e-4
    a_g, a_32 = t[:, 1] + 1j * t[:, 2], t32[:, 1] + 1j * t32[:, 2]
    assert np.max(np.abs(a_g - a_32) / np.abs(a_32)) < 1e-4  # step size (Im ~ 0)


def test_power_method_and_fista_match_oracle():
    case = CASES[0]
    N, C = case[0], case[1]
    traj, X, smaps, _ = radial_problem(N, C, case[4], case[5], seed=21)
    o32 = orc.OracleOp(N, overgrid=case[2], dwin=case[3], precision="f32")
    d = o32.image_to_kspace(X, smaps, traj)
    op = GridFFT(N, ncoils=C, overgrid=case[2], dwin=case[3])
    op.set_traj(*traj_dev(traj))
    op.set_smaps(to_dev(smaps))
    op.set_kdata(to_dev(d))
    step = op.power_method(15)
    step32, _ = o32.power_method(smaps, traj, iters=15)
    assert abs(step - step32) / step32 < 1e-5
    for graph in (True, False):
        Xg, res = op.fista(6, step32, graph=graph)
        X32, r32 = o32.fista(smaps, d, traj, 6, step32)
        e = rel(Xg[0].cpu().numpy(), X32)
        print("FISTA graph=%s vs f32 %.2e" % (graph, e))
        assert e <= 1e-5
        assert np.max(np.abs(res.numpy() - r32) / r32) < 1e-4


def test_frames_of_uneven_length():
    """Frames are independent operators sharing coil maps; FISTA over frames
    with one step equals per-frame FISTA."""
    N, C = 16, 2
    from mri_recon_amd import synth
    frames = synth.gated_radial_frames(3, 60, 16, N, seed=3)
    smaps = synth.coil_maps(C, N, seed=4).numpy()
    op = GridFFT(N, ncoils=C, nframes=3, overgrid=1.5)
    op.set_smaps(to_dev(smaps))
    o32 = orc.OracleOp(N, overgrid=1.5, precision="f32")
    Xs, ds = [], []
    for f, tr in enumerate(frames):
        tr = [t.numpy() for t in tr]
        op.set_traj(*traj_dev(tr), frame=f)
        X = crandn((N, N, N), 30 + f)
        d = o32.image_to_kspace(X, smaps, tr)
        op.set_kdata(to_dev(d), frame=f)
        y = op.backward(to_dev(X), frame=f).cpu().numpy()
        assert rel(y, d) <= TOL32
        Xs.append(tr)
        ds.append(d)
    Xg, _ = op.fista(4, 1e-5, graph=True)
    for f, tr in enumerate(Xs):
        Xf, _ = o32.fista(smaps, ds[f], tr, 4, 1e-5)
        assert rel(Xg[f].cpu().numpy(), Xf) <= TOL32


def test_autotune_keeps_results():
    case = CASES[3]
    op, o32, _, traj, X, smaps, d = make(case)
    before = op.forward(to_dev(d), accumulate=False).cpu().numpy()
    op.autotune(cache=None)
    after = op.forward(to_dev(d), accumulate=False).cpu().numpy()
    # the tuned shapes sum in another order (bins, FFT pass form)
    assert rel(after, before) <= 1e-5
    assert rel(after, o32.kspace_to_image(d, smaps, traj)) <= TOL32


@pytest.mark.gpu
@pytest.mark.parametrize("C,dwin,mode", [(16, 2.5, 1), (8, 2.0, 1), (16, 1.5, 1), (8, 3.0, 1), (16, 2.0, 1),
                                         (32, 2.5, 1), (16, 3.0, 1), (16, 2.5, 2), (16, 1.5, 2), (16, 2.0, 2),
                                         (32, 2.5, 2), (32, 2.0, 2)])
def test_tile_staged_gather_equals_direct_and_oracle(C, dwin, mode):
    """The opt-in gathers of the dense tiles (mode 1: coil blocks of 8 or
    16, LDS region per dense tile; mode 2: the box matrix-core gather of
    16-coil blocks; the direct gather for the samples of sparse tiles; this
    radial set has both kinds) match the direct gather and the FP32 oracle.
    With 8-coil blocks A sums the same products in the same order (the same
    bits); with 16-coil blocks the matrix cores sum them in another order
    (FP32 rounding, <= 1e-6).  dwin 2.0 puts the k = 0 samples exactly on
    the window edge (5 taps on an axis), the tap-by-tap path."""
    N = 32
    traj, X, smaps, d = radial_problem(N, C, 900, 40, seed=41)
    ops = []
    for tiles in (mode, False):
        op = GridFFT(N, ncoils=C, overgrid=1.25, dwin=dwin, gather_tiles=tiles)
        op.set_traj(*traj_dev(traj))
        op.set_smaps(to_dev(smaps))
        ops.append(op)
    Xd = to_dev(X)
    a_tile = ops[0].backward(Xd).cpu().numpy()
    a_dir = ops[1].backward(Xd).cpu().numpy()
    n_tile = ops[0].normal(Xd).cpu().numpy()
    n_dir = ops[1].normal(Xd).cpu().numpy()
    print("tile gather %d C=%d dwin=%.1f: A %.2e  AHA %.2e" % (mode, C, dwin, rel(a_tile, a_dir), rel(n_tile, n_dir)))
    if C == 8:
        assert np.array_equal(a_tile, a_dir)
    else:
        assert rel(a_tile, a_dir) <= 1e-6
    # E'E also runs the scatter, whose window-edge samples add with atomics
    assert rel(n_tile, n_dir) <= 1e-6
    o = orc.OracleOp(N, overgrid=1.25, dwin=dwin, precision="f32")
    ref = o.image_to_kspace(X, smaps, traj)
    assert rel(a_tile, ref) < 1e-5


@pytest.mark.parametrize("C,dwin,max_visits", [(16, 2.5, 0), (32, 2.0, 0), (16, 1.5, 64), (32, 2.5, 64),
                                               (16, 3.0, 0)])
def test_matrix_core_scatter_matches_valu_and_oracle(C, dwin, max_visits):
    """The matrix-core scatter (16-coil blocks) sums the same products as the
    VALU scatter in another order: Aᴴ agrees with it to FP32 rounding and with
    the oracle to the usual bar.  max_visits=64 splits the dense tiles over
    several work items, whose partial sums are added with atomics."""
    N = 32
    traj, X, smaps, d = radial_problem(N, C, 900, 40, seed=43)
    out = []
    for valu in (False, True):
        op = GridFFT(N, ncoils=C, overgrid=1.25, dwin=dwin, scatter_valu=valu, max_visits=max_visits)
        op.set_traj(*traj_dev(traj))
        op.set_smaps(to_dev(smaps))
        out.append(op.forward(to_dev(d), accumulate=False).cpu().numpy())
    assert rel(out[0], out[1]) <= 1e-6
    o32 = orc.OracleOp(N, overgrid=1.25, dwin=dwin, precision="f32")
    o64 = orc.OracleOp(N, overgrid=1.25, dwin=dwin, precision="f64")
    X32 = o32.kspace_to_image(d, smaps, traj)
    X64 = o64.kspace_to_image(d.astype(np.complex128), smaps.astype(np.complex128), traj)
    e32, e64 = rel(out[0], X32), rel(out[0], X64)
    print("matrix-core AH C=%d dwin=%.1f: vs f32 %.2e vs f64 %.2e" % (C, dwin, e32, e64))
    assert e32 <= TOL32 and e64 <= TOL64
