"""Pins for the CPU oracle (oracle/gridfft_oracle.c).

The reference ships no golden vectors for gridFFT, fft3, CG or FISTA and
cannot be built here (Blitz++, Armadillo and FFTW are absent), so each oracle
function is pinned by something independent of its own formula:
closed forms, scipy's I0, numpy's FFT and brute-force DFT/NDFT sums, the
adjoint identity, special cases that reduce to a plain DFT, and dense FP64
solves of the normal equations.
"""
import math

import numpy as np
import pytest
import scipy.special

from oracle import oracle as orc

RNG = np.random.default_rng(1234)


def crandn(*shape, rng=RNG):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / math.sqrt(2)


# ------------------------------------------------------------------ geometry
@pytest.mark.parametrize("N,os_,S,og", [
    (128, 1.5, 192, 32),   # 128^3 x 8 coils, os 1.5
    (256, 1.25, 320, 32),  # 256^3 x 32 coils, os 1.25
    (160, 1.5, 240, 40),   # 160^3 gated frames
    (192, 1.5, 288, 48),   # 192^3 cones
    (320, 1.5, 480, 80),   # 320^3 x 48 coils
    (16, 1.5, 32, 8),      # rounding to 16 raises os to 2.0
    (32, 1.5, 48, 8),
    (256, 1.5, 384, 64),
])
def test_geometry_kat(N, os_, S, og):
    op = orc.OracleOp(N, overgrid=os_, precision="f32")
    g = op.geometry()
    assert tuple(g["S"]) == (S, S, S)
    assert tuple(g["og"]) == (og, og, og)
    assert abs(g["grid"][0] - S / N) < 1e-6


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("N,os_,S,og", [
    (36, 1.5, 64, 14),  # float grid 64/36 rounds up: 36 * 0.77777779 / 2 = 14.0000002
    (48, 1.5, 80, 15),  # float grid 80/48 rounds down: 48 * 0.66666663 / 2 = 15.999999
    (40, 1.5, 64, 12),
])
def test_geometry_float_arithmetic_in_both_builds(N, os_, S, og, precision):
    """grid_x is a float member in the reference (gridFFT.h), so the grid
    offset og = (int)((float)N * (grid - 1) / 2.0) truncates the float value
    in the FP64 oracle too: N = 48 gives 15, not the centred 16, and N = 36
    gives 14 where double arithmetic would give 13."""
    g32 = np.float32(16.0 * math.ceil(np.float32(os_) * np.float32(N) / 16.0) / N)
    assert int(np.float32(N) * (g32 - np.float32(1)) / 2.0) == og
    g = orc.OracleOp(N, overgrid=os_, precision=precision).geometry()
    assert tuple(g["S"]) == (S, S, S) and tuple(g["og"]) == (og, og, og)


@pytest.mark.parametrize("N,os_,dwin,beta", [
    (128, 1.5, 2.0, 7.779),    # beta_min dominates
    (256, 1.25, 2.5, 12.175),  # beta_min dominates (textbook Beatty with W=5 gives 8.996)
    (160, 1.5, 2.5, 9.999),
])
def test_beta_kat(N, os_, dwin, beta):
    op = orc.OracleOp(N, overgrid=os_, dwin=dwin, precision="f32")
    assert abs(op.geometry()["beta"][0] - beta) < 1.5e-3
    # half-width quirk: Beatty's formula with the full width 2*dwin differs
    g = op.geometry()["grid"][0]
    W = 2 * dwin
    beatty = math.pi * math.sqrt(W * W / (g * g) * (g - 0.5) ** 2 - 0.8)
    if N == 256:
        assert abs(beatty - 8.996) < 1e-3


def test_bessi0_vs_scipy():
    xs = np.concatenate([np.linspace(0, 3.7499, 400), np.linspace(3.75, 30, 400)])
    got = np.array([orc.bessi0(x) for x in xs])
    ref = scipy.special.i0(xs)
    assert np.max(np.abs(got / ref - 1)) < 2.5e-7  # A&S 9.8.1/9.8.2 bounds


def _kb_exact(u, beta, a):
    u = np.abs(u)
    s = np.sqrt(np.clip(1 - (u / a) ** 2, 0, None))
    return np.where(u <= a, scipy.special.i0(beta * s) / scipy.special.i0(beta), 0.0)


@pytest.mark.parametrize("dwin,os_", [(2.5, 1.25), (2.0, 1.5), (2.5, 1.5)])
def test_kb_table_closed_form(dwin, os_):
    op = orc.OracleOp(64, overgrid=os_, dwin=dwin, precision="f64")
    g = op.geometry()
    L, beta = int(g["L"][0]), g["beta"][0]
    f = op.table(0)
    assert len(f) == L + 10
    assert abs(f.sum() - 300.0) < 1e-9          # normalised to 0.5 * 600
    assert np.all(f[L + 1:] == 0)                # 9 zero pad entries
    raw = _kb_exact(np.arange(L + 1) / 600.0, beta, dwin)
    assert np.max(np.abs(f[:L + 1] / f[0] - raw / raw[0])) < 3e-7


def test_triangle_and_poly_tables():
    op = orc.OracleOp(32, overgrid=1.0, kernel="triangle", dwin=1.0, precision="f64")
    f = op.table(0)
    L = 600
    assert np.allclose(f[:L + 1] / f[0], 1 - np.arange(L + 1) / L, atol=1e-12)
    assert abs(f.sum() - 300) < 1e-9
    op = orc.OracleOp(32, kernel="poly", precision="f64")
    f = op.table(0)
    assert len(f) == 1210 and abs(f.sum() - 300) < 1e-9
    x = np.arange(1210) / 1210.0
    poly = np.polyval([-1.63447764, 2.82448025, 1.12417012, -3.36020304, 0.04522831, 1.0], x)
    assert np.allclose(f / f[0], poly / poly[0], rtol=1e-12)


@pytest.mark.parametrize("N,os_,dwin", [(32, 1.5, 2.5), (64, 1.25, 2.5), (40, 1.5, 2.0)])
def test_window_vs_analytic_kb_transform(N, os_, dwin):
    """Deapodisation = 1 / (discrete kernel transform).  Trapezoid rule turns
    the reference's sum (DC counted twice) into 600*Khat(nu) + f[0] - f[L]cos,
    with Khat the closed-form Kaiser-Bessel transform."""
    op = orc.OracleOp(N, overgrid=os_, dwin=dwin, precision="f64")
    g = op.geometry()
    S, og, L, beta = int(g["S"][0]), int(g["og"][0]), int(g["L"][0]), g["beta"][0]
    f = op.table(0)
    norm = f[0] / _kb_exact(0.0, beta, dwin)
    i = np.arange(S)
    nu = (i - S / 2) / S
    z = beta ** 2 - (2 * math.pi * dwin * nu) ** 2
    khat = np.where(z > 0, np.sinh(np.sqrt(np.abs(z))) / np.sqrt(np.abs(z)),
                    np.sin(np.sqrt(np.abs(z))) / np.sqrt(np.abs(z)))
    khat = khat * 2 * dwin * norm / scipy.special.i0(beta)
    denom = 600 * khat + f[0] - f[L] * np.cos(2 * math.pi * nu * L / 600)
    w = 600 / denom * (2 * (i % 2) - 1) / S
    w[(i < og) | (i >= og + N)] = 0
    w /= w.max()
    got = op.window(0)
    assert np.max(np.abs(got - w)) < 2e-6
    assert got.max() == 1.0
    assert np.all(got[:og] == 0) and np.all(got[og + N:] == 0)


# ----------------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 7, 8, 12, 15, 16, 30, 48, 49, 64, 96, 192, 240,
                               288, 320, 384, 480])
def test_fft1d_vs_numpy(n):
    x = crandn(3, n)
    ref = np.fft.fft(x, axis=-1) / math.sqrt(n)
    got = orc.fft1d(x, -1, "f64")
    assert np.max(np.abs(got - ref)) < 1e-12 * max(1, math.log2(n))
    inv = orc.fft1d(got, +1, "f64")
    assert np.max(np.abs(inv - x)) < 1e-12
    got32 = orc.fft1d(x, -1, "f32")
    assert np.linalg.norm(got32 - ref) / np.linalg.norm(ref) < 2e-6


@pytest.mark.parametrize("n", [6, 10, 24, 45])
def test_fft1d_vs_bruteforce(n):
    x = crandn(n)
    k = np.arange(n)
    F = np.exp(-2j * np.pi * np.outer(k, k) / n) / math.sqrt(n)
    assert np.max(np.abs(orc.fft1d(x, -1) - F @ x)) < 1e-13
    assert np.max(np.abs(orc.fft1d(x, +1) - F.conj() @ x)) < 1e-13


def test_fft3_pruned_equals_full():
    """do_fft transforms only the crop lines in z and y: identical to a full
    unitary 3-D DFT of the zero-padded grid."""
    op = orc.OracleOp(16, overgrid=1.5, precision="f64")
    Sz, Sy, Sx = op.grid_shape()
    og = op.og[0]
    g = np.zeros((Sz, Sy, Sx), complex)
    g[og:og + 16, og:og + 16, og:og + 16] = crandn(16, 16, 16)
    got = op.fft3(g, -1)
    ref = np.fft.fftn(g) / math.sqrt(Sx * Sy * Sz)
    assert np.max(np.abs(got - ref)) < 1e-12
    back = op.fft3(got, +1)
    assert np.max(np.abs(back - g)) < 1e-12


# ------------------------------------------------------------------ operator
def rand_traj(n, N, rng=RNG, kw=None):
    k = rng.uniform(-N / 2, N / 2, size=(3, n)).astype(np.float32)
    w = np.ones(n, np.float32) if kw is None else kw
    return k[0], k[1], k[2], w


def test_adjoint_dot_fp64():
    op = orc.OracleOp(16, overgrid=1.5, k_rad=1e30, precision="f64")
    C, M = 2, 1500
    traj = rand_traj(M, 16)
    x = crandn(16, 16, 16)
    sm = crandn(C, 16, 16, 16)
    y = crandn(C, M)
    Ax = op.image_to_kspace(x, sm, traj)
    AHy = op.kspace_to_image(y, sm, traj)
    lhs = np.vdot(Ax, y)
    rhs = np.vdot(x, AHy)
    assert abs(lhs - rhs) / abs(lhs) < 1e-12


def test_linearity_fp64():
    op = orc.OracleOp(16, overgrid=1.5, precision="f64")
    traj = rand_traj(800, 16)
    x1, x2 = crandn(16, 16, 16), crandn(16, 16, 16)
    a, b = 0.7 - 0.2j, -1.3 + 0.4j
    y = op.backward(a * x1 + b * x2, None, traj)
    y2 = a * op.backward(x1, None, traj) + b * op.backward(x2, None, traj)
    assert np.max(np.abs(y - y2)) < 1e-12 * np.max(np.abs(y))


def test_on_grid_fast_grid_reduces_to_centred_dft():
    """-fast_grid (os 1, triangle, dwin 1) with integer k: gridding is exact and
    A x = f0^3 N^-3/2 * sum_r |w(r)| x(r) exp(-2 pi i k.(r - N/2)/N)."""
    N = 16
    op = orc.OracleOp(N, overgrid=1.0, kernel="triangle", dwin=1.0, precision="f64")
    assert op.S == (N, N, N)
    M = 400
    k = RNG.integers(-N // 2, N // 2, size=(3, M))
    traj = (k[0].astype(np.float32), k[1].astype(np.float32), k[2].astype(np.float32),
            np.ones(M, np.float32))
    x = crandn(N, N, N)
    w = [np.abs(op.window(a)) for a in range(3)]
    xw = x * w[2][:, None, None] * w[1][None, :, None] * w[0][None, None, :]
    r = np.arange(N) - N / 2
    ex = np.exp(-2j * np.pi * np.outer(k[0], r) / N)
    ey = np.exp(-2j * np.pi * np.outer(k[1], r) / N)
    ez = np.exp(-2j * np.pi * np.outer(k[2], r) / N)
    ndft = np.einsum("mz,my,mx,zyx->m", ez, ey, ex, xw)
    got = op.backward(x, None, traj)
    f0 = op.table(0)[0]
    assert np.max(np.abs(got - f0 ** 3 * N ** -1.5 * ndft)) < 1e-12 * np.max(np.abs(got))


def test_kb_gridding_close_to_ndft():
    """Off-grid KB gridding approximates c * NDFT(x) with a real constant c;
    away from the faces (where the reference clips taps instead of wrapping)
    the scheme's discretisation error is ~3e-3, mostly the DC double count."""
    N = 16
    op = orc.OracleOp(N, overgrid=2.0, dwin=2.5, k_rad=1e30, precision="f64")
    k = RNG.uniform(-N / 2 + 2, N / 2 - 2, size=(3, 500)).astype(np.float32)
    traj = (k[0], k[1], k[2], np.ones(500, np.float32))
    x = crandn(N, N, N)
    r = np.arange(N) - N / 2
    kx, ky, kz = traj[0].astype(float), traj[1].astype(float), traj[2].astype(float)
    ndft = np.einsum("mz,my,mx,zyx->m", np.exp(-2j * np.pi * np.outer(kz, r) / N),
                     np.exp(-2j * np.pi * np.outer(ky, r) / N),
                     np.exp(-2j * np.pi * np.outer(kx, r) / N), x)
    got = op.backward(x, None, traj)
    c = np.vdot(ndft, got) / np.vdot(ndft, ndft)
    assert abs(c.imag) < 1e-3 * abs(c.real)
    rel = np.linalg.norm(got - c.real * ndft) / np.linalg.norm(got)
    assert rel < 5e-3


def test_gather_tap_weights_delta_probe():
    """A delta on the grid picks out one tap: its weight must be
    sgn(lx)sgn(ly)sgn(lz) K(dx)K(dy)K(dz) with K the Kaiser-Bessel function
    (scipy I0) normalised like the table; cells out of reach give 0."""
    N = 32
    op = orc.OracleOp(N, overgrid=1.5, dwin=2.5, precision="f64")
    g = op.geometry()
    S, beta, gr = op.S[0], g["beta"][0], g["grid"][0]
    f = op.table(0)
    norm = f[0]
    K = lambda u: norm * _kb_exact(u, beta, 2.5)
    sgn = lambda l: 2 * (l % 2) - 1
    rng = np.random.default_rng(7)
    for trial in range(20):
        kk = rng.uniform(-N / 2 + 3, N / 2 - 3, 3).astype(np.float32)
        d = kk.astype(np.float64) * gr + S // 2
        base = np.floor(d).astype(int)
        lx, ly, lz = base + rng.integers(-3, 4, 3)
        grid = np.zeros(op.grid_shape(), complex)
        grid[lz, ly, lx] = 1.0
        traj = (np.array([kk[0]]), np.array([kk[1]]), np.array([kk[2]]), np.ones(1, np.float32))
        got = op.stage_gather(grid, traj)[0]
        dist = np.abs(d - np.array([lx, ly, lz]))
        if np.any(dist > 2.5):
            assert got == 0
        else:
            exp = np.prod([sgn(l) * K(di) for l, di in zip((lx, ly, lz), dist)])
            assert abs(got - exp) < 2e-6 * abs(norm) ** 3


def test_edge_clipping_no_wrap():
    """k = -N/2 puts the sample on grid cell 0: taps 0..2 only; k = +N/2 on
    cell S: taps S-2, S-1 only (no periodic wrap)."""
    N = 32
    op = orc.OracleOp(N, overgrid=1.5, dwin=2.5, precision="f64")
    S = op.S[0]
    for kval, cells in ((-N / 2, [0, 1, 2]), (N / 2, [S - 2, S - 1])):
        hit = []
        for l in range(S):
            grid = np.zeros(op.grid_shape(), complex)
            grid[S // 2, S // 2, l] = 1
            traj = (np.array([kval], np.float32), np.zeros(1, np.float32), np.zeros(1, np.float32),
                    np.ones(1, np.float32))
            if op.stage_gather(grid, traj)[0] != 0:
                hit.append(l)
        assert hit == cells


def test_kw_zero_and_out_of_range_samples():
    op = orc.OracleOp(16, overgrid=1.5, precision="f64")
    x = crandn(16, 16, 16)
    traj = (np.array([1.0, 2.0, 40.0], np.float32), np.array([0.5, 0.0, 0.0], np.float32),
            np.zeros(3, np.float32), np.array([1.0, 0.0, 1.0], np.float32))
    sub = crandn(1, 3)
    y = op.image_to_kspace(x, None, traj, sub=sub)[0]
    assert y[0] != 0 and y[1] == 0 and y[2] == 0  # residual not applied to skipped samples
    X = op.kspace_to_image(np.ones((1, 3), complex), None,
                           (traj[0][1:], traj[1][1:], traj[2][1:], traj[3][1:]))
    assert np.all(X == 0)


def test_centre_sample_adjoint_is_flat_coil_map():
    """One sample at k=0: the deapodisation cancels the kernel apodisation, so
    A^H gives a real constant times conj(S_c) up to the DC double count."""
    N = 32
    op = orc.OracleOp(N, overgrid=1.5, precision="f64")
    sm = crandn(1, N, N, N)
    traj = (np.zeros(1, np.float32),) * 3 + (np.ones(1, np.float32),)
    X = op.kspace_to_image(np.ones((1, 1), complex), sm, traj)
    ratio = X / np.conj(sm[0])
    c = ratio.mean()
    assert np.max(np.abs(ratio.imag)) < 1e-10 * abs(c)
    assert np.max(np.abs(ratio - c)) / abs(c) < 1e-2


def test_taper_applied_in_adjoint_only():
    N = 16
    kk = np.array([5.0], np.float32)
    traj = (kk, np.array([3.0], np.float32), np.array([-2.0], np.float32), np.ones(1, np.float32))
    a = orc.OracleOp(N, overgrid=1.5, k_rad=4.0, precision="f64")
    b = orc.OracleOp(N, overgrid=1.5, k_rad=1e30, precision="f64")
    Xa = a.kspace_to_image(np.ones((1, 1), complex), None, traj)
    Xb = b.kspace_to_image(np.ones((1, 1), complex), None, traj)
    kr = 25 + 9 + 4
    assert np.allclose(Xa, Xb * math.exp(-kr / (2 * 16.0)), rtol=1e-12, atol=0)
    x = crandn(N, N, N)
    assert np.array_equal(a.backward(x, None, traj), b.backward(x, None, traj))


# ------------------------------------------------------------- CG / FISTA
def _dense(op, sm, traj, N):
    nv = N ** 3
    C = sm.shape[0]
    E = np.zeros((C * len(traj[0]), nv), complex)
    for j in range(nv):
        e = np.zeros(nv, complex)
        e[j] = 1
        E[:, j] = op.image_to_kspace(e.reshape(N, N, N), sm, traj).reshape(-1)
    return E


def test_cg_converges_to_dense_solution():
    """SENSE-CG equals textbook CG on the dense normal matrix iterate by
    iterate, and converges to the dense FP64 least-squares solution."""
    N, C, M = 8, 2, 2000
    op = orc.OracleOp(N, overgrid=1.5, k_rad=1e30, precision="f64")
    rng = np.random.default_rng(3)
    kk = rng.uniform(-N / 2, N / 2, size=(3, M)).astype(np.float32)
    kw = rng.uniform(0.5, 2.0, M).astype(np.float32)
    traj = (kk[0], kk[1], kk[2], kw)
    sm = crandn(C, N, N, N, rng=rng)
    d = crandn(C, M, rng=rng)
    E = _dense(op, sm, traj, N)
    W = np.tile(kw.astype(float), C)
    H = E.conj().T @ (W[:, None] * E)
    b = E.conj().T @ (W * d.reshape(-1))
    xstar = np.linalg.solve(H, b)
    x, r = np.zeros_like(b), b.copy()
    p = r.copy()
    X5, trace = op.cg(sm, d, traj, iters=5)
    for it in range(5):
        Hp = H @ p
        rr = np.vdot(r, r).real
        alpha = rr / np.vdot(p, Hp)
        x = x + alpha * p
        r = r - alpha * Hp
        beta = np.vdot(r, r).real / rr
        p = r + beta * p
        assert abs(trace[it, 0] - rr) < 1e-9 * rr
        assert abs(complex(trace[it, 1], trace[it, 2]) - alpha) < 1e-9 * abs(alpha)
    assert np.linalg.norm(X5.reshape(-1) - x) / np.linalg.norm(x) < 1e-9
    X, trace = op.cg(sm, d, traj, iters=160)
    assert np.linalg.norm(X.reshape(-1) - xstar) / np.linalg.norm(xstar) < 1e-7


def test_power_method_and_fista_vs_dense():
    N, C, M = 8, 2, 300
    op = orc.OracleOp(N, overgrid=1.5, k_rad=1e30, precision="f64")
    rng = np.random.default_rng(5)
    kk = rng.uniform(-N / 2, N / 2, size=(3, M)).astype(np.float32)
    kw = rng.uniform(0.5, 2.0, M).astype(np.float32)
    traj = (kk[0], kk[1], kk[2], kw)
    sm = crandn(C, N, N, N, rng=rng)
    E = _dense(op, sm, traj, N)
    W = np.tile(kw.astype(float), C)
    H = E.conj().T @ (W[:, None] * E)
    lmax = np.linalg.eigvalsh(H).max()
    step, eig = op.power_method(sm, traj, iters=200)
    assert abs(eig[-1] - lmax) / lmax < 1e-6
    d = crandn(C, M, rng=rng)
    X, resid = op.fista(sm, d, traj, iters=3, step=step)
    # three iterations: momentum (A,B) = (0,1), (1,0), (1.5,-0.5)
    b = E.conj().T @ (W * d.reshape(-1))
    x_old, x = np.zeros_like(b), np.zeros_like(b)
    for it in range(3):
        A, B = 1 + (it - 1) / (it + 1), -(it - 1) / (it + 1)
        x, x_old = A * x + B * x_old, x
        x = x - step * (H @ x - b)
    assert np.linalg.norm(X.reshape(-1) - x) / np.linalg.norm(x) < 1e-6


def test_forward_sos_is_squared_no_map_forward():
    """forward_sos (deapp, then X += image * conj(image)) equals the squared
    magnitude of the no-map forward, which accumulates the deapodised image."""
    N = 16
    op = orc.OracleOp(N, overgrid=1.5, precision="f64")
    rng = np.random.default_rng(11)
    M = 800
    traj = [rng.uniform(-N / 2, N / 2, M).astype(np.float32) for _ in range(3)]
    traj.append(rng.uniform(0.2, 1.0, M).astype(np.float32))
    d = crandn(M, rng=rng)
    Xf = op.forward(np.zeros((N, N, N), np.complex128), None, d, traj)
    X0 = crandn(N, N, N, rng=rng)
    Xs = op.forward_sos(X0.copy(), d, traj)
    assert np.allclose(Xs - X0, np.abs(Xf) ** 2, rtol=1e-12, atol=1e-12 * np.max(np.abs(Xf) ** 2))


def test_cauchy_step_is_the_exact_line_search():
    """The Cauchy step of the gradient iteration (recon_lib.cpp:2013-2041)
    minimises f(x - a R) = 1/2 sum kw |A(x - a R) - d|^2 along the gradient
    R = A^H kw (A x - d) (taper off): f at the oracle's step lies below f at
    the step scaled by 1 +- 1e-3 and the parabola through the three values
    has its vertex there; the first iteration starts from x = 0."""
    import math
    from tests.helpers import radial_problem
    N, C = 12, 2
    traj, X, smaps, d = radial_problem(N, C, 60, 12, seed=3)
    o = orc.OracleOp(N, overgrid=1.5, precision="f64", k_rad=math.inf)
    sm, dd = smaps.astype(np.complex128), d.astype(np.complex128)
    X1, _, steps = o.fista_cauchy(sm, dd, traj, 1, 0.0, 5)
    a = steps[0]
    kw = traj[3].astype(np.float64)
    R = -X1 / a  # x1 = 0 - a R

    def f(alpha):
        r = o.image_to_kspace(-alpha * R, sm, traj) - dd
        return 0.5 * float(np.sum(kw * np.abs(r) ** 2))

    f0, fp, fm = f(a), f(a * (1 + 1e-3)), f(a * (1 - 1e-3))
    assert f0 < fp and f0 < fm
    # vertex of the parabola through (1-e, 1, 1+e): at 1 within rounding
    e = 1e-3
    vertex = 1 - e * (fp - fm) / (2 * (fp - 2 * f0 + fm))
    assert abs(vertex - 1) < 1e-6
    assert abs(a.imag) < 1e-6 * abs(a.real)  # R^H A^H W A R is real for a Hermitian E'WE


def test_mkl_line_transforms_match_the_textbook_fft():
    """The CPU timing baseline may route the oracle's line transforms through
    MKL's FFTW3 interface; it computes the same unitary 3-D transform as the
    oracle's textbook passes (pinned against numpy above) to FP32 rounding."""
    o = orc.OracleOp(24, overgrid=1.5, precision="f32")
    rng = np.random.default_rng(2)
    g = (rng.standard_normal(o.grid_shape()) + 1j * rng.standard_normal(o.grid_shape())).astype(np.complex64)
    ref = [o.fft3(g, -1), o.fft3(g, 1)]
    if not orc.use_mkl_fft(True, "f32"):
        pytest.skip("libmkl_rt not found")
    try:
        got = [o.fft3(g, -1), o.fft3(g, 1)]
    finally:
        orc.use_mkl_fft(False, "f32")
    for a, b in zip(got, ref):
        assert np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(b)
