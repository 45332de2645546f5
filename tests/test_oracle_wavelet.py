"""Pins for the oracle's L1 wavelet proximal step (wavelet3D.cpp, threshold.cpp,
recon_lib.cpp L1_threshold), independent of the restated formulas:

* the 3-D transform is orthogonal (norm preserved) and backward inverts forward,
  for any shift and for levels limited by divisibility;
* the analysis filters are the 4-tap Daubechies pair: the lowpass sums to
  sqrt(2), the highpass annihilates constants and linear ramps (two vanishing
  moments), so a reversed or sign-flipped filter fails;
* one level along one axis acts on every line separately (separability), and a
  shift is a circular shift of the input;
* the fraction threshold lands within the 0.1% count band of the requested
  quantile of |c|; the soft threshold is c * max(0, 1 - t/|c|);
* the L1 step with one spin zeroes about `fraction` of the wavelet coefficients.
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc

RNG = np.random.default_rng(77)


def crandn(*shape):
    return (RNG.standard_normal(shape) + 1j * RNG.standard_normal(shape)) / math.sqrt(2)


@pytest.mark.parametrize("shape,levels,shift", [
    ((16, 32, 48), (4, 4, 4), (0, 0, 0)),
    ((16, 32, 48), (4, 4, 4), (1, 2, 3)),
    ((8, 12, 20), (4, 4, 4), (3, 0, 1)),     # 12 and 20 allow 2 levels, 8 allows 3
    ((10, 16, 16), (2, 3, 1), (2, 2, 2)),
])
def test_wavelet_orthogonal_and_invertible(shape, levels, shift):
    x = crandn(*shape)
    y = orc.wavelet3d(x, levels, shift)
    assert abs(np.linalg.norm(y) / np.linalg.norm(x) - 1) < 1e-13
    z = orc.wavelet3d(y, levels, shift, inverse=True)
    assert np.linalg.norm(z - x) / np.linalg.norm(x) < 1e-13


def test_levels_limited_by_divisibility():
    assert orc.wavelet_levels((16, 32, 48)) == (4, 4, 4)
    assert orc.wavelet_levels((8, 12, 20)) == (2, 2, 3)   # (x, y, z)
    assert orc.wavelet_levels((10, 16, 16), (2, 3, 1)) == (2, 3, 1)
    assert orc.wavelet_levels((10, 16, 16), (4, 4, 4)) == (4, 4, 1)


def _one_level_x(x, shift=0):
    return orc.wavelet3d(x, (1, 0, 0), (shift, 0, 0))


def test_daubechies4_moments():
    n = 32
    # constant lines: all detail coefficients vanish; approximation = sqrt(2) * c
    x = np.ones((1, 1, n), complex)
    y = _one_level_x(x)[0, 0]
    assert np.max(np.abs(y[n // 2:])) < 1e-14
    assert np.allclose(y[: n // 2], math.sqrt(2))
    # a linear ramp: the details vanish away from the periodic wrap
    x = np.arange(n, dtype=float)[None, None, :] + 0j
    d = _one_level_x(x)[0, 0, n // 2:]
    assert np.max(np.abs(d[2:])) < 1e-12
    assert np.max(np.abs(d[:2])) > 1.0  # the wrap (0 after n-1) is not a ramp


def test_separable_and_shift_is_circular():
    x = crandn(3, 5, 16)
    y = _one_level_x(x)
    for k in range(3):
        for j in range(5):
            line = np.zeros((1, 1, 16), complex)
            line[0, 0] = x[k, j]
            assert np.allclose(_one_level_x(line)[0, 0], y[k, j], atol=1e-14)
    # shifting by s reads the input s samples later
    y1 = _one_level_x(x, shift=1)
    y0 = _one_level_x(np.roll(x, -1, axis=2), shift=0)
    assert np.allclose(y1, y0, atol=1e-14)


@pytest.mark.parametrize("fraction", [0.1, 0.5, 0.9])
def test_fraction_threshold_quantile(fraction):
    c = crandn(40000)
    t = orc.fraction_threshold(c, fraction)
    a = np.abs(c)
    count = int(np.sum(a < t))
    # within the 0.1% band of the target count (the bisection's stop rule)
    assert abs(count - int(fraction * a.size)) <= int(0.001 * a.size) + 1
    # so the threshold sits between the neighbouring quantiles
    assert np.quantile(a, fraction - 0.002) <= t <= np.quantile(a, fraction + 0.002)


def test_fraction_threshold_tiny_target_is_zero():
    assert orc.fraction_threshold(crandn(10), 0.1) == 0.0


def test_soft_threshold_closed_form():
    c = crandn(5000)
    t = 0.7
    got = orc.soft_threshold(c, t)
    a = np.abs(c)
    ref = np.where(a <= t, 0, c * (1 - t / np.maximum(a, 1e-300)))
    assert np.allclose(got, ref, rtol=1e-14, atol=1e-15)


def test_l1_step_sparsifies_about_fraction():
    x = crandn(1, 16, 16, 32)
    frac = 0.6
    y, t = orc.l1_wavelet(x, frac, (4, 4, 4), ((1, 3, 2),))
    w = orc.wavelet3d(y[0], (4, 4, 4), (1, 3, 2))
    zero = np.mean(np.abs(w) < 1e-12)
    assert abs(zero - frac) < 0.01
    # threshold taken on the forward coefficients of x, spins = 1
    assert abs(t - orc.fraction_threshold(orc.wavelet3d(x[0], (4, 4, 4), (1, 3, 2)), frac)) < 1e-12


def test_spin_shifts_in_range_and_deterministic():
    a = orc.spin_shifts(5, 300)
    assert a.min() >= 0 and a.max() <= 3 and len(set(a.tolist())) == 4
    assert np.array_equal(a, orc.spin_shifts(5, 300))
    assert not np.array_equal(a, orc.spin_shifts(6, 300))


def test_fista_l1_reduces_to_fista_without_threshold():
    """A fraction whose target count is below 2 gives threshold 0 (the
    reference returns early), so the proximal step is the identity and the
    L1 loop must reproduce plain FISTA; with a real fraction it differs."""
    from mri_recon_amd import synth
    N = 16
    kx, ky, kz, kw = [t.numpy() for t in synth.radial_traj(80, 12, N)]
    traj = [kx, ky, kz, kw]
    smaps = synth.coil_maps(2, N, seed=3).numpy()
    d = crandn(2, kx.size)
    op = orc.OracleOp(N, overgrid=1.5, precision="f64")
    step, _ = op.power_method(smaps, traj, iters=10)
    X0, r0 = op.fista(smaps, d, traj, 4, step)
    X1, r1, t1 = op.fista_l1(smaps, d, traj, 4, step, fraction=1e-9, spins=2, seed=3)
    assert np.all(t1 == 0)
    assert np.linalg.norm(X1 - X0) / np.linalg.norm(X0) < 1e-12
    assert np.allclose(r1, r0, rtol=1e-12)
    X2, r2, t2 = op.fista_l1(smaps, d, traj, 4, step, fraction=0.5, spins=2, seed=3)
    assert np.all(t2 > 0) and np.linalg.norm(X2 - X0) / np.linalg.norm(X0) > 1e-3
