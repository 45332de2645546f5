"""Pins for the oracle's Pipe DCF (recon_lib.cpp:854-917, DCFgridFFT.cpp),
independent of the restated loops:

* the table is the published fifth-order polynomial evaluated at i / len (len
  = 600 * dwin + 100 entries) and sums to 0.5 * 600 after normalisation;
* the grid extents are trunc(os * N + 16), with the kx extent taken from Nz;
* one sample on a grid point: its taps sit at whole-cell distances, so the
  weights are table entries 0, 600, 1200 and the DCF is 1 / (sum of squared
  weights), reached after one iteration and kept after that;
* duplicating every sample halves every weight (A^T A doubles);
* the iteration drives A^T A kw towards 1 (the Pipe fixed point);
* scale_kw multiplies by (|k| / 128 * (acc - 1) + 1)^3;
* the FP32 restatement agrees with the FP64 one.
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from mri_recon_amd import synth

COEF = [1., 0.04522831, -3.36020304, 1.12417012, 2.82448025, -1.63447764]


def poly(x):
    return sum(c * x ** i for i, c in enumerate(COEF))


def test_table_is_polynomial_at_i_over_len_and_normalised():
    t = orc.dcf_table(2.1, 0, "f64")
    L = len(t)
    assert L == int(np.float32(2.1) * np.float32(600.0)) + 100 == 1360
    assert abs(t.sum() - 300.0) < 1e-9
    i = np.arange(L)
    ref = np.array([poly(v) for v in i / L])
    assert np.allclose(t / t[0], ref / ref[0], rtol=0, atol=1e-12)
    # the argument is i / len, not i / (600 * dwin): at the window edge the
    # kernel has not reached its root yet
    assert t[1260] / t[0] > 0.01
    t32 = orc.dcf_table(2.1, 0, "f32")
    assert np.allclose(t32, t, rtol=2e-7)


def test_grid_extents_quirk():
    assert orc.dcf_grid_size(orc.dcf_params(256)) == (553, 553, 553)
    # Nx=128, Ny=96, Nz=64: the x extent comes from Nz
    assert orc.dcf_grid_size(orc.dcf_params((128, 96, 64))) == (150, 217, 284)


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_single_sample_on_grid_point(precision):
    t = orc.dcf_table(2.1, 0, "f64")
    s1 = t[0] ** 2 + 2 * t[600] ** 2 + 2 * t[1200] ** 2
    want = 1.0 / s1 ** 3
    z = np.zeros(1, np.float32)
    for iters in (1, 3):
        kw, _ = orc.dcf_calc(orc.dcf_params(32, iters=iters), z, z, z, precision)
        assert abs(kw[0] / want - 1) < (2e-6 if precision == "f32" else 1e-12)


def _radial(nspokes=300, npts=24, N=32, seed=0):
    kx, ky, kz, _ = [t.numpy() for t in synth.radial_traj(nspokes, npts, N)]
    return kx, ky, kz


def test_duplicated_samples_halve_the_weights():
    kx, ky, kz = _radial()
    p = orc.dcf_params(32, iters=6)
    kw1, _ = orc.dcf_calc(p, kx, ky, kz, "f64")
    kw2, _ = orc.dcf_calc(p, np.r_[kx, kx], np.r_[ky, ky], np.r_[kz, kz], "f64")
    M = kx.size
    assert np.allclose(kw2[:M], kw1 / 2, rtol=1e-12, atol=0)
    assert np.allclose(kw2[M:], kw1 / 2, rtol=1e-12, atol=0)


def test_iteration_approaches_fixed_point():
    kx, ky, kz = _radial()
    res = []
    for iters in (2, 10, 30):
        p = orc.dcf_params(32, iters=iters)
        kw, _ = orc.dcf_calc(p, kx, ky, kz, "f64")
        r = orc.dcf_backward(p, orc.dcf_forward(p, kw, kx, ky, kz, "f64"), kx, ky, kz, "f64")
        res.append(np.median(np.abs(r[kw > 0] - 1)))
    assert res[0] > res[1] > res[2]
    assert res[2] < 0.02


def test_scale_kw_acc():
    kx, ky, kz = _radial(40, 8)
    p = orc.dcf_params(32, iters=0, acc=2.0)
    kw, _ = orc.dcf_calc(p, kx, ky, kz, "f64")
    kr = np.sqrt(kx.astype(np.float64) ** 2 + ky ** 2 + kz ** 2)
    assert np.allclose(kw, (kr / 128 + 1) ** 3, rtol=1e-6)


def test_wider_kernel_with_acc_reaches_more_cells():
    """acc > 1 widens the kernel with |k|: an outer sample's footprint grows."""
    k = np.array([14.0], np.float32)
    z = np.zeros(1, np.float32)
    X1 = orc.dcf_forward(orc.dcf_params(32, acc=1.0), np.ones(1), k, z, z, "f64")
    X4 = orc.dcf_forward(orc.dcf_params(32, acc=4.0), np.ones(1), k, z, z, "f64")
    assert np.count_nonzero(X4) > np.count_nonzero(X1)


def test_f32_matches_f64():
    kx, ky, kz = _radial()
    p = orc.dcf_params(32, iters=10)
    a, sa = orc.dcf_calc(p, kx, ky, kz, "f32")
    b, sb = orc.dcf_calc(p, kx, ky, kz, "f64")
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-5
    assert np.allclose(sa, sb, rtol=1e-4)
