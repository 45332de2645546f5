"""FP64 reference for the trajectory derivative of the gather (test infrastructure).

The gather of chop_grid_backward (gridFFT.cpp:1166-1322) evaluated in float64
with the exact Kaiser-Bessel kernel in place of the 600-point table:

    y[c, m] = sum_{l in taps(m)} prod_a s_a(l_a) w_a(dk_a - l_a) G_c[l]
    dk_a    = k_a * g_a + S_a / 2
    w_a(u)  = knorm_a * I0(beta_a sqrt(1 - (u/dwin_a)^2)) / I0(beta_a),  |u| <= dwin_a
    s_a(l)  = -1 for even l (the fftshift sign), +1 otherwise

and its analytic derivative dy/dk_a = g_a * sum_l w_a'(.) prod_{b != a} w_b(.) s G_c[l],
with w'(u) = -knorm/I0(beta) * beta^2 * (I1(z)/z) * u / dwin^2, z = beta sqrt(1 - (u/dwin)^2).
Taps run over max(ceil(dk - dwin), 0) .. min(floor(dk + dwin), S - 1) (no periodic wrap,
gridFFT.cpp:1211-1271).  G_c is the oversampled grid of coil c after the pruned forward FFT,
taken from the FP64 oracle (stage_pre_fft).  knorm is the table's normalisation
300 / sum(f), which is the table's first entry (f[0] = 1).

Pinned in tests/test_autograd.py against the oracle's table-mode gather (the two
kernels differ only by the table's interpolation error) and against central
finite differences of this model.
"""
import math

import numpy as np
import scipy.special


def _i1_over_z(z):
    z = np.asarray(z, dtype=np.float64)
    out = np.full_like(z, 0.5)
    nz = z > 1e-8
    out[nz] = scipy.special.i1(z[nz]) / z[nz]
    return out


class ExactKBGather:
    def __init__(self, oracle_op):
        geo = oracle_op.geometry()
        self.S = [int(s) for s in geo["S"]]
        self.g = [float(v) for v in geo["grid"]]
        self.dwin = [float(v) for v in geo["dwin"]]
        self.beta = [float(v) for v in geo["beta"]]
        self.knorm = [float(oracle_op.table(a)[0]) for a in range(3)]
        self.NT = max(int(math.floor(2 * d)) + 1 for d in self.dwin)

    def _axis(self, a, k):
        """Per-axis tap indices (M, NT), weights and derivatives w.r.t. k (zero off-range)."""
        dk = np.asarray(k, np.float64) * self.g[a] + self.S[a] // 2
        s = np.maximum(np.ceil(dk - self.dwin[a]), 0).astype(np.int64)
        e = np.minimum(np.floor(dk + self.dwin[a]), self.S[a] - 1).astype(np.int64)
        l = s[:, None] + np.arange(self.NT)[None, :]
        valid = l <= e[:, None]
        u = dk[:, None] - l
        x = np.abs(u) / self.dwin[a]
        sq = np.sqrt(np.clip(1.0 - x * x, 0.0, None))
        i0b = scipy.special.i0(self.beta[a])
        w = self.knorm[a] * scipy.special.i0(self.beta[a] * sq) / i0b
        dw = (-self.knorm[a] / i0b * self.beta[a] ** 2 * _i1_over_z(self.beta[a] * sq)
              * u / self.dwin[a] ** 2)
        sign = np.where(l % 2 == 0, -1.0, 1.0)
        w = np.where(valid, w * sign, 0.0)
        dw = np.where(valid, dw * sign * self.g[a], 0.0)
        return np.clip(l, 0, self.S[a] - 1), w, dw

    def apply(self, grids, k):
        """grids (C, Sz, Sy, Sx) complex128, k (M, 3) -> (y (C, M), dy/dk (C, M, 3))."""
        lx, wx, dx = self._axis(0, k[:, 0])
        ly, wy, dy = self._axis(1, k[:, 1])
        lz, wz, dz = self._axis(2, k[:, 2])
        C = grids.shape[0]
        M = k.shape[0]
        y = np.zeros((C, M), np.complex128)
        dyk = np.zeros((C, M, 3), np.complex128)
        for c in range(C):
            V = grids[c][lz[:, :, None, None], ly[:, None, :, None], lx[:, None, None, :]]
            y[c] = np.einsum("mz,my,mx,mzyx->m", wz, wy, wx, V)
            dyk[c, :, 0] = np.einsum("mz,my,mx,mzyx->m", wz, wy, dx, V)
            dyk[c, :, 1] = np.einsum("mz,my,mx,mzyx->m", wz, dy, wx, V)
            dyk[c, :, 2] = np.einsum("mz,my,mx,mzyx->m", dz, wy, wx, V)
        return y, dyk
