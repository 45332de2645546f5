"""Pins for the oracle's locally low-rank step (clear.cpp:233-556), checked
against LAPACK (numpy.linalg.svd) and against properties that fix the block
geometry, not against the restated loops:

* the one-sided Jacobi SVD gives LAPACK's singular values (tall, wide,
  rank-deficient);
* on a volume that is one block, a pass equals numpy's singular value soft
  threshold of the (voxels x frames) matrix;
* tau = 0 is the identity for any shift; a huge tau zeroes exactly the voxels
  the blocks cover (the rest, when a size is not a multiple of the block, is
  untouched);
* a shifted pass is the unshifted pass of the circularly shifted volume;
* update_threshold's block maxima are LAPACK's largest singular values, its
  median is over the nonzero ones (mean of the middle two for an even count),
  scaled by 1 + amp*exp(-iter/t2) in float;
* dim = 1 (encodes) is dim = 0 with the two frame indices swapped;
* the block shifts lie in [-b/2, b - 1 - b/2], depend on the seed and the call,
  and are 0 when block_iter is 1.
"""
import numpy as np
import pytest

from oracle import oracle as orc

RNG = np.random.default_rng(5)


def crandn(*shape):
    return (RNG.standard_normal(shape) + 1j * RNG.standard_normal(shape)) / np.sqrt(2)


@pytest.mark.parametrize("shape,rank", [((64, 10), None), ((8, 20), None), ((64, 12), 3), ((128, 32), None)])
def test_svd_matches_lapack(shape, rank):
    A = crandn(*shape)
    if rank:
        A = crandn(shape[0], rank) @ crandn(rank, shape[1])
    s = orc.svd_values(A)
    ref = np.linalg.svd(A, compute_uv=False)
    k = min(shape)
    assert np.allclose(s[:k], ref, rtol=1e-12, atol=1e-12 * ref[0])
    assert np.all(np.abs(s[k:]) < 1e-12 * ref[0])


def svt(A, tau):
    U, s, Vh = np.linalg.svd(A, full_matrices=False)
    return (U * np.maximum(s - tau, 0)) @ Vh


@pytest.mark.parametrize("Nt,Ne,dim", [(6, 1, 0), (5, 2, 0), (3, 4, 1)])
def test_single_block_pass_is_lapack_svt(Nt, Ne, dim):
    X = crandn(Nt, Ne, 4, 4, 4)
    tau = 1.7
    got = orc.llr_pass(X, dim, (0, 0, 0), tau)
    if dim == 0:  # columns t, rows (e, voxel)
        A = X.reshape(Nt, Ne * 64).T
        ref = svt(A, tau).T.reshape(X.shape)
    else:  # columns e, rows (t, voxel)
        A = X.transpose(1, 0, 2, 3, 4).reshape(Ne, Nt * 64).T
        ref = svt(A, tau).T.reshape(Ne, Nt, 4, 4, 4).transpose(1, 0, 2, 3, 4)
    assert np.allclose(got, ref, rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("shift", [(0, 0, 0), (-2, 1, -1), (1, -2, 0)])
def test_zero_tau_identity_and_huge_tau_zeroes_covered(shift):
    X = crandn(4, 1, 10, 8, 9)  # 10 x 8 x 9 voxels (z, y, x): x and z not multiples of 4
    assert np.allclose(orc.llr_pass(X, 0, shift, 0.0), X, atol=1e-12)
    Z = orc.llr_pass(X, 0, shift, 1e9)
    cov = np.zeros((10, 8, 9), bool)
    for z in range(8):
        for y in range(8):
            for x in range(8):
                cov[(z + shift[2]) % 10, (y + shift[1]) % 8, (x + shift[0]) % 9] = True
    assert np.all(Z[:, :, cov] == 0)
    assert np.array_equal(Z[:, :, ~cov], X[:, :, ~cov])


def test_shift_is_circular():
    X = crandn(5, 1, 8, 8, 12)
    sh = (1, -2, -1)  # (x, y, z)
    got = orc.llr_pass(X, 0, sh, 0.8)
    rolled = np.roll(X, shift=(-sh[2], -sh[1], -sh[0]), axis=(2, 3, 4))
    ref = np.roll(orc.llr_pass(rolled, 0, (0, 0, 0), 0.8), shift=(sh[2], sh[1], sh[0]), axis=(2, 3, 4))
    assert np.allclose(got, ref, atol=1e-12)


def test_block_maxima_and_median():
    X = crandn(6, 1, 8, 4, 8)
    X[:, :, 4:, :, 4:] = 0  # one zero block (z >= 4, x >= 4): excluded from the median
    sb = orc.llr_block_smax(X, 0)
    assert sb.size == 4
    ref = []
    for i in range(2):  # block order: x block outer, z block inner
        for k in range(2):
            A = X[:, 0, 4 * k:4 * k + 4, :, 4 * i:4 * i + 4].reshape(6, 64).T
            ref.append(np.linalg.svd(A, compute_uv=False)[0] if np.any(A) else 0.0)
    assert np.allclose(sb, ref, rtol=1e-12)
    nz = np.sort([v for v in ref if v != 0])
    assert abs(orc.median_nonzero(sb) - np.median(nz)) < 1e-12
    assert orc.median_nonzero([0.0, 0.0]) == 0.0
    p = orc.llr_params(alpha_time=0.1, amp=4.0, t2=6.0)
    for it in (0, 3):
        amp = np.float32(1) + np.float32(4.0) * np.exp(np.float32(-it) / np.float32(6.0), dtype=np.float32)
        want = np.float32(np.median(nz) * np.float64(np.float32(amp)))
        assert orc.llr_update_threshold(p, X, 0, it) == pytest.approx(float(want), rel=1e-6)


def test_encode_dim_is_time_dim_transposed():
    X = crandn(3, 5, 4, 8, 4)
    p = orc.llr_params(alpha_time=0.2, alpha_encode=0.2, block_iter=2, seed=4)
    a, sa = orc.llr(p, X, 1, 2, 7)
    b, sb = orc.llr(p, X.transpose(1, 0, 2, 3, 4), 0, 2, 7)
    assert sa == pytest.approx(sb, rel=1e-12)
    assert np.allclose(a, b.transpose(1, 0, 2, 3, 4), atol=1e-12)


def test_llr_off_when_alpha_zero_and_shrinks_rank():
    X = crandn(8, 1, 8, 8, 8)
    p0 = orc.llr_params(alpha_time=0.0)
    Y, smax = orc.llr(p0, X, 0, 0, 0)
    assert np.array_equal(Y, X) and smax > 0
    # one unshifted pass: every block's singular values become max(s - tau, 0),
    # tau = alpha * smax (block_iter 1), and the noise above rank 2 goes
    base = np.einsum("tr,rzyx->tzyx", crandn(8, 2), crandn(2, 8, 8, 8))[:, None]
    Xn = base + 0.05 * crandn(8, 1, 8, 8, 8)
    Y, smax = orc.llr(orc.llr_params(alpha_time=0.05, block_iter=1), Xn, 0, 20, 0)
    tau = np.float64(np.float32(0.05)) * smax
    for z0, y0, x0 in [(0, 0, 0), (4, 0, 4), (0, 4, 4)]:
        blk = (slice(None), 0, slice(z0, z0 + 4), slice(y0, y0 + 4), slice(x0, x0 + 4))
        s_in = np.linalg.svd(Xn[blk].reshape(8, 64).T, compute_uv=False)
        s_out = np.linalg.svd(Y[blk].reshape(8, 64).T, compute_uv=False)
        assert np.allclose(s_out, np.maximum(s_in - tau, 0), atol=1e-9 * s_in[0])
        assert s_out[2] == pytest.approx(0, abs=1e-9 * s_in[0])


def test_shift_stream():
    p = orc.llr_params(seed=9, block_iter=4)
    sh = np.array([orc.llr_shift(p, (4, 4, 4), c, k) for c in range(20) for k in range(4)])
    assert sh.min() >= -2 and sh.max() <= 1 and len(np.unique(sh)) == 4
    assert np.array_equal(sh[0], orc.llr_shift(p, (4, 4, 4), 0, 0))
    assert not np.array_equal(sh[:8], np.array([orc.llr_shift(orc.llr_params(seed=10, block_iter=4), (4, 4, 4), c, k)
                                                for c in range(2) for k in range(4)]))
    assert np.all(orc.llr_shift(orc.llr_params(seed=9, block_iter=1), (4, 4, 4), 3, 0) == 0)
    odd = np.array([orc.llr_shift(p, (3, 5, 2), c, 0) for c in range(40)])
    assert odd[:, 0].min() >= -1 and odd[:, 0].max() <= 1 and odd[:, 1].min() >= -2 and odd[:, 1].max() <= 2
