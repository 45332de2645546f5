"""mrn_llr_noalloc (the locally low-rank step with its scratch in a fixed
device array) against mrn_llr through the C ABI: two calls in a row, the
second on a smaller image, give the same bits and the same smax as the
allocating entry point; sizes beyond the fixed scratch are rejected."""
import ctypes

import numpy as np
import pytest
import torch

from mri_recon_amd import llr_params
from mri_recon_amd._ffi import MRN_EINVAL, check, lib

pytestmark = pytest.mark.gpu


def _frames(Nt, n, seed):
    rng = np.random.default_rng(seed)
    base = rng.standard_normal((2, n, n, n)) + 1j * rng.standard_normal((2, n, n, n))
    w = rng.standard_normal((Nt, 2))
    X = np.einsum("tr,rzyx->tzyx", w, base) + 0.1 * (rng.standard_normal((Nt, n, n, n)) +
                                                     1j * rng.standard_normal((Nt, n, n, n)))
    return torch.from_numpy(X.astype(np.complex64)[:, None]).cuda().contiguous()


def _call(fn, p, X, it, call):
    Nt, Ne, Nz, Ny, Nx = X.shape
    N = (ctypes.c_int32 * 3)(Nx, Ny, Nz)
    smax = ctypes.c_float()
    rc = fn(ctypes.byref(p), ctypes.c_void_p(X.data_ptr()), Nt, Ne, N, 0, it, call, ctypes.byref(smax),
            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return rc, smax.value


def test_noalloc_matches_allocating_entry_point():
    L = lib()
    p = llr_params(alpha_time=0.05, block=(4, 4, 4), block_iter=2, seed=3)
    for it, (Nt, n) in enumerate([(8, 64), (5, 24)]):  # the second call is smaller
        X = _frames(Nt, n, seed=10 + it)
        Xa, Xb = X.clone(), X.clone()
        rca, sa = _call(L.mrn_llr, p, Xa, it, it)
        rcb, sb = _call(L.mrn_llr_noalloc, p, Xb, it, it)
        check(rca)
        check(rcb)
        torch.cuda.synchronize()
        assert sa == sb and sa > 0
        assert torch.equal(Xa, Xb)
        assert not torch.equal(Xa, X)  # the threshold changed the image


def test_noalloc_rejects_sizes_beyond_its_scratch():
    L = lib()
    p = llr_params(alpha_time=0.05, block=(1, 1, 1), block_iter=1)
    X = torch.zeros((2, 1, 128, 128, 128), dtype=torch.complex64, device="cuda")  # 2M blocks of 1 voxel
    rc, _ = _call(L.mrn_llr_noalloc, p, X, 0, 0)
    assert rc == MRN_EINVAL
    assert b"scratch" in L.mrn_last_error(None)
