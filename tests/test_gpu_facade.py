"""The drop-in boundary from C++ and the remaining gridFFT entry points.

* tests/cpp/facade_cg (built with the library) runs SENSE-CG the way
  recon_lib.cpp:1523-1759 writes it, a per-coil loop of gridding.backward /
  gridding.forward calls on the C++ facade (include/mrn_gridfft.hpp), and
  compares it with the batched, graph-captured solver on the same inputs.
* forward_sos (gridFFT.cpp:722-748) and a changed k_rad (the taper radius the
  reference's callers set between calls) against the FP32 oracle.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as orc
from tests.helpers import radial_problem, rel, to_dev, traj_dev

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "facade_cg")


@pytest.mark.gpu
def test_cpp_facade_cg_matches_batched_graph_cg():
    assert os.path.exists(BIN), "tests/cpp/facade_cg is not built (python -m mri_recon_amd.build)"
    r = subprocess.run([BIN, "32", "4", "600", "32", "5"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    err = float(r.stdout.strip().split("rel_err=")[1])
    assert err < 1e-4


@pytest.mark.gpu
def test_coil_forward_sos_vs_oracle():
    import torch

    from mri_recon_amd import GridFFT
    N = 24
    traj, X, smaps, d = radial_problem(N, 1, 300, 24, seed=21)
    op = GridFFT(N, ncoils=1, overgrid=1.5, form="plain")
    Xg = to_dev(np.zeros((N, N, N), np.complex64))
    t = traj_dev(traj)
    op.coil_forward_sos(Xg, to_dev(d[0]), t)
    torch.cuda.synchronize()
    o = orc.OracleOp(N, overgrid=1.5, precision="f32")
    ref = o.forward_sos(np.zeros((N, N, N), np.complex64), d[0], traj)
    got = Xg.cpu().numpy()
    assert rel(got, ref) < 1e-5
    assert np.max(np.abs(got.imag)) == 0.0


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["batched", "plain"])
def test_set_k_rad_changes_taper(form):
    from mri_recon_amd import GridFFT
    N, C = 24, 2
    traj, X, smaps, d = radial_problem(N, C, 300, 24, seed=22)
    op = GridFFT(N, ncoils=C, overgrid=1.5, form=form)
    op.set_traj(*traj_dev(traj))
    op.set_smaps(to_dev(smaps))
    for k_rad in (4.0, math.inf, 9999.0):
        op.set_k_rad(k_rad)
        got = op.forward(to_dev(d), accumulate=False).cpu().numpy()
        o = orc.OracleOp(N, overgrid=1.5, precision="f32", k_rad=k_rad)
        ref = o.kspace_to_image(d, smaps, traj)
        assert rel(got, ref) < 1e-5, k_rad
