"""Golden fixtures (tests/golden/, written by make_golden.py from the oracle).

CPU: the files match their SHA-256 manifest and the oracle still reproduces
them (the FP32 geometry bit for bit, the FP64 end-to-end vectors to rounding).
GPU: the HIP operator reproduces the tables and windows of every benchmark
configuration and the tiny end-to-end A, A^H, E'E and 5-iteration CG vectors
(relative L2 <= 1e-5 against FP64; the bar for the product is 1e-4 against
FP64 and 1e-5 against the FP32 restatement).
"""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as orc

from tests.helpers import rel

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def _configs():
    g = _load("geometry.npz")
    names = sorted({k.split("/")[0] for k in g.files})
    return [(n, g) for n in names]


def test_manifest_sha256():
    with open(os.path.join(GOLD, "MANIFEST.json")) as f:
        man = json.load(f)
    assert set(man) == {"geometry.npz", "tiny_e2e.npz"}
    for fn, digest in man.items():
        with open(os.path.join(GOLD, fn), "rb") as f:
            assert hashlib.sha256(f.read()).hexdigest() == digest, fn


@pytest.mark.parametrize("name,g", _configs(), ids=lambda v: v if isinstance(v, str) else "")
def test_oracle_geometry_reproduces_golden(name, g):
    N, os_, dwin = g[name + "/params"]
    op = orc.OracleOp(int(N), overgrid=float(os_), dwin=float(dwin), precision="f32")
    geo = op.geometry()
    assert np.array_equal(geo["S"], g[name + "/S"])
    assert np.array_equal(geo["og"], g[name + "/og"])
    assert np.array_equal(geo["L"], g[name + "/L"])
    assert np.allclose(geo["beta"], g[name + "/beta"], rtol=0, atol=1e-6)
    assert np.array_equal(op.table(0).astype(np.float32), g[name + "/table"])
    assert np.array_equal(op.window(0).astype(np.float32), g[name + "/window"])


def _tiny_traj(t):
    return [t["kx"], t["ky"], t["kz"], t["kw"]]


def test_oracle_tiny_reproduces_golden():
    t = _load("tiny_e2e.npz")
    N, C, _, _, os_, dwin, iters, _ = t["params"]
    op = orc.OracleOp(int(N), overgrid=float(os_), dwin=float(dwin), precision="f64")
    traj = _tiny_traj(t)
    X = t["X"].astype(np.complex128)
    assert rel(op.image_to_kspace(X, t["smaps"], traj), t["A"]) < 1e-12
    assert rel(op.kspace_to_image(t["d"].astype(np.complex128), t["smaps"], traj), t["AH"]) < 1e-12
    assert rel(op.normal(X, t["smaps"], traj), t["EE"]) < 1e-12
    Xcg, tr = op.cg(t["smaps"], t["d"].astype(np.complex128), traj, int(iters))
    assert rel(Xcg, t["cg_X"]) < 1e-12
    gtr = t["cg_trace"]
    assert np.allclose(tr[:, [0, 3]], gtr[:, [0, 3]], rtol=1e-12, atol=0)
    # alpha is real up to rounding for this Hermitian system: compare it as a complex number
    a_o = tr[:, 1] + 1j * tr[:, 2]
    a_r = gtr[:, 1] + 1j * gtr[:, 2]
    assert np.max(np.abs(a_o - a_r) / np.abs(a_r)) < 1e-12
    # the FP32 restatement sits within FP32 rounding of the FP64 vectors
    o32 = orc.OracleOp(int(N), overgrid=float(os_), dwin=float(dwin), precision="f32")
    assert rel(o32.image_to_kspace(t["X"], t["smaps"], traj), t["A"]) < 1e-5
    assert rel(o32.kspace_to_image(t["d"], t["smaps"], traj), t["AH"]) < 1e-5


# ------------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("name,g", _configs(), ids=lambda v: v if isinstance(v, str) else "")
def test_gpu_geometry_matches_golden(name, g):
    from mri_recon_amd import GridFFT
    N, os_, dwin = g[name + "/params"]
    op = GridFFT(int(N), ncoils=1, overgrid=float(os_), dwin=float(dwin))
    assert tuple(op.S) == tuple(int(s) for s in g[name + "/S"])
    assert tuple(op.og) == tuple(int(s) for s in g[name + "/og"])
    assert np.allclose(op.beta, g[name + "/beta"], rtol=1e-6, atol=0)
    tab = op.table(0).numpy()
    win = op.window(0).numpy()
    gt, gw = g[name + "/table"], g[name + "/window"]
    assert tab.shape == gt.shape and win.shape == gw.shape
    assert np.max(np.abs(tab - gt)) <= 2e-7 * np.max(np.abs(gt))
    assert np.max(np.abs(win - gw)) <= 2e-6
    op.close()


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["batched", "plain"])
def test_gpu_tiny_matches_golden(form):
    import torch

    from mri_recon_amd import GridFFT

    from tests.helpers import to_dev, traj_dev
    t = _load("tiny_e2e.npz")
    N, C, _, _, os_, dwin, iters, _ = t["params"]
    op = GridFFT(int(N), ncoils=int(C), overgrid=float(os_), dwin=float(dwin), form=form)
    op.set_traj(*traj_dev(_tiny_traj(t)))
    op.set_smaps(to_dev(t["smaps"]))
    X = to_dev(t["X"])
    d = to_dev(t["d"])
    A = op.backward(X).cpu().numpy()
    AH = op.forward(d, accumulate=False).cpu().numpy()
    EE = op.normal(X).cpu().numpy()
    assert rel(A, t["A"]) < 1e-5
    assert rel(AH, t["AH"]) < 1e-5
    assert rel(EE, t["EE"]) < 1e-5
    op.set_kdata(d)
    Xcg, tr = op.cg(int(iters), graph=(form == "batched"))
    torch.cuda.synchronize()
    assert rel(Xcg[0].cpu().numpy(), t["cg_X"]) < 1e-4
    gtr = t["cg_trace"]
    assert np.allclose(tr[:, 0].numpy(), gtr[:, 0], rtol=1e-4)
    assert np.allclose(tr[:, 3].numpy(), gtr[:, 3], rtol=1e-4)
    a_g = tr[:, 1].numpy() + 1j * tr[:, 2].numpy()
    a_r = gtr[:, 1] + 1j * gtr[:, 2]
    assert np.max(np.abs(a_g - a_r) / np.abs(a_r)) < 1e-4
    op.close()
