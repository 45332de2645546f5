"""The autotuner (mrn_autotune): every candidate it can choose passes parity,
its choice is cached and a cache hit skips the timing, and a failure part way
restores the previous launch shapes with the op still usable.

Candidates per knob (one knob at a time from the defaults, each an op built
with that setting): scatter bin edge {1, 2}, samples per work item {1024,
4096, 16384} and a small value that splits tiles over work items, matrix-core
or vector-ALU tile accumulate, gather coils per lane {1, 2, 4, 8}, gather XCD
run {16, 64, 256}, FFT XCD run {1, 4, 8, 16}, four-step or Stockham FFT;
and the op's own gather form, direct or box matrix-core for the dense tiles.
"""
import json
import os

import numpy as np
import torch
import pytest

from oracle import oracle as orc
from mri_recon_amd import GridFFT, MrnError
from tests.helpers import radial_problem, rel, to_dev, traj_dev

pytestmark = pytest.mark.gpu

N, C = 32, 16
CANDIDATES = ([dict(bin_edge=b) for b in (1, 2)] + [dict(max_visits=v) for v in (64, 1024, 4096, 16384)] +
              [dict(scatter_valu=v) for v in (False, True)] + [dict(gather_coils=c) for c in (1, 2, 4, 8)] +
              [dict(xcd_run_gather=r) for r in (16, 64, 256)] + [dict(xcd_run_fft=r) for r in (1, 4, 8, 16)] +
              [dict(fft_stockham=s) for s in (False, True)] + [dict(gather_tiles=g) for g in (0, 2)])


@pytest.fixture(scope="module")
def problem():
    traj, X, smaps, d = radial_problem(N, C, 700, 32, seed=12)
    o32 = orc.OracleOp(N, overgrid=1.5, dwin=2.5, precision="f32")
    return traj, X, smaps, d, o32.image_to_kspace(X, smaps, traj), o32.kspace_to_image(d, smaps, traj)


def _op(problem, **kw):
    traj, X, smaps, d, _, _ = problem
    op = GridFFT(N, ncoils=C, overgrid=1.5, dwin=2.5, **kw)
    op.set_traj(*traj_dev(traj))
    op.set_smaps(to_dev(smaps))
    return op


def _parity(op, problem):
    traj, X, smaps, d, A32, H32 = problem
    a = op.backward(to_dev(X)).cpu().numpy()
    h = op.forward(to_dev(d), accumulate=False).cpu().numpy()
    return rel(a, A32), rel(h, H32)


@pytest.mark.parametrize("kw", CANDIDATES, ids=lambda k: "%s=%s" % next(iter(k.items())))
def test_every_candidate_passes_parity(problem, kw):
    op = _op(problem, **kw)
    ea, eh = _parity(op, problem)
    print("%s: A %.2e A^H %.2e" % (kw, ea, eh))
    assert ea <= 1e-5 and eh <= 1e-5
    op.close()


def test_autotune_cache_hit_skips_timing(problem, tmp_path):
    cache = str(tmp_path / "tune.json")
    op = _op(problem)
    t1 = op.autotune(cache=cache)
    assert t1["from_cache"] == 0 and t1["normal_ms"] > 0
    ea, eh = _parity(op, problem)
    assert ea <= 1e-5 and eh <= 1e-5
    entries = json.load(open(cache))
    assert len(entries) == 1
    (key, val), = entries.items()
    assert "N=32x32x32" in key and "C=16" in key and val["bin_edge"] == t1["bin_edge"]
    op.close()
    op2 = _op(problem)
    t2 = op2.autotune(cache=cache)
    assert t2["from_cache"] == 1
    same = lambda t: {k: v for k, v in t.items() if k not in ("from_cache", "normal_ms")}
    assert same(t2) == same(t1)
    assert abs(t2["normal_ms"] - t1["normal_ms"]) < 1e-3  # kept in whole microseconds
    ea, eh = _parity(op2, problem)
    assert ea <= 1e-5 and eh <= 1e-5
    op2.close()
    # an unreadable cache is ignored and rewritten
    open(cache, "w").write("{not json")
    op3 = _op(problem)
    assert op3.autotune(cache=cache)["from_cache"] == 0
    assert len(json.load(open(cache))) == 1
    op3.close()


def test_autotune_failure_restores_previous_shapes(problem):
    traj, X, smaps, d, A32, H32 = problem
    op = _op(problem, bin_edge=2, max_visits=1024)
    before = op.tuning()
    a0 = op.backward(to_dev(X)).cpu().numpy()
    os.environ["MRN_TUNE_FAIL_AFTER"] = "3"
    try:
        with pytest.raises(MrnError, match="injected"):
            op.autotune(cache=None)
    finally:
        del os.environ["MRN_TUNE_FAIL_AFTER"]
    assert op.tuning() == before
    a1 = op.backward(to_dev(X)).cpu().numpy()
    assert np.array_equal(a0, a1)
    h = op.forward(to_dev(d), accumulate=False).cpu().numpy()
    assert rel(h, H32) <= 1e-5
    op.close()


def test_box_gather_in_captured_cg(problem):
    """The box gather runs on a side stream (fork and join by events, two
    branches of the captured CG graph).  Three graph replays, with eager
    calls and synchronisations between them, each equal the eager CG with the
    direct gather (FP32 summation order)."""
    traj, X, smaps, d, A32, H32 = problem
    box, direct = _op(problem, gather_tiles=2), _op(problem)
    for op in (box, direct):
        op.set_kdata(to_dev(d))
    assert box.tuning()["gather_tiles"] == 2 and direct.tuning()["gather_tiles"] == 0
    ref, tref = direct.cg(4, graph=False)
    ref = ref.cpu().numpy()
    for rep in range(3):
        Xg, tg = box.cg(4, graph=True)
        e = rel(Xg.cpu().numpy(), ref)
        print("box gather, graph replay %d: CG vs eager direct %.1e" % (rep, e))
        assert e <= 1e-6 and np.allclose(tg.numpy(), tref.numpy(), rtol=1e-5, atol=1e-6 * np.abs(tref.numpy()).max())
        box.backward(to_dev(X))  # eager work between replays
        torch.cuda.synchronize()
    box.close()
    direct.close()
