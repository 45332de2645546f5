"""FISTA with the Cauchy step (-iterative_step_type cauchy,
recon_lib.cpp:2013-2041, 2093-2100) against the oracle's loop: the complex
step of each iteration, its hold after cauchy_update_number updates, and the
image; graph and no graph give the same bits."""
import numpy as np
import pytest

from oracle import oracle as orc
from mri_recon_amd import GridFFT, MrnError
from tests.helpers import radial_problem, rel, to_dev, traj_dev

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("C", [2, 16])
def test_fista_cauchy_matches_oracle(C):
    N = 16
    traj, X, smaps, d = radial_problem(N, C, 160, 16, seed=21)
    op = GridFFT(N, ncoils=C, overgrid=1.5)
    op.set_traj(*traj_dev(traj))
    op.set_smaps(to_dev(smaps))
    op.set_kdata(to_dev(d))
    op.set_fista_step("cauchy", 3)
    outs = []
    for graph in (True, False):
        Xg, res = op.fista(6, 0.0, graph=graph)
        outs.append((Xg[0].cpu().numpy(), res.numpy(), op.fista_steps(6).numpy()))
    assert np.array_equal(outs[0][0], outs[1][0])
    Xg, res, steps = outs[0]
    o32 = orc.OracleOp(N, overgrid=1.5, precision="f32")
    Xo, ro, so = o32.fista_cauchy(smaps, d, traj, 6, 0.0, 3)
    print("C=%d: image %.2e, steps %s vs %s" % (C, rel(Xg, Xo), steps[:4], so[:4]))
    assert np.allclose(steps, so, rtol=1e-4, atol=0)
    assert steps[3] == steps[2] and steps[5] == steps[2]  # held after 3 updates
    assert rel(Xg, Xo) <= 1e-4 and np.allclose(res, ro, rtol=1e-4)
    # back to the fixed step
    op.set_fista_step("maxeig")
    Xf, _ = op.fista(3, 1e-3)
    Xp, _ = o32.fista(smaps, d, traj, 3, 1e-3)
    assert rel(Xf[0].cpu().numpy(), Xp) <= 1e-5
    with pytest.raises(MrnError):
        op.set_fista_step("cauchy", -1)
