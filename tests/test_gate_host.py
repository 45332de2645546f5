"""The library's host-side gate-time scaling (mrn_gate_scale_times, no GPU
needed) against the oracle's restatement of GATING::init_time_resolved
(gating.cpp:880-1078), and the -gating flags."""
import numpy as np
import pytest

from oracle import oracle as orc
from mri_recon_amd import MrnError, gate_params, gate_scale_times

RNG = np.random.default_rng(17)


@pytest.mark.parametrize("gate,vs,extra", [("time", "tornado", {}), ("time", "none", {}), ("ecg", "hist", {}),
                                           ("retro_ecg", "tornado", {}),
                                           ("retro_ecg", "none", {"ecg_outlier_reject": False}),
                                           ("resp", "tornado", {"min_quantile": 0.05, "max_quantile": 0.95}),
                                           ("prep", "hist", {})])
def test_scaling_matches_oracle(gate, vs, extra):
    t = np.r_[RNG.uniform(0, 900, 777), [4000.0]] if "ecg" in gate else RNG.standard_normal(777) * 3 + 10
    got, info = gate_scale_times(gate_params(12, gate=gate, vs=vs, **extra), t)
    ref, rinfo = orc.gate_scale(orc.gate_params(12, gate=gate, vs=vs, **extra), t)
    assert np.array_equal(got, ref)
    assert np.allclose(info, rinfo, rtol=0, atol=0)


def test_hist_ties_keep_readout_order():
    t = np.repeat(np.arange(10.0), 6)  # every time six times
    got, _ = gate_scale_times(gate_params(4, vs="hist"), t)
    ranks = np.argsort(t, kind="stable").argsort()
    assert np.array_equal(got, (ranks * 4) // 60)


def test_flags_and_errors():
    p = gate_params(1, args=["-viewshare_type", "tornado", "-gating_type", "resp", "-vs_wdth_low", "2",
                             "-vs_wdth_high", "6", "-rcframes", "15", "-gate_min_quantile", "0.1",
                             "-wo_ecg_out_reject"])
    assert (p.vs_type, p.gate_type, p.wdth_low, p.wdth_high, p.frames, p.ecg_outlier_reject) == (0, 3, 2, 6, 15, 0)
    with pytest.raises(MrnError):
        gate_params(1, args=["-viewshare_type", "sideways"])
    with pytest.raises(MrnError, match="range"):
        gate_scale_times(gate_params(4), np.ones(10))
