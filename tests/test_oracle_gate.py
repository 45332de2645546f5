"""Pins for the oracle's gated frame selection (gating.cpp:880-1200,
recon_lib.cpp:675-776) against numpy and against properties that fix it:

* the respiratory quantile is Hyndman & Fan's method 5 (numpy "hazen");
* scaling maps the time range onto [0, frames): min -> 0, max -> frames(1-1e-9);
  retrospective ECG uses 2 x median (after dropping times >= 1.3 x that), the
  respiratory range the quantiles widened by half a frame;
* hist mode bins readouts by rank into equal groups; no sharing floors;
* selection without sharing partitions the in-range samples by frame, in
  sample order, and scales kw by 1e6 / count (or sets 1 with reset_dens);
* a flat tornado of width 1 selects the same samples as plain binning (for
  times off the frame edges), a VIPR tornado shares more at larger |k|;
* a readout with respiratory weight 0 is in no frame.
"""
import numpy as np
import pytest

from oracle import oracle as orc

RNG = np.random.default_rng(13)


def test_quantile_is_hazen():
    for n in (1, 2, 7, 100, 1001):
        v = RNG.standard_normal(n)
        for q in (0.0, 0.03, 0.25, 0.5, 0.9, 1.0):
            assert orc.quantile5(v, q) == pytest.approx(np.quantile(v, q, method="hazen"), abs=1e-12)


def test_scaling_branches():
    t = RNG.uniform(3.0, 17.0, 500)
    s, info = orc.gate_scale(orc.gate_params(10, gate="time", vs="tornado"), t)
    assert s.min() == pytest.approx(0.0, abs=1e-12) and s.max() == pytest.approx(10 * (1 - 1e-9), rel=1e-12)
    assert info[0] == t.min() and info[2] == pytest.approx((t.max() - t.min()) / 10)
    # retrospective ECG: range [0, 2 median), outliers above 1.3 x (2 median) dropped first
    ecg = np.r_[RNG.uniform(0, 800, 300), [5000.0, 6000.0]]
    s, info = orc.gate_scale(orc.gate_params(20, gate="retro_ecg", vs="tornado"), ecg)
    m1 = 2 * np.median(ecg)
    mx = 2 * np.median(ecg[ecg < 1.3 * m1])
    assert info[0] == 0.0 and info[1] == pytest.approx(np.float32(20) / mx * (1 - 1e-9), rel=1e-12)
    # respiratory: quantiles widened by half a frame on each side
    resp = RNG.standard_normal(400)
    p = orc.gate_params(5, gate="resp", vs="tornado", min_quantile=0.1, max_quantile=0.9)
    s, info = orc.gate_scale(p, resp)
    # the bounds are float members: 0.1f and 0.9f
    v0, v1 = np.quantile(resp, [float(np.float32(0.1)), float(np.float32(0.9))], method="hazen")
    fw = (v1 - v0) / 4
    assert info[0] == pytest.approx(v0 - fw / 2, abs=1e-12)
    assert info[1] == pytest.approx(5 / (v1 - v0 + fw) * (1 - 1e-9), rel=1e-12)


def test_hist_mode_ranks_and_none_floors():
    t = RNG.permutation(240).astype(float) * 0.37 + 5
    s, _ = orc.gate_scale(orc.gate_params(8, vs="hist"), t)
    ranks = np.argsort(np.argsort(t))
    assert np.array_equal(s, (ranks * 8) // 240)
    assert np.all(np.bincount(s.astype(int)) == 30)
    s2, _ = orc.gate_scale(orc.gate_params(8, vs="none"), t)
    lin = (t - t.min()) * (np.float32(8) / (t.max() - t.min()) * (1 - 1e-9))
    assert np.array_equal(s2, np.floor(lin))


def problem(R=60, xres=16, C=3):
    kx, ky, kz = [RNG.uniform(-20, 20, (R, xres)).astype(np.float32) for _ in range(3)]
    kw = RNG.uniform(0.5, 2, (R, xres)).astype(np.float32)
    kd = (RNG.standard_normal((C, R, xres)) + 1j * RNG.standard_normal((C, R, xres))).astype(np.complex64)
    return kx, ky, kz, kw, kd


def test_binning_partitions_samples():
    R, xres, C, T = 60, 16, 3, 5
    kx, ky, kz, kw, kd = problem(R, xres, C)
    gt = RNG.uniform(-0.5, 5.5, R)  # some readouts fall outside [0, T)
    p = orc.gate_params(T, vs="none")
    counts, ox, oy, oz, ow, od = orc.pregate(p, gt, kx, ky, kz, kw, kd)
    fr = np.floor(gt)
    base = 0
    for t in range(T):
        sel = np.repeat(fr == t, xres)
        n = int(sel.sum())
        assert counts[t] == n
        assert np.array_equal(ox[base:base + n], kx.ravel()[sel])
        assert np.array_equal(ow[base:base + n], np.float32(1e6 / np.float32(n)) * kw.ravel()[sel])
        blk = od[C * base:C * (base + n)].reshape(C, n)
        assert np.array_equal(blk, kd.reshape(C, -1)[:, sel])
        base += n
    assert counts.sum() == np.sum((fr >= 0) & (fr < T)) * xres
    _, _, _, _, ow1, _ = orc.pregate(orc.gate_params(T, vs="none", reset_dens=True), gt, kx, ky, kz, kw, kd)
    assert np.all(ow1 == 1.0)


def test_tornado_flat_equals_binning_and_vipr_shares_more():
    R, xres, T = 80, 12, 6
    kx, ky, kz, kw, kd = problem(R, xres, 1)
    gt = RNG.uniform(0.01, T - 0.01, R)
    gt = np.where(np.abs(gt - np.round(gt)) < 1e-3, gt + 0.01, gt)
    a = orc.pregate(orc.gate_params(T, vs="none"), gt, kx, ky, kz, kw, kd)
    b = orc.pregate(orc.gate_params(T, vs="tornado", shape="flat", wdth_low=1, wdth_high=1), gt, kx, ky, kz, kw, kd)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    v = orc.pregate(orc.gate_params(T, vs="tornado", shape="vipr", kmax=35.0), gt, kx, ky, kz, kw, kd)
    assert np.all(v[0] >= a[0]) and v[0].sum() > a[0].sum()
    # a sample at k = 0 keeps the narrow width: it is shared like the flat case
    z = np.zeros((R, xres), np.float32)
    c0 = orc.pregate(orc.gate_params(T, vs="tornado", shape="vipr", kmax=35.0), gt, z, z, z, kw, kd)[0]
    assert np.array_equal(c0, a[0])


def test_resp_weight_excludes_readouts():
    R, xres, T = 40, 8, 4
    kx, ky, kz, kw, kd = problem(R, xres, 2)
    gt = RNG.uniform(0, T, R)
    resp = np.ones(R, np.float32)
    resp[::3] = 0.0
    c = orc.pregate(orc.gate_params(T, vs="none"), gt, kx, ky, kz, kw, kd, resp=resp)[0]
    fr = np.floor(gt)
    for t in range(T):
        assert c[t] == np.sum((fr == t) & (resp > 0)) * xres
