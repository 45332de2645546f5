#!/usr/bin/env python3
"""SENSE-CG benchmark of the MI355X NUFFT/SENSE operator.

Workload (one GPU): 256^3 image, 32 coils, 3-D radial kooshball of 20000
golden-means spokes x 256 centre-out points (5.12 M samples per coil),
oversampling 1.25 (grid 320^3), Kaiser-Bessel half width 2.5 (W = 5),
complex64 data simulated by the operator from an i.i.d. CN(0,1) image plus 1%
noise.  A step is one SENSE-CG iteration (E'E over all coils + FP64 scalars +
vector updates).  `value` is the batched, hipGraph-captured form's iter/s with
every input resident in HBM; the plain per-coil GPU form (the reference's
structure) is timed on the same inputs beside it, and the CPU oracle (the C
restatement of the reference, OpenMP) on a bounded sample of coils.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("batched-graph vs plain SENSE-CG iter/s, 256³×32ch; gridding GB/s; "
          "rel err≤1e-4")
HBM_PEAK_GBPS = 8000.0
# The gather's other bound: every tap read passes the texture path (address
# unit and vector L1), one 64-byte access per clock per CU; 256 CUs at the
# 2.4 GHz peak engine clock
CU_COUNT = 256
ENGINE_GHZ = 2.4
TEX_BYTES_PER_CLK = 64.0
# The CPU leg's FFT, in the order of preference of the reference's build:
# FFTW3f (absent from the image), MKL's FFTW3 interface (libmkl_rt, its
# sequential layer per line under the oracle's OpenMP loop over lines), else
# the oracle's own mixed-radix Stockham passes
CPU_FFT_BACKEND = "oracle Stockham (radix 2/3/4/5, OpenMP over lines)"


def cpu_fft_setup(o32):
    """Switch the oracle's line transforms to MKL's FFTW3 interface when it is
    present and warm it up at the grid size (its first call initialises)."""
    global CPU_FFT_BACKEND
    import numpy as np

    from oracle import oracle as orc
    if orc.use_mkl_fft(True, "f32"):
        CPU_FFT_BACKEND = ("MKL FFTW3 interface (libmkl_rt, sequential per line, OpenMP over lines); "
                           "FFTW3f absent from the image")
        o32.fft3(np.zeros(o32.grid_shape(), np.complex64), -1)
    return CPU_FFT_BACKEND
# device kernel behind each stage of GridFFT.stage_times (as rocprofv3 names it)
KERNEL_NAMES = {
    "pre_z": "k_fft_pre_z", "fft_y": "k_fft_lines (y)", "fft_x": "k_fft_lines (x)",
    "gather": "k_gather", "scatter_zero": "k_zero_tiles", "scatter": "k_scatter_mm",
    "ifft_x": "k_fft_lines (x, inverse)", "ifft_y": "k_fft_lines (y, inverse)", "post_z": "k_fft_post_z",
}


# the config lines beside the headline, in the order the full run takes them
CONFIG_LINES = {"sense_cg_128cube_8coil": "small_cg",
                "gated_llr_fista_160cube_24coil_10frame": "gated_fista",
                "cones_gradients_192cube_16coil": "cones_gradients",
                "l1_fista_320cube_48coil": "large_l1_fista"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rocprof_summary(root):
    """Condense rocprofv3 CSV output under `root` (kernel stats and counter
    collections) into a short text table on stdout."""
    import collections
    import csv
    import glob
    out = []
    for f in sorted(glob.glob(os.path.join(root, "**", "*kernel_stats.csv"), recursive=True)):
        rows = sorted(csv.DictReader(open(f)), key=lambda r: -float(r["TotalDurationNs"]))
        out.append("%-64s %6s %11s %10s %6s" % ("kernel (kernel_stats.csv)", "calls", "avg_us", "total_ms", "pct"))
        for r in rows[:40]:
            out.append("%-64s %6s %11.1f %10.2f %6.2f" % (r["Name"][:64], r["Calls"], float(r["AverageNs"]) / 1e3,
                                                          float(r["TotalDurationNs"]) / 1e6, float(r["Percentage"])))
    for f in sorted(glob.glob(os.path.join(root, "**", "*counter_collection.csv"), recursive=True)):
        acc = collections.defaultdict(list)
        for r in csv.DictReader(open(f)):
            acc[(r["Kernel_Name"][:64], r["Counter_Name"])].append(float(r["Counter_Value"]))
        out.append("%-64s %-11s %5s %15s %15s" % ("kernel (counter_collection.csv)", "counter", "n",
                                                  "mean/dispatch", "max/dispatch"))
        for (k, c), v in sorted(acc.items(), key=lambda kv: -sum(kv[1]))[:80]:
            out.append("%-64s %-11s %5d %15.1f %15.1f" % (k, c, len(v), sum(v) / len(v), max(v)))
        # HBM counters: scale factor from the profile run's plain 1 GiB stream
        # (torch's neg kernel, which reads and writes 1048576 KiB)
        for ctr, verb in (("FETCH_SIZE", "reads"), ("WRITE_SIZE", "writes")):
            best = None
            for r in csv.DictReader(open(f)):
                if r["Counter_Name"] == ctr and "neg_kernel" in r["Kernel_Name"]:
                    v = float(r["Counter_Value"])
                    if best is None or v > best[1]:
                        best = (r.get("Dispatch_Id", "?"), v)
            if best and best[1] > 0:
                out.append("# calibration %s factor %.4f (1 GiB stream, dispatch %s, %s %.6f KiB)"
                           % (ctr, 1048576.0 / best[1], best[0], verb, best[1]))
    print("\n".join(out))


def stage_bytes(N, C, M, S):
    """Algorithmic bytes per launch of each batched stage over all C coils
    (compulsory reads + writes; 16*M for the kx, ky, kz, kw trajectory)."""
    return {
        "pre_z": 8 * N**3 + 8 * N**3 * C + 8 * N * N * S * C,
        "fft_y": 8 * N * N * S * C + 8 * N * S * S * C,
        "fft_x": 8 * N * S * S * C + 8 * S**3 * C,
        "gather": 8 * S**3 * C + 8 * M * C + 16 * M,
        "scatter_zero": 0,  # tiles shared by several work items; its bytes count in the scatter's
        "scatter": 8 * M * C + 16 * M + 8 * S**3 * C,
        "ifft_x": 8 * S**3 * C + 8 * N * S * S * C,
        "ifft_y": 8 * N * S * S * C + 8 * N * N * S * C,
        "post_z": 8 * N * N * S * C + 8 * N**3 * C + 16 * N**3,
    }


def roofline_of(st, b):
    """Dominant stage of `st` (stage -> ms) with its achieved HBM rate on the
    algorithmic bytes `b`."""
    dom = max(st, key=lambda k: st[k])
    ach = b[dom] / (st[dom] * 1e-3) / 1e9
    return {"kernel": KERNEL_NAMES[dom], "stage": dom, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBPS,
            "unit": "GB/s", "frac": ach / HBM_PEAK_GBPS, "traffic": None, "algorithmic_bytes": b[dom],
            "avg_ms": st[dom]}


def texture_path(taps_per_s):
    """The gather's tap reads (8 bytes per tap and coil, every tap of every
    sample) against the texture path's 64 bytes per clock per CU."""
    per_clk = taps_per_s * 8.0 / (CU_COUNT * ENGINE_GHZ * 1e9)
    return {"achieved": per_clk, "peak": TEX_BYTES_PER_CLK, "unit": "B/clk/CU", "frac": per_clk / TEX_BYTES_PER_CLK,
            "bytes_model": "8 * taps * coils per sample, %d CUs at %.1f GHz" % (CU_COUNT, ENGINE_GHZ)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="CPU baseline work budget")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-next", action="store_true", help="skip the DCF / LLR / gating / L1-FISTA timings")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the 128^3 x 8 CG, 160^3 x 24 x 10-frame FISTA+LLR, 192^3 x 16 cones gradient "
                         "and 320^3 x 48 L1-FISTA lines")
    ap.add_argument("--next-only", action="store_true",
                    help="only the DCF / LLR / gating / L1-FISTA timings (no CPU leg; for rocprofv3)")
    ap.add_argument("--profile", action="store_true", help="batched iterations only (for rocprofv3)")
    ap.add_argument("--gather-tiles", type=int, default=0, choices=(0, 1, 2),
                    help="gather of the dense tiles: 0 direct, 1 LDS-staged, 2 box matrix-core")
    ap.add_argument("--small", action="store_true", help="128^3 x 8 coils, 4096 x 64 radial (os 1.5, W=4)")
    ap.add_argument("--rocprof-summary", metavar="DIR", help="summarise rocprofv3 CSV output and exit")
    ap.add_argument("--config", choices=sorted(CONFIG_LINES),
                    help="only this config line (its JSON object), e.g. for rocprofv3")
    a = ap.parse_args()
    if a.rocprof_summary:
        rocprof_summary(a.rocprof_summary)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(coil sharding over RCCL is design only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from mri_recon_amd import GridFFT, synth

    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)
    if a.config:
        print(json.dumps({a.config: globals()[CONFIG_LINES[a.config]](a, log)}))
        return
    dev = torch.device("cuda")
    if a.small:
        N, C, NSP, NPT, OS, DW = 128, 8, 4096, 64, 1.5, 2.0
        workload = "SENSE-CG 128^3 x 8 coils, 3-D radial 4096x64, os 1.5, KB W=4 (grid 192^3)"
    else:
        N, C, NSP, NPT, OS, DW = 256, 32, 20000, 256, 1.25, 2.5
        workload = "SENSE-CG 256^3 x 32 coils, 3-D radial 20000x256, os 1.25, KB W=5 (grid 320^3)"
    seed = 1002
    t_setup = time.time()
    kx, ky, kz, kw = synth.radial_traj(NSP, NPT, N, device=dev)
    M = kx.numel()
    smaps = synth.coil_maps(C, N, seed=seed, device=dev)
    xtrue = synth.cn_image((N, N, N), seed=seed + 1, device=dev)
    op = GridFFT(N, ncoils=C, overgrid=OS, dwin=DW, form="batched", gather_tiles=a.gather_tiles)
    op.set_traj(kx, ky, kz, kw)
    op.set_smaps(smaps)
    data = synth.add_noise(op.backward(xtrue), 0.01, seed=seed + 2)
    op.set_kdata(data)
    torch.cuda.synchronize()
    log("setup %.1fs (S=%s, M=%d per coil)" % (time.time() - t_setup, op.S, M))
    if not a.no_autotune:
        t0 = time.time()
        op.autotune()
        log("autotune %.1fs" % (time.time() - t0))

    if a.next_only:
        a.no_cpu = True
        print(json.dumps({"next_rows": next_rows(a, op, N, kx, ky, kz, M, log)}))
        return

    # ---- timed region: batched, graph-captured CG --------------------------
    X = op.cg_init()
    op.cg_iterate(X, a.warmup, graph=True, trace=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    op.cg_iterate(X, a.steps, graph=True, trace=False)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / a.steps
    value = a.steps / (t1 - t0)
    log("batched+graph CG: %.2f ms/iter, %.2f iter/s" % (ms_step, value))
    if a.profile:
        # a plain 1 GiB stream (torch's neg kernel reads and writes 1 GiB, and
        # nothing else here runs it): its HBM counters (FETCH_SIZE, WRITE_SIZE)
        # calibrate those of the operator's kernels
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst
        st = op.stage_times(xtrue, reps=3)
        log("stages ms: " + ", ".join("%s %.2f" % (k, v) for k, v in st.items()))
        print(json.dumps({"metric": METRIC, "value": value, "unit": "iter/s", "ms_per_step": ms_step,
                          "profile": True, "stage_ms": st, "tuning": op.tuning()}))
        return

    res = {}
    # non-graph batched form, same kernels
    Xn = op.cg_init()
    op.cg_iterate(Xn, 1, graph=False, trace=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    op.cg_iterate(Xn, 3, graph=False, trace=False)
    torch.cuda.synchronize()
    res["batched_nograph_iter_s"] = 3 / (time.perf_counter() - t0)
    del Xn

    # ---- per-stage device times (HIP events on the launch stream) ---------
    st = op.stage_times(xtrue, reps=5)
    S = op.S[0]
    CC = C  # all coils per launch
    b = stage_bytes(N, C, M, S)
    stages = {k: {"ms": st[k], "GBps": b[k] / (st[k] * 1e-3) / 1e9 if b[k] else None} for k in st}
    dom = max(st, key=lambda k: st[k])
    fft_ms = sum(st[k] for k in ("pre_z", "fft_y", "fft_x"))
    ifft_ms = sum(st[k] for k in ("ifft_x", "ifft_y", "post_z"))
    scat_ms = st["scatter_zero"] + st["scatter"]
    res["gridding"] = {
        "gather_GBps": stages["gather"]["GBps"],
        "scatter_GBps": b["scatter"] / (scat_ms * 1e-3) / 1e9,
        "fft_GBps": (2 * 16 * S**3 * CC) / ((fft_ms + ifft_ms) * 1e-3) / 1e9,
        "gather_taps_per_s": M * C * (2 * DW) ** 3 / (st["gather"] * 1e-3),
        "scatter_taps_per_s": M * C * (2 * DW) ** 3 / (scat_ms * 1e-3),
        "stages": stages,
        "bytes_model": "per launch over all coils: gather/scatter 8*S^3*C + 8*M*C + 16*M; "
                       "FFT 16*S^3*C per 3-D transform; passes = compulsory reads+writes",
        "tuning": op.tuning(),
    }
    log("stages ms: " + ", ".join("%s %.2f" % (k, v) for k, v in st.items()))
    # float atomics of the scatter: split tiles' flushes and window-edge samples
    fs = op.frame_stats()
    res["atomics"] = {
        "scatter_atomics_per_launch": fs["scatter_atomics"], "scatter_ms": scat_ms,
        "scatter_atomics_per_s": fs["scatter_atomics"] / (scat_ms * 1e-3),
        "split_items": fs["split_items"], "split_tiles": fs["split_tiles"], "scatter_items": fs["scatter_items"],
        "window_edge_samples": fs["wide"], "tile_visits": fs["tile_visits"],
        "note": "upper bound: zero sums skip their atomics; the rest of the scatter writes each tile once"}
    log("scatter: %d work items (%d split, %d window-edge samples), <= %.3g atomics -> %.3g /s"
        % (fs["scatter_items"], fs["split_items"], fs["wide"], fs["scatter_atomics"],
           res["atomics"]["scatter_atomics_per_s"]))

    # ---- weighted adjoint identity on the same operator -------------------
    y = synth.cn_image((C, M), seed=seed + 3, device=dev)
    Ax = op.backward(xtrue)
    AHy = op.forward(y, accumulate=False)
    kr = (kx.double() ** 2 + ky.double() ** 2 + kz.double() ** 2)
    taper = torch.exp(-kr / (2.0 * 9999.0 ** 2))
    lhs = torch.sum(torch.conj(Ax.to(torch.complex128)) * (kw.double() * taper) * y.to(torch.complex128))
    rhs = torch.sum(torch.conj(xtrue.to(torch.complex128)) * AHy.to(torch.complex128))
    res["adjoint_rel_err"] = float(abs(lhs - rhs) / abs(lhs))
    del y, Ax, AHy, kr, taper
    log("adjoint identity rel err %.2e" % res["adjoint_rel_err"])

    # ---- plain per-coil GPU form ------------------------------------------
    if not a.no_plain:
        op_p = GridFFT(N, ncoils=C, overgrid=OS, dwin=DW, form="plain")
        op_p.set_traj(kx, ky, kz, kw)
        op_p.set_smaps(smaps)
        op_p.set_kdata(data)
        Xp = op_p.cg_init()
        op_p.cg_iterate(Xp, 1, graph=False, trace=False)
        torch.cuda.synchronize()
        log("plain warm-up iteration done")
        t0 = time.perf_counter()
        op_p.cg_iterate(Xp, a.plain_steps, graph=False, trace=False)
        torch.cuda.synchronize()
        tp = (time.perf_counter() - t0) / a.plain_steps
        res["plain_iter_s"] = 1.0 / tp
        res["speedup_vs_plain"] = value / res["plain_iter_s"]
        log("plain CG: %.1f ms/iter -> speedup %.1fx" % (tp * 1e3, res["speedup_vs_plain"]))
        # same number of iterations in both forms from the same start
        nit = 1 + a.plain_steps
        Xb = op.cg_init()
        op.cg_iterate(Xb, nit, graph=True, trace=False)
        torch.cuda.synchronize()
        res["cg_rel_err_batched_vs_plain"] = float(torch.linalg.vector_norm(Xb - Xp) /
                                                   torch.linalg.vector_norm(Xp))
        log("CG image batched vs plain after %d iters: %.2e" % (nit, res["cg_rel_err_batched_vs_plain"]))
        del op_p, Xp, Xb
        torch.cuda.empty_cache()

    # ---- CPU baseline: the oracle (C/OpenMP restatement) on a coil sample --
    cpu = None
    if not a.no_cpu:
        from oracle import oracle as orc
        o32 = orc.OracleOp(N, overgrid=OS, dwin=DW, precision="f32")
        cpu_fft_setup(o32)
        traj_h = [t.cpu().numpy() for t in (kx, ky, kz, kw)]
        P = xtrue.cpu().numpy()
        sm_h = smaps.cpu().numpy()
        t0 = time.perf_counter()
        L1 = o32.normal(P, sm_h[:1], traj_h)
        t_coil = time.perf_counter() - t0
        k = 1
        if t_coil < a.cpu_seconds / 2:
            k2 = int(min(C - 1, max(1, (a.cpu_seconds - t_coil) / t_coil)))
            t0 = time.perf_counter()
            Lk = o32.normal(P, sm_h[1:1 + k2], traj_h)
            t_more = time.perf_counter() - t0
            t_coil = (t_coil + t_more) / (1 + k2)
            k += k2
            L1 = L1 + Lk
        t_iter = t_coil * C
        # GPU E'E over the same coils, for the error against the oracle
        op_k = GridFFT(N, ncoils=k, overgrid=OS, dwin=DW, form="batched")
        op_k.set_traj(kx, ky, kz, kw)
        op_k.set_smaps(smaps[:k].contiguous())
        Lg = op_k.normal(xtrue).cpu().numpy()
        err = float(np.linalg.norm(Lg - L1) / np.linalg.norm(L1))
        del op_k
        cpu = {"value": 1.0 / t_iter, "unit": "iter/s", "cores": orc.threads(), "kind": "port",
               "fft_backend": CPU_FFT_BACKEND,
               "sample": "E'E (A then A^H) of %d of %d coils at the full 256^3 / 5.12M-sample size "
                         "on the FP32 oracle, %.2f s per coil, scaled x%d to one CG iteration "
                         "(vector updates excluded)" % (k, C, t_coil, C),
               "gpu_vs_cpu_sample_rel_err": err}
        log("CPU oracle: %.2f s/coil over %d coils -> %.4f iter/s; GPU vs oracle E'E %.2e"
            % (t_coil, k, 1.0 / t_iter, err))

    # ---- rows beside the operator: Pipe DCF and L1-wavelet FISTA ----------
    if not a.no_next:
        res["next_rows"] = next_rows(a, op, N, kx, ky, kz, M, log)

    grid_S = list(op.S)
    if not a.no_configs:
        op.close()
        del data, smaps
        torch.cuda.empty_cache()
        res["configs"] = {name: globals()[fn](a, log) for name, fn in CONFIG_LINES.items()}

    dom_b = b[dom]
    achieved = dom_b / (st[dom] * 1e-3) / 1e9
    traffic = profiled_traffic(KERNEL_NAMES[dom])
    out = {
        "metric": METRIC, "value": value, "unit": "iter/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 (fp64 CG scalars)", "data": "synthetic",
        "config": {"workload": workload, "N": N, "coils": C, "samples_per_coil": M, "grid": grid_S,
                   "overgrid": OS, "kb_half_width": DW, "form": "batched, hipGraph-captured",
                   "parallelism": "single GPU"},
        "roofline": {"kernel": KERNEL_NAMES[dom], "stage": dom, "bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS,
                     "traffic": traffic["bytes"] if traffic else None, "traffic_source": traffic,
                     "algorithmic_bytes": dom_b, "avg_ms": st[dom]},
        "cpu_baseline": cpu,
    }
    if dom == "gather":
        out["roofline"]["texture_path"] = texture_path(res["gridding"]["gather_taps_per_s"])
    out.update(res)
    print(json.dumps(out))


def small_cg(a, log):
    """SENSE-CG 128^3 x 8 coils, 4096 x 64 radial, os 1.5, KB half width 2
    (grid 192^3): batched + graph iter/s, the plain per-coil form, the
    dominant stage's roofline and the CPU oracle."""
    import numpy as np
    import torch

    from mri_recon_amd import GridFFT, synth
    N, C, NSP, NPT, OS, DW, seed = 128, 8, 4096, 64, 1.5, 2.0, 1001
    dev = torch.device("cuda")
    kx, ky, kz, kw = synth.radial_traj(NSP, NPT, N, device=dev)
    M = kx.numel()
    smaps = synth.coil_maps(C, N, seed=seed + 1, device=dev)
    xtrue = synth.ellipsoid_phantom(N, seed=seed, device=dev)
    op = GridFFT(N, ncoils=C, overgrid=OS, dwin=DW)
    op.set_traj(kx, ky, kz, kw)
    op.set_smaps(smaps)
    data = synth.add_noise(op.backward(xtrue), 0.01, seed=seed + 2)
    op.set_kdata(data)
    if not a.no_autotune:
        op.autotune()
    steps = 20
    X = op.cg_init()
    op.cg_iterate(X, 3, graph=True, trace=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    op.cg_iterate(X, steps, graph=True, trace=False)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    st = op.stage_times(xtrue, reps=5)
    out = {"workload": "SENSE-CG 128^3 x 8 coils, 3-D radial 4096x64, os 1.5, KB W=4 (grid 192^3)",
           "value": 1.0 / dt, "unit": "iter/s", "ms_per_step": dt * 1e3, "steps": steps,
           "roofline": roofline_of(st, stage_bytes(N, C, M, op.S[0])), "stage_ms": st, "tuning": op.tuning()}
    pl = GridFFT(N, ncoils=C, overgrid=OS, dwin=DW, form="plain")
    pl.set_traj(kx, ky, kz, kw)
    pl.set_smaps(smaps)
    pl.set_kdata(data)
    Xp = pl.cg_init()
    pl.cg_iterate(Xp, 1, graph=False, trace=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pl.cg_iterate(Xp, 3, graph=False, trace=False)
    torch.cuda.synchronize()
    out["plain_iter_s"] = 3 / (time.perf_counter() - t0)
    out["speedup_vs_plain"] = out["value"] / out["plain_iter_s"]
    pl.close()
    if not a.no_cpu:
        from oracle import oracle as orc
        o32 = orc.OracleOp(N, overgrid=OS, dwin=DW, precision="f32")
        cpu_fft_setup(o32)
        th = [t.cpu().numpy() for t in (kx, ky, kz, kw)]
        sm = smaps.cpu().numpy()
        P = xtrue.cpu().numpy()
        t0 = time.perf_counter()
        o32.normal(P, sm[:2], th)
        tc = (time.perf_counter() - t0) / 2
        out["cpu_baseline"] = {"value": 1.0 / (tc * C), "unit": "iter/s", "cores": orc.threads(), "kind": "port",
                               "fft_backend": CPU_FFT_BACKEND,
                               "sample": "E'E of 2 of 8 coils on the FP32 oracle, %.3f s per coil, x8 to one "
                                         "iteration (vector updates excluded)" % tc}
    log("128^3 x 8 CG: %.1f iter/s (plain %.2f, x%.1f), dominant %s %.2f ms (%.1f%% of HBM)"
        % (out["value"], out["plain_iter_s"], out["speedup_vs_plain"], out["roofline"]["stage"],
           out["roofline"]["avg_ms"], 100 * out["roofline"]["frac"]))
    op.close()
    torch.cuda.empty_cache()
    return out


def gated_fista(a, log):
    """Gated low-rank FISTA 160^3 x 24 coils x 10 frames (Poisson spoke counts,
    mean 3000 x 256), grid 240^3, local low rank on 8^3 blocks over the
    frames: ms per FISTA+LLR iteration of the graph-captured form, the
    dominant stage's roofline on frame 0, and the CPU oracle."""
    import numpy as np
    import torch

    from mri_recon_amd import GridFFT, synth
    from mri_recon_amd.llr import llr_params
    N, C, F, seed = 160, 24, 10, 1003
    dev = torch.device("cuda")
    frames = synth.gated_radial_frames(F, 3000, 256, N, seed=seed, device=dev)
    smaps = synth.coil_maps(C, N, seed=seed + 1, device=dev)
    base = synth.ellipsoid_phantom(N, seed=seed + 2, device=dev)
    blob = synth.ellipsoid_phantom(N, n_ellipsoids=2, seed=seed + 3, device=dev)
    op = GridFFT(N, ncoils=C, nframes=F)
    for f in range(F):
        op.set_traj(*frames[f], frame=f)
    op.set_smaps(smaps)
    for f in range(F):
        truth = base + (0.5 + 0.5 * math.sin(2 * math.pi * f / F)) * blob
        op.set_kdata(synth.add_noise(op.backward(truth.contiguous(), frame=f), 0.01, seed=seed + 10 + f), frame=f)
    if not a.no_autotune:
        op.autotune()
    step = 0.5 * op.power_method(8)
    op.set_fista_llr(llr_params(alpha_time=0.02, block=(8, 8, 8), block_iter=4, seed=23))
    op.fista(1, step, graph=True)
    torch.cuda.synchronize()
    tt = []
    for k in (1, 4):
        t0 = time.perf_counter()
        op.fista(k, step, graph=True)
        torch.cuda.synchronize()
        tt.append(time.perf_counter() - t0)
    per = (tt[1] - tt[0]) / 3
    Ms = [fr[0].numel() for fr in frames]
    st = op.stage_times(base, reps=3, frame=0)
    out = {"workload": "FISTA + local low rank (8^3 blocks, 4 shifted passes), 160^3 x 24 coils x 10 frames, "
                       "3-D radial Poisson(3000) x 256 spokes per frame, os 1.5, KB W=5 (grid 240^3)",
           "value": 1.0 / per, "unit": "iter/s", "ms_per_step": per * 1e3, "samples_per_frame": Ms,
           "roofline": roofline_of(st, stage_bytes(N, C, Ms[0], op.S[0])), "stage_ms_frame0": st,
           "tuning": op.tuning()}
    if not a.no_cpu:
        from oracle import oracle as orc
        o32 = orc.OracleOp(N, overgrid=1.5, precision="f32")
        cpu_fft_setup(o32)
        th = [t.cpu().numpy() for t in frames[0]]
        sm = smaps[:1].cpu().numpy()
        P = base.cpu().numpy()
        d0 = np.zeros((1, Ms[0]), np.complex64)
        t0 = time.perf_counter()
        r = o32.image_to_kspace(P, sm, th, sub=d0)
        o32.kspace_to_image(r, sm, th)
        tg = time.perf_counter() - t0
        n = 32
        Xs = (np.random.default_rng(6).standard_normal((F, 1, n, n, n)) + 0j).astype(np.complex64)
        t0 = time.perf_counter()
        orc.llr(orc.llr_params(alpha_time=0.02, block=(8, 8, 8), block_iter=4), Xs, 0, 0, 0, "f32")
        tl = (time.perf_counter() - t0) * (N / n) ** 3
        t_it = tg * C * F + tl
        out["cpu_baseline"] = {"value": 1.0 / t_it, "unit": "iter/s", "cores": orc.threads(), "kind": "port",
                               "fft_backend": CPU_FFT_BACKEND,
                               "sample": "FP32 oracle gradient (A - d, then A^H) of 1 coil of frame 0, %.2f s, "
                                         "x%d coil-frames, plus the oracle LLR step on a %d^3 x %d-frame image "
                                         "scaled by voxels (%.1f s)" % (tg, C * F, n, F, tl)}
    log("160^3 x 24 x 10 FISTA+LLR: %.1f ms/iter, dominant %s %.2f ms (%.1f%% of HBM)"
        % (per * 1e3, out["roofline"]["stage"], out["roofline"]["avg_ms"], 100 * out["roofline"]["frac"]))
    op.close()
    torch.cuda.empty_cache()
    return out


def cones_gradients(a, log):
    """Differentiation workload 192^3 x 16 coils, 8192 x 512 stack-of-cones
    (grid 288^3, exact Kaiser-Bessel): ms per evaluation of the loss
    0.5 ||W^1/2 (A(x; k) - y)||^2 with its gradients w.r.t. x and k through
    SenseEncoding (the trajectory bound anew each evaluation, as in a
    trajectory optimisation), the operator's dominant stage and the CPU
    oracle's A and A^H."""
    import numpy as np
    import torch

    from mri_recon_amd import SenseEncoding, synth
    N, C, seed = 192, 16, 1004
    kx, ky, kz, kw = synth.cones_traj(8192, 512, N, device="cuda")
    k = torch.stack([kx, ky, kz], dim=1).contiguous()
    M = k.shape[0]
    smaps = synth.coil_maps(C, N, seed=seed, device="cuda")
    x = synth.cn_image((N, N, N), seed=seed + 1, device="cuda")
    y = synth.cn_image((C, M), seed=seed + 2, device="cuda")
    w = torch.clamp(kw / kw.max(), min=0.05)
    enc = SenseEncoding(N, smaps, overgrid=1.5, kernel_mode="exact_kb")

    def evaluate():
        xt = x.clone().requires_grad_(True)
        kt = k.clone().requires_grad_(True)
        r = enc(xt, kt) - y
        (0.5 * torch.sum(w * (r.real ** 2 + r.imag ** 2))).backward()
        return xt.grad, kt.grad

    evaluate()
    torch.cuda.synchronize()
    n = 5
    t0 = time.perf_counter()
    for _ in range(n):
        evaluate()
    torch.cuda.synchronize()
    per = (time.perf_counter() - t0) / n
    t0 = time.perf_counter()
    for _ in range(n):
        enc.set_traj(k)
    torch.cuda.synchronize()
    bind = (time.perf_counter() - t0) / n
    st = enc.op.stage_times(x, reps=3)
    out = {"workload": "loss 0.5||W^1/2 (A(x;k) - y)||^2 and its gradients in x and k, 192^3 x 16 coils, "
                       "stack-of-cones 8192 x 512, os 1.5, exact KB W=5 (grid 288^3)",
           "value": 1.0 / per, "unit": "evaluations/s", "ms_per_step": per * 1e3,
           "trajectory_bind_ms": bind * 1e3,
           "roofline": roofline_of(st, stage_bytes(N, C, M, enc.op.S[0])), "stage_ms": st}
    if not a.no_cpu:
        from oracle import oracle as orc
        o32 = orc.OracleOp(N, overgrid=1.5, precision="f32")
        cpu_fft_setup(o32)
        th = [t.cpu().numpy() for t in (kx, ky, kz, kw)]
        sm = smaps[:1].cpu().numpy()
        xh = x.cpu().numpy()
        t0 = time.perf_counter()
        r = o32.image_to_kspace(xh, sm, th)
        o32.kspace_to_image(r, sm, th)
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": 1.0 / (tc * C), "unit": "evaluations/s", "cores": orc.threads(),
                               "kind": "port", "fft_backend": CPU_FFT_BACKEND,
                               "sample": "FP32 oracle A then A^H of 1 of 16 coils, %.2f s, x16, with table "
                                         "taps (the oracle has no exact-KB taps and no trajectory gradient), "
                                         "so this bounds the CPU evaluation from below" % tc}
    log("192^3 x 16 cones: %.1f ms per loss + gradients evaluation (trajectory bind %.1f ms)"
        % (per * 1e3, bind * 1e3))
    enc.close()
    torch.cuda.empty_cache()
    return out


def large_l1_fista(a, log):
    """Large L1-wavelet FISTA: 320^3 x 48 coils, 60000 x 320 radial, os 1.5
    (grid 480^3, 42 GB of coil grids), 50% fraction threshold on 3-level db4
    with 4 cycle spins: ms per graph-captured iteration, the dominant stage
    and the CPU oracle on a bounded sample."""
    import numpy as np
    import torch

    from mri_recon_amd import GridFFT, synth
    N, C, seed = 320, 48, 1005
    traj = synth.radial_traj(60000, 320, N, device="cuda")
    M = traj[0].numel()
    smaps = synth.coil_maps(C, N, seed=seed, device="cuda")
    ph = synth.ellipsoid_phantom(N, seed=seed + 1, device="cuda")
    op = GridFFT(N, ncoils=C, overgrid=1.5, dwin=2.5)
    op.set_traj(*traj)
    op.set_smaps(smaps)
    op.set_kdata(synth.add_noise(op.backward(ph.contiguous()), 0.01, seed=seed + 2))
    step = 0.5 * op.power_method(6)
    op.set_fista_l1(0.5, levels=(3, 3, 3), spins=4, seed=seed)
    op.fista(1, step, graph=True)
    torch.cuda.synchronize()
    tt = []
    for kk in (1, 3):
        t0 = time.perf_counter()
        op.fista(kk, step, graph=True)
        torch.cuda.synchronize()
        tt.append(time.perf_counter() - t0)
    per = (tt[1] - tt[0]) / 2
    st = op.stage_times(ph.contiguous(), reps=2)
    out = {"workload": "L1-wavelet FISTA (fraction 0.5, 3-level db4, 4 cycle spins), 320^3 x 48 coils, "
                       "3-D radial 60000 x 320, os 1.5, KB W=5 (grid 480^3)",
           "value": 1.0 / per, "unit": "iter/s", "ms_per_step": per * 1e3, "samples_per_coil": M,
           "roofline": roofline_of(st, stage_bytes(N, C, M, op.S[0])), "stage_ms": st}
    if not a.no_cpu:
        from oracle import oracle as orc
        o32 = orc.OracleOp(N, overgrid=1.5, dwin=2.5, precision="f32")
        cpu_fft_setup(o32)
        th = [t.cpu().numpy() for t in traj]
        sm = smaps[:1].cpu().numpy()
        P = ph.cpu().numpy()
        d0 = np.zeros((1, M), np.complex64)
        t0 = time.perf_counter()
        r = o32.image_to_kspace(P, sm, th, sub=d0)
        o32.kspace_to_image(r, sm, th)
        tg = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": 1.0 / (tg * C), "unit": "iter/s", "cores": orc.threads(), "kind": "port",
                               "fft_backend": CPU_FFT_BACKEND,
                               "sample": "FP32 oracle gradient (A - d, then A^H) of 1 of 48 coils, %.1f s, x48; "
                                         "the L1 step and the vector updates are left out" % tg}
    log("320^3 x 48 L1 FISTA: %.1f ms/iter, dominant %s %.2f ms" % (per * 1e3, out["roofline"]["stage"],
                                                                  out["roofline"]["avg_ms"]))
    op.close()
    torch.cuda.empty_cache()
    return out


def next_rows(a, op, N, kx, ky, kz, M, log):
    """Device time of the Pipe DCF (default settings, 20 iterations) on the
    bench trajectory with the oracle timed on a bounded sample beside it, and
    of one L1-wavelet FISTA iteration beside one plain FISTA iteration."""
    import numpy as np
    import torch

    from mri_recon_amd import dcf_calc, dcf_params
    from mri_recon_amd.dcf import dcf_grid_size
    out = {}
    p = dcf_params(N)
    dcf_calc(kx, ky, kz, params=p)  # first call pays the allocations
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    w = dcf_calc(kx, ky, kz, params=p)
    torch.cuda.synchronize()
    t_dcf = time.perf_counter() - t0
    from mri_recon_amd.dcf import dcf_last_atomics
    nat = dcf_last_atomics()
    dcf = {"ms": t_dcf * 1e3, "iters": p.iters, "samples": M, "grid": list(dcf_grid_size(p)),
           "ms_per_iter": t_dcf * 1e3 / p.iters, "nonzero_weights": int(torch.count_nonzero(w).item()),
           "grid_atomics_per_iter": nat / p.iters,
           "grid_atomics_per_s_whole_iteration": nat / t_dcf}
    log("Pipe DCF (%d iterations, grid %s): %.1f ms" % (p.iters, dcf["grid"], t_dcf * 1e3))
    if not a.no_cpu:
        from oracle import oracle as orc
        h = [t.cpu().numpy() for t in (kx, ky, kz)]
        # one oracle iteration on m and 2m samples: t = clear + m * per-sample
        m = max(1, M // 32)
        ts = []
        for mm in (m, 2 * m):
            t0 = time.perf_counter()
            orc.dcf_calc(orc.dcf_params(N, iters=1), h[0][:mm], h[1][:mm], h[2][:mm], "f32")
            ts.append(time.perf_counter() - t0)
        per = max(ts[1] - ts[0], 1e-9) / m
        fixed = max(ts[0] - per * m, 0.0)
        t_full = p.iters * (fixed + per * M)
        dcf["cpu_baseline"] = {"ms": t_full * 1e3, "cores": orc.threads(), "kind": "port",
                               "sample": "one FP32-oracle iteration on the first %d and %d samples, "
                                         "fitted as grid clearing + per-sample cost and scaled to %d "
                                         "samples x %d iterations" % (m, 2 * m, M, p.iters)}
        log("Pipe DCF oracle estimate: %.1f s" % t_full)
    out["dcf"] = dcf
    # L1-wavelet FISTA: per-iteration time from calls of 1 and 3 iterations
    step = op.power_method(iters=4)
    per = {}
    for frac in (0.0, 0.5):
        op.set_fista_l1(frac, spins=4, seed=1)
        op.fista(1, step)
        torch.cuda.synchronize()
        tt = []
        for k in (1, 3):
            t0 = time.perf_counter()
            op.fista(k, step)
            torch.cuda.synchronize()
            tt.append(time.perf_counter() - t0)
        per[frac] = (tt[1] - tt[0]) / 2
    op.set_fista_l1(0.0)
    out["llr"] = llr_timing(a, N, log)
    out["pregate"] = pregate_timing(a, kx, ky, kz, M, log)
    out["fista"] = {"plain_ms_per_iter": per[0.0] * 1e3, "l1_wavelet_ms_per_iter": per[0.5] * 1e3,
                    "l1": "fraction 0.5, 4 cycle spins, Daubechies-4 levels (4,4,4)"}
    log("FISTA: %.1f ms/iter plain, %.1f ms/iter with the L1 wavelet step" % (per[0.0] * 1e3, per[0.5] * 1e3))
    return out


def llr_timing(a, N, log):
    """Locally low-rank step (defaults: 4^3 blocks, 4 shifted passes) on an
    N^3 x 8-frame image, with the oracle on a 32^3 x 8 sample scaled by voxels."""
    import numpy as np
    import torch

    from mri_recon_amd import llr_threshold
    F = 8
    g = torch.Generator(device="cuda").manual_seed(5)
    X = torch.randn((F, 1, N, N, N), dtype=torch.complex64, device="cuda", generator=g)
    llr_threshold(X, alpha_time=0.05, call=0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 3
    for r in range(reps):
        llr_threshold(X, alpha_time=0.05, iteration=r, call=r + 1)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    res = {"ms": ms, "frames": F, "image": [N, N, N], "blocks": (N // 4) ** 3, "passes": 5,
           "GBps": 13 * 8 * F * N ** 3 / (ms * 1e-3) / 1e9,
           "bytes_model": "update_threshold reads the image once; each of the 4 shifted passes reads it twice "
                          "(Gram matrix, then the product) and writes it once"}
    del X
    log("LLR (%d^3 x %d frames, 4 passes): %.1f ms" % (N, F, ms))
    if not a.no_cpu:
        from oracle import oracle as orc
        n = 32
        Xs = (np.random.default_rng(6).standard_normal((F, 1, n, n, n)) + 0j).astype(np.complex64)
        t0 = time.perf_counter()
        orc.llr(orc.llr_params(alpha_time=0.05), Xs, 0, 0, 0, "f32")
        t = (time.perf_counter() - t0) * (N / n) ** 3
        res["cpu_baseline"] = {"ms": t * 1e3, "cores": orc.threads(), "kind": "port",
                               "sample": "one oracle call on a %d^3 x %d-frame image, scaled by voxels" % (n, F)}
        log("LLR oracle estimate: %.1f s" % t)
    return res


def pregate_timing(a, kx, ky, kz, M, log):
    """Frame selection of the bench trajectory into 20 frames (no sharing and
    tornado sharing) with 32 coils of k-space."""
    import numpy as np
    import torch

    from mri_recon_amd import gate_params, gate_scale_times, pregate
    R = M // 256
    C = 32
    kd = torch.zeros((C, R, 256), dtype=torch.complex64, device="cuda")
    t = np.linspace(0.0, 1.0, R)
    trj = [v.view(R, 256) for v in (kx, ky, kz)]
    kw = torch.ones((R, 256), dtype=torch.float32, device="cuda")
    res = {"readouts": R, "samples": M, "coils": C, "frames": 20}
    for vs in ("none", "tornado"):
        p = gate_params(20, vs=vs, kmax=128.0)
        gt = torch.from_numpy(gate_scale_times(p, t)[0]).cuda()
        ms = float("inf")
        for rep in range(4):  # the first call sizes the output allocations
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            n, fr = pregate(p, gt, *trj, kw, kd)
            torch.cuda.synchronize()
            if rep:
                ms = min(ms, (time.perf_counter() - t0) * 1e3)
            del fr
        kept = sum(n)
        res[vs] = {"ms": ms, "kept_samples": kept,
                   "GBps_out": kept * (16 + 8 * C) / (ms * 1e-3) / 1e9}
        log("pregate %s: %.1f ms, %d samples kept" % (vs, ms, kept))
    del kd
    if not a.no_cpu:
        from oracle import oracle as orc
        Rs = R // 20
        h = [v[:Rs].cpu().numpy() for v in trj]
        kds = (np.zeros((C, Rs, 256)) + 0j).astype(np.complex64)
        p = orc.gate_params(20, vs="none")
        gt, _ = orc.gate_scale(p, t[:Rs])
        t0 = time.perf_counter()
        orc.pregate(p, gt, *h, np.ones_like(h[0]), kds)
        tc = (time.perf_counter() - t0) * 20
        res["cpu_baseline"] = {"ms": tc * 1e3, "cores": 1, "kind": "port",
                               "sample": "oracle selection of %d of %d readouts (no sharing), x20" % (Rs, R)}
        log("pregate oracle estimate: %.1f s" % tc)
    return res


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summaries (FETCH_SIZE, WRITE_SIZE in KiB per dispatch), each scaled by the
    same counter's reading for a plain 1 GiB stream; None if the summaries do
    not cover the kernel."""
    here = os.path.dirname(os.path.abspath(__file__))
    base = os.path.join(here, "profiles", "sense_cg_256cube_32coil_")
    vals, cal = {}, {}
    name = kernel.split(" ")[0]
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        path = base + ctr.lower() + ".txt"
        if not os.path.exists(path):
            return None
        for line in open(path):
            if line.startswith("# calibration") and ctr in line:
                cal[ctr] = float(line.split("factor")[1].split()[0])
            f = line.split()
            if ctr in f and name in line and "(" in line:
                i = f.index(ctr)
                vals[ctr] = float(f[i + 2])  # mean per dispatch, KiB
    if len(vals) < 2 or len(cal) < 2:
        return None
    by = 1024.0 * (vals["FETCH_SIZE"] * cal["FETCH_SIZE"] + vals["WRITE_SIZE"] * cal["WRITE_SIZE"])
    return {"bytes": by, "fetch_kib": vals["FETCH_SIZE"], "write_kib": vals["WRITE_SIZE"],
            "calibration": cal, "files": ["profiles/sense_cg_256cube_32coil_fetch_size.txt",
                                          "profiles/sense_cg_256cube_32coil_write_size.txt"]}


if __name__ == "__main__":
    main()
