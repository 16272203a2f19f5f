#!/usr/bin/env python3
"""Benchmark of the split-step hot path on one MI355X.

Workload (default): the Talbot-Lau three-grating chain at 46 keV, one
simulation per step: plane wave at G0, G0 Au / G1 Si / G2 Au gratings
(pitch 4.2 um, height 60 um, 64 slices each), 0.2 m start to start, detector
at the end of G2; N = 2^26 samples of complex64 with FP64 phase arguments,
dx = 2 nm, 1 um pixels over 1 mm.  One simulation is 194 propagation steps of
the reference (3 x 64 slices + 2 hops).

Prints one JSON line: sims/hour (value), Gsample-steps/s, HBM roofline of the
dominant pass kernel (HIP events on the library's stream, algorithmic bytes
16 B per sample per pass), and the complex128 oracle timed on a bounded
sample on the host.
"""
import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "rave-sim_amd"))

HBM_PEAK_GBS = 8000.0
WL_46 = None


def cfg2_members(ravesim, N, slices=64, plain=False):
    wl = ravesim.energy_to_wavelength(46000.0)
    h, p = 60e-6, 4.2e-6
    g0 = ravesim.Grating(pitch=p, dc=(0.5, 0.5), z_start=0.0, thickness=h, nr_steps=slices,
                         db_a=complex(1.1e-6, 4.0e-8))
    g1 = ravesim.Grating(pitch=p, dc=(0.5, 0.5), z_start=0.2, thickness=h, nr_steps=slices,
                         db_a=complex(2.6e-7, 1.0e-10))
    g2 = ravesim.Grating(pitch=p, dc=(0.5, 0.5), z_start=0.4, thickness=h, nr_steps=slices,
                         db_a=complex(1.1e-6, 4.0e-8))
    m = ravesim.core.BatchMember(wl=wl, gratings=[g0, g1, g2], cutoff_angles=[math.pi / 2] * 4,
                                 plane_wave=True, source_z=0.0)
    return [m], 0.4 + h, 3 * slices + 2


def cfg3_members(ravesim, N, n_energies=64, slices=32):
    """Polychromatic batch: 64 energies 20-70 keV on the single-grating chain
    (Si phase grating p 4.2 um, h 30 um, 32 slices, 0.1 m hop), weights of a
    70 kVp tungsten spectrum (Kramers (70-E)/E times 3 mm Al, mu ~ E^-3,
    0.5/cm at 50 keV), delta ~ E^-2 and beta ~ E^-4 from the 25 keV values."""
    import numpy as np
    e = np.linspace(20.0, 70.0, n_energies)
    w = (70.0 - e) / e * np.exp(-0.5 * (e / 50.0) ** -3 * 0.3)
    w = np.clip(w, 0, None)
    w = w / w.sum()
    h = 30e-6
    ms = []
    for ek, wk in zip(e, w):
        sc = 25.0 / ek
        g = ravesim.Grating(pitch=4.2e-6, dc=(0.5, 0.5), z_start=0.0, thickness=h, nr_steps=slices,
                            db_a=complex(7.7e-7 * sc ** 2, 3.5e-10 * sc ** 4))
        ms.append(ravesim.core.BatchMember(wl=ravesim.energy_to_wavelength(ek * 1000.0), gratings=[g],
                                           cutoff_angles=[math.pi / 2] * 2, weight=float(wk), plane_wave=True))
    return ms, h + 0.1, slices + 1


def batch_vs_plain(ravesim, ctx, lgN=24, reps=2):
    """Batched fused complex64 form vs the plain complex128 form run one
    simulation at a time, on the polychromatic batch (sims/hour each)."""
    import numpy as np
    from ravesim.core import run_batch
    N = 2 ** lgN
    ms, zdet, steps = cfg3_members(ravesim, N)
    kw = dict(det_size=1e-3, det_px=1e-6, det_py=1.0, z_detector=zdet)
    tuned = ctx.autotune(N, np.complex64, members=len(ms))
    run_batch(ctx, N, 2e-9, ms[:2], dtype=np.complex64, **kw)
    rep_s = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fused = run_batch(ctx, N, 2e-9, ms, dtype=np.complex64, **kw)
        rep_s.append(time.perf_counter() - t0)
    tf = sum(rep_s) / reps
    # the baseline at the faster of its default split and the tuned one
    plain_runs = {}
    for label in ("default", "tuned"):
        if label == "tuned":
            plain_runs["tuned_split"] = ctx.autotune(N, np.complex128, members=1)
        run_batch(ctx, N, 2e-9, ms[:1], dtype=np.complex128, plain=True, max_resident=1, **kw)
        t0 = time.perf_counter()
        plain = run_batch(ctx, N, 2e-9, ms, dtype=np.complex128, plain=True, max_resident=1, **kw)
        plain_runs[label] = time.perf_counter() - t0
    tp = min(plain_runs["default"], plain_runs["tuned"])
    err = float(np.linalg.norm(fused - plain) / np.linalg.norm(plain))
    return {"workload": f"64 energies x N=2^{lgN}, {steps} steps each, weighted detector sum",
            "fused_c64_batched_sims_per_hour": round(len(ms) / tf * 3600, 1),
            "plain_c128_one_at_a_time_sims_per_hour": round(len(ms) / tp * 3600, 1),
            "speedup": round(tp / tf, 2), "rel_l2_fused_vs_plain_c128": err, "split": tuned,
            "fused_rep_s": [round(x, 4) for x in rep_s], "plain_s": round(tp, 4),
            "plain_s_by_split": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in plain_runs.items()}}


def serving_run(ravesim, ctx, qps, duration_s, seed=7, sizes=(22, 24, 26), max_energies=4):
    """Poisson arrivals of Talbot-Lau phase-stepping requests (8 G2 positions,
    3 x 64 slices) at a fixed rate; N log-uniform over `sizes`, 1..max_energies
    energies uniform on 20-70 keV.  Returns latency percentiles and throughput."""
    import numpy as np
    from ravesim.serve import Server, phase_sweep
    rng = np.random.default_rng(seed)
    arrivals, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / qps)
        if t > duration_s:
            break
        n = 2 ** int(rng.choice(sizes))
        ne = int(rng.integers(1, max_energies + 1))
        arrivals.append((t, phase_sweep(n, list(rng.uniform(20e3, 70e3, ne)), n_steps=8, slices=64)))
    srv = Server(ctx)
    t0 = time.perf_counter()
    ids = []
    for at, req in arrivals:
        dt = at - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        ids.append((srv.submit(req), len(req.energies_ev)))
    res = [srv.result(i, timeout_s=600) for i, _ in ids]
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    lat = np.array([r.latency_s for r in res])
    sims = sum(ne for _, ne in ids)
    return {"qps": qps, "requests": len(res), "sims": sims, "duration_s": round(wall, 2),
            "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
            "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
            "sims_per_hour": round(sims / wall * 3600, 1),
            "mix": f"N in 2^{list(sizes)}, 1-{max_energies} energies, 8 phase steps, 650 steps per energy",
            "jobs_per_launch": round(st["jobs_per_launch"], 2), "max_resident_jobs": st["max_resident_jobs"]}


def cpu_baseline(N, budget_s=25.0):
    """Oracle (complex128 numpy/scipy restatement, 1 scipy worker) on whole
    grating slices of G0 at the benchmark N; scaled to sims/hour."""
    import numpy as np
    from oracle import bigwave as bw
    wl = bw.convert_energy_wavelength(46000.0)
    dx = 2e-9
    p = bw.SimParams(N=N, dx=dx, z_detector=0.4 + 60e-6, detector_size=1e-3,
                     detector_pixel_size_x=1e-6, detector_pixel_size_y=1.0, wl=wl, chunk_size=N)
    g = bw.Grating(pitch=4.2e-6, dc=(0.5, 0.5), z_start=0.0, thickness=60e-6 / 64, nr_steps=1,
                   x_positions=np.array([0.0]), substrate_thickness=0.0, db_a=complex(1.1e-6, 4.0e-8))
    f = bw.Fields(np.ones(N, dtype=np.complex128))
    n, t0 = 0, time.perf_counter()
    while True:
        bw.grating_apply(f, g, p, 1e30, 0)
        n += 1
        el = time.perf_counter() - t0
        if el > budget_s or n >= 8:
            break
    per_step = el / n
    return {"value": 3600.0 / (194 * per_step), "unit": "sims/hour", "cores": 1, "kind": "port",
            "sample": f"{n} grating slice(s) (mask + FFT, propagator, cutoff, iFFT) of G0 at N=2^{int(math.log2(N))} "
                      f"complex128, scipy.fft 1 worker, {per_step:.2f} s/step, extrapolated x194 steps per sim"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--lgN", type=int, default=26)
    ap.add_argument("--slices", type=int, default=64)
    ap.add_argument("--dtype", default="c8", choices=["c8", "c16"])
    ap.add_argument("--plain", action="store_true", help="one kernel per operation (baseline form)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the digit-group autotuner")
    ap.add_argument("--split", default="", help="fixed digit-group split 'lgR,lgA' instead of tuning")
    ap.add_argument("--traffic", default=str(ROOT / "profiles" / "traffic_cfg2.json"))
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-vs-plain comparison")
    ap.add_argument("--serve-qps", type=float, default=1.0,
                    help="serving run arrival rate in requests/s (0 = skip); the mix saturates near 1.7/s")
    ap.add_argument("--serve-seconds", type=float, default=30.0)
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="after the timed run, copy one field this many times (calibrates HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every configuration of this path runs on a single "
              "MI355X (see DESIGN.md for how it would shard); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import ravesim
    from ravesim.core import run_batch

    N = 2 ** args.lgN
    dtype = np.complex64 if args.dtype == "c8" else np.complex128
    members, zdet, steps_per_sim = cfg2_members(ravesim, N, args.slices)
    ctx = ravesim.Context(0)
    kw = dict(det_size=1e-3, det_px=1e-6, det_py=1.0, z_detector=zdet, dtype=dtype, plain=args.plain)

    if args.split:
        lgR, lgA = (int(v) for v in args.split.split(","))
        ctx.set_split(N, dtype, lgR, lgA)
        tuned = {"lgR": lgR, "lgA": lgA, "fixed": True}
    else:
        tuned = None if args.plain or args.no_tune else ctx.autotune(N, dtype)
    for _ in range(args.warmup):
        run_batch(ctx, N, 2e-9, members, **kw)
    ctx.synchronize()
    ctx.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = run_batch(ctx, N, 2e-9, members, **kw)
    ctx.synchronize()
    el = time.perf_counter() - t0
    stats = ctx.pass_stats()
    ctx.set_timing(False)
    ms_per_step = el * 1e3 / args.steps
    sims_per_hour = args.steps / el * 3600.0
    gss = N * steps_per_sim * args.steps / el / 1e9

    dom = max(stats, key=lambda k: stats[k][1])
    launches, tot_ms, byts = stats[dom]
    per_launch_ms = tot_ms / max(launches, 1)
    per_launch_bytes = byts / max(launches, 1)
    achieved = per_launch_bytes / (per_launch_ms * 1e-3) / 1e9
    traffic = None
    tp = Path(args.traffic)
    if tp.exists():
        try:
            rec = json.loads(tp.read_text()).get(dom)
            traffic = rec.get("hbm_bytes_per_launch") if rec else None
        except Exception:
            traffic = None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": f"{dom} pass", "launches": launches, "avg_ms": round(per_launch_ms, 4),
                "bytes_per_launch": per_launch_bytes,
                "passes": {k: {"launches": v[0], "avg_ms": round(v[1] / max(v[0], 1), 4),
                               "GBs": round(v[2] / max(v[1], 1e-12) / 1e6, 1)} for k, v in stats.items() if v[0]}}

    batch = None if (args.no_batch or args.plain) else batch_vs_plain(ravesim, ctx)
    serving = None
    if args.serve_qps > 0 and not args.plain:
        serving = serving_run(ravesim, ctx, args.serve_qps, args.serve_seconds)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(N)

    line = {
        "metric": "sims/hour & Gsample-steps/s at 1 GPU (%HBM roofline); p50/p99 latency at fixed QPS",
        "value": round(sims_per_hour, 2), "unit": "sims/hour", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c64+fp64phase" if args.dtype == "c8" else "c128",
        "data": "synthetic",
        "config": {"workload": f"Talbot-Lau 3-grating chain, 46 keV, N=2^{args.lgN}, "
                               f"{args.slices} slices/grating, {steps_per_sim} propagation steps/sim, "
                               f"{'plain' if args.plain else 'fused'} form",
                   "N": N, "steps_per_sim": steps_per_sim, "form": "plain" if args.plain else "fused",
                   "split": tuned},
        "gsample_steps_per_s": round(gss, 3),
        "hbm_frac_whole_sim": round(gss * 1e9 * (32 if args.dtype == "c8" else 64) / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched_vs_plain": batch,
        "serving": serving,
        "detector_checksum": float(np.sum(out)),
    }
    print(json.dumps(line), flush=True)
    if args.copy_probe:
        a = ravesim.Field(ctx, N, 2e-9, dtype=dtype)
        b = ravesim.Field(ctx, N, 2e-9, dtype=dtype)
        for _ in range(args.copy_probe):
            a.copy_to(b)
        ctx.synchronize()
        a.close()
        b.close()
    ctx.close()


if __name__ == "__main__":
    main()
