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

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "rave-sim_amd"))

HBM_PEAK_GBS = 8000.0
TIMING_STRIDE = 7
WL_46 = None


def talbot_lau_members(ravesim, N, slices=64, plain=False):
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


def poly_members(ravesim, N, n_energies=64, slices=32):
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


TUNE_TABLE = ROOT / "rave-sim_amd" / "ravesim" / "data" / "tuned_mi355x.txt"
TUNE_LOG = []


def ensure_tuned(ctx, N, dtype, members=1, retune=False):
    """The tuning entry for exactly (N, dtype, members): taken from the loaded
    table when it has one, else measured by the autotuner now (its seconds
    are logged in TUNE_LOG)."""
    t = None if retune else ctx.tuned(N, dtype, members)
    if t is not None and t["members"] == members:
        return {**t, "tuned_now_s": 0.0}
    t0 = time.perf_counter()
    ctx.autotune(N, dtype, members=members)
    dt = time.perf_counter() - t0
    TUNE_LOG.append({"N": N, "dtype": np.dtype(dtype).name, "members": members, "s": round(dt, 3)})
    return {**ctx.tuned(N, dtype, members), "tuned_now_s": round(dt, 3)}


def batch_vs_plain(ravesim, ctx, lgN=24, reps=2, plain_leg=True):
    """Batched fused complex64 form vs the plain complex128 form run one
    simulation at a time, on the polychromatic batch (sims/hour each);
    plain_leg False: the fused form only (for counter runs)."""
    import numpy as np
    from ravesim.core import run_batch
    N = 2 ** lgN
    ms, zdet, steps = poly_members(ravesim, N)
    kw = dict(det_size=1e-3, det_px=1e-6, det_py=1.0, z_detector=zdet)
    tuned = ensure_tuned(ctx, N, np.complex64, len(ms))
    run_batch(ctx, N, 2e-9, ms[:2], dtype=np.complex64, **kw)
    rep_s = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fused = run_batch(ctx, N, 2e-9, ms, dtype=np.complex64, **kw)
        rep_s.append(time.perf_counter() - t0)
    te = sum(rep_s) / reps
    # the batched form as a captured hipGraph over both lanes: the first run
    # captures it, the timed runs replay it with their argument records
    fused_graph = run_batch(ctx, N, 2e-9, ms, dtype=np.complex64, use_graph=True, **kw)
    graph_s = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fused_graph = run_batch(ctx, N, 2e-9, ms, dtype=np.complex64, use_graph=True, **kw)
        graph_s.append(time.perf_counter() - t0)
    tf = sum(graph_s) / reps
    if not plain_leg:
        return {"workload": f"64 energies x N=2^{lgN}, {steps} steps each, weighted detector sum",
                "fused_c64_batched_sims_per_hour": round(len(ms) / tf * 3600, 1), "split": tuned,
                "fused_graph_rep_s": [round(x, 4) for x in graph_s],
                "fused_eager_rep_s": [round(x, 4) for x in rep_s], "plain": "skipped (--no-plain)"}
    # the baseline at the faster of its default split and the tuned one
    plain_runs = {}
    for label in ("default", "tuned"):
        if label == "tuned":
            plain_runs["tuned_split"] = ensure_tuned(ctx, N, np.complex128, 1)
        run_batch(ctx, N, 2e-9, ms[:1], dtype=np.complex128, plain=True, max_resident=1, **kw)
        t0 = time.perf_counter()
        plain = run_batch(ctx, N, 2e-9, ms, dtype=np.complex128, plain=True, max_resident=1, **kw)
        plain_runs[label] = time.perf_counter() - t0
    tp = min(plain_runs["default"], plain_runs["tuned"])
    err = float(np.linalg.norm(fused - plain) / np.linalg.norm(plain))
    return {"workload": f"64 energies x N=2^{lgN}, {steps} steps each, weighted detector sum",
            "fused_c64_batched_sims_per_hour": round(len(ms) / tf * 3600, 1),
            "plain_c128_one_at_a_time_sims_per_hour": round(len(ms) / tp * 3600, 1),
            "speedup": round(tp / tf, 2), "speedup_eager": round(tp / te, 2),
            "graph_rows_equal_eager": bool(np.array_equal(fused_graph, fused)),
            "rel_l2_fused_vs_plain_c128": err, "split": tuned,
            "fused_graph_rep_s": [round(x, 4) for x in graph_s], "fused_eager_rep_s": [round(x, 4) for x in rep_s],
            "plain_s": round(tp, 4),
            "plain_s_by_split": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in plain_runs.items()}}


def uneven_mix_sims(ravesim, n_sims=256, seed=11, slices=32):
    """The uneven mix: n_sims simulations, N = 2^k with k uniform on 16..24
    (log-uniform sizes), energy uniform on 20-70 keV, grating period uniform
    on 2-6 um with duty cycle 0.4-0.6 (the single-grating chain: Si, 30 um,
    `slices` slices, delta ~ E^-2 and beta ~ E^-4 from the 25 keV values,
    0.1 m hop), 1-16 point sources 5 m before the grating at lateral offsets
    uniform within +-0.5 um, their intensities summed incoherently.
    Returns [(lgN, [sources of sim])]."""
    import numpy as np
    rng = np.random.default_rng(seed)
    h = 30e-6
    sims = []
    for _ in range(n_sims):
        lgN = int(rng.integers(16, 25))
        e = float(rng.uniform(20.0, 70.0))
        pitch = float(rng.uniform(2e-6, 6e-6))
        dc = float(rng.uniform(0.4, 0.6))
        nsrc = int(rng.integers(1, 17))
        xs = rng.uniform(-0.5e-6, 0.5e-6, nsrc)
        sc = 25.0 / e
        g = ravesim.Grating(pitch=pitch, dc=(dc, dc), z_start=0.0, thickness=h, nr_steps=slices,
                            db_a=complex(7.7e-7 * sc ** 2, 3.5e-10 * sc ** 4))
        wl = ravesim.energy_to_wavelength(e * 1000.0)
        sims.append((lgN, [ravesim.core.BatchMember(wl=wl, gratings=[g], cutoff_angles=[math.pi / 2] * 2,
                                                    source_x=float(x), source_z=-5.0) for x in xs]))
    return sims, h + 0.1, slices + 1


def uneven_mix(ravesim, ctx, n_sims=256, retune=False):
    """All simulations in one mixed call (rs_batch_run_mixed: one lockstep
    part per grid size, split and tiles tuned per size and batch size, each
    simulation's sources summed into its own detector row), eager and as one
    captured hipGraph over both lanes, vs the plain complex128 form one
    simulation at a time."""
    import numpy as np
    from ravesim.core import BatchSim, run_batch, run_mixed
    sims, zdet, steps = uneven_mix_sims(ravesim, n_sims)
    dx = 2e-9
    by_n = {}
    for i, (lgN, srcs) in enumerate(sims):
        by_n.setdefault(lgN, []).append(i)

    def kw(lgN):
        return dict(det_size=(2 ** lgN) * dx / 4, det_px=1e-6, det_py=1.0, z_detector=zdet)

    mixed = [BatchSim(N=2 ** lgN, dx=dx, members=srcs, **kw(lgN)) for lgN, srcs in sims]

    t0 = time.perf_counter()
    for lgN, idx in sorted(by_n.items()):
        ensure_tuned(ctx, 2 ** lgN, np.complex64, sum(len(sims[i][1]) for i in idx), retune)
    tune_s = time.perf_counter() - t0
    run_mixed(ctx, mixed)  # warm: plans, tables
    t0 = time.perf_counter()
    rows = run_mixed(ctx, mixed)
    tf = time.perf_counter() - t0
    fused = dict(enumerate(rows))
    run_mixed(ctx, mixed, use_graph=True)
    t0 = time.perf_counter()
    grows = run_mixed(ctx, mixed, use_graph=True)
    tg = time.perf_counter() - t0
    graph_equal = all(np.array_equal(a, b) for a, b in zip(grows, rows))
    t0 = time.perf_counter()
    err_num = err_den = 0.0
    for i, (lgN, srcs) in enumerate(sims):
        for m in srcs:
            m.out_index = 0
        row = run_batch(ctx, 2 ** lgN, dx, srcs, dtype=np.complex128, plain=True, max_resident=1,
                        **kw(lgN))[0]
        err_num += float(np.sum((fused[i] - row) ** 2))
        err_den += float(np.sum(row ** 2))
    tp = time.perf_counter() - t0
    n_src = sum(len(s) for _, s in sims)
    samples = sum(len(s) * 2 ** lgN for lgN, s in sims)
    tb = min(tf, tg)
    return {"workload": f"{n_sims} sims, N = 2^16..2^24 log-uniform, 1-16 point sources each "
                        f"({n_src} in all), {steps} steps per source, E 20-70 keV, p 2-6 um, dc 0.4-0.6",
            "form": "one mixed call; graph replay" if tg <= tf else "one mixed call; eager",
            "fused_c64_batched_sims_per_hour": round(n_sims / tb * 3600, 1),
            "fused_gsample_steps_per_s": round(samples * steps / tb / 1e9, 2),
            "plain_c128_one_at_a_time_sims_per_hour": round(n_sims / tp * 3600, 1),
            "speedup": round(tp / tb, 2), "speedup_eager": round(tp / tf, 2), "speedup_graph": round(tp / tg, 2),
            "rel_l2_fused_vs_plain_c128": math.sqrt(err_num / err_den),
            "fused_s": round(tf, 4), "fused_graph_s": round(tg, 4), "graph_rows_equal_eager": graph_equal,
            "plain_s": round(tp, 4), "tune_s": round(tune_s, 2),
            "sims_per_size": {f"2^{k}": len(v) for k, v in sorted(by_n.items())}}


def rocfft_comparison_child(N):
    """rocfft_comparison in a child process: torch carries its own HIP runtime,
    which cannot share a process with the one this library loaded."""
    import subprocess
    r = subprocess.run([sys.executable, str(Path(__file__).resolve()), "--vendor-fft-only", str(N)],
                       capture_output=True, text=True, timeout=240)
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    if r.returncode != 0 or not lines:
        return {"error": f"exit {r.returncode}: {r.stderr.strip()[-300:]}"}
    return json.loads(lines[-1])


def simdir_end_to_end(ravesim, ctx, N, slices=64, reps=2):
    """The Talbot-Lau chain through the sim-dir boundary (`run_single_simulation`,
    the fastwave / big-wave entry point) with a vector source of ones read
    from a complex64 .npy: YAML parsing, the .npy read, the upload over
    PCIe, the simulation and the detector write, by wall clock."""
    import tempfile

    import numpy as np
    from ravesim.multisim import run_single_simulation
    from ravesim.simdir import write_simdir
    h, p, dx = 60e-6, 4.2e-6, 2e-9

    def grating(z, mat):
        return {"type": "grating", "pitch": p, "dc": [0.5, 0.5], "z_start": z, "thickness": h,
                "nr_steps": slices, "x_positions": [0.0], "substrate_thickness": 0.0, "mat_a": mat,
                "mat_b": None, "mat_substrate": None}

    with tempfile.TemporaryDirectory(prefix="rs_simdir_") as tmp:
        src = Path(tmp) / "ones.npy"
        np.save(src, np.ones(N, dtype=np.complex64))
        cfg = {"sim_params": {"N": N, "dx": dx, "z_detector": 0.4 + h, "detector_size": 1e-3,
                              "detector_pixel_size_x": 1e-6, "detector_pixel_size_y": 1.0, "chunk_size": N},
               "use_disk_vector": False, "save_final_u_vectors": False, "dtype": "c8",
               "elements": [grating(0.0, ["Au", 19.32]), grating(0.2, ["Si", 2.34]), grating(0.4, ["Au", 19.32])]}
        sub = {"energy": 46000.0, "source": {"type": "vector", "input_path": str(src), "z": 0.0},
               "deltabeta_table": [[["Au", 19.32], [1.1e-6, 4.0e-8]], [["Si", 2.34], [2.6e-7, 1.0e-10]]]}
        sd = write_simdir(Path(tmp) / "sim", cfg, [sub], [math.pi / 2] * 4)
        run_single_simulation(sd, 0, ctx=ctx)
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            rows = run_single_simulation(sd, 0, ctx=ctx)
            times.append(time.perf_counter() - t0)
    t = min(times)
    return {"what": "Talbot-Lau chain (N=2^26) through run_single_simulation: YAML, 512 MiB complex64 vector source read and "
                    "uploaded, 194 steps, detected.npy written", "ms_per_sim": round(t * 1e3, 2),
            "sims_per_hour": round(3600.0 / t, 1), "detector_checksum": float(np.sum(rows))}


def rocfft_comparison(N=2 ** 26, reps=20):
    """One propagation step FFT * propagator * iFFT on N complex64 samples with
    the vendor FFT (torch.fft -> hipFFT -> rocFFT), as a reference point for
    this library's fused propagation step."""
    import torch
    if not torch.cuda.is_available():
        return None
    dev = torch.device("cuda", 0)
    u = torch.ones(N, dtype=torch.complex64, device=dev)
    f = torch.fft.fftfreq(N, d=2e-9, device=dev, dtype=torch.float64)
    wl, dz = 2.695e-11, 1e-6
    ph = (math.pi * wl * dz) * f * f - 2 * math.pi * dz / wl
    H = torch.polar(torch.ones_like(ph), torch.remainder(ph, 2 * math.pi)).to(torch.complex64)
    del f, ph
    for _ in range(3):
        u = torch.fft.ifft(torch.fft.fft(u) * H)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        u = torch.fft.ifft(torch.fft.fft(u) * H)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    e0.record()
    for _ in range(reps):
        v = torch.fft.fft(u)
    e1.record()
    torch.cuda.synchronize()
    fft_ms = e0.elapsed_time(e1) / reps
    del u, H, v
    torch.cuda.empty_cache()
    return {"what": f"torch.fft (hipFFT/rocFFT) fft * H * ifft, N=2^{int(math.log2(N))} complex64",
            "ms_per_propagation_step": round(ms, 4), "ms_per_fft": round(fft_ms, 4),
            "fft_GBs_at_16B_per_sample": round(16 * N / (fft_ms * 1e-3) / 1e9, 1)}


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
    for k, (at, req) in enumerate(arrivals):
        dt = at - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        ids.append((srv.submit(req), len(req.energies_ev)))
        print(f"[serve] t={time.perf_counter() - t0:.1f}s submitted {k}: N=2^{int(math.log2(req.N))}, "
              f"{len(req.energies_ev)} energies", file=sys.stderr, flush=True)
    res = []
    for k, (i, _) in enumerate(ids):
        res.append(srv.result(i, timeout_s=900))
        print(f"[serve] t={time.perf_counter() - t0:.1f}s request {k} done, latency {res[-1].latency_s:.2f}s",
              file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    st = srv.stats()
    # error vs FP64 per request: every request up to 2^26 and the 2^28
    # request with the fewest energies served again in complex128
    errs = []
    big = sorted((k for k in range(len(arrivals)) if arrivals[k][1].N == 2 ** 28),
                 key=lambda k: len(arrivals[k][1].energies_ev))[:1]
    for k in [k for k in range(len(arrivals)) if arrivals[k][1].N <= 2 ** 26] + big:
        r = arrivals[k][1]
        ref = srv.result(srv.submit(phase_sweep(r.N, r.energies_ev, n_steps=8, slices=64,
                                                dtype=np.complex128)), timeout_s=600).rows
        errs.append({"request": k, "N": r.N, "energies": len(r.energies_ev),
                     "rel_l2": float(np.linalg.norm(res[k].rows - ref) / np.linalg.norm(ref)),
                     "fringe_phase_err_rad": fringe_phase_error(res[k].rows, ref)})
    srv.close()
    lat = np.array([r.latency_s for r in res])
    sims = sum(ne for _, ne in ids)
    by_size = {}
    for (at, req), r in zip(arrivals, res):
        by_size.setdefault(f"2^{int(math.log2(req.N))}", []).append(round(r.latency_s, 3))
    return {"qps": qps, "seed": seed, "requests": len(res), "sims": sims, "duration_s": round(wall, 2),
            "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
            "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
            "sims_per_hour": round(sims / wall * 3600, 1),
            "requests_by_size": [{"N": f"2^{int(math.log2(q.N))}", "energies": len(q.energies_ev), "at_s": round(at, 2),
                                  "latency_s": round(r.latency_s, 3)} for (at, q), r in zip(arrivals, res)],
            "mix": f"N in 2^{list(sizes)}, 1-{max_energies} energies, 8 phase steps, 650 steps per energy",
            "jobs_per_launch": round(st["jobs_per_launch"], 2), "max_resident_jobs": st["max_resident_jobs"],
            "error_vs_fp64": errs}


def fringe_phase_error(rows, ref):
    """Largest fringe-phase difference (rad) of a phase-stepping scan: per
    pixel -atan2(sum_k I_k sin(2 pi k / n), sum_k I_k cos(2 pi k / n)), over the
    pixels whose first-harmonic amplitude exceeds 1e-3 of the mean intensity."""
    import numpy as np
    n = rows.shape[0]
    w = np.exp(2j * np.pi * np.arange(n) / n)[:, None]
    a, b = (w * rows).sum(0), (w * ref).sum(0)
    ok = np.abs(b) > 1e-3 * np.abs(ref).mean() * n
    if not ok.any():
        return None
    return float(np.max(np.abs(np.angle(a[ok] * np.conj(b[ok])))))


def fringe_bin_phase_error(row, ref):
    """Fringe-phase difference (rad) of one detector row: the phase of the
    row's DFT at the bin where the reference's spectrum peaks (DC excluded),
    against the reference's phase at that bin."""
    import numpy as np
    a = np.fft.rfft(row - row.mean())
    b = np.fft.rfft(ref - ref.mean())
    k = 1 + int(np.argmax(np.abs(b[1:])))
    return float(abs(np.angle(a[k] * np.conj(b[k])))), k


def chain_precision(ravesim, ctx, N, members, kw, out):
    """The timed chain's complex64 + FP64-phase detector row against the
    same simulation in the plain complex128 form (the oracle's precision;
    the oracle itself is checked against that form at smaller N)."""
    import numpy as np
    from ravesim.core import run_batch
    t0 = time.perf_counter()
    ref = run_batch(ctx, N, 2e-9, members, **{**kw, "dtype": np.complex128, "plain": True})
    ctx.synchronize()
    dphi, k = fringe_bin_phase_error(out[0], ref[0])
    return {"reference": "plain complex128 form, same simulation",
            "rel_l2": float(np.linalg.norm(out - ref) / np.linalg.norm(ref)),
            "fringe_phase_err_rad": dphi, "fringe_bin": k, "reference_s": round(time.perf_counter() - t0, 3)}


def cpu_baseline(N, budget_s=25.0, workers=1):
    """Oracle (complex128 numpy/scipy restatement, `workers` scipy.fft
    threads; big-wave uses the default single one) on whole grating slices of
    G0 at the benchmark N; scaled to sims/hour."""
    import numpy as np
    import scipy.fft
    from oracle import bigwave as bw
    wl = bw.convert_energy_wavelength(46000.0)
    dx = 2e-9
    p = bw.SimParams(N=N, dx=dx, z_detector=0.4 + 60e-6, detector_size=1e-3,
                     detector_pixel_size_x=1e-6, detector_pixel_size_y=1.0, wl=wl, chunk_size=N)
    g = bw.Grating(pitch=4.2e-6, dc=(0.5, 0.5), z_start=0.0, thickness=60e-6 / 64, nr_steps=1,
                   x_positions=np.array([0.0]), substrate_thickness=0.0, db_a=complex(1.1e-6, 4.0e-8))
    f = bw.Fields(np.ones(N, dtype=np.complex128))
    n, t0 = 0, time.perf_counter()
    with scipy.fft.set_workers(workers):
        while True:
            bw.grating_apply(f, g, p, 1e30, 0)
            n += 1
            el = time.perf_counter() - t0
            if el > budget_s or n >= 8:
                break
    per_step = el / n
    cpu = ""
    try:
        with open("/proc/cpuinfo") as fh:
            cpu = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": 3600.0 / (194 * per_step), "unit": "sims/hour", "cores": workers, "kind": "port",
            "cpu_model": cpu, "nproc": os.cpu_count(),
            "sample": f"{n} grating slice(s) (mask + FFT, propagator, cutoff, iFFT) of G0 at N=2^{int(math.log2(N))} "
                      f"complex128, scipy.fft {workers} worker(s), {per_step:.2f} s/step, extrapolated x194 steps per sim"}


T_START = time.perf_counter()


def progress(what):
    """One line per finished leg on stderr (the JSON line stays alone on stdout)."""
    print(f"[bench] {time.perf_counter() - T_START:7.1f} s  {what}", file=sys.stderr, flush=True)


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
    ap.add_argument("--tune-table", default=str(TUNE_TABLE),
                    help="tuning table to load (splits and column tiles per N, precision and batch size); "
                         "configurations it lacks are tuned in the run")
    ap.add_argument("--retune", action="store_true", help="tune every configuration again, ignoring the table")
    ap.add_argument("--tune-save", default="", help="write the context's tuning table here at the end")
    ap.add_argument("--traffic", default=str(ROOT / "profiles" / "traffic_talbot_lau_n26.json"))
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-vs-plain comparisons")
    ap.add_argument("--no-plain", action="store_true",
                    help="batch comparison: run the fused form only (no plain complex128 leg)")
    ap.add_argument("--mix-sims", type=int, default=256, help="simulations of the uneven mix (0 = skip)")
    ap.add_argument("--no-rocfft", action="store_true", help="skip the vendor FFT comparison")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the chain's complex128 reference run")
    ap.add_argument("--no-simdir", action="store_true", help="skip the sim-dir (PCIe-inclusive) run")
    ap.add_argument("--serve-qps", type=float, default=0.2,
                    help="serving run arrival rate in requests/s (0 = skip)")
    ap.add_argument("--serve-seconds", type=float, default=30.0)
    ap.add_argument("--serve-sizes", default="22,26,28,30", help="log2 grid sizes of served requests")
    ap.add_argument("--serve-max-energies", type=int, default=8, help="energies per request: 1..this")
    ap.add_argument("--serve-seed", type=int, default=32,
                    help="seed of the arrival process (32: eight requests from 2^22 to 2^30, up to 8 energies, "
                         "about 90 s of device work)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="after the timed run, copy one field this many times (calibrates HBM counters)")
    ap.add_argument("--vendor-fft-only", type=int, default=0, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.vendor_fft_only:
        print(json.dumps(rocfft_comparison(args.vendor_fft_only)), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: every configuration of this path runs on a single "
              "MI355X, and separate simulations scale out as one process per GPU; run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import ravesim
    from ravesim.core import run_batch

    N = 2 ** args.lgN
    dtype = np.complex64 if args.dtype == "c8" else np.complex128
    members, zdet, steps_per_sim = talbot_lau_members(ravesim, N, args.slices)
    ctx = ravesim.Context(0)
    kw = dict(det_size=1e-3, det_px=1e-6, det_py=1.0, z_detector=zdet, dtype=dtype, plain=args.plain)

    table = {"path": None, "entries": 0}
    if args.tune_table and Path(args.tune_table).exists() and not args.retune:
        table = {"path": str(Path(args.tune_table).relative_to(ROOT)) if Path(args.tune_table).is_relative_to(ROOT)
                 else args.tune_table, "entries": ctx.tune_load(args.tune_table)}
    if args.split:
        lgR, lgA = (int(v) for v in args.split.split(","))
        ctx.set_split(N, dtype, lgR, lgA)
        tuned = {"lgR": lgR, "lgA": lgA, "fixed": True}
    else:
        tuned = {"default": True} if args.plain or args.no_tune else ensure_tuned(ctx, N, dtype, 1, args.retune)
    progress(f"split for the chain: {tuned}")
    for _ in range(args.warmup):
        run_batch(ctx, N, 2e-9, members, **kw)
    ctx.synchronize()
    # HIP events on the library's stream, recorded by the dispatches of every
    # TIMING_STRIDE-th launch of each pass class (a timed dispatch costs a few
    # microseconds; 7 is prime to the launch patterns, whose periods are powers
    # of two)
    ctx.set_timing(True, stride=TIMING_STRIDE)
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

    # Roofline kernel: the top column pass, the one pass of a step that
    # covers the whole 512 MiB field (twice the 256 MiB last-level cache),
    # so its reads and writes stream from and to HBM; the middle and row
    # passes run over cache-sized spans and read largely from that cache.
    # Beside it: a streaming copy of the same field size in the same run.
    dom = "top" if stats.get("top", (0,))[0] else max(stats, key=lambda k: stats[k][1])
    launches, tot_ms, byts = stats[dom]
    per_launch_ms = tot_ms / max(launches, 1)
    per_launch_bytes = byts / max(launches, 1)
    achieved = per_launch_bytes / (per_launch_ms * 1e-3) / 1e9
    copy_gbs = ctx.copy_bandwidth(N * (8 if args.dtype == "c8" else 16), reps=20)
    split_now = None if args.plain or args.no_tune else [tuned.get("lgR"), tuned.get("lgA"), tuned.get("tileA", 0),
                                                          tuned.get("tileB", 0)]
    traffic, traffic_split = None, None
    tp = Path(args.traffic)
    if tp.exists():
        try:
            prof = json.loads(tp.read_text())
            traffic_split = prof.get("_split")
            rec = prof.get(dom)
            # counters taken at another split describe other kernels
            if rec and traffic_split == split_now:
                traffic = rec.get("hbm_bytes_per_launch")
        except Exception:
            traffic = None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "achievable": round(copy_gbs, 1), "frac_of_achievable": round(achieved / copy_gbs, 4),
                "kernel": f"{dom} pass", "launches": launches, "timing_stride": TIMING_STRIDE,
                "avg_ms": round(per_launch_ms, 4),
                "bytes_per_launch": per_launch_bytes, "traffic_profiled_at_split": traffic_split,
                "passes": {k: {"launches": v[0], "avg_ms": round(v[1] / max(v[0], 1), 4),
                               "GBs": round(v[2] / max(v[1], 1e-12) / 1e6, 1)} for k, v in stats.items() if v[0]}}

    progress(f"timed chain: {sims_per_hour:.0f} sims/h; top pass {achieved:.0f} GB/s, copy {copy_gbs:.0f} GB/s")
    precision = None
    if not args.plain and args.dtype == "c8" and not args.no_precision:
        precision = chain_precision(ravesim, ctx, N, members, kw, out)
        progress(f"chain vs plain complex128: rel-L2 {precision['rel_l2']:.2e}, "
                 f"fringe phase {precision['fringe_phase_err_rad']:.2e} rad")
    batch = None if (args.no_batch or args.plain) else batch_vs_plain(ravesim, ctx, plain_leg=not args.no_plain)
    if batch:
        progress(f"polychromatic batch: {batch.get('speedup', 'fused only')}x")
    simdir = None
    if not (args.no_simdir or args.plain):
        try:
            simdir = simdir_end_to_end(ravesim, ctx, N, args.slices)
            simdir["vs_device_resident"] = round(simdir["sims_per_hour"] / sims_per_hour, 4)
        except Exception as e:  # a comparison only: report why it is missing
            simdir = {"error": repr(e)}
        progress("sim-dir boundary run")
    mix = None
    if not (args.no_batch or args.plain) and args.mix_sims > 0:
        mix = uneven_mix(ravesim, ctx, args.mix_sims, args.retune)
        progress(f"uneven mix: fused {mix['fused_s']} s, graph {mix['fused_graph_s']} s, {mix['speedup']}x")
    ours_step_ms = ms_per_step / steps_per_sim
    vendor = None
    if not args.no_rocfft and not args.plain:
        try:
            vendor = rocfft_comparison_child(N)
        except Exception as e:  # a comparison only: report why it is missing
            vendor = {"error": repr(e)}
        if vendor and "ms_per_propagation_step" in vendor:
            vendor["this_library_ms_per_step_incl_grating"] = round(ours_step_ms, 4)
        progress("vendor FFT comparison")
    serving = None
    if args.serve_qps > 0 and not args.plain:
        serving = serving_run(ravesim, ctx, args.serve_qps, args.serve_seconds,
                              sizes=tuple(int(v) for v in args.serve_sizes.split(",")),
                              max_energies=args.serve_max_energies, seed=args.serve_seed)

    cpu = None
    if not args.no_cpu:
        progress("CPU baseline starts (about 40 s)")
        cpu = cpu_baseline(N)
        # the same with every visible core on the FFTs (at most 16, the box's
        # share), a reported extra: big-wave itself runs scipy.fft single-threaded
        nthr = min(os.cpu_count() or 1, 16)
        if nthr > 1:
            cpu["multicore"] = cpu_baseline(N, budget_s=12.0, workers=nthr)

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
        "precision": precision,
        "cpu_baseline": cpu,
        "batched_vs_plain": batch,
        "uneven_mix": mix,
        "vendor_fft": vendor,
        "serving": serving,
        "simdir_end_to_end": simdir,
        "tuning": {"table": table, "tuned_in_run": TUNE_LOG, "tune_s": round(sum(t["s"] for t in TUNE_LOG), 3)},
        "detector_checksum": float(np.sum(out)),
    }
    if args.tune_save:
        ctx.tune_save(args.tune_save)
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
