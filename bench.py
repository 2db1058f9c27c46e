"""Benchmark: FP64 ADI time step of the synthetic meandering river
(980 x 1048 cells, dx = dy = 20 m, dT = 5 s, hydrodynamics only) on one
MI355X, in wet-cell-steps per second.

A step is one full ADI step (UZ + VZ half-steps) of the whole river; the
timed region replays K captured steps (hd_run) with the state resident in
HBM, after W warm-up steps.  A clone of the state at the start of the timed
window then runs the same K steps again with HIP events around every launch
on the library's stream, so the roofline of the dominant kernel (the kernel
function with the largest total time per step, its UZ and VZ launches
together) is measured on exactly the timed steps; its HBM traffic per launch
comes from the rocprofv3 counter runs recorded in profiles/traffic.json.
One JSON line is printed, with the CPU checker timed on this host (median of
three runs, -O3 -march=native built here) and the batched ensemble (64
seeded rivers, one launch per kernel over all of them).

    python bench.py [--gpus 1] [--steps 720] [--warmup 20] [--no-cpu] [--no-ensemble]

--workload sediment | ensemble | cnn | surrogate prints the line of another subsystem
instead (the river with suspended sediment and bed change every step; the
ensemble alone, with per-kernel times; the surrogate network's forward pass
and training step; the surrogate-accelerated bed change against plain FP64
ADI and an FP64 twin), for profiling each subsystem with rocprofv3.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-hydist_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
B_STEP = 256               # algorithmic bytes per wet-cell-step (hydrodynamics)
B_STEP_SED = 304           # ... with suspended sediment

# Algorithmic bytes per wet cell of the per-launch kernels (compulsory FP64
# reads/writes of the kernel's inputs and outputs, each element counted once).
KERNEL_BYTES = {
    # fused assembly + Thomas + extraction of one half-step: u, v, z, Htdu,
    # Htdv, H_moi, Kx1|Ky1, VISCOIDX, Tsxw|Tsyw read, t_z and t_u|t_v written
    "adi_spec_uz": 9 * 8 + 2 * 8,
    "adi_spec_vz": 9 * 8 + 2 * 8,
    # the same work with the serial Thomas on one lane (HYDIST_SOLVER=run)
    "adi_run_uz": 9 * 8 + 2 * 8,
    "adi_run_vz": 9 * 8 + 2 * 8,
    # explicit v (u): v, u, t_u, Htdv, Ky1, H_moi, z, VISCOIDX, Tsyw read
    # (mirror for u), t_v (t_u) written
    "explicit_margins_v": 9 * 8 + 8,
    "explicit_margins_u": 9 * 8 + 8,
    "explicit_rows_u": 9 * 8 + 8,
    # smoothing of t_v with explicit u in one pass: t_v, khouot read, the
    # partner written, plus explicit u's 80 B, plus the drying marks (htaiz,
    # t_z, z read, the drying flag written)
    "smooth_v_explicit_u": 8 + 4 + 8 + 9 * 8 + 8 + 3 * 8 + 1,
    "smooth_u_explicit_v": 8 + 4 + 8 + 9 * 8 + 8 + 3 * 8 + 1,
    # state copy + drying + face depths: t_z, t_u, t_v, H_moi, h read;
    # z, u, v, Htdu, Htdv written
    "uvz_dry_htd": 5 * 8 + 5 * 8,
    # smoothing: t_u (t_v) and khouot read, the partner buffer written; the
    # second pass also reads htaiz, t_z and writes H_moi
    "smooth_u": 8 + 4 + 8,
    "smooth_v": 8 + 4 + 8,
    "smooth_v_hmoi": 8 + 4 + 8 + 2 * 8 + 8,
    "smooth_u_hmoi": 8 + 4 + 8 + 2 * 8 + 8,
    # the second smoothing (velocity read, partner written) with update_h_moi
    # (htaiz, t_z read, H_moi written) beside the wetting lines
    "smooth_v_wetting_rows": 8 + 8 + 2 * 8 + 8,
    "smooth_u_wetting_cols": 8 + 8 + 2 * 8 + 8,
    # wet-run tables: khouot read
    "segments_fs_edges": 4,
    # the two in one launch (state copy + tables)
    "uvz_segments": 5 * 8 + 5 * 8 + 4,
    # the same work, one wave per line (HYDIST_SOLVER=line)
    "adi_line_uz": 9 * 8 + 2 * 8,
    "adi_line_vz": 9 * 8 + 2 * 8,
    # two-kernel form (HYDIST_SOLVER=rows): rows of A,B,C,D round-trip HBM
    "adi_solve_uz": 2 * 4 * 8 + 2 * 8,
    "adi_solve_vz": 2 * 4 * 8 + 2 * 8,
    "adi_coef_uz": 9 * 8 + 2 * 4 * 8,
    "adi_coef_vz": 9 * 8 + 2 * 4 * 8,
    # suspended sediment: Find_VTH + hesoK (t_u, t_v, H_moi, htaiz, Htdu,
    # Htdv, h read; VTH, Kx, Ky written)
    "vth_hesok": 7 * 8 + 3 * 8,
    # the concentration sweeps (FS, H_moi, t_u, t_v, VTH, Htdu, Htdv, Fw, Kx,
    # Ky read; tFS written)
    "fs_spec_j": 10 * 8 + 8,
    "fs_spec_i": 10 * 8 + 8,
    # Update_FS: tFS read, FS written
    "update_fs": 2 * 8,
    # bed-load flux: Fw, VTH, H_moi, t_u, t_v and khouot read; Qbx, Qby written
    "calc_qb": 5 * 8 + 4 + 2 * 8,
    # bed change: FS, H_moi, t_u, t_v, VTH, Htdu, Htdv, Fw, Kx, Ky, Qbx, Qby,
    # dH and khouot read; dH written
    "bedload": 13 * 8 + 4 + 8,
}

# Surrogate network: multiply-adds per pixel of one forward pass (3 x 3
# convolutions 8 -> 64, six 64 -> 64, 64 -> 3; two flops each); a training
# step (forward, data and weight gradients) is three times that.
CNN_FLOP_FWD = 2 * 9 * (8 * 64 + 6 * 64 * 64 + 64 * 3)
MFMA_BF16_PEAK = 2.5e15


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=720)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--profile-steps", type=int, default=20)
    p.add_argument("--cpu-steps", type=int, default=400, help="steps per CPU run (all host threads)")
    p.add_argument("--cpu-steps-single", type=int, default=200, help="steps per single-thread CPU run")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--ensemble-steps", type=int, default=720)
    p.add_argument("--no-ensemble", action="store_true")
    p.add_argument("--no-tune", action="store_true", help="keep the default launch geometry (no autotuning)")
    p.add_argument("--workload", default="river", choices=("river", "sediment", "ensemble", "cnn", "surrogate"),
                   help="river: the headline line (default); the others print their own line "
                        "(sediment: the river with suspended sediment and bed change every step; "
                        "ensemble: the 64-river batch; cnn: the surrogate network's passes; surrogate: the "
                        "surrogate-accelerated bed change against plain FP64 ADI, about two minutes)")
    p.add_argument("--cnn-reps", type=int, default=20)
    p.add_argument("--sur-pairs", type=int, default=60, help="surrogate: training pairs per inflow variant")
    p.add_argument("--sur-epochs", type=int, default=6)
    p.add_argument("--sur-days", type=float, default=1.0, help="surrogate: days of bed change compared")
    return p.parse_args()


def cpu_threads() -> int:
    """Host threads for the CPU leg: the cores this process may run on,
    at most 16 (one GPU's share of the host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def roofline(kt, wet_mean, traffic_file=None):
    """The roofline object of the dominant kernel from per-label HIP-event
    totals kt = {label: (total ms, launches)} and the mean wet cells per
    step.  The two sweeps of one kernel function (labels *_uz and *_vz)
    count as one kernel: each alone is about as long per step as the state
    copy's two launches, so a per-label pick would flip between runs.  The
    traffic is the launch-weighted mean of the labels' HBM bytes per launch
    in traffic_file (profiles/traffic.json), or None."""
    groups = {}
    for k, (ms, n) in kt.items():
        g_name = k[:-3] if k.endswith(("_uz", "_vz")) else k
        groups.setdefault(g_name, []).append((k, ms, n))
    dom_name, members = max(groups.items(), key=lambda kv: sum(ms for _, ms, _ in kv[1]))
    dom_ms = sum(ms for _, ms, _ in members)
    dom_n = sum(n for _, _, n in members)
    avg_s = dom_ms / dom_n * 1e-3
    per_cells = {KERNEL_BYTES.get(k) for k, _, _ in members}
    per_cell = per_cells.pop() if len(per_cells) == 1 else None
    roof = {"kernel": dom_name, "labels": sorted(k for k, _, _ in members), "bound": "hbm",
            "unit": "GB/s", "peak": HBM_PEAK / 1e9, "avg_launch_us": avg_s * 1e6, "launches": dom_n}
    if per_cell is not None:
        bytes_launch = per_cell * wet_mean
        roof.update(achieved=bytes_launch / avg_s / 1e9, frac=bytes_launch / avg_s / HBM_PEAK,
                    bytes_per_wet_cell=per_cell, algorithmic_bytes_per_launch=bytes_launch)
    else:
        roof.update(achieved=None, frac=None)
    traffic = None
    tf = traffic_file or os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            tj = json.load(f)
        if all(k in tj for k, _, _ in members):
            traffic = round(sum(tj[k] * n for k, _, n in members) / dom_n)
    roof["traffic"] = traffic
    return roof


def cpu_run(case, steps):
    """One timed run of the CPU checker: `steps` steps of the workload from
    t = 0 (initialisation excluded); returns (wet-cell-steps/s, seconds)."""
    import oracle
    o = oracle.Oracle(case)
    o.init()
    opts = case.opts()
    wet = 0
    t = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        wet += o.wet_count()
        t = o.step(t, opts)
    dt = time.perf_counter() - t0
    o.close()
    return wet / dt, dt


def cpu_baseline(case, steps, threads, runs=3):
    """The CPU checker (same FP64 method, OpenMP over independent lines),
    median of `runs` runs of `steps` steps each on `threads` threads; returns
    (median wet-cell-steps/s, seconds per run, threads used)."""
    import oracle
    used = oracle.set_threads(threads)
    out = [cpu_run(case, steps) for _ in range(runs)]
    rates = sorted(r for r, _ in out)
    return rates[len(rates) // 2], [round(d, 2) for _, d in out], used


def cpu_host():
    """(logical CPUs of this host, model name)."""
    return os.cpu_count(), cpu_model()


def ensemble_bench(nsteps, profile=False, tune=True):
    """The batched-ensemble workload: 64 seeded meandering rivers (sides
    256-1024, own discharge, roughness, tide; sediment on), one launch per
    kernel over all of them, nsteps steps (720 = one simulated hour)."""
    import numpy as np
    from hydist_amd import synth
    from hydist_amd._lib import Ensemble, Solver

    def integrates(c):
        """The member alone on the GPU for the same steps: no error and
        velocities below 5 m/s (a stable integration of the method)."""
        g = Solver(c, device=0)
        try:
            g.init()
            g.run(0.0, nsteps, c.opts())
            u = g.get("u")
            return bool(np.isfinite(u).all() and np.abs(u).max() < 5.0)
        except Exception:
            return False
        finally:
            g.close()

    t0 = time.perf_counter()
    cases = synth.ensemble_cases(accept=integrates)
    ms = []
    for c in cases:
        g = Solver(c, device=0)
        g.init()
        ms.append(g)
    setup = time.perf_counter() - t0
    op = cases[0].opts()
    launch = None
    if tune:      # launch geometry for the batch: from the cache, or timed with the states restored after
        from hydist_amd import tuning
        launch = tuning.tune_ensemble(ms, 0.0, op, steps=10)
        launch.pop("key", None)
    ens = Ensemble(ms)
    t = ens.run(0.0, 2, op)                      # capture the step graphs
    for m in ms:
        m.sync()
    wc0 = sum(m.wet_count() for m in ms)
    ens.reset_wet_counter()
    t1 = time.perf_counter()
    t = ens.run(t, nsteps, op, counted=True)
    el = time.perf_counter() - t1
    acc = ens.wet_cell_steps()
    wc1 = sum(m.wet_count() for m in ms)
    wet_steps = acc + wc0 - wc1
    finite = sum(int(np.isfinite(m.get("z")).all() and np.isfinite(m.get("u")).all()) for m in ms)
    cells = sum((c.N + 3) * (c.M + 3) for c in cases)
    rate = wet_steps / el
    redrawn = sum(1 for c in cases if not c.name.split("_")[0].endswith(".0"))
    kt = None
    if profile:       # per-kernel launch durations over 20 more steps
        ens.profile(True)
        ens.run(t, 20, op)
        kt = ens.kernel_times()
        ens.profile(False)
    out = {"workload": f"ensemble_{len(cases)}_rivers_256-1024", "scenarios": len(cases), "padded_cells": cells,
           "redrawn_members": redrawn,
           "steps": nsteps, "sim_hours": nsteps * cases[0].dT / 3600.0, "value": rate,
           "unit": "wet-cell-steps/s", "ms_per_step": el * 1e3 / nsteps,
           "mean_wet_cells": wet_steps / nsteps, "members_finite": finite,
           "hbm_frac_step_algorithmic": rate * B_STEP_SED / HBM_PEAK, "setup_s": round(setup, 1),
           "launch": launch}
    if profile:
        out["kt"] = kt
    ens.close()
    for m in ms:
        m.close()
    return out


def timed_window(case, opts, warmup, steps, tune=True):
    """Launch geometry tuned (hydist_amd.tuning: from its cache, or timed on
    clones over the first steps from t = 0), warm-up steps, then `steps`
    steps replayed from captured HIP graphs with the state in HBM, timed; a
    clone of the state taken at the start of the timed window then runs the
    same steps again with HIP events around every launch (on the library's
    stream), so the per-kernel durations belong to exactly the timed steps.
    Returns (timing dict, solver)."""
    from hydist_amd import tuning
    from hydist_amd._lib import Solver
    g = Solver(case, device=0)
    g.init()
    launch = None
    if tune:
        launch = tuning.tune(g, 0.0, opts, steps=max(10, min(steps, 20)))
        launch.pop("key", None)
    t = g.run(0.0, warmup, opts) if warmup else 0.0
    g.sync()
    shadow = g.clone()
    t_start = t
    wc0 = g.wet_count()
    g.reset_wet_counter()
    g.sync()
    t0 = time.perf_counter()
    t = g.run(t, steps, opts, counted=True)
    g.sync()
    elapsed = time.perf_counter() - t0
    acc = g.wet_cell_steps()                     # end states of the timed steps
    wc1 = g.wet_count()
    wet_steps = acc + wc0 - wc1                  # start states of the timed steps
    shadow.profile(True)
    shadow.run(t_start, steps, opts)
    kt = shadow.kernel_times()
    shadow.profile(False)
    shadow.close()
    return {"value": wet_steps / elapsed, "ms_per_step": elapsed * 1e3 / steps, "wet_mean": wet_steps / steps,
            "kt": kt, "t": t, "launch": launch}, g


def kernel_table(kt, wet_mean):
    """Per kernel label: average launch (us) and achieved algorithmic GB/s."""
    out = {}
    for k, (ms, n) in kt.items():
        us = ms / n * 1e3
        row = {"us": round(us, 2), "launches": n}
        if k in KERNEL_BYTES:
            row["GB/s"] = round(KERNEL_BYTES[k] * wet_mean / (us * 1e-6) / 1e9, 1)
        out[k] = row
    return out


def river_line(a):
    from hydist_amd import synth
    case = synth.config(2)
    opts = case.opts()
    w, g = timed_window(case, opts, a.warmup, a.steps, tune=not a.no_tune)
    g.close()
    roof = roofline(w["kt"], w["wet_mean"])
    kernels = {k: round(ms / n * 1e3, 2) for k, (ms, n) in w["kt"].items()}

    cpu = None
    if not a.no_cpu:
        import oracle
        try:
            lib_used = os.path.basename(oracle.use_native_build())
            build = "-O3 -march=native built on this host"
        except Exception as e:     # no compiler on the host: the prebuilt checker
            lib_used, build = "liboracle.so", f"-O2 prebuilt (native build failed: {type(e).__name__})"
        ncpu, model = cpu_host()
        rate, secs, used = cpu_baseline(case, a.cpu_steps, cpu_threads())
        cpu = {"value": rate, "unit": "wet-cell-steps/s", "cores": used, "kind": "port",
               "sample": f"median of 3 runs of the first {a.cpu_steps} steps of {case.name} from t=0 "
                         f"({secs} s), CPU checker (FP64 restatement, OpenMP over lines, {used} threads; "
                         f"nproc {ncpu}, {model}; {lib_used}, {build})"}
        rate1, secs1, _ = cpu_baseline(case, a.cpu_steps_single, 1)
        cpu["single_thread"] = {"value": rate1, "cores": 1,
                                "sample": f"median of 3 runs of the first {a.cpu_steps_single} steps of "
                                          f"{case.name} from t=0 ({secs1} s), one thread"}

    ens = None
    if not a.no_ensemble:
        try:
            ens = ensemble_bench(a.ensemble_steps, tune=not a.no_tune)
        except Exception as e:  # reported, the river line still prints
            ens = {"error": f"{type(e).__name__}: {e}"}

    return {
        "metric": "ADI wet-cell-steps/s (FP64)",
        "value": w["value"],
        "unit": "wet-cell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": w["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": case.name, "N": case.N, "M": case.M, "dx_m": 20.0, "dT_s": case.dT,
                   "mean_wet_cells": round(w["wet_mean"], 1), "physics": "hydrodynamics",
                   "hbm_frac_step_algorithmic": w["value"] * B_STEP / HBM_PEAK, "launch": w["launch"]},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernel_us": kernels,
        "ensemble": ens,
    }


def sediment_line(a):
    """The 980 x 1048 river with suspended sediment and bed change every
    step, both from t = 0."""
    from hydist_amd import synth
    from hydist_amd.case import BED_EVERY_STEP
    case = synth.config(3)
    case.sediment_start = 0.0
    case.bed_change_start = 0.0
    case.bed_mode = BED_EVERY_STEP
    opts = case.opts()
    w, g = timed_window(case, opts, a.warmup, a.steps, tune=not a.no_tune)
    g.close()
    return {"metric": "ADI wet-cell-steps/s with sediment and bed change (FP64)", "value": w["value"],
            "unit": "wet-cell-steps/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
            "ms_per_step": w["ms_per_step"], "higher_is_better": True, "dtype": "f64", "data": "synthetic",
            "config": {"workload": case.name, "physics": "hydrodynamics + suspended sediment + bed change",
                       "mean_wet_cells": round(w["wet_mean"], 1),
                       "hbm_frac_step_algorithmic": w["value"] * (B_STEP_SED + 16) / HBM_PEAK},
            "roofline": roofline(w["kt"], w["wet_mean"]), "kernels": kernel_table(w["kt"], w["wet_mean"])}


def ensemble_line(a):
    e = ensemble_bench(a.ensemble_steps, profile=True, tune=not a.no_tune)
    kt = e.pop("kt")
    return {"metric": "ensemble wet-cell-steps/s (FP64)", **e, "n_gpus": 1, "higher_is_better": True,
            "dtype": "f64", "data": "synthetic",
            "roofline": roofline(kt, e["mean_wet_cells"]), "kernels": kernel_table(kt, e["mean_wet_cells"])}


def cnn_line(a):
    """The surrogate network on the river's padded grid (batch 1): forward
    with the loss (hd_cnn_eval_samples) and a training step (forward, data
    and weight gradients, Adam: hd_cnn_train_samples) on a device-resident
    sample captured from the solver."""
    from hydist_amd import cnn, synth
    from hydist_amd._lib import Solver
    case = synth.config(2)
    g = Solver(case, device=0)
    g.init()
    g.run(0.0, 10, case.opts())
    H, W = case.N + 3, case.M + 3
    net = cnn.Network(H, W, batch=1)
    net.data_alloc(1)
    net.capture(0, 0, g)
    net.capture(0, 1, g)
    out = {}
    for name, fn, mult in (("forward", lambda: net.eval_samples([0]), 1),
                           ("train_step", lambda: net.train_samples([0], 1e-6), 3)):
        fn()
        g.sync()
        t0 = time.perf_counter()
        for _ in range(a.cnn_reps):
            fn()
        el = (time.perf_counter() - t0) / a.cnn_reps
        flop = CNN_FLOP_FWD * mult * H * W
        out[name] = {"ms": el * 1e3, "TFLOP/s": flop / el / 1e12, "frac_bf16_dense_peak": flop / el / MFMA_BF16_PEAK}
    net.close()
    g.close()
    f = out["forward"]
    return {"metric": "surrogate network TFLOP/s (BF16 MFMA, FP32 accumulate)", "value": f["TFLOP/s"],
            "unit": "TFLOP/s", "n_gpus": 1, "higher_is_better": True, "dtype": "bf16", "data": "synthetic",
            "config": {"workload": f"cnn_8layer_w64_{H}x{W}", "batch": 1, "reps": a.cnn_reps,
                       "flop_per_pixel_forward": CNN_FLOP_FWD},
            "roofline": {"bound": "mfma", "achieved": f["TFLOP/s"], "peak": MFMA_BF16_PEAK / 1e12,
                         "unit": "TFLOP/s", "frac": f["frac_bf16_dense_peak"], "traffic": None},
            "passes": out}


def surrogate_line(a):
    """The surrogate-accelerated bed change in a bounded form: train the
    surrogate on (state, state + 60 steps) pairs from 8 inflow variants of
    the 980 x 1048 sediment river, then run a day
    of bed change every step three ways from the same state at hour 2 --
    plain FP64 ADI, the surrogate with an FP64 check after every call
    (sediment every step, and with morphological factor 6), and an FP64 twin
    whose inflow is scaled by 1 + 1e-12 -- and report the simulated hours per
    wall second against plain ADI with the bed difference of each."""
    from hydist_amd import surrogate
    r = surrogate.evaluate(scale=1.0, variants=8, pairs=a.sur_pairs, K=60, epochs=a.sur_epochs, days=a.sur_days,
                           check_every=1, log=lambda m: print(m, file=sys.stderr, flush=True))
    modes = r["modes"]
    best = max(modes, key=lambda k: modes[k].get("speedup", 0.0))
    m = modes[best]
    return {"metric": "surrogate speed-up (simulated hours per wall second vs plain FP64 ADI) at the stated bed "
                      "accuracy", "value": m.get("speedup"), "unit": "x", "n_gpus": 1, "higher_is_better": True,
            "vs_baseline": None, "dtype": "bf16+f64", "data": "synthetic",
            "config": {"workload": r["workload"], "mode": best, "K": 60, "samples": r["train"]["samples"],
                       "epochs": a.sur_epochs, "days": a.sur_days},
            "plain_sim_hours_per_wall_s": r["plain_sim_hours_per_wall_s"],
            "bed_error_vs_plain": m.get("bed_end"), "checks": m.get("checks"), "accepted": m.get("accepted"),
            "bed_floor_fp64_twin": r["twin_bed_end"], "plain_check_statistics": r["plain_check_statistics"],
            "bounds": r["bounds"], "val_loss": r["train"]["val_loss"],
            "val_loss_persistence": r["train"]["val_loss_persistence"], "modes": modes}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              f"(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)
    line = {"river": river_line, "sediment": sediment_line, "ensemble": ensemble_line, "cnn": cnn_line,
            "surrogate": surrogate_line}[a.workload](a)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
