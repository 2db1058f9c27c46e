"""Surrogate-accelerated morphodynamics: a residual CNN (hydist_amd.cnn, BF16
MFMA kernels) stands in for the FP64 ADI hydrodynamics during a bed-change
phase, and the plain ADI step judges it.

The plain path is the drop-in (Solver.run: hydrodynamics + sediment + bed
change every step).  The accelerated path, per macro step of K steps:

1. the network predicts u, v, z at t + K dT from the state at t and the
   boundary forcing at t + K dT (hd_cnn_predict_into_ctx writes them into
   the context and refreshes the derived depths);
2. the sediment half of K steps (FS transport, bed change) runs on that
   frozen flow (Solver.sediment_run);
3. every `check_every` network calls one plain FP64 ADI step runs from the
   surrogate state; the fraction of wet cells whose water level that step
   moves by more than `z_bound` (the state's departure from the FP64
   dynamics: a plain step moves z by about 1 mm except at the wetting and
   drying cells, up to about 1 cm) and the relative change of the wet volume
   (mass balance) are compared with the bounds; past either bound the
   driver reverts to plain ADI for `fallback_macro` macro steps.

The flow does not depend on the bed: the reference accumulates the bed
change dH (Sediment_Transport.cu:504) and never feeds it back into the depth
h, so the bed enters the network's input only scaled out (s_b).  That also
keeps the bank cells whose dH runs away (a one-step velocity spike of tens
of m/s over a few centimetres of water drives FS far above 1, and the
hindered settling velocity wss (1 - FS)^4 of Sediment_Transport.cu:230 then
grows as FS^4) from spoiling the prediction of their neighbours.

Training data are (state, state + K steps) pairs from plain ADI runs of the
river with seeded inflow variants, kept on the device (hd_cnn_data_alloc /
hd_cnn_capture); training uses the library's own forward, backward and Adam
kernels.
"""
from __future__ import annotations

import dataclasses
import math
import time
from typing import List, Optional

import numpy as np

from . import cnn, synth
from ._lib import HydistError, Solver
from .case import BED_EVERY_STEP


def series_at(series: np.ndarray, t: float) -> float:
    """Hourly series value at time t, linear in the hour (the boundary
    kernel's interpolation, support_funcs.cu:547-565)."""
    t1 = int(t / 3600.0)
    t2 = (t - 3600.0 * t1) / 3600.0
    return float(series[t1] * (1.0 - t2) + series[t1 + 1] * t2)


def forcing(case, t: float):
    """(discharge, water level) boundary forcing of a river case at t."""
    a = case.arrays
    if "bc_left" in a and np.any(a["bc_left"]):
        return series_at(a["bc_left"], t), series_at(a["bc_right"], t)
    return series_at(a["bc_down"], t), series_at(a["bc_up"], t)


# wet runs per line allowed (the header's segment_limit; the library's
# largest): over days the wet mask fragments past the river's 20, more in
# some runs than in others
SEGMENT_LIMIT = 24

SCALES = dict(s_u=1.0, s_v=1.0, s_z=1.0, s_b=1e30, s_q=3000.0, s_h=10.0, d_u=0.02, d_v=0.02, d_z=0.01)


def river_variant(k: int, scale: float = 1.0, hours: float = 30.0, seed: int = 50):
    """The 980 x 1048 sediment river (bed change every step from hour 2)
    with a seeded inflow hydrograph: mean, amplitude and period drawn per
    variant."""
    rng = np.random.default_rng(seed + k)
    qm = 3000.0 * scale * float(rng.uniform(0.75, 1.25))
    qa = qm * float(rng.uniform(0.15, 0.4))
    c = synth.config(3, scale=scale, hours=hours, q_mean=qm, q_amp=qa, SL=SEGMENT_LIMIT)
    c.bed_mode = BED_EVERY_STEP
    c.name = f"{c.name}_v{k}"
    return c


@dataclasses.dataclass
class TrainReport:
    samples: int
    epochs: int
    train_loss: List[float]
    val_loss: float
    val_loss_persistence: float
    gen_s: float
    train_s: float


def build_dataset(net: cnn.Network, variants: int, pairs: int, K: int, warm_h: float = 2.0,
                  scale: float = 1.0, log=print):
    """Fill the network's sample store with `variants * pairs` (state,
    state + K steps) pairs from plain FP64 ADI runs; returns the number of
    samples."""
    n = variants * pairs
    net.data_alloc(n)
    idx = 0
    for v in range(variants):
        c = river_variant(v, scale=scale)
        g = Solver(c)
        g.init()
        op = c.opts()
        t = g.run(0.0, int(round(warm_h * 3600.0 / c.dT)), op)
        for _ in range(pairs):
            q, zb = forcing(c, t + K * c.dT)
            net.capture(idx, 0, g, q, zb)
            t = g.run(t, K, op)
            net.capture(idx, 1, g)
            idx += 1
        g.close()
        log(f"  variant {v}: {pairs} pairs up to t = {t / 3600.0:.2f} h")
    return n


def train(net: cnn.Network, n: int, epochs: int, batch: int, lr: float, val_frac: float = 0.1,
          seed: int = 0, log=print):
    """Adam on shuffled mini-batches of the sample store; the last val_frac
    of every variant block is held out (here: the last samples)."""
    rng = np.random.default_rng(seed)
    nval = max(1, int(n * val_frac))
    tr = np.arange(n - nval)
    va = np.arange(n - nval, n)
    hist = []
    for ep in range(epochs):
        rng.shuffle(tr)
        tot, cnt = 0.0, 0
        lr_e = lr * (0.5 * (1 + math.cos(math.pi * ep / max(1, epochs))))
        for s in range(0, len(tr) - batch + 1, batch):
            tot += net.train_samples([int(i) for i in tr[s:s + batch]], lr_e)
            cnt += 1
        hist.append(tot / max(1, cnt))
        log(f"  epoch {ep}: train loss {hist[-1]:.3e}")
    val = float(np.mean([net.eval_samples([int(i)]) for i in va]))
    return hist, val, va


def persistence_loss(net: cnn.Network, idx) -> float:
    """Loss of the 'no change' prediction (network output forced to zero)
    on the given samples: the method without the network."""
    saved = [net.get_layer(l) for l in range(net.layers)]
    w, b = saved[-1]
    net.set_layer(net.layers - 1, np.zeros_like(w), np.zeros_like(b))
    val = float(np.mean([net.eval_samples([int(i)]) for i in idx]))
    net.set_layer(net.layers - 1, w, b)
    return val


@dataclasses.dataclass
class RunReport:
    sim_hours: float
    wall_s: float
    sim_hours_per_wall_s: float
    macro_steps: int = 0
    surrogate_calls: int = 0
    checks: int = 0
    fallbacks: int = 0
    adi_steps: int = 0
    accepted: int = 0


def plain_run(case, t0: float, t1: float, g: Solver, log=None) -> RunReport:
    """The plain FP64 path over [t0, t1] (replayed graphs, one simulated day
    per call)."""
    op = case.opts()
    n = int(round((t1 - t0) / case.dT))
    day = int(round(86400.0 / case.dT))
    g.sync()
    w0 = time.perf_counter()
    t, done = t0, 0
    while done < n:
        k = min(day, n - done)
        t = g.run(t, k, op)
        done += k
        if log:
            log(f"    plain ADI: {done * case.dT / 3600.0:.0f} h")
    g.sync()
    w = time.perf_counter() - w0
    return RunReport((t1 - t0) / 3600.0, w, (t1 - t0) / 3600.0 / w, adi_steps=n)


def wet_volume(g: Solver) -> float:
    z, kh, hz = g.get("z"), g.get("khouot"), g.get("htaiz")
    m = kh == 0
    return float((hz[m] + z[m]).sum())


def surrogate_run(case, t0: float, t1: float, g: Solver, net: cnn.Network, K: int, check_every: int = 10,
                  z_bound: float = 5e-3, z_frac: float = 1e-3, mass_bound: float = 1e-4, fallback_macro: int = 5,
                  morph_substeps: int = 0, log=None) -> RunReport:
    """Bed-change phase from t0 to t1 with the network standing in for the
    hydrodynamics (module docstring).  A check passes when at most a
    fraction `z_frac` of the wet cells move by more than `z_bound` in the
    FP64 check step and the wet volume changes by at most `mass_bound`
    (relative).  The last state that passed a check is kept on the device (a
    clone); a failed check restores it and runs plain ADI from there for
    `fallback_macro` macro steps.  morph_substeps = m > 0: of the K sediment
    steps on the predicted flow only m run and the bed change of the last of
    them is repeated K - m times (morphological acceleration over the
    quasi-steady flow, hd_bed_extrapolate)."""
    op = case.opts()
    rep = RunReport(0, 0, 0)
    t = t0
    calls_since_check = 0
    adi_left = 0
    g.sync()
    w0 = time.perf_counter()
    good = g.clone()          # last verified state
    t_good = t
    while t < t1 - 1e-9:
        k = min(K, int(round((t1 - t) / case.dT)))
        rep.macro_steps += 1
        if adi_left > 0 or k < K:
            t = g.run(t, k, op)
            rep.adi_steps += k
            adi_left = max(0, adi_left - 1)
            if adi_left == 0:
                good.copy_state_from(g)
                t_good = t
            continue
        q, zb = forcing(case, t + K * case.dT)
        net.predict_into(g, q, zb)
        rep.surrogate_calls += 1
        calls_since_check += 1
        if morph_substeps > 0:
            ts = t
            if morph_substeps > 1:
                t = g.sediment_run(t, morph_substeps - 1, op)
            g.bed_snapshot()
            g.sediment_run(t, 1, op)
            g.bed_extrapolate(K - morph_substeps)
            t = ts + K * case.dT
        else:
            t = g.sediment_run(t, K, op)
        if calls_since_check >= check_every:
            calls_since_check = 0
            rep.checks += 1
            g.z_snapshot()
            t = g.run(t, 1, op)
            rep.adi_steps += 1
            dz, frac, v1, v0 = g.z_check(z_bound)
            dv = abs(v1 - v0) / max(abs(v0), 1e-30)
            ok = frac <= z_frac and dv <= mass_bound      # NaN fails too
            if log:
                log(f"    check at {t / 3600.0:.2f} h: max dz {dz:.3e} m, over {z_bound:g} m {frac:.2e}, "
                    f"dV {dv:.2e} -> {'pass' if ok else 'fail'}")
            if ok:
                rep.accepted += 1
                good.copy_state_from(g)
                t_good = t
            else:
                rep.fallbacks += 1
                g.copy_state_from(good)
                t = t_good
                adi_left = fallback_macro
    g.sync()
    good.close()
    rep.wall_s = time.perf_counter() - w0
    rep.sim_hours = (t - t0) / 3600.0
    rep.sim_hours_per_wall_s = rep.sim_hours / rep.wall_s
    return rep


def perfect_flow_run(case, t0: float, t1: float, g: Solver, K: int, morph_substeps: int = 0) -> RunReport:
    """The accelerated scheme with a perfect network: each macro step takes
    the true FP64 flow at t + K dT (a clone of the state stepped with plain
    ADI; the whole state but the concentration and the bed, FS / tFS / dH)
    in place of the prediction, then runs the sediment half on that frozen
    flow as surrogate_run does.  Slower than plain ADI (it runs it); the
    bed's departure from a plain run measures the scheme alone."""
    op = case.opts()
    rep = RunReport(0, 0, 0)
    t = t0
    g.sync()
    w0 = time.perf_counter()
    while t < t1 - 1e-9:
        k = min(K, int(round((t1 - t) / case.dT)))
        rep.macro_steps += 1
        sed = {f: g.get(f) for f in ("FS", "tFS", "dH")}
        truth = g.clone()
        truth.run(t, k, op)
        g.copy_state_from(truth)
        truth.close()
        for f, a in sed.items():
            g.set(f, a)
        if morph_substeps > 0 and k == K:
            ts = t
            if morph_substeps > 1:
                t = g.sediment_run(t, morph_substeps - 1, op)
            g.bed_snapshot()
            g.sediment_run(t, 1, op)
            g.bed_extrapolate(K - morph_substeps)
            t = ts + K * case.dT
        else:
            t = g.sediment_run(t, k, op)
    g.sync()
    rep.wall_s = time.perf_counter() - w0
    rep.sim_hours = (t - t0) / 3600.0
    rep.sim_hours_per_wall_s = rep.sim_hours / rep.wall_s
    return rep


def check_statistics(case, g: Solver, t: float, z_bound: float = 5e-3):
    """What the surrogate's check measures, for one plain FP64 step from the
    current state (left unchanged): (max |dz|, fraction of wet cells over
    z_bound, relative wet-volume change)."""
    c = g.clone()
    c.z_snapshot()
    c.run(t, 1, case.opts())
    dz, frac, v1, v0 = c.z_check(z_bound)
    c.close()
    return dz, frac, abs(v1 - v0) / max(abs(v0), 1e-30)


def bed_difference(d_ref: np.ndarray, d: np.ndarray, wet: np.ndarray) -> dict:
    """Bed-change difference over the wet cells: max, 99th and 99.9th
    percentiles of |d - d_ref| and the fraction of cells within 1 mm."""
    e = np.abs(d[wet] - d_ref[wet])
    e = np.where(np.isfinite(e), e, np.inf)
    return {"max_m": float(e.max()), "p99_m": float(np.percentile(e, 99)),
            "p999_m": float(np.percentile(e, 99.9)), "frac_within_1mm": float((e <= 1e-3).mean())}


def evaluate(scale: float = 1.0, variants: int = 8, pairs: int = 250, K: int = 60, epochs: int = 6, batch: int = 4,
             lr: float = 5e-4, days: float = 7.0, layers: int = 8, morph: int = 6, check_every: int = 10,
             z_bound: float = 5e-3, z_frac: float = 1e-3, mass_bound: float = 1e-4, log_checks: bool = False,
             extra_z_fracs=(), perfect_flow: bool = False, log=print) -> dict:
    """Train the surrogate on plain-ADI pairs and compare the accelerated
    bed-change run with the plain one on the 980 x 1048 sediment river (bed
    change every step, from hour 2): simulated hours per wall second of both, and the bed
    (dH) difference after one day and after `days` days.  Beside them, the
    FP64 method against itself: a twin run whose inflow series is scaled by
    1 + 1e-12, the smallest difference any other method could have, and the
    check's statistics on plain steps.  extra_z_fracs: further runs of the
    morphological mode with the check's fraction bound set to each value
    (the trade-off of speed against the bed's accuracy).  perfect_flow: also
    the scheme with the true flow in place of the network (perfect_flow_run)
    over the first day, with the sediment every step and with the
    morphological factor."""
    warm = 2.0
    case = synth.config(3, scale=scale, hours=warm + days * 24.0 + 1.0, SL=SEGMENT_LIMIT)
    case.bed_mode = BED_EVERY_STEP
    H, W = case.N + 3, case.M + 3
    net = cnn.Network(H, W, batch=batch, layers=layers, seed=1)
    net.set_scales(**SCALES)
    t0 = time.perf_counter()
    n = build_dataset(net, variants, pairs, K, scale=scale, log=log)
    gen_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    hist, val, va = train(net, n, epochs, batch, lr, log=log)
    train_s = time.perf_counter() - t0
    pers = persistence_loss(net, va)
    log(f"  validation loss {val:.3e} (persistence {pers:.3e})")
    tb = warm * 3600.0
    marks = [tb + min(1.0, days) * 86400.0, tb + days * 86400.0]
    op = case.opts()
    w0 = int(round(tb / case.dT))

    def fresh(c):
        g = Solver(c)
        g.init()
        g.run(0.0, w0, c.opts())
        return g

    g1 = fresh(case)
    plain_beds, plain_wall, t = [], 0.0, tb
    for m in marks:
        r = plain_run(case, t, m, g1, log=log)
        plain_wall += r.wall_s
        t = m
        plain_beds.append((g1.get("dH"), g1.get("khouot") == 0))
    plain_rate = (marks[-1] - tb) / 3600.0 / plain_wall
    log(f"  plain ADI: {(marks[-1] - tb) / 3600.0:.1f} h in {plain_wall:.1f} s")
    twin = synth.config(3, scale=scale, hours=warm + days * 24.0 + 1.0, SL=SEGMENT_LIMIT)
    twin.bed_mode = BED_EVERY_STEP
    for key in ("bc_left", "bc_down"):
        if key in twin.arrays and np.any(twin.arrays[key]):
            twin.arrays[key] = twin.arrays[key] * (1.0 + 1e-12)
            break
    g3 = fresh(twin)
    t, stats, twin_diff, twin_stop = tb, [], [None] * len(marks), None
    step6 = 6 * 3600.0
    while t < marks[-1] - 1e-9:
        try:
            t = g3.run(t, int(round(min(step6, marks[-1] - t) / case.dT)), twin.opts())
        except HydistError as e:          # e.g. more wet runs on a line than the limit
            twin_stop = f"stopped after {(t - tb) / 3600.0:.0f} h: {e}"
            log(f"  twin {twin_stop}")
            break
        stats.append(check_statistics(twin, g3, t, z_bound))
        for k, m in enumerate(marks):
            if abs(t - m) < 1e-6:
                twin_diff[k] = bed_difference(plain_beds[k][0], g3.get("dH"), plain_beds[k][1])
    g3.close()
    st = np.array(stats)
    plain_checks = {"max_dz_m": float(st[:, 0].max()), "frac_over_max": float(st[:, 1].max()),
                    "dV_max": float(st[:, 2].max()), "samples": len(stats), "every_h": 6}
    log(f"  plain check statistics: {plain_checks}")
    modes = {}
    runs = [("full_sediment_steps", 0, z_frac), (f"morph_{morph}", morph, z_frac)]
    runs += [(f"morph_{morph}_z_frac_{zf:g}", morph, zf) for zf in extra_z_fracs]
    for label, msub, zf in runs:
        g2 = fresh(case)
        t, wall, diffs, calls, acc, fb, adi, checks = tb, 0.0, [], 0, 0, 0, 0, 0
        stop = None
        for k, m in enumerate(marks):
            try:
                sur = surrogate_run(case, t, m, g2, net, K, check_every=check_every, z_bound=z_bound,
                                    z_frac=zf, mass_bound=mass_bound, morph_substeps=msub,
                                    log=log if log_checks else None)
            except HydistError as e:
                stop = f"stopped in the segment ending at {(m - tb) / 3600.0:.0f} h: {e}"
                log(f"  surrogate ({label}) {stop}")
                break
            t = m
            wall += sur.wall_s
            calls += sur.surrogate_calls
            acc += sur.accepted
            fb += sur.fallbacks
            adi += sur.adi_steps
            checks += sur.checks
            diffs.append(bed_difference(plain_beds[k][0], g2.get("dH"), plain_beds[k][1]))
        if stop:
            modes[label] = {"stopped": stop, "bed": diffs}
            g2.close()
            continue
        rate = (marks[-1] - tb) / 3600.0 / wall
        modes[label] = {"z_frac": zf, "sim_hours_per_wall_s": rate, "speedup": rate / plain_rate,
                        "surrogate_calls": calls,
                        "checks": checks, "accepted": acc, "fallbacks": fb, "adi_steps": adi,
                        "bed_1d": diffs[0], "bed_end": diffs[-1]}
        log(f"  surrogate ({label}): {wall:.1f} s, speed-up {rate / plain_rate:.2f}, {calls} calls, "
            f"{acc}/{checks} checks passed, bed at the end {diffs[-1]}")
        g2.close()
    perfect = {}
    if perfect_flow:
        for label, msub in (("sediment_every_step", 0), (f"morph_{morph}", morph)):
            g4 = fresh(case)
            perfect_flow_run(case, tb, marks[0], g4, K, morph_substeps=msub)
            perfect[label] = bed_difference(plain_beds[0][0], g4.get("dH"), plain_beds[0][1])
            log(f"  perfect flow ({label}): bed after 1 day {perfect[label]}")
            g4.close()
    g1.close()
    net.close()
    return {
        "perfect_flow_bed_1d": perfect,
        "workload": f"{case.name} bed change {days:g} d (every step), K={K}",
        "plain_sim_hours_per_wall_s": plain_rate, "plain_check_statistics": plain_checks,
        "twin_bed_1d": twin_diff[0], "twin_bed_end": twin_diff[-1], "twin_stopped": twin_stop, "modes": modes,
        "bounds": {"z_bound_m": z_bound, "z_frac": z_frac, "mass_bound": mass_bound, "check_every": check_every},
        "train": {"samples": n, "epochs": epochs, "loss": hist, "val_loss": val,
                  "val_loss_persistence": pers, "gen_s": gen_s, "train_s": train_s},
    }


if __name__ == "__main__":
    import argparse
    import json

    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--variants", type=int, default=8)
    ap.add_argument("--pairs", type=int, default=250)
    ap.add_argument("--K", type=int, default=60)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--days", type=float, default=7.0)
    ap.add_argument("--morph", type=int, default=6)
    ap.add_argument("--layers", type=int, default=8)
    ap.add_argument("--check-every", type=int, default=10)
    ap.add_argument("--z-bound", type=float, default=5e-3)
    ap.add_argument("--z-frac", type=float, default=1e-3)
    ap.add_argument("--mass-bound", type=float, default=1e-4)
    ap.add_argument("--log-checks", action="store_true")
    ap.add_argument("--extra-z-fracs", default="", help="comma-separated further check bounds (morphological mode)")
    ap.add_argument("--perfect-flow", action="store_true", help="also the scheme with the true flow (first day)")
    a = ap.parse_args()
    r = evaluate(scale=a.scale, variants=a.variants, pairs=a.pairs, K=a.K, epochs=a.epochs, batch=a.batch,
                 days=a.days, layers=a.layers, morph=a.morph, check_every=a.check_every, z_bound=a.z_bound,
                 z_frac=a.z_frac, mass_bound=a.mass_bound, log_checks=a.log_checks,
                 extra_z_fracs=tuple(float(x) for x in a.extra_z_fracs.split(",") if x),
                 perfect_flow=a.perfect_flow,
                 log=lambda s: print(s, flush=True))
    print(json.dumps(r))
