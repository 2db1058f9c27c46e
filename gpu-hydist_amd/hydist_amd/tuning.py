"""Launch-geometry tuning of a case or an ensemble, cached per case.

The knobs are the library's two launch choices (hd_set_launch): the solver
grid (blocks per line, `rb_mult`) and the block width of the per-row
kernels (`bw`).  Results do not depend on them (every geometry gives the
same bits); the time per step does, by case shape and wet mask.  A tuning
is timed on copies of the state (hd_autotune clones the context; the
ensemble form restores its members afterwards), so the run it precedes is
unchanged.

The cache is a JSON file keyed by grid size, segment limit, a digest of the
wet mask at tuning time, the solver form and the target (gfx950)
($HYDIST_TUNE_CACHE, default ~/.cache/hydist_amd/launch.json); a case seen
before is set from it without timing.  An unwritable cache is skipped.
"""
from __future__ import annotations

import hashlib
import json
import os
import time
from typing import Dict, List, Optional

import numpy as np

from ._lib import Ensemble, Solver

RB_CHOICES = (1, 2, 4)
BW_CHOICES = (64, 128, 256)


def cache_path() -> str:
    return os.environ.get("HYDIST_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "hydist_amd", "launch.json"))


def case_key(g: Solver) -> str:
    """Grid, segment limit, wet-mask digest and solver form of a context."""
    wet = g.get("khouot") == 0
    digest = hashlib.sha1(np.packbits(wet).tobytes()).hexdigest()[:16]
    form = os.environ.get("HYDIST_SOLVER", "spec")
    return f"{g.N}x{g.M}/SL{g.params().segment_limit}/{digest}/{form}/gfx950"


def _load() -> Dict[str, dict]:
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(key: str, entry: dict) -> None:
    try:
        data = _load()
        data[key] = entry
        os.makedirs(os.path.dirname(cache_path()), exist_ok=True)
        tmp = cache_path() + f".{os.getpid()}"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, cache_path())
    except OSError:
        pass


def tune(g: Solver, t: float, opts, steps: int = 10, use_cache: bool = True) -> dict:
    """Set the context's launch geometry: from the cache, or by hd_autotune
    (every rb_mult x bw pair timed over `steps` steps on a clone)."""
    key = case_key(g)
    if use_cache:
        hit = _load().get(key)
        if hit:
            g.set_launch(hit["rb_mult"], hit["bw"])
            return dict(hit, cached=True, key=key)
    rb, bw, ms = g.autotune(t, opts, steps)
    entry = {"rb_mult": rb, "bw": bw, "ms_per_step": ms, "steps": steps}
    if use_cache:
        _store(key, entry)
    return dict(entry, cached=False, key=key)


def tune_ensemble(members: List[Solver], t: float, opts, steps: int = 10, use_cache: bool = True,
                  log=None) -> dict:
    """The same for an ensemble: every pair is set on all members, an
    ensemble over them captures its step graphs and times `steps` steps, and
    the members' states are restored from copies before the next pair.  The
    best pair stays set on the members (build the ensemble after this)."""
    key = "ensemble:" + hashlib.sha1("|".join(case_key(m) for m in members).encode()).hexdigest()[:16] + \
          f"/{len(members)}"
    if use_cache:
        hit = _load().get(key)
        if hit:
            for m in members:
                m.set_launch(hit["rb_mult"], hit["bw"])
            return dict(hit, cached=True, key=key)
    saved = [m.clone() for m in members]
    best: Optional[tuple] = None
    trials = []
    try:
        for rb in RB_CHOICES:
            for bw in BW_CHOICES:
                for m in members:
                    m.set_launch(rb, bw)
                ens = Ensemble(members)
                try:
                    tt = ens.run(t, 2, opts)           # capture the step graphs
                    for m in members:
                        m.sync()
                    w0 = time.perf_counter()
                    ens.run(tt, steps, opts)
                    for m in members:
                        m.sync()
                    ms = (time.perf_counter() - w0) * 1e3 / steps
                finally:
                    ens.close()
                trials.append({"rb_mult": rb, "bw": bw, "ms_per_step": ms})
                if log:
                    log(f"  ensemble launch rb_mult {rb} bw {bw}: {ms:.3f} ms per step")
                if best is None or ms < best[2]:
                    best = (rb, bw, ms)
                for m, s in zip(members, saved):
                    m.copy_state_from(s)
    finally:
        for s in saved:
            s.close()
    rb, bw, ms = best
    for m in members:
        m.set_launch(rb, bw)
    entry = {"rb_mult": rb, "bw": bw, "ms_per_step": ms, "steps": steps, "trials": trials}
    if use_cache:
        _store(key, entry)
    return dict(entry, cached=False, key=key)
