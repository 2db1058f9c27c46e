"""GPU parity over the workloads' real windows, bit for bit against the CPU
checker at hourly checkpoints (the reference's own output interval,
Engine.py:398-420):

* the 300 x 60 channel, all 3600 steps;
* the 980 x 1048 river, all 4320 steps (6 simulated hours);
* the 980 x 1048 river with suspended sediment from hour 1 and bed change
  from hour 2, through hour 3, with the bed change once per period as
  Engine.py:364 triggers it and every step;
* four members of the seeded 64-river ensemble draw (sides 256-1024),
  stepped together as one ensemble with sediment and bed change, each
  member against its own checker.
"""
import sys

import pytest

import oracle
from hydist_amd import synth
from hydist_amd._lib import Ensemble, Solver
from hydist_amd.case import BED_EVERY_STEP, BED_PERIODIC

from _parity import assert_same_state

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_compare(case, chunks, label):
    """Step the checker and the library chunk by chunk (steps per chunk),
    comparing every field after each chunk."""
    o, g = oracle.Oracle(case), Solver(case)
    o.init()
    g.init()
    op = case.opts()
    t = 0.0
    done = 0
    for n in chunks:
        t_o = o.run(t, n, op)
        t_g = g.run(t, n, op)
        assert t_o == t_g
        t = t_o
        done += n
        say(f"{label}: step {done}, t = {t:.0f} s")
        assert_same_state(o, g, f"{label} step {done}")
    o.close()
    g.close()
    return t


def test_channel_300x60_full_window():
    case = synth.config(1)
    t = run_compare(case, [900] * 4, "channel 300x60")
    assert t == 3600 * case.dT


def test_river_980x1048_full_window():
    case = synth.config(2)
    t = run_compare(case, [720] * 6, "river 980x1048")
    assert t == 6 * 3600.0


@pytest.mark.parametrize("bed_mode", [BED_PERIODIC, BED_EVERY_STEP])
def test_sediment_river_from_hour1_and_hour2(bed_mode):
    case = synth.config(3)
    case.sediment_start = 3600.0
    case.bed_change_start = 7200.0
    case.bed_mode = bed_mode
    t = run_compare(case, [720, 720, 720], f"river+sediment bed_mode {bed_mode}")
    assert t == 3 * 3600.0


def test_ensemble_members_against_checker():
    """Members 0-3 of the seeded draw, one ensemble launch per kernel, 72
    steps with sediment and bed change every step from step 31."""
    cases = [synth.ensemble_member(k) for k in range(4)]
    sides = [(c.N, c.M) for c in cases]
    assert all(256 <= n <= 1024 and 256 <= m <= 1024 for n, m in sides)
    members = [Solver(c) for c in cases]
    for m in members:
        m.init()
    ens = Ensemble(members)
    op = cases[0].opts(bed_mode=BED_EVERY_STEP, bed_change_start=150.0)
    t = 0.0
    checkers = []
    for c in cases:
        o = oracle.Oracle(c)
        o.init()
        checkers.append(o)
    for chunk in (36, 36):
        t_g = ens.run(t, chunk, op)
        for k, (o, m) in enumerate(zip(checkers, members)):
            t_o = o.run(t, chunk, op)
            assert t_o == t_g
            assert_same_state(o, m, f"member {k} {sides[k]} at t = {t_o:.0f} s")
        t = t_g
        say(f"ensemble members {sides}: t = {t:.0f} s")
    ens.close()
    for o, m in zip(checkers, members):
        o.close()
        m.close()
