"""GPU parity: libhydist (hand-written gfx950 kernels, through the C ABI)
against the CPU checker, bit for bit, on seeded synthetic cases small enough
for the checker to finish in seconds; plus error paths and run == step."""
import os

import numpy as np
import pytest

import oracle
from hydist_amd import synth
from hydist_amd._lib import HydistError, Solver, thomas_batch
from hydist_amd.case import BED_EVERY_STEP, QB_REFERENCE

from _parity import assert_same_state

pytestmark = pytest.mark.gpu


def small_channel():
    return synth.channel_case(N=40, M=12, nsteps=60, name="channel_40x12")


def small_river():
    return synth.config(2, scale=0.1)


def small_river_sediment(qb_mode=None):
    c = synth.config(3, scale=0.1)
    c.sediment_start = 0.0
    c.bed_change_start = 10.0
    c.bed_mode = BED_EVERY_STEP
    if qb_mode is not None:
        c.qb_mode = qb_mode
    return c


def pair(case):
    o = oracle.Oracle(case)
    g = Solver(case)
    o.init()
    g.init()
    return o, g


@pytest.mark.parametrize("spec", [False, True])
def test_thomas_batch_bitwise(spec):
    rng = np.random.default_rng(0)
    systems = []
    for n in (1, 2, 3, 7, 33, 34, 64, 333, 510, 511, 1961):
        A = rng.normal(size=n); C = rng.normal(size=n)
        B = 3 + np.abs(A) + np.abs(C); D = rng.normal(size=n)
        systems.append((A, B, C, D))
    for x, s in zip(thomas_batch(systems, spec=spec), systems):
        assert np.array_equal(x, oracle.thomas(*s))


@pytest.mark.parametrize("cscale,margin", [(0.05, 0.05), (1.0, 2.0), (1.0, 100.0)])
def test_thomas_spec_bitwise_decaying_rhs(cscale, margin):
    """Right-hand sides that vanish after the first rows, as ahead of the
    waves of a flow starting from rest: bp decays geometrically, the
    speculative lanes' estimates never merge with it bit for bit, and almost
    every lane of hd_thomas_batch_spec is repaired with its divisors known
    (the reciprocal form of the division: bp stays normal over all 1961
    rows at the slowest decay and reaches the underflow range, where that
    form divides instead, within a few hundred rows at the others).  The
    result must still be the serial solve's."""
    rng = np.random.default_rng(1)
    systems = []
    for n in (200, 509, 1961):
        A = rng.uniform(0.5, 1.0, size=n) * rng.choice([-1.0, 1.0], size=n)
        C = cscale * rng.uniform(0.5, 1.0, size=n) * rng.choice([-1.0, 1.0], size=n)
        B = (np.abs(A) + np.abs(C)) * (1.0 + margin) * rng.uniform(1.0, 1.05, size=n)
        D = np.zeros(n)
        D[:4] = rng.normal(size=4)
        systems.append((A, B, C, D))
    for x, s in zip(thomas_batch(systems, spec=True), systems):
        assert np.array_equal(x, oracle.thomas(*s))


@pytest.mark.parametrize("make", [small_channel, small_river, small_river_sediment])
def test_init_parity(make):
    case = make()
    o, g = pair(case)
    assert_same_state(o, g, "after init")
    assert o.hmax() == g.hmax()


@pytest.mark.parametrize("make,nsteps", [(small_channel, 40), (small_river, 40), (small_river_sediment, 30)])
def test_halfsteps_and_steps_bitwise(make, nsteps):
    case = make()
    o, g = pair(case)
    op = case.opts()
    t = 0.0
    for which in (0, 1):
        t = t + case.dT * 0.5
        o.halfstep(which, t, op)
        g.halfstep(which, t, op)
        assert_same_state(o, g, f"half-step {which}")
    for k in range(nsteps):
        tn = o.step(t, op)
        g.step(t, op)
        t = tn
        if k % 10 == 9 or k == nsteps - 1:
            assert_same_state(o, g, f"step {k + 2}")


def test_sediment_starting_mid_run_bitwise():
    """Suspended sediment switched on after 12 steps and the bed change
    after 24, in one hd_run: the first sweep and the first bed-load flux run
    over the whole grid, later ones over the active tiles; every field,
    including VTH, Kx/Ky, Qbx/Qby, FS/tFS on land, matches the checker."""
    case = synth.config(3, scale=0.25)
    case.sediment_start = 12 * case.dT
    case.bed_change_start = 24 * case.dT
    case.bed_mode = BED_EVERY_STEP
    o, g = pair(case)
    op = case.opts()
    o.run(0.0, 40, op)
    g.run(0.0, 40, op)
    assert_same_state(o, g, "sediment from step 13, bed change from step 25")


def test_fs_edges_reading_new_rows_bitwise():
    """The FS edge interpolations with velocities that make row 2 read row N
    (after row N was interpolated), row N read row 2 (before row 2 was), and
    the column-M cells of rows N and 2 read those rows' new values: the
    reference's kernel order (row N, row 2, column M) against the checker's
    segment finder, through hd_prepare."""
    case = synth.channel_case(N=16, M=40, nsteps=10, name="channel_16x40")  # column M spans rows 1..N+2
    o, g = pair(case)
    N, M = case.N, case.M
    P = g.params()
    rng = np.random.default_rng(11)
    fs = rng.uniform(0.0, 1.0, (N + 3, M + 3))
    t_u = g.get("t_u").reshape(N + 3, M + 3).copy()
    t_v = g.get("t_v").reshape(N + 3, M + 3).copy()

    def speed(cells_per_step, dx):
        # u = t_u = s: travel (u + t_u) dT / 4 = cells_per_step * dx
        return 2.0 * cells_per_step * dx / P.dT

    j = np.arange(1, M)
    # row 2: sources N - 1 and N on alternate columns (reads row N)
    t_u[2, 1:M] = np.where(j % 2 == 0, speed(N - 2.5, P.dX), speed(N - 1.5, P.dX))
    # row N: sources 2 (reads row 2) on every third column, one cell back elsewhere
    t_u[N, 1:M] = np.where(j % 3 == 0, speed(N - 3.5, P.dX), speed(0.5, P.dX))
    t_u[N, 4] = -1.0  # a column where row N does not move
    # column M, rows N and 2: sources M - 2 (reads the rows' columns M-2, M-1)
    t_v[:, M] = speed(0.25, P.dY)
    t_v[N, M] = t_v[2, M] = speed(0.5, P.dY)
    src2 = ((2 + (t_u[2, 1:M] * 0.5 * P.dT) / P.dX).astype(int))
    assert {N - 1, N} <= set(src2.tolist())
    for name, arr in (("FS", fs), ("t_u", t_u), ("u", t_u), ("t_v", t_v), ("v", t_v)):
        o.set(name, arr)
        g.set(name, arr)
    o.find_limits()
    g.prepare()
    assert np.array_equal(o.get("FS"), g.get("FS"))
    for name in ("daui", "cuoii", "moci", "dauj", "cuoij", "mocj"):
        assert np.array_equal(o.get(name), g.get(name)), name


def short_edge_runs_case():
    """A channel with dry cells (3, 8) and (4, 12): columns 8 and 12 start
    with the runs [2, 2] (followed by a run from 4) and [2, 3] at the open
    inflow edge, whose margins the position-1 cell writes in order."""
    case = synth.channel_case(N=24, M=30, nsteps=40, name="short_edge_runs")
    h = case.arrays["h"].copy()
    h[2:4, 7:9] = 0.0
    h[3:5, 11:13] = 0.0
    case.arrays["h"] = synth.remove_pinches(h)
    return case


def test_short_runs_at_open_edge_bitwise():
    """Margins of runs [2, 2] and [2, 3] at an open edge (2 E(2) - E(3) with
    the old values of the run's own last cell and of the next run's first
    margin) match the checker step by step."""
    case = short_edge_runs_case()
    o, g = pair(case)
    SL = o.params().segment_limit
    mocj, dauj, cuoij = o.get("mocj"), o.get("dauj").reshape(-1, SL), o.get("cuoij").reshape(-1, SL)
    assert mocj[8] >= 2 and (dauj[8, 0], cuoij[8, 0], dauj[8, 1]) == (2, 2, 4)
    assert mocj[12] >= 1 and (dauj[12, 0], cuoij[12, 0]) == (2, 3)
    op = case.opts()
    t = 0.0
    for k in range(30):
        tn = o.step(t, op)
        g.step(t, op)
        t = tn
        if k % 10 == 9:
            assert_same_state(o, g, f"short edge runs, step {k + 1}")


def test_bed_change_reference_mode():
    case = small_river_sediment(qb_mode=QB_REFERENCE)
    o, g = pair(case)
    op = case.opts()
    t = o.run(0.0, 6, op)
    g.run(0.0, 6, op)
    assert_same_state(o, g, "reference Qb mode")


def test_run_equals_step():
    case = small_river()
    g1, g2 = Solver(case), Solver(case)
    g1.init()
    g2.init()
    op = case.opts()
    t = 0.0
    for _ in range(25):
        t = g1.step(t, op)
    t2 = g2.run(0.0, 25, op)
    assert t == t2
    assert_same_state(g1, g2, "graph replay vs step")


def open_basin(types, N=30, M=26, nsteps=40):
    """A basin wet up to all four edges, with the boundary series and no
    narrow-channel mode: each edge line's first run is the whole edge, so
    every corner cell is written by two edges, and the sweeps' runs write the
    boundary values themselves.  types: boundary type of the edges up
    (column M), down (column 2), left (row 2), right (row N); 2 is a
    discharge (with a sediment concentration), anything else a level."""
    rng = np.random.default_rng(5)
    hn = 4.0 + rng.normal(0.0, 0.05, (N, M))
    total_time = 4
    hours = np.arange(total_time, dtype=np.float64)
    series = {}
    for e, (key, cc) in enumerate((("bc_up", "CC_u"), ("bc_down", "CC_d"), ("bc_left", "CC_l"),
                                   ("bc_right", "CC_r"))):
        if types[e] == 2:
            series[key] = 0.2 + 0.05 * hours
            series[cc] = 0.01 + 0.002 * hours
        else:
            series[key] = 0.05 * np.sin(hours + e)
    header = synth.make_case_header(10.0, 10.0, 2.0, 20, H_TINH=0.02, heso=0.94)
    case = synth._finish("open_basin", N, M, 20, header, hn, 0.025, 1e-6, types, series, total_time,
                         nsteps * 2.0)
    case.sediment_start = 0.0
    return case


@pytest.mark.parametrize("types", [(0, 2, 2, 1), (1, 1, 1, 1), (2, 2, 2, 2)])
def test_open_edges_and_corners_bitwise(types):
    """Boundary values written by the sweeps' runs (no k_boundary launch)
    against the checker's Update_Boundary_Value, at every edge and at corners
    where a level meets a discharge, two levels or two discharges meet, with
    the time interpolated within the hour: each half-step, then 30 steps in
    one run."""
    case = open_basin(types)
    o, g = pair(case)
    op = case.opts()
    t = 0.0
    g.profile(True)
    for which in (0, 1):
        t = t + case.dT * 0.5
        o.halfstep(which, t, op)
        g.halfstep(which, t, op)
        assert_same_state(o, g, f"half-step {which}")
    times = g.kernel_times()
    g.profile(False)
    assert "boundary" not in times and times["adi_spec_uz"][1] == 1
    o.run(t, 30, op)
    g.run(t, 30, op)
    assert_same_state(o, g, f"open basin {types}, 31 steps")


def overflow_case():
    """A channel whose rows are cut by three dry walls into four wet runs,
    with segment_limit = 3."""
    case = synth.channel_case(N=30, M=40, nsteps=10, SL=3, name="overflow")
    h = case.arrays["h"].copy()
    for k in (10, 18, 26):
        h[:, k:k + 2] = 0.0
    case.arrays["h"] = synth.remove_pinches(h)
    case.arrays["boundary_type"] = np.zeros((4, 3), np.int32)
    case.arrays["bienQ"][:] = 0
    return case


def test_segment_overflow_raises():
    case = overflow_case()
    o = oracle.Oracle(case)
    with pytest.raises(oracle.OracleError):
        o.init()
    g = Solver(case)
    with pytest.raises(HydistError) as e:
        g.init()
    assert e.value.code == -5


def test_boundary_series_range_error():
    case = small_channel()
    g = Solver(case)
    g.init()
    with pytest.raises(HydistError) as e:
        g.run(case.total_time * 3600.0, 1, case.opts())
    assert e.value.code == -7


def test_analytic_channel_boundaries():
    case = small_channel()
    case.boundary_mode = 0   # the reference rule: channel flag -> analytic tide
    o, g = pair(case)
    op = case.opts()
    o.run(0.0, 10, op)
    g.run(0.0, 10, op)
    assert_same_state(o, g, "analytic channel mode")


def test_long_runs_bitwise():
    """Runs of up to 297 cells (the 300 x 60 channel's UZ lines): the run
    solver's multi-chunk path (510 rows per chunk) against the checker."""
    case = synth.channel_case(N=300, M=60, nsteps=40, name="channel_300x60")
    o, g = pair(case)
    op = case.opts()
    t = 0.0
    for k in range(20):
        tn = o.step(t, op)
        g.step(t, op)
        t = tn
    assert_same_state(o, g, "channel 300x60, 20 steps")


def test_long_sediment_runs_bitwise():
    """A straight river 700 x 60 along i with sediment and bed change from
    t = 0: column runs of about 700 cells, so the speculative sediment solve
    along columns takes three or more 256-row chunks and the UZ sweep three
    510-row chunks, against the checker bit for bit."""
    case = synth.river_case(N=700, M=60, hours=1.0, amp=0.0, wmin=20.0, wmax=30.0, n_islands=0,
                            q_mean=300.0, q_amp=0.0, tide_amp=0.2, sediment=True, name="straight_700x60_sed")
    case.sediment_start = 0.0
    case.bed_change_start = 10.0
    case.bed_mode = BED_EVERY_STEP
    o, g = pair(case)
    wet_rows = int((o.get("khouot")[:, 30] == 0).sum())
    assert wet_rows > 600, wet_rows
    op = case.opts()
    t = 0.0
    for k in range(12):
        tn = o.step(t, op)
        g.step(t, op)
        t = tn
    assert_same_state(o, g, "straight river 700x60 with sediment, 12 steps")


def test_full_river_bitwise():
    """The 980 x 1048 river itself, a few steps, against the checker."""
    oracle.set_threads(min(16, len(os.sched_getaffinity(0))))
    case = synth.config(2)
    o, g = pair(case)
    op = case.opts()
    o.run(0.0, 4, op)
    g.run(0.0, 4, op)
    assert_same_state(o, g, "river 980x1048, 4 steps")


def test_full_river_sediment_bitwise():
    """The 980 x 1048 river with suspended sediment from t = 0 and the bed
    updated every step, 120 steps (10 simulated minutes), against the
    checker: speculative Thomas in both sweeps, the fused state copy and
    tables, FS edges and bed change at full size."""
    oracle.set_threads(min(16, len(os.sched_getaffinity(0))))
    case = synth.config(3)
    case.sediment_start = 0.0
    case.bed_change_start = 0.0
    case.bed_mode = BED_EVERY_STEP
    o, g = pair(case)
    op = case.opts()
    o.run(0.0, 120, op)
    g.run(0.0, 120, op)
    assert_same_state(o, g, "river 980x1048 with sediment and bed change, 120 steps")


def full_river():
    return synth.config(2)


def long_channel():
    # 600 cells along the UZ lines: runs of 1200 rows span three LDS chunks
    return synth.channel_case(N=600, M=60, nsteps=200)


@pytest.mark.parametrize("form,make,steps", [("line", small_river, 30), ("rows", small_river, 30),
                                             ("line", small_river_sediment, 30), ("rows", small_river_sediment, 30),
                                             ("run", small_river, 30), ("run", full_river, 60),
                                             ("run", long_channel, 60)])
def test_solver_forms_identical(form, make, steps, monkeypatch):
    """Each alternative solver form (one wave per run with the serial
    Thomas, one wave per line, rows kernel + lane per line), hydrodynamics
    and sediment, against the CPU checker, bit for bit (the default form,
    one wave per run with the speculative parallel Thomas, is checked the
    same way by the tests above)."""
    case = make()
    ref = oracle.Oracle(case)
    monkeypatch.setenv("HYDIST_SOLVER", form)
    alt = Solver(case)
    monkeypatch.delenv("HYDIST_SOLVER")
    ref.init()
    alt.init()
    op = case.opts()
    ref.run(0.0, steps, op)
    alt.run(0.0, steps, op)
    assert_same_state(ref, alt, f"checker vs {form}")


@pytest.mark.parametrize("form,make,steps", [("window", small_river, 40), ("scan", small_river, 40),
                                             ("window", full_river, 40), ("scan", full_river, 20),
                                             ("window", small_river_sediment, 30)])
def test_wetting_forms_identical(form, make, steps, monkeypatch):
    """The wetting pass's three ways -- candidates visited with their own
    loads (up to 8 per line), replayed from prefetched windows, or the whole
    line scanned -- give the checker's bits (HYDIST_WETTING forces one way for
    every line), each against the CPU checker, bit for bit."""
    case = make()
    ref = oracle.Oracle(case)
    monkeypatch.setenv("HYDIST_WETTING", form)
    alt = Solver(case)
    monkeypatch.delenv("HYDIST_WETTING")
    ref.init()
    alt.init()
    op = case.opts()
    ref.run(0.0, steps, op)
    alt.run(0.0, steps, op)
    assert_same_state(ref, alt, f"checker vs wetting {form}")


def test_clone_copies_state_and_steps_identically():
    case = small_river()
    g = Solver(case)
    g.init()
    op = case.opts()
    t = g.run(0.0, 7, op)
    c = g.clone()
    assert_same_state(g, c, "clone")
    g.run(t, 9, op)
    c.run(t, 9, op)
    assert_same_state(g, c, "clone after 9 more steps")


def test_autotuned_launch_geometry_same_bits():
    case = small_river()
    a, b = Solver(case), Solver(case)
    a.init()
    b.init()
    op = case.opts()
    t = a.run(0.0, 5, op)
    b.run(0.0, 5, op)
    rb, bw, ms = a.autotune(t, op, steps=4)
    assert rb in (1, 2, 4) and bw in (64, 128, 256) and ms > 0
    assert_same_state(a, b, "autotune leaves the state alone")
    b.set_launch(4, 64)
    a.run(t, 12, op)
    b.run(t, 12, op)
    assert_same_state(a, b, "tuned vs other geometry")
