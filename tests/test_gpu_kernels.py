"""Per-kernel GPU parity: every hd_k_<reference kernel> entry point against
the checker's restatement of that kernel, bit for bit, on a state both reach
by stepping from the same inputs (and, for the mask kernels, after the same
perturbation).  The step tests cover the kernels in sequence; these pin each
one alone."""
import numpy as np
import pytest

import oracle
from hydist_amd import synth
from hydist_amd._lib import Solver
from hydist_amd.case import BED_EVERY_STEP, QB_FIXED, QB_REFERENCE

from _parity import diff_fields, diff_segments

pytestmark = pytest.mark.gpu


def river_sediment():
    c = synth.config(3, scale=0.1)
    c.sediment_start = 0.0
    c.bed_change_start = 10.0
    c.bed_mode = BED_EVERY_STEP
    return c


def stepped(make=river_sediment, steps=12):
    case = make()
    o, g = oracle.Oracle(case), Solver(case)
    o.init()
    g.init()
    op = case.opts()
    t = o.run(0.0, steps, op)
    g.run(0.0, steps, op)
    g.sync()
    assert same(o, g, ("u", "v", "z", "FS", "dH", "khouot")) == []
    return case, o, g, t


def same(o, g, names):
    return diff_fields(o.get, g.get, names)


def both(o, g, name, arr):
    o.set(name, arr)
    g.set(name, arr)


def test_onetime_init():
    case, o, g, t = stepped()
    o.kernel("onetime_init")
    g.kernel("Onetime_init")
    assert same(o, g, ("khouot", "H_moi", "htaiz", "htaiz_bd", "Kx1", "Ky1")) == []


def test_find_calculation_limits_and_fs_edges():
    case, o, g, t = stepped()
    kh = o.get("khouot")
    rng = np.random.default_rng(3)
    wet = np.argwhere(kh == 0)
    for i, j in wet[rng.choice(len(wet), 40, replace=False)]:
        kh[i, j] = 1                        # dried cells split runs
    both(o, g, "khouot", kh)
    o.kernel("find_limits")
    g.kernel("Find_Calculation_limits")
    assert diff_segments(o.get, g.get) == []
    assert same(o, g, ("moci", "mocj", "FS", "tFS")) == []


def test_htuongdoi():
    case, o, g, t = stepped()
    z = o.get("z")
    z += 0.01 * np.sin(np.arange(z.size)).reshape(z.shape)
    both(o, g, "z", z)
    o.kernel("htuongdoi")
    g.kernel("Htuongdoi")
    assert same(o, g, ("Htdu", "Htdv")) == []


def test_preprocess_data():
    case, o, g, t = stepped()
    o.kernel("preprocess")
    g.kernel("preprocess_data")
    assert same(o, g, ("hi",)) == [] and o.hmax() == g.hmax()


@pytest.mark.parametrize("analytic", [0, 1])
def test_update_boundary_value(analytic):
    case, o, g, t = stepped()
    tb = t + 1234.5
    o.kernel("channel_boundary" if analytic else "boundary", t=tb)
    g.kernel("Update_Boundary_Value", tb, analytic)
    assert same(o, g, ("t_z", "ubt", "ubp", "vbt", "vbd", "FS", "t_u", "t_v")) == []


@pytest.mark.parametrize("isU,shift", [(1, 0), (1, 1), (0, 0), (0, 1)])
def test_half_step_solve(isU, shift):
    case, o, g, t = stepped()
    N, M = case.N, case.M
    if isU:
        o.kernel("uz_solve", 2 + shift, M - shift)
        g.kernel("UZSolver", shift)
    else:
        o.kernel("vz_solve", 2 + shift, N)
        g.kernel("VZSolver", shift)
    assert same(o, g, ("t_z", "t_u" if isU else "t_v")) == []


@pytest.mark.parametrize("isU", [1, 0])
def test_matrix_coeff_rows(isU):
    """The interleaved continuity / momentum rows with their end rows and
    the system sizes, compared on every system row."""
    case, o, g, t = stepped()
    N, M, SL = case.N, case.M, case.segment_limit
    o.kernel("uz_solve" if isU else "vz_solve", 2, M if isU else N)
    AA, BB, CC, DD, SN = g.rows("UZSolver_calculate_matrix_coeff" if isU else "VZSolver_calculate_matrix_coeff", 0)
    ref = {k: o.scratch(k) for k in ("AA", "BB", "CC", "DD")}
    oSN = o.get("SN", like=np.zeros((max(N, M) + 3) * SL, np.int32))
    stride = 2 * (N if isU else M) + 1
    moc, dau, cuoi = (o.get("mocj"), o.get("dauj"), o.get("cuoij")) if isU else (o.get("moci"), o.get("daui"), o.get("cuoii"))
    nrows = 0
    for line in range(2, (M if isU else N) + 1):
        for s in range(moc[line]):
            first = dau[line, s]
            sn = oSN[line * SL + s]
            assert SN[line * SL + s] == sn, (line, s)
            for r in range(sn + 1):
                k = line * stride + 2 * first + r
                for name, got in (("AA", AA), ("BB", BB), ("CC", CC), ("DD", DD)):
                    if r == 0 and name == "AA" or r == sn and name == "CC":
                        continue            # never read by the Thomas solve
                    assert got[k] == ref[name][k], (name, line, s, r)
                nrows += 1
    assert nrows > 1000


@pytest.mark.parametrize("isU", [1, 0])
def test_normalize(isU):
    case, o, g, t = stepped()
    o.kernel("normalize", isU)
    g.kernel("Normalize", isU)
    assert same(o, g, ("t_u", "t_v")) == []


@pytest.mark.parametrize("isU", [1, 0])
def test_explicit_velocity_and_margins(isU):
    case, o, g, t = stepped()
    if isU:
        o.kernel("solve_u", 2, case.M)
        o.kernel("margins_u", 2, case.M)
        g.kernel("solveU")
    else:
        o.kernel("solve_v", 2, case.N)
        o.kernel("margins_v", 2, case.N)
        g.kernel("solveV")
    assert same(o, g, ("t_u", "t_v")) == []


@pytest.mark.parametrize("horizontal", [1, 0])
def test_reset_states_wetting_and_drying(horizontal):
    """update_h_moi, Reset_states, update_uvz and Htuongdoi after a level
    perturbation that both opens dried cells and dries shallow ones."""
    case, o, g, t = stepped()
    kh, z, htaiz = o.get("khouot"), o.get("z"), o.get("htaiz")
    rng = np.random.default_rng(5 + horizontal)
    wet = np.argwhere(kh == 0)
    pick = wet[rng.choice(len(wet), 60, replace=False)]
    for i, j in pick[:30]:
        kh[i, j] = 1                         # dried, re-wettable
    tz = z + 0.02 * rng.standard_normal(z.shape)
    for i, j in pick[30:]:
        tz[i, j] = -htaiz[i, j] + 0.01       # shallow: dries
    both(o, g, "khouot", kh)
    both(o, g, "t_z", tz)
    o.kernel("update_h_moi")
    o.kernel("reset_states_h" if horizontal else "reset_states_v")
    o.kernel("update_uvz", 0, 0)
    o.kernel("htuongdoi")
    g.kernel("Reset_states", horizontal, 0, 0)
    names = ("khouot", "t_z", "H_moi", "FS", "t_u", "t_v", "z", "u", "v", "Htdu", "Htdv")
    assert same(o, g, names) == []
    assert (o.get("khouot") != kh).sum() > 20          # cells opened and closed


def sediment_ready(case, o, g):
    o.kernel("find_vth")
    o.kernel("heso_k")
    g.kernel("Find_VTH_hesoK")
    assert same(o, g, ("VTH", "Kx", "Ky")) == []


def test_find_vth_hesok():
    case, o, g, t = stepped()
    sediment_ready(case, o, g)


@pytest.mark.parametrize("which,ketdinh", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_fs_sweep(which, ketdinh):
    case, o, g, t = stepped()
    sediment_ready(case, o, g)
    s_start = case.bed_change_start
    o.kernel("sediment_j" if which == 0 else "sediment_i", ketdinh, 0, int(s_start), t=t)
    g.kernel("FSj_solve" if which == 0 else "FSi_solve", t, s_start, ketdinh)
    assert same(o, g, ("FS", "tFS")) == []


@pytest.mark.parametrize("which", [0, 1])
def test_fs_rows(which):
    case, o, g, t = stepped()
    sediment_ready(case, o, g)
    N, M, SL = case.N, case.M, case.segment_limit
    s_start = case.bed_change_start
    o.kernel("sediment_j" if which == 0 else "sediment_i", 0, 0, int(s_start), t=t)
    rows = g.rows("Scan_FSj" if which == 0 else "Scan_FSi", t, s_start, 0)
    ref = [o.scratch(k) for k in ("AA", "BB", "CC", "DD")]
    stride = (N if which == 0 else M) + 3
    moc, dau, cuoi = (o.get("mocj"), o.get("dauj"), o.get("cuoij")) if which == 0 else \
        (o.get("moci"), o.get("daui"), o.get("cuoii"))
    n = 0
    for line in range(3, (M if which == 0 else N)):
        for s in range(moc[line]):
            f, l = dau[line, s], cuoi[line, s]
            for k in range(f + 1, l):
                idx = line * stride + k
                for a, b in zip(rows, ref):
                    assert a[idx] == b[idx], (line, k)
                n += 1
    assert n > 500


@pytest.mark.parametrize("mode", [QB_FIXED, QB_REFERENCE])
def test_calculate_qb_and_bed_load(mode):
    case, o, g, t = stepped()
    sediment_ready(case, o, g)
    o.kernel("calculate_qb", mode)
    g.kernel("Calculate_Qb", mode)
    assert same(o, g, ("Qbx", "Qby")) == []
    for ketdinh in (0, 1):
        o.kernel("bed_load", ketdinh, 3, case.N - 1)
        g.kernel("BedLoad", ketdinh, 3, case.N - 1)
        assert same(o, g, ("dH",)) == []
