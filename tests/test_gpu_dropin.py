"""The reference's Python driver surface on the GPU path, against the CPU
checker: rep.gpu_init + Pointers + Engine.hydraulic_Calculation (Engine.py:
49-469, rep.py:58-142) with its hourly and debug extractions, the pick-up
writer and restart (rep.py:43-55), the Argument_Pointers / Array_Pointers
mirrors read back from the device (constant.cuh:6-29), and the narrow-channel
index shifts of kenhhepd / kenhhepng (Engine.py:152-158, 273-279)."""
import ctypes

import numpy as np
import pytest

import oracle
from hydist_amd import synth
from hydist_amd._lib import Solver, load
from hydist_amd.engine import hydraulic_Calculation
from hydist_amd.rep import gpu_init, save_pick_up

from _parity import assert_same_state

pytestmark = pytest.mark.gpu


def channel():
    return synth.channel_case(N=40, M=12, nsteps=60, name="channel_40x12")


def checker(case):
    o = oracle.Oracle(case)
    o.init()
    return o


def test_engine_interval_outputs_match_checker():
    """hydraulic_Calculation returns u, v, z every `interval` seconds."""
    case = channel()
    pointers, ctx, argp, arrp, supmod = gpu_init(0, False, None, case)
    Tmax = 40 * case.dT
    ul, vl, zl, fl = hydraulic_Calculation(Tmax, pointers, argp, arrp, supmod, ctx, kenhhepng=case.kenhhepng,
                                           debug=False, interval=10, save_grid=False, boundary_mode="series")
    o = checker(case)
    op = case.opts()
    t = 0.0
    assert len(ul) == len(vl) == len(zl) == 8
    for k in range(8):
        t = o.run(t, 5, op)
        for got, name in ((ul[k], "u"), (vl[k], "v"), (zl[k], "z")):
            assert np.array_equal(got, o.get(name)), (k, name)
    ctx.detach()


def test_engine_debug_records_both_half_steps():
    """With debug the reference appends t_u, t_v, t_z after each half-step
    (Engine.py:245-257, 369-381): two entries per step."""
    case = channel()
    pointers, ctx, argp, arrp, supmod = gpu_init(0, False, None, case)
    ul, vl, zl, fl = hydraulic_Calculation(3 * case.dT, pointers, argp, arrp, supmod, ctx,
                                           kenhhepng=case.kenhhepng, debug=True, interval=0, save_grid=False,
                                           boundary_mode="series")
    assert len(ul) == len(vl) == len(zl) == 6
    o = checker(case)
    op = case.opts()
    t = 0.0
    for k in range(6):
        t += case.dT * 0.5
        o.halfstep(k % 2, t, op)
        for got, name in ((ul[k], "t_u"), (vl[k], "t_v"), (zl[k], "t_z")):
            assert np.array_equal(got, o.get(name)), (k, name)
    ctx.detach()


def test_pick_up_round_trip_and_restart(tmp_path):
    """save_pick_up writes the files load_intial_condition reads; a restart
    from them loads the same u, v, z, khouot, FS and then steps exactly as
    the checker restarted from the same files."""
    case = synth.config(3, scale=0.1)
    case.sediment_start = 0.0
    pointers, ctx, argp, arrp, supmod = gpu_init(0, False, None, case)
    op = case.opts()
    t = supmod.run(0.0, 30, op)
    save_pick_up(str(tmp_path), pointers, t)
    N, M = case.N, case.M
    saved = {k: supmod.get(k) for k in ("u", "v", "z", "khouot", "FS")}
    p2, ctx2, a2, r2, s2 = gpu_init(0, True, str(tmp_path), case)
    for k, a in saved.items():
        b = s2.get(k)
        assert np.array_equal(b[:N + 2, :M + 2], a[:N + 2, :M + 2]), k
        if k == "khouot":
            assert (b[N + 2, :] == 2).all() and (b[:, M + 2] == 2).all()
        else:
            assert np.array_equal(b[N + 2, :M + 2], b[N + 1, :M + 2]), k     # edge padding
    o = checker(case)
    for k in ("u", "v", "z", "khouot", "FS"):
        o.set(k, s2.get(k))
    o.find_limits()
    o.kernel("htuongdoi")
    o.kernel("preprocess")
    t2 = s2.run(t, 10, op)
    assert o.run(t, 10, op) == t2
    assert_same_state(o, s2, "10 steps after the restart")
    ctx.detach()
    ctx2.detach()


ARG_FIELDS = ("bienQ", "daui", "dauj", "cuoii", "cuoij", "moci", "mocj", "khouot", "boundary_type", "h", "v", "u",
              "z", "t_u", "t_v", "t_z", "Htdu", "Htdv", "H_moi", "htaiz", "htaiz_bd", "ubt", "ubp", "vbt", "vbd",
              "hsnham", "VISCOIDX", "Kx1", "Ky1", "Tsyw", "Tsxw", "bc_up", "bc_down", "bc_left", "bc_right", "hi",
              "FS", "tFS", "CC_u", "CC_d", "CC_l", "CC_r", "VTH", "Kx", "Ky", "Fw", "Qbx", "Qby", "dH")


def hip():
    load()
    H = ctypes.CDLL("libamdhip64.so.7")
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return H


def test_argument_pointers_mirror_reads_back():
    """The 432-byte Argument_Pointers mirror: M, N, the four hmax values and
    49 device pointers in the reference's member order, each pointing at the
    named field's data; Array_Pointers holds 20 pointers."""
    case = synth.config(3, scale=0.1)
    g = Solver(case)
    g.init()
    g.run(0.0, 5, case.opts())
    arg, arr = g.struct_ptrs()
    H = hip()
    raw = np.zeros(432, np.uint8)
    assert H.hipMemcpy(raw.ctypes.data, arg, 432, 2) == 0
    M, N = np.frombuffer(raw[:8].tobytes(), np.int32)
    assert (M, N) == (case.M, case.N)
    assert list(np.frombuffer(raw[8:40].tobytes(), np.float64)) == g.hmax()
    ptrs = np.frombuffer(raw[40:].tobytes(), np.uint64)
    assert len(ptrs) == len(ARG_FIELDS) == 49
    for p, name in zip(ptrs, ARG_FIELDS):
        want = g.get(name)
        got = np.empty_like(want)
        assert H.hipMemcpy(got.ctypes.data, int(p), got.nbytes, 2) == 0
        assert np.array_equal(got, want, equal_nan=want.dtype.kind == "f"), name
    rawr = np.zeros(20 * 8, np.uint8)
    assert H.hipMemcpy(rawr.ctypes.data, arr, rawr.nbytes, 2) == 0
    rp = np.frombuffer(rawr.tobytes(), np.uint64)
    assert rp[11] != 0 and rp[19] != 0          # AA and SN
    g.close()


@pytest.mark.parametrize("kd,kn,bmode", [(1, 0, "series"), (0, 1, "series"), (1, 0, "analytic"),
                                         (0, 1, "analytic")])
def test_narrow_channel_modes(kd, kn, bmode):
    """kenhhepd = 1 starts the UZ lines at 3 and ends them at M - 1
    (Engine.py:154-156) and zeroes u in update_uvz; kenhhepng = 1 starts the
    VZ lines at 3 (Engine.py:277-279) and zeroes v; with the analytic tide
    boundary the reference forces for these modes (Engine.py:29-43)."""
    case = synth.channel_case(N=30, M=24, nsteps=40, name="channel_30x24")
    case.kenhhepd, case.kenhhepng = kd, kn
    case.boundary_mode = {"series": 1, "analytic": 2}[bmode]
    o, g = checker(case), Solver(case)
    g.init()
    op = case.opts()
    t = 0.0
    for chunk in (10, 10, 10):
        t_o = o.run(t, chunk, op)
        g.run(t, chunk, op)
        t = t_o
        assert_same_state(o, g, f"kenhhepd {kd} kenhhepng {kn} {bmode} at t = {t}")
    assert not np.any(o.get("u" if kd else "v"))
    g.close()
