"""Physics the whole restated step must reproduce, independent of how any
kernel is written: the fundamental seiche of a closed frictionless basin and
the steady flow of a straight channel.  They pin the half-steps' momentum and
continuity rows together (gravity, friction and flux terms, signs and the
half-step time levels, Engine.py:144-366), beyond the per-kernel closed forms.
"""
import math

import numpy as np
import pytest

import oracle
from oracle import PLANTS, planted
from hydist_amd import synth
from hydist_amd.case import make_case_header


def closed_basin(N, M, along_i, dx=100.0, dt=10.0, depth=10.0, amp=0.01):
    """Frictionless, inviscid, unsmoothed basin with walls (zero node depth)
    on all four sides, its water surface in the first mode along i or j."""
    case = synth.channel_case(N=N, M=M, dx=dx, dt=dt, nsteps=10, noise=0.0)
    case.header = make_case_header(dx, dx, dt, 20, heso=1.0, H_TINH=0.02)
    hn = np.full((N, M), depth)
    hn[:2, :] = 0; hn[-2:, :] = 0; hn[:, :2] = 0; hn[:, -2:] = 0
    a = case.arrays
    a["h"] = synth.pad_map(hn)
    a["hsnham"][:] = 0.0
    a["VISCOINDX"][:] = 0.0
    a["boundary_type"][:] = 0
    a["bienQ"][:] = 0
    case.kenhhepng = 0
    o = oracle.Oracle(case)
    o.init()
    wet = o.get("khouot") == 0
    idx = np.indices(wet.shape)[0 if along_i else 1]
    lo, hi = int(idx[wet].min()), int(idx[wet].max())
    ncell = hi - lo + 1
    z0 = np.where(wet, amp * np.cos(math.pi * (idx - lo + 0.5) / ncell), 0.0)
    o.set("z", z0)
    o.set("t_z", z0)
    o.kernel("find_limits")
    o.kernel("htuongdoi")
    return case, o, lo, ncell


def seiche_period(along_i, N=100, M=9):
    Nn, Mm = (N, M) if along_i else (M, N)
    case, o, lo, ncell = closed_basin(Nn, Mm, along_i)
    p = case.params
    # long-wave speed of the section: sqrt(g A / B), A / B the mean depth of
    # the faces across the basin (the cells beside a wall are half as deep)
    mid = ncell // 2 + lo
    face = o.get("Htdu")[mid, :] if along_i else o.get("Htdv")[:, mid]
    wet_faces = face[face > 0]
    H = wet_faces.mean()
    L = ncell * (p.dX if along_i else p.dY)
    T = 2 * L / math.sqrt(p.g * H)
    cell = (lo, Mm // 2) if along_i else (Nn // 2, lo)
    op = case.opts()
    t, ts, zs = 0.0, [], []
    for _ in range(int(2.3 * T / p.dT)):
        t = o.step(t, op)
        ts.append(t)
        zs.append(o.get("z")[cell])
    ts, zs = np.array(ts), np.array(zs)
    k = np.nonzero(zs[:-1] * zs[1:] < 0)[0]
    cross = ts[k] - zs[k] * (ts[k + 1] - ts[k]) / (zs[k + 1] - zs[k])
    return 2 * np.mean(np.diff(cross)), T, len(cross)


@pytest.mark.parametrize("along_i", [True, False])
def test_seiche_period_of_closed_basin(along_i):
    """T = 2 L / sqrt(g H) for the first mode of a closed basin (L the length
    between the walls, H the section's mean depth) to 1 %."""
    period, T, n = seiche_period(along_i)
    assert n >= 4
    assert period == pytest.approx(T, rel=0.01)


def test_seiche_pin_catches():
    with planted(PLANTS["F4_FACTOR"]):
        try:
            period, T, n = seiche_period(True)
        except Exception:
            return                      # the planted mistake broke the run outright
    assert not (n >= 4 and period == pytest.approx(T, rel=0.01))


def steady_channel(Q=60.0, depth=3.0, n_manning=0.025, steps=7200):
    """Straight channel (150 x 20 cells, dx = 10 m, flat bed, no smoothing):
    constant discharge Q at the upstream edge, level 0 downstream, run to a
    steady state."""
    case = synth.channel_case(N=150, M=20, dx=10.0, dt=2.0, nsteps=steps, noise=0.0, slope=0.0,
                              depth=depth, Q=Q, n_manning=n_manning)
    case.header = make_case_header(10.0, 10.0, 2.0, 20, H_TINH=0.02, heso=1.0)
    o = oracle.Oracle(case)
    o.init()
    o.run(0.0, steps, case.opts())
    return case, o


@pytest.fixture(scope="module")
def channel():
    return steady_channel()


def section_discharges(case, o):
    u, Htdu, kh = o.get("u"), o.get("Htdu"), o.get("khouot")
    N, M, dY = case.N, case.M, case.params.dY
    q = [np.sum((Htdu[i, 2:M + 1] * u[i, 2:M + 1])[kh[i, 2:M + 1] == 0]) * dY for i in range(3, N - 1)]
    return np.array(q)


def test_steady_channel_carries_the_inflow_through_every_section(channel):
    """At steady state continuity makes every cross-section carry the flux
    of the inflow face (its velocities times the face depths); that flux is
    Q to within the upstream level over the depth.  (The last face, beside
    the open water-level edge, takes its velocity from the edge's own rule.)"""
    case, o = channel
    q = section_discharges(case, o)
    ubt, Htdu = o.get("ubt"), o.get("Htdu")
    M = case.M
    inflow = np.sum(Htdu[1, 2:M + 1] * ubt[2:M + 1]) * case.params.dY
    assert np.all(np.abs(q / inflow - 1) < 1e-5)
    assert inflow == pytest.approx(60.0, rel=0.01)


def friction_balance(case, o):
    """Surface slope along the centreline against the friction slope of the
    reference's Manning-type law, S_f = n^2 |V| u / (hbar^(2 mu) H), from the
    roughness map and the depths (hbar the face's mean node depth, H the
    face's total depth), over the middle half of the channel."""
    p = case.params
    N, M = case.N, case.M
    u, z, Htdu, h = o.get("u"), o.get("z"), o.get("Htdu"), o.get("h")
    n = case.arrays["hsnham"]
    jc = M // 2
    rows = np.arange(N // 4, 3 * N // 4)
    zc = (z[rows, jc] + z[rows + 1, jc]) / 2          # level at the u-faces
    slope = np.polyfit(rows * p.dX, zc, 1)[0]
    hbar = (h[rows, jc - 1] + h[rows, jc]) / 2
    nbar = (n[rows, jc - 1] + n[rows, jc]) / 2
    uu = u[rows, jc]
    Sf = np.mean(nbar ** 2 * np.abs(uu) * uu / (hbar ** (2 * p.mu_mn) * Htdu[rows, jc]))
    return -slope, Sf


def test_steady_channel_surface_slope_balances_friction(channel):
    case, o = channel
    S, Sf = friction_balance(case, o)
    assert Sf > 1e-6
    assert S == pytest.approx(Sf, rel=0.02)


def test_friction_pin_catches():
    with planted(PLANTS["UZ_FRICTION"]):
        case, o = steady_channel()
    S, Sf = friction_balance(case, o)
    assert S != pytest.approx(Sf, rel=0.02)
