"""Properties the mathematics fixes for the CPU checker's full step, so a
dropped term, a wrong sign or a transposed index in the restated kernels
shows up without the reference running:

* lake at rest: still water, no forcing -> every field stays exactly zero;
* closed basin: with walls all round and no wet/dry change, the continuity
  fluxes telescope, so the water volume is conserved to round-off;
* discharge boundary: the per-cell velocities of a Q edge carry exactly Q
  (sum of (Q hi_i) hbar_i dX = Q, support_funcs.cu:451-511);
* sediment rows: every concentration row satisfies the diffusion row-sum
  identity AA+BB+CC = 2/dT (BB+CC at a first row folded with BB += AA,
  AA+BB at a last row folded with BB += CC, BB alone when both fold), the
  identity all 448 rows of the
  reference's ABCD.txt / VBABCD.txt dumps satisfy.
"""
import os

import numpy as np
import pytest

import oracle
from hydist_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))


def closed_basin(N=40, M=30, depth=4.0, hump=0.05, seed=3):
    case = synth.channel_case(N=N, M=M, nsteps=10, name="basin", noise=0.0)
    a = case.arrays
    h = np.full((N, M), depth) + np.random.default_rng(seed).normal(0, 0.02, (N, M))
    h[:2, :] = 0; h[-2:, :] = 0; h[:, :2] = 0; h[:, -2:] = 0
    a["h"] = synth.remove_pinches(synth.pad_map(h))
    a["boundary_type"][:] = 0
    a["bienQ"][:] = 0
    for k in ("bc_up", "bc_down", "bc_left", "bc_right"):
        a[k][:] = 0
    case.kenhhepng = 0
    return case, hump


def test_lake_at_rest_stays_exactly_at_rest():
    case, _ = closed_basin()
    o = oracle.Oracle(case)
    o.init()
    t = o.run(0.0, 40, case.opts())
    for name in ("u", "v", "z", "t_u", "t_v", "t_z"):
        assert not np.any(o.get(name)), name
    assert t == 40 * case.dT


def test_closed_basin_conserves_volume():
    case, hump = closed_basin()
    N, M = case.N, case.M
    i = np.arange(N + 3)[:, None]
    j = np.arange(M + 3)[None, :]
    z0 = hump * np.exp(-((i - N / 2) ** 2 + (j - M / 3) ** 2) / 20.0)
    o = oracle.Oracle(case)
    o.init()
    kh = o.get("khouot")
    wet = kh == 0
    z0 = np.where(wet, z0, 0.0)
    o.set("z", z0)
    o.set("t_z", z0)
    o.find_limits()
    vol0 = z0[wet].sum()
    o.run(0.0, 60, case.opts())
    z = o.get("z")
    assert np.array_equal(o.get("khouot"), kh)           # no wet/dry event
    assert abs(z.max()) > 1e-4                            # the hump moved
    assert abs(z[wet].sum() - vol0) < 1e-11 * np.abs(z0).sum()


def test_discharge_edge_carries_Q():
    case = synth.config(1, scale=0.2)
    o = oracle.Oracle(case)
    o.init()
    op = case.opts()
    o.halfstep(0, case.dT * 0.5, op)
    ubt = o.get("ubt")
    h = o.get("h")
    Q = case.arrays["bc_left"][0]
    n, d, c = o.get("moci"), o.get("daui"), o.get("cuoii")
    js = np.arange(d[2, 0], c[2, 0] + 1)
    hbar = (h[1, js] + h[1, js - 1]) / 2.0
    flux = np.sum(ubt[js] * hbar * case.params.dX)
    assert flux == pytest.approx(Q, rel=1e-12)


def _row_identity_ok(A, B, C, H, tol=1e-12):
    """A+B+C = H for a plain row; B+C = H after BB += AA (first row); A+B = H
    after BB += CC (last row); B = H after both (a one-row system)."""
    return any(abs(s - H) < tol * abs(H) for s in (A + B + C, B + C, A + B, B))


def test_reference_fs_dumps_satisfy_row_identity():
    for name in ("ABCD.txt", "VBABCD.txt"):
        rows = np.loadtxt(os.path.join(HERE, "golden", name))
        H = 1.0  # 2/dT of the TanChau case, dT = 2
        ok = [_row_identity_ok(r[2], r[3], r[4], H, 1e-12) for r in rows]
        assert all(ok), name


def test_checker_fs_rows_satisfy_row_identity():
    case = synth.config(3, scale=0.1)
    case.sediment_start = 0.0
    case.bed_change_start = 0.0
    o = oracle.Oracle(case)
    o.init()
    op = case.opts()
    t = o.run(0.0, 3, op)
    o.halfstep(0, t + case.dT * 0.5, op)        # ends with the FSj sweep
    p = case.params
    N, M, SL = case.N, case.M, case.segment_limit
    L = max(N, M)
    size = (L + 3) * (2 * L + 4)
    AA = o.get("AA", like=np.zeros(size)); BB = o.get("BB", like=np.zeros(size)); CC = o.get("CC", like=np.zeros(size))
    moc, dau, cuoi = o.get("mocj"), o.get("dauj"), o.get("cuoij")
    nrows = 0
    for jj in range(3, M):
        for s in range(moc[jj]):
            f, l = dau[jj, s], cuoi[jj, s]
            for i in range(f + 1, l):
                k = jj * (N + 3) + i
                assert _row_identity_ok(AA[k], BB[k], CC[k], p.HaiChiadT), (jj, i)
                nrows += 1
    assert nrows > 100
