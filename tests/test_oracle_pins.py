"""Closed forms that pin the CPU checker's restated kernels one by one.

Each test feeds a kernel of the checker inputs for which the mathematics fixes
the answer, independently of how the kernel is written:

* momentum assembly (UZ / VZ preindex, UVZSolver_multithread.cu:295-407) and
  the explicit velocities (vSolver / uSolver, :727-815) on quadratic fields,
  where the central, one-sided and second differences of the stencils are
  exact, so every p / q term equals its calculus value;
* the per-line elimination, assembly, Thomas solve and extraction
  (:409-723) against a dense solve of the unreduced (z, u) system the
  continuity and momentum rows state, for every end condition;
* the run margins (:970-1052) as linear extrapolation or wall zeros;
* Normalize (support_funcs.cu:682-729): the identity at heso = 1 and the
  neighbour mean otherwise;
* wet/dry (support_funcs.cu:75-186): hand-built one-line opening, cascade
  and drying scenarios;
* sediment closed forms: VTH, hesoK, the cohesive and non-cohesive sources,
  the bed-load flux and the bed change of a hand-computed patch
  (Sediment_Transport.cu:12-60, 341-568).

Every pin is shown to catch a plausible mistake: the same check runs with a
mistake planted in the checker (oracle.planted) and must fail.
"""
import math

import numpy as np
import pytest

import oracle
from oracle import PLANTS, planted
from hydist_amd import synth
from hydist_amd.case import make_case_header


def make(N=10, M=10, dx=10.0, dy=8.0, dt=2.0, depth=6.0, **over):
    """A checker instance on an N x M grid with every node at `depth` (all
    cells wet) and the case parameters `over`."""
    case = synth.channel_case(N=N, M=M, dx=dx, dt=dt, nsteps=10, noise=0.0)
    case.header = make_case_header(dx, dy, dt, 20, **over)
    case.arrays["h"] = np.full((N + 3, M + 3), depth)
    o = oracle.Oracle(case)
    return case, o


def full(case, value):
    return np.full((case.N + 3, case.M + 3), float(value))


def coords(case):
    p = case.params
    i = np.arange(case.N + 3, dtype=float)[:, None] * np.ones((1, case.M + 3))
    j = np.ones((case.N + 3, 1)) * np.arange(case.M + 3, dtype=float)[None, :]
    return i * p.dX, j * p.dY


def rel_err(a, b, scale):
    return abs(a - b) / scale


# ----------------------------------------------------------------- preindex

UZ_PARAMS = dict(CORIOLIS_FORCE=1.1e-4, heso=1.0)


def uz_setup(v0=0.4, dry=(), visc=0.5):
    """u(y) = 0.3 + 2e-3 y + 1e-5 y^2 (quadratic across the UZ line), v = v0,
    face depths Htdu linear in x, Htdv quadratic in y."""
    case, o = make(**UZ_PARAMS)
    x, y = coords(case)
    u = 0.3 + 2e-3 * y + 1e-5 * y ** 2
    Htdv = 3.0 + 0.01 * y + 1e-4 * y ** 2
    fields = dict(u=u, v=full(case, v0), z=0.05 + 1e-3 * x, Htdu=4.0 + 0.02 * x, Htdv=Htdv,
                  H_moi=full(case, 5.0), VISCOIDX=full(case, visc), Kx1=full(case, 3e-3),
                  Tsxw=full(case, 0.01))
    for (i, j) in dry:
        fields["H_moi"][i, j] = 0.01
    for k, a in fields.items():
        o.set(k, a)
    return case, o, fields


def uz_expected(case, f, i, j, branch):
    """The UZ momentum row -2u + dT (vbar du/dy - f vbar - nu d2u/dy2
    - (Wx - Tsx)/H) and the continuity row z - dT/(2 dY) [F(j) - F(j-1)],
    F = Htdv v the y-flux through the cell's two faces, with the analytic
    derivatives of the quadratic u(y).  branch 'none' drops the advection
    and viscous y terms (dry side with the wrong flow direction)."""
    p = case.params
    y = j * p.dY
    u = f["u"][i, j]
    vbar = f["v"][i, j]
    uy = 2e-3 + 2e-5 * y
    uyy = 2e-5
    nu = f["VISCOIDX"][i, j]
    H = f["Htdu"][i, j]
    adv = 0.0 if branch == "none" else vbar * uy
    vis = 0.0 if branch == "none" else nu * uyy
    f5 = -2 * u + p.dT * adv - p.dT * p.CORIOLIS_FORCE * vbar - p.dT * vis - p.dT * (0.0 - f["Tsxw"][i, j]) / H
    speed = math.hypot(u, vbar)
    f2 = -(2 + f["Kx1"][i, j] * p.dT * speed / H + 2 * p.dT * nu / p.dX ** 2)
    f1 = p.dT / (2 * p.dX) * u + nu * p.dT / p.dX ** 2
    f3 = nu * p.dT / p.dX ** 2 - p.dT / (2 * p.dX) * u
    a2 = -p.dT / (2 * p.dX) * f["Htdu"][i - 1, j]
    c2 = p.dT / (2 * p.dX) * H
    F = lambda yy: (3.0 + 0.01 * yy + 1e-4 * yy ** 2) * vbar
    d2 = f["z"][i, j] - p.dT / (2 * p.dY) * (F(y) - F(y - p.dY))
    return dict(f1=f1, f2=f2, f3=f3, f5=f5, a2=a2, c2=c2, d2=d2)


def uz_check(o, case, f, i, j, branch, tol=1e-12):
    o.kernel("uz_preindex", i, j)
    exp = uz_expected(case, f, i, j, branch)
    sw = case.N + 2
    bad = []
    for k, e in exp.items():
        got = o.scratch(k)[j * sw + i]
        scale = max(abs(e), 1e-3) if k != "f5" else 1e-3   # f5: scale of its dT terms
        if k == "f5":
            got, e = got + 2 * f["u"][i, j], e + 2 * f["u"][i, j]
        if rel_err(got, e, scale) > tol:
            bad.append((k, got, e))
    return bad


@pytest.mark.parametrize("v0,dry,branch", [
    (0.4, (), "central"),
    (-0.4, ((5, 4),), "central"),             # H_moi[j-1] dry, vbar < 0: forward one-sided
    (0.4, ((5, 4),), "none"),                 # H_moi[j-1] dry, vbar > 0: no y terms
    (0.4, ((5, 6),), "central"),              # H_moi[j+1] dry, [j-2] wet, vbar > 0: backward one-sided
    (-0.4, ((5, 6),), "none"),
    (0.4, ((5, 6), (5, 3)), "none"),          # H_moi[j+1] and [j-2] dry
])
def test_uz_preindex_on_quadratic_field(v0, dry, branch):
    case, o, f = uz_setup(v0=v0, dry=dry)
    assert uz_check(o, case, f, 5, 5, branch) == []


@pytest.mark.parametrize("plant", ["UZ_ADV_INDEX", "UZ_FRICTION", "UZ_D2_INDEX"])
def test_uz_preindex_pin_catches(plant):
    case, o, f = uz_setup()
    with planted(PLANTS[plant]):
        assert uz_check(o, case, f, 5, 5, "central") != []


def vz_setup(u0=0.4, dry=(), visc=0.5):
    """Mirror of uz_setup: v(x) quadratic along the VZ line, u = u0."""
    case, o = make(**UZ_PARAMS)
    x, y = coords(case)
    fields = dict(v=0.3 + 2e-3 * x + 1e-5 * x ** 2, u=full(case, u0), z=0.05 + 1e-3 * y,
                  Htdv=4.0 + 0.02 * y, Htdu=3.0 + 0.01 * x + 1e-4 * x ** 2,
                  H_moi=full(case, 5.0), VISCOIDX=full(case, visc), Ky1=full(case, 3e-3),
                  Tsyw=full(case, 0.01))
    for (i, j) in dry:
        fields["H_moi"][i, j] = 0.01
    for k, a in fields.items():
        o.set(k, a)
    return case, o, fields


def vz_expected(case, f, i, j, branch):
    """The VZ momentum row -2v + dT (ubar dv/dx + f ubar - nu d2v/dx2
    - (Wy - Tsy)/H) and continuity z - dT/(2 dX) [F(i) - F(i-1)], F = Htdu u."""
    p = case.params
    x = i * p.dX
    v = f["v"][i, j]
    ubar = f["u"][i, j]
    vx = 2e-3 + 2e-5 * x
    vxx = 2e-5
    nu = f["VISCOIDX"][i, j]
    H = f["Htdv"][i, j]
    adv = 0.0 if branch == "none" else ubar * vx
    vis = 0.0 if branch == "none" else nu * vxx
    f5 = -2 * v + p.dT * adv + p.dT * p.CORIOLIS_FORCE * ubar - p.dT * vis - p.dT * (0.0 - f["Tsyw"][i, j]) / H
    speed = math.hypot(v, ubar)
    f2 = -(2 + f["Ky1"][i, j] * p.dT * speed / H + 2 * p.dT * nu / p.dY ** 2)
    f1 = p.dT / (2 * p.dY) * v + nu * p.dT / p.dY ** 2
    f3 = nu * p.dT / p.dY ** 2 - p.dT / (2 * p.dY) * v
    a2 = -p.dT / (2 * p.dY) * f["Htdv"][i, j - 1]
    c2 = p.dT / (2 * p.dY) * H
    F = lambda xx: (3.0 + 0.01 * xx + 1e-4 * xx ** 2) * ubar
    d2 = f["z"][i, j] - p.dT / (2 * p.dX) * (F(x) - F(x - p.dX))
    return dict(f1=f1, f2=f2, f3=f3, f5=f5, a2=a2, c2=c2, d2=d2)


def vz_check(o, case, f, i, j, branch, tol=1e-12):
    o.kernel("vz_preindex", i, j)
    exp = vz_expected(case, f, i, j, branch)
    sw = case.M + 2
    bad = []
    for k, e in exp.items():
        got = o.scratch(k)[i * sw + j]
        scale = max(abs(e), 1e-3) if k != "f5" else 1e-3
        if k == "f5":
            got, e = got + 2 * f["v"][i, j], e + 2 * f["v"][i, j]
        if rel_err(got, e, scale) > tol:
            bad.append((k, got, e))
    return bad


@pytest.mark.parametrize("u0,dry,branch", [
    (0.4, (), "central"),
    (-0.4, ((4, 5),), "central"),
    (0.4, ((4, 5),), "none"),
    (0.4, ((6, 5),), "central"),
    (-0.4, ((6, 5),), "none"),
])
def test_vz_preindex_on_quadratic_field(u0, dry, branch):
    case, o, f = vz_setup(u0=u0, dry=dry)
    assert vz_check(o, case, f, 5, 5, branch) == []


def test_vz_preindex_pin_catches():
    case, o, f = vz_setup()
    with planted(PLANTS["VZ_CORIOLIS_SIGN"]):
        assert vz_check(o, case, f, 5, 5, "central") != []


# ---------------------------------------------------------- explicit u / v

def usolver_setup(v0=0.35, dry=()):
    case, o = make(**UZ_PARAMS)
    x, y = coords(case)
    u = 0.3 + 3e-3 * x - 2e-5 * x ** 2 + 2e-3 * y + 1e-5 * y ** 2
    z = 0.05 + 1e-4 * x + 3e-7 * x ** 2
    fields = dict(u=u, v=full(case, v0), t_v=full(case, v0), z=z, Htdu=full(case, 4.0),
                  H_moi=full(case, 5.0), VISCOIDX=full(case, 0.5), Kx1=full(case, 3e-3),
                  Tsxw=full(case, 0.01))
    for (i, j) in dry:
        fields["H_moi"][i, j] = 0.01
    for k, a in fields.items():
        o.set(k, a)
    return case, o, fields


def usolver_expected(case, f, i, j, ybranch):
    """uSolver as the semi-implicit x-momentum balance: t_u (2/dT + du/dx +
    K |V| / H) = 2u/dT - vbar* du/dy + f vbar* - g dz/dx|face + nu (u_xx +
    u_yy) + (Wx - Tsx)/H, the surface slope taken at the u-face between
    cells i and i+1; derivatives of the quadratic fields analytic."""
    p = case.params
    x, y = i * p.dX, j * p.dY
    u = f["u"][i, j]
    ux = 3e-3 - 4e-5 * x
    uxx = -4e-5
    uy = 2e-3 + 2e-5 * y
    uyy = 2e-5
    zx_face = 1e-4 + 6e-7 * (x + p.dX / 2)
    vb = f["v"][i, j]
    nu, H = f["VISCOIDX"][i, j], f["Htdu"][i, j]
    adv_y, vis_y = (0.0, 0.0) if ybranch == "none" else (vb * uy, uyy)
    rhs = 2 * u / p.dT - adv_y + p.CORIOLIS_FORCE * vb - p.g * zx_face + nu * (uxx + vis_y) + (0.0 - f["Tsxw"][i, j]) / H
    lhs = 2 / p.dT + ux + f["Kx1"][i, j] * math.hypot(vb, u) / H
    return rhs / lhs


@pytest.mark.parametrize("v0,dry,branch", [
    (0.35, (), "central"), (-0.35, ((5, 4),), "central"), (0.35, ((5, 4),), "none"),
    (0.35, ((5, 6),), "central"), (-0.35, ((5, 6),), "none")])
def test_u_solver_on_quadratic_field(v0, dry, branch):
    case, o, f = usolver_setup(v0, dry)
    o.kernel("u_solver", 5, 5)
    got = o.get("t_u")[5, 5]
    assert got == pytest.approx(usolver_expected(case, f, 5, 5, branch), rel=1e-12)


def test_u_solver_pin_catches():
    case, o, f = usolver_setup()
    with planted(PLANTS["US_ZGRAD_INDEX"]):
        o.kernel("u_solver", 5, 5)
    assert o.get("t_u")[5, 5] != pytest.approx(usolver_expected(case, f, 5, 5, "central"), rel=1e-9)


def vsolver_setup(u0=0.35, dry=()):
    case, o = make(**UZ_PARAMS)
    x, y = coords(case)
    v = 0.3 + 3e-3 * y - 2e-5 * y ** 2 + 2e-3 * x + 1e-5 * x ** 2
    z = 0.05 + 1e-4 * y + 3e-7 * y ** 2
    fields = dict(v=v, u=full(case, u0), t_u=full(case, u0), z=z, Htdv=full(case, 4.0),
                  H_moi=full(case, 5.0), VISCOIDX=full(case, 0.5), Ky1=full(case, 3e-3),
                  Tsyw=full(case, 0.01))
    for (i, j) in dry:
        fields["H_moi"][i, j] = 0.01
    for k, a in fields.items():
        o.set(k, a)
    return case, o, fields


def vsolver_expected(case, f, i, j, xbranch):
    """vSolver: t_v (2/dT + dv/dy + K |V| / H) = 2v/dT - ubar* dv/dx - f ubar*
    - g dz/dy|face + nu (v_xx + v_yy) + (Wy - Tsy)/H."""
    p = case.params
    x, y = i * p.dX, j * p.dY
    v = f["v"][i, j]
    vy = 3e-3 - 4e-5 * y
    vyy = -4e-5
    vx = 2e-3 + 2e-5 * x
    vxx = 2e-5
    zy_face = 1e-4 + 6e-7 * (y + p.dY / 2)
    ub = f["u"][i, j]
    nu, H = f["VISCOIDX"][i, j], f["Htdv"][i, j]
    adv_x, vis_x = (0.0, 0.0) if xbranch == "none" else (ub * vx, vxx)
    rhs = 2 * v / p.dT - adv_x - p.CORIOLIS_FORCE * ub - p.g * zy_face + nu * (vis_x + vyy) + (0.0 - f["Tsyw"][i, j]) / H
    lhs = 2 / p.dT + vy + f["Ky1"][i, j] * math.hypot(ub, v) / H
    return rhs / lhs


@pytest.mark.parametrize("u0,dry,branch", [
    (0.35, (), "central"), (-0.35, ((4, 5),), "central"), (0.35, ((4, 5),), "none"),
    (0.35, ((6, 5),), "central"), (-0.35, ((6, 5),), "none")])
def test_v_solver_on_quadratic_field(u0, dry, branch):
    case, o, f = vsolver_setup(u0, dry)
    o.kernel("v_solver", 5, 5)
    assert o.get("t_v")[5, 5] == pytest.approx(vsolver_expected(case, f, 5, 5, branch), rel=1e-12)


def test_v_solver_pin_catches():
    case, o, f = vsolver_setup()
    with planted(PLANTS["VS_VISC_DROP"]):
        o.kernel("v_solver", 5, 5)
    assert o.get("t_v")[5, 5] != pytest.approx(vsolver_expected(case, f, 5, 5, "central"), rel=1e-9)


# ------------------------------------------- the unreduced (z, u) line system

def dense_line_solution(f, first, last, start, end, fixed):
    """Solve the unreduced half-step system of one run directly: continuity
    a2_i u_{i-1} + z_i + c2_i u_i = d2_i at every cell i of [first, last] and
    momentum f1 u_{i-1} + f2 u_i + f3 u_{i+1} + f4 (z_i - z_{i+1}) = f5 at
    every face i of [first, last-1], with the end conditions: a wall face
    carries u = 0, a discharge face its given velocity, an open water-level
    end its given z.  Returns {('z'|'u', index): value}."""
    known = dict(fixed)
    unk = []
    for i in range(first - 1, last + 1):
        if ("u", i) not in known:
            unk.append(("u", i))
    for i in range(first, last + 1):
        if ("z", i) not in known:
            unk.append(("z", i))
    col = {k: n for n, k in enumerate(unk)}
    rows, rhs = [], []

    def row(terms, r):
        a = np.zeros(len(unk))
        for coef, key in terms:
            if key in col:
                a[col[key]] += coef
            else:
                r -= coef * known[key]
        rows.append(a)
        rhs.append(r)

    for i in range(first, last + 1):
        row([(f["a2"][i], ("u", i - 1)), (1.0, ("z", i)), (f["c2"][i], ("u", i))], f["d2"][i])
    for i in range(first, last):
        row([(f["f1"][i], ("u", i - 1)), (f["f2"][i], ("u", i)), (f["f3"][i], ("u", i + 1)),
             (f["f4"], ("z", i)), (-f["f4"], ("z", i + 1))], f["f5"][i])
    A = np.array(rows)
    assert A.shape[0] == A.shape[1], (start, end, A.shape)
    sol = np.linalg.solve(A, np.array(rhs))
    out = dict(known)
    out.update({k: sol[n] for k, n in col.items()})
    return out


def line_case(kind_start, kind_end, N=14, M=7):
    """A UZ column j = 4 whose one wet run spans rows 2..N, with a smooth
    non-trivial state; the run's ends are walls ('wall': zero node depths
    on the edge row) or open edges ('z' water level, 'q' discharge)."""
    case, o = make(N=N, M=M, dx=12.0, dy=9.0, dt=3.0, heso=1.0, CORIOLIS_FORCE=1e-4)
    W = M + 3
    h = np.full((N + 3, W), 7.0)
    h += 0.3 * np.sin(np.arange(N + 3))[:, None]
    if kind_start == "wall":
        h[1, :] = 0.0
    if kind_end == "wall":
        h[N, :] = 0.0
    o.set("h", h)
    bq = np.zeros(4, np.int32)
    bq[2] = 1 if kind_start == "q" else 0
    bq[3] = 1 if kind_end == "q" else 0
    o.set("bienQ", bq)
    o.init()
    x, y = coords(case)
    rng = np.random.default_rng(7)
    o.set("u", 0.2 + 0.05 * np.sin(x / 40.0) + 0.01 * rng.random(x.shape))
    o.set("v", 0.1 + 0.03 * np.cos(y / 30.0) + 0.01 * rng.random(x.shape))
    z = 0.1 + 0.02 * np.sin(x / 50.0 + y / 70.0)
    o.set("z", z)
    tz = z.copy()
    tz[2, :] = 0.17        # water levels of the open ends (TZ_l, TZ_r)
    tz[N, :] = -0.06
    o.set("t_z", tz)
    o.set("ubt", np.full(M + 2, 0.45))
    o.set("ubp", np.full(M + 2, 0.38))
    o.set("VISCOINDX", full(case, 0.3))
    o.kernel("htuongdoi")
    return case, o


@pytest.mark.parametrize("ends", [("wall", "wall"), ("wall", "z"), ("z", "wall"), ("z", "z"),
                                  ("q", "wall"), ("q", "z"), ("wall", "q")])
def test_uz_line_solve_equals_unreduced_system(ends):
    """Elimination (abcd), interleaved assembly with its end rows, Thomas and
    extraction of one UZ run against the dense solution of the unreduced
    system.  A run starting at a discharge edge and ending at an open edge is
    solved with its end face carrying the end edge's velocity array and the
    last face velocity then taken from continuity (the reference tests the
    start edge's discharge flag at the end row); an open end keeps its given
    level."""
    case, o = line_case(*ends)
    N, M, j = case.N, case.M, 4
    first, last = o.get("dauj")[j, 0], o.get("cuoij")[j, 0]
    assert o.get("mocj")[j] == 1 and first == 2 and last == N
    o.kernel("uz_solve", j, j)
    sw = N + 2
    f = {k: o.scratch(k)[j * sw: j * sw + N + 2] for k in ("f1", "f2", "f3", "f5", "a2", "c2", "d2")}
    p = case.params
    f["f4"] = 2 * p.g * p.dTchia2dX
    tz0 = o.get("t_z")
    start, end = ends
    fixed = {}
    if start == "wall":
        fixed[("u", first - 1)] = 0.0
    elif start == "q":
        fixed[("u", first - 1)] = 0.45
    else:
        fixed[("z", first)] = 0.17
    d22 = start == "q" and end == "z"
    if end == "wall":
        fixed[("u", last)] = 0.0
    elif end == "q" or d22:
        fixed[("u", last)] = 0.38
    else:
        fixed[("z", last)] = -0.06
    sol = dense_line_solution(f, first, last, start, end, fixed)
    t_z, t_u = o.get("t_z"), o.get("t_u")
    for i in range(first, last + 1):
        want = sol[("z", i)]
        if d22 and i == last:
            want = tz0[last, j]        # the open end keeps its level
        assert t_z[i, j] == pytest.approx(want, rel=1e-10, abs=1e-12), ("z", i)
    for i in range(first - 1, last + 1):
        want = sol[("u", i)]
        if d22 and i == last:          # from continuity with the kept level
            want = (f["d2"][last] - f["a2"][last] * t_u[last - 1, j] - tz0[last, j]) / f["c2"][last]
        if end == "q" and i == last and not d22:
            # the end row uses the start edge's discharge flag: a wall / open
            # start with a discharge end solves the end as an open edge
            continue
        assert t_u[i, j] == pytest.approx(want, rel=1e-10, abs=1e-12), ("u", i)


def test_line_solve_pin_catches():
    case, o = line_case("wall", "wall")
    with planted(PLANTS["ABCD_C1_SIGN"]):
        o.kernel("uz_solve", 4, 4)
    N, j = case.N, 4
    sw = N + 2
    f = {k: o.scratch(k)[j * sw: j * sw + N + 2] for k in ("f1", "f2", "f3", "f5", "a2", "c2", "d2")}
    f["f4"] = 2 * case.params.g * case.params.dTchia2dX
    sol = dense_line_solution(f, 2, N, "wall", "wall", {("u", 1): 0.0, ("u", N): 0.0})
    t_z = o.get("t_z")
    assert not all(t_z[i, j] == pytest.approx(sol[("z", i)], rel=1e-8) for i in range(2, N + 1))


# ------------------------------------------------------------------ margins

def test_margins_open_edges_extrapolate_and_walls_are_zero():
    """update_margin_elem_U: a run's outer faces are zero at a wall and the
    linear extrapolation of the run's first / last two faces at an open edge
    (exact for a linear profile)."""
    case, o = make(N=12, M=8)
    N, M = case.N, case.M
    h = np.full((N + 3, M + 3), 6.0)
    h[N, 5] = 0.0; h[N, 4] = 0.0          # column 5: wall at the last edge
    o.set("h", h)
    x, _ = coords(case)
    o.set("t_u", 0.2 + 0.01 * x)
    dauj = np.zeros((M + 2, 20), np.int32); cuoij = np.zeros((M + 2, 20), np.int32)
    mocj = np.zeros(M + 2, np.int32)
    for j in (3, 5):
        dauj[j, 0], cuoij[j, 0], mocj[j] = 2, N, 1
    dauj[6, 0], cuoij[6, 0], mocj[6] = 4, 9, 1    # inner run: walls at both ends
    for k, a in (("dauj", dauj), ("cuoij", cuoij), ("mocj", mocj)):
        o.set(k, a)

    def check():
        o.kernel("margins_u", 2, M)
        t = o.get("t_u")
        lin = lambda i: 0.2 + 0.01 * i * case.params.dX
        ok = abs(t[1, 3] - lin(1)) < 1e-14 and abs(t[N, 3] - lin(N)) < 1e-14
        ok &= abs(t[1, 5] - lin(1)) < 1e-14 and t[N, 5] == 0.0
        ok &= t[3, 6] == 0.0 and t[9, 6] == 0.0
        return ok

    assert check()
    o.set("t_u", 0.2 + 0.01 * x)
    with planted(PLANTS["MARGIN_EXTRAP"]):
        assert not check()


# ---------------------------------------------------------------- Normalize

def normalize_setup(heso, seed=5):
    case, o = make(N=9, M=8, heso=heso)
    rng = np.random.default_rng(seed)
    val = rng.normal(size=(case.N + 3, case.M + 3))
    val[rng.random(val.shape) < 0.2] = 0.0
    kh = rng.choice([0, 0, 0, 1, 2], size=val.shape).astype(np.int32)
    o.set("t_u", val)
    o.set("t_v", val.T.copy()[: case.N + 3, : case.M + 3] if case.N == case.M else val)
    o.set("khouot", kh)
    return case, o, val, kh


def test_normalize_identity_at_heso_one():
    case, o, val, _ = normalize_setup(1.0)
    for isU in (1, 0):
        o.kernel("normalize", isU)
    assert np.array_equal(o.get("t_u"), val) and np.array_equal(o.get("t_v"), val)


def neighbour_mean(val, kh, i, j, isU):
    """Mean of the two neighbours along the velocity's own direction (rows
    for u, columns for v) and of those across it that are wet."""
    along = [(i - 1, j), (i + 1, j)] if isU else [(i, j - 1), (i, j + 1)]
    across = [(i, j - 1), (i, j + 1)] if isU else [(i - 1, j), (i + 1, j)]
    vals = [val[a] for a in along] + [val[a] for a in across if kh[a] == 0]
    return sum(vals) / len(vals)


@pytest.mark.parametrize("isU", [1, 0])
def test_normalize_is_weighted_neighbour_mean(isU):
    def check():
        case, o, val, kh = normalize_setup(0.8)
        o.kernel("normalize", isU)
        got = o.get("t_u" if isU else "t_v")
        N, M = case.N, case.M
        for i in range(case.N + 3):
            for j in range(case.M + 3):
                inside = 2 <= i <= N and 2 <= j <= M
                if not inside or val[i, j] == 0:
                    want = val[i, j]
                else:
                    want = 0.8 * val[i, j] + 0.2 * neighbour_mean(val, kh, i, j, isU)
                if abs(got[i, j] - want) > 1e-14 * max(1.0, abs(want)):
                    return False
        return True

    assert check()
    with planted(PLANTS["NORM_COUNT"]):
        assert not check()


# ------------------------------------------------------------------ wet/dry

def wetdry_line(kh_row, tz_row, z_row, hm_row=None, fs_row=None, htaiz=1.0):
    """One interior row i = 3 of a 5 x 12 grid; every other row is land."""
    case, o = make(N=5, M=12)
    W = case.M + 3
    kh = np.full((case.N + 3, W), 2, np.int32)
    kh[3, 1:1 + len(kh_row)] = kh_row
    tz = np.zeros((case.N + 3, W)); tz[3, 1:1 + len(tz_row)] = tz_row
    z = np.zeros((case.N + 3, W)); z[3, 1:1 + len(z_row)] = z_row
    hm = np.full((case.N + 3, W), 2.0)
    if hm_row is not None:
        hm[3, 1:1 + len(hm_row)] = hm_row
    fs = np.zeros((case.N + 3, W))
    if fs_row is not None:
        fs[3, 1:1 + len(fs_row)] = fs_row
    for k, a in (("khouot", kh), ("t_z", tz), ("z", z), ("H_moi", hm), ("FS", fs),
                 ("htaiz", np.full((case.N + 3, W), htaiz)), ("t_u", np.ones((case.N + 3, W))),
                 ("t_v", np.ones((case.N + 3, W)))):
        o.set(k, a)
    return case, o


def row3(o, name):
    return list(o.get(name)[3, 1:13])


def scenario_open():
    # columns 1..12; wet 2..4 rising (t_z 0.5 > z 0.3); 5 dried (1) with a
    # higher old level, 6 land (2), 7..12 land
    case, o = wetdry_line([2, 0, 0, 0, 1, 2, 2, 2, 2, 2, 2, 2],
                          [0, .5, .5, .5, .1, 0, 0, 0, 0, 0, 0, 0],
                          [0, .3, .3, .3, .6, .2, 0, 0, 0, 0, 0, 0],
                          fs_row=[0, 7, 7, 9, 0, 0, 0, 0, 0, 0, 0, 0])
    o.kernel("reset_states_h")
    return (row3(o, "khouot"), row3(o, "t_z"), row3(o, "FS"), row3(o, "H_moi"))


def test_wetting_opens_a_dried_neighbour():
    kh, tz, fs, hm = scenario_open()
    # cell 5 opens from cell 4 (t_z, FS copied; H_moi = htaiz + t_z); its own
    # old level 0.6 is above 0.5 so it passes nothing on; land (2) never opens
    assert kh == [2, 0, 0, 0, 0, 2, 2, 2, 2, 2, 2, 2]
    assert tz[4] == 0.5 and fs[4] == 9 and hm[4] == 1.5


def test_wetting_cascades_along_the_scan():
    case, o = wetdry_line([2, 0, 0, 1, 1, 1, 2, 2, 2, 2, 2, 2],
                          [0, .5, .5, 0, 0, 0, 0, 0, 0, 0, 0, 0],
                          [.3] * 12, fs_row=[0, 4, 5, 0, 0, 0, 0, 0, 0, 0, 0, 0])
    o.kernel("reset_states_h")
    # 3 opens 4, which (now rising above its old level) opens 5, then 6;
    # 7 is land
    assert row3(o, "khouot") == [2, 0, 0, 0, 0, 0, 2, 2, 2, 2, 2, 2]
    assert row3(o, "t_z")[3:6] == [.5, .5, .5] and row3(o, "FS")[3:6] == [5, 5, 5]


def test_wetting_opens_backwards_by_one_cell_only():
    case, o = wetdry_line([2, 1, 1, 0, 0, 2, 2, 2, 2, 2, 2, 2],
                          [0, 0, 0, .5, .5, 0, 0, 0, 0, 0, 0, 0],
                          [.3] * 12)
    o.kernel("reset_states_h")
    # column 4 (rising) opens 3 behind the scan; 2 was visited before and
    # stays dried
    assert row3(o, "khouot") == [2, 1, 0, 0, 0, 2, 2, 2, 2, 2, 2, 2]


def test_drying_closes_a_shallow_cell():
    case, o = wetdry_line([2, 0, 0, 0, 0, 2, 2, 2, 2, 2, 2, 2],
                          [0, .1] * 6, [0, .2] * 6,
                          hm_row=[2, 2, .01, 2, 2, 2, 2, 2, 2, 2, 2, 2],
                          fs_row=[1] * 12)
    o.kernel("reset_states_h")
    kh, t_u, t_v = o.get("khouot"), o.get("t_u"), o.get("t_v")
    assert kh[3, 3] == 1 and o.get("FS")[3, 3] == 0
    # the four faces of cell (3, 3): u at rows 2 and 3, v at columns 2 and 3
    assert t_u[2, 3] == 0 and t_u[3, 3] == 0 and t_v[3, 2] == 0 and t_v[3, 3] == 0
    assert t_u[4, 3] == 1 and t_v[3, 4] == 1 and kh[3, 2] == 0 and kh[3, 4] == 0


@pytest.mark.parametrize("plant,test", [
    ("WET_OPENS_LAND", test_wetting_opens_a_dried_neighbour),
    ("WET_NO_CASCADE", test_wetting_cascades_along_the_scan),
    ("DRY_FACE", test_drying_closes_a_shallow_cell)])
def test_wet_dry_pins_catch(plant, test):
    with planted(PLANTS[plant]):
        with pytest.raises(AssertionError):
            test()


def test_vertical_wet_dry_mirrors_horizontal():
    """Reset_states_vertical on a column is Reset_states_horizontal on the
    transposed row (the mask, levels and FS; the face velocities swap)."""
    kh_row = [2, 0, 0, 1, 1, 0, 0, 0, 1, 2, 2, 2]
    tz_row = [0, .5, .5, 0, 0, .4, .4, .25, 0, 0, 0, 0]
    z_row = [.3] * 12
    case, oh = wetdry_line(kh_row, tz_row, z_row, hm_row=[2, 2, 2, 2, 2, 2, .01, 2, 2, 2, 2, 2])
    oh.kernel("reset_states_h")
    c2, ov = make(N=12, M=5)
    W = c2.M + 3
    kh = np.full((15, W), 2, np.int32); kh[1:13, 3] = kh_row
    tz = np.zeros((15, W)); tz[1:13, 3] = tz_row
    z = np.zeros((15, W)); z[1:13, 3] = z_row
    hm = np.full((15, W), 2.0); hm[7, 3] = .01
    for k, a in (("khouot", kh), ("t_z", tz), ("z", z), ("H_moi", hm),
                 ("htaiz", np.ones((15, W))), ("FS", np.zeros((15, W)))):
        ov.set(k, a)
    ov.kernel("reset_states_v")
    assert list(ov.get("khouot")[1:13, 3]) == row3(oh, "khouot")
    assert list(ov.get("t_z")[1:13, 3]) == row3(oh, "t_z")


# ----------------------------------------------------------------- sediment

SED = dict(Toe=0.15, Tod=0.06, Mbochat=1e-5, ros=2000, ro=1000, hstoe=0.5, ghtoe=3)


def test_cohesive_source_closed_forms():
    case, o = make(**SED)
    p = case.params
    # erosion: bed stress ro fw V^2 = 0.3 = 2 Toe -> S = Mbochat / ros
    assert o.source_chs(0.01, 0.2, 0.002, 0.0012, 0.5, 1.0) == pytest.approx(p.Mbochat / p.ros, rel=1e-14)
    # deeper than ghtoe: critical stress Toe (1 + hstoe (H - ghtoe)) = 0.3
    assert o.source_chs(0.01, 0.2, 0.002, 0.0024, 0.5, 5.0) == pytest.approx(p.Mbochat / p.ros, rel=1e-14)
    # still water: Pd = 1, beta = 1 + Zf / 6, S = -wsm beta fs
    assert o.source_chs(0.01, 0.6, 0.002, 0.0012, 0.0, 1.0) == pytest.approx(-0.01 * 1.1 * 0.002, rel=1e-14)
    # between Tod and Toe: no exchange
    assert o.source_chs(0.01, 0.6, 0.002, 0.001, 10 ** -0.5, 1.0) == 0.0
    # deposition at Pd = 1/2 (tob = Tod / 2): beta = 1 + Zf / (1.25 + 4.75 / 2^2.5)
    tob = 0.03
    fw = tob / (1000 * 0.25)
    beta = 1 + 0.6 / (1.25 + 4.75 * 0.5 ** 2.5)
    assert o.source_chs(0.01, 0.6, 0.002, fw, 0.5, 1.0) == pytest.approx(-0.01 * beta * 0.002 * 0.5, rel=1e-13)
    with planted(PLANTS["CHS_EROSION"]):
        assert o.source_chs(0.01, 0.2, 0.002, 0.0012, 0.5, 1.0) != pytest.approx(p.Mbochat / p.ros, rel=1e-6)


def test_noncohesive_source_closed_forms():
    case, o = make(dm=2e-4, Sx=2.65)
    p = o.params()                      # with the derived Ufr, Dxr
    assert p.Ufr > 0 and p.Dxr > 0
    # at the critical shear velocity no pick-up: S = -wsm gamac fs
    for Zf in (0.0, 0.1, 0.3):
        gamac = 0.434 + 5.975 * Zf + 2.888 * Zf ** 2
        assert o.source_nchs(0.02, Zf, p.Ufr, 2.0, 0.001) == pytest.approx(-0.02 * gamac * 0.001, rel=1e-13)
    # transport stage T = 1 (Uf^2 = 2 Ufr^2): Cac = 0.015 d T^1.5 D*^-0.3 / (0.05 H)
    Uf = math.sqrt(2.0) * p.Ufr
    cac = 0.015 * p.dm * p.Dxr ** -0.3 / (0.05 * 2.0)
    assert o.source_nchs(0.02, 0.0, Uf, 2.0, 0.0) == pytest.approx(0.02 * cac, rel=1e-12)
    with planted(PLANTS["NCHS_GAMAC"]):
        assert o.source_nchs(0.02, 0.1, p.Ufr, 2.0, 0.001) != pytest.approx(
            -0.02 * (0.434 + 0.5975 + 0.02888) * 0.001, rel=1e-6)


def sed_patch(N=6, M=6, **over):
    """All cells of the padded grid wet (khouot 0 inside the ghost frame)."""
    case, o = make(N=N, M=M, **dict(SED, **over))
    kh = np.full((N + 3, M + 3), 2, np.int32)
    kh[1:N + 2, 1:M + 2] = 0
    o.set("khouot", kh)
    return case, o


def test_vth_is_speed_of_face_means():
    case, o = sed_patch()
    x, y = coords(case)
    tu = 0.2 + 0.002 * x + 0.001 * y          # mean of faces (i, j), (i-1, j)
    tv = 0.1 + 0.001 * x + 0.003 * y          # mean of faces (i, j), (i, j-1)
    hm = full(case, 2.0); hm[4, 4] = 0.01
    for k, a in (("t_u", tu), ("t_v", tv), ("H_moi", hm), ("htaiz", full(case, 2.0))):
        o.set(k, a)

    def check():
        o.kernel("find_vth")
        V = o.get("VTH")
        p = case.params
        for i in range(2, case.N + 1):
            for j in range(2, case.M + 1):
                ub = 0.2 + 0.002 * (i - 0.5) * p.dX + 0.001 * j * p.dY
                vb = 0.1 + 0.001 * i * p.dX + 0.003 * (j - 0.5) * p.dY
                want = 0.0 if (i, j) == (4, 4) else math.hypot(ub, vb)
                if abs(V[i, j] - want) > 1e-14:
                    return False
        return True

    assert check()
    with planted(PLANTS["VTH_AVG"]):
        assert not check()


def test_hesok_closed_form_and_clamps():
    case, o = sed_patch()
    p = case.params
    # Cz = 18 ln(12 H / Ks) = 36 at H = Ks e^2 / 12; Kx = 5.93 sqrt(g) hbar/2 |t_u| / Cz
    H = p.Ks * math.exp(2.0) / 12
    h = full(case, 8.0)
    h[:, 3] = 4.0                       # node depths (i, j-1) and (i, j) of the faces of column 4 differ
    tu = full(case, -4.0); tu[2, 4] = 0.5; tu[3, 4] = 90.0
    for k, a in (("Htdu", full(case, H)), ("Htdv", full(case, 0.0)), ("h", h), ("t_u", tu)):
        o.set(k, a)

    def kx(i, j):
        o.kernel("heso_k")
        return o.get("Kx")[i, j]

    want = 5.93 * math.sqrt(p.g) * (4.0 + 8.0) * 0.25 * 4.0 / 36.0
    assert kx(4, 4) == pytest.approx(want, rel=1e-14)
    assert kx(2, 4) == 5.0 and kx(3, 4) == 100.0
    with planted(PLANTS["HESOK_DEPTH"]):
        assert kx(4, 4) != pytest.approx(want, rel=1e-6)


def qb_setup(mode):
    case, o = sed_patch(dm=2e-4, Sx=2.65)
    x, y = coords(case)
    # face velocities vary in both directions; their cell means are 0.3, 0.4
    tu = 0.3 + 0.004 * (x - 4 * case.params.dX) + 0.01 * np.sin(y)
    tv = 0.4 + 0.003 * (y - 4 * case.params.dY) + 0.01 * np.sin(x)
    o.set("t_u", tu)
    o.set("t_v", tv)
    o.set("H_moi", full(case, 2.0))
    o.set("htaiz", full(case, 2.0))
    o.kernel("find_vth")
    V = o.get("VTH")
    o.set("Fw", 0.3 / (1000 * np.where(V > 0, V, 1.0) ** 2))   # bed stress 2 Toe: T = 1
    o.kernel("calculate_qb", mode)
    return case, o, tu, tv, V


def test_bed_load_flux_closed_form():
    """Calculate_Qb at transport stage T = 1: |Qb| = 0.053 sqrt((s-1) g)
    d^1.5 D*^-0.3 along the cell-mean velocity; the ghost rows and columns
    copy their neighbours."""
    case, o, tu, tv, V = qb_setup(1)
    p = o.params()
    mag = 0.053 * math.sqrt((p.Sx - 1) * p.g) * p.dm ** 1.5 * p.Dxr ** -0.3
    Qx, Qy = o.get("Qbx"), o.get("Qby")
    N, M = case.N, case.M
    for i in range(2, N + 1):
        for j in range(2, M + 1):
            ub = (tu[i, j] + tu[i - 1, j]) / 2
            vb = (tv[i, j] + tv[i, j - 1]) / 2
            assert Qx[i, j] == pytest.approx(mag * ub / V[i, j], rel=1e-12)
            assert Qy[i, j] == pytest.approx(mag * vb / V[i, j], rel=1e-12)
    assert np.array_equal(Qx[1, 2:M + 1], Qx[2, 2:M + 1]) and np.array_equal(Qx[N + 1, 2:M + 1], Qx[N, 2:M + 1])
    assert np.array_equal(Qy[2:N + 1, 1], Qy[2:N + 1, 2]) and np.array_equal(Qy[2:N + 1, M + 1], Qy[2:N + 1, M])
    assert Qx[1, 1] == Qx[2, 2] and Qx[N + 1, M + 1] == Qx[N, M]
    with planted(PLANTS["QB_AVG_INDEX"]):
        case, o, tu, tv, V = qb_setup(1)
        ub = (tu[3, 3] + tu[2, 3]) / 2
        assert o.get("Qbx")[3, 3] != pytest.approx(mag * ub / V[3, 3], rel=1e-6)


def test_bed_load_flux_reference_mode_reads_as_written():
    """qb_mode REFERENCE evaluates the flux only where VTH == 0: zero where
    the flow moves, 0/0 where it is still."""
    case, o, tu, tv, V = qb_setup(0)
    Qx = o.get("Qbx")
    assert np.all(Qx[2:case.N + 1, 2:case.M + 1] == 0)
    o.set("VTH", full(case, 0.0))
    o.kernel("calculate_qb", 0)
    assert np.isnan(o.get("Qbx")[3, 3])


def bed_patch(alpha=2e-6, gamma=3e-9, beta=-1e-6, c2x=1e-7, c2y=2e-7, dried=None, tu_flat1=1.0):
    """A 5 x 5 interior (rows and columns 2..6), every cell wet: Qbx =
    alpha x + gamma x^2, Qby = beta y, FS = 0.01 + c2x x^2 + c2y y^2,
    Htdu Kx = 4 * 10, Htdv Ky = 4 * 5, and a bed stress between Tod and Toe
    (no suspended-sediment exchange)."""
    case, o = sed_patch(N=6, M=6, Dorong=0.4)
    x, y = coords(case)
    kh = o.get("khouot")
    if dried is not None:
        kh[dried] = 1
    o.set("khouot", kh)
    o.kernel("find_limits")
    for k, a in (("Qbx", alpha * x + gamma * x ** 2), ("Qby", beta * y),
                 ("FS", 0.01 + c2x * x ** 2 + c2y * y ** 2), ("Htdu", full(case, 4.0)),
                 ("Htdv", full(case, 4.0)), ("Kx", full(case, 10.0)), ("Ky", full(case, 5.0)),
                 ("H_moi", full(case, 2.0)), ("VTH", full(case, 0.01)), ("Fw", full(case, 1.0)),
                 ("dH", full(case, 0.0))):
        o.set(k, a)
    tu = full(case, -0.2)
    tu.flat[1] = tu_flat1
    o.set("t_u", tu)
    o.set("t_v", full(case, 0.1))
    return case, o, x, y


def test_bed_change_of_a_hand_computed_patch():
    """_bed_load with ketdinh: dH += dT / (1 - porosity) (S + dQbx/dx +
    dQby/dy + d/dx(H Kx dFS/dx) + d/dy(H Ky dFS/dy)); S = 0 here (ro Fw V^2
    between Tod and Toe) and every difference is exact on these fields."""
    def check():
        case, o, x, y = bed_patch()
        p = case.params
        o.kernel("bed_load", 1, 2, case.N)
        dH = o.get("dH")
        for i in range(2, case.N):            # BedLoad rows sidx .. eidx-1
            for j in range(2, case.M + 1):
                div = 2e-6 + 2 * 3e-9 * x[i, j] - 1e-6
                diff = 40.0 * 2 * 1e-7 + 20.0 * 2 * 2e-7
                want = p.dT / (1 - 0.4) * (div + diff)
                if abs(dH[i, j] - want) > 1e-12 * abs(want):
                    return False
        return True

    assert check()
    with planted(PLANTS["BED_DIV_SIGN"]):
        assert not check()


def test_bed_change_one_sided_next_to_a_dried_cell():
    """Upstream (i-1) neighbour dried with u < 0: the x-divergence is the
    second-order forward difference (exact for the quadratic Qbx); the
    downstream (i+1) neighbour dried: the backward difference, taken only
    when t_u at flat index 1 is non-zero (the reference tests t_u[pos > 0])."""
    case, o, x, y = bed_patch(dried=(3, 4), c2x=0.0, c2y=0.0)
    p = case.params
    o.kernel("bed_load", 1, 2, case.N)
    dH = o.get("dH")
    want = lambda i, j: p.dT / (1 - 0.4) * (2e-6 + 2 * 3e-9 * x[i, j] - 1e-6)
    assert dH[4, 4] == pytest.approx(want(4, 4), rel=1e-12)     # neighbour (3, 4) dried: forward
    assert dH[2, 4] == pytest.approx(want(2, 4), rel=1e-12)     # neighbour (3, 4) dried: backward
    case, o, x, y = bed_patch(dried=(3, 4), c2x=0.0, c2y=0.0, tu_flat1=0.0)
    o.kernel("bed_load", 1, 2, case.N)
    assert o.get("dH")[2, 4] == pytest.approx(p.dT / (1 - 0.4) * (-1e-6), rel=1e-12)
