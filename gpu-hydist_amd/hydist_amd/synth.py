"""Seeded synthetic cases in the reference's array conventions.

Every generator returns a :class:`Case` whose host arrays are exactly the
globals that ``Global_Variables.py`` builds from input files
(Global_Variables.py:30-133): node depths ``h`` of shape (N+3, M+3) obtained by
``pad(((1, 2), (1, 2)), 'edge')`` of an N x M node map plus NANGDAY, the
roughness / viscosity / friction maps padded the same way, hourly boundary
series laid out ``seg * total_time + hour`` (Global_Variables.py:39-53),
``boundary_type`` (4 x segment_limit, edges up/down/left/right), ``bienQ``,
``khouot`` = 2 everywhere and zero state arrays.

Walls are encoded the reference's way, by zero node depth (constant.cuh:147-150).
Each open edge gets exactly one wet run, because only segment 0 of an edge line
receives a boundary value (support_funcs.cu:555-559).
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, Optional

import numpy as np

from .case import (BED_EVERY_STEP, BED_PERIODIC, BND_SERIES, QB_FIXED, HdParams,
                   HdStepOpts, make_case_header, params_from_dict, parse_case_header)

F64 = np.float64


@dataclasses.dataclass
class Case:
    name: str
    N: int
    M: int
    header: str
    arrays: Dict[str, np.ndarray]
    total_time: int
    Tmax: float
    kenhhepd: int = 0
    kenhhepng: int = 0
    boundary_mode: int = BND_SERIES
    ketdinh: int = 0
    qb_mode: int = QB_FIXED
    bed_mode: int = BED_PERIODIC
    sediment_start: float = 1e30
    bed_change_start: float = 1e30

    @property
    def params_dict(self):
        return parse_case_header(self.header)

    @property
    def params(self) -> HdParams:
        return params_from_dict(self.params_dict)

    @property
    def dT(self) -> float:
        return float(self.params_dict["dT"])

    @property
    def segment_limit(self) -> int:
        return int(self.params_dict["segment_limit"])

    @property
    def series_len(self) -> int:
        return int(self.arrays["bc_up"].size)

    @property
    def nsteps(self) -> int:
        return int(round(self.Tmax / self.dT))

    def opts(self, **over) -> HdStepOpts:
        o = HdStepOpts()
        o.kenhhepd = over.get("kenhhepd", self.kenhhepd)
        o.kenhhepng = over.get("kenhhepng", self.kenhhepng)
        o.boundary_mode = over.get("boundary_mode", self.boundary_mode)
        o.ketdinh = over.get("ketdinh", self.ketdinh)
        o.qb_mode = over.get("qb_mode", self.qb_mode)
        o.bed_mode = over.get("bed_mode", self.bed_mode)
        o.sediment_start = over.get("sediment_start", self.sediment_start)
        o.bed_change_start = over.get("bed_change_start", self.bed_change_start)
        return o


def pad_map(node_map: np.ndarray, nangday: float = 0.0) -> np.ndarray:
    """Global_Variables.py:71-73: pad ((1, 2), (1, 2)) with 'edge' (the
    transpose/flip of the file read is already applied by the generators)."""
    return np.pad(np.asarray(node_map, dtype=F64), ((1, 2), (1, 2)), "edge") + nangday


def empty_state(N: int, M: int, SL: int) -> Dict[str, np.ndarray]:
    shape = (N + 3, M + 3)
    a: Dict[str, np.ndarray] = {}
    for k in ("u", "v", "z", "t_u", "t_v", "t_z", "Kx1", "Ky1", "Htdu", "Htdv", "htaiz", "htaiz_bd",
              "Tsxw", "Tsyw", "H_moi", "VTH", "Qbx", "Qby", "FS", "tFS", "Kx", "Ky", "dH"):
        a[k] = np.zeros(shape, F64)
    a["khouot"] = np.full(shape, 2, np.int32)
    a["moci"] = np.zeros(N + 2, np.int32)
    a["mocj"] = np.zeros(M + 2, np.int32)
    a["daui"] = np.zeros((N + 2, SL), np.int32)
    a["cuoii"] = np.zeros((N + 2, SL), np.int32)
    a["dauj"] = np.zeros((M + 2, SL), np.int32)
    a["cuoij"] = np.zeros((M + 2, SL), np.int32)
    a["hi"] = np.zeros(2 * (M + N + 6), F64)
    a["ubt"] = np.zeros(M + 2, F64)
    a["ubp"] = np.zeros(M + 2, F64)
    a["vbt"] = np.zeros(N + 2, F64)
    a["vbd"] = np.zeros(N + 2, F64)
    return a


def _series(values: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(values, dtype=F64).reshape(-1))


def remove_pinches(h: np.ndarray, fill: float = 0.1, max_iter: int = 64) -> np.ndarray:
    """Raise zero nodes that leave a wall face between two wet cells.

    A face whose two nodes have zero depth has zero total depth (Htdu/Htdv,
    support_funcs.cu:351-352) while the cells on both sides can still be wet
    (their 4-node mean is positive, support_funcs.cu:42).  The reference then
    divides by that depth (UVZSolver_multithread.cu:325, :377, :738, :785) and
    the run turns to NaN.  Real bathymetry avoids such one-node pinches; the
    synthetic maps are made regular the same way."""
    h = h.copy()
    for _ in range(max_iter):
        wet = (h[:-1, :-1] + h[:-1, 1:] + h[1:, :-1] + h[1:, 1:]) * 0.25 > 0  # cell (i+1, j+1)
        n_fix = 0
        # u-face of cell (i, j): nodes (i, j-1), (i, j); cells (i, j) and (i+1, j)
        zu = (h[1:-1, :-1] + h[1:-1, 1:]) == 0                       # face at node row r=1..R-2
        pin_u = zu & wet[:-1, :] & wet[1:, :]
        for r, c in np.argwhere(pin_u):
            h[r + 1, c + 1] = fill
            n_fix += 1
        wet = (h[:-1, :-1] + h[:-1, 1:] + h[1:, :-1] + h[1:, 1:]) * 0.25 > 0
        # v-face of cell (i, j): nodes (i-1, j), (i, j); cells (i, j) and (i, j+1)
        zv = (h[:-1, 1:-1] + h[1:, 1:-1]) == 0
        pin_v = zv & wet[:, :-1] & wet[:, 1:]
        for r, c in np.argwhere(pin_v):
            h[r + 1, c + 1] = fill
            n_fix += 1
        if n_fix == 0:
            return h
    raise RuntimeError("could not remove wall pinches from the depth map")


def _finish(name, N, M, SL, header, hn, n_manning, visc, btype, series, total_time, Tmax, **kw) -> Case:
    a = empty_state(N, M, SL)
    a["h"] = remove_pinches(pad_map(hn))
    a["hsnham"] = pad_map(np.full((N, M), n_manning) if np.isscalar(n_manning) else n_manning)
    a["VISCOINDX"] = pad_map(np.full((N, M), visc))
    # friction factor of the cohesive source terms, g n^2 / h^(1/3), with the
    # depth floored at 1 m (no case header gives it; an unbounded value on
    # near-dry banks drives the bed-load transport to meaningless rates)
    depth = np.maximum(hn, 1.0)
    fw = 9.81 * (n_manning if np.isscalar(n_manning) else n_manning) ** 2 / depth ** (1.0 / 3.0)
    a["Fw"] = pad_map(fw)
    bt = np.zeros((4, SL), np.int32)
    for e, t in enumerate(btype):
        bt[e, 0] = t
    a["boundary_type"] = bt
    bienQ = np.zeros(4, np.int32)
    bienQ[np.where(bt == 2)[0]] = 1          # Global_Variables.py:66
    a["bienQ"] = bienQ
    for k in ("bc_up", "bc_down", "bc_left", "bc_right", "CC_u", "CC_d", "CC_l", "CC_r"):
        a[k] = _series(series.get(k, np.zeros(total_time)))
        assert a[k].size == total_time
    return Case(name=name, N=N, M=M, header=header, arrays=a, total_time=total_time, Tmax=Tmax, **kw)


def channel_case(N: int = 300, M: int = 60, dx: float = 10.0, dt: float = 2.0, nsteps: int = 3600,
                 Q: float = 500.0, seed: int = 1, slope: float = 1e-4, depth: float = 5.0,
                 noise: float = 0.05, n_manning: float = 0.025, SL: int = 20,
                 name: str = "channel") -> Case:
    """Straight rectangular channel along i: Q at the left edge (row 2),
    water level 0 at the right edge (row N), side walls two node columns thick
    on both sides (so the edge columns 2 and M stay dry), narrow-channel
    horizontal mode (kenhhepng = 1) with series boundaries."""
    rng = np.random.default_rng(seed)
    x = np.arange(N, dtype=F64)[:, None] * dx
    hn = depth + slope * x + rng.normal(0.0, noise, (N, M))
    hn = np.broadcast_to(hn, (N, M)).copy()
    hn[:, :2] = 0.0
    hn[:, -2:] = 0.0
    Tmax = nsteps * dt
    total_time = int(math.ceil(Tmax / 3600.0)) + 3
    series = {"bc_left": np.full(total_time, Q), "bc_right": np.zeros(total_time)}
    header = make_case_header(dx, dx, dt, SL, H_TINH=0.02, heso=0.94)
    # edges: up (j=M), down (j=2), left (i=2), right (i=N)
    return _finish(name, N, M, SL, header, hn, n_manning, 1e-6, (0, 0, 2, 1), series, total_time, Tmax,
                   kenhhepng=1, boundary_mode=BND_SERIES)


def river_nodes(N: int, M: int, amp: float, wavelength: float, wmin: float, wmax: float, depth: float,
                noise: float, n_islands: int, rmin: float, rmax: float, seed: int,
                edge_margin: int = 20, open_floor: float = 0.0) -> np.ndarray:
    """Node depth map (N x M) of a meandering channel running along j: the
    centreline i_c(j) = N/2 + amp sin(2 pi j / wavelength), local width
    varying between wmin and wmax (seeded-phase sinusoid), parabolic cross
    section of maximum depth `depth` measured across the channel axis, seeded
    Gaussian bed noise, and seeded circular dry islands kept at least
    `edge_margin` cells from the open edges j = 0 and j = M-1.  With
    open_floor > 0 the straight reaches next to the open edges get that
    minimum depth across the whole channel width (steep banks), so a large
    tide does not dry and re-wet cells on the open boundary itself, and the
    islands stay out of those reaches."""
    rng = np.random.default_rng(seed)
    j = np.arange(M, dtype=F64)[None, :]
    i = np.arange(N, dtype=F64)[:, None]
    k = 2.0 * math.pi / wavelength
    # straight inlet / outlet reaches: the meander amplitude tapers to zero
    # over `taper` cells next to the open edges, so the channel meets both
    # open edges at right angles
    taper = max(4.0, 0.1 * M)
    dist = np.minimum(j, (M - 1) - j)
    x = np.clip((dist - 0.5 * taper) / (0.5 * taper), 0.0, 1.0)
    env = x * x * (3.0 - 2.0 * x)
    denv = np.where((x > 0) & (x < 1), 6.0 * x * (1.0 - x) / (0.5 * taper), 0.0) * np.where(j < (M - 1) / 2.0, 1.0, -1.0)
    ic = N / 2.0 + amp * env * np.sin(k * j)
    s = amp * (env * k * np.cos(k * j) + denv * np.sin(k * j))
    phase = rng.uniform(0.0, 2.0 * math.pi)
    w = 0.5 * (wmin + wmax) + 0.5 * (wmax - wmin) * np.sin(2.0 * math.pi * j / (0.73 * wavelength) + phase)
    d = np.abs(i - ic) / np.sqrt(1.0 + s * s)
    inside = d < 0.5 * w
    hn = np.where(inside, depth * (1.0 - (2.0 * d / w) ** 2), 0.0)
    nz = rng.normal(0.0, noise, (N, M))
    nz[:, :3] = 0.0
    nz[:, -3:] = 0.0
    hn = np.where(inside, np.maximum(hn + nz, 0.0), 0.0)
    if open_floor > 0.0:
        reach = (dist < taper) & inside
        hn = np.where(reach, np.maximum(hn, open_floor), hn)
    jj = np.arange(M)[None, :]
    ii = np.arange(N)[:, None]
    # with open_floor the islands also stay out of the straight open-edge reaches
    margin = max(edge_margin, taper) if open_floor > 0.0 else edge_margin
    for _ in range(n_islands):
        j0 = rng.uniform(margin + rmax, M - 1 - margin - rmax)
        xd = min(max((min(j0, (M - 1) - j0) - 0.5 * taper) / (0.5 * taper), 0.0), 1.0)
        ic0 = N / 2.0 + amp * xd * xd * (3.0 - 2.0 * xd) * math.sin(k * j0)
        w0 = 0.5 * (wmin + wmax) + 0.5 * (wmax - wmin) * math.sin(2.0 * math.pi * j0 / (0.73 * wavelength) + phase)
        i0 = ic0 + rng.uniform(-0.2, 0.2) * w0
        r = rng.uniform(rmin, rmax)
        hn[(ii - i0) ** 2 + (jj - j0) ** 2 < r * r] = 0.0
    return hn


def river_case(N: int = 980, M: int = 1048, dx: float = 20.0, dt: float = 5.0, hours: float = 6.0,
               amp: float = 150.0, wavelength: float = 400.0, wmin: float = 30.0, wmax: float = 60.0,
               depth: float = 8.0, noise: float = 0.1, n_islands: int = 20, rmin: float = 3.0,
               rmax: float = 10.0, n_manning: float = 0.03, seed: int = 2, SL: int = 20,
               q_mean: float = 3000.0, q_amp: float = 1000.0, q_period_h: float = 6.0,
               tide_amp: float = 0.5, tide_period_h: float = 12.42, H_TINH: float = 0.05,
               sediment: bool = False, name: str = "river", axis: str = "i", open_floor: float = 0.0,
               **header_over) -> Case:
    """Meandering river: discharge hydrograph q_mean - q_amp cos(2 pi t / q_period)
    entering at the down edge (column 2), tide tide_amp sin(2 pi t / tide_period)
    at the up edge (column M), solid left/right edges.  Series are hourly
    samples, the reference's boundary format (support_funcs.cu:547-565)."""
    if axis == "i":
        hn = river_nodes(M, N, amp, wavelength, wmin, wmax, depth, noise, n_islands, rmin, rmax, seed,
                         edge_margin=max(3, int(round(20 * min(1.0, N / 980.0)))), open_floor=open_floor).T.copy()
    else:
        hn = river_nodes(N, M, amp, wavelength, wmin, wmax, depth, noise, n_islands, rmin, rmax, seed,
                         edge_margin=max(3, int(round(20 * min(1.0, M / 1048.0)))), open_floor=open_floor)
    Tmax = hours * 3600.0
    total_time = int(math.ceil(hours)) + 3
    th = np.arange(total_time, dtype=F64)
    qs = q_mean - q_amp * np.cos(2.0 * math.pi * th / q_period_h)
    zs = tide_amp * np.sin(2.0 * math.pi * th / tide_period_h)
    if axis == "i":
        series = {"bc_left": qs, "bc_right": zs}
        btype = (0, 0, 2, 1)
        cc_key = "CC_l"
    else:
        series = {"bc_down": qs, "bc_up": zs}
        btype = (1, 2, 0, 0)
        cc_key = "CC_d"
    # heso: weight of a face velocity against its neighbour mean in the
    # smoothing pass (Normalize, support_funcs.cu:682-703); at the channel
    # case's 0.94 this river's run fragments its wet mask within hours, at
    # 0.8 it integrates stably for days
    over = dict(H_TINH=H_TINH, heso=0.8)
    kw = {}
    if sediment:
        over.update(dm=2e-4, ros=2650.0, Dorong=0.4, Sx=2.65, NDnen=0.0, Toe=1.0)
        series[cc_key] = np.full(total_time, 0.3)
        kw.update(sediment_start=3600.0, bed_change_start=7200.0, ketdinh=0)
    over.update(header_over)
    header = make_case_header(dx, dx, dt, SL, **over)
    case = _finish(name, N, M, SL, header, hn, n_manning, 1e-6, btype, series, total_time, Tmax,
                   boundary_mode=BND_SERIES, **kw)
    if sediment:
        fs0 = 0.1 / 2650.0
        case.arrays["FS"][:] = fs0
        case.arrays["tFS"][:] = fs0
    return case


def config(n: int, scale: float = 1.0, hours: Optional[float] = None, **kw) -> Case:
    """The workloads: 1 straight channel 300x60 (3600 steps of 2 s),
    2 meandering river 980x1048 (6 h of 5 s steps), 3 = 2 with sediment and
    bed change (24 h).  `scale` shrinks the grid (and the meander) for tests."""
    if n == 1:
        N = max(8, int(round(300 * scale)))
        M = max(8, int(round(60 * scale))) if scale < 1 else 60
        steps = kw.pop("nsteps", 3600)
        return channel_case(N=N, M=M, nsteps=steps, name="channel_300x60" if scale == 1 else f"channel_{N}x{M}", **kw)
    if n in (2, 3):
        N = int(round(980 * scale))
        M = int(round(1048 * scale))
        h = hours if hours is not None else (6.0 if n == 2 else 24.0)
        name = ("river_980x1048" if scale == 1 else f"river_{N}x{M}") + ("_sed" if n == 3 else "")
        return river_case(N=N, M=M, hours=h, amp=150 * scale, wavelength=400 * scale,
                          wmin=max(6.0, 30 * scale), wmax=max(10.0, 60 * scale),
                          rmin=max(1.5, 3 * scale), rmax=max(2.5, 10 * scale),
                          n_islands=kw.pop("n_islands", 20 if scale >= 0.5 else 4),
                          q_mean=kw.pop("q_mean", 3000.0 * scale), q_amp=kw.pop("q_amp", 1000.0 * scale),
                          sediment=(n == 3), name=name, **kw)
    raise ValueError(f"no generator for config {n}")


def ensemble_member(k: int, seed: int = 4, attempt: int = 0, hours: float = 1.0, side_min: int = 256,
                    side_max: int = 1024, sediment: bool = True, tide_max: float = 1.0) -> Case:
    """Member k of the batched-ensemble workload (draw `attempt` of it): a
    meandering river with grid sides N, M ~ U{side_min..side_max} and a
    constant discharge Q ~ U[500, 5000] m3/s, Manning n ~ U[0.02, 0.04] and
    tide amplitude ~ U[0, tide_max] m; dx = 20 m, dT = 5 s, suspended
    sediment from t = 0 (bed change after half an hour).  The mask is the
    980 x 1048 river's: meander amplitude and wavelength scaled together with
    the grid, channel widths 30-60 cells, straight open-edge reaches at least
    2 m deep and free of islands."""
    rng = np.random.default_rng([seed, k, attempt])
    N = int(rng.integers(side_min, side_max + 1))
    M = int(rng.integers(side_min, side_max + 1))
    q = float(rng.uniform(500.0, 5000.0))
    nm = float(rng.uniform(0.02, 0.04))
    tide = float(rng.uniform(0.0, 1.0)) * tide_max
    si, sj = N / 980.0, M / 1048.0
    c = river_case(N=N, M=M, hours=hours, amp=150.0 * min(si, sj), wavelength=400.0 * si,
                   n_islands=max(4, int(round(20 * si * sj))), n_manning=nm, q_mean=q, q_amp=0.0,
                   tide_amp=tide, seed=1000 + 97 * k + attempt, sediment=sediment,
                   name=f"ens{k:02d}.{attempt}_{N}x{M}", open_floor=2.0, SL=24)
    if sediment:
        c.sediment_start = 0.0
        c.bed_change_start = 1800.0
    return c


def ensemble_cases(n: int = 64, seed: int = 4, accept=None, max_attempts: int = 8, **kw):
    """The n members of the batched-ensemble workload.  With `accept` (a
    callable on a Case), a member the callable rejects is redrawn (attempt
    1, 2, ...): the reference's FP64 method goes unstable on some seeded
    geometries (the CPU oracle reproduces it), and a benchmark runs the
    members that it integrates."""
    out = []
    for k in range(n):
        for attempt in range(max_attempts):
            c = ensemble_member(k, seed=seed, attempt=attempt, **kw)
            if accept is None or accept(c):
                break
        out.append(c)
    return out
