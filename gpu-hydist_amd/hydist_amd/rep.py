"""Command-line driver with the reference's ``rep.py`` options.

rep.py:23-42 defines --GPU --Device --mins --hours --plot --test --sediment
--bed_change --pick_up --pickup_dirs --kenhhepng --kenhhepd --include_dirs;
``gpu_init`` (rep.py:58-142) creates the context, the Pointers registry and
the device structs, uploads the inputs and runs the init kernels; ``main``
(rep.py:145-173) runs hydraulic_Calculation and prints the wall time.

The same names are kept, with types (the reference leaves --plot,
--sediment, --bed_change, --pick_up and --kenhhep* untyped, so a value given
on the command line arrives as a string).  New options select the case
(--case: a synthetic workload, or a DirsOfInputs file with --constants) and
the explicit modes (--boundary, --qb_mode, --bed_mode).
"""
from __future__ import annotations

import argparse
import os
import sys
import timeit

import numpy as np

from . import synth
from .engine import hydraulic_Calculation
from .pointers import Pointers, PointersStruct
from ._lib import Solver


def _bool(s):
    if isinstance(s, bool):
        return s
    return str(s).lower() in ("1", "true", "yes", "y", "on")


def make_parser():
    p = argparse.ArgumentParser(description="HYDIST ADI step on MI355X (libhydist)")
    p.add_argument("--GPU", type=int, default=1)
    p.add_argument("--Device", type=int, default=0)
    p.add_argument("--mins", type=int, default=0)
    p.add_argument("--hours", type=int, default=10)
    p.add_argument("--plot", type=_bool, default=False)
    p.add_argument("--test", type=_bool, default=False)
    p.add_argument("--sediment", type=float, default=10)
    p.add_argument("--bed_change", type=float, default=20)
    p.add_argument("--pick_up", type=_bool, default=False)
    p.add_argument("--pickup_dirs", default="/Input/Initial_Condition")
    p.add_argument("--kenhhepng", type=_bool, default=False)
    p.add_argument("--kenhhepd", type=_bool, default=False)
    p.add_argument("--include_dirs", default=".")
    p.add_argument("--case", default="river",
                   help="channel | river | river_sed | path to a DirsOfInputs file")
    p.add_argument("--constants", default=None, help="constant.cuh-format case header")
    p.add_argument("--scale", type=float, default=1.0, help="grid scale of a synthetic case")
    p.add_argument("--boundary", default="auto", choices=("auto", "series", "analytic"))
    p.add_argument("--qb_mode", default="fixed", choices=("fixed", "reference"))
    p.add_argument("--bed_mode", default="periodic", choices=("periodic", "every_step"))
    return p


class Context:
    """Stands where the reference's PyCUDA context is (rep.py:59-61): it owns
    the libhydist context of one case on one device."""

    def __init__(self, device, case):
        self.solver = Solver(params=case.params, N=case.N, M=case.M, series_len=case.series_len,
                             total_time=case.total_time, device=device)
        self.solver.case = case

    def synchronize(self):
        self.solver.sync()

    def detach(self):
        self.solver.close()


def load_intial_condition(dirs, pointers):
    """Pick-up restart (rep.py:43-55): u, v, z, FS (text, one header row) and
    khouot, edge-padded (0, 1) / padded with 2, then uploaded."""
    def txt(name, **kw):
        return np.loadtxt(os.path.join(dirs, name), **kw)
    u = np.pad(txt("u.txt", skiprows=1), ((0, 1), (0, 1)), "edge")
    v = np.pad(txt("v.txt", skiprows=1), ((0, 1), (0, 1)), "edge")
    z = np.pad(txt("z.txt", skiprows=1), ((0, 1), (0, 1)), "edge")
    kh = np.pad(txt("khouot.txt", dtype=np.int32), ((0, 1), (0, 1)), "constant", constant_values=((2, 2), (2, 2)))
    FS = np.pad(txt("FS.txt"), ((0, 1), (0, 1)), "edge")
    for k, a in (("u", u), ("v", v), ("z", z), ("khouot", kh.astype(np.int32)), ("FS", FS)):
        pointers.host_ptrs[k] = a
    pointers.toDevice(["u", "v", "z", "khouot", "FS"])


def save_pick_up(dirs, pointers, t=0.0):
    """The files load_intial_condition reads (rep.py:43-55), which nothing in
    the reference writes: u, v, z with one header line, khouot and FS, each
    the (N+2) x (M+2) leading block of the padded field, written with 17
    significant digits so that a restart reproduces the doubles exactly."""
    os.makedirs(dirs, exist_ok=True)
    s = pointers._solver()
    N, M = s.N, s.M
    for name in ("u", "v", "z"):
        a = s.get(name)[:N + 2, :M + 2]
        np.savetxt(os.path.join(dirs, name + ".txt"), a, fmt="%.17g", header=f"{name} N={N} M={M} t={t!r}",
                   comments="")
    np.savetxt(os.path.join(dirs, "khouot.txt"), s.get("khouot")[:N + 2, :M + 2], fmt="%d")
    np.savetxt(os.path.join(dirs, "FS.txt"), s.get("FS")[:N + 2, :M + 2], fmt="%.17g")


def gpu_init(device_no, pick_up, pickup_dirs, case):
    ctx = Context(device_no, case)
    a = case.arrays
    kwargs = {k: a[k] for k in ("h", "hsnham", "VISCOINDX", "H_moi", "bienQ", "moci", "mocj", "dauj", "daui",
                                "cuoii", "cuoij", "Tsxw", "Tsyw", "khouot", "boundary_type", "u", "v", "z",
                                "t_u", "t_v", "t_z", "Htdu", "Htdv", "Kx1", "Ky1", "htaiz", "htaiz_bd",
                                "bc_up", "bc_down", "bc_left", "bc_right", "ubt", "ubp", "vbt", "vbd", "hi",
                                "FS", "tFS", "CC_u", "CC_d", "CC_l", "CC_r", "VTH", "Kx", "Ky", "Fw", "Qbx",
                                "Qby", "dH")}
    pointers = Pointers(ctx, dtype=np.float64, **kwargs)
    pointers.alloc_on_device_only(case.N, case.M)
    pointers.alloc()
    arg_struct = PointersStruct(None, structtype="ARG", ctx=ctx)
    arr_struct = PointersStruct(None, structtype="ARR", ctx=ctx)
    pointers.toDevice(["h", "hsnham", "VISCOINDX", "bienQ", "Tsyw", "Tsxw", "boundary_type", "bc_up", "bc_down",
                       "bc_left", "bc_right", "u", "v", "z", "CC_u", "CC_d", "CC_l", "CC_r", "Fw",
                       "FS", "tFS"])
    ctx.solver.init()
    if pick_up:
        load_intial_condition(pickup_dirs, pointers)
        ctx.solver.prepare()
    arg_struct = PointersStruct(None, structtype="ARG", ctx=ctx)
    arr_struct = PointersStruct(None, structtype="ARR", ctx=ctx)
    return pointers, ctx, arg_struct.ptr, arr_struct.ptr, ctx.solver


def resolve_case(args):
    if args.case in ("channel", "river", "river_sed"):
        n = {"channel": 1, "river": 2, "river_sed": 3}[args.case]
        return synth.config(n, scale=args.scale) if n != 1 or args.scale != 1 else synth.config(1)
    from .inputs import load_case
    hdr = args.constants or os.path.join(args.include_dirs, "constant.cuh")
    with open(hdr) as f:
        return load_case(args.case, f.read(), name=os.path.basename(args.case))


def main(argv=None):
    args = make_parser().parse_args(argv)
    hours = args.hours
    mins = args.mins
    secs = 0
    Tmax = hours * 3600 + mins * 60 + secs
    plot = args.plot
    debug = False
    sediment_start = args.sediment * 3600
    bedchange_start = args.bed_change * 3600
    interval = 3600
    export = False
    save_grid = False
    case = resolve_case(args)
    kenhhepd, kenhhepng = int(args.kenhhepd), int(args.kenhhepng)
    if args.case == "channel" and not (args.kenhhepd or args.kenhhepng):
        kenhhepng = case.kenhhepng
    boundary = args.boundary
    if args.case in ("channel", "river", "river_sed") and boundary == "auto":
        boundary = "series"
    pointers, ctx, arg_struct_ptr, arr_struct_ptr, supmod = gpu_init(args.Device, args.pick_up, args.pickup_dirs,
                                                                     case)
    start_time = timeit.default_timer()
    ulist, vlist, zlist, fslist = hydraulic_Calculation(
        Tmax, pointers, arg_struct_ptr, arr_struct_ptr, supmod, ctx, kenhhepd=kenhhepd, kenhhepng=kenhhepng,
        save_grid=save_grid, sediment_start=sediment_start, debug=debug, plot=plot, interval=interval,
        export=export, bed_change_start=bedchange_start, boundary_mode=boundary, qb_mode=args.qb_mode,
        bed_mode=args.bed_mode, ketdinh=case.ketdinh if args.case.startswith("river") else True)
    stop_time = timeit.default_timer()
    print("total running time of ", hours, "hours is ", stop_time - start_time)
    ctx.detach()
    return ulist, vlist, zlist, fslist


if __name__ == "__main__":
    main()
