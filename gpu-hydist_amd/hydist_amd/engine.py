"""Time loop with the reference's ``hydraulic_Calculation`` interface.

Engine.py:49-469 advances ``t`` by two half-steps per iteration, launching the
reference kernels one by one with a host synchronisation after almost every
launch, extracts u, v, z (and FS) every ``interval`` seconds, stops on NaN,
and returns ``(u_list, v_list, z_list, fslist)``.  This version keeps the
signature, the return value, the extraction schedule and the NaN stop, and
runs the steps between two host events as one libhydist call (``hd_run``,
a replayed HIP graph per step, no host synchronisation inside).

``supmod`` / ``ctx`` are the library context returned by ``rep.gpu_init``;
``arg_struct_ptr`` / ``arr_struct_ptr`` are accepted for call compatibility.
New keyword options: ``boundary_mode`` ('auto' = the reference rule, channel
flags force the analytic tide, Engine.py:29-46; 'series'; 'analytic'),
``qb_mode`` ('fixed' or 'reference', Sediment_Transport.cu:365) and
``bed_mode`` ('periodic', Engine.py:364, or 'every_step').  With ``debug``
the half-steps run one by one so that t_u, t_v, t_z are recorded after each,
as the reference does; where the reference calls exit() on a NaN in debug
mode (Engine.py:383-393) this returns the lists gathered so far.
"""
from __future__ import annotations

import copy
import math
import os

import numpy as np

from .case import (BED_EVERY_STEP, BED_PERIODIC, BND_ANALYTIC, BND_AUTO, BND_SERIES, QB_FIXED,
                   QB_REFERENCE, HdStepOpts)

_BND = {"auto": BND_AUTO, "series": BND_SERIES, "analytic": BND_ANALYTIC}
_QB = {"fixed": QB_FIXED, "reference": QB_REFERENCE}
_BED = {"periodic": BED_PERIODIC, "every_step": BED_EVERY_STEP}


def step_options(kenhhepd=0, kenhhepng=0, ketdinh=True, sediment_start=36000, bed_change_start=72000,
                 boundary_mode="auto", qb_mode="fixed", bed_mode="periodic") -> HdStepOpts:
    o = HdStepOpts()
    o.kenhhepd = int(bool(kenhhepd))
    o.kenhhepng = int(bool(kenhhepng))
    o.boundary_mode = _BND[boundary_mode] if isinstance(boundary_mode, str) else int(boundary_mode)
    o.ketdinh = int(bool(ketdinh))
    o.qb_mode = _QB[qb_mode] if isinstance(qb_mode, str) else int(qb_mode)
    o.bed_mode = _BED[bed_mode] if isinstance(bed_mode, str) else int(bed_mode)
    o.sediment_start = float(sediment_start)
    o.bed_change_start = float(bed_change_start)
    return o


def _bed_trigger(t, o: HdStepOpts) -> bool:
    """Whether the step ending at t changes the bed (Engine.py:364; every
    step from bed_change_start in the every_step mode)."""
    if o.bed_mode == BED_EVERY_STEP:
        return t >= o.bed_change_start
    if t != int(t) or o.bed_change_start == 0:
        return False
    return math.fmod(int(t), o.bed_change_start) == 0


def _event(t, interval, sediment_start, save_grid, export, debug):
    """True when the reference does host work at the end of the step ending
    at t (Engine.py:245-257, 369-464)."""
    if debug:
        return True
    integral = t - int(t) == 0
    if interval and integral and int(t) % interval == 0:
        return True
    if save_grid and t >= sediment_start and integral and int(t) % 360 == 0:
        return True
    if export and integral and int(t) % 3600 == 0:
        return True
    return False


def hydraulic_Calculation(Tmax, pointers, arg_struct_ptr, arr_struct_ptr, supmod, ctx, ketdinh=True,
                          blk_size=1024, kenhhepd=0, kenhhepng=0, debug=True, interval=1,
                          sediment_start=36000, t_start=0, sec=1, plot=False, export=False, save_grid=True,
                          bed_change_start=72000, boundary_mode="auto", qb_mode="fixed",
                          bed_mode="periodic", output_dir="Outputs"):
    u_list, v_list, z_list, fslist = [], [], [], []
    solver = ctx.solver
    N, M = solver.N, solver.M
    dT = solver.params().dT
    opts = step_options(kenhhepd, kenhhepng, ketdinh, sediment_start, bed_change_start, boundary_mode,
                        qb_mode, bed_mode)
    print("Simulation time = ", int(Tmax) // 3600, "hours")
    print("N, M : ", N, " ", M)
    shape = (N + 3, M + 3)
    gpu_u, gpu_v, gpu_z = np.zeros(shape), np.zeros(shape), np.zeros(shape)
    FS = np.zeros(shape)
    t = float(t_start)
    while t < Tmax:
        if debug:
            # the reference records t_u, t_v, t_z after each half-step
            # (Engine.py:245-257, 369-393) and stops on a NaN in u, v or z
            for which in (0, 1):
                t = t + dT * 0.5
                solver.halfstep(which, t, opts)
                if which == 1 and _bed_trigger(t, opts):
                    solver.bed_change(t, opts)
                pointers.extract({"u": gpu_u, "v": gpu_v, "z": gpu_z})
                u_list.append(solver.get("t_u").copy())
                v_list.append(solver.get("t_v").copy())
                z_list.append(solver.get("t_z").copy())
                if np.isnan(gpu_u).any() or np.isnan(gpu_v).any() or np.isnan(gpu_z).any():
                    return u_list, v_list, z_list, fslist
        else:
            # steps until the next host event (or Tmax): replayed without syncs
            n, tt = 0, t
            while tt < Tmax:
                tt = tt + dT * 0.5
                tt = tt + dT * 0.5
                n += 1
                if _event(tt, interval, sediment_start, save_grid, export, debug):
                    break
            t = solver.run(t, n, opts)
        integral = t - int(t) == 0
        if interval and integral and int(t) % interval == 0:
            pointers.extract({"u": gpu_u, "v": gpu_v, "z": gpu_z, "FS": FS})
            print("hour: ", int(t) // 3600)
            if np.isnan(gpu_u).any() or np.isnan(gpu_v).any() or np.isnan(gpu_z).any():
                return u_list, v_list, z_list, fslist
            if plot:
                _visualize(gpu_u[1:N + 1, 1:M + 1], gpu_v[1:N + 1, 1:M + 1])
            v_list.append(copy.deepcopy(gpu_v))
            u_list.append(copy.deepcopy(gpu_u))
            z_list.append(copy.deepcopy(gpu_z))
        if save_grid and t >= sediment_start and integral and int(t) % 360 == 0:
            from .outputs import save_fs_grid
            pointers.extract({"FS": FS})
            save_fs_grid(FS, os.path.join(output_dir, "FS_GPU", "fs_" + str(t) + ".grd"), ros=solver.params().ros)
            fslist.append(copy.deepcopy(FS))
        if export and integral and int(t) % 3600 == 0:
            from .outputs import export_vth_and_level
            pointers.extract({"u": gpu_u, "v": gpu_v, "z": gpu_z})
            export_vth_and_level(t, gpu_u, gpu_v, gpu_z, solver.get("htaiz"), solver.params().NANGDAY,
                                 os.path.join(output_dir, "Song_Luy"))
    return u_list, v_list, z_list, fslist


def _visualize(u, v):
    try:
        import matplotlib.pyplot as plt
    except ImportError:
        return
    plt.figure(figsize=(10, 10))
    plt.quiver(u[::4, ::4], v[::4, ::4])
    plt.show()
