"""Output writers of the reference's driver.

* ``save_fs_grid``: the Surfer DSAA text grid of FS * ros (Visualizor.py:32-47):
  header ``DSAA``, ``N-1 M-1``, ``2 N``, ``2 M``, the value range, then rows
  i = 2..N-1 of columns j = 2..M-1 (the ranges the reference writes).
* ``export_vth_and_level``: the hourly depth-averaged speed / direction and
  water-level text files of Engine.py:433-464.
"""
from __future__ import annotations

import math
import os

import numpy as np


def save_fs_grid(FS, filename, ros, mode="DSAA", val_range=(0, 10)):
    N, M = FS.shape[0] - 3, FS.shape[1] - 3
    os.makedirs(os.path.dirname(filename) or ".", exist_ok=True)
    with open(filename, "w") as f:
        f.write(mode + "\n")
        f.write(str(N - 1) + " " + str(M - 1) + "\n")
        f.write(str(2) + " " + str(N) + "\n")
        f.write(str(2) + " " + str(M) + "\n")
        f.write(str(val_range[0]) + " " + str(val_range[1]) + "\n")
        for i in range(2, N):
            f.write(" ".join(str(FS[i, j] * ros) for j in range(2, M)) + " \n")


def _angle(ut, vt):
    """Direction in degrees as Engine.py:448-460 computes it (the third
    quadrant adds 180, the value the reference's `angl += 180` intends)."""
    if ut != 0:
        a = 180 * math.atan(abs(vt / ut)) / math.pi
        if vt < 0 and ut > 0:
            a = 360 - a
        elif vt < 0 and ut < 0:
            a += 180
        elif vt > 0 and ut < 0:
            a = 180 - a
        return a
    if vt > 0:
        return 90.0
    if vt < 0:
        return 270.0
    return 0.0


def export_vth_and_level(t, u, v, z, htaiz, nangday, outdir):
    N, M = u.shape[0] - 3, u.shape[1] - 3
    vdir = os.path.join(outdir, "velocity")
    zdir = os.path.join(outdir, "water_level")
    os.makedirs(vdir, exist_ok=True)
    os.makedirs(zdir, exist_ok=True)
    with open(os.path.join(vdir, "VTH" + str(t) + "s.data"), "w") as vf, \
            open(os.path.join(zdir, "mucnuoc" + str(t) + "s.grid"), "w") as zf:
        for i in range(1, N + 1):
            for j in range(1, M + 1):
                vt = (v[i, j - 1] + v[i, j]) * 0.5
                ut = (u[i, j] + u[i - 1, j]) * 0.5
                vth = np.sqrt(vt * vt + ut * ut)
                if htaiz[i, j] <= nangday:
                    vth = 0
                vf.write("%.1f %d %d %.2f %.2f" % (t, i, j, vth, _angle(ut, vt)))
                zf.write("%.2f " % (z[i, j]))
            zf.write("\n")
