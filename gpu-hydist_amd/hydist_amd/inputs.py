"""Real-case input loader with the transforms of Global_Variables.py.

Global_Variables.py:27-133 reads a key -> path map (DirsOfInputs_1.txt), the
boundary types (one line per edge up/down/left/right, 'NA' for none), the
hourly boundary and concentration series (transposed so that a segment's
series is contiguous), and the topography / roughness / viscosity / friction
maps, each transposed, flipped along axis 1 and edge-padded ((1, 2), (1, 2));
N, M are taken after the transpose.  ``total_time`` is the length of the first
non-solid edge's series divided by its number of typed segments.
"""
from __future__ import annotations

import os
import warnings
from typing import Dict, Optional

import numpy as np

from .case import parse_case_header
from .synth import Case, empty_state


def read_dirs(path: str) -> Dict[str, str]:
    base = os.path.dirname(os.path.abspath(path))
    out = {}
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 2:
                p = parts[1]
                out[parts[0]] = p if os.path.isabs(p) else os.path.join(base, p)
    return out


def _map(path: str, nangday: float = 0.0) -> np.ndarray:
    a = np.loadtxt(path)
    a = np.transpose(a)
    a = np.flip(a, 1)
    return np.pad(a, ((1, 2), (1, 2)), "edge") + nangday


def device_view(a: np.ndarray, shape) -> np.ndarray:
    """What a kernel reads from a map uploaded at its own size and indexed
    with the grid's row width (Pointers.py:29-33 allocates each array at its
    nbytes; the kernels read flat index i*(M+3)+j): the first
    (N+3)*(M+3) values of the row-major map, reshaped to the grid."""
    flat = np.ascontiguousarray(a).reshape(-1)
    n = shape[0] * shape[1]
    if flat.size < n:
        raise ValueError(f"map of {flat.size} values is smaller than the {shape[0]}x{shape[1]} grid")
    return flat[:n].reshape(shape).copy()


def load_case(dirs_path: str, header_text: str, segment_limit: Optional[int] = None, name: str = "case",
              hours: float = 1.0, mismatched_maps: str = "reference") -> Case:
    """The case a DirsOfInputs file describes.  A map whose shape differs
    from the topography's (TanChau's viscosity map is 335 x 465 against
    324 x 464) is, with mismatched_maps="reference", taken as the reference's
    device reads it (device_view, with a warning); "error" raises instead."""
    pdict = parse_case_header(header_text)
    SL = int(segment_limit or pdict["segment_limit"])
    nangday = float(pdict["NANGDAY"])
    D = read_dirs(dirs_path)
    bt = np.zeros((4, SL), np.int32)
    with open(D["boundary_type"]) as f:
        for i, line in enumerate(f):
            inp = line.strip()
            if i < 4 and inp and inp != "NA":
                vals = [int(v) for v in inp.split()]
                bt[i, :len(vals)] = vals
    series = {}
    for key, name_ in (("up", "bc_up"), ("down", "bc_down"), ("left", "bc_left"), ("right", "bc_right"),
                       ("CC_u", "CC_u"), ("CC_d", "CC_d"), ("CC_l", "CC_l"), ("CC_r", "CC_r")):
        series[name_] = np.ascontiguousarray(np.transpose(np.loadtxt(D[key])), dtype=np.float64)
    blist = [series["bc_up"], series["bc_down"], series["bc_left"], series["bc_right"]]
    e = int(np.where(bt != 0)[0][0])
    total_time = len(blist[e]) // int(np.count_nonzero(bt[e]))
    if blist[e].ndim == 1:
        total_time = blist[e].size // int(np.count_nonzero(bt[e]))
    h = _map(D["topography"], nangday)
    N, M = h.shape[0] - 3, h.shape[1] - 3
    a = empty_state(N, M, SL)
    a["h"] = h
    for key, field in (("roughness_map", "hsnham"), ("visco_map", "VISCOINDX"), ("friction_map", "Fw")):
        m = _map(D[key])
        if m.shape != h.shape:
            msg = (f"{key} has shape {m.shape[0] - 3}x{m.shape[1] - 3} after the transpose, the topography "
                   f"{N}x{M}")
            if mismatched_maps != "reference":
                raise ValueError(msg)
            warnings.warn(msg + "; taking the values the reference's kernels read (its first "
                                f"{(N + 3) * (M + 3)} values, row-major)")
            m = device_view(m, h.shape)
        a[field] = m
    a["boundary_type"] = bt
    bienQ = np.zeros(4, np.int32)
    bienQ[np.where(bt == 2)[0]] = 1
    a["bienQ"] = bienQ
    n = max(v.size for v in series.values())
    for k, v in series.items():
        flat = v.reshape(-1)
        out = np.zeros(n)
        out[:flat.size] = flat
        a[k] = out
    return Case(name=name, N=N, M=M, header=header_text, arrays=a, total_time=total_time,
                Tmax=hours * 3600.0)
