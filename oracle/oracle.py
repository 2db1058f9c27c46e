"""ctypes binding of the CPU checker (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the hydist_amd package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG = os.path.join(os.path.dirname(_HERE), "gpu-hydist_amd")
if _PKG not in sys.path:
    sys.path.insert(0, _PKG)

from hydist_amd.case import HdParams, HdStepOpts  # noqa: E402

LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile the checker with its Makefile (g++, FP contraction off)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(LIB_PATH)
        vp, cp, sz = ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t
        L.or_create.restype = vp
        L.or_create.argtypes = [ctypes.POINTER(HdParams), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_destroy.argtypes = [vp]
        L.or_last_error.restype = cp
        L.or_last_error.argtypes = [vp]
        L.or_get_params.argtypes = [vp, ctypes.POINTER(HdParams)]
        L.or_set.argtypes = [vp, cp, vp, sz]
        L.or_get.argtypes = [vp, cp, vp, sz]
        L.or_get_hmax.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
        L.or_init.argtypes = [vp]
        L.or_step.argtypes = [vp, ctypes.c_double, ctypes.POINTER(HdStepOpts), ctypes.POINTER(ctypes.c_double)]
        L.or_run.argtypes = [vp, ctypes.c_double, ctypes.c_int, ctypes.POINTER(HdStepOpts),
                             ctypes.POINTER(ctypes.c_double)]
        L.or_halfstep.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.POINTER(HdStepOpts)]
        L.or_bed_change.argtypes = [vp, ctypes.c_double, ctypes.POINTER(HdStepOpts), ctypes.c_int, ctypes.c_int]
        L.or_find_limits.argtypes = [vp]
        L.or_wet_count.restype = ctypes.c_longlong
        L.or_wet_count.argtypes = [vp]
        dp = ctypes.POINTER(ctypes.c_double)
        L.or_thomas.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp]
        ip = ctypes.POINTER(ctypes.c_int)
        L.or_scan_line.argtypes = [ip, ctypes.c_int, ctypes.c_int, ip, ip]
        for fn in ("or_hd_log", "or_hd_exp", "or_hd_cos", "or_hd_sin"):
            getattr(L, fn).restype = ctypes.c_double
            getattr(L, fn).argtypes = [ctypes.c_double]
        L.or_set_threads.restype = ctypes.c_int
        L.or_set_threads.argtypes = [ctypes.c_int]
        L.or_hd_pow.restype = ctypes.c_double
        L.or_hd_pow.argtypes = [ctypes.c_double, ctypes.c_double]
        _lib = L
    return _lib


def set_threads(n: int) -> int:
    """OpenMP threads of the checker's loops over independent lines / cells;
    returns the count in effect (results are identical for any count)."""
    return int(lib().or_set_threads(int(n)))


class OracleError(RuntimeError):
    pass


class Oracle:
    """The CPU restatement of the reference step, driven like libhydist."""

    def __init__(self, case, upload_keys=None):
        L = lib()
        self.case = case
        self.N, self.M = case.N, case.M
        self._p = case.params
        self.h = L.or_create(ctypes.byref(self._p), case.N, case.M, case.series_len, case.total_time)
        if not self.h:
            raise OracleError("or_create failed")
        for k, arr in case.arrays.items():
            self.set(k, arr)

    def close(self):
        if self.h:
            lib().or_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise OracleError(f"oracle error {rc}: {lib().or_last_error(self.h).decode()}")

    def set(self, name, arr):
        a = np.ascontiguousarray(arr)
        self._check(lib().or_set(self.h, name.encode(), a.ctypes.data, a.nbytes))

    def get(self, name, like=None):
        c = self.case
        if like is None:
            like = c.arrays[name] if name in c.arrays else c.arrays["VISCOINDX" if name == "VISCOIDX" else "u"]
        out = np.empty_like(like)
        self._check(lib().or_get(self.h, name.encode(), out.ctypes.data, out.nbytes))
        return out

    def params(self) -> HdParams:
        p = HdParams()
        lib().or_get_params(self.h, ctypes.byref(p))
        return p

    def hmax(self):
        out = (ctypes.c_double * 4)()
        lib().or_get_hmax(self.h, out)
        return list(out)

    def init(self):
        self._check(lib().or_init(self.h))

    def step(self, t, opts):
        tout = ctypes.c_double()
        self._check(lib().or_step(self.h, t, ctypes.byref(opts), ctypes.byref(tout)))
        return tout.value

    def run(self, t, nsteps, opts):
        tout = ctypes.c_double()
        self._check(lib().or_run(self.h, t, nsteps, ctypes.byref(opts), ctypes.byref(tout)))
        return tout.value

    def halfstep(self, which, t, opts):
        self._check(lib().or_halfstep(self.h, which, t, ctypes.byref(opts)))

    def find_limits(self):
        self._check(lib().or_find_limits(self.h))

    def wet_count(self):
        return int(lib().or_wet_count(self.h))


def thomas(A, B, C, D):
    """Reference Thomas (UVZSolver_multithread.cu:203-228) on one system."""
    A, B, C, D = (np.ascontiguousarray(a, dtype=np.float64) for a in (A, B, C, D))
    n = A.size
    x = np.zeros(n)
    dp = ctypes.POINTER(ctypes.c_double)
    lib().or_thomas(n - 1, A.ctypes.data_as(dp), B.ctypes.data_as(dp), C.ctypes.data_as(dp),
                    D.ctypes.data_as(dp), x.ctypes.data_as(dp))
    return x


def scan_line(kh, lim, SL=20):
    """Reference segment scan of one line (support_funcs.cu:252-279)."""
    k = np.ascontiguousarray(kh, dtype=np.int32)
    dau = np.zeros(SL, np.int32)
    cuoi = np.zeros(SL, np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    n = lib().or_scan_line(k.ctypes.data_as(ip), lim, SL, dau.ctypes.data_as(ip), cuoi.ctypes.data_as(ip))
    return n, dau, cuoi
