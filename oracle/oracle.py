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
NATIVE_PATH = os.path.join(_HERE, "_build", "liboracle_native.so")
_lib = None
_path = LIB_PATH


def use_native_build() -> str:
    """Build (make native: -O3 -march=native for this host) and select the
    host-tuned checker for the CPU baseline; must precede the first use of
    the checker in the process.  Returns the library path in use."""
    global _path
    if _lib is not None:
        raise RuntimeError("the checker library is already loaded")
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "native"], check=True)
    _path = NATIVE_PATH
    return _path


def build(force: bool = False) -> str:
    """Compile the checker with its Makefile (g++, FP contraction off)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if _path == LIB_PATH:
            build()
        L = ctypes.CDLL(_path)
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
        L.or_kernel.argtypes = [vp, cp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        d = ctypes.c_double
        L.or_source_chs.restype = d
        L.or_source_chs.argtypes = [vp, d, d, d, d, d, d]
        L.or_source_nchs.restype = d
        L.or_source_nchs.argtypes = [vp, d, d, d, d, d]
        L.or_plant.restype = ctypes.c_int
        L.or_plant.argtypes = [ctypes.c_int]
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

    def kernel(self, name, a=0, b=0, c=0, t=0.0):
        """Run one restated kernel on the current state (or_kernel)."""
        self._check(lib().or_kernel(self.h, name.encode(), int(a), int(b), int(c), float(t)))

    def scratch(self, name):
        """A per-line assembly scratch array (a1..f5: (L+3)^2 doubles;
        AA..DD, x: (L+3)(2L+4) doubles, L = max(N, M))."""
        L = max(self.N, self.M)
        n = (L + 3) * (2 * L + 4) if name in ("AA", "BB", "CC", "DD", "x") else (L + 3) * (L + 3)
        return self.get(name, like=np.zeros(n))

    def source_chs(self, wsm, Zf, fs, fw, vth, h_moi):
        return float(lib().or_source_chs(self.h, wsm, Zf, fs, fw, vth, h_moi))

    def source_nchs(self, wsm, Zf, Uf, h_moi, fs):
        return float(lib().or_source_nchs(self.h, wsm, Zf, Uf, h_moi, fs))


class planted:
    """Context manager: plant mistake `code` (hydist_oracle.cpp enum Plant)
    in the checker's kernels for the duration of a block, so a test can show
    that its pin fails on that mistake."""

    def __init__(self, code):
        self.code = code

    def __enter__(self):
        self.prev = lib().or_plant(self.code)
        return self

    def __exit__(self, *exc):
        lib().or_plant(self.prev)
        return False


PLANTS = dict(UZ_ADV_INDEX=1, VZ_CORIOLIS_SIGN=2, US_ZGRAD_INDEX=3, VS_VISC_DROP=4, NORM_COUNT=5,
              WET_OPENS_LAND=6, QB_AVG_INDEX=7, BED_DIV_SIGN=8, CHS_EROSION=9, NCHS_GAMAC=10,
              HESOK_DEPTH=11, VTH_AVG=12, MARGIN_EXTRAP=13, UZ_FRICTION=14, UZ_D2_INDEX=15,
              ABCD_C1_SIGN=16, WET_NO_CASCADE=17, DRY_FACE=18, F4_FACTOR=19)


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
