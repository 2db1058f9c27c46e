"""ctypes binding of libhydist.so (include/hydist.h).

The library is built in-tree (hydist_amd/_build/libhydist.so) by
``hydist_amd.build.build()``.  There is no fallback: if the shared library is
missing or no HIP device is present, every call raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .case import HdParams, HdStepOpts

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libhydist.so")

# every entry point declared in include/hydist.h
EXPORTS = (
    "hd_version", "hd_error_string", "hd_create", "hd_destroy", "hd_last_error", "hd_get_params",
    "hd_field_info", "hd_upload", "hd_download", "hd_get_hmax", "hd_init", "hd_prepare", "hd_step", "hd_run",
    "hd_halfstep", "hd_bed_change", "hd_run_counted", "hd_reset_wet_counter", "hd_wet_cell_steps",
    "hd_wet_count", "hd_sync", "hd_struct_ptrs", "hd_profile", "hd_kernel_times", "hd_thomas_batch",
    "hd_thomas_batch_spec",
    "hd_ensemble_create", "hd_ensemble_destroy", "hd_ensemble_last_error", "hd_ensemble_run",
    "hd_ensemble_profile", "hd_ensemble_kernel_times", "hd_flow_refresh", "hd_sediment_run",
    "hd_clone", "hd_autotune", "hd_set_launch", "hd_bed_snapshot", "hd_bed_extrapolate", "hd_copy_state",
    "hd_z_snapshot", "hd_z_change", "hd_z_check", "hd_thomas_batch_noalloc",
    "hd_k_Onetime_init", "hd_k_Find_Calculation_limits", "hd_k_Htuongdoi", "hd_k_preprocess_data",
    "hd_k_Update_Boundary_Value", "hd_k_UZSolver", "hd_k_VZSolver", "hd_k_UZSolver_calculate_matrix_coeff",
    "hd_k_VZSolver_calculate_matrix_coeff", "hd_k_Normalize", "hd_k_solveV", "hd_k_solveU", "hd_k_Reset_states",
    "hd_k_Find_VTH_hesoK", "hd_k_FSj_solve", "hd_k_FSi_solve", "hd_k_Scan_FSj", "hd_k_Scan_FSi",
    "hd_k_Calculate_Qb", "hd_k_BedLoad",
)

_lib = None


class HydistError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libhydist error {code}: {msg}")
        self.code = code


def load() -> ctypes.CDLL:
    """Load libhydist.so and declare its signatures (no device call)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is not built; run hydist_amd.build.build()")
    L = ctypes.CDLL(LIB_PATH)
    vp, cp, sz, ci, cd = ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_double
    pp, po = ctypes.POINTER(HdParams), ctypes.POINTER(HdStepOpts)
    dp, ip, ll = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_longlong)
    sig = {
        "hd_version": (cp, []),
        "hd_error_string": (cp, [ci]),
        "hd_create": (ci, [ci, pp, ci, ci, ci, ci, ctypes.POINTER(vp)]),
        "hd_destroy": (ci, [vp]),
        "hd_last_error": (cp, [vp]),
        "hd_get_params": (ci, [vp, pp]),
        "hd_field_info": (ci, [vp, cp, ctypes.POINTER(sz), ip]),
        "hd_upload": (ci, [vp, cp, vp, sz]),
        "hd_download": (ci, [vp, cp, vp, sz]),
        "hd_get_hmax": (ci, [vp, dp]),
        "hd_init": (ci, [vp]),
        "hd_prepare": (ci, [vp]),
        "hd_step": (ci, [vp, cd, po, dp]),
        "hd_run": (ci, [vp, cd, ci, po, dp]),
        "hd_halfstep": (ci, [vp, ci, cd, po]),
        "hd_bed_change": (ci, [vp, cd, po]),
        "hd_run_counted": (ci, [vp, cd, ci, po, dp]),
        "hd_reset_wet_counter": (ci, [vp]),
        "hd_wet_cell_steps": (ci, [vp, ll]),
        "hd_wet_count": (ci, [vp, ll]),
        "hd_sync": (ci, [vp]),
        "hd_struct_ptrs": (ci, [vp, ctypes.POINTER(vp), ctypes.POINTER(vp)]),
        "hd_profile": (ci, [vp, ci]),
        "hd_kernel_times": (ci, [vp, ci, ctypes.POINTER(cp), dp, ip]),
        "hd_thomas_batch": (ci, [ci, ci, ip, ip, dp, dp, dp, dp, dp, ci]),
        "hd_thomas_batch_spec": (ci, [ci, ci, ip, ip, dp, dp, dp, dp, dp, ci]),
        "hd_thomas_batch_noalloc": (ci, [ci, ci, vp, vp, vp, vp, vp, vp, vp, ci, vp]),
        "hd_k_Onetime_init": (ci, [vp]),
        "hd_k_Find_Calculation_limits": (ci, [vp]),
        "hd_k_Htuongdoi": (ci, [vp]),
        "hd_k_preprocess_data": (ci, [vp]),
        "hd_k_Update_Boundary_Value": (ci, [vp, cd, ci]),
        "hd_k_UZSolver": (ci, [vp, ci]),
        "hd_k_VZSolver": (ci, [vp, ci]),
        "hd_k_UZSolver_calculate_matrix_coeff": (ci, [vp, ci, vp, vp, vp, vp, vp]),
        "hd_k_VZSolver_calculate_matrix_coeff": (ci, [vp, ci, vp, vp, vp, vp, vp]),
        "hd_k_Normalize": (ci, [vp, ci]),
        "hd_k_solveV": (ci, [vp]),
        "hd_k_solveU": (ci, [vp]),
        "hd_k_Reset_states": (ci, [vp, ci, ci, ci]),
        "hd_k_Find_VTH_hesoK": (ci, [vp]),
        "hd_k_FSj_solve": (ci, [vp, cd, cd, ci]),
        "hd_k_FSi_solve": (ci, [vp, cd, cd, ci]),
        "hd_k_Scan_FSj": (ci, [vp, cd, cd, ci, vp, vp, vp, vp]),
        "hd_k_Scan_FSi": (ci, [vp, cd, cd, ci, vp, vp, vp, vp]),
        "hd_k_Calculate_Qb": (ci, [vp, ci]),
        "hd_k_BedLoad": (ci, [vp, ci, ci, ci]),
        "hd_flow_refresh": (ci, [vp]),
        "hd_clone": (ci, [vp, ctypes.POINTER(vp)]),
        "hd_autotune": (ci, [vp, cd, po, ci, ip, ip, dp]),
        "hd_set_launch": (ci, [vp, ci, ci]),
        "hd_bed_snapshot": (ci, [vp]),
        "hd_copy_state": (ci, [vp, vp]),
        "hd_z_snapshot": (ci, [vp]),
        "hd_z_change": (ci, [vp, dp, dp, dp]),
        "hd_z_check": (ci, [vp, cd, dp, dp, dp, dp]),
        "hd_bed_extrapolate": (ci, [vp, cd]),
        "hd_sediment_run": (ci, [vp, cd, ci, po, dp]),
        "hd_ensemble_create": (ci, [ctypes.POINTER(vp), ci, ctypes.POINTER(vp)]),
        "hd_ensemble_destroy": (ci, [vp]),
        "hd_ensemble_last_error": (cp, [vp]),
        "hd_ensemble_run": (ci, [vp, cd, ci, po, ci, dp]),
        "hd_ensemble_profile": (ci, [vp, ci]),
        "hd_ensemble_kernel_times": (ci, [vp, ci, ctypes.POINTER(cp), dp, ip]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc: int, ctx=None):
    if rc != 0:
        L = load()
        msg = L.hd_last_error(ctx).decode()
        raise HydistError(rc, msg or L.hd_error_string(rc).decode())


class Solver:
    """One case on one GPU: the device state of the reference's Pointers
    (rep.py:62-111) plus the step of Engine.hydraulic_Calculation."""

    def __init__(self, case=None, *, params: HdParams = None, N: int = None, M: int = None,
                 series_len: int = None, total_time: int = None, device: int = 0, arrays=None):
        L = load()
        if case is not None:
            params, N, M = case.params, case.N, case.M
            series_len, total_time = case.series_len, case.total_time
            arrays = case.arrays if arrays is None else arrays
        self.N, self.M = N, M
        self._params = params
        self.h = ctypes.c_void_p()
        check(L.hd_create(device, ctypes.byref(params), N, M, series_len, total_time, ctypes.byref(self.h)))
        self.case = case
        if arrays:
            for k, a in arrays.items():
                self.set(k, a)

    def close(self):
        if self.h:
            load().hd_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        check(rc, self.h)

    def field_info(self, name):
        nb, dt = ctypes.c_size_t(), ctypes.c_int()
        self._chk(load().hd_field_info(self.h, name.encode(), ctypes.byref(nb), ctypes.byref(dt)))
        return nb.value, (np.float64 if dt.value == 0 else np.int32)

    def set(self, name, arr):
        nb, dt = self.field_info(name)
        a = np.ascontiguousarray(arr, dtype=dt)
        self._chk(load().hd_upload(self.h, name.encode(), a.ctypes.data, a.nbytes))

    def get(self, name, like=None):
        nb, dt = self.field_info(name)
        if like is not None:
            out = np.empty(like.shape, dtype=dt)
        elif self.case is not None and name in self.case.arrays:
            out = np.empty(self.case.arrays[name].shape, dtype=dt)
        else:
            out = np.empty(nb // np.dtype(dt).itemsize, dtype=dt)
            if out.size == (self.N + 3) * (self.M + 3):
                out = out.reshape(self.N + 3, self.M + 3)
        self._chk(load().hd_download(self.h, name.encode(), out.ctypes.data, out.nbytes))
        return out

    def params(self) -> HdParams:
        p = HdParams()
        self._chk(load().hd_get_params(self.h, ctypes.byref(p)))
        return p

    def hmax(self):
        out = (ctypes.c_double * 4)()
        self._chk(load().hd_get_hmax(self.h, out))
        return list(out)

    def init(self):
        self._chk(load().hd_init(self.h))

    def prepare(self):
        self._chk(load().hd_prepare(self.h))

    def step(self, t, opts):
        tout = ctypes.c_double()
        self._chk(load().hd_step(self.h, t, ctypes.byref(opts), ctypes.byref(tout)))
        return tout.value

    def run(self, t, nsteps, opts, counted=False):
        tout = ctypes.c_double()
        fn = load().hd_run_counted if counted else load().hd_run
        self._chk(fn(self.h, t, nsteps, ctypes.byref(opts), ctypes.byref(tout)))
        return tout.value

    def halfstep(self, which, t, opts):
        self._chk(load().hd_halfstep(self.h, which, t, ctypes.byref(opts)))

    def bed_change(self, t, opts):
        self._chk(load().hd_bed_change(self.h, t, ctypes.byref(opts)))

    def sync(self):
        self._chk(load().hd_sync(self.h))

    def kernel(self, name, *args):
        """One reference kernel on the current state: hd_k_<name>(ctx, *args)
        (per-kernel entry points of include/hydist.h, for unit parity)."""
        self._chk(getattr(load(), "hd_k_" + name)(self.h, *args))

    def rows(self, name, *args):
        """hd_k_<name> of a row-assembly kernel (UZSolver_/VZSolver_
        calculate_matrix_coeff, Scan_FSj, Scan_FSi): its AA, BB, CC, DD in
        the reference's layout (only the systems' rows are defined) and, for
        the momentum/continuity rows, SN."""
        N, M, L = self.N, self.M, max(self.N, self.M)
        if name.startswith(("UZ", "VZ")):
            n = (M + 3) * (2 * N + 1) if name.startswith("UZ") else (N + 3) * (2 * M + 1)
        else:
            n = (M + 3) * (N + 3)
        out = [np.full(n, np.nan) for _ in range(4)]
        ptrs = [a.ctypes.data for a in out]
        if name.startswith(("UZ", "VZ")):
            sn = np.zeros((L + 3) * self.params().segment_limit, np.int32)
            self.kernel(name, *args, *ptrs, sn.ctypes.data)
            return out + [sn]
        self.kernel(name, *args, *ptrs)
        return out

    def clone(self):
        """A second Solver holding a copy of this one's device state."""
        h = ctypes.c_void_p()
        self._chk(load().hd_clone(self.h, ctypes.byref(h)))
        other = Solver.__new__(Solver)
        other.N, other.M, other._params, other.case = self.N, self.M, self._params, self.case
        other.h = h
        return other

    def z_snapshot(self):
        self._chk(load().hd_z_snapshot(self.h))

    def z_change(self):
        """(max |z - z_snapshot| on wet cells, wet volume now, at the snapshot)."""
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._chk(load().hd_z_change(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def z_check(self, bound):
        """(max |z - z_snapshot|, fraction of wet cells above `bound`, wet
        volume now, at the snapshot)."""
        a, f, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._chk(load().hd_z_check(self.h, bound, ctypes.byref(a), ctypes.byref(f), ctypes.byref(b),
                                    ctypes.byref(c)))
        return a.value, f.value, b.value, c.value

    def copy_state_from(self, other):
        """Overwrite this context's device state with other's (same case)."""
        self._chk(load().hd_copy_state(self.h, other.h))

    def autotune(self, t, opts, steps=10):
        """Pick the launch geometry on a clone; returns (rb_mult, bw, ms/step)."""
        rb, bw, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        self._chk(load().hd_autotune(self.h, t, ctypes.byref(opts), steps, ctypes.byref(rb), ctypes.byref(bw),
                                     ctypes.byref(ms)))
        return rb.value, bw.value, ms.value

    def set_launch(self, rb_mult, bw):
        self._chk(load().hd_set_launch(self.h, rb_mult, bw))

    def bed_snapshot(self):
        self._chk(load().hd_bed_snapshot(self.h))

    def bed_extrapolate(self, factor):
        self._chk(load().hd_bed_extrapolate(self.h, float(factor)))

    def flow_refresh(self):
        self._chk(load().hd_flow_refresh(self.h))

    def sediment_run(self, t, nsteps, opts):
        tout = ctypes.c_double()
        self._chk(load().hd_sediment_run(self.h, t, nsteps, ctypes.byref(opts), ctypes.byref(tout)))
        return tout.value

    def wet_count(self):
        n = ctypes.c_longlong()
        self._chk(load().hd_wet_count(self.h, ctypes.byref(n)))
        return n.value

    def reset_wet_counter(self):
        self._chk(load().hd_reset_wet_counter(self.h))

    def wet_cell_steps(self):
        n = ctypes.c_longlong()
        self._chk(load().hd_wet_cell_steps(self.h, ctypes.byref(n)))
        return n.value

    def profile(self, enable=True):
        self._chk(load().hd_profile(self.h, int(enable)))

    def kernel_times(self, n=64):
        names = (ctypes.c_char_p * n)()
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_int * n)()
        k = load().hd_kernel_times(self.h, n, names, ms, cnt)
        if k < 0:
            self._chk(k)
        return {names[i].decode(): (ms[i], cnt[i]) for i in range(k)}

    def struct_ptrs(self):
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        self._chk(load().hd_struct_ptrs(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value


def thomas_batch(systems, device=0, spec=False):
    """Solve a list of (A, B, C, D) tridiagonal systems with the GPU Thomas
    kernel (spec: the run solvers' speculative parallel form,
    hd_thomas_batch_spec); returns the list of solutions."""
    L = load()
    off, sn, rows = [], [], 0
    for A, *_ in systems:
        off.append(rows)
        sn.append(len(A) - 1)
        rows += len(A)
    cat = [np.ascontiguousarray(np.concatenate([s[k] for s in systems]), dtype=np.float64) for k in range(4)]
    x = np.zeros(rows)
    off_a = np.asarray(off, np.int32)
    sn_a = np.asarray(sn, np.int32)
    dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    fn = L.hd_thomas_batch_spec if spec else L.hd_thomas_batch
    check(fn(device, len(systems), off_a.ctypes.data_as(ip), sn_a.ctypes.data_as(ip),
             *(c.ctypes.data_as(dp) for c in cat), x.ctypes.data_as(dp), rows))
    return [x[o:o + n + 1] for o, n in zip(off, sn)]


class Ensemble:
    """Independent cases stepped together on one device: every kernel of a
    step is one launch over all members (hd_ensemble_*).  Members are
    Solver objects, already initialised; read them back with Solver.get."""

    def __init__(self, members):
        L = load()
        self.members = list(members)
        arr = (ctypes.c_void_p * len(self.members))(*[m.h.value for m in self.members])
        self.h = ctypes.c_void_p()
        rc = L.hd_ensemble_create(arr, len(self.members), ctypes.byref(self.h))
        if rc != 0:
            raise HydistError(rc, L.hd_ensemble_last_error(None).decode())

    def _chk(self, rc):
        if rc != 0:
            raise HydistError(rc, load().hd_ensemble_last_error(self.h).decode())

    def close(self):
        if self.h:
            load().hd_ensemble_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, t, nsteps, opts, counted=False):
        tout = ctypes.c_double()
        self._chk(load().hd_ensemble_run(self.h, t, nsteps, ctypes.byref(opts), 1 if counted else 0,
                                         ctypes.byref(tout)))
        return tout.value

    def wet_cell_steps(self):
        return sum(m.wet_cell_steps() for m in self.members)

    def reset_wet_counter(self):
        for m in self.members:
            m.reset_wet_counter()

    def profile(self, on=True):
        self._chk(load().hd_ensemble_profile(self.h, 1 if on else 0))

    def kernel_times(self):
        n = 64
        names = (ctypes.c_char_p * n)()
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_int * n)()
        k = load().hd_ensemble_kernel_times(self.h, n, names, ms, cnt)
        if k < 0:
            self._chk(k)
        return {names[i].decode(): (ms[i], cnt[i]) for i in range(k)}
