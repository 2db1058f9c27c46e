"""Host/device array registry with the reference's ``Pointers`` interface.

Pointers.py:15-166 keeps host numpy arrays in ``host_ptrs`` and PyCUDA
allocations in ``arg_list`` / ``device_only_ptrs``, and moves data with
``toDevice`` / ``toHost`` / ``extract`` / ``update``.  Here the device memory is
owned by libhydist (one allocation per field, exact sizes), so ``alloc`` binds
the registry to a library context and ``arg_list`` maps each key to a
:class:`DeviceField` handle.  Method names, keyword arguments, dict names and
the dtype conversions (float64 -> dtype, int64 -> int32; Pointers.py:21-35)
are kept; ``PointersStruct`` returns the device addresses of the
Argument_Pointers / Array_Pointers mirrors the library builds
(Pointers.py:169-191).
"""
from __future__ import annotations

import numpy as np

INT_SIZE = 4
DOUBLE_SIZE = 8
segment_limit = 20

# device-only scratch names of Pointers.alloc_on_device_only (Pointers.py:55-79);
# the library keeps the coefficients in registers, so only these exist
_DEVICE_ONLY = ("x", "AA", "BB", "CC", "DD", "Ap", "Bp", "ep", "SN",
                "a1", "b1", "c1", "d1", "a2", "c2", "d2", "f1", "f2", "f3", "f5")


class DeviceField:
    """Handle of one named device array of a library context."""

    def __init__(self, ctx, name, nbytes):
        self.ctx, self.name, self.nbytes = ctx, name, nbytes

    def __repr__(self):
        return f"DeviceField({self.name!r}, {self.nbytes} B)"


def _convert(value, dtype):
    value = np.asarray(value)
    if value.dtype == np.float64 and dtype != np.float64:
        value = value.astype(dtype)
    if value.dtype == np.int64 or value.dtype == np.dtype(int):
        value = value.astype(np.int32)
    return value


class Pointers:
    # store pointers in host and corresponding pointers to device
    def __init__(self, ctx, dtype=np.float32, **kwargs):
        self.arg_list = {}          # device handles
        self.host_ptrs = {}         # host arrays
        self.device_only_ptrs = {}
        self.dtype = dtype
        if dtype != np.float64:
            raise ValueError("the ADI step is FP64 only (Engine.py:26); pass dtype=np.float64")
        for key, value in kwargs.items():
            self.host_ptrs[key] = _convert(value, dtype)
        self.ctx = ctx

    def _solver(self):
        return self.ctx.solver

    def alloc(self):
        """Bind every host key to the matching library field."""
        s = self._solver()
        for key in self.host_ptrs:
            name = "VISCOIDX" if key == "VISCOINDX" else key
            nb, _ = s.field_info(name)
            self.arg_list[key] = DeviceField(self.ctx, name, nb)
        return self.arg_list

    def alloc_on_device_only(self, m, n):
        """The reference allocates Thomas / coefficient scratch here
        (Pointers.py:47-81); the library owns its own scratch, sized at
        context creation, so this only records the names."""
        for k in _DEVICE_ONLY:
            self.device_only_ptrs[k] = None
        return self.device_only_ptrs

    def add_host_pts(self, **kwargs):
        for key, value in kwargs.items():
            self.host_ptrs[key] = _convert(value, self.dtype)
            name = "VISCOIDX" if key == "VISCOINDX" else key
            nb, _ = self._solver().field_info(name)
            self.arg_list[key] = DeviceField(self.ctx, name, nb)

    def add_device_only_pts(self, *args):
        for a in args:
            self.device_only_ptrs[a] = None

    def update(self, keyList=None):
        self.toDevice(keyList)

    def toDevice(self, copyList=None):
        keys = list(self.arg_list.keys()) if copyList is None else copyList
        for key in keys:
            self._solver().set(self.arg_list[key].name, self.host_ptrs[key])
        self.ctx.synchronize()

    def toHost(self, keyList=None, all=False):
        keys = list(self.host_ptrs.keys()) if keyList is None else keyList
        for key in keys:
            dst = self.host_ptrs[key]
            self.host_ptrs[key][...] = self._solver().get(self.arg_list[key].name, like=dst).reshape(dst.shape)
        self.ctx.synchronize()

    def extract(self, argmap):
        for key, value in argmap.items():
            if key in self.arg_list:
                value[...] = self._solver().get(self.arg_list[key].name, like=value).reshape(value.shape)

    def debug(self, keyMap, direction="VZ"):
        """Pointers.py:158-162 copies the device-only Thomas scratch.  Here the
        coefficients live in registers and LDS during a sweep, so the rows are
        assembled on request from the current state by the library's row
        kernel (hd_k_UZ/VZSolver_calculate_matrix_coeff): keys AA, BB, CC, DD
        and SN, in the reference's layout (line * (2M+1) + row for VZ, the
        direction of the last assembly of a step without sediment).  The
        reference's buffers instead hold the rows of the last half-step's
        starting state.  Other scratch names (x, Ap, Bp, ep, a1..f5) have no
        device counterpart and raise KeyError."""
        s = self._solver()
        name = ("UZ" if direction == "UZ" else "VZ") + "Solver_calculate_matrix_coeff"
        AA, BB, CC, DD, SN = s.rows(name, 0)
        got = {"AA": AA, "BB": BB, "CC": CC, "DD": DD, "SN": SN}
        for key, des in keyMap.items():
            if key not in got:
                raise KeyError(f"{key}: no device counterpart of this scratch array")
            src = got[key]
            n = min(des.size, src.size)
            des.reshape(-1)[:n] = src[:n]


class PointersStruct:
    """Pointers.py:169-191 packs the device structs by hand; the library
    builds identical Argument_Pointers / Array_Pointers mirrors, whose device
    addresses this object carries (``.ptr``)."""

    def __init__(self, ptrList, struct_ptr=None, structtype="ARG", ctx=None):
        solver = ctx.solver if ctx is not None else getattr(struct_ptr, "solver", None)
        if solver is None:
            raise ValueError("PointersStruct needs the library context (ctx=...)")
        arg, arr = solver.struct_ptrs()
        self.ptr = arg if structtype == "ARG" else arr
