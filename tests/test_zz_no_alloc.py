"""hd_thomas_batch_noalloc (device inputs and outputs, factor scratch in a
static device array) against hd_thomas_batch (host buffers, allocates) on the
same systems: identical bits, on two calls in a row, the second smaller, and
the size limit reported as an error."""
import ctypes

import numpy as np
import pytest

from hydist_amd import _lib

pytestmark = pytest.mark.gpu


def systems(rng, nsys, max_len):
    """nsys diagonally dominant systems packed one after another (with a gap
    row between them), as the ADI half-steps build them."""
    sn = rng.integers(-1, max_len, size=nsys).astype(np.int32)
    off = np.zeros(nsys, np.int32)
    pos = 0
    for s in range(nsys):
        off[s] = pos
        pos += max(int(sn[s]), 0) + 2
    nrows = pos
    A = rng.normal(size=nrows); C = rng.normal(size=nrows)
    B = 3.0 + np.abs(A) + np.abs(C) + rng.random(nrows)
    D = rng.normal(size=nrows)
    return off, sn, A, B, C, D, nrows


def host_solve(off, sn, A, B, C, D, nrows):
    L = _lib.load()
    x = np.full(nrows, -7.0)
    ip, dp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    rc = L.hd_thomas_batch(0, off.size, off.ctypes.data_as(ip), sn.ctypes.data_as(ip), A.ctypes.data_as(dp),
                           B.ctypes.data_as(dp), C.ctypes.data_as(dp), D.ctypes.data_as(dp), x.ctypes.data_as(dp),
                           nrows)
    assert rc == 0, L.hd_last_error(None)
    return x


def hip():
    """The HIP runtime libhydist.so is linked against (already loaded in
    this process), for the test's own device buffers."""
    _lib.load()
    H = ctypes.CDLL("libamdhip64.so.7")
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    H.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
    H.hipFree.argtypes = [vp]
    H.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
    H.hipDeviceSynchronize.argtypes = []
    H.hipSetDevice.argtypes = [ctypes.c_int]
    return H


def device_solve(off, sn, A, B, C, D, nrows):
    L, H = _lib.load(), hip()
    assert H.hipSetDevice(0) == 0
    x0 = np.full(nrows, -7.0)
    host = dict(off=off, sn=sn, A=A, B=B, C=C, D=D, x=x0)
    dev = {}
    try:
        for k, v in host.items():
            p = ctypes.c_void_p()
            assert H.hipMalloc(ctypes.byref(p), v.nbytes) == 0
            dev[k] = p
            assert H.hipMemcpy(p, v.ctypes.data, v.nbytes, 1) == 0          # host to device
        rc = L.hd_thomas_batch_noalloc(0, off.size, dev["off"], dev["sn"], dev["A"], dev["B"], dev["C"], dev["D"],
                                       dev["x"], nrows, None)
        assert rc == 0, L.hd_last_error(None)
        assert H.hipDeviceSynchronize() == 0
        out = np.empty(nrows)
        assert H.hipMemcpy(out.ctypes.data, dev["x"], out.nbytes, 2) == 0   # device to host
        return out
    finally:
        for p in dev.values():
            H.hipFree(p)


def test_noalloc_matches_allocating_entry_point_twice():
    rng = np.random.default_rng(11)
    big = systems(rng, 700, 600)           # about 2e5 rows
    small = systems(rng, 40, 50)
    for sysset in (big, small):
        want = host_solve(*sysset)
        got = device_solve(*sysset)
        assert np.array_equal(got, want)


def test_noalloc_rejects_too_many_rows():
    L = _lib.load()
    rc = L.hd_thomas_batch_noalloc(0, 1, 1, 1, 1, 1, 1, 1, 1, (1 << 20) + 1, None)
    assert rc == -1
    assert b"exceeds the scratch limit" in L.hd_last_error(None)
