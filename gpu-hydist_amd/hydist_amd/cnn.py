"""ctypes binding of the flow surrogate's network (include/hydist_cnn.h).

The network lives in libhydist.so next to the solver; its kernels are the
project's own BF16 MFMA convolutions.  No fallback: every call raises when
the library or a GPU is missing.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

CIN, COUT = 8, 3

EXPORTS = (
    "hd_cnn_create", "hd_cnn_destroy", "hd_cnn_last_error", "hd_cnn_set_scales", "hd_cnn_layer_dims",
    "hd_cnn_get_layer", "hd_cnn_set_layer", "hd_cnn_get_grad", "hd_cnn_load_batch", "hd_cnn_forward",
    "hd_cnn_backward", "hd_cnn_adam", "hd_cnn_train_step", "hd_cnn_data_alloc", "hd_cnn_capture",
    "hd_cnn_train_samples", "hd_cnn_eval_samples", "hd_cnn_predict_into_ctx", "hd_cnn_mfma_check",
    "hd_cnn_set_sparse",
)


class Scales(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in ("s_u", "s_v", "s_z", "s_b", "s_q", "s_h", "d_u", "d_v", "d_z")]


_declared = False


def lib():
    global _declared
    L = _lib.load()
    if not _declared:
        vp, ci, cf, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_double
        fp, ip, dp = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
        sig = {
            "hd_cnn_create": (ci, [ci, ci, ci, ci, ci, ctypes.c_uint64, ctypes.POINTER(vp)]),
            "hd_cnn_destroy": (ci, [vp]),
            "hd_cnn_last_error": (ctypes.c_char_p, [vp]),
            "hd_cnn_set_scales": (ci, [vp, ctypes.POINTER(Scales)]),
            "hd_cnn_layer_dims": (ci, [vp, ci, ip, ip]),
            "hd_cnn_get_layer": (ci, [vp, ci, fp, fp]),
            "hd_cnn_set_layer": (ci, [vp, ci, fp, fp]),
            "hd_cnn_get_grad": (ci, [vp, ci, fp, fp]),
            "hd_cnn_load_batch": (ci, [vp, ci, fp, fp, fp]),
            "hd_cnn_forward": (ci, [vp, fp]),
            "hd_cnn_backward": (ci, [vp, dp]),
            "hd_cnn_adam": (ci, [vp, cf]),
            "hd_cnn_train_step": (ci, [vp, cf, dp]),
            "hd_cnn_data_alloc": (ci, [vp, ci]),
            "hd_cnn_capture": (ci, [vp, ci, ci, vp, cf, cf]),
            "hd_cnn_train_samples": (ci, [vp, ip, ci, cf, dp]),
            "hd_cnn_eval_samples": (ci, [vp, ip, ci, dp]),
            "hd_cnn_predict_into_ctx": (ci, [vp, vp, cf, cf]),
            "hd_cnn_mfma_check": (ci, [ci, fp, fp, fp]),
            "hd_cnn_set_sparse": (ci, [vp, ci]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _declared = True
    return L


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def mfma_check(A, B, device=0):
    """D = A @ B through one BF16 MFMA tile on the GPU (A 16x32, B 32x16)."""
    a, pa = _f32(A)
    b, pb = _f32(B)
    d = np.zeros((16, 16), np.float32)
    rc = lib().hd_cnn_mfma_check(device, pa, pb, d.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    if rc:
        raise _lib.HydistError(rc, lib().hd_cnn_last_error(None).decode())
    return d


class Network:
    """The residual CNN on an H x W grid (device-resident weights, Adam
    state, activations for `batch` images and an optional sample store)."""

    def __init__(self, H, W, batch=1, layers=8, seed=0, device=0):
        L = lib()
        self.H, self.W, self.batch, self.layers = H, W, batch, layers
        self.h = ctypes.c_void_p()
        rc = L.hd_cnn_create(device, H, W, batch, layers, seed, ctypes.byref(self.h))
        if rc:
            raise _lib.HydistError(rc, L.hd_cnn_last_error(None).decode())

    def _chk(self, rc):
        if rc:
            raise _lib.HydistError(rc, lib().hd_cnn_last_error(self.h).decode())

    def close(self):
        if self.h:
            lib().hd_cnn_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_scales(self, **kw):
        kw = {"d_u": 1.0, "d_v": 1.0, "d_z": 1.0, **kw}
        s = Scales(**{k: float(v) for k, v in kw.items()})
        self._chk(lib().hd_cnn_set_scales(self.h, ctypes.byref(s)))

    def dims(self, l):
        ci, co = ctypes.c_int(), ctypes.c_int()
        self._chk(lib().hd_cnn_layer_dims(self.h, l, ctypes.byref(ci), ctypes.byref(co)))
        return ci.value, co.value

    def get_layer(self, l):
        ci, co = self.dims(l)
        w = np.zeros((9, co, ci), np.float32)
        b = np.zeros(co, np.float32)
        self._chk(lib().hd_cnn_get_layer(self.h, l, _f32(w)[1], _f32(b)[1]))
        return w, b

    def set_layer(self, l, w, b):
        w, pw = _f32(w)
        b, pb = _f32(b)
        self._chk(lib().hd_cnn_set_layer(self.h, l, pw, pb))

    def get_grad(self, l):
        ci, co = self.dims(l)
        w = np.zeros((9, co, ci), np.float32)
        b = np.zeros(co, np.float32)
        self._chk(lib().hd_cnn_get_grad(self.h, l, _f32(w)[1], _f32(b)[1]))
        return w, b

    def load(self, x8, tgt=None, msk=None):
        """Host images: x8 [b,H,W,8], tgt [b,H,W,3], msk [b,H,W] (FP32)."""
        x, px = _f32(x8)
        py = pm = None
        if tgt is not None:
            t, py = _f32(tgt)
        if msk is not None:
            m, pm = _f32(msk)
        self._chk(lib().hd_cnn_load_batch(self.h, x.shape[0], px, py, pm))
        self._n = x.shape[0]

    def forward(self):
        """Prediction of the loaded batch, [b,H,W,3] (normalised u, v, z)."""
        out = np.zeros((self._n, self.H, self.W, 3), np.float32)
        self._chk(lib().hd_cnn_forward(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def backward(self):
        loss = ctypes.c_double()
        self._chk(lib().hd_cnn_backward(self.h, ctypes.byref(loss)))
        return loss.value

    def adam(self, lr):
        self._chk(lib().hd_cnn_adam(self.h, lr))

    def train_step(self, lr):
        loss = ctypes.c_double()
        self._chk(lib().hd_cnn_train_step(self.h, lr, ctypes.byref(loss)))
        return loss.value

    # sample store ---------------------------------------------------------
    def data_alloc(self, count):
        self._chk(lib().hd_cnn_data_alloc(self.h, count))

    def capture(self, idx, role, solver, q=0.0, zb=0.0):
        self._chk(lib().hd_cnn_capture(self.h, idx, role, solver.h, q, zb))

    def train_samples(self, idx, lr):
        a = (ctypes.c_int * len(idx))(*idx)
        loss = ctypes.c_double()
        self._chk(lib().hd_cnn_train_samples(self.h, a, len(idx), lr, ctypes.byref(loss)))
        return loss.value

    def eval_samples(self, idx):
        a = (ctypes.c_int * len(idx))(*idx)
        loss = ctypes.c_double()
        self._chk(lib().hd_cnn_eval_samples(self.h, a, len(idx), ctypes.byref(loss)))
        return loss.value

    def predict_into(self, solver, q=0.0, zb=0.0):
        self._chk(lib().hd_cnn_predict_into_ctx(self.h, solver.h, q, zb))

    def set_sparse(self, on: bool):
        """Tile-sparse prediction (the default): only the image tiles within
        layers - 1 cells of a wet cell; same outputs at the wet cells."""
        self._chk(lib().hd_cnn_set_sparse(self.h, 1 if on else 0))
