"""CPU restatement of the flow surrogate's network (the project's own model,
include/hydist_cnn.h): 3x3 zero-padded convolutions on NHWC images, layer 0
ReLU, residual ReLU layers, linear output layer added to the input u, v, z;
masked mean-squared error; exact FP64 backward.  `emulate_bf16` rounds the
weights, the input image and the stored activations to bfloat16 the way the
gfx950 kernels store them, so that the GPU's forward pass can be compared
within bfloat16 tolerance.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the hydist_amd package.
Parity of this restatement is pinned by central finite differences of the
loss (tests/test_cnn_oracle.py).
"""
from __future__ import annotations

import numpy as np

CIN, CIN_PAD, WIDTH, COUT, COUT_PAD = 8, 32, 64, 3, 32


def bf16(x):
    """Round to bfloat16 (round to nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def dims(L):
    """(cin, cout) padded channel counts of each layer."""
    return [(CIN_PAD if l == 0 else WIDTH, COUT_PAD if l == L - 1 else WIDTH) for l in range(L)]


def conv(X, Wt, b=None):
    """Y[b,h,w,co] = sum_{tap,ci} X[b,h+dy,w+dx,ci] Wt[tap,co,ci] (+ b[co]),
    tap = 3(dy+1)+(dx+1), zero padding."""
    B, H, W, C = X.shape
    Xp = np.zeros((B, H + 2, W + 2, C))
    Xp[:, 1:-1, 1:-1] = X
    Y = np.zeros((B, H, W, Wt.shape[1]))
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        Y += Xp[:, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W] @ Wt[tap].T
    if b is not None:
        Y += b
    return Y


def conv_wgrad(X, dY):
    """dW[tap,co,ci] = sum_p dY[p,co] X[p+off(tap),ci]."""
    B, H, W, C = X.shape
    Xp = np.zeros((B, H + 2, W + 2, C))
    Xp[:, 1:-1, 1:-1] = X
    dW = np.zeros((9, dY.shape[-1], C))
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        xs = Xp[:, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W].reshape(-1, C)
        dW[tap] = dY.reshape(-1, dY.shape[-1]).T @ xs
    return dW


def conv_dgrad(dY, Wt):
    """dX of conv: conv of dY with the tap-flipped, transposed weights."""
    WT = np.transpose(Wt[::-1], (0, 2, 1))
    return conv(dY, WT)


def forward(params, x8, emulate_bf16=False):
    """params: list of (W [9,cout,cin], b [cout]) with padded counts;
    x8: [B,H,W,8].  Returns (prediction [B,H,W,3], cache)."""
    q = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    L = len(params)
    B, H, W, _ = x8.shape
    x0 = np.zeros((B, H, W, CIN_PAD))
    x0[..., :CIN] = q(x8)
    base = np.asarray(x8[..., :3], dtype=np.float64)
    if emulate_bf16:
        base = np.asarray(x8[..., :3], dtype=np.float32).astype(np.float64)
    Ws = [q(w) for w, _ in params]
    acts, rels = [x0], [None]
    a = q(np.maximum(conv(x0, Ws[0], params[0][1]), 0.0))
    acts.append(a)
    for l in range(1, L - 1):
        r = np.maximum(conv(a, Ws[l], params[l][1]), 0.0)
        rels.append(q(r))
        a = q(a + r)
        acts.append(a)
    out = conv(a, Ws[L - 1], params[L - 1][1])
    pred = base + out[..., :COUT]
    return pred, dict(acts=acts, rels=rels, Ws=Ws, out=out)


def loss_and_grads(params, x8, tgt, msk):
    """Masked MSE (mean over wet pixels and the 3 outputs) and its exact
    gradients [(dW, db)] in FP64."""
    pred, c = forward(params, x8)
    L = len(params)
    norm = 3.0 * msk.sum()
    d = pred - tgt
    loss = float((msk[..., None] * d * d).sum() / norm)
    dout = np.zeros(c["out"].shape)
    dout[..., :COUT] = 2.0 * msk[..., None] * d / norm
    grads = [None] * L
    acts, Ws = c["acts"], c["Ws"]
    grads[L - 1] = (conv_wgrad(acts[L - 1], dout), dout.sum(axis=(0, 1, 2)))
    dA = conv_dgrad(dout, Ws[L - 1])
    for l in range(L - 2, 0, -1):
        z = conv(acts[l], Ws[l], params[l][1])
        dZ = dA * (z > 0)
        grads[l] = (conv_wgrad(acts[l], dZ), dZ.sum(axis=(0, 1, 2)))
        dA = dA + conv_dgrad(dZ, Ws[l])
    z0 = conv(acts[0], Ws[0], params[0][1])
    dZ = dA * (z0 > 0)
    grads[0] = (conv_wgrad(acts[0], dZ), dZ.sum(axis=(0, 1, 2)))
    return loss, grads


def random_params(L, seed=0, scale=0.3):
    """Random weights with the padded layout; padded input channels and
    padded outputs zero, as the library initialises them."""
    rng = np.random.default_rng(seed)
    ps = []
    for l, (ci, co) in enumerate(dims(L)):
        w = np.zeros((9, co, ci))
        cr = CIN if l == 0 else WIDTH
        orr = COUT if l == L - 1 else WIDTH
        w[:, :orr, :cr] = rng.uniform(-1, 1, (9, orr, cr)) * scale / np.sqrt(9 * cr)
        b = np.zeros(co)
        b[:orr] = rng.uniform(-0.1, 0.1, orr)
        ps.append((w, b))
    return ps
