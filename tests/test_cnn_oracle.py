"""The CPU restatement of the surrogate network pinned by what the
mathematics fixes: central finite differences of the loss against its
analytic gradients, the convolution against a direct loop, and the
tap-flipped transpose as the adjoint of the convolution."""
import numpy as np

import cnn_oracle as co


def test_conv_against_direct_loop():
    rng = np.random.default_rng(0)
    X = rng.normal(size=(1, 4, 5, 3))
    Wt = rng.normal(size=(9, 2, 3))
    Y = co.conv(X, Wt)
    for h in range(4):
        for w in range(5):
            for o in range(2):
                s = 0.0
                for tap in range(9):
                    hh, ww = h + tap // 3 - 1, w + tap % 3 - 1
                    if 0 <= hh < 4 and 0 <= ww < 5:
                        s += X[0, hh, ww] @ Wt[tap, o]
                assert abs(Y[0, h, w, o] - s) < 1e-12


def test_dgrad_is_adjoint_of_conv():
    rng = np.random.default_rng(1)
    X = rng.normal(size=(2, 5, 6, 4))
    Wt = rng.normal(size=(9, 3, 4))
    dY = rng.normal(size=(2, 5, 6, 3))
    lhs = (co.conv(X, Wt) * dY).sum()
    rhs = (X * co.conv_dgrad(dY, Wt)).sum()
    assert abs(lhs - rhs) < 1e-9 * abs(lhs)
    # and the weight gradient is the other adjoint
    dW = co.conv_wgrad(X, dY)
    assert abs((dW * Wt).sum() - lhs) < 1e-9 * abs(lhs)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(2)
    L = 4
    ps = co.random_params(L, seed=3, scale=1.0)
    x8 = rng.normal(size=(1, 5, 6, 8))
    tgt = rng.normal(size=(1, 5, 6, 3))
    msk = (rng.random((1, 5, 6)) > 0.3).astype(float)
    loss, g = co.loss_and_grads(ps, x8, tgt, msk)
    eps = 1e-6
    checked = 0
    for l in range(L):
        w, b = ps[l]
        cr = co.CIN if l == 0 else co.WIDTH
        orr = co.COUT if l == L - 1 else co.WIDTH
        for _ in range(6):
            idx = (rng.integers(9), rng.integers(orr), rng.integers(cr))
            w[idx] += eps
            lp, _ = co.loss_and_grads(ps, x8, tgt, msk)
            w[idx] -= 2 * eps
            lm, _ = co.loss_and_grads(ps, x8, tgt, msk)
            w[idx] += eps
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g[l][0][idx]) < 1e-6 + 1e-4 * abs(fd), (l, idx, fd, g[l][0][idx])
            checked += 1
        k = rng.integers(orr)
        b[k] += eps
        lp, _ = co.loss_and_grads(ps, x8, tgt, msk)
        b[k] -= 2 * eps
        lm, _ = co.loss_and_grads(ps, x8, tgt, msk)
        b[k] += eps
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[l][1][k]) < 1e-6 + 1e-4 * abs(fd)
    assert checked == 6 * L


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-3])
    r = co.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0            # tie to even
    assert r[2] == 1.0 + 2 ** -7                  # rounds up
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-3) < 3.0e-3 * 2 ** -8
