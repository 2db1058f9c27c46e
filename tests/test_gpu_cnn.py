"""GPU: the surrogate network's BF16 MFMA kernels against the CPU
restatement (oracle/cnn_oracle.py): the MFMA operand layout, the forward
pass within bfloat16 tolerance, the gradients of every layer, and a short
training run that lowers the loss."""
import numpy as np
import pytest

import cnn_oracle as co
from hydist_amd import cnn

pytestmark = pytest.mark.gpu


def test_mfma_tile_layout_exact():
    rng = np.random.default_rng(0)
    A = rng.integers(-4, 5, size=(16, 32)).astype(np.float32)
    B = rng.integers(-4, 5, size=(32, 16)).astype(np.float32)
    D = cnn.mfma_check(A, B)
    assert np.array_equal(D, A @ B)


def setup_net(H, W, nimg, L, seed):
    rng = np.random.default_rng(seed)
    ps = co.random_params(L, seed=seed, scale=1.5)
    ps = [(w.astype(np.float32).astype(np.float64), b.astype(np.float32).astype(np.float64)) for w, b in ps]
    net = cnn.Network(H, W, batch=nimg, layers=L, seed=seed)
    for l, (w, b) in enumerate(ps):
        net.set_layer(l, w, b)
    x8 = rng.normal(size=(nimg, H, W, 8)).astype(np.float32)
    x8[..., 4] = rng.random((nimg, H, W)) > 0.5
    tgt = (x8[..., :3] + 0.2 * rng.normal(size=(nimg, H, W, 3))).astype(np.float32)
    msk = (rng.random((nimg, H, W)) > 0.25).astype(np.float32)
    return net, ps, x8, tgt, msk


@pytest.mark.parametrize("H,W,nimg", [(7, 70, 2), (5, 13, 1)])
def test_forward_matches_oracle(H, W, nimg):
    net, ps, x8, tgt, msk = setup_net(H, W, nimg, 4, seed=1)
    net.load(x8, tgt, msk)
    pred = net.forward()
    ref, _ = co.forward(ps, x8.astype(np.float64), emulate_bf16=True)
    out_ref = ref - x8[..., :3].astype(np.float64)
    scale = np.abs(out_ref).max()
    err = np.abs(pred - ref).max()
    assert err <= 1e-2 * scale + 1e-4, (err, scale)


def test_gradients_match_oracle():
    H, W, nimg, L = 6, 67, 2, 4
    net, ps, x8, tgt, msk = setup_net(H, W, nimg, L, seed=2)
    net.load(x8, tgt, msk)
    net.forward()
    loss = net.backward()
    loss_ref, g = co.loss_and_grads(ps, x8.astype(np.float64), tgt.astype(np.float64), msk.astype(np.float64))
    assert abs(loss - loss_ref) <= 2e-2 * loss_ref
    for l in range(L):
        dw, db = net.get_grad(l)
        rw = np.linalg.norm(dw - g[l][0]) / np.linalg.norm(g[l][0])
        rb = np.linalg.norm(db - g[l][1]) / max(np.linalg.norm(g[l][1]), 1e-12)
        assert rw < 5e-2 and rb < 5e-2, (l, rw, rb)


def test_training_lowers_loss():
    H, W, L = 16, 64, 4
    rng = np.random.default_rng(3)
    net = cnn.Network(H, W, batch=2, layers=L, seed=4)
    x8 = rng.normal(size=(2, H, W, 8)).astype(np.float32)
    # target: the input u, v, z plus a smooth function of two other channels
    d = 0.3 * np.tanh(x8[..., 3:4] + 0.5 * x8[..., 7:8])
    tgt = (x8[..., :3] + d).astype(np.float32)
    msk = np.ones((2, H, W), np.float32)
    net.load(x8, tgt, msk)
    first = net.train_step(1e-3)
    for _ in range(150):
        last = net.train_step(1e-3)
    assert last < 0.2 * first, (first, last)


@pytest.mark.parametrize("layers", [3, 8])
def test_sparse_prediction_equals_dense(layers):
    """hd_cnn_predict_into_ctx over only the tiles within layers - 1 cells of
    a wet cell writes the same u, v, z as the dense pass, bit for bit, with
    stale values left in the skipped tiles of every layer buffer."""
    from hydist_amd import synth
    from hydist_amd import surrogate as S
    from hydist_amd._lib import Solver
    case = synth.config(3, scale=0.2, hours=3.0)
    op = case.opts()
    g = Solver(case)
    g.init()
    g.run(0.0, 120, op)
    before = g.get("z")
    net = cnn.Network(case.N + 3, case.M + 3, batch=1, layers=layers, seed=7)
    net.set_scales(**S.SCALES)
    rng = np.random.default_rng(3)
    # fill every layer buffer with a dense pass on noise first, so that the
    # tiles the sparse pass skips hold values of another image
    net.load(rng.normal(size=(1, case.N + 3, case.M + 3, 8)).astype(np.float32),
             np.zeros((1, case.N + 3, case.M + 3, 3), np.float32), np.ones((1, case.N + 3, case.M + 3), np.float32))
    net.forward()
    g2 = g.clone()
    net.set_sparse(False)
    net.predict_into(g, 600.0, 0.2)
    net.set_sparse(True)
    net.predict_into(g2, 600.0, 0.2)
    wet = g.get("khouot") == 0
    assert wet.sum() > 100
    assert not np.array_equal(g.get("z")[wet], before[wet])
    for f in ("u", "v", "z", "t_u", "t_v", "t_z", "H_moi"):
        assert np.array_equal(g.get(f), g2.get(f), equal_nan=True), f
    g.close()
    g2.close()
