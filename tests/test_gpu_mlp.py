"""The HIP kernels of the proposer network (rp_mlp_forward, rp_mlp_loss_grad,
rp_mlp_adam, rp_mlp_infer) against the oracle's double-precision loops."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NP = 5122


def setup(ra, rng, n, scale=0.3):
    p = rng.normal(0, scale, NP).astype(np.float32)
    X = rng.normal(0, 1, (n, 12)).astype(np.float32)
    T = rng.normal(0, 1, (n, 2)).astype(np.float32)
    W = rng.uniform(0, 1, (n, 2)).astype(np.float32)
    return p, X, T, W


@pytest.mark.parametrize("n", [1, 5, 64, 1000])
def test_forward(gpu, orc, n):
    rng = np.random.default_rng(n)
    p, X, _, _ = setup(gpu, rng, n)
    Y = gpu.mlp_forward(gpu.DeviceArray(p), gpu.DeviceArray(X)).get()
    ref = orc.mlp_forward(p.astype(np.float64), X.astype(np.float64))
    assert np.abs(Y - ref).max() < 2e-5 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("n,weighted", [(1, True), (7, False), (129, True), (3000, True)])
def test_loss_grad(gpu, orc, n, weighted):
    rng = np.random.default_rng(100 + n)
    p, X, T, W = setup(gpu, rng, n)
    dp, dX, dT = gpu.DeviceArray(p), gpu.DeviceArray(X), gpu.DeviceArray(T)
    dW = gpu.DeviceArray(W) if weighted else None
    g = gpu.DeviceArray.zeros(NP)
    loss = gpu.mlp_loss_grad(dp, dX, dT, dW, g)
    gg = g.get()
    lref, gref = orc.mlp_loss_grad(p.astype(np.float64), X.astype(np.float64),
                                   T.astype(np.float64), W.astype(np.float64) if weighted else None)
    print("n=%d loss %.8g ref %.8g  grad err %.3g (scale %.3g)" % (
        n, loss, lref, np.abs(gg - gref).max(), np.abs(gref).max()), flush=True)
    assert loss == pytest.approx(lref, rel=1e-5)
    assert np.abs(gg - gref).max() < 1e-4 * np.abs(gref).max()
    # deterministic: a second call gives the same bits
    g2 = gpu.DeviceArray.zeros(NP)
    loss2 = gpu.mlp_loss_grad(dp, dX, dT, dW, g2)
    assert loss2 == loss and np.array_equal(g2.get(), gg)


def test_adam_matches_oracle(gpu, orc):
    rng = np.random.default_rng(11)
    p, X, T, W = setup(gpu, rng, 200)
    dp, dX, dT, dW = (gpu.DeviceArray(a) for a in (p, X, T, W))
    g, m, v = gpu.DeviceArray.zeros(NP), gpu.DeviceArray.zeros(NP), gpu.DeviceArray.zeros(NP)
    q = p.astype(np.float64)
    mo, vo = np.zeros(NP), np.zeros(NP)
    X64, T64, W64 = X.astype(np.float64), T.astype(np.float64), W.astype(np.float64)
    losses = []
    for step in range(1, 11):
        losses.append(gpu.mlp_loss_grad(dp, dX, dT, dW, g))
        gpu.mlp_adam(dp, g, m, v, step, lr=3e-3, wd=1e-4)
        _, go = orc.mlp_loss_grad(q, X64, T64, W64)
        orc.mlp_adam(q, go, mo, vo, 3e-3, 0.9, 0.999, 1e-8, 1e-4, step)
    err = np.abs(dp.get() - q).max()
    print("10 AdamW steps: max param diff %.3g; loss %.5g -> %.5g" % (err, losses[0], losses[-1]),
          flush=True)
    # after 10 steps of size <= lr the trajectories agree to far below lr
    assert err < 3e-4
    assert losses[-1] < losses[0]


def test_infer_matches_forward(gpu):
    rng = np.random.default_rng(12)
    p, X, _, _ = setup(gpu, rng, 64)
    dp = gpu.DeviceArray(p)
    Y = gpu.mlp_forward(dp, gpu.DeviceArray(X)).get()
    assert np.array_equal(gpu.mlp_infer(dp, X), Y)
    assert np.array_equal(gpu.mlp_infer(dp, X[3]), Y[3:4])
    with pytest.raises(gpu.RpError):
        gpu.mlp_infer(dp, np.zeros((65, 12), np.float32))


def test_training_fits_smooth_target(gpu):
    # the network learns y0 = sum(x[:3]) / 2, y1 = x[4] * x[5] from 2000
    # samples by full-batch AdamW on the GPU
    rng = np.random.default_rng(13)
    X = rng.uniform(-1, 1, (2000, 12)).astype(np.float32)
    T = np.stack([X[:, :3].sum(1) / 2, X[:, 4] * X[:, 5]], 1).astype(np.float32)
    p = (rng.normal(0, 1, NP) * np.concatenate([
        np.full(768, 1 / np.sqrt(12)), np.zeros(64), np.full(4096, 1 / 8), np.zeros(64),
        np.full(128, 1 / 8), np.zeros(2)])).astype(np.float32)
    dp, dX, dT = gpu.DeviceArray(p), gpu.DeviceArray(X), gpu.DeviceArray(T)
    g, m, v = gpu.DeviceArray.zeros(NP), gpu.DeviceArray.zeros(NP), gpu.DeviceArray.zeros(NP)
    l0 = None
    for step in range(1, 601):
        loss = gpu.mlp_loss_grad(dp, dX, dT, None, g)
        l0 = l0 or loss
        gpu.mlp_adam(dp, g, m, v, step, lr=3e-3)
    print("fit: loss %.4g -> %.4g" % (l0, loss), flush=True)
    assert loss < 0.02 * l0
