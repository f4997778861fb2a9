"""The MLP's HIP kernels against the numpy oracle (forward, loss, gradient),
BF16 matrix-core inference against FP32, and training on a toy target."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net():
    from quack_amd import gw
    return gw.MLP(nin=14, width=128, depth=3, seed=7)


def data(n=333, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, 14)).astype(np.float32)
    T = (np.sin(X[:, 0]) + 0.3 * X[:, 1] * X[:, 2] - 0.1 * X[:, 3] ** 2).astype(np.float32)
    return X, T


def test_forward_and_gradient_vs_oracle(net):
    print("mlp gradient check", flush=True)
    X, T = data()
    p = net.get_params()
    y = net.forward(X)
    y0, _, _ = mo.forward(p, X, 14, 128, 3)
    assert np.abs(y - y0).max() < 1e-4 * max(1.0, np.abs(y0).max())
    loss, g = net.loss_grad(X, T)
    loss0, g0 = mo.loss_grad(p, X, T, 14, 128, 3)
    assert abs(loss - loss0) < 1e-5 * max(1.0, loss0)
    assert np.abs(g - g0).max() < 1e-4 * max(1e-3, np.abs(g0).max())
    # the oracle gradient itself against a central finite difference
    k = int(np.argmax(np.abs(g0)))
    h = 1e-4
    pp, pm = p.astype(np.float64), p.astype(np.float64)
    pp = pp.copy(); pm = pm.copy()
    pp[k] += h; pm[k] -= h
    fd = (mo.loss_grad(pp, X, T, 14, 128, 3)[0] - mo.loss_grad(pm, X, T, 14, 128, 3)[0]) / (2 * h)
    assert abs(fd - g0[k]) < 1e-5 * max(1.0, abs(g0[k]))


def test_bf16_inference_matches_fp32(net):
    print("mlp bf16 inference", flush=True)
    X, _ = data(1000, 3)
    y32 = net.forward(X)
    y16 = net.forward(X, bf16=True)
    scale = max(1.0, np.abs(y32).max())
    assert np.abs(y16 - y32).max() < 3e-2 * scale
    assert np.corrcoef(y16, y32)[0, 1] > 0.999


def test_training_reduces_loss():
    from quack_amd import gw
    print("mlp training", flush=True)
    X, T = data(4096, 1)
    m = gw.MLP(nin=14, width=128, depth=3, seed=1)
    l0, _ = m.loss_grad(X, T)
    hist = m.train(X, T, epochs=60, batch=512, lr=2e-3, seed=3)
    l1, _ = m.loss_grad(X, T)
    assert l1 < 0.1 * l0 and hist[-1] < hist[0]
    Xv, Tv = data(1000, 9)
    yv = m.forward(Xv)
    assert np.mean((yv - Tv) ** 2) < 0.2 * np.mean((Tv - Tv.mean()) ** 2)
