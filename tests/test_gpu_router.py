"""Router MLP kernels on the GPU against the FP64 oracle (marked gpu):
forward outputs, loss gradient and parameter gradients (within FP32
tolerance), Adam step, and a short training run that lowers the loss."""
import numpy as np
import pytest

from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    from sable_amd import runtime
    from sable_amd.router import mlp
    runtime.init()
    return runtime, mlp


def _data(n=300, F=6, K=3, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (n, F)).astype(np.float32)
    T = rng.normal(0, 1, (n, K)).astype(np.float32)
    M = (rng.random((n, K)) < 0.8).astype(np.float32)
    return X, T, M


def test_forward_and_gradient_match_oracle(mods):
    rt, mlp = mods
    F, H, K = 6, 32, 3
    p = mlp.init_params(1)
    X, T, M = _data()
    n = len(X)
    net = mlp.DeviceMLP(p)
    Xd, Td, Md = rt.DeviceArray.from_numpy(X), rt.DeviceArray.from_numpy(T), rt.DeviceArray.from_numpy(M)
    print("case: mlp forward/backward", flush=True)
    out = net.gradient(Xd, Td, Md, n, 1.0 / n)
    o64, _ = MO.forward(p, X, F, H, K)
    assert np.allclose(out.numpy(), o64, rtol=1e-4, atol=1e-5)
    _, g64 = MO.grad(p, X, T, M, F, H, K, 1.0 / n)
    g = net.g.numpy()
    assert np.max(np.abs(g - g64)) <= 1e-4 * max(1.0, np.max(np.abs(g64)))


def test_adam_step_matches_oracle(mods):
    rt, mlp = mods
    p = mlp.init_params(2)
    X, T, M = _data(seed=3)
    n = len(X)
    net = mlp.DeviceMLP(p)
    Xd, Td, Md = rt.DeviceArray.from_numpy(X), rt.DeviceArray.from_numpy(T), rt.DeviceArray.from_numpy(M)
    net.gradient(Xd, Td, Md, n, 1.0 / n)
    g = net.g.numpy().astype(np.float64)
    net.adam(lr=1e-2)
    p1, _, _ = MO.adam(p.astype(np.float64), g, np.zeros_like(g), np.zeros_like(g), 1e-2, 0.9, 0.999, 1e-8, 1)
    assert np.allclose(net.params(), p1, rtol=1e-5, atol=1e-6)


def test_training_lowers_loss(mods):
    rt, mlp = mods
    rng = np.random.default_rng(5)
    X = rng.normal(0, 1, (4000, 6)).astype(np.float32)
    W = rng.normal(0, 1, (6, 3))
    T = (np.tanh(X @ W) + 0.1 * X[:, :3]).astype(np.float32)
    M = np.ones_like(T)
    net = mlp.DeviceMLP(mlp.init_params(0))
    print("case: mlp training", flush=True)
    hist = net.train(X, T, M, steps=300, batch=512, lr=1e-2, log_every=100)
    assert hist[-1][1] < 0.5 * hist[0][1]
