"""The router's 2-layer MLP on the GPU (library kernels: forward, masked
MSE gradient, backward, Adam).  Weights are saved as .npz (no pickles)."""
from __future__ import annotations

import numpy as np

from .._lib import check, lib
from ..runtime import DeviceArray, init

N_FEATURES = 6
N_HIDDEN = 32
N_ROUTES = 3          # dense FP32, sparse, dense MFMA
N_SHAPES = 2          # workgroup shapes (see labels.SHAPES)
N_OUT = N_ROUTES * N_SHAPES   # output k = shape * N_ROUTES + route


def n_params(F=N_FEATURES, H=N_HIDDEN, K=N_OUT):
    return H * F + H + K * H + K


def outputs_of(params, F=N_FEATURES, H=N_HIDDEN):
    """The output count K of a parameter vector of a 2-layer F-H-K MLP."""
    k, r = divmod(np.asarray(params).size - H * F - H, H + 1)
    if r or k <= 0:
        raise ValueError("parameter vector does not match an MLP with these F, H")
    return int(k)


def init_params(seed=0, F=N_FEATURES, H=N_HIDDEN, K=N_OUT):
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, np.sqrt(2.0 / F), (H, F))
    b1 = np.zeros(H)
    W2 = rng.normal(0, np.sqrt(1.0 / H), (K, H))
    b2 = np.zeros(K)
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2]).astype(np.float32)


class DeviceMLP:
    """Parameters and optimizer state on the device."""

    def __init__(self, params, F=N_FEATURES, H=N_HIDDEN, K=None):
        init()
        K = outputs_of(params, F, H) if K is None else K
        self.F, self.H, self.K = F, H, K
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == n_params(F, H, K)
        self.p = DeviceArray.from_numpy(p)
        self.g = DeviceArray.zeros(p.size)
        self.m = DeviceArray.zeros(p.size)
        self.v = DeviceArray.zeros(p.size)
        self.step = 0
        self._bufs = {}

    def _buf(self, name, shape):
        b = self._bufs.get(name)
        if b is None or b.shape != shape:
            b = DeviceArray(shape)
            self._bufs[name] = b
        return b

    def params(self):
        return self.p.numpy()

    def forward(self, Xd, n):
        out = self._buf("out", (n, self.K))
        hid = self._buf("hid", (n, self.H))
        check(lib.sable_mlp_forward(self.p.ptr, self.F, self.H, self.K, Xd.ptr, n, out.ptr, hid.ptr))
        return out, hid

    def predict(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Xd = DeviceArray.from_numpy(X)
        out, _ = self.forward(Xd, X.shape[0])
        return out.numpy()

    def gradient(self, Xd, Td, Md, n, scale):
        """Loss gradient w.r.t. the parameters into self.g (device)."""
        out, hid = self.forward(Xd, n)
        dout = self._buf("dout", (n, self.K))
        check(lib.sable_mlp_mse_grad(out.ptr, Td.ptr, Md.ptr, n * self.K, scale, dout.ptr))
        check(lib.sable_mlp_backward(self.p.ptr, self.F, self.H, self.K, Xd.ptr, n, hid.ptr, dout.ptr,
                                     self.g.ptr))
        return out

    def adam(self, lr=3e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step += 1
        check(lib.sable_adam_step(self.p.ptr, self.g.ptr, self.m.ptr, self.v.ptr, self.p.size, lr, b1, b2, eps,
                                  self.step))

    def train(self, X, T, M, steps=2000, batch=4096, lr=3e-3, seed=0, log_every=0):
        """Mini-batch Adam on masked MSE; the batches are drawn on the host,
        every arithmetic step runs on the device."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        M = np.ascontiguousarray(M, dtype=np.float32)
        n = X.shape[0]
        rng = np.random.default_rng(seed)
        batch = min(batch, n)
        Xb, Tb, Mb = DeviceArray((batch, self.F)), DeviceArray((batch, self.K)), DeviceArray((batch, self.K))
        hist = []
        for s in range(steps):
            idx = rng.integers(0, n, batch)
            Xb.copy_from(X[idx])
            Tb.copy_from(T[idx])
            Mb.copy_from(M[idx])
            denom = max(float(M[idx].sum()), 1.0)
            out = self.gradient(Xb, Tb, Mb, batch, 1.0 / denom)
            self.adam(lr=lr * (0.5 ** (s / max(steps / 4, 1))))
            if log_every and (s % log_every == 0 or s == steps - 1):
                o = out.numpy()
                hist.append((s, float(np.sum(M[idx] * (o - T[idx]) ** 2) / denom)))
        return hist


def save(path, params, mean, std, meta):
    np.savez(path, params=np.asarray(params, np.float32), mean=np.asarray(mean, np.float64),
             std=np.asarray(std, np.float64), meta=np.array(str(meta)))


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return z["params"], z["mean"], z["std"], str(z["meta"])
