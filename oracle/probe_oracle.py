"""CPU fp64 restatement of the learned probe-budget model (test infrastructure).

The model has no counterpart in the reference; this restates the GPU kernels'
definition (vectorcache_amd/csrc/vc_probe.hip): an MLP 96 -> 256 -> 256 -> 1
with ReLU, the loss mean((y - t)^2) over a minibatch, Adam. With
``bf16=True`` every value that enters a matrix product on the GPU is rounded
to bf16 at the same point (the features, W1, W2, the first hidden layer, the
backward signals dZ2 and dZ1); everything else is exact fp64 arithmetic on
the fp32 inputs. With ``bf16=False`` it is the plain fp64 network, whose
gradients the tests check against central differences.

Parameter vector layout (fp32 on the GPU): W1 [256][96], b1 [256],
W2 [256][256], b2 [256], w3 [256], b3 (+3 pad).
"""
from __future__ import annotations

import numpy as np

NI, NH = 96, 256
O_W1 = 0
O_B1 = O_W1 + NH * NI
O_W2 = O_B1 + NH
O_B2 = O_W2 + NH * NH
O_W3 = O_B2 + NH
O_B3 = O_W3 + NH
N_PARAMS = O_B3 + 4


def bf16(a) -> np.ndarray:
    """Round float32 values to bf16 (nearest, ties to even), as float32."""
    a = np.ascontiguousarray(a, np.float32).copy()
    u = a.view(np.uint32)
    u[:] = ((u.astype(np.uint64) + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return a


def unpack(w):
    w = np.asarray(w)
    if w.dtype != np.float64:
        w = w.astype(np.float32)
    return (w[O_W1:O_B1].reshape(NH, NI), w[O_B1:O_W2], w[O_W2:O_B2].reshape(NH, NH), w[O_B2:O_W3],
            w[O_W3:O_B3], w[O_B3])


def forward(w, X, bf16_points: bool = True):
    W1, b1, W2, b2, w3, b3 = unpack(w)
    r = bf16 if bf16_points else (lambda a: np.asarray(a))
    Xb = r(X).astype(np.float64)
    W1b = r(W1).astype(np.float64)
    W2b = r(W2).astype(np.float64)
    Z1 = Xb @ W1b.T + b1.astype(np.float64)
    H1 = np.maximum(Z1, 0.0)
    H1b = r(H1.astype(np.float32)).astype(np.float64) if bf16_points else H1
    Z2 = H1b @ W2b.T + b2.astype(np.float64)
    H2 = np.maximum(Z2, 0.0)
    y = H2 @ w3.astype(np.float64) + np.float64(b3)
    return dict(Xb=Xb, W1b=W1b, W2b=W2b, H1b=H1b, H2=H2, y=y)


def loss(w, X, t, bf16_points: bool = False) -> float:
    y = forward(w, X, bf16_points)["y"]
    return float(np.mean((y - np.asarray(t, np.float64)) ** 2))


def grads(w, X, t, bf16_points: bool = True) -> np.ndarray:
    """Gradient of the minibatch loss, in the parameter layout."""
    W1, b1, W2, b2, w3, b3 = unpack(w)
    f = forward(w, X, bf16_points)
    B = X.shape[0]
    dy = 2.0 * (f["y"] - np.asarray(t, np.float64)) / B
    g = np.zeros(N_PARAMS, np.float64)
    g[O_W3:O_B3] = f["H2"].T @ dy
    g[O_B3] = dy.sum()
    dZ2 = (f["H2"] > 0) * dy[:, None] * w3.astype(np.float64)[None, :]
    g[O_B2:O_W3] = dZ2.sum(0)
    dZ2m = bf16(dZ2.astype(np.float32)).astype(np.float64) if bf16_points else dZ2
    g[O_W2:O_B2] = (dZ2m.T @ f["H1b"]).ravel()
    dZ1 = (dZ2m @ f["W2b"]) * (f["H1b"] > 0)
    g[O_B1:O_W2] = dZ1.sum(0)
    dZ1m = bf16(dZ1.astype(np.float32)).astype(np.float64) if bf16_points else dZ1
    g[O_W1:O_B1] = (dZ1m.T @ f["Xb"]).ravel()
    return g


def adam_step(w, m, v, g, step: int, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
    """One Adam update (step counts from 1) in fp64: returns (w, m, v)."""
    w = np.asarray(w, np.float64)
    m = b1 * np.asarray(m, np.float64) + (1 - b1) * g
    v = b2 * np.asarray(v, np.float64) + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v
