"""CPU oracle of the ladder network (include/drift_mlp.h) -- TEST INFRASTRUCTURE ONLY.

A float64 numpy restatement of the network, its softplus-cumsum head, loss and backward pass, and
of BF16 rounding (round to nearest even on the top 16 bits of a float32).  The analytic gradients
are themselves checked against central finite differences in the tests, so this file pins the
GPU kernels without depending on them.
"""

from __future__ import annotations

import numpy as np

IN, HID, OUT = 16, 128, 32
SHAPES = [("W1", (IN, HID)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
          ("W3", (HID, HID)), ("b3", (HID,)), ("W4", (HID, OUT)), ("b4", (OUT,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p: np.ndarray) -> dict:
    out, off = {}, 0
    for name, shape in SHAPES:
        size = int(np.prod(shape))
        out[name] = p[off:off + size].reshape(shape)
        off += size
    return out


def bf16_round(x: np.ndarray) -> np.ndarray:
    """float32 -> nearest-even bfloat16 -> float32, as the device does."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def _sigmoid(x):
    return 0.5 * (1.0 + np.tanh(0.5 * x))


def _silu(x):
    return x * _sigmoid(x)


def _dsilu(x):
    s = _sigmoid(x)
    return s * (1.0 + x * (1.0 - s))


def forward(params, X, bf16=True, cache=False):
    P = unpack(np.asarray(params, dtype=np.float64))
    W = {k: (bf16_round(v).astype(np.float64) if bf16 and k.startswith("W") else v)
         for k, v in P.items()}
    X = np.asarray(X, dtype=np.float64)
    p1 = X @ W["W1"] + W["b1"]
    a1 = _silu(p1)
    p2 = a1 @ W["W2"] + W["b2"]
    a2 = _silu(p2)
    p3 = a2 @ W["W3"] + W["b3"]
    a3 = _silu(p3)
    z = a3 @ W["W4"] + W["b4"]
    d = np.logaddexp(0.0, z) + 1e-4
    S = np.concatenate([np.zeros((z.shape[0], 1)), np.cumsum(d[:, 1:], axis=1)], axis=1)
    x = S / S[:, -1:]
    if cache:
        return x, dict(W=W, X=X, p1=p1, a1=a1, p2=p2, a2=a2, p3=p3, a3=a3, z=z, S=S)
    return x


def loss(params, X, T, bf16=True):
    x = forward(params, X, bf16)
    e = (x - T)[:, 1:OUT - 1]
    return float((e ** 2).sum() / (x.shape[0] * (OUT - 2)))


def gradients(params, X, T, bf16=True):
    """Analytic gradient with respect to the master weights (straight through the BF16 rounding)."""
    x, c = forward(params, X, bf16, cache=True)
    B = x.shape[0]
    g = np.zeros_like(x)
    g[:, 1:OUT - 1] = 2.0 * (x - T)[:, 1:OUT - 1] / (B * (OUT - 2))
    S = c["S"]
    C = S[:, -1:]
    tail = np.cumsum(g[:, ::-1], axis=1)[:, ::-1]            # sum_{j >= k} g_j
    gS = (g * S).sum(axis=1, keepdims=True)
    gd = tail / C - gS / C ** 2
    gd[:, 0] = 0.0
    gz = gd * _sigmoid(c["z"])
    W = c["W"]
    grads = {"W4": c["a3"].T @ gz, "b4": gz.sum(0)}
    g3 = (gz @ W["W4"].T) * _dsilu(c["p3"])
    grads.update(W3=c["a2"].T @ g3, b3=g3.sum(0))
    g2 = (g3 @ W["W3"].T) * _dsilu(c["p2"])
    grads.update(W2=c["a1"].T @ g2, b2=g2.sum(0))
    g1 = (g2 @ W["W2"].T) * _dsilu(c["p1"])
    grads.update(W1=c["X"].T @ g1, b1=g1.sum(0))
    return np.concatenate([grads[name].ravel() for name, _ in SHAPES])
