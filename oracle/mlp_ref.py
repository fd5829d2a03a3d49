"""CPU reference of the learned cost model (TEST INFRASTRUCTURE ONLY).

float64 restatement of the 8 -> 32 -> 32 -> 3 MLP of lpa_amd/csrc/hip/mlp_kernels.hpp
(ReLU hidden layers, linear outputs, masked mean-squared error).  The model
has no reference counterpart; its gradient is pinned by central finite
differences (tests/test_oracle_mlp.py).
Parameter layout: W1[32][8] b1[32] W2[32][32] b2[32] W3[3][32] b3[3].
"""
import numpy as np

IN, H, OUT = 8, 32, 3
SIZES = [("W1", (H, IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (OUT, H)), ("b3", (OUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    p = np.asarray(p, np.float64)
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = p[o:o + k].reshape(shape)
        o += k
    return out


def forward(p, x):
    w = unpack(p)
    x = np.asarray(x, np.float64)
    z1 = x @ w["W1"].T + w["b1"]
    h1 = np.maximum(z1, 0)
    z2 = h1 @ w["W2"].T + w["b2"]
    h2 = np.maximum(z2, 0)
    y = h2 @ w["W3"].T + w["b3"]
    return y, (x, z1, h1, z2, h2)


def loss(p, x, t, mask):
    y, _ = forward(p, x)
    e = np.where(mask, y - t, 0.0)
    return float((e * e).sum() / len(x))


def loss_grad(p, x, t, mask):
    w = unpack(p)
    y, (x, z1, h1, z2, h2) = forward(p, x)
    n = len(x)
    e = np.where(mask, y - t, 0.0)
    dy = 2.0 * e / n
    g = {"W3": dy.T @ h2, "b3": dy.sum(0)}
    dh2 = (dy @ w["W3"]) * (z2 > 0)
    g["W2"], g["b2"] = dh2.T @ h1, dh2.sum(0)
    dh1 = (dh2 @ w["W2"]) * (z1 > 0)
    g["W1"], g["b1"] = dh1.T @ x, dh1.sum(0)
    flat = np.concatenate([g[name].reshape(-1) for name, _ in SIZES])
    return float((e * e).sum() / n), flat
