"""FP64 reference of the cutoff surrogate network (6 -> 128 -> 128 -> 1, GELU).

TEST INFRASTRUCTURE ONLY: the checker for the GPU network's forward pass,
gradients and Adam step.  The network has no counterpart in the reference;
its maths is fixed here and pinned by finite differences.

Parameter vector layout (17,537 values):
  W1[128][6] | b1[128] | W2[128][128] | b2[128] | w3[128] | b3[1]
"""
import math

import numpy as np

D_IN, H = 6, 128
SIZES = [("W1", (H, D_IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("w3", (H,)), ("b3", (1,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(theta):
    out, o = {}, 0
    for name, shp in SIZES:
        n = int(np.prod(shp))
        out[name] = theta[o:o + n].reshape(shp)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SIZES])


_erf = np.vectorize(math.erf)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def forward(theta, X):
    p = unpack(np.asarray(theta, np.float64))
    Z1 = X @ p["W1"].T + p["b1"]
    H1 = gelu(Z1)
    Z2 = H1 @ p["W2"].T + p["b2"]
    H2 = gelu(Z2)
    y = H2 @ p["w3"] + p["b3"][0]
    return y, (Z1, H1, Z2, H2)


def loss_and_grad(theta, X, t):
    """Mean-squared error over the batch and its gradient."""
    p = unpack(np.asarray(theta, np.float64))
    y, (Z1, H1, Z2, H2) = forward(theta, X)
    B = len(t)
    r = y - t
    L = float(np.mean(r * r))
    dy = 2.0 * r / B
    g = {}
    g["w3"] = H2.T @ dy
    g["b3"] = np.array([dy.sum()])
    dZ2 = np.outer(dy, p["w3"]) * gelu_grad(Z2)
    g["W2"] = dZ2.T @ H1
    g["b2"] = dZ2.sum(0)
    dZ1 = (dZ2 @ p["W2"]) * gelu_grad(Z1)
    g["W1"] = dZ1.T @ X
    g["b1"] = dZ1.sum(0)
    return L, pack(g)


def adam_step(theta, m, v, g, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """Adam (bias-corrected), step counted from 1."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(seed=0):
    """He-uniform-like initialisation used by both sides (host-generated)."""
    rng = np.random.default_rng(seed)
    d = {}
    d["W1"] = rng.uniform(-1, 1, (H, D_IN)) * math.sqrt(6.0 / D_IN)
    d["b1"] = np.zeros(H)
    d["W2"] = rng.uniform(-1, 1, (H, H)) * math.sqrt(6.0 / H)
    d["b2"] = np.zeros(H)
    d["w3"] = rng.uniform(-1, 1, H) * math.sqrt(3.0 / H)
    d["b3"] = np.zeros(1)
    return pack(d)


def bf16_round(a):
    """Round-to-nearest-even to bfloat16, returned as float64."""
    a32 = np.asarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)
