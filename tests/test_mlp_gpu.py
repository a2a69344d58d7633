"""GPU surrogate network (BF16 MFMA, FP32 accumulate) against the FP64
reference: forward pass, gradients, one Adam step, and training."""
import numpy as np
import pytest

from amps_amd import _lib as L
from amps_amd.surrogate import Surrogate
from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu


def _data(n, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1, 1, (n, 6)).astype(np.float32)
    t = (np.sin(2 * X[:, 0]) * X[:, 3] + 0.5 * X[:, 4] ** 2 - 0.3 * X[:, 5]).astype(np.float32)
    return X, t


def _rounded_forward(theta, X):
    """FP64 forward on BF16-rounded weights, inputs and hidden activations:
    what the MFMA path computes up to FP32 accumulation order."""
    p = MO.unpack(np.asarray(theta, np.float64))
    Xb = MO.bf16_round(X)
    Z1 = Xb @ MO.bf16_round(p["W1"]).T + p["b1"]
    H1 = MO.bf16_round(MO.gelu(Z1))
    Z2 = H1 @ MO.bf16_round(p["W2"]).T + p["b2"]
    H2 = MO.bf16_round(MO.gelu(Z2))
    return H2 @ p["w3"] + p["b3"][0]


def test_forward_matches_reference(gpu_ctx):
    th = MO.init_params(3).astype(np.float32)
    sur = Surrogate(gpu_ctx, theta=th)
    X, _ = _data(1000)
    y = sur.predict(X)
    yr = _rounded_forward(th.astype(np.float64), X.astype(np.float64))
    y64, _ = MO.forward(th.astype(np.float64), X.astype(np.float64))
    scale = np.abs(y64).max()
    print("forward: max |gpu - bf16 reference| / scale =", np.abs(y - yr).max() / scale,
          " vs FP64:", np.abs(y - y64).max() / scale)
    assert np.abs(y - yr).max() <= 2e-3 * scale
    assert np.abs(y - y64).max() <= 3e-2 * scale


def test_gradient_matches_reference(gpu_ctx):
    th = MO.init_params(4).astype(np.float32)
    sur = Surrogate(gpu_ctx, theta=th)
    X, t = _data(4096, 1)
    sur.set_data(X, t)
    g, loss = sur.grad(4096)
    L64, g64 = MO.loss_and_grad(th.astype(np.float64), X.astype(np.float64), t.astype(np.float64))
    assert abs(loss - L64) <= 2e-2 * L64
    errs = {}
    for name, (o, n) in {"W1": (0, 768), "b1": (768, 128), "W2": (896, 16384), "b2": (17280, 128),
                         "w3": (17408, 128), "b3": (17536, 1)}.items():
        a, b = g[o:o + n], g64[o:o + n]
        errs[name] = np.linalg.norm(a - b) / np.linalg.norm(b)
    print("gradient relative errors vs FP64:", errs)
    assert max(errs.values()) <= 5e-2


def test_adam_step_and_training(gpu_ctx):
    th = MO.init_params(5).astype(np.float32)
    sur = Surrogate(gpu_ctx, theta=th)
    X, t = _data(1 << 14, 2)
    sur.set_data(X, t)
    g, _ = sur.grad(4096)
    sur.train(1, batch=4096, lr=1e-3)
    th1, steps = sur.params()
    assert steps == 1
    want, _, _ = MO.adam_step(th.astype(np.float64), 0, 0, g.astype(np.float64), 1, lr=1e-3)
    assert np.allclose(th1, want, rtol=0, atol=2e-6)
    losses = sur.train(300, batch=4096, lr=1e-3, record_loss=True)
    print("loss first/last:", losses[0], losses[-1])
    assert losses[-1] < 0.2 * losses[0]


def test_training_step_follows_new_data(gpu_ctx):
    """Training with one data set, then a larger one (new device buffers) at
    the same batch and rate: the captured training step is rebuilt, so the
    next Adam step is the one of the new data's gradient."""
    th = MO.init_params(6).astype(np.float32)
    sur = Surrogate(gpu_ctx, theta=th)
    X, t = _data(1 << 12, 3)
    sur.set_data(X, t)
    sur.train(3, batch=4096, lr=1e-3)
    sur = Surrogate(gpu_ctx, theta=th)
    X2, t2 = _data(1 << 15, 4)
    sur.set_data(X2, t2)
    g, _ = sur.grad(4096)
    sur.train(1, batch=4096, lr=1e-3)
    th1, steps = sur.params()
    assert steps == 1
    want, _, _ = MO.adam_step(th.astype(np.float64), 0, 0, g.astype(np.float64), 1, lr=1e-3)
    assert np.allclose(th1, want, rtol=0, atol=2e-6)
