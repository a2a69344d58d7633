"""CPU pins of the FP64 surrogate reference: finite-difference gradient and
the Adam step's closed form on a one-parameter problem."""
import numpy as np

from oracle import mlp_oracle as MO


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    th = MO.init_params(1) * 0.5
    X = rng.normal(size=(32, 6))
    t = rng.normal(size=32)
    L0, g = MO.loss_and_grad(th, X, t)
    idx = rng.choice(MO.NPARAM, 60, replace=False)
    idx = np.concatenate([idx, [MO.NPARAM - 1, 0, 768, 768 + 128 + 5]])
    for j in idx:
        e = np.zeros(MO.NPARAM)
        h = 1e-6 * max(1.0, abs(th[j]))
        e[j] = h
        Lp, _ = MO.loss_and_grad(th + e, X, t)
        Lm, _ = MO.loss_and_grad(th - e, X, t)
        fd = (Lp - Lm) / (2 * h)
        assert abs(fd - g[j]) <= 1e-6 * max(1.0, abs(g[j])), (j, fd, g[j])


def test_adam_first_step_is_lr_times_sign():
    th = np.array([1.0, -2.0, 3.0])
    g = np.array([0.5, -0.25, 1e-3])
    t2, m, v = MO.adam_step(th, np.zeros(3), np.zeros(3), g, 1, lr=1e-3)
    # bias-corrected first step moves every parameter by lr * g/(|g|+eps')
    assert np.allclose(t2, th - 1e-3 * np.sign(g), rtol=0, atol=1e-8)


def test_bf16_rounding():
    assert MO.bf16_round(1.0) == 1.0
    assert MO.bf16_round(1.0 + 2 ** -9) == 1.0           # ties to even
    assert MO.bf16_round(1.0 + 3 * 2 ** -9) == 1.0 + 2 ** -7
    assert MO.bf16_round(-3.14159) == -3.140625


def test_bracket_bins_fixed_width_inside_grid():
    from amps_amd import surrogate as SG
    pb = np.array([0, 1, 2, 3, 10, 44, 45, 46, 47])
    lo, hi = SG.bracket_bins(pb, 48, half=3)
    assert np.all(hi - lo == 6) and lo.min() == 0 and hi.max() == 47
    inner = (pb >= 3) & (pb <= 44)
    assert np.array_equal(lo[inner], pb[inner] - 3)
    assert np.all((pb >= lo) & (pb <= hi))
