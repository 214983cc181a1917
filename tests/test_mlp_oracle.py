"""Pins the warm-start MLP oracle (oracle/mlp_oracle.cpp): its analytic gradient against
central finite differences of its own forward pass in FP64, its BF16 rounding against the
round-to-nearest-even definition, and its Adam step against the closed form of the first step."""
import numpy as np
import pytest

from oracle_lib import MLP_SLICES, mlp_adam, mlp_forward, mlp_loss_grad, orc


def random_params(seed=0, scale=0.3):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(orc().mlp_param_count()) * scale).astype(np.float32)


def test_param_count():
    assert orc().mlp_param_count() == 64 * 16 + 64 + 2 * (64 * 64 + 64) + 64 + 1 == 9473
    assert MLP_SLICES["b4"].stop == 9473


def test_bf16_round_to_nearest_even():
    b = orc().mlp_bf16
    assert b(1.0) == 1.0
    assert b(1.0 + 2**-8) == 1.0  # exact tie -> even (mantissa ...0)
    assert b(1.0 + 3 * 2**-8) == 1.0 + 2**-6  # tie -> even (rounds up)
    assert b(1.0 + 2**-8 + 2**-20) == 1.0 + 2**-7  # above the tie
    assert b(-2.5) == -2.5
    x = np.float32(3.14159265)
    assert abs(b(float(x)) - float(x)) <= 2**-7 * 2


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = random_params(1)
    x = rng.standard_normal((24, 16)).astype(np.float32)
    t = rng.standard_normal(24).astype(np.float32)
    loss, g = mlp_loss_grad(p, x, t, bf16=False)
    idx = np.concatenate([rng.choice(np.arange(s.start, s.stop), size=min(12, s.stop - s.start), replace=False) for s in MLP_SLICES.values()])
    for i in idx:
        h = 1e-3 * max(1.0, abs(float(p[i])))
        pp = p.astype(np.float64).copy()
        # perturb in float32 storage exactly: use the representable neighbours
        p1 = p.copy()
        p2 = p.copy()
        p1[i] = np.float32(p[i] + h)
        p2[i] = np.float32(p[i] - h)
        d = float(p1[i]) - float(p2[i])
        l1 = np.mean((mlp_forward(p1, x, bf16=False) - t.astype(np.float64)) ** 2)
        l2 = np.mean((mlp_forward(p2, x, bf16=False) - t.astype(np.float64)) ** 2)
        fd = (l1 - l2) / d
        assert g[i] == pytest.approx(fd, rel=2e-4, abs=1e-7), f"param {i}"
    assert loss == pytest.approx(np.mean((mlp_forward(p, x, bf16=False) - t) ** 2), rel=1e-12)


def test_bf16_forward_close_to_fp64():
    rng = np.random.default_rng(2)
    p = random_params(2)
    x = rng.standard_normal((64, 16)).astype(np.float32)
    y64 = mlp_forward(p, x, bf16=False)
    y16 = mlp_forward(p, x, bf16=True)
    assert np.abs(y16 - y64).max() < 0.05 * np.abs(y64).max()


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(3)
    n = orc().mlp_param_count()
    p = rng.standard_normal(n).astype(np.float32)
    g = rng.standard_normal(n).astype(np.float32)
    z = np.zeros(n, np.float32)
    lr, eps = 1e-3, 1e-8
    p1, m1, v1 = mlp_adam(p, z, z, g, 1, lr)
    # step 1: m_hat = g, v_hat = g^2 -> p - lr * g / (|g| + eps)
    expect = p.astype(np.float64) - lr * g / (np.abs(g.astype(np.float64)) + eps)
    np.testing.assert_allclose(p1, expect, rtol=0, atol=1e-6)
    b1, b2 = float(np.float32(0.9)), float(np.float32(0.999))  # the betas arrive as fp32
    np.testing.assert_allclose(m1, (1 - b1) * g.astype(np.float64), rtol=1e-6)
    np.testing.assert_allclose(v1, (1 - b2) * g.astype(np.float64) ** 2, rtol=1e-6)
