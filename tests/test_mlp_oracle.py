"""Pins the warm-start MLP oracle (oracle/mlp_oracle.cpp): its FP64 forward pass against the same
network assembled from torch's float64 linear layers and hard-swish and against closed forms (zero
weights, a single path through the layers, hard-swish's identity above 3 and zero below -3); its
analytic gradient against torch autograd and central finite differences; its BF16 rounding against the round-to-nearest-even
definition; and its Adam step against the closed form of the first step."""

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
        # hard-swish has kinks at -3 and 3; a small step keeps the central difference from
        # straddling one (the loss is a piecewise polynomial, exact to rounding between kinks)
        h = 1e-5 * max(1.0, abs(float(p[i])))
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


# ---- the forward pass against library routines and closed forms ----------------------------
def torch_net(p):
    """The network as the packing defines it (W row-major [out][in], oracle/mlp_oracle.cpp),
    built from torch's float64 linear layers and hard-swish."""
    import torch

    t = torch.tensor(np.asarray(p, np.float64), requires_grad=True)
    S = MLP_SLICES
    W1 = t[S["W1"]].reshape(64, 16)
    W2 = t[S["W2"]].reshape(64, 64)
    W3 = t[S["W3"]].reshape(64, 64)
    W4 = t[S["W4"]].reshape(1, 64)

    def f(x):
        F = torch.nn.functional
        h1 = F.hardswish(F.linear(x, W1, t[S["b1"]]))
        h2 = F.hardswish(F.linear(h1, W2, t[S["b2"]]))
        h3 = F.hardswish(F.linear(h2, W3, t[S["b3"]]))
        return F.linear(h3, W4, t[S["b4"]])[:, 0]

    return t, f


def test_forward_matches_torch_float64():
    """Layer order, weight layout and activation: the oracle's FP64 forward equals the same
    network assembled from torch.nn.functional.linear / hardswish."""
    import torch

    rng = np.random.default_rng(11)
    p = random_params(11)
    x = rng.standard_normal((40, 16)).astype(np.float32)
    _, f = torch_net(p)
    want = f(torch.tensor(x.astype(np.float64))).detach().numpy()
    got = mlp_forward(p, x, bf16=False)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


def test_gradient_matches_torch_autograd():
    import torch

    rng = np.random.default_rng(12)
    p = random_params(12)
    x = rng.standard_normal((24, 16)).astype(np.float32)
    t = rng.standard_normal(24).astype(np.float32)
    tp, f = torch_net(p)
    y = f(torch.tensor(x.astype(np.float64)))
    loss_t = ((y - torch.tensor(t.astype(np.float64))) ** 2).mean()
    loss_t.backward()
    loss, g = mlp_loss_grad(p, x, t, bf16=False)
    assert loss == pytest.approx(loss_t.item(), rel=1e-12)
    want = tp.grad.numpy()
    np.testing.assert_allclose(g, want, rtol=1e-10, atol=1e-12 * np.abs(want).max())


def test_forward_closed_forms():
    """All-zero weights give the output bias; with one path through the network (one unit per
    layer) the output is a(a(a(w x0 + c))) with hard-swish a(z) = z min(max(z + 3, 0), 6) / 6
    (a(1) = 2/3, a(-1) = -1/3), since a(0) = 0 keeps every other unit silent; a(z) = z above 3 and
    0 below -3."""
    n = orc().mlp_param_count()
    x = np.zeros((3, 16), np.float32)
    x[:, 0] = [0.5, -1.25, 2.0]
    p = np.zeros(n, np.float32)
    p[MLP_SLICES["b4"]] = 0.375
    np.testing.assert_array_equal(mlp_forward(p, x, bf16=False), [0.375] * 3)

    def hswish(z):
        return z * min(max(z + 3.0, 0.0), 6.0) / 6.0

    assert hswish(1.0) == pytest.approx(2.0 / 3.0) and hswish(-1.0) == pytest.approx(-1.0 / 3.0)

    w, c = 1.5, -0.25
    p = np.zeros(n, np.float32)
    p[MLP_SLICES["W1"].start + 0 * 16 + 0] = w  # W1[0][0]
    p[MLP_SLICES["b1"].start] = c
    p[MLP_SLICES["W2"].start] = 1.0  # W2[0][0]
    p[MLP_SLICES["W3"].start] = 1.0  # W3[0][0]
    p[MLP_SLICES["W4"].start] = 1.0  # W4[0]
    want = [hswish(hswish(hswish(w * float(x0) + c))) for x0 in x[:, 0]]
    np.testing.assert_allclose(mlp_forward(p, x, bf16=False), want, rtol=1e-14)
    # a unit in the second slot of layer 1 that W2 reads from its column 1 (W2[0][1])
    p = np.zeros(n, np.float32)
    p[MLP_SLICES["W1"].start + 1 * 16 + 3] = w  # W1[1][3]: hidden unit 1 reads input 3
    p[MLP_SLICES["W2"].start + 0 * 64 + 1] = 1.0  # W2[0][1]
    p[MLP_SLICES["W3"].start] = 1.0
    p[MLP_SLICES["W4"].start] = 1.0
    x2 = np.zeros((2, 16), np.float32)
    x2[:, 3] = [0.75, -2.0]
    want = [hswish(hswish(hswish(w * float(v)))) for v in x2[:, 3]]
    np.testing.assert_allclose(mlp_forward(p, x2, bf16=False), want, rtol=1e-14)
    # pre-activations above 3: hard-swish is the identity; below -3 it is 0
    p = np.zeros(n, np.float32)
    p[MLP_SLICES["W1"].start] = 1.0
    p[MLP_SLICES["W2"].start] = 1.0
    p[MLP_SLICES["W3"].start] = 1.0
    p[MLP_SLICES["W4"].start] = 1.0
    x3 = np.zeros((1, 16), np.float32)
    x3[0, 0] = 60.0
    assert mlp_forward(p, x3, bf16=False)[0] == pytest.approx(60.0, rel=1e-20, abs=1e-12)
    x3[0, 0] = -7.5
    assert mlp_forward(p, x3, bf16=False)[0] == 0.0


def test_gradient_at_the_kinks_follows_torch():
    """Hard-swish's derivative where it is not differentiable: a single path through the network
    whose first pre-activation is exactly -3, 0, 3 or 4 (the other layers' identity weights keep
    the later pre-activations away from the kinks or at 0).  The oracle's gradient equals torch's
    autograd, which takes 0 up to and including -3, z / 3 + 1/2 between and 1 from 3 on."""
    import torch

    n = orc().mlp_param_count()
    p = np.zeros(n, np.float32)
    p[MLP_SLICES["W1"].start] = 1.0  # W1[0][0]
    p[MLP_SLICES["W2"].start] = 1.0
    p[MLP_SLICES["W3"].start] = 1.0
    p[MLP_SLICES["W4"].start] = 1.0
    for z0 in (-3.0, 0.0, 3.0, 4.0):
        x = np.zeros((1, 16), np.float32)
        x[0, 0] = z0
        t = np.array([0.5], np.float32)
        tp, f = torch_net(p)
        y = f(torch.tensor(x.astype(np.float64)))
        ((y - torch.tensor(t.astype(np.float64))) ** 2).mean().backward()
        _, g = mlp_loss_grad(p, x, t, bf16=False)
        np.testing.assert_allclose(g, tp.grad.numpy(), rtol=1e-12, atol=1e-14, err_msg=f"z = {z0}")
