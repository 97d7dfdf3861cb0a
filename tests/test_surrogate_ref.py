"""Pin the CPU reference of the closure surrogate (tests/surrogate_ref.py):
its hand-written backward equals torch autograd of the same network in
float64, autograd agrees with central finite differences, the periodic
padding wraps, and the Adam update matches its closed form."""
import numpy as np
import pytest
import torch

from tests import surrogate_ref as R


def autograd_loss(theta, feat, target, perx, pery):
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    layers = []
    o = 0
    for co, _, ci in R.layer_shapes():
        w = th[o:o + co * 9 * ci].reshape(co, 9, ci)
        o += co * 9 * ci
        b = th[o:o + co]
        o += co
        layers.append((w, b))
    x = torch.as_tensor(feat, dtype=torch.float64)
    for li, (w, b) in enumerate(layers):
        x = R.conv(x, w, b, perx, pery)
        if li < 3:
            x = torch.relu(x)
    loss = ((x - torch.as_tensor(target)) ** 2).mean()
    loss.backward()
    return float(loss.detach()), th.grad.numpy()


def data(seed, B=2, X=6, Y=5):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(B, 4, X, Y)), rng.normal(size=(B, 2, X, Y))


@pytest.mark.parametrize("perx,pery", [(True, False), (False, True), (True, True), (False, False)])
def test_manual_backward_equals_autograd(perx, pery):
    theta = R.init_params(3)
    f, t = data(1)
    l1, g1, _ = R.loss_and_grad(theta, f, t, perx, pery, bf=False)
    l2, g2 = autograd_loss(theta, f, t, perx, pery)
    assert abs(l1 - l2) <= 1e-12 * abs(l2)
    assert np.max(np.abs(g1 - g2)) <= 1e-12 * np.max(np.abs(g2))


def test_autograd_matches_finite_differences():
    theta = R.init_params(5)
    f, t = data(2, B=1, X=5, Y=4)
    _, g = autograd_loss(theta, f, t, True, False)
    rng = np.random.default_rng(0)
    idx = rng.choice(theta.size, 25, replace=False)
    h = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (autograd_loss(tp, f, t, True, False)[0] - autograd_loss(tm, f, t, True, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), i


def test_periodic_padding_wraps():
    x = torch.arange(12, dtype=torch.float64).reshape(1, 1, 3, 4)
    p = R.pad(x, True, False)
    assert torch.equal(p[0, 0, 0, 1:-1], x[0, 0, -1]) and torch.equal(p[0, 0, -1, 1:-1], x[0, 0, 0])
    assert torch.all(p[0, 0, :, 0] == 0) and torch.all(p[0, 0, :, -1] == 0)


def test_adam_closed_form():
    th, g = np.array([1.0, -2.0]), np.array([0.5, -0.25])
    t1, m, v = R.adam(th, np.zeros(2), np.zeros(2), g, 1, 1e-3)
    # first step: mhat = g, vhat = g^2 -> step = lr * sign(g) (up to eps)
    assert np.allclose(t1, th - 1e-3 * np.sign(g), atol=1e-10)


def test_bf16_rounding_is_small():
    theta = R.init_params(7)
    f, t = data(3)
    lb, gb, _ = R.loss_and_grad(theta, f, t, True, False, bf=True)
    l, g, _ = R.loss_and_grad(theta, f, t, True, False, bf=False)
    assert abs(lb - l) <= 2e-2 * l
    assert np.linalg.norm(gb - g) <= 5e-2 * np.linalg.norm(g)
