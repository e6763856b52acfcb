"""Pins the float64 MLP oracle: its analytic gradient against central finite
differences of its own loss (the mathematics fixes the gradient)."""
import numpy as np

from oracle import mlp_oracle as M


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    P = M.H * M.D + M.H + M.H * M.H + 2 * M.H + 1
    p = rng.normal(scale=0.3, size=P)
    x = rng.normal(size=(20, M.D))
    t = rng.normal(size=20)
    _, g = M.loss_and_grad(p, x, t)
    idx = rng.choice(P, 200, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros(P)
        e[i] = h
        fd = (M.loss_and_grad(p + e, x, t)[0] - M.loss_and_grad(p - e, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), i


def test_forward_is_the_loss_argument():
    rng = np.random.default_rng(1)
    P = M.H * M.D + M.H + M.H * M.H + 2 * M.H + 1
    p = rng.normal(scale=0.3, size=P)
    x = rng.normal(size=(7, M.D))
    y = M.forward(p, x)
    loss, _ = M.loss_and_grad(p, x, y)
    assert loss == 0.0
