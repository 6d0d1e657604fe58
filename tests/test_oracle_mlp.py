"""The FP64 MLP oracle pinned by central finite differences (no GPU)."""
import numpy as np

from oracle import mlp as M


def _setup(seed=0, F=4, H=8, L=2, n=7):
    rng = np.random.default_rng(seed)
    nf = 2 * F
    n_params = nf * H + H + (L - 1) * (H * H + H) + H + 1
    theta = rng.normal(size=n_params) * 0.5
    B = rng.normal(size=(F, 3)) * 2
    x = rng.uniform(-1, 1, size=(n, 3))
    y = rng.normal(size=n)
    return theta, B, x, y, H, L


def test_parameter_gradient_matches_finite_differences():
    theta, B, x, y, H, L = _setup()
    loss, g = M.gradients(theta, B, x, y, H, L)
    eps = 1e-6
    num = np.zeros_like(theta)
    for i in range(len(theta)):
        t1 = theta.copy(); t1[i] += eps
        t2 = theta.copy(); t2[i] -= eps
        num[i] = (M.gradients(t1, B, x, y, H, L)[0] - M.gradients(t2, B, x, y, H, L)[0]) / (2 * eps)
    np.testing.assert_allclose(g, num, atol=1e-8)


def test_input_gradient_matches_finite_differences():
    theta, B, x, y, H, L = _setup(1, L=3)
    c, gx = M.input_gradient(theta, B, x, H, L)
    eps = 1e-6
    for k in range(3):
        x1 = x.copy(); x1[:, k] += eps
        x2 = x.copy(); x2[:, k] -= eps
        num = (M.forward(theta, B, x1, H, L)[0] - M.forward(theta, B, x2, H, L)[0]) / (2 * eps)
        np.testing.assert_allclose(gx[:, k], num, atol=1e-8)


def test_bf16_rounding_is_round_to_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e38], np.float32)
    r = M.to_bf16(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0                    # halfway, ties to even (mantissa ...0)
    assert r[2] == 1.0 + 2 ** -7          # above halfway rounds up
    assert r[3] == -2.5
