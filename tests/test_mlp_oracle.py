"""Pins the CPU oracle of the surrogate network (no GPU): its analytic gradient
against central finite differences of its own float64 loss, a single-layer case
with a closed form, the BF16 rounding against hand-checked bit patterns, and
the BF16 gradient against the float64 one."""
import numpy as np

from oracle import mlp_oracle as mo


def test_bf16_rounding_known_values():
    # BF16 keeps 7 fraction bits: the spacing above 1.0 is 2^-7
    x = np.array([1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e38, 0.0],
                 np.float32)
    # exact; exact; tie -> even 1.0; tie -> even 1+2^-6; above the midpoint -> 1+2^-7;
    # exact; 0x7F61B1E6 rounds up; zero
    want = [0x3F80, 0x3F81, 0x3F80, 0x3F82, 0x3F81, 0xC020, 0x7F62, 0x0000]
    assert mo.bf16_bits(x).tolist() == want


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    n_in, hid, n_out, rows = 5, 8, 2, 7
    p = mo.init_params(n_in, hid, n_out, 1).astype(np.float64)
    p += rng.normal(0, 0.05, p.size)              # non-zero biases
    x = rng.normal(size=(rows, n_in))
    t = rng.normal(size=(rows, n_out))
    loss, g = mo.loss_grad(p, x, t, n_in, hid, n_out, quant=False)
    eps = 1e-6
    idx = rng.choice(p.size, 60, replace=False)
    for k in idx:
        e = np.zeros_like(p)
        e[k] = eps
        lp, _ = mo.loss_grad(p + e, x, t, n_in, hid, n_out, quant=False)
        lm, _ = mo.loss_grad(p - e, x, t, n_in, hid, n_out, quant=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_linear_output_layer_closed_form():
    # all hidden units active with identity-like weights: Y = X W1 W2 W3 W4 + ...
    n_in, hid, n_out = 2, 2, 1
    layers = [(np.eye(2), np.array([5.0, 5.0])), (np.eye(2), np.zeros(2)), (np.eye(2), np.zeros(2)),
              (np.array([[2.0], [-1.0]]), np.array([0.5]))]
    p = mo.pack(layers)
    x = np.array([[1.0, 2.0], [3.0, -1.0]])
    y, _ = mo.forward(p, x, n_in, hid, n_out, quant=False)
    # hidden = x + 5 (all positive), y = 2 h0 - h1 + 0.5
    want = 2 * (x[:, 0] + 5) - (x[:, 1] + 5) + 0.5
    assert np.allclose(y[:, 0], want)
    t = np.zeros((2, 1))
    loss, g = mo.loss_grad(p, x, t, n_in, hid, n_out, quant=False)
    assert np.isclose(loss, np.mean(want ** 2))
    # dL/db4 = mean over rows of 2 y
    assert np.isclose(g[-1], np.sum(2 * want / 2))


def test_bf16_gradient_close_to_exact():
    rng = np.random.default_rng(3)
    n_in, hid, n_out, rows = 24, 64, 2, 256
    p = mo.init_params(n_in, hid, n_out, 2)
    x = rng.normal(size=(rows, n_in)).astype(np.float32)
    t = rng.normal(size=(rows, n_out)).astype(np.float32)
    _, gq = mo.loss_grad(p, x, t, n_in, hid, n_out, quant=True)
    _, ge = mo.loss_grad(p.astype(np.float64), x, t, n_in, hid, n_out, quant=False)
    assert np.linalg.norm(gq - ge) / np.linalg.norm(ge) < 3e-2


def test_weighted_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(7)
    n_in, hid, n_out, rows = 4, 6, 2, 5
    p = mo.init_params(n_in, hid, n_out, 3).astype(np.float64)
    p += rng.normal(0, 0.05, p.size)                 # non-zero biases
    x = rng.normal(size=(rows, n_in))
    t = rng.normal(size=(rows, n_out))
    loss, g = mo.loss_grad(p, x, t, n_in, hid, n_out, quant=False, w_over=5.0)
    y, _ = mo.forward(p, x, n_in, hid, n_out, quant=False)
    d = y - t
    assert np.isclose(loss, np.mean(np.where(d > 0, 5.0, 1.0) * d * d))
    eps = 1e-6
    for k in rng.choice(p.size, 40, replace=False):
        e = np.zeros_like(p)
        e[k] = eps
        fd = (mo.loss_grad(p + e, x, t, n_in, hid, n_out, quant=False, w_over=5.0)[0]
              - mo.loss_grad(p - e, x, t, n_in, hid, n_out, quant=False, w_over=5.0)[0]) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])
