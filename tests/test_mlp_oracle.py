"""Pins the MLP oracle (oracle/mlp_ref.py) on the CPU: its gradient against
central finite differences of its own loss, and closed forms."""
import numpy as np

import oracle.mlp_ref as ref


def _params(n_in, seed, scale=0.1):
    rng = np.random.default_rng(seed)
    return rng.normal(0, scale, ref.num_params(n_in))


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    n_in, n = 5, 7
    p = _params(n_in, 4)
    x = rng.normal(0, 1, (n, n_in))
    y = rng.normal(0, 1, n)
    loss, g = ref.loss_grad(p, x, y)
    h = 1e-6
    idx = rng.choice(p.size, 300, replace=False)
    # include every bias and output weight, where a mistake would hide least
    idx = np.unique(np.concatenate([idx, np.arange(p.size - (ref.H * ref.OUT + ref.OUT), p.size)]))
    for i in idx:
        q = p.copy()
        q[i] += h
        lp, _ = ref.loss_grad(q, x, y)
        q[i] -= 2 * h
        lm, _ = ref.loss_grad(q, x, y)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_closed_form_with_only_output_bias():
    """All weights zero: out = b3 for every row, so the loss and its bias
    gradient have closed forms."""
    n_in = 3
    p = np.zeros(ref.num_params(n_in))
    mu, s = 0.75, -0.4
    p[-2:] = [mu, s]
    y = np.array([1.0, -2.0, 0.5, 3.0])
    x = np.ones((4, n_in))
    loss, g = ref.loss_grad(p, x, y)
    r = y - mu
    w = np.exp(-2 * s)
    assert np.isclose(loss, np.mean(0.5 * r * r * w + s))
    assert np.isclose(g[-2], np.mean(-r * w))
    assert np.isclose(g[-1], np.mean(1 - r * r * w))
    assert np.all(g[:-2] == 0)  # ReLU of zero pre-activations passes no gradient


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.00390625, 1.0078125, 3.0e-3, -2.5], dtype=np.float32)
    out = ref.bf16(vals)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties go to the even mantissa (1.0)
    assert out[0] == 1.0 and out[1] == 1.0 and out[2] == np.float32(1.0078125)
    assert abs(out[3] - 3.0e-3) < 3.0e-3 * 2 ** -8 and out[4] == -2.5


def test_adam_weight_decay_pulls_towards_zero():
    """With a zero data gradient (all weights zero, targets at the output bias's
    optimum) only the decay term moves the parameters: each Adam step moves a
    parameter by about lr towards zero (the first step exactly lr * sign)."""
    n_in = 2
    p = np.zeros(ref.num_params(n_in))
    p[:n_in * ref.H] = 0.5  # W1 only; W2 = 0 makes every gradient of W1 zero
    # targets equal to mu = 0 with s = 0 give a zero output-bias gradient for mu,
    # and d/ds = 1 - r^2 = 1 at r = 0; keep it by comparing W1 alone
    x = np.ones((3, n_in))
    y = np.zeros(3)
    out, _ = ref.adam_full_batch(p, x, y, steps=1, lr=1e-3, wd=0.1)
    assert np.allclose(out[:n_in * ref.H], 0.5 - 1e-3, atol=1e-9)
    out0, _ = ref.adam_full_batch(p, x, y, steps=1, lr=1e-3, wd=0.0)
    assert np.allclose(out0[:n_in * ref.H], 0.5)
