"""CPU oracle of the cardinality estimator's network. TEST INFRASTRUCTURE ONLY:
the tests use it to check the GPU forward, backward and BF16 inference; the
product library never calls it.

Network (float64 here): h1 = relu(x W1 + b1), h2 = relu(h1 W2 + b2),
out = h2 W3 + b3 with out = (mu, s); loss per row
0.5 (y - mu)^2 exp(-2 s) + s, averaged over the rows. Parameters are one flat
array W1 (in x H), b1, W2 (H x H), b2, W3 (H x 2), b3, row-major.

The reference has no counterpart (it estimates with samples and the
max-of-inputs rule, engine/query/QueryGraph.cpp:53-68); this oracle is pinned
by finite differences of its own loss and by closed forms (tests/test_mlp_oracle.py).
"""
import numpy as np

H = 256
OUT = 2


def split(params, n_in):
    p = np.asarray(params, dtype=np.float64)
    o = 0
    shapes = [(n_in, H), (H,), (H, H), (H,), (H, OUT), (OUT,)]
    out = []
    for s in shapes:
        n = int(np.prod(s))
        out.append(p[o:o + n].reshape(s))
        o += n
    assert o == p.size, (o, p.size)
    return out


def num_params(n_in):
    return n_in * H + H + H * H + H + H * OUT + OUT


def forward(params, x):
    w1, b1, w2, b2, w3, b3 = split(params, x.shape[1])
    h1 = np.maximum(x @ w1 + b1, 0.0)
    h2 = np.maximum(h1 @ w2 + b2, 0.0)
    return h1, h2, h2 @ w3 + b3


def loss_grad(params, x, y):
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    n = x.shape[0]
    w1, b1, w2, b2, w3, b3 = split(params, x.shape[1])
    h1, h2, out = forward(params, x)
    mu, s = out[:, 0], out[:, 1]
    r = y - mu
    w = np.exp(-2.0 * s)
    loss = np.mean(0.5 * r * r * w + s)
    dout = np.stack([-r * w, 1.0 - r * r * w], axis=1) / n
    gw3 = h2.T @ dout
    gb3 = dout.sum(0)
    d2 = (dout @ w3.T) * (h2 > 0)
    gw2 = h1.T @ d2
    gb2 = d2.sum(0)
    d1 = (d2 @ w2.T) * (h1 > 0)
    gw1 = x.T @ d1
    gb1 = d1.sum(0)
    grads = np.concatenate([gw1.ravel(), gb1, gw2.ravel(), gb2, gw3.ravel(), gb3])
    return loss, grads


def bf16(a):
    """Round float32 values to bfloat16 (nearest even) and back."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def forward_bf16(params, x):
    """Inference as the BF16 path computes it: BF16 inputs, weights and hidden
    activations, FP32 (here float64) accumulation; W3/b3 stay FP32."""
    w1, b1, w2, b2, w3, b3 = split(params, x.shape[1])
    xb = bf16(x).astype(np.float64)
    h1 = bf16(np.maximum(xb @ bf16(w1).astype(np.float64) + b1, 0.0)).astype(np.float64)
    h2 = bf16(np.maximum(h1 @ bf16(w2).astype(np.float64) + b2, 0.0)).astype(np.float64)
    return h2 @ w3 + b3


def adam_full_batch(params, x, y, steps, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """`steps` Adam updates (Kingma & Ba, 2015, Algorithm 1) on the gradient
    of the whole set plus wd x params (L2 weight decay); returns the
    parameters and the loss before each step."""
    p = np.asarray(params, dtype=np.float64).copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    losses = []
    for t in range(1, steps + 1):
        loss, g = loss_grad(p, x, y)
        losses.append(loss)
        g = g + wd * p
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        p -= lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, losses
