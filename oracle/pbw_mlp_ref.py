"""FP64 CPU check of the prefilter network (test infrastructure only).

Restates the forward and backward of the GPU network at its rounding points:
inputs and weights are BF16, the ReLU outputs h1 and h2 are rounded to BF16,
the backpropagated dh2 and dh1 are rounded to BF16 where the GPU feeds them to
the matrix cores; everything else is exact FP64 here (FP32 on the GPU).

    a1 = x W1 + b1, h1 = bf16(relu(a1)); a2 = h1 W2 + b2, h2 = bf16(relu(a2));
    z = h2 W3 + b3; loss = mean_i w_i BCE(sigmoid(z_i), y_i)

Parameter layout: W1[32][128] | b1[128] | W2[128][128] | b2[128] | W3[128] | b3.

`rounding=False` drops every BF16 rounding: the exact FP64 network, which the
tests pin against torch's autograd and central finite differences.  `adam_step`
is Adam (Kingma & Ba 2015, Algorithm 1) with bias-corrected moments, as the GPU
applies it to its FP32 master parameters.
"""
import numpy as np

IN, H = 32, 128
OFF_W1, OFF_B1 = 0, IN * H
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
NP = OFF_B3 + 1


def bf16(a):
    """Round to BF16 (nearest even) and return the value as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _exact(a):
    return np.asarray(a, dtype=np.float64)


def unpack(theta, rounding=True):
    r = bf16 if rounding else _exact
    t = np.asarray(theta, dtype=np.float32 if rounding else np.float64)
    W1 = r(t[OFF_W1:OFF_B1]).reshape(IN, H)
    W2 = r(t[OFF_W2:OFF_B2]).reshape(H, H)
    W3 = r(t[OFF_W3:OFF_B3])
    return W1, t[OFF_B1:OFF_W2].astype(np.float64), W2, t[OFF_B2:OFF_W3].astype(np.float64), W3, float(t[OFF_B3])


def forward(theta, x, rounding=True):
    """x: float array [n, 32] (already BF16 values). Returns z and the intermediates."""
    r = bf16 if rounding else _exact
    W1, b1, W2, b2, W3, b3 = unpack(theta, rounding)
    a1 = x @ W1 + b1
    h1 = r(np.maximum(a1, 0.0))
    a2 = h1 @ W2 + b2
    h2 = r(np.maximum(a2, 0.0))
    z = h2 @ W3 + b3
    return z, (a1, h1, a2, h2)


def loss_of(theta, x, y, pos_weight=1.0, rounding=True):
    z, _ = forward(theta, x, rounding)
    w = np.where(y > 0.5, pos_weight, 1.0)
    return np.sum(w * (np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z))))) / len(x)


def adam_step(theta, m, v, g, t, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update at step t >= 1; returns (theta, m, v) as new float64 arrays."""
    m = beta1 * np.asarray(m, np.float64) + (1.0 - beta1) * g
    v = beta2 * np.asarray(v, np.float64) + (1.0 - beta2) * g * g
    mh = m / (1.0 - beta1 ** t)
    vh = v / (1.0 - beta2 ** t)
    return np.asarray(theta, np.float64) - lr * mh / (np.sqrt(vh) + eps), m, v


def gradient(theta, x, y, pos_weight=1.0, rounding=True):
    """Mean weighted BCE-with-logits and its gradient (layout of theta)."""
    r = bf16 if rounding else _exact
    W1, b1, W2, b2, W3, b3 = unpack(theta, rounding)
    z, (a1, h1, a2, h2) = forward(theta, x, rounding)
    n = len(x)
    w = np.where(y > 0.5, pos_weight, 1.0)
    loss = np.sum(w * (np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z))))) / n
    dz = w * (1.0 / (1.0 + np.exp(-z)) - y) / n
    g = np.zeros(NP)
    g[OFF_W3:OFF_B3] = h2.T @ dz
    g[OFF_B3] = dz.sum()
    dh2 = r(np.outer(dz, W3) * (a2 > 0))
    g[OFF_W2:OFF_B2] = (h1.T @ dh2).ravel()
    g[OFF_B2:OFF_W3] = dh2.sum(0)
    dh1 = r((dh2 @ W2.T) * (a1 > 0))
    g[OFF_W1:OFF_B1] = (x.T @ dh1).ravel()
    g[OFF_B1:OFF_W2] = dh1.sum(0)
    return g, loss
