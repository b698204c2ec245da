"""The atom-pair network's loss and gradients in float64 numpy (test
infrastructure: the checker of calchem_amd/csrc/mlp.hip).

y = W3 s(W2 s(W1 x + b1) + b2) + b3 with s(z) = z / (1 + e^-z); loss
L = sum m (y - t)^2 / sum m over the batch; gradients by the chain rule.
"""
import numpy as np


def unpack(p, d_in, h, d_out):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    out = []
    for shape in ((h, d_in), (h,), (h, h), (h,), (d_out, h), (d_out,)):
        n = int(np.prod(shape))
        out.append(p[o:o + n].reshape(shape))
        o += n
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, x, d_in, h, d_out):
    W1, b1, W2, b2, W3, b3 = unpack(p, d_in, h, d_out)
    z1 = x @ W1.T + b1
    a1 = silu(z1)
    z2 = a1 @ W2.T + b2
    a2 = silu(z2)
    return a2 @ W3.T + b3, (z1, a1, z2, a2)


def loss_grad(p, x, t, m, d_in, h, d_out):
    W1, b1, W2, b2, W3, b3 = unpack(p, d_in, h, d_out)
    x, t, m = (np.asarray(a, dtype=np.float64) for a in (x, t, m))
    y, (z1, a1, z2, a2) = forward(p, x, d_in, h, d_out)
    cnt = m.sum()
    r = m * (y - t)
    loss = float(np.sum(r * r) / cnt)
    d3 = 2.0 * r * m / cnt
    gW3, gb3 = d3.T @ a2, d3.sum(axis=0)
    d2 = (d3 @ W3) * silu_grad(z2)
    gW2, gb2 = d2.T @ a1, d2.sum(axis=0)
    d1 = (d2 @ W2) * silu_grad(z1)
    gW1, gb1 = d1.T @ x, d1.sum(axis=0)
    return loss, np.concatenate([g.ravel() for g in (gW1, gb1, gW2, gb2, gW3, gb3)])
