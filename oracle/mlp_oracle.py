"""CPU oracle of the learned preconditioner network -- TEST INFRASTRUCTURE ONLY.

FP64 numpy forward/backward of the 24 -> 128 -> 128 -> 1 SiLU network with
bounded factor g = 1 + G_AMP tanh(out) (G_AMP = 1/2) and loss mean((g m - y)^2).  With emulate=True every matrix-product operand
is rounded to bfloat16 (round to nearest even) and feature standardisation is
done in FP32, at the same points as the GPU kernels, so the two agree to
accumulation-order differences.  The backward pass is pinned by finite
differences (tests/test_mlp_oracle.py).
"""
import numpy as np

F, H = 24, 128
G_AMP = 0.5
NPARAM = F * H + H + H * H + H + H + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + F * H].reshape(F, H); o += F * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H].reshape(H, 1); o += H
    b3 = p[o:o + 1]
    return W1, b1, W2, b2, W3, b3


def bf16(x):
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _mm(a, b, emulate):
    if emulate:
        return bf16(a) @ bf16(b)
    return a @ b


def _silu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return z * s, s


def forward(p, X, mean, std, emulate=False):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    if emulate:
        Xn = ((np.asarray(X, np.float32) - np.asarray(mean, np.float32)) / np.asarray(std, np.float32)).astype(np.float64)
    else:
        Xn = (np.asarray(X, np.float64) - mean) / std
    Z1 = _mm(Xn, W1, emulate) + b1
    H1, S1 = _silu(Z1)
    Z2 = _mm(H1, W2, emulate) + b2
    H2, S2 = _silu(Z2)
    out = (_mm(H2, W3, emulate) + b3)[:, 0]
    return out, (Xn, Z1, H1, S1, Z2, H2, S2)


def loss_grad(p, X, m, y, mean, std, emulate=False):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    out, (Xn, Z1, H1, S1, Z2, H2, S2) = forward(p, X, mean, std, emulate)
    n = X.shape[0]
    th = np.tanh(out)
    r = (1.0 + G_AMP * th) * m - y
    loss = float(np.mean(r * r))
    dout = (2.0 * r * m * G_AMP * (1.0 - th * th) / n)[:, None]
    gW3 = _mm(H2.T, dout, emulate)
    gb3 = dout.sum(axis=0)
    dH2 = _mm(dout, W3.T, emulate)
    dZ2 = dH2 * S2 * (1.0 + Z2 * (1.0 - S2))
    gW2 = _mm(H1.T, dZ2, emulate)
    gb2 = dZ2.sum(axis=0)
    dH1 = _mm(dZ2, W2.T, emulate)
    dZ1 = dH1 * S1 * (1.0 + Z1 * (1.0 - S1))
    gW1 = _mm(Xn.T, dZ1, emulate)
    gb1 = dZ1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g
