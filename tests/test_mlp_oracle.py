"""The MLP oracle's backward pass against central finite differences."""
import numpy as np

from oracle import mlp_oracle as mo


def test_oracle_gradient_is_finite_difference():
    rng = np.random.default_rng(0)
    p = rng.standard_normal(mo.NPARAM) * 0.1
    X = rng.standard_normal((40, mo.F))
    m, y = rng.standard_normal(40), rng.standard_normal(40)
    mu, sd = rng.standard_normal(mo.F) * 0.1, 1.0 + rng.random(mo.F)
    _, g = mo.loss_grad(p, X, m, y, mu, sd)
    # every block of the parameter vector: W1, b1, W2, b2, W3, b3
    blocks = [0, mo.F * mo.H, mo.F * mo.H + mo.H, mo.F * mo.H + mo.H + mo.H * mo.H,
              mo.F * mo.H + 2 * mo.H + mo.H * mo.H, mo.NPARAM - 1]
    idx = np.concatenate([blocks, rng.choice(mo.NPARAM, 30, replace=False)])
    h = 1e-6
    for i in idx:
        pp = p.copy()
        pp[i] += h
        lp, _ = mo.loss_grad(pp, X, m, y, mu, sd)
        pp[i] -= 2 * h
        lm, _ = mo.loss_grad(pp, X, m, y, mu, sd)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -9, -2.5, 3.0e-20], dtype=np.float32)
    r = mo.bf16(vals)
    assert r[0] == 1.0 and r[4] == -2.5
    # halfway 1 + 2^-8 rounds to even (1.0); 1 + 3*2^-9 rounds up to 1 + 2^-7
    assert mo.bf16(np.float32(1.0 + 2 ** -8)) == 1.0
    assert mo.bf16(np.float32(1.0 + 3 * 2 ** -9)) == 1.0 + 2 ** -7
    assert mo.bf16(np.float32(-2.5)) == -2.5
