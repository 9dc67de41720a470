"""Pins the CPU oracle of the warm-start network: its gradients (without bf16
rounding) match central finite differences of its loss, and bf16 rounding is
round-to-nearest-even on known values."""
import numpy as np

import mlp_oracle as mo


def test_bf16_rounding_known_values():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.14159265], dtype=np.float32)
    r = mo.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0            # tie rounds to even
    assert r[2] == np.float32(1.0 + 2 ** -7)      # above the tie rounds up
    assert r[3] == -2.5 and abs(r[4] - 3.140625) < 1e-7


def test_mlp_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    shapes = [(5, 6), (5, 5), (4, 5)]
    P = sum(o * i + o for o, i in shapes)
    p = rng.standard_normal(P) * 0.5
    x = rng.standard_normal((7, 6))
    t = rng.standard_normal((7, 4))
    loss, g, _ = mo.forward_backward(p, shapes, x, t, round_bf16=False)
    h = 1e-6
    for k in range(P):
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (mo.forward_backward(pp, shapes, x, t, False)[0] - mo.forward_backward(pm, shapes, x, t, False)[0]) / (2 * h)
        assert abs(g[k] - fd) < 1e-6 * max(1.0, abs(fd))
