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


def _torch_net(p, shapes, x, t):
    """The same network written with torch's own layers (float64, CPU): linear layers,
    exact erf GELU, mean-squared error; its autograd gives the gradient."""
    import torch
    import torch.nn.functional as F
    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    h = torch.tensor(x, dtype=torch.float64)
    off = 0
    for li, (o, i) in enumerate(shapes):
        W = pt[off:off + o * i].reshape(o, i)
        off += o * i
        b = pt[off:off + o]
        off += o
        h = F.linear(h, W, b)
        if li + 1 < len(shapes):
            h = F.gelu(h, approximate="none")
    loss = F.mse_loss(h, torch.tensor(t, dtype=torch.float64))
    loss.backward()
    return float(loss), pt.grad.numpy(), h.detach().numpy()


def test_mlp_oracle_forward_and_gradient_match_torch():
    """Forward outputs, loss and gradients of the oracle (no bf16 rounding) against torch's
    F.linear / F.gelu(approximate="none") / F.mse_loss and autograd, at a small shape and
    at the warm-start network's shape (512 -> 3 x 1024 -> 168) on a few samples."""
    rng = np.random.default_rng(3)
    for shapes, B in (([(9, 6), (9, 9), (4, 9)], 5), ([(1024, 512), (1024, 1024), (1024, 1024), (168, 1024)], 3)):
        P = sum(o * i + o for o, i in shapes)
        p = rng.standard_normal(P) * 0.05
        x = rng.standard_normal((B, shapes[0][1]))
        t = rng.standard_normal((B, shapes[-1][0]))
        loss, g, y = mo.forward_backward(p, shapes, x, t, round_bf16=False)
        lt, gt, yt = _torch_net(p, shapes, x, t)
        assert np.max(np.abs(y - yt)) <= 1e-12 * max(1.0, np.max(np.abs(yt)))
        assert abs(loss - lt) <= 1e-12 * abs(lt)
        assert np.max(np.abs(g - gt)) <= 1e-12 * max(1e-300, np.max(np.abs(gt)))


def test_mlp_oracle_tanh_gelu_would_fail():
    """The check above separates the exact GELU from its tanh approximation (the two
    differ by up to ~1e-3), so a swapped activation cannot pass it."""
    import torch
    import torch.nn.functional as F
    z = np.linspace(-4, 4, 101)
    exact = mo.gelu(z)
    assert np.max(np.abs(exact - F.gelu(torch.tensor(z), approximate="none").numpy())) < 1e-15
    assert np.max(np.abs(exact - F.gelu(torch.tensor(z), approximate="tanh").numpy())) > 1e-5
