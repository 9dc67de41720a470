"""The warm-start network's HIP forward / backward / Adam kernels against the CPU
oracle (oracle/mlp_oracle.py, which mirrors the bf16 roundings).  Bar: outputs and
per-layer gradients within 2e-2 relative L2 error (bf16 inputs, FP32 vs FP64
accumulation order), and training lowers the loss."""
import sys

import numpy as np
import pytest
import torch

import mlp_oracle as mo
from squander_hip import Context
from squander_hip.mlp import WarmStartMLP

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("dims", [(64, 128, 2, 40), (512, 256, 3, 168), (512, 1024, 3, 168)])
def test_forward_and_gradients_match_oracle(dims):
    say(f"mlp dims {dims}")
    ind, hid, nh, out = dims
    ctx = Context(0)
    net = WarmStartMLP(ctx, ind, hid, nh, out, seed=3)
    p = net.get_params()
    shapes = net.layer_shapes()
    assert p.size == sum(o * i + o for o, i in shapes)
    rng = np.random.default_rng(1)
    B = 100
    x = rng.standard_normal((B, ind)).astype(np.float32)
    t = rng.standard_normal((B, out)).astype(np.float32)
    loss_g, g = net.gradients(x, t)
    loss_o, g_o, y_o = mo.forward_backward(p.astype(np.float64), shapes, x, t, round_bf16=True)
    say(f"  loss gpu {loss_g:.6f} oracle {loss_o:.6f}")
    assert abs(loss_g - loss_o) < 2e-2 * loss_o
    off = 0
    for (o, i) in shapes:
        nW = o * i
        eW = rel(g[off:off + nW], g_o[off:off + nW])
        eb = rel(g[off + nW:off + nW + o], g_o[off + nW:off + nW + o])
        say(f"  layer {o}x{i}: dW rel err {eW:.2e}, db rel err {eb:.2e}")
        assert eW < 2e-2 and eb < 2e-2
        off += nW + o
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty((B, out), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    net.forward_dev(xd.data_ptr(), B, yd.data_ptr())
    ctx.synchronize()
    assert rel(yd.cpu().numpy(), y_o) < 2e-2


def test_training_lowers_loss():
    say("mlp training")
    ctx = Context(0)
    net = WarmStartMLP(ctx, 64, 128, 2, 40, seed=5)
    rng = np.random.default_rng(2)
    A = rng.standard_normal((64, 40)).astype(np.float32) / 8.0
    x = rng.standard_normal((512, 64)).astype(np.float32)
    t = np.tanh(x @ A).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    td = torch.from_numpy(t).cuda()
    torch.cuda.synchronize()
    first = net.train_step_dev(xd.data_ptr(), td.data_ptr(), 512, 1e-3)
    for _ in range(300):
        last = net.train_step_dev(xd.data_ptr(), td.data_ptr(), 512, 1e-3, want_loss=False)
    last = net.train_step_dev(xd.data_ptr(), td.data_ptr(), 512, 1e-3)
    say(f"  loss {first:.4f} -> {last:.4f}")
    assert last < 0.2 * first


def test_backward_step_with_mse_gradient_equals_train_step():
    """An Adam step from the caller-supplied output gradient equals the MSE training
    step when the supplied gradient is the MSE gradient 2 (y - t) / (B out)."""
    from squander_hip import _native as nat
    from squander_hip.mlp import WarmStartMLP
    ctx = nat.Context(0)
    a = WarmStartMLP(ctx, 64, 128, 2, 40, seed=3)
    b = WarmStartMLP(ctx, 64, 128, 2, 40, seed=3)
    rng = np.random.default_rng(0)
    B = 96
    x = torch.from_numpy(rng.standard_normal((B, 64)).astype(np.float32)).cuda()
    t = torch.from_numpy(rng.standard_normal((B, 40)).astype(np.float32)).cuda()
    y = torch.empty((B, 40), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    b.forward_dev(x.data_ptr(), B, y.data_ptr())
    ctx.synchronize()
    dy = (2.0 * (y - t) / (B * 40)).contiguous()
    torch.cuda.synchronize()
    a.train_step_dev(x.data_ptr(), t.data_ptr(), B, 1e-3)
    b.backward_step_dev(x.data_ptr(), dy.data_ptr(), B, 1e-3)
    ctx.synchronize()
    pa, pb = a.get_params(), b.get_params()
    assert np.max(np.abs(pa - pb)) <= 1e-6 * max(1.0, np.max(np.abs(pa)))
