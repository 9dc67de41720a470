"""CPU oracle of the warm-start network (TEST INFRASTRUCTURE ONLY).

A numpy restatement of the network the library trains: in -> hidden (GELU, erf
form) x n -> out, mean-squared error, with the same BF16 roundings the GPU path
makes (weights, layer inputs and back-propagated deltas rounded to bfloat16,
round-to-nearest-even; products accumulated in float64 here, float32 on the GPU).
Gradients of the unrounded network are pinned against central finite
differences in the tests.
"""
import numpy as np
from scipy.special import erf


def bf16(a):
    """Round float32 values to bfloat16 (nearest, ties to even) and back to float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def unflatten(p, shapes):
    out, off = [], 0
    for (o, i) in shapes:
        W = p[off:off + o * i].reshape(o, i)
        off += o * i
        b = p[off:off + o]
        off += o
        out.append((W.astype(np.float64), b.astype(np.float64)))
    return out


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward_backward(p, shapes, x, t, round_bf16=True):
    """(loss, flat gradient, outputs) of the network on (x, t)."""
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unflatten(p, shapes)
    Hs, Zs = [np.asarray(r(x), dtype=np.float64)], []
    for li, (W, b) in enumerate(layers):
        z = Hs[-1] @ np.asarray(r(W), dtype=np.float64).T + b
        Zs.append(z)
        if li + 1 < len(layers):
            Hs.append(np.asarray(r(gelu(z)), dtype=np.float64))
    y = Zs[-1]
    B, out = y.shape
    d = y - t
    loss = float(np.sum(d * d) / (B * out))
    dz = 2.0 * d / (B * out)
    grads = []
    for li in range(len(layers) - 1, -1, -1):
        W, b = layers[li]
        dzr = np.asarray(r(dz), dtype=np.float64)
        grads.append((dzr.T @ Hs[li], dz.sum(axis=0)))
        if li > 0:
            dh = dzr @ np.asarray(r(W), dtype=np.float64)
            dz = dh * gelu_grad(Zs[li - 1])
    grads = grads[::-1]
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat, y
