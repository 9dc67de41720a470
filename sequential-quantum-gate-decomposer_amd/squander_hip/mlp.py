"""The warm-start network (sqh_mlp_*): BF16 MLP trained on the GPU by the library's
own forward / backward / Adam kernels."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as nat


class WarmStartMLP:
    """in -> hidden (GELU) x n_hidden -> out; parameters as a flat float32 array
    (per layer W (out x in) then b)."""

    def __init__(self, ctx: nat.Context, in_dim=512, hidden=1024, n_hidden=3, out_dim=168, seed=0):
        self.ctx = ctx
        self.in_dim, self.hidden, self.n_hidden, self.out_dim = in_dim, hidden, n_hidden, out_dim
        h = ctypes.c_void_p()
        ctx.check(ctx._L.sqh_mlp_create(ctx.h, in_dim, hidden, n_hidden, out_dim, seed, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.ctx._L.sqh_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def param_count(self) -> int:
        return int(self.ctx._L.sqh_mlp_param_count(self.h))

    def get_params(self) -> np.ndarray:
        out = np.empty(self.param_count(), dtype=np.float32)
        self.ctx.check(self.ctx._L.sqh_mlp_get_params(self.h, out.ctypes.data))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self.ctx.check(self.ctx._L.sqh_mlp_set_params(self.h, p.ctypes.data))

    def layer_shapes(self):
        dims = [self.in_dim] + [self.hidden] * self.n_hidden + [self.out_dim]
        return [(dims[i + 1], dims[i]) for i in range(len(dims) - 1)]

    def forward_dev(self, x_ptr: int, B: int, y_ptr: int):
        self.ctx.check(self.ctx._L.sqh_mlp_forward_dev(self.h, x_ptr, B, y_ptr))

    def train_step_dev(self, x_ptr: int, t_ptr: int, B: int, lr: float, want_loss=True):
        loss = ctypes.c_double(0.0)
        self.ctx.check(self.ctx._L.sqh_mlp_train_step_dev(self.h, x_ptr, t_ptr, B, lr,
                                                           ctypes.byref(loss) if want_loss else None))
        return loss.value

    def backward_step_dev(self, x_ptr: int, dy_ptr: int, B: int, lr: float):
        """One Adam step from dLoss/dOutput (B x out FP32 on the device)."""
        self.ctx.check(self.ctx._L.sqh_mlp_backward_step_dev(self.h, x_ptr, dy_ptr, B, lr))

    def gradients(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.empty(self.param_count(), dtype=np.float32)
        loss = ctypes.c_double(0.0)
        self.ctx.check(self.ctx._L.sqh_mlp_gradients(self.h, x.ctypes.data, t.ctypes.data, x.shape[0], g.ctypes.data,
                                                      ctypes.byref(loss)))
        return loss.value, g
