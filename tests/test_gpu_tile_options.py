"""Launch options of the register-tile kernels (sqh_set_tile_config): wavefronts per
workgroup of the forward pass (or the packed layout) and the gradient pass's program
staged in LDS.  Every option gives bit-identical costs and gradients, the run-time
tuner picks one of them, and the results match the oracle."""
import sys

import numpy as np
import pytest

from squander_hip import _native as nat
from squander_hip import workloads as W

pytestmark = pytest.mark.gpu

OPTIONS = [(w, s) for w in (1, 2, 4, -1) for s in (0, 1)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _costs(ctx, ps, prec):
    c = np.empty(ps.shape[0])
    ctx.check(ctx._L.sqh_optimization_problem_batch(ctx.h, np.ascontiguousarray(ps).ctypes.data, ps.shape[0], prec,
                                                     c.ctypes.data))
    return c


def _combined(ctx, ps, prec):
    B, P = ps.shape
    f = np.empty(B)
    g = np.empty((B, P))
    ctx.check(ctx._L.sqh_optimization_problem_combined_batch(ctx.h, np.ascontiguousarray(ps).ctypes.data, B, prec,
                                                              f.ctypes.data, g.ctypes.data))
    return f, g


@pytest.mark.parametrize("prec", [nat.PREC_F64, nat.PREC_F32])
@pytest.mark.parametrize("n", [3, 5, 6, 8, 10])
def test_tile_launch_options_are_bit_identical(oracle, n, prec):
    gates = W.mixed_gate_circuit(n, 3 * n + 6, 4300 + n)
    P = W.param_count(gates)
    d = 1 << n
    U = W.haar_unitary(d, 4400 + n)[:, :max(1, d - 5)].copy()
    B = 6
    ps = W.random_params(P, 4500 + n, batch=B)
    ctx = nat.Context(0)
    ctx.set_circuit(n, gates)
    ctx.set_target(U)
    ctx.set_fused_cost(False)
    ctx.set_jit(False)
    f32 = prec == nat.PREC_F32
    tc = _costs(ctx, ps, prec)
    tf, tg = _combined(ctx, ps, prec)
    say(f"tile options n={n} prec={prec}: tuner picked forward {ctx.tile_config(1, f32, B)}, "
        f"gradient {ctx.tile_config(2, f32, B)}")
    for w, st in OPTIONS:
        ctx.set_tile_config(w, st)
        c = _costs(ctx, ps, prec)
        f, g = _combined(ctx, ps, prec)
        assert np.array_equal(c, tc), (w, st)
        assert np.array_equal(f, tf) and np.array_equal(g, tg), (w, st)
        wpg, staged = ctx.tile_config(1, f32, B)
        assert wpg in (-1, 1, 2, 4) and staged in (0, 1)
        if w == -1 or n <= 4:
            assert wpg == -1
    ctx.set_tile_config(0, -1)
    tol = 1e-5 if f32 else 1e-12
    for b in range(2):
        fo, go = oracle.optimization_problem_combined(gates, U, n, ps[b])
        assert abs(tc[b] - fo) < tol and abs(tf[b] - fo) < tol and np.abs(tg[b] - go).max() < tol
    with pytest.raises(nat.SqhError):
        ctx.set_tile_config(3, 0)
    with pytest.raises(nat.SqhError):
        ctx.set_tile_config(1, 2)
