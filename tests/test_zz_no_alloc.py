"""The allocation-free evaluation entry point
(sqh_optimization_problem_combined_batch_noalloc_dev) against the allocating one
(sqh_optimization_problem_combined_batch_dev) through the C ABI: costs and gradients
within 1e-12 (FP64) and 1e-6 (complex64; the same kernels, so in practice equal) on
two calls in a row, the second with a smaller batch (the scratch array keeps the
first call's contents), with and without a circuit-specialised kernel compiled by an
earlier call; a batch beyond the scratch limit is rejected."""
import sys

import numpy as np
import pytest
import torch

from squander_hip import _native as nat
from squander_hip import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _eval(ctx, noalloc, ps, prec, grad):
    """Costs (and gradients) through the allocation-free entry point, or through the
    allocating one: sqh_optimization_problem_combined_batch_dev, or for costs only
    sqh_optimization_problem_batch_dev."""
    B, P = ps.shape
    dp = torch.from_numpy(ps).cuda()
    f = torch.full((B,), np.nan, dtype=torch.float64, device="cuda")
    g = torch.full((B, P), np.nan, dtype=torch.float64, device="cuda") if grad else None
    torch.cuda.synchronize()
    L = ctx._L
    if noalloc:
        ctx.check(L.sqh_optimization_problem_combined_batch_noalloc_dev(ctx.h, dp.data_ptr(), B, prec, f.data_ptr(),
                                                                        g.data_ptr() if grad else None))
    elif grad:
        ctx.check(L.sqh_optimization_problem_combined_batch_dev(ctx.h, dp.data_ptr(), B, prec, f.data_ptr(),
                                                                g.data_ptr()))
    else:
        ctx.check(L.sqh_optimization_problem_batch_dev(ctx.h, dp.data_ptr(), B, prec, f.data_ptr()))
    ctx.synchronize()
    return f.cpu().numpy(), (g.cpu().numpy() if grad else None)


@pytest.mark.parametrize("n,jit,prec,grad,batches", [
    (8, False, nat.PREC_F64, True, (100, 37)),
    (8, True, nat.PREC_F64, True, (64, 9)),
    (8, False, nat.PREC_F64, False, (512, 100)),
    (4, False, nat.PREC_F32, True, (300, 5)),
    (6, True, nat.PREC_F32, False, (200, 17)),
])
def test_noalloc_matches_allocating_entry_point(n, jit, prec, grad, batches):
    say(f"no-alloc evaluation n={n} jit={jit} prec={prec} grad={grad} batches={batches}")
    gates = W.ansatz_8q_40() if n == 8 else W.ansatz_layers(n, W.random_adjacent_pairs(n, 3 * n, n))
    P = W.param_count(gates)
    ctx = nat.Context(0)
    ctx.set_circuit(n, gates)
    ctx.set_target(W.haar_unitary(1 << n, 77 + n).conj().T.copy())
    ctx.set_jit(jit)
    tol = 1e-12 if prec == nat.PREC_F64 else 1e-6
    for k, B in enumerate(batches):
        ps = W.random_params(P, 1000 * n + k, batch=B)
        fa, ga = _eval(ctx, False, ps, prec, grad)
        fn, gn = _eval(ctx, True, ps, prec, grad)
        say(f"  B={B}: max|df| {np.abs(fa - fn).max():.2e}" + (f", max|dg| {np.abs(ga - gn).max():.2e}" if grad else ""))
        assert np.all(np.isfinite(fn)) and np.abs(fa - fn).max() <= tol
        if grad:
            assert np.all(np.isfinite(gn)) and np.abs(ga - gn).max() <= tol


def test_noalloc_rejects_batches_beyond_the_scratch_limit():
    gates = W.ansatz_8q_40()
    P = W.param_count(gates)
    ctx = nat.Context(0)
    ctx.set_circuit(8, gates)
    ctx.set_target(W.haar_unitary(256, 5).conj().T.copy())
    ps = W.random_params(P, 6, batch=400)
    dp = torch.from_numpy(ps).cuda()
    f = torch.empty(400, dtype=torch.float64, device="cuda")
    g = torch.empty((400, P), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(nat.SqhError, match="limit"):
        ctx.check(ctx._L.sqh_optimization_problem_combined_batch_noalloc_dev(ctx.h, dp.data_ptr(), 400, nat.PREC_F64,
                                                                             f.data_ptr(), g.data_ptr()))
