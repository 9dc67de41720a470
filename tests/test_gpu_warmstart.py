"""The learned warm start's predict -> refine -> fallback path at reduced scale, and
what the trained network gains over the method without it.

The network (512 -> 3 x 1024 GELU -> 168, the configured shape) is trained on 20k
generated pairs; 256 held-out reachable targets are decomposed from its predicted
starts with the batched mixed-precision optimiser, the misses again from random
starts.  Checked: every reported solve is a solve by the FP64 oracle's cost of the
returned parameters (within 1e-12), the fallback runs exactly the misses, the total
success is the union of the two stages, and the gain of the predicted starts over
random starts (success from prediction / success from a random start, same
optimiser) is measured and printed.  The gain is reported, not required to exceed 1:
on this 12-layer ansatz the network's starts do not beat random ones."""
import sys

import numpy as np
import pytest

from squander_hip import _native as nat, decompose_batch, default_options, initial_guess
from squander_hip.solve_bench import wilson, with_fallback
from squander_hip.warmstart import CircuitBatch, WarmStart, generate_pairs, target_features
from squander_hip.workloads import ansatz_4q_12

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_predict_refine_fallback(oracle):
    tol = 1e-8
    ctx = nat.Context(0)
    gates = ansatz_4q_12()
    cb = CircuitBatch(ctx, 4, gates)
    P = cb.P
    X, Y, _, _, vcost = generate_pairs(cb, 20_000, 5)
    assert np.max(np.abs(vcost)) < 1e-12
    ws = WarmStart(ctx, X.shape[1], P, hidden=1024, n_hidden=3, seed=5)
    losses = ws.train(X, Y, epochs=3, batch=1024, lr=1e-3, seed=5)
    assert losses[-1] < losses[0]
    n = 256
    _, _, _, V, _ = generate_pairs(cb, n, 6)
    U = np.conj(np.transpose(V, (0, 2, 1))).copy()
    x_pred = ws.predict(target_features(V))
    assert x_pred.shape == (n, P) and np.all(np.isfinite(x_pred))
    x_rand = np.array([initial_guess(P, 40 + i) for i in range(n)])
    ctx.set_circuit(4, gates)
    ok_pred, ok_total, _, res = with_fallback(ctx, U, x_pred, x_rand, tol, mixed=True)
    for i in np.flatnonzero(ok_pred):
        fo = oracle.optimization_problem(gates, U[i], 4, res["x"][i])
        assert abs(fo - res["f"][i]) < 1e-12 and fo <= tol + 1e-12
    # the second stage ran exactly the misses: a fresh run from random starts on them
    miss = np.flatnonzero(~ok_pred)
    r2 = decompose_batch(ctx, U[miss], np.arange(miss.size, dtype=np.int32), x_rand[miss],
                         default_options(4, optimization_tolerance=tol, mixed_precision=True, switch_cost=1e-4))
    assert np.array_equal(ok_total[miss], r2["f"] <= tol)
    assert np.array_equal(ok_total[ok_pred], ok_pred[ok_pred])
    # the method without the network: the same optimiser from random starts
    rr = decompose_batch(ctx, U, np.arange(n, dtype=np.int32), x_rand,
                         default_options(4, optimization_tolerance=tol, mixed_precision=True, switch_cost=1e-4))
    s_rand = float(np.mean(rr["f"] <= tol))
    s_pred = float(ok_pred.mean())
    gain = s_pred / max(s_rand, 1e-30)
    say(f"  success from predicted starts {s_pred:.3f} {wilson(int(ok_pred.sum()), n)}, from random starts "
        f"{s_rand:.3f}, gain {gain:.2f}; with the random-start fallback {ok_total.mean():.3f}")
    assert ok_total.mean() >= s_pred
