"""Uneven persistent batch: independent decompositions of 3..6-qubit targets with
their own circuits, run as one batch (sqh_decompose_problems): by default the
persistent solver (whole optimisations in one launch per qubit-count group, the groups
side by side), or the lockstep rounds with engine="rounds".

Half of the problems are Haar targets, half are ansatz-reachable targets built from
seeded random angles; every problem's circuit is the U3 layer + L random adjacent
CNOT layers ansatz with L uniform in 4..20 (a reachable target is decomposed with
the circuit that built it).  Complex64 descent, FP64 final iterations and FP64
acceptance (mixed precision), every reported cost the FP64 cost at the returned
parameters.
"""
from __future__ import annotations

import time

import numpy as np

from . import _native as nat
from .circuit import Circuit
from .optimizer import decompose_problems, default_options, initial_guess
from .workloads import ansatz_layers, haar_unitary, param_count, random_adjacent_pairs, random_params


def make_problems(count=256, seed=2028, qubits=(3, 6), cnots=(4, 20), matrix=None):
    """[(qbit_num, gates, Umtx, x0, kind)] with seeded circuits and targets.
    matrix(n, gates, theta) -> V builds reachable targets (default: the library's
    get_Matrix on the GPU)."""
    rng = np.random.default_rng(seed)
    if matrix is None:
        def matrix(n, gates, theta):
            return Circuit.from_gate_array(n, gates).get_Matrix(theta)
    out = []
    for i in range(count):
        n = int(rng.integers(qubits[0], qubits[1] + 1))
        L = int(rng.integers(cnots[0], cnots[1] + 1))
        gates = ansatz_layers(n, random_adjacent_pairs(n, L, rng))
        P = param_count(gates)
        if i % 2 == 0:
            V = haar_unitary(1 << n, rng)
            kind = "haar"
        else:
            V = matrix(n, gates, random_params(P, rng))
            kind = "reachable"
        x0 = initial_guess(P, seed + 1 + i)
        out.append((n, gates, np.conj(V.T).copy(), x0, kind))
    return out


def run(count=256, seed=2028, tol=1e-8, mixed=True, log=print, engine="auto"):
    ctx = nat.Context(0)
    probs = make_problems(count, seed)
    opts = default_options(4, optimization_tolerance=tol, mixed_precision=mixed, switch_cost=1e-4, engine=engine)
    t0 = time.perf_counter()
    res = decompose_problems(ctx, [(n, g, U, x0) for n, g, U, x0, _ in probs], opts)
    dt = time.perf_counter() - t0
    kinds = np.array([p[4] for p in probs])
    ok = res["f"] <= tol
    out = {"problems": count, "seconds": dt, "solves_per_s": float(ok.sum() / dt), "rounds": res["rounds"],
           "evals_p50": float(np.median(res["evals"])), "evals_p99": float(np.percentile(res["evals"], 99)),
           "mixed_precision": bool(mixed), "engine": engine}
    for k in ("reachable", "haar"):
        sel = kinds == k
        out[k] = {"count": int(sel.sum()), "success": float(ok[sel].mean()),
                  "evals_p50": float(np.median(res["evals"][sel])),
                  "evals_p99": float(np.percentile(res["evals"][sel], 99))}
    qn = np.array([p[0] for p in probs])
    out["success_by_qubits"] = {int(n): float(ok[qn == n].mean()) for n in np.unique(qn)}
    log(f"uneven batch: {count} problems in {dt:.2f}s, success reachable {out['reachable']['success']:.3f} "
        f"haar {out['haar']['success']:.3f}, {out['solves_per_s']:.1f} solves/s")
    return out, probs, res


if __name__ == "__main__":
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(description="uneven batch of independent decompositions on the GPU")
    ap.add_argument("--count", type=int, default=256)
    ap.add_argument("--seed", type=int, default=2028)
    ap.add_argument("--fp64", action="store_true", help="FP64 throughout (default: complex64 descent)")
    ap.add_argument("--engine", default="auto", choices=["auto", "rounds", "persistent"])
    a = ap.parse_args()
    res, _, _ = run(count=a.count, seed=a.seed, mixed=not a.fp64, log=lambda *m: print(*m, file=sys.stderr),
                    engine=a.engine)
    print(json.dumps(res))
