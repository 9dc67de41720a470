"""4-qubit decompositions solved per second: plain FP64 random starts versus the
batched mixed-precision optimiser started from the learned warm start (with random
starts as the fallback for runs that miss the tolerance).

Ansatz: U3 on 4 qubits + 12 CNOT layers on alternating adjacent pairs, each followed
by U3 on both qubits (84 raw parameters).  Evaluation set: seeded held-out
ansatz-reachable targets plus Haar targets (a generic 4-qubit unitary needs >= 61
CNOTs, so Haar targets are reported separately and are not expected to solve).
"""
from __future__ import annotations

import time

import numpy as np

from . import _native as nat
from .optimizer import decompose_batch, default_options, initial_guess
from .warmstart import CircuitBatch, WarmStart, generate_pairs, target_features
from .workloads import ansatz_4q_12, haar_unitary


def _solve(ctx, Umtx, x0, tol, mixed, max_outer=4):
    opts = default_options(4, optimization_tolerance=tol, mixed_precision=mixed, switch_cost=1e-4,
                           max_outer_iterations=max_outer)
    t0 = time.perf_counter()
    res = decompose_batch(ctx, Umtx, np.arange(Umtx.shape[0], dtype=np.int32), x0, opts)
    return res, time.perf_counter() - t0


def wilson(k: int, n: int, z: float = 1.96):
    """95 % Wilson score interval of a success rate k / n."""
    if n == 0:
        return [0.0, 0.0]
    p = k / n
    den = 1.0 + z * z / n
    mid = (p + z * z / (2 * n)) / den
    half = z * np.sqrt(p * (1 - p) / n + z * z / (4 * n * n)) / den
    return [float(max(0.0, mid - half)), float(min(1.0, mid + half))]


def _leg(res, seconds, tol, extra=None):
    ok = res["f"] <= tol
    d = {"success": float(ok.mean()), "success_ci95": wilson(int(ok.sum()), ok.size), "solved": int(ok.sum()),
         "seconds": seconds, "solves_per_s": float(ok.sum() / seconds), "rounds": int(res["rounds"]),
         "evals_p50": float(np.median(res["evals"])), "evals_p99": float(np.percentile(res["evals"], 99))}
    if extra:
        d.update(extra)
    return d


def with_fallback(ctx, Umtx, x_first, x_second, tol, mixed):
    """One batched run from x_first; the runs that miss the tolerance run again from
    x_second.  Returns (ok first, ok total, seconds, first result)."""
    res1, t1 = _solve(ctx, Umtx, x_first, tol, mixed=mixed)
    ok1 = res1["f"] <= tol
    miss = np.flatnonzero(~ok1)
    ok = ok1.copy()
    t2 = 0.0
    if miss.size:
        res2, t2 = _solve(ctx, Umtx[miss], x_second[miss], tol, mixed=mixed)
        ok[miss] = res2["f"] <= tol
    return ok1, ok, t1 + t2, res1


def plain_single(ctx, Umtx, x0, tol):
    """The plain method on the GPU: one FP64 decomposition per call from a random start
    (what Start_Decomposition does for one target), targets one after another."""
    fs, evals, rounds = [], [], 0
    t0 = time.perf_counter()
    for i in range(Umtx.shape[0]):
        res, _ = _solve(ctx, Umtx[i:i + 1], x0[i:i + 1], tol, mixed=False)
        fs.append(res["f"][0])
        evals.append(res["evals"][0])
        rounds += res["rounds"]
    dt = time.perf_counter() - t0
    return {"f": np.array(fs), "evals": np.array(evals), "rounds": rounds}, dt


def run(n_train=200_000, n_reach=4096, n_haar=1024, epochs=20, tol=1e-8, seed=2026, log=print,
        hidden=1024, n_hidden=3, n_plain=128, lr=1e-3, ft_epochs=0, ft_lr=1e-4, jit=True, ft_stream_steps=0,
        keep=None):
    """Returns a dict with the legs per target set, each component isolated:
    plain_single (the reference flow on the GPU: one FP64 decomposition per call,
    measured on the first n_plain targets), batched_fp64 and batched_mixed (all
    targets in one batch from the same random starts), two_random_starts (batched
    mixed, misses rerun from a second random start: the control for the fallback),
    and learned_mixed (network start, batched mixed precision, random-start fallback).
    Success rates carry 95 % Wilson intervals.  With a dict as keep, keep[target set]
    receives the targets (Umtx), the reported costs and the parameters of the batched
    and learned legs, for an outside check."""
    ctx = nat.Context(0)
    ctx.set_jit(jit)
    gates = ansatz_4q_12()
    cb = CircuitBatch(ctx, 4, gates)
    P = cb.P
    out = {"ansatz": "4 qubits, U3 layer + 12 alternating adjacent CNOT layers", "params": P, "tol": tol}

    # training data: seeded angles, target V = C(theta*), label = canonical(theta*),
    # every pair verified with the decomposer's FP64 cost at the label
    t0 = time.perf_counter()
    Xtr, Ytr, _, Vtr, vcost = generate_pairs(cb, n_train, seed)
    out["train_pairs"] = n_train
    out["train_pairs_max_verify_cost"] = float(np.max(np.abs(vcost)))
    out["datagen_s"] = time.perf_counter() - t0
    log(f"generated {n_train} pairs in {out['datagen_s']:.1f}s, max verification cost {out['train_pairs_max_verify_cost']:.1e}")

    ws = WarmStart(ctx, Xtr.shape[1], P, hidden=hidden, n_hidden=n_hidden, seed=seed)
    t0 = time.perf_counter()
    losses = ws.train(Xtr, Ytr, epochs=epochs, batch=1024, lr=lr, seed=seed, log=log)
    out["train_s"] = time.perf_counter() - t0
    out["train_loss_first_last"] = [losses[0], losses[-1]]
    if ft_epochs > 0:
        # second stage: the decomposition cost itself as the loss (phase-blind)
        t0 = time.perf_counter()
        Utr = np.conj(np.transpose(Vtr, (0, 2, 1)))
        fc = ws.finetune_cost(Xtr, Utr, gates, 4, epochs=ft_epochs, lr=ft_lr, seed=seed, log=log)
        out["finetune_s"] = time.perf_counter() - t0
        out["finetune_cost_first_last"] = [fc[0], fc[-1]]
        del Utr
    if ft_stream_steps > 0:
        t0 = time.perf_counter()
        fc = ws.finetune_cost_stream(gates, 4, steps=ft_stream_steps, lr=ft_lr, seed=seed + 5, log=log)
        out["stream_finetune_s"] = time.perf_counter() - t0
        out["stream_finetune_cost"] = fc
    del Xtr, Ytr, Vtr

    _, _, _, Vr, _ = generate_pairs(cb, n_reach, seed + 1)
    # compile the circuit-specialised kernels (both precisions) outside the timed legs
    t0 = time.perf_counter()
    Uw = np.conj(np.transpose(Vr[:1], (0, 2, 1))).copy()
    xw = np.array([initial_guess(P, seed)])
    for mixed in (False, True):
        _solve(ctx, Uw, xw, tol, mixed=mixed)
    out["kernel_warmup_s"] = time.perf_counter() - t0
    Vh = np.array([haar_unitary(16, seed + 10_000 + i) for i in range(n_haar)]) if n_haar else None

    results = {}
    for name, V in (("reachable", Vr), ("haar", Vh)):
        if V is None:
            continue
        Umtx = np.conj(np.transpose(V, (0, 2, 1))).copy()
        n = V.shape[0]
        x_rand = np.array([initial_guess(P, seed + 7 + i) for i in range(n)])
        x_rand2 = np.array([initial_guess(P, seed + 7 + n + i) for i in range(n)])
        k = min(n_plain, n)
        res_s, t_s = plain_single(ctx, Umtx[:k], x_rand[:k], tol)
        res_b, t_b = _solve(ctx, Umtx, x_rand, tol, mixed=False)
        res_m, t_m = _solve(ctx, Umtx, x_rand, tol, mixed=True)
        ok2_first, ok2, t_2, _ = with_fallback(ctx, Umtx, x_rand, x_rand2, tol, mixed=True)
        t0 = time.perf_counter()
        x_pred = ws.predict(target_features(V))
        t_pred = time.perf_counter() - t0
        ok_l, ok_total, t_lf, res_l = with_fallback(ctx, Umtx, x_pred, x_rand, tol, mixed=True)
        t_total = t_pred + t_lf
        # start quality: FP64 cost at the predicted and at the random start points
        c_pred = cb.verify(x_pred, V)
        c_rand = cb.verify(x_rand, V)
        c_pred_pf = cb.verify(x_pred, V, phase_free=True)
        c_rand_pf = cb.verify(x_rand, V, phase_free=True)
        r = {"targets": n,
             "plain_single": _leg(res_s, t_s, tol, {"targets_measured": k}),
             "batched_fp64": _leg(res_b, t_b, tol),
             "batched_mixed": _leg(res_m, t_m, tol),
             "two_random_starts": {"success": float(ok2.mean()), "success_ci95": wilson(int(ok2.sum()), n),
                                   "success_first_start": float(ok2_first.mean()), "seconds": t_2,
                                   "solves_per_s": float(ok2.sum() / t_2)},
             "learned_mixed": {"success": float(ok_total.mean()), "success_ci95": wilson(int(ok_total.sum()), n),
                               "solved": int(ok_total.sum()),
                               "success_from_prediction": float(ok_l.mean()),
                               "success_from_prediction_ci95": wilson(int(ok_l.sum()), n), "seconds": t_total,
                               "solves_per_s": float(ok_total.sum() / t_total), "rounds": int(res_l["rounds"]),
                               "evals_p50": float(np.median(res_l["evals"])),
                               "evals_p99": float(np.percentile(res_l["evals"], 99)),
                               "predict_s": t_pred, "fallback_runs": int((~ok_l).sum())},
             "start_cost_median": {"predicted": float(np.median(c_pred)), "random": float(np.median(c_rand)),
                                   "predicted_phase_free": float(np.median(c_pred_pf)),
                                   "random_phase_free": float(np.median(c_rand_pf))}}
        r["learned_gain_over_random_start"] = (r["learned_mixed"]["success_from_prediction"]
                                               / max(r["batched_mixed"]["success"], 1e-30))
        r["speedup_vs_plain_single"] = r["learned_mixed"]["solves_per_s"] / max(r["plain_single"]["solves_per_s"], 1e-30)
        r["speedup_vs_batched_fp64"] = r["learned_mixed"]["solves_per_s"] / max(r["batched_fp64"]["solves_per_s"], 1e-30)
        r["max_fp64_cost_of_learned_solves"] = (float(np.max(cb.verify(res_l["x"][ok_l], V[ok_l])))
                                                if ok_l.any() else None)
        log(f"{name}: plain single {r['plain_single']['success']:.3f} @ {r['plain_single']['solves_per_s']:.1f}/s; "
            f"batched fp64 {r['batched_fp64']['success']:.3f} @ {r['batched_fp64']['solves_per_s']:.0f}/s; "
            f"batched mixed {r['batched_mixed']['success']:.3f} @ {r['batched_mixed']['solves_per_s']:.0f}/s; "
            f"two random starts {r['two_random_starts']['success']:.3f} @ {r['two_random_starts']['solves_per_s']:.0f}/s; "
            f"learned+mixed {r['learned_mixed']['success']:.3f} (pred {r['learned_mixed']['success_from_prediction']:.3f}) "
            f"@ {r['learned_mixed']['solves_per_s']:.0f}/s; start cost pred {r['start_cost_median']['predicted']:.3f} "
            f"rand {r['start_cost_median']['random']:.3f}")
        results[name] = r
        if keep is not None:
            keep[name] = {"Umtx": Umtx, "batched": (res_b["f"], res_b["x"]), "learned": (res_l["f"], res_l["x"])}
    out["results"] = results
    return out


if __name__ == "__main__":
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(description="4-qubit solves/s: plain vs batched vs learned + batched")
    ap.add_argument("--train", type=int, default=200_000)
    ap.add_argument("--reach", type=int, default=4096)
    ap.add_argument("--haar", type=int, default=1024)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--plain", type=int, default=128, help="targets timed on the one-at-a-time plain path")
    ap.add_argument("--ft-epochs", type=int, default=0, help="epochs of training on the decomposition cost")
    ap.add_argument("--ft-lr", type=float, default=1e-4)
    ap.add_argument("--ft-stream", type=int, default=0, help="steps of cost training on fresh targets")
    ap.add_argument("--lr", type=float, default=1e-3)
    a = ap.parse_args()
    res = run(n_train=a.train, n_reach=a.reach, n_haar=a.haar, epochs=a.epochs, n_plain=a.plain, lr=a.lr,
              ft_epochs=a.ft_epochs, ft_lr=a.ft_lr, ft_stream_steps=a.ft_stream,
              log=lambda *m: print(*m, file=sys.stderr, flush=True))
    print(json.dumps(res))
