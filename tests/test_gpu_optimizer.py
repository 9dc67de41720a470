"""GPU batched decomposition (sqh_decompose_batch) against the oracle's restated
driver (or_decompose_custom: BFGS_Powell + Wolfe line search + restarts) from the
same mt19937 start points.  Trajectories are chaotic in rounding, so the bar is:
when the oracle converges below the tolerance the GPU run does too (and vice
versa on these well-conditioned cases), and every GPU-reported cost equals the FP64
oracle's cost of the returned parameters within 1e-12."""
import sys

import numpy as np
import pytest

import oracle as orc
from squander_hip import Circuit, N_Qubit_Decomposition_custom, decompose_batch, default_options, initial_guess
from squander_hip import _native as nat
from squander_hip import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def reachable_targets(n, gates, count, seed):
    o = orc.Oracle()
    P = W.param_count(gates)
    rng = np.random.default_rng(seed)
    Ts = []
    for _ in range(count):
        th = rng.uniform(0, 2 * np.pi, P)
        V = o.apply_to(gates, th, np.eye(1 << n))
        Ts.append(V.conj().T)
    return np.array(Ts)


@pytest.mark.parametrize("n,layers", [(2, 3), (3, 4), (3, 8)])
def test_batch_matches_oracle_driver(oracle, n, layers):
    say(f"decompose_batch vs oracle n={n} layers={layers}")
    gates = W.ansatz_layers(n, W.alternating_pairs(n, layers))
    P = W.param_count(gates)
    T = reachable_targets(n, gates, 6, 100 + n)
    seeds = list(range(1, 7))
    x0 = np.array([initial_guess(P, s, "RANDOM") for s in seeds])
    ctx = nat.Context(0)
    ctx.set_circuit(n, gates)
    opts = default_options(n, optimization_tolerance=1e-8, max_outer_iterations=4)
    res = decompose_batch(ctx, T, np.arange(6, dtype=np.int32), x0, opts)
    for r in range(6):
        fo, xo, co = oracle.decompose_custom(gates, T[r], n, seeds[r], max_outer_iterations=4,
                                             optimization_tolerance=1e-8)
        fg = res["f"][r]
        f_check = oracle.optimization_problem(gates, T[r], n, res["x"][r])
        say(f"  run {r}: gpu f={fg:.3e} evals={res['evals'][r]}  oracle f={fo:.3e} evals={co}")
        assert abs(f_check - fg) < 1e-12
        assert (fo < 1e-8) == (fg < 1e-8) or abs(fo - fg) < 1e-10


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_start_decomposition_matches_oracle_driver(oracle, seed):
    """N_Qubit_Decomposition_custom.Start_Decomposition with config random_seed gives
    the oracle driver's result from the same seed (a random start may end in a local
    minimum; both must end in the same one)."""
    say(f"Start_Decomposition 3q seed {seed}")
    gates = W.ansatz_layers(3, W.alternating_pairs(3, 6))
    T = reachable_targets(3, gates, 1, 7)[0]
    dec = N_Qubit_Decomposition_custom(T, config={"random_seed": seed, "optimization_tolerance": 1e-10})
    dec.set_Gate_Structure(Circuit.from_gate_array(3, gates))
    dec.Start_Decomposition()
    err = dec.get_Decomposition_Error()
    fo, xo, co = oracle.decompose_custom(gates, T, 3, seed, max_outer_iterations=4, optimization_tolerance=1e-10)
    say(f"  error {err:.3e}, evals {dec.last_run['evals'][0]}; oracle {fo:.3e}, evals {co}")
    assert abs(err - oracle.optimization_problem(gates, T, 3, dec.get_Optimized_Parameters())) < 1e-12
    assert (fo < 1e-10) == (err < 1e-10) or abs(err - fo) < 1e-10


def test_mixed_precision_batch_4q(oracle):
    say("mixed-precision batch 4q")
    n = 4
    gates = W.ansatz_4q_12()
    P = W.param_count(gates)
    T = reachable_targets(n, gates, 16, 44)
    rng = np.random.default_rng(5)
    x0 = rng.uniform(0, 2 * np.pi, (32, P))
    run_target = np.repeat(np.arange(16, dtype=np.int32), 2)
    ctx = nat.Context(0)
    ctx.set_circuit(n, gates)
    res64 = decompose_batch(ctx, T, run_target, x0, default_options(n, optimization_tolerance=1e-8))
    resmx = decompose_batch(ctx, T, run_target, x0,
                            default_options(n, optimization_tolerance=1e-8, mixed_precision=True, switch_cost=1e-4))
    for res in (res64, resmx):
        for r in range(32):
            assert abs(oracle.optimization_problem(gates, T[run_target[r]], n, res["x"][r]) - res["f"][r]) < 1e-12
    s64 = np.mean(res64["f"] < 1e-8)
    smx = np.mean(resmx["f"] < 1e-8)
    say(f"  success FP64 {s64:.2f} ({res64['rounds']} rounds), mixed {smx:.2f} ({resmx['rounds']} rounds)")
    assert s64 > 0 and smx > 0
