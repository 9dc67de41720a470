"""GPU batched decomposition (sqh_decompose_batch) against the oracle's restated
driver (or_decompose_custom: BFGS_Powell + Wolfe line search + iteration loops with
restart jitter + outer restarts) from the same mt19937 start points.  The bar:
every GPU-reported cost equals the FP64 oracle's cost of the returned parameters
within 1e-12; a run that takes as many evaluations as the oracle's (the same
trajectory up to rounding) ends within 1e-10 of the oracle's final cost; a run whose
trajectory forked on rounding reaches the same outcome (both below the tolerance or
both above); and at least half of the runs take the oracle's evaluation count."""
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


def check_against_oracle(pairs, tol):
    """pairs: (gpu cost, gpu evaluations, oracle cost, oracle evaluations) per run."""
    same = 0
    for fg, eg, fo, eo in pairs:
        if eg == eo:
            same += 1
            assert abs(fo - fg) <= 1e-10, (fg, fo, eg)
        else:
            assert (fo < tol) == (fg < tol), (fg, fo, eg, eo)
    say(f"  {same} of {len(pairs)} runs took the oracle's evaluation count")
    assert 2 * same >= len(pairs)


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


ENGINES = ["persistent", "rounds", "persistent_jit"]


def engine_opts(ctx, engine):
    """The engine option, and for persistent_jit the circuit-specialised evaluation
    (run-time compiled, up to 4 qubits) switched on."""
    if engine == "persistent_jit":
        ctx.set_jit(True)
        return "persistent"
    return engine


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("n,layers,loops", [(2, 3, 3), (2, 2, 1), (3, 4, 1), (3, 8, 1), (3, 4, 2), (4, 3, 1), (5, 3, 1)])
def test_batch_matches_oracle_driver(oracle, n, layers, loops, engine):
    """Iteration loops: 3 at 2 qubits is the reference's default (BFGS.cpp:65-83 with
    Optimization_Interface.cpp:149); 2 at 3 qubits stands for config max_iteration_loops.
    Both engines: the persistent solver (one wavefront per run, one launch) and the
    lockstep rounds."""
    from squander_hip.optimizer import restart_jitter
    say(f"decompose_batch vs oracle n={n} layers={layers} loops={loops} engine={engine}")
    gates = W.ansatz_layers(n, W.alternating_pairs(n, layers))
    P = W.param_count(gates)
    T = reachable_targets(n, gates, 6, 100 + n)
    seeds = list(range(1, 7))
    x0 = np.array([initial_guess(P, s, "RANDOM") for s in seeds])
    jitter = np.array([restart_jitter(P, s, "RANDOM", 4 * loops) for s in seeds]) if loops > 1 else None
    ctx = nat.Context(0)
    ctx.set_circuit(n, gates)
    opts = default_options(n, optimization_tolerance=1e-8, max_outer_iterations=4, iteration_loops=loops,
                           engine=engine_opts(ctx, engine))
    res = decompose_batch(ctx, T, np.arange(6, dtype=np.int32), x0, opts, jitter)
    pairs = []
    for r in range(6):
        fo, xo, co = oracle.decompose_custom(gates, T[r], n, seeds[r], max_outer_iterations=4,
                                             iteration_loops=loops, optimization_tolerance=1e-8)
        fg = res["f"][r]
        f_check = oracle.optimization_problem(gates, T[r], n, res["x"][r])
        say(f"  run {r}: gpu f={fg:.3e} evals={res['evals'][r]}  oracle f={fo:.3e} evals={co}")
        assert abs(f_check - fg) < 1e-12
        pairs.append((fg, res["evals"][r], fo, co))
    check_against_oracle(pairs, 1e-8)


@pytest.mark.parametrize("engine", ENGINES)
def test_mixed_gate_circuit_4q_matches_oracle_driver(oracle, engine):
    """Every gate kind (U3, RX, RY, RZ, CRY, CNOT, CZ; controls below and above the
    target) at 4 qubits, square target: the engines (the circuit-specialised persistent
    solver splits each column over four lanes here) against the oracle driver."""
    n = 4
    say(f"mixed-gate 4q circuit vs oracle engine={engine}")
    gates = W.mixed_gate_circuit(n, 26, 4711)
    P = W.param_count(gates)
    T = reachable_targets(n, gates, 6, 4712)
    seeds = list(range(21, 27))
    x0 = np.array([initial_guess(P, s, "RANDOM") for s in seeds])
    ctx = nat.Context(0)
    ctx.set_circuit(n, gates)
    opts = default_options(n, optimization_tolerance=1e-8, max_outer_iterations=2, engine=engine_opts(ctx, engine))
    res = decompose_batch(ctx, T, np.arange(6, dtype=np.int32), x0, opts)
    pairs = []
    for r in range(6):
        fo, xo, co = oracle.decompose_custom(gates, T[r], n, seeds[r], max_outer_iterations=2,
                                             optimization_tolerance=1e-8)
        fg = res["f"][r]
        say(f"  run {r}: gpu f={fg:.3e} evals={res['evals'][r]}  oracle f={fo:.3e} evals={co}")
        assert abs(oracle.optimization_problem(gates, T[r], n, res["x"][r]) - fg) < 1e-12
        pairs.append((fg, res["evals"][r], fo, co))
    check_against_oracle(pairs, 1e-8)


@pytest.mark.parametrize("n,layers,fused", [(4, 3, True), (5, 3, True), (7, 3, True)])
def test_rounds_with_fused_gradient_match_oracle_driver(oracle, n, layers, fused):
    """The lockstep rounds evaluating through the fused blocks' adjoint (switched on):
    reported costs equal the oracle's cost of the returned parameters, runs match the
    oracle driver, and the mixed-precision rounds (complex64 fused blocks, FP64 finish)
    report FP64 costs."""
    say(f"rounds with fused gradient n={n} layers={layers} fused={fused}")
    gates = W.ansatz_layers(n, W.alternating_pairs(n, layers))
    P = W.param_count(gates)
    T = reachable_targets(n, gates, 6, 300 + n)
    seeds = list(range(11, 17))
    x0 = np.array([initial_guess(P, s, "RANDOM") for s in seeds])
    ctx = nat.Context(0)
    ctx.set_circuit(n, gates)
    ctx.set_fused_cost(fused)
    res = decompose_batch(ctx, T, np.arange(6, dtype=np.int32), x0,
                          default_options(n, optimization_tolerance=1e-8, max_outer_iterations=2, engine="rounds"))
    pairs = []
    for r in range(6):
        fo, xo, co = oracle.decompose_custom(gates, T[r], n, seeds[r], max_outer_iterations=2,
                                             optimization_tolerance=1e-8)
        fg = res["f"][r]
        say(f"  run {r}: gpu f={fg:.3e} evals={res['evals'][r]}  oracle f={fo:.3e} evals={co}")
        assert abs(oracle.optimization_problem(gates, T[r], n, res["x"][r]) - fg) < 1e-12
        pairs.append((fg, res["evals"][r], fo, co))
    check_against_oracle(pairs, 1e-8)
    resmx = decompose_batch(ctx, T, np.arange(6, dtype=np.int32), x0,
                            default_options(n, optimization_tolerance=1e-8, max_outer_iterations=2, engine="rounds",
                                            mixed_precision=True, switch_cost=1e-4))
    for r in range(6):
        assert abs(oracle.optimization_problem(gates, T[r], n, resmx["x"][r]) - resmx["f"][r]) < 1e-12
    say(f"  mixed: success {np.mean(resmx['f'] < 1e-8):.2f} (FP64 {np.mean(res['f'] < 1e-8):.2f})")


@pytest.mark.parametrize("n,seed", [(3, 1), (3, 2), (3, 3), (2, 4), (2, 5)])
def test_start_decomposition_matches_oracle_driver(oracle, n, seed):
    """N_Qubit_Decomposition_custom.Start_Decomposition with config random_seed gives
    the oracle driver's result from the same seed: the start point drawn after the
    constructor's id draw, and at 2 qubits three iteration loops with the restart
    jitter from the same generator (a random start may end in a local minimum; both
    must end in the same one)."""
    say(f"Start_Decomposition {n}q seed {seed}")
    gates = W.ansatz_layers(n, W.alternating_pairs(n, 6 if n == 3 else 2))
    T = reachable_targets(n, gates, 1, 7)[0]
    dec = N_Qubit_Decomposition_custom(T, config={"random_seed": seed, "optimization_tolerance": 1e-10})
    dec.set_Gate_Structure(Circuit.from_gate_array(n, gates))
    dec.Start_Decomposition()
    err = dec.get_Decomposition_Error()
    fo, xo, co = oracle.decompose_custom(gates, T, n, seed, max_outer_iterations=4,
                                         iteration_loops=3 if n == 2 else 1, optimization_tolerance=1e-10)
    eg = int(dec.last_run["evals"][0])
    say(f"  error {err:.3e}, evals {eg}; oracle {fo:.3e}, evals {co}")
    assert abs(err - oracle.optimization_problem(gates, T, n, dec.get_Optimized_Parameters())) < 1e-12
    if eg == co:
        assert abs(err - fo) <= 1e-10
    else:
        assert (fo < 1e-10) == (err < 1e-10)


def test_start_decomposition_above_five_qubits_needs_bfgs():
    """Above 5 qubits the reference's custom decomposition runs ADAM
    (N_Qubit_Decomposition_custom.cpp:63-75), which is not built: Start_Decomposition
    raises until set_Optimizer("BFGS") selects BFGS, which then runs."""
    gates = W.ansatz_layers(6, W.alternating_pairs(6, 1))
    T = reachable_targets(6, gates, 1, 3)[0]
    dec = N_Qubit_Decomposition_custom(T, config={"random_seed": 1, "max_outer_iterations": 1,
                                                  "max_inner_iterations": 30})
    dec.set_Gate_Structure(Circuit.from_gate_array(6, gates))
    with pytest.raises(Exception, match="ADAM"):
        dec.Start_Decomposition()
    dec.set_Optimizer("BFGS")
    dec.Start_Decomposition()
    assert dec.get_Decomposition_Error() < 1.0
    with pytest.raises(Exception):
        dec.set_Optimizer("NEWTON")


def test_decompose_ibm_challenge_fixture(oracle):
    """The reference's data fixture (4th IBM Challenge unitary, data/Umtx.mat, 4 qubits,
    tests/decomposition/test_IBM.py:47-110) decomposed on the GPU through the Python
    mirror with the phase-invariant HS cost: the run from config random_seed 1 reaches the
    oracle driver's result from the same seed, and the phase-corrected error the
    reference test computes stays below its 1e-3 bar."""
    import os
    say("IBM challenge fixture, 4 qubits, HS cost")
    U = np.load(os.path.join(os.path.dirname(__file__), "golden", "ibm_challenge_umtx.npy"))
    gates = W.ansatz_layers(4, [0, 1, 2] * 4)
    dec = N_Qubit_Decomposition_custom(U.conj().T, config={"random_seed": 1, "max_outer_iterations": 8,
                                                           "optimization_tolerance": 1e-8})
    dec.set_Cost_Function_Variant(nat.COST_HILBERT_SCHMIDT)
    dec.set_Gate_Structure(Circuit.from_gate_array(4, gates))
    dec.Start_Decomposition()
    x = dec.get_Optimized_Parameters()
    f = dec.get_Decomposition_Error()
    fo, xo, _ = oracle.decompose_custom(gates, U.conj().T, 4, seed=1, max_outer_iterations=8,
                                        optimization_tolerance=1e-8, variant=orc.COST_HS)
    say(f"  gpu cost {f:.3e}, oracle {fo:.3e}")
    assert abs(f - oracle.optimization_problem(gates, U.conj().T, 4, x, variant=orc.COST_HS)) < 1e-12
    assert abs(f) < 1e-7 and abs(fo) < 1e-7
    C = oracle.apply_to(gates, x, np.eye(16))
    prod = U @ C.conj().T
    prod = prod * np.exp(-1j * np.angle(prod[0, 0]))
    err = np.real(np.trace(np.eye(16) * 2 - prod - prod.conj().T)) / 2
    assert err < 1e-3


@pytest.mark.parametrize("engine", ENGINES)
def test_mixed_precision_batch_4q(oracle, engine):
    say(f"mixed-precision batch 4q engine={engine}")
    n = 4
    gates = W.ansatz_4q_12()
    P = W.param_count(gates)
    T = reachable_targets(n, gates, 16, 44)
    rng = np.random.default_rng(5)
    x0 = rng.uniform(0, 2 * np.pi, (32, P))
    run_target = np.repeat(np.arange(16, dtype=np.int32), 2)
    ctx = nat.Context(0)
    ctx.set_circuit(n, gates)
    eng = engine_opts(ctx, engine)
    res64 = decompose_batch(ctx, T, run_target, x0, default_options(n, optimization_tolerance=1e-8, engine=eng))
    resmx = decompose_batch(ctx, T, run_target, x0,
                            default_options(n, optimization_tolerance=1e-8, mixed_precision=True, switch_cost=1e-4,
                                            engine=eng))
    for res in (res64, resmx):
        for r in range(32):
            assert abs(oracle.optimization_problem(gates, T[run_target[r]], n, res["x"][r]) - res["f"][r]) < 1e-12
    s64 = np.mean(res64["f"] < 1e-8)
    smx = np.mean(resmx["f"] < 1e-8)
    say(f"  success FP64 {s64:.2f} ({res64['rounds']} rounds), mixed {smx:.2f} ({resmx['rounds']} rounds)")
    assert s64 > 0 and smx > 0


@pytest.mark.parametrize("engine", ENGINES[:2])
def test_decompose_problems_matches_oracle_driver(oracle, engine):
    """Problems with their own circuits and qubit counts in one batch, each compared
    with the oracle's driver from the same mt19937 start point."""
    from squander_hip import decompose_problems
    specs = [(2, 2, 1), (3, 4, 2), (4, 3, 3), (5, 2, 4), (3, 6, 5), (4, 5, 6), (2, 3, 7), (5, 4, 8), (6, 2, 9)]
    probs, ref = [], []
    for n, layers, seed in specs:
        say(f"  problem n={n} layers={layers} seed={seed}")
        gates = W.ansatz_layers(n, W.random_adjacent_pairs(n, layers, seed))
        P = W.param_count(gates)
        T = reachable_targets(n, gates, 1, 300 + seed)[0]
        probs.append((n, gates, T, initial_guess(P, seed)))
        ref.append(oracle.decompose_custom(gates, T, n, seed, max_outer_iterations=4, optimization_tolerance=1e-8))
    ctx = nat.Context(0)
    res = decompose_problems(ctx, probs, default_options(4, optimization_tolerance=1e-8, max_outer_iterations=4,
                                                         engine=engine))
    pairs = []
    for i, ((n, gates, T, _), (fo, xo, co)) in enumerate(zip(probs, ref)):
        fg = res["f"][i]
        f_check = oracle.optimization_problem(gates, T, n, res["x"][i])
        say(f"  problem {i} (n={n}): gpu f={fg:.3e} evals={res['evals'][i]}  oracle f={fo:.3e} evals={co}")
        assert res["x"][i].shape == (W.param_count(gates),)
        assert abs(f_check - fg) < 1e-12
        pairs.append((fg, res["evals"][i], fo, co))
    check_against_oracle(pairs, 1e-8)


@pytest.mark.parametrize("engine", ENGINES[:2])
def test_decompose_problems_iteration_loops_match_oracle(oracle, engine):
    """Uneven problems with three iteration loops and per-problem restart jitter."""
    from squander_hip import decompose_problems
    from squander_hip.optimizer import restart_jitter
    specs = [(2, 2, 1), (3, 3, 2), (2, 3, 3), (4, 2, 4)]
    probs, jit, ref = [], [], []
    for n, layers, seed in specs:
        gates = W.ansatz_layers(n, W.random_adjacent_pairs(n, layers, seed))
        P = W.param_count(gates)
        T = reachable_targets(n, gates, 1, 400 + seed)[0]
        probs.append((n, gates, T, initial_guess(P, seed)))
        jit.append(restart_jitter(P, seed, "RANDOM", 4 * 3))
        ref.append(oracle.decompose_custom(gates, T, n, seed, max_outer_iterations=4, iteration_loops=3,
                                           optimization_tolerance=1e-8))
    ctx = nat.Context(0)
    res = decompose_problems(ctx, probs, default_options(4, optimization_tolerance=1e-8, max_outer_iterations=4,
                                                         iteration_loops=3, engine=engine), jit)
    pairs = []
    for i, ((n, gates, T, _), (fo, xo, co)) in enumerate(zip(probs, ref)):
        say(f"  problem {i} (n={n}): gpu f={res['f'][i]:.3e} evals={res['evals'][i]}  oracle f={fo:.3e} evals={co}")
        assert abs(oracle.optimization_problem(gates, T, n, res["x"][i]) - res["f"][i]) < 1e-12
        pairs.append((res["f"][i], res["evals"][i], fo, co))
    check_against_oracle(pairs, 1e-8)


def test_decompose_problems_mixed_precision_uneven(oracle):
    """The uneven workload generator with complex64 descent and FP64 acceptance:
    every reported cost is the FP64 oracle cost of the returned parameters."""
    from squander_hip import decompose_problems
    from squander_hip.uneven_bench import make_problems

    def matrix(n, gates, theta):
        return oracle.apply_to(gates, theta, np.eye(1 << n))

    probs = make_problems(24, seed=11, matrix=matrix)
    say(f"  uneven mixed batch of {len(probs)}: qubits {[p[0] for p in probs]}")
    ctx = nat.Context(0)
    res = decompose_problems(ctx, [p[:4] for p in probs],
                             default_options(4, optimization_tolerance=1e-8, mixed_precision=True))
    for i, (n, gates, U, _, kind) in enumerate(probs):
        f_check = oracle.optimization_problem(gates, U, n, res["x"][i])
        assert abs(f_check - res["f"][i]) < 1e-12, (i, f_check, res["f"][i])
    reach = np.array([p[4] == "reachable" for p in probs])
    say(f"  reachable success {np.mean(res['f'][reach] < 1e-8):.2f}, haar {np.mean(res['f'][~reach] < 1e-8):.2f}")
    assert np.mean(res["f"][reach] < 1e-8) > 0


def test_decompose_problems_rejects_bad_problem():
    from squander_hip import decompose_problems
    ctx = nat.Context(0)
    gates = W.ansatz_layers(3, [0, 1])
    P = W.param_count(gates)
    bad = np.array([[W.GATE_U3, 5, -1, 0]], dtype=np.int32)  # qubit 5 of a 3-qubit problem
    with pytest.raises(nat.SqhError):
        decompose_problems(ctx, [(3, gates, np.eye(8), np.zeros(P)), (3, bad, np.eye(8), np.zeros(3))],
                           default_options(3))
    with pytest.raises(ValueError):
        decompose_problems(ctx, [(3, gates, np.eye(4), np.zeros(P))], default_options(3))


@pytest.mark.parametrize("n,layers,seed,hs", [(2, 2, 1, False), (3, 4, 2, False), (2, 2, 1, True), (3, 4, 2, True)])
def test_agents_matches_oracle(oracle, n, layers, seed, hs):
    """AGENTS with the batched GPU cost follows the oracle's AGENTS from the same
    generator seed and guess: same iteration count, same minimum, same parameters.
    FROBENIUS cost (three-point line search) and HILBERT_SCHMIDT_TEST cost (five-point
    fit minimised by the one-variable BFGS_Powell)."""
    from squander_hip import agents, agents_options
    say(f"AGENTS n={n} layers={layers} seed={seed} hs={hs}")
    gates = W.ansatz_layers(n, W.alternating_pairs(n, layers))
    P = W.param_count(gates)
    T = reachable_targets(n, gates, 1, 500 + seed)[0]
    x0 = np.random.default_rng(seed + 100).uniform(0, 2 * np.pi, P)
    ctx = nat.Context(0)
    ctx.set_circuit(n, gates)
    ctx.set_target(T)
    variant = orc.COST_HS if hs else orc.COST_FROBENIUS
    if hs:
        ctx.set_cost_function(nat.COST_HILBERT_SCHMIDT)
    fg, xg, itg, evg = agents(ctx, x0, seed, agents_options(max_iterations=3000, optimization_tolerance=1e-8))
    fo, xo, ito, evo = oracle.agents(gates, T, n, seed, x0, variant=variant, agent_num=64, max_iterations=3000,
                                     lifetime=1000, tolerance=1e-8)
    say(f"  gpu f={fg:.3e} it={itg} evals={evg}; oracle f={fo:.3e} it={ito} evals={evo}")
    assert abs(oracle.optimization_problem(gates, T, n, xg, variant=variant) - fg) < 1e-10
    assert itg == ito and evg == evo
    assert abs(fg - fo) < 1e-10
    d = np.mod(xg - xo + np.pi, 2 * np.pi) - np.pi
    assert np.max(np.abs(d)) < 1e-6


@pytest.mark.parametrize("n,layers,jit,hs", [(3, 4, False, False), (4, 4, True, False), (8, 2, False, False),
                                             (3, 4, False, True)])
def test_agents_batch_equals_single_runs(n, layers, jit, hs):
    """Runs against several targets advanced in lockstep (one launch per stage) give
    exactly the results of one sqh_agents call per target: the same parameters,
    minimum, iteration and evaluation counts.  Also with a trace offset."""
    from squander_hip import agents, agents_batch, agents_options
    gates = W.ansatz_layers(n, W.alternating_pairs(n, layers))
    P = W.param_count(gates)
    K = 5
    d = 1 << n
    off = 1 if n == 4 else 0
    m = d - off
    say(f"AGENTS batch n={n} layers={layers} jit={jit} hs={hs} K={K} offset={off}")
    Ts = np.array([T[:, :m] for T in reachable_targets(n, gates, K, 700 + n)])
    x0 = np.random.default_rng(n + 900).uniform(0, 2 * np.pi, (K, P))
    seeds = np.array([11 + 7 * k for k in range(K)], dtype=np.uint32)
    # the runs stop at different iterations (lifetime 7 exercises the period steps)
    opts = agents_options(agent_num=16, max_iterations=40 + 15 * n, optimization_tolerance=1e-8, agent_lifetime=7,
                          agent_randomization_rate=0.6)
    ctx = nat.Context(0)
    ctx.set_circuit(n, gates)
    ctx.set_trace_offset(off)
    ctx.set_jit(jit)
    if hs:
        ctx.set_cost_function(nat.COST_HILBERT_SCHMIDT)
    fb, xb, itb, evb = agents_batch(ctx, Ts, x0, seeds, opts)
    for k in range(K):
        ctx.set_target(Ts[k])
        fs, xs, its, evs = agents(ctx, x0[k], int(seeds[k]), opts)
        say(f"  run {k}: f={fs:.3e} it={its} evals={evs}")
        assert (its, evs) == (itb[k], evb[k])
        assert fs == fb[k]
        assert np.array_equal(xs, xb[k])


def test_agents_batch_rejects_bad_shapes():
    from squander_hip import agents_batch
    gates = W.ansatz_layers(3, [0, 1])
    P = W.param_count(gates)
    ctx = nat.Context(0)
    ctx.set_circuit(3, gates)
    with pytest.raises(nat.SqhError):  # 4 rows for a 3-qubit circuit
        agents_batch(ctx, np.zeros((2, 4, 4), complex), np.zeros((2, P)), [1, 2])
    ctx.set_trace_offset(2)
    with pytest.raises(nat.SqhError):  # 7 columns + offset 2 > 8 rows
        agents_batch(ctx, np.zeros((2, 8, 7), complex), np.zeros((2, P)), [1, 2])
    with pytest.raises(ValueError):
        agents_batch(ctx, np.zeros((2, 8, 6), complex), np.zeros((2, P + 1)), [1, 2])


def test_agents_rejects_other_costs():
    from squander_hip import agents
    gates = W.ansatz_layers(2, [0])
    ctx = nat.Context(0)
    ctx.set_circuit(2, gates)
    ctx.set_target(np.eye(4))
    ctx.set_cost_function(nat.COST_INFIDELITY)
    with pytest.raises(nat.SqhError):
        agents(ctx, np.zeros(W.param_count(gates)), 1)


def test_combined_targets_matches_oracle(oracle):
    """Per-element targets: element b against targets[index[b]], cost and gradient."""
    import torch
    say("combined cost+grad with per-element targets")
    for n, variant in ((3, nat.COST_ABS_TRACE), (4, nat.COST_FROBENIUS), (6, nat.COST_HILBERT_SCHMIDT)):
        gates = W.ansatz_layers(n, W.alternating_pairs(n, 5))
        P = W.param_count(gates)
        T = np.array([W.haar_unitary(1 << n, 60 + i) for i in range(5)])
        idx = np.array([3, 0, 4, 4, 1, 2, 0], dtype=np.int32)
        ps = W.random_params(P, 61, batch=len(idx))
        for jit in (False, True):
            ctx = nat.Context(0)
            ctx.set_circuit(n, gates)
            ctx.set_cost_function(variant)
            ctx.set_jit(jit)
            Td, Id, Pd = (torch.from_numpy(a).cuda() for a in (T, idx, ps))
            f = torch.empty(len(idx), dtype=torch.float64, device="cuda")
            g = torch.empty((len(idx), P), dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            ctx.check(ctx._L.sqh_optimization_problem_combined_targets_dev(
                ctx.h, Td.data_ptr(), len(T), Id.data_ptr(), Pd.data_ptr(), len(idx), nat.PREC_F64, f.data_ptr(),
                g.data_ptr()))
            ctx.synchronize()
            for b, i in enumerate(idx):
                fo, go = oracle.optimization_problem_combined(gates, T[i], n, ps[b], variant=variant)
                assert abs(f[b].item() - fo) < 1e-12
                assert np.max(np.abs(g[b].cpu().numpy() - go)) < 1e-12


def test_persistent_engine_with_trace_offset_and_hs_cost(oracle):
    """The persistent solver with a rectangular target and trace offset (FROBENIUS) and
    with the Hilbert-Schmidt cost, against the oracle's driver; the correction variants
    are refused by the persistent engine and run in lockstep rounds under AUTO."""
    n = 3
    gates = W.ansatz_layers(n, W.alternating_pairs(n, 5))
    P = W.param_count(gates)
    T = reachable_targets(n, gates, 4, 810)
    seeds = [21, 22, 23, 24]
    x0 = np.array([initial_guess(P, s) for s in seeds])
    for variant, off, cols in ((nat.COST_FROBENIUS, 2, 6), (nat.COST_HILBERT_SCHMIDT, 0, 8)):
        say(f"  persistent variant={variant} offset={off} cols={cols}")
        Tv = np.ascontiguousarray(T[:, :, :cols])
        ctx = nat.Context(0)
        ctx.set_circuit(n, gates)
        ctx.set_cost_function(variant)
        ctx.set_trace_offset(off)
        res = decompose_batch(ctx, Tv, np.arange(4, dtype=np.int32), x0,
                              default_options(n, optimization_tolerance=1e-8, engine="persistent"))
        pairs = []
        for r in range(4):
            fo, xo, co = oracle.decompose_custom(gates, Tv[r], n, seeds[r], max_outer_iterations=4,
                                                 optimization_tolerance=1e-8, variant=variant, trace_offset=off)
            assert abs(oracle.optimization_problem(gates, Tv[r], n, res["x"][r], variant, off) - res["f"][r]) < 1e-12
            pairs.append((res["f"][r], res["evals"][r], fo, co))
        check_against_oracle(pairs, 1e-8)
    ctx.set_cost_function(nat.COST_FROBENIUS_CORRECTION1)
    ctx.set_trace_offset(0)
    with pytest.raises(nat.SqhError):
        decompose_batch(ctx, T, np.arange(4, dtype=np.int32), x0, default_options(n, engine="persistent"))
    decompose_batch(ctx, T, np.arange(4, dtype=np.int32), x0, default_options(n, max_outer_iterations=1))
