"""GPU parity of the correction cost variants (FROBENIUS_NORM_CORRECTION1/2,
HILBERT_SCHMIDT_TEST_CORRECTION1/2) against the CPU oracle.

Bars: FP64 costs and gradients within 1e-12 absolute of the oracle; complex64 within
1e-5; the batched optimiser's reported costs equal the oracle's cost of the returned
parameters within 1e-12.  Every evaluation goes through the C ABI (libsqhip.so).
"""
import sys

import numpy as np
import pytest

import oracle as orc
from squander_hip import Circuit, N_Qubit_Decomposition_custom, PREC_F32, decompose_batch, default_options
from squander_hip import initial_guess
from squander_hip import _native as nat
from squander_hip import workloads as W

pytestmark = pytest.mark.gpu

CORR = [nat.COST_FROBENIUS_CORRECTION1, nat.COST_FROBENIUS_CORRECTION2, nat.COST_HILBERT_SCHMIDT_CORRECTION1,
        nat.COST_HILBERT_SCHMIDT_CORRECTION2]
WEIGHTS = (0.36, 0.3, 0.7)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def circuit(n, seed):
    gates = W.ansatz_layers(n, W.random_adjacent_pairs(n, 2 * n, seed) if n > 1 else [])
    if n >= 3:
        extra = W.mixed_gate_circuit(n, 6, seed=seed + 1).copy()
        P0 = W.param_count(gates)
        for row in extra:
            if row[0] in (7, 8, 9, 10, 14):
                row[3] += P0
        gates = np.vstack([gates, extra])
    return gates


@pytest.mark.parametrize("weights", [None, WEIGHTS])
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 10])
@pytest.mark.parametrize("variant", CORR)
def test_cost_and_gradient_match_oracle(oracle, variant, n, weights):
    say(f"correction variant={variant} n={n} weights={weights}")
    gates = circuit(n, 300 + n)
    P = W.param_count(gates)
    U = W.haar_unitary(1 << n, 310 + n).conj().T.copy()
    B = 3
    ps = W.random_params(P, 320 + n, batch=B)
    dec = N_Qubit_Decomposition_custom(U)
    dec.set_Gate_Structure(Circuit.from_gate_array(n, gates))
    dec.set_Cost_Function_Variant(variant)
    oracle.correction = orc.DEFAULT_CORRECTION
    if weights is not None:
        dec.set_Correction_Weights(*weights)
        oracle.correction = weights
    try:
        costs = dec.Optimization_Problem_Batch(ps)
        f, g = dec.Optimization_Problem_Combined_Batch(ps)
        for b in range(B):
            fr, gr = oracle.optimization_problem_combined(gates, U, n, ps[b], variant)
            assert abs(costs[b] - fr) < 1e-12
            assert abs(f[b] - fr) < 1e-12
            assert np.abs(g[b] - gr).max() < 1e-12
        assert abs(dec.Optimization_Problem(ps[0]) - costs[0]) < 1e-15
        c32 = dec.Optimization_Problem_Batch(ps, precision=PREC_F32)
        f32, g32 = dec.Optimization_Problem_Combined_Batch(ps, precision=PREC_F32)
        assert np.abs(c32 - costs).max() < 1e-5
        assert np.abs(f32 - costs).max() < 1e-5
        assert np.abs(g32 - g).max() < 1e-5
    finally:
        oracle.correction = orc.DEFAULT_CORRECTION


@pytest.mark.parametrize("variant", [nat.COST_FROBENIUS_CORRECTION1, nat.COST_FROBENIUS_CORRECTION2])
def test_frobenius_corrections_with_trace_offset_8q(oracle, variant):
    """The reference's 8-qubit trace-offset shape (256 x 160 target, offset 80)."""
    say(f"correction variant={variant} 8q offset 80")
    gates = W.ansatz_layers(8, W.random_adjacent_pairs(8, 12, 44))
    P = W.param_count(gates)
    U = W.haar_unitary(256, 45).conj().T.copy()[:, :160].copy()
    ps = W.random_params(P, 46, batch=2)
    dec = N_Qubit_Decomposition_custom(U)
    dec.set_Trace_Offset(80)
    dec.set_Gate_Structure(Circuit.from_gate_array(8, gates))
    dec.set_Cost_Function_Variant(variant)
    f, g = dec.Optimization_Problem_Combined_Batch(ps)
    for b in range(2):
        fr, gr = oracle.optimization_problem_combined(gates, U, 8, ps[b], variant, 80)
        assert abs(f[b] - fr) < 1e-12
        assert np.abs(g[b] - gr).max() < 1e-12


@pytest.mark.parametrize("variant", [nat.COST_FROBENIUS, nat.COST_HILBERT_SCHMIDT, nat.COST_INFIDELITY] + CORR)
def test_streaming_path_costs_11q(oracle, variant):
    """Above the register tile the cost comes from the streamed matrices and the
    channel extraction kernel (11 qubits, 2048 x 8 target)."""
    say(f"correction variant={variant} 11q streaming")
    n = 11
    gates = W.ansatz_layers(n, W.random_adjacent_pairs(n, 3, 47))
    P = W.param_count(gates)
    rng = np.random.default_rng(48)
    U = (rng.standard_normal((2048, 8)) + 1j * rng.standard_normal((2048, 8))) / 8
    ps = W.random_params(P, 49, batch=2)
    dec = N_Qubit_Decomposition_custom(U)
    dec.set_Gate_Structure(Circuit.from_gate_array(n, gates))
    dec.set_Cost_Function_Variant(variant)
    costs = dec.Optimization_Problem_Batch(ps)
    for b in range(2):
        assert abs(costs[b] - oracle.optimization_problem(gates, U, n, ps[b], variant)) < 1e-12


def test_reachable_target_zero_at_solution(oracle):
    """At the generating angles C.U = I: every correction variant is 0 and its gradient
    vanishes (the cost the reference test_parametric_circuit.py:234 asserts on)."""
    say("correction variants at the solution")
    n = 4
    gates = W.ansatz_layers(n, W.random_adjacent_pairs(n, 6, 50))
    P = W.param_count(gates)
    th = W.random_params(P, 51)
    U = oracle.apply_to(gates, th, np.eye(16)).conj().T.copy()
    dec = N_Qubit_Decomposition_custom(U)
    dec.set_Gate_Structure(Circuit.from_gate_array(n, gates))
    for v in CORR:
        dec.set_Cost_Function_Variant(v)
        f, g = dec.Optimization_Problem_Combined(th)
        assert abs(f) < 1e-12
        assert np.abs(g).max() < 1e-12


@pytest.mark.parametrize("variant", CORR)
def test_batched_optimiser_corrections(oracle, variant):
    """Every correction variant in the batched optimiser (the Hilbert-Schmidt ones with
    a forward pass for the seed weights each round): reported costs equal the oracle's
    cost of the returned parameters."""
    say(f"decompose_batch correction variant={variant}")
    n = 3
    gates = W.ansatz_layers(n, W.alternating_pairs(n, 4))
    P = W.param_count(gates)
    rng = np.random.default_rng(52)
    T = np.array([oracle.apply_to(gates, rng.uniform(0, 2 * np.pi, P), np.eye(8)).conj().T for _ in range(4)])
    x0 = np.array([initial_guess(P, s, "RANDOM") for s in range(1, 5)])
    ctx = nat.Context(0)
    ctx.set_circuit(n, gates)
    ctx.set_cost_function(variant)
    res = decompose_batch(ctx, T, np.arange(4, dtype=np.int32), x0,
                          default_options(n, optimization_tolerance=1e-8, max_outer_iterations=2))
    for r in range(4):
        f_check = oracle.optimization_problem(gates, T[r], n, res["x"][r], variant)
        say(f"  run {r}: f={res['f'][r]:.3e} evals={res['evals'][r]}")
        assert abs(f_check - res["f"][r]) < 1e-12


def test_batched_optimiser_hs_correction_mixed_precision(oracle):
    """Hilbert-Schmidt correction with the mixed-precision optimiser (complex64 and
    complex128 evaluations in one round share the seed buffer)."""
    say("decompose_batch HS correction 2, mixed precision")
    n = 3
    gates = W.ansatz_layers(n, W.alternating_pairs(n, 4))
    P = W.param_count(gates)
    rng = np.random.default_rng(53)
    T = np.array([oracle.apply_to(gates, rng.uniform(0, 2 * np.pi, P), np.eye(8)).conj().T for _ in range(3)])
    x0 = np.array([initial_guess(P, s, "RANDOM") for s in range(11, 14)])
    ctx = nat.Context(0)
    ctx.set_circuit(n, gates)
    ctx.set_cost_function(nat.COST_HILBERT_SCHMIDT_CORRECTION2)
    res = decompose_batch(ctx, T, np.arange(3, dtype=np.int32), x0,
                          default_options(n, optimization_tolerance=1e-8, max_outer_iterations=2,
                                          mixed_precision=1))
    for r in range(3):
        f_check = oracle.optimization_problem(gates, T[r], n, res["x"][r], nat.COST_HILBERT_SCHMIDT_CORRECTION2)
        say(f"  run {r}: f={res['f'][r]:.3e} evals={res['evals'][r]}")
        assert abs(f_check - res["f"][r]) < 1e-12


def test_correction_weights_validated():
    ctx = nat.Context(0)
    with pytest.raises(nat.SqhError):
        ctx.set_correction_weights(-1.0, 0.5, 0.5)
