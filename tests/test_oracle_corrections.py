"""Pins the oracle's correction cost variants (FROBENIUS_NORM_CORRECTION1/2 and
HILBERT_SCHMIDT_TEST_CORRECTION1/2: N_Qubit_Decomposition_Cost_Function.cpp:191-380,
564-664; Optimization_Interface.cpp:676-716, 1391-1420) without a GPU.

Checks that do not restate the oracle's loops:
- closed forms on permutation matrices: an X string on k qubits puts all of its
  weight in the k-flip channel, so the cost is 1 - w_k exactly;
- the bit-flip traces against an operator construction (Kronecker products of Pauli X
  and a shift matrix, ``brute_flip_trace``) on random matrices, square and offset;
- gradients against the exact parameter-shift difference of the oracle's own cost
  (the FROBENIUS corrections are linear in C.U) and against central differences
  (the Hilbert-Schmidt corrections are quadratic);
- non-default weights (prev_cost, scale1, scale2), so a dropped square root or a
  swapped scale fails.
"""
import itertools

import numpy as np
import pytest

import oracle as orc
from squander_hip import workloads as W

CORR = [orc.COST_FROBENIUS_C1, orc.COST_FROBENIUS_C2, orc.COST_HS_C1, orc.COST_HS_C2]
WEIGHTS = (0.36, 0.3, 0.7)  # prev_cost, scale1, scale2 -> w1 = 0.18, w2 = 0.42


def x_string(n, qubits):
    d = 1 << n
    P = np.zeros((d, d), dtype=complex)
    mask = sum(1 << q for q in qubits)
    P[np.arange(d) ^ mask, np.arange(d)] = 1.0
    return P


@pytest.fixture
def weighted(oracle):
    oracle.correction = WEIGHTS
    yield oracle
    oracle.correction = orc.DEFAULT_CORRECTION


@pytest.mark.parametrize("n", [2, 3, 5])
def test_closed_forms_on_x_strings(weighted, n):
    prev, s1, s2 = WEIGHTS
    w1, w2 = np.sqrt(prev) * s1, np.sqrt(prev) * s2
    d = 1 << n
    eye = np.eye(d, dtype=complex)
    for v in CORR:
        assert abs(weighted.cost(eye, v)) < 1e-15
    for q in range(n):
        X1 = x_string(n, [q])
        assert weighted.cost(X1, orc.COST_FROBENIUS_C1) == pytest.approx(1.0 - w1, abs=1e-15)
        assert weighted.cost(X1, orc.COST_FROBENIUS_C2) == pytest.approx(1.0 - w1, abs=1e-15)
        assert weighted.cost(X1, orc.COST_HS_C1) == pytest.approx(1.0 - w1, abs=1e-15)
        assert weighted.cost(X1, orc.COST_FROBENIUS) == pytest.approx(1.0, abs=1e-15)
    for q, q2 in itertools.combinations(range(n), 2):
        X2 = x_string(n, [q, q2])
        assert weighted.cost(X2, orc.COST_FROBENIUS_C1) == pytest.approx(1.0, abs=1e-15)
        assert weighted.cost(X2, orc.COST_FROBENIUS_C2) == pytest.approx(1.0 - w2, abs=1e-15)
        assert weighted.cost(X2, orc.COST_HS_C2) == pytest.approx(1.0 - w2, abs=1e-15)
    # three flips fall outside both channels
    if n >= 3:
        X3 = x_string(n, [0, 1, 2])
        assert weighted.cost(X3, orc.COST_FROBENIUS_C2) == pytest.approx(1.0, abs=1e-15)
    # a phase on the one-flip channel: FROBENIUS reads Re, HS reads |.|^2
    Xi = 1j * x_string(n, [0])
    assert weighted.cost(Xi, orc.COST_FROBENIUS_C1) == pytest.approx(1.0, abs=1e-15)
    assert weighted.cost(Xi, orc.COST_HS_C1) == pytest.approx(1.0 - w1, abs=1e-15)


def test_offset_selects_shifted_flips(weighted):
    """Rows offset .. offset + m of an X string: the offset trace of the FROBENIUS
    corrections finds the flips again (cost 1 - w1); the Hilbert-Schmidt ones read the
    un-offset diagonal, as get_trace_with_correction does."""
    n, off, m = 4, 4, 8
    prev, s1, _ = WEIGHTS
    w1 = np.sqrt(prev) * s1
    X = x_string(n, [2])[:, off:off + m]
    assert X.shape == (16, 8)
    assert weighted.cost(X, orc.COST_FROBENIUS_C1, off) == pytest.approx(1.0 - w1, abs=1e-15)
    # without the offset the same columns give a different split between channels
    assert weighted.cost(X, orc.COST_FROBENIUS_C1, 0) == pytest.approx(
        orc.brute_cost(X, orc.COST_FROBENIUS_C1, 0, WEIGHTS), abs=1e-15)
    assert abs(weighted.cost(X, orc.COST_FROBENIUS_C1, 0) - (1.0 - w1)) > 0.1


@pytest.mark.parametrize("shape,off", [((8, 8), 0), ((32, 32), 0), ((32, 12), 20), ((64, 40), 7)])
def test_flip_traces_match_operator_form(oracle, shape, off):
    rng = np.random.default_rng(shape[0] + off)
    M = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    for bits in (1, 2):
        assert abs(oracle.trace_flips(M, bits, off) - orc.brute_flip_trace(M, bits, off)) < 1e-12
    for v in CORR:
        o = off if v in (orc.COST_FROBENIUS_C1, orc.COST_FROBENIUS_C2) else 0
        assert abs(oracle.cost(M, v, o) - orc.brute_cost(M, v, o)) < 1e-12


@pytest.mark.parametrize("variant", CORR)
def test_gradients_pinned(weighted, variant):
    n = 3
    gates = W.mixed_gate_circuit(n, 14, seed=61)
    P = W.param_count(gates)
    U = W.haar_unitary(1 << n, 62).conj().T.copy()
    p = W.random_params(P, 63)
    off = 0
    f0, g = weighted.optimization_problem_combined(gates, U, n, p, variant, off)
    assert abs(f0 - weighted.optimization_problem(gates, U, n, p, variant, off)) < 1e-14
    M = weighted.apply_to(gates, p, U)
    assert abs(f0 - orc.brute_cost(M, variant, off, WEIGHTS)) < 1e-12
    for k in range(P):
        if variant in (orc.COST_FROBENIUS_C1, orc.COST_FROBENIUS_C2):
            pp, pm = p.copy(), p.copy()
            pp[k] += np.pi / 2
            pm[k] -= np.pi / 2
            ref = (weighted.optimization_problem(gates, U, n, pp, variant, off)
                   - weighted.optimization_problem(gates, U, n, pm, variant, off)) / 2
            assert abs(g[k] - ref) < 1e-12
        else:
            h = 1e-6
            pp, pm = p.copy(), p.copy()
            pp[k] += h
            pm[k] -= h
            ref = (weighted.optimization_problem(gates, U, n, pp, variant, off)
                   - weighted.optimization_problem(gates, U, n, pm, variant, off)) / (2 * h)
            assert abs(g[k] - ref) < 1e-8


def test_frobenius_correction_gradient_with_offset(weighted):
    n, off = 4, 4
    gates = W.ansatz_layers(n, W.random_adjacent_pairs(n, 4, 5))
    P = W.param_count(gates)
    U = W.haar_unitary(16, 6).conj().T.copy()[:, :10]
    p = W.random_params(P, 7)
    for variant in (orc.COST_FROBENIUS_C1, orc.COST_FROBENIUS_C2):
        f0, g = weighted.optimization_problem_combined(gates, U, n, p, variant, off)
        M = weighted.apply_to(gates, p, U)
        assert abs(f0 - orc.brute_cost(M, variant, off, WEIGHTS)) < 1e-12
        for k in range(P):
            pp, pm = p.copy(), p.copy()
            pp[k] += np.pi / 2
            pm[k] -= np.pi / 2
            ref = (weighted.optimization_problem(gates, U, n, pp, variant, off)
                   - weighted.optimization_problem(gates, U, n, pm, variant, off)) / 2
            assert abs(g[k] - ref) < 1e-12


def test_plain_variants_unchanged_by_weights(weighted, oracle):
    n = 3
    gates = W.mixed_gate_circuit(n, 10, seed=8)
    U = W.haar_unitary(8, 9).conj().T.copy()
    p = W.random_params(W.param_count(gates), 10)
    M = weighted.apply_to(gates, p, U)
    for v in (orc.COST_FROBENIUS, orc.COST_HS, orc.COST_INFIDELITY):
        assert abs(weighted.cost(M, v) - orc.brute_cost(M, v)) < 1e-13
