"""Edge shapes of the hot path against the oracle: state vectors (one column, the
reference's cols == 1 dispatch, Gate.cpp:1707-1765), circuits without gates or without
parameters, a one-column cost target, and a batch of one."""
import sys

import numpy as np
import pytest

from conftest import rel_fro
from squander_hip import Circuit, N_Qubit_Decomposition_custom
from squander_hip import _native as nat
from squander_hip import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 10, 12])
def test_state_vector_apply_to(oracle, n):
    say(f"state vector n={n}")
    gates = W.mixed_gate_circuit(n, 5 * n + 6, seed=2500 + n)
    p = W.random_params(W.param_count(gates), 2600 + n)
    d = 1 << n
    rng = np.random.default_rng(n)
    x = (rng.standard_normal((d, 1)) + 1j * rng.standard_normal((d, 1))).astype(np.complex128)
    ref = oracle.apply_to(gates, p, x)
    c = Circuit.from_gate_array(n, gates)
    y = x.copy()
    c.apply_to(p, y)
    assert rel_fro(y, ref) < 1e-12
    y32 = x.astype(np.complex64)
    c.apply_to(p.astype(np.float32), y32, is_f32=True)
    assert rel_fro(y32, ref) < 1e-5
    # a 1-D state is taken as one column
    z = x[:, 0].copy()
    c.apply_to(p, z)
    assert rel_fro(z[:, None], ref) < 1e-12


def test_empty_and_parameter_free_circuits(oracle):
    say("empty and parameter-free circuits")
    X = (np.random.default_rng(1).standard_normal((8, 8)) + 0j)
    empty = Circuit(3)
    Y = X.copy()
    empty.apply_to(np.zeros(0), Y)
    assert np.array_equal(Y, X)
    assert np.array_equal(empty.get_Matrix(np.zeros(0)), np.eye(8))
    assert empty.apply_derivate_to(np.zeros(0), X.copy()) == []
    g = np.array([[W.GATE_CNOT, 0, 1, 0], [W.GATE_CZ, 2, 0, 0], [W.GATE_CNOT, 2, 1, 0]], dtype=np.int32)
    c = Circuit.from_gate_array(3, g)
    assert c.get_Parameter_Num() == 0
    Y = X.copy()
    c.apply_to(np.zeros(0), Y)
    assert rel_fro(Y, oracle.apply_to(g, np.zeros(0), X)) < 1e-15
    # cost and (empty) gradient of a parameter-free circuit
    U = W.haar_unitary(8, 3).conj().T.copy()
    dec = N_Qubit_Decomposition_custom(U, config={})
    dec.set_Gate_Structure(c)
    f, grad = dec.Optimization_Problem_Combined(np.zeros(0))
    assert grad.shape == (0,)
    assert abs(f - oracle.optimization_problem(g, U, 3, np.zeros(0))) < 1e-14


@pytest.mark.parametrize("n", [3, 6, 9])
def test_one_column_target_and_batch_of_one(oracle, n):
    """A d x 1 cost target (FROBENIUS: 1 - Re M[off, 0]) with a trace offset, and the
    batched entry points called with one vector."""
    say(f"one-column target n={n}")
    gates = W.ansatz_layers(n, W.alternating_pairs(n, 3))
    P = W.param_count(gates)
    d = 1 << n
    U = W.haar_unitary(d, 40 + n)[:, :1].copy()
    for off in (0, d - 1):
        dec = N_Qubit_Decomposition_custom(U, config={})
        dec.set_Gate_Structure(Circuit.from_gate_array(n, gates))
        dec.set_Trace_Offset(off)
        p = W.random_params(P, 50 + n)
        f1 = dec.Optimization_Problem_Batch(p[None, :])
        f, g = dec.Optimization_Problem_Combined(p)
        fo, go = oracle.optimization_problem_combined(gates, U, n, p, trace_offset=off)
        assert f1.shape == (1,)
        assert abs(f1[0] - fo) < 1e-12 and abs(f - fo) < 1e-12
        assert np.max(np.abs(g - go)) < 1e-12
