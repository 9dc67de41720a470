"""Concurrent callers.  The reference calls the path from TBB workers on distinct
matrices at once (Optimization_Interface.cpp:1018-1025, Gates_block.cpp:1142-1147); the
library keeps one stream per context and a mutex around the context's state, and ctypes
releases the GIL during each call, so these threads really overlap inside the library.
Every concurrent result must equal the serial result bit for bit."""
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from squander_hip import Circuit, N_Qubit_Decomposition_custom
from squander_hip import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_concurrent_calls_share_one_context(oracle):
    """Eight threads evaluate batched costs and cost+gradient on one decomposition
    object (one context) at once."""
    say("threads on one context")
    n = 5
    gates = W.ansatz_layers(n, W.alternating_pairs(n, 6))
    P = W.param_count(gates)
    T = W.haar_unitary(1 << n, 77).conj().T.copy()
    dec = N_Qubit_Decomposition_custom(T, config={})
    dec.set_Gate_Structure(Circuit.from_gate_array(n, gates))
    params = [W.random_params(P, 200 + s, batch=16) for s in range(8)]
    serial_f = [dec.Optimization_Problem_Batch(p) for p in params]
    serial_fg = [dec.Optimization_Problem_Combined_Batch(p) for p in params]

    def work(i):
        p = params[i % 8]
        if i % 2:
            return dec.Optimization_Problem_Batch(p)
        return dec.Optimization_Problem_Combined_Batch(p)

    with ThreadPoolExecutor(8) as ex:
        res = list(ex.map(work, range(48)))
    for i, r in enumerate(res):
        if i % 2:
            assert np.array_equal(r, serial_f[i % 8]), i
        else:
            assert np.array_equal(r[0], serial_fg[i % 8][0]) and np.array_equal(r[1], serial_fg[i % 8][1]), i
    f0 = oracle.optimization_problem(gates, T, n, params[3][5])
    assert abs(serial_f[3][5] - f0) < 1e-12


def test_concurrent_calls_on_separate_contexts(oracle):
    """Four threads, each with its own circuit object (its own context and stream), run
    apply_to on their own matrices at once."""
    say("threads on separate contexts")
    jobs = []
    for k, n in enumerate((4, 6, 7, 9)):
        gates = W.mixed_gate_circuit(n, 30, seed=2100 + k)
        p = W.random_params(W.param_count(gates), 2200 + k)
        X = np.random.default_rng(k).standard_normal((1 << n, 24)) + 0j
        jobs.append((n, gates, p, X, Circuit.from_gate_array(n, gates)))

    def work(i):
        n, gates, p, X, c = jobs[i % 4]
        Y = X.copy()
        c.apply_to(p, Y)
        return Y

    serial = [work(i) for i in range(4)]
    with ThreadPoolExecutor(4) as ex:
        res = list(ex.map(work, range(32)))
    for i, Y in enumerate(res):
        assert np.array_equal(Y, serial[i % 4]), i
    for i, (n, gates, p, X, _) in enumerate(jobs):
        ref = oracle.apply_to(gates, p, X)
        assert np.linalg.norm(serial[i] - ref) / np.linalg.norm(ref) < 1e-12
    # fresh circuit objects whose first use (context creation) is concurrent
    jobs[:] = [(n, gates, p, X, Circuit.from_gate_array(n, gates)) for n, gates, p, X, _ in jobs]
    with ThreadPoolExecutor(8) as ex:
        res = list(ex.map(work, range(16)))
    for i, Y in enumerate(res):
        assert np.array_equal(Y, serial[i % 4]), i
