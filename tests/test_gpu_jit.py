"""Run-time specialised kernels on the GPU against the FP64 oracle (and against the
generic kernels): costs and gradients of all four cost variants, apply_to, complex64
costs, and the batched optimiser."""
import sys

import numpy as np
import pytest

import oracle as orc
from squander_hip import decompose_batch, default_options, initial_guess
from squander_hip import _native as nat
from squander_hip import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _ctx(n, gates, U, jit):
    """A context on the register-tile kernels (the fused-block path off), specialised
    to the circuit when jit."""
    ctx = nat.Context(0)
    ctx.set_circuit(n, gates)
    ctx.set_target(U)
    ctx.set_jit(jit)
    ctx.set_fused_cost(False)
    return ctx


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 8, 10])
def test_jit_cost_and_gradient_match_oracle(oracle, n):
    say(f"jit cost+grad n={n}")
    gates = W.mixed_gate_circuit(n, 24, 700 + n)
    P = W.param_count(gates)
    d = 1 << n
    U = W.haar_unitary(d, 800 + n).conj().T.copy()
    B = 6
    ps = W.random_params(P, 900 + n, batch=B)
    for variant in (nat.COST_FROBENIUS, nat.COST_HILBERT_SCHMIDT, nat.COST_INFIDELITY, nat.COST_ABS_TRACE):
        ctx = _ctx(n, gates, U, True)
        ctx.set_cost_function(variant)
        f = np.empty(B)
        g = np.empty((B, P))
        ctx.check(ctx._L.sqh_optimization_problem_combined_batch(ctx.h, ps.ctypes.data, B, nat.PREC_F64,
                                                                 f.ctypes.data, g.ctypes.data))
        c = np.empty(B)
        ctx.check(ctx._L.sqh_optimization_problem_batch(ctx.h, ps.ctypes.data, B, nat.PREC_F64, c.ctypes.data))
        for b in range(B):
            fo, go = oracle.optimization_problem_combined(gates, U, n, ps[b], variant=variant)
            assert abs(f[b] - fo) < 1e-12 and abs(c[b] - fo) < 1e-12
            assert np.max(np.abs(g[b] - go)) < 1e-12
        ms, compiles = ctx.jit_stats()
        assert compiles >= 1
    say(f"  compile {ms:.0f} ms")


@pytest.mark.parametrize("n", [2, 4, 7, 9])
def test_jit_apply_to_matches_oracle(oracle, n):
    say(f"jit apply_to n={n}")
    import torch
    gates = W.mixed_gate_circuit(n, 30, 1100 + n)
    P = W.param_count(gates)
    d = 1 << n
    m = max(1, d // 2 + 1)
    ctx = nat.Context(0)
    ctx.set_circuit(n, gates)
    ctx.set_jit(True)
    rng = np.random.default_rng(n)
    B = 3
    X = rng.standard_normal((B, d, m)) + 1j * rng.standard_normal((B, d, m))
    ps = W.random_params(P, 1200 + n, batch=B)
    Xd = torch.from_numpy(X.copy()).cuda()
    pd = torch.from_numpy(ps).cuda()
    torch.cuda.synchronize()
    ctx.check(ctx._L.sqh_apply_to_dev(ctx.h, pd.data_ptr(), P, Xd.data_ptr(), d, m, m, d * m, B, 0))
    ctx.synchronize()
    R = Xd.cpu().numpy()
    for b in range(B):
        ref = oracle.apply_to(gates, ps[b], X[b])
        assert np.linalg.norm(R[b] - ref) / np.linalg.norm(ref) < 1e-12


def test_jit_complex64_cost(oracle):
    say("jit complex64 cost n=8")
    n = 8
    gates = W.ansatz_8q_40(2026)
    P = W.param_count(gates)
    U = W.haar_unitary(256, 2026).conj().T.copy()
    ctx = _ctx(n, gates, U, True)
    ps = W.random_params(P, 7, batch=4)
    c = np.empty(4)
    ctx.check(ctx._L.sqh_optimization_problem_batch(ctx.h, ps.ctypes.data, 4, nat.PREC_F32, c.ctypes.data))
    ref = oracle.optimization_problem_batched(gates, U, n, ps, nthreads=4)
    assert np.max(np.abs(c - ref)) < 1e-5


def test_jit_decompose_batch_matches_oracle_driver(oracle):
    say("jit decompose_batch 3q")
    n = 3
    gates = W.ansatz_layers(n, W.alternating_pairs(n, 4))
    P = W.param_count(gates)
    rng = np.random.default_rng(31)
    T = np.array([oracle.apply_to(gates, rng.uniform(0, 2 * np.pi, P), np.eye(8)).conj().T for _ in range(4)])
    x0 = np.array([initial_guess(P, s) for s in range(1, 5)])
    ctx = nat.Context(0)
    ctx.set_circuit(n, gates)
    ctx.set_jit(True)
    # the circuit-specialised kernels serve the lockstep rounds
    res = decompose_batch(ctx, T, np.arange(4, dtype=np.int32), x0,
                          default_options(n, optimization_tolerance=1e-8, mixed_precision=True, engine="rounds"))
    for r in range(4):
        fo, xo, co = oracle.decompose_custom(gates, T[r], n, r + 1, optimization_tolerance=1e-8)
        assert abs(oracle.optimization_problem(gates, T[r], n, res["x"][r]) - res["f"][r]) < 1e-12
        assert (fo < 1e-8) == (res["f"][r] < 1e-8) or abs(fo - res["f"][r]) < 1e-10
    assert ctx.jit_stats()[1] >= 1


def test_jit_gradient_8q_ansatz_matches_oracle(oracle):
    """The 264-parameter 8-qubit circuit: specialised cost+gradient against the oracle."""
    say("jit 8q ansatz gradient")
    n = 8
    gates = W.ansatz_8q_40(2026)
    P = W.param_count(gates)
    U = W.haar_unitary(256, 2026).conj().T.copy()
    ctx = _ctx(n, gates, U, True)
    ps = W.random_params(P, 5, batch=2)
    f = np.empty(2)
    g = np.empty((2, P))
    ctx.check(ctx._L.sqh_optimization_problem_combined_batch(ctx.h, ps.ctypes.data, 2, nat.PREC_F64, f.ctypes.data,
                                                             g.ctypes.data))
    for b in range(2):
        fo, go = oracle.optimization_problem_combined(gates, U, n, ps[b])
        assert abs(f[b] - fo) < 1e-12
        assert np.max(np.abs(g[b] - go)) < 1e-12
    say(f"  compile {ctx.jit_stats()[0]:.0f} ms")


def test_jit_rectangular_trace_offset_matches_oracle(oracle):
    """The reference's 8-qubit 256 x 160 target with trace offset 80
    (test_optmization_problem_combined.py shape), specialised kernels."""
    say("jit 8q 256x160 offset 80")
    n = 8
    gates = W.ansatz_layers(n, W.random_adjacent_pairs(n, 12, 42))
    P = W.param_count(gates)
    rng = np.random.default_rng(42)
    U = W.haar_unitary(256, 43)[:, :160].copy()
    ctx = nat.Context(0)
    ctx.set_circuit(n, gates)
    ctx.set_target(U)
    ctx.set_trace_offset(80)
    ctx.set_jit(True)
    ps = rng.uniform(0, 2 * np.pi, (3, P))
    f = np.empty(3)
    g = np.empty((3, P))
    ctx.check(ctx._L.sqh_optimization_problem_combined_batch(ctx.h, ps.ctypes.data, 3, nat.PREC_F64, f.ctypes.data,
                                                             g.ctypes.data))
    c = np.empty(3)
    ctx.check(ctx._L.sqh_optimization_problem_batch(ctx.h, ps.ctypes.data, 3, nat.PREC_F64, c.ctypes.data))
    for b in range(3):
        fo, go = oracle.optimization_problem_combined(gates, U, n, ps[b], trace_offset=80)
        assert abs(f[b] - fo) < 1e-12 and abs(c[b] - fo) < 1e-12
        assert np.max(np.abs(g[b] - go)) < 1e-12


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 8, 10])
def test_jit_apply_derivate_to_matches_oracle(oracle, n):
    """Circuit-specialised derivative matrices (Gates_block::apply_derivate_to) against
    the oracle's derivative matrices, square and rectangular inputs (1e-12 relative)."""
    say(f"jit apply_derivate_to n={n}")
    gates = W.mixed_gate_circuit(n, 6 * n + 6, 1200 + n)
    P = W.param_count(gates)
    d = 1 << n
    p = W.random_params(P, 1300 + n)
    rng = np.random.default_rng(1400 + n)
    for m in sorted({d, max(1, d // 2 + 3)}):
        X = rng.standard_normal((d, m)) + 1j * rng.standard_normal((d, m))
        ref = oracle.apply_derivate_to(gates, p, X)
        for jit in (True, False):
            ctx = nat.Context(0)
            ctx.set_circuit(n, gates)
            ctx.set_jit(jit)
            ctx.set_fused_cost(False)  # the register-tile store kernels (the fused path: test_gpu_fused_cost)
            out = np.empty((P, d, m), dtype=np.complex128)
            ctx.check(ctx._L.sqh_apply_derivate_to(ctx.h, p.ctypes.data, X.ctypes.data, d, m, m, out.ctypes.data))
            for k in range(P):
                err = np.linalg.norm(out[k] - ref[k]) / max(np.linalg.norm(ref[k]), 1e-300)
                assert err < 1e-12, (jit, m, k, err)
            if jit:
                assert ctx.jit_stats()[1] == 1


def test_decomposition_config_enables_specialised_kernels(oracle):
    """N_Qubit_Decomposition_custom(config={"specialised_kernels": True}) evaluates
    through the run-time compiled kernels, with the oracle's results."""
    from squander_hip import Circuit, N_Qubit_Decomposition_custom
    say("decomposition class with specialised kernels")
    n = 5
    gates = W.mixed_gate_circuit(n, 24, 1500)
    P = W.param_count(gates)
    U = W.haar_unitary(32, 1501).conj().T.copy()
    p = W.random_params(P, 1502)
    dec = N_Qubit_Decomposition_custom(U, config={"specialised_kernels": True})
    dec.set_Gate_Structure(Circuit.from_gate_array(n, gates))
    f, g = dec.Optimization_Problem_Combined(p)
    fo, go = oracle.optimization_problem_combined(gates, U, n, p)
    assert abs(f - fo) < 1e-12 and np.abs(g - go).max() < 1e-12
    assert dec._ctx.jit_stats()[1] >= 1
