"""GPU parity of the hot path against the CPU oracle (run on an MI355X: -m gpu).

Bars (written into each assertion): complex128 matrices within 1e-12 relative
Frobenius norm; costs and gradients within 1e-12 absolute (FP64); complex64
results within the FP32 bar stated per test.  Every call goes through the C ABI
(libsqhip.so) via the reference-shaped Python objects.
"""
import sys

import numpy as np
import pytest

import oracle as orc
from conftest import rel_fro
from squander_hip import Circuit, N_Qubit_Decomposition_custom, PREC_F32, PREC_F64
from squander_hip import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def circuit_of(n, gates):
    return Circuit.from_gate_array(n, gates)


@pytest.mark.parametrize("n", list(range(1, 11)))
def test_apply_to_matches_oracle(oracle, n):
    say(f"apply_to n={n}")
    gates = W.mixed_gate_circuit(n, 6 * n + 10, seed=500 + n)
    P = W.param_count(gates)
    p = W.random_params(P, n)
    d = 1 << n
    rng = np.random.default_rng(n)
    X = (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))).astype(np.complex128)
    ref = oracle.apply_to(gates, p, X)
    c = circuit_of(n, gates)
    Y = X.copy()
    c.apply_to(p, Y)
    assert rel_fro(Y, ref) < 1e-12
    # rectangular, non-power-of-two column count
    m = max(1, (3 * d) // 4 - 1)
    Z = np.ascontiguousarray(X[:, :m])
    c.apply_to(p, Z)
    assert rel_fro(Z, ref[:, :m]) < 1e-12
    # complex64 path (FP32 arithmetic): bar 1e-5 relative
    Z32 = X.astype(np.complex64)
    c.apply_to(p.astype(np.float32), Z32, is_f32=True)
    assert rel_fro(Z32, ref) < 1e-5


@pytest.mark.parametrize("n", [11, 12, 13, 14, 16])
def test_apply_to_streaming_beyond_tile(oracle, n):
    say(f"apply_to streaming n={n}")
    gates = W.mixed_gate_circuit(n, 24, seed=600 + n)
    p = W.random_params(W.param_count(gates), n)
    d = 1 << n
    X = np.random.default_rng(n).standard_normal((d, 40)).astype(np.complex128)
    ref = oracle.apply_to(gates, p, X)
    c = circuit_of(n, gates)
    c.apply_to(p, X)
    assert rel_fro(X, ref) < 1e-12


def test_get_matrix_and_dtype_errors(oracle):
    gates = W.ansatz_layers(3, [0, 1, 0])
    p = W.random_params(W.param_count(gates), 3)
    c = circuit_of(3, gates)
    assert rel_fro(c.get_Matrix(p), oracle.apply_to(gates, p, np.eye(8))) < 1e-12
    with pytest.raises(TypeError):
        c.apply_to(p.astype(np.float32), np.eye(8, dtype=np.complex128))
    with pytest.raises(TypeError):
        c.apply_to(p, np.eye(8, dtype=np.complex64))


@pytest.mark.parametrize("n", [2, 3, 4, 6])
def test_apply_derivate_to_matches_oracle(oracle, n):
    say(f"apply_derivate_to n={n}")
    gates = W.mixed_gate_circuit(n, 5 * n + 6, seed=700 + n)
    P = W.param_count(gates)
    p = W.random_params(P, 7 * n)
    d = 1 << n
    X = W.haar_unitary(d, n)[:, : max(1, d - 1)].copy()
    ref = oracle.apply_derivate_to(gates, p, X)
    got = circuit_of(n, gates).apply_derivate_to(p, X)
    assert len(got) == P
    for k in range(P):
        assert rel_fro(got[k], ref[k]) < 1e-12


def test_apply_derivate_to_8q_ansatz(oracle):
    say("apply_derivate_to 8q ansatz, 264 derivative matrices")
    gates = W.ansatz_8q_40()
    P = W.param_count(gates)
    p = W.random_params(P, 11)
    U = W.haar_unitary(256, 12)
    ref = oracle.apply_derivate_to(gates, p, U)
    got = circuit_of(8, gates).apply_derivate_to(p, U)
    worst = max(rel_fro(got[k], ref[k]) for k in range(P))
    say(f"  worst relative Frobenius error over {P} matrices: {worst:.3e}")
    assert worst < 1e-12


VARIANTS = [orc.COST_FROBENIUS, orc.COST_HS, orc.COST_INFIDELITY, orc.COST_ABS_TRACE]


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10])
@pytest.mark.parametrize("variant", VARIANTS)
def test_cost_and_gradient_batch_match_oracle(oracle, n, variant):
    say(f"cost+grad n={n} variant={variant}")
    layers = 3 * n
    gates = W.ansatz_layers(n, W.random_adjacent_pairs(n, layers, n) if n > 1 else [])
    if n >= 3:  # add controlled rotations and CZ on non-adjacent pairs
        extra = W.mixed_gate_circuit(n, 8, seed=800 + n)
        P0 = W.param_count(gates)
        extra = extra.copy()
        for row in extra:
            if row[0] in (7, 8, 9, 10, 14):
                row[3] += P0
        gates = np.vstack([gates, extra])
    P = W.param_count(gates)
    d = 1 << n
    U = W.haar_unitary(d, 900 + n).conj().T.copy()
    B = 5
    ps = W.random_params(P, 17 * n, batch=B)
    dec = N_Qubit_Decomposition_custom(U)
    dec.set_Gate_Structure(circuit_of(n, gates))
    dec.set_Cost_Function_Variant(variant)
    costs = dec.Optimization_Problem_Batch(ps)
    f, g = dec.Optimization_Problem_Combined_Batch(ps)
    for b in range(B):
        fr, gr = oracle.optimization_problem_combined(gates, U, n, ps[b], variant)
        assert abs(costs[b] - fr) < 1e-12
        assert abs(f[b] - fr) < 1e-12
        assert np.abs(g[b] - gr).max() < 1e-12
    # batch == single (test_optmization_problem_combined.py:214-219)
    assert abs(dec.Optimization_Problem(ps[0]) - costs[0]) < 1e-15
    # complex64 evaluation: error vs the FP64 oracle stays within 1e-5
    c32 = dec.Optimization_Problem_Batch(ps, precision=PREC_F32)
    f32, g32 = dec.Optimization_Problem_Combined_Batch(ps, precision=PREC_F32)
    assert np.abs(c32 - costs).max() < 1e-5
    assert np.abs(g32 - g).max() < 1e-5


def test_c2_config_8q_cost_and_gradient(oracle):
    """8 qubits, U3 layer + 40 random adjacent CNOT layers (264 angles): cost and the
    full gradient against the oracle's derivative-matrix route to 1e-12."""
    say("8q 264-parameter cost + gradient")
    gates = W.ansatz_8q_40()
    P = W.param_count(gates)
    U = W.haar_unitary(256, 2026).conj().T.copy()
    p = W.random_params(P, 1)
    dec = N_Qubit_Decomposition_custom(U)
    dec.set_Gate_Structure(circuit_of(8, gates))
    f, g = dec.Optimization_Problem_Combined(p)
    fr, gr = oracle.optimization_problem_combined(gates, U, 8, p)
    say(f"  |df|={abs(f - fr):.3e} max|dg|={np.abs(g - gr).max():.3e}")
    assert abs(f - fr) < 1e-12
    assert np.abs(g - gr).max() < 1e-12
    Uc, D = dec.Optimization_Problem_Combined_Unitary(p)
    assert rel_fro(Uc, oracle.apply_to(gates, p, U)) < 1e-12


def test_trace_offset_rectangular_8q_reachable(oracle):
    """The reference's 8-qubit trace-offset shape (test_optmization_problem_combined.py:
    123-184): rows 80..240 of C(theta*) as a 256 x 160 adjoint target, offset 80."""
    say("8q trace offset 80, 256x160 target")
    gates = W.ansatz_layers(8, W.random_adjacent_pairs(8, 20, 42))
    P = W.param_count(gates)
    theta = W.random_params(P, 42)
    V = oracle.apply_to(gates, theta, np.eye(256))
    Ucut = V[80:240, :].conj().T.copy()
    dec = N_Qubit_Decomposition_custom(Ucut)
    dec.set_Trace_Offset(80)
    dec.set_Gate_Structure(circuit_of(8, gates))
    f0, g0 = dec.Optimization_Problem_Combined(theta)
    assert abs(f0) < 1e-8
    assert np.abs(g0).max() < 1e-10
    p = W.random_params(P, 43)
    f, g = dec.Optimization_Problem_Combined(p)
    fr, gr = oracle.optimization_problem_combined(gates, Ucut, 8, p, orc.COST_FROBENIUS, 80)
    assert abs(f - fr) < 1e-12 and np.abs(g - gr).max() < 1e-12


def test_wrong_parameter_count_raises():
    gates = W.ansatz_layers(3, [0, 1])
    dec = N_Qubit_Decomposition_custom(np.eye(8, dtype=np.complex128))
    dec.set_Gate_Structure(circuit_of(3, gates))
    with pytest.raises(Exception):
        dec.Optimization_Problem(np.zeros(W.param_count(gates) + 1))


def _dense_block_full(n, F, qubits):
    """Full 2^n operator of a dense block on sorted qubits (local bit j <-> qubits[j])."""
    d = 1 << n
    k = len(qubits)
    full = np.zeros((d, d), dtype=complex)
    for r in range(d):
        li = sum(((r >> q) & 1) << j for j, q in enumerate(qubits))
        base = r & ~sum(1 << q for q in qubits)
        for lo in range(1 << k):
            ro = base | sum(((lo >> j) & 1) << q for j, q in enumerate(qubits))
            full[ro, r] += F[lo, li]
    return full


@pytest.mark.parametrize("n,k", [(3, 1), (4, 2), (5, 3), (6, 4), (8, 4), (7, 5)])
def test_large_kernel_matches_dense_reference(n, k):
    import torch
    from squander_hip import Context
    say(f"dense block n={n} k={k}")
    rng = np.random.default_rng(10 * n + k)
    qubits = np.sort(rng.choice(n, size=k, replace=False)).astype(np.int32)
    F = (rng.standard_normal((1 << k, 1 << k)) + 1j * rng.standard_normal((1 << k, 1 << k)))
    d = 1 << n
    B, m = 3, max(1, d // 2 + 3)
    X = rng.standard_normal((B, d, m)) + 1j * rng.standard_normal((B, d, m))
    ref = np.einsum("ij,bjk->bik", _dense_block_full(n, F, list(qubits)), X)
    ctx = Context(0)
    for f32 in (False, True):
        dt = np.complex64 if f32 else np.complex128
        Fd = torch.from_numpy(F.astype(dt)).cuda()
        Xd = torch.from_numpy(X.astype(dt)).cuda()
        torch.cuda.synchronize()
        ctx.check(ctx._L.sqh_apply_large_kernel_dev(ctx.h, Fd.data_ptr(), qubits.ctypes.data, k, Xd.data_ptr(), d, m,
                                                    m, d * m, B, 1 if f32 else 0))
        ctx.synchronize()
        err = rel_fro(Xd.cpu().numpy(), ref)
        assert err < (1e-5 if f32 else 1e-12), (f32, err)


@pytest.mark.parametrize("n", [4, 6, 8, 10, 12, 14])
def test_fused_streaming_matches_oracle(oracle, n):
    import torch
    from squander_hip import Context
    say(f"fused streaming n={n}")
    gates = W.mixed_gate_circuit(n, 5 * n + 10, seed=1100 + n)
    P = W.param_count(gates)
    d = 1 << n
    m = 48 if n >= 10 else d
    B = 3
    ps = W.random_params(P, n, batch=B)
    X = np.random.default_rng(n).standard_normal((B, d, m)) + 0j
    ctx = Context(0)
    ctx.set_circuit(n, gates)
    # complex64 blocks run on v_mfma_f32_16x16x4f32 with FP32 accumulation
    for f32, tol in ((False, 1e-12), (True, 2e-5)):
        for per_b in (True, False):
            pd = torch.from_numpy(ps if per_b else ps[:1].copy()).cuda()
            Xd = torch.from_numpy(X.astype(np.complex64 if f32 else np.complex128)).cuda()
            torch.cuda.synchronize()
            ctx.check(ctx._L.sqh_apply_to_fused_dev(ctx.h, pd.data_ptr(), P if per_b else 0, Xd.data_ptr(), d, m, m,
                                                    d * m, B, int(f32)))
            ctx.synchronize()
            got = Xd.cpu().numpy()
            for b in range(B):
                ref = oracle.apply_to(gates, ps[b] if per_b else ps[0], X[b])
                err = rel_fro(got[b], ref)
                assert err < tol, (f32, per_b, b, err)


def test_fused_streaming_list_order_plan_17_qubits(oracle):
    """Above 16 qubits the fused path uses the list-order block plan."""
    import torch
    from squander_hip import Context
    n, m, B = 17, 4, 2
    say(f"fused streaming n={n} (list-order plan)")
    gates = W.mixed_gate_circuit(n, 40, seed=3000)
    P = W.param_count(gates)
    d = 1 << n
    ps = W.random_params(P, 31, batch=B)
    X = np.random.default_rng(31).standard_normal((B, d, m)) + 1j * np.random.default_rng(32).standard_normal((B, d, m))
    ctx = Context(0)
    ctx.set_circuit(n, gates)
    pd = torch.from_numpy(ps).cuda()
    Xd = torch.from_numpy(X.copy()).cuda()
    torch.cuda.synchronize()
    ctx.check(ctx._L.sqh_apply_to_fused_dev(ctx.h, pd.data_ptr(), P, Xd.data_ptr(), d, m, m, d * m, B, 0))
    ctx.synchronize()
    got = Xd.cpu().numpy()
    for b in range(B):
        assert rel_fro(got[b], oracle.apply_to(gates, ps[b], X[b])) < 1e-12


def test_block_launch_shapes_give_identical_results(oracle):
    """Every launch shape of the 4-qubit matrix-core block kernel (tiles per wavefront x
    threads per workgroup) computes the same bits, including a column count that is not a
    multiple of the tile group; the tuner picks one of them."""
    import torch
    from squander_hip import Context
    n, m, B = 7, 75, 3
    d = 1 << n
    gates = W.mixed_gate_circuit(n, 40, seed=1500)
    P = W.param_count(gates)
    ps = W.random_params(P, 15, batch=B)
    X = np.random.default_rng(15).standard_normal((B, d, m)) + 1j * np.random.default_rng(16).standard_normal((B, d, m))
    ctx = Context(0)
    ctx.set_circuit(n, gates)
    pd = torch.from_numpy(ps).cuda()
    for f32 in (False, True):
        say(f"block launch shapes f32={f32}")
        dt = np.complex64 if f32 else np.complex128
        outs = {}
        for cfg in ((1, 256), (2, 256), (4, 256), (1, 64), (2, 64), (4, 64), (0, 0)):
            ctx.set_block_config(*cfg)
            Xd = torch.from_numpy(X.astype(dt)).cuda()
            torch.cuda.synchronize()
            ctx.check(ctx._L.sqh_apply_to_fused_dev(ctx.h, pd.data_ptr(), P, Xd.data_ptr(), d, m, m, d * m, B,
                                                    int(f32)))
            ctx.synchronize()
            outs[cfg] = Xd.cpu().numpy()
        for cfg, got in outs.items():
            assert np.array_equal(got, outs[(1, 256)]), cfg
        for b in range(B):
            assert rel_fro(outs[(1, 256)][b], oracle.apply_to(gates, ps[b], X[b])) < (2e-5 if f32 else 1e-12)
        tiles, threads, ms = ctx.block_config(d, m, B, f32)
        say(f"  tuned: {tiles} tiles x {threads} threads ({ms:.1f} ms of tuning so far)")
        assert tiles in (1, 2, 4) and threads in (64, 256)
    with pytest.raises(Exception):
        ctx.set_block_config(3, 256)


def test_streaming_float32_faster_than_float64():
    """Counterpart of the reference's tests/gates/test_float32_performance.py:23-116
    (U3 in float32 >= 2x faster than float64, CNOT >= 0.25x, 12 qubits x 256 columns):
    on the GPU the unfused per-gate pass is HBM-bound, so complex64 moves half the
    bytes.  Bounds are lenient (1.4x, 0.25x) for a shared device."""
    import time
    import torch
    from squander_hip import Context
    say("streaming f32 vs f64, 12 qubits x 256 columns")
    n, m, B = 12, 256, 16
    d = 1 << n
    res = {}
    for name, gate in (("U3", [[W.GATE_U3, 5, -1, 0]]), ("CNOT", [[W.GATE_CNOT, 5, 7, 0]])):
        gates = np.array(gate * 8, dtype=np.int32)
        gates[:, 3] = np.arange(8) * (3 if name == "U3" else 0)
        P = W.param_count(gates)
        ctx = Context(0)
        ctx.set_circuit(n, gates)
        p = torch.from_numpy(W.random_params(max(P, 1), 3)).cuda()
        times = {}
        for f32 in (False, True):
            dt = torch.complex64 if f32 else torch.complex128
            M = torch.zeros((B, d, m), dtype=dt, device="cuda")
            torch.cuda.synchronize()
            call = lambda: ctx.check(ctx._L.sqh_apply_to_stream_dev(ctx.h, p.data_ptr(), 0, M.data_ptr(), d, m, m,
                                                                    d * m, B, int(f32)))
            call()
            ctx.synchronize()
            t0 = time.perf_counter()
            for _ in range(5):
                call()
            ctx.synchronize()
            times[f32] = (time.perf_counter() - t0) / 5
        res[name] = times[False] / times[True]
        say(f"  {name}: f64 {times[False] * 1e3:.3f} ms, f32 {times[True] * 1e3:.3f} ms, speedup {res[name]:.2f}")
    assert res["U3"] >= 1.4
    assert res["CNOT"] >= 0.25
