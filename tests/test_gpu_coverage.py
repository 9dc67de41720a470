"""GPU parity for configurations and layouts beyond the core suite:

* the 10-qubit / 60-layer / 390-angle circuit of the batched benchmark (cost and full
  gradient, complex128 and complex64, generic and circuit-specialised kernels), checked
  against the oracle's costs: the gradient by the exact parameter-shift identity on a
  sample of components (FROBENIUS cost, linear in every gate's 2x2 matrix, each of
  whose entries is a sinusoid of period 2 pi in its raw parameter) and by a central
  difference of the cost along a random direction for the whole vector;
* row strides larger than the column count -- the reference's matrix_base layout
  (matrix_base.hpp:45-46, 127-137) -- through sqh_apply_to, the device entry points
  (register tile, unfused streaming, fused blocks), sqh_set_target and
  sqh_apply_derivate_to: a rows x m view inside a wider buffer, with the columns
  outside the view left untouched;
* the derivative matrices that optimization_problem_combined_unitary returns
  (Optimization_Interface.cpp:1525-1561) against the oracle's apply_derivate_to.
Bars: complex128 within 1e-12 (relative Frobenius for matrices, absolute for costs and
gradients), complex64 within 1e-5.
"""
import sys

import numpy as np
import pytest
import torch

import oracle as orc
from conftest import rel_fro
from squander_hip import Circuit, N_Qubit_Decomposition_custom
from squander_hip import _native as nat
from squander_hip import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ten_qubit():
    gates = W.ansatz_10q_60()
    P = W.param_count(gates)
    U = W.haar_unitary(1024, 1060).conj().T.copy()
    return gates, P, U


@pytest.mark.parametrize("jit", [False, True])
def test_10q_60_layer_cost_and_gradient(oracle, ten_qubit, jit):
    gates, P, U = ten_qubit
    assert P == 390 and len(gates) == 190
    B = 3
    ps = W.random_params(P, 1061, batch=B)
    say(f"10q 60-layer cost+grad, B={B}, jit={jit}")
    ctx = nat.Context(0)
    ctx.set_circuit(10, gates)
    ctx.set_target(U)
    ctx.set_jit(jit)
    f = np.zeros(B)
    g = np.zeros((B, P))
    ctx.check(ctx._L.sqh_optimization_problem_combined_batch(ctx.h, ps.ctypes.data, B, nat.PREC_F64,
                                                              f.ctypes.data, g.ctypes.data))
    fo = oracle.optimization_problem_batched(gates, U, 10, ps, nthreads=16)
    say(f"  max|df| = {np.abs(f - fo).max():.2e}")
    assert np.abs(f - fo).max() < 1e-12
    # complex64 costs and gradients against the FP64 values
    f32 = np.zeros(B)
    g32 = np.zeros((B, P))
    ctx.check(ctx._L.sqh_optimization_problem_combined_batch(ctx.h, ps.ctypes.data, B, nat.PREC_F32,
                                                              f32.ctypes.data, g32.ctypes.data))
    say(f"  complex64: max|df| = {np.abs(f32 - fo).max():.2e}, max|dg| = {np.abs(g32 - g).max():.2e}")
    assert np.abs(f32 - fo).max() < 1e-5
    assert np.abs(g32 - g).max() < 1e-5
    # parameter-shift components of element 0
    rng = np.random.default_rng(1062)
    ks = rng.choice(P, 16, replace=False)
    shifted = []
    for k in ks:
        for s in (+0.5 * np.pi, -0.5 * np.pi):
            q = ps[0].copy()
            q[k] += s
            shifted.append(q)
    fs = oracle.optimization_problem_batched(gates, U, 10, np.array(shifted), nthreads=16).reshape(-1, 2)
    ref = 0.5 * (fs[:, 0] - fs[:, 1])
    say(f"  parameter shift on {len(ks)} components: max|dg| = {np.abs(g[0, ks] - ref).max():.2e}")
    assert np.abs(g[0, ks] - ref).max() < 1e-12
    # the whole vector along a random direction
    v = rng.standard_normal(P)
    v /= np.linalg.norm(v)
    h = 1e-4
    fp, fm = oracle.optimization_problem_batched(gates, U, 10, np.array([ps[1] + h * v, ps[1] - h * v]), nthreads=16)
    dd = (fp - fm) / (2 * h)
    say(f"  directional derivative {g[1] @ v:.12f} vs central difference {dd:.12f}")
    assert abs(g[1] @ v - dd) < 1e-8


def _strided_host_case(n, m, col0, seed):
    d = 1 << n
    rng = np.random.default_rng(seed)
    buf = (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))).astype(np.complex128)
    return buf, buf[:, col0:col0 + m].copy()


@pytest.mark.parametrize("n,m,col0", [(8, 160, 0), (8, 100, 37), (5, 7, 20), (11, 300, 500)])
def test_apply_to_with_row_stride(oracle, n, m, col0):
    say(f"apply_to strided n={n} m={m} first column {col0}")
    d = 1 << n
    gates = W.mixed_gate_circuit(n, 4 * n + 8, seed=900 + n)
    p = W.random_params(W.param_count(gates), 901 + n)
    buf, view = _strided_host_case(n, m, col0, 902 + n)
    ref = oracle.apply_to(gates, p, view)
    before = buf.copy()
    ctx = nat.Context(0)
    ctx.set_circuit(n, gates)
    ptr = buf.ctypes.data + col0 * 16
    ctx.check(ctx._L.sqh_apply_to(ctx.h, p.ctypes.data, ptr, d, m, d, 0))
    assert rel_fro(buf[:, col0:col0 + m], ref) < 1e-12
    assert np.array_equal(buf[:, :col0], before[:, :col0])
    assert np.array_equal(buf[:, col0 + m:], before[:, col0 + m:])


@pytest.mark.parametrize("path", ["tile", "stream", "fused"])
def test_device_apply_with_row_stride(oracle, path):
    """Two 256 x 160 views (first columns 0 and 50) of 256 x 256 device buffers, batch
    stride one buffer, per-matrix parameters."""
    n, m, col0, B = 8, 160, 50, 2
    d = 1 << n
    say(f"device apply strided path={path}")
    gates = W.ansatz_8q_40()
    P = W.param_count(gates)
    ps = W.random_params(P, 910, batch=B)
    rng = np.random.default_rng(911)
    bufs = (rng.standard_normal((B, d, d)) + 1j * rng.standard_normal((B, d, d))).astype(np.complex128)
    refs = [oracle.apply_to(gates, ps[b], bufs[b][:, col0:col0 + m]) for b in range(B)]
    ctx = nat.Context(0)
    ctx.set_circuit(n, gates)
    dbuf = torch.from_numpy(bufs.copy()).cuda()
    dps = torch.from_numpy(ps).cuda()
    torch.cuda.synchronize()
    fn = {"tile": ctx._L.sqh_apply_to_dev, "stream": ctx._L.sqh_apply_to_stream_dev,
          "fused": ctx._L.sqh_apply_to_fused_dev}[path]
    ctx.check(fn(ctx.h, dps.data_ptr(), P, dbuf.data_ptr() + col0 * 16, d, m, d, d * d, B, 0))
    ctx.synchronize()
    out = dbuf.cpu().numpy()
    for b in range(B):
        assert rel_fro(out[b][:, col0:col0 + m], refs[b]) < 1e-12
        assert np.array_equal(out[b][:, :col0], bufs[b][:, :col0])
        assert np.array_equal(out[b][:, col0 + m:], bufs[b][:, col0 + m:])


def test_set_target_with_row_stride(oracle):
    """A 256 x 160 target stored with row stride 256 (first column 40) gives the costs
    and gradients of its contiguous copy, with and without a trace offset."""
    n, m, col0 = 8, 160, 40
    say("set_target strided 256x160")
    gates = W.ansatz_layers(n, W.random_adjacent_pairs(n, 6, 920))
    P = W.param_count(gates)
    buf = W.haar_unitary(256, 921)
    T = buf[:, col0:col0 + m].copy()
    ps = W.random_params(P, 922, batch=3)
    for off in (0, 96):
        ctx = nat.Context(0)
        ctx.set_circuit(n, gates)
        ctx.set_trace_offset(off)
        ctx.check(ctx._L.sqh_set_target(ctx.h, buf.ctypes.data + col0 * 16, 256, m, 256))
        f = np.zeros(3)
        g = np.zeros((3, P))
        ctx.check(ctx._L.sqh_optimization_problem_combined_batch(ctx.h, ps.ctypes.data, 3, nat.PREC_F64,
                                                                  f.ctypes.data, g.ctypes.data))
        for b in range(3):
            fo, go = oracle.optimization_problem_combined(gates, T, n, ps[b], orc.COST_FROBENIUS, off)
            assert abs(f[b] - fo) < 1e-12 and np.abs(g[b] - go).max() < 1e-12
    with pytest.raises(nat.SqhError):  # stride below the column count
        ctx.check(ctx._L.sqh_set_target(ctx.h, buf.ctypes.data, 256, m, m - 1))


def test_apply_derivate_to_with_row_stride(oracle):
    """Input 64 x 40 inside a 64 x 64 buffer (first column 11); the P outputs are dense
    64 x 40 matrices."""
    n, m, col0 = 6, 40, 11
    d = 1 << n
    say("apply_derivate_to strided 64x40")
    gates = W.mixed_gate_circuit(n, 30, seed=930)
    P = W.param_count(gates)
    p = W.random_params(P, 931)
    buf, view = _strided_host_case(n, m, col0, 932)
    ref = oracle.apply_derivate_to(gates, p, view)
    ctx = nat.Context(0)
    ctx.set_circuit(n, gates)
    out = np.zeros((P, d, m), dtype=np.complex128)
    ctx.check(ctx._L.sqh_apply_derivate_to(ctx.h, p.ctypes.data, buf.ctypes.data + col0 * 16, d, m, d,
                                           out.ctypes.data))
    worst = max(rel_fro(out[k], ref[k]) for k in range(P))
    say(f"  worst relative error over {P} matrices: {worst:.2e}")
    assert worst < 1e-12
    # the device entry point reads the strided input in place
    dbuf = torch.from_numpy(buf.copy()).cuda()
    dps = torch.from_numpy(p).cuda()
    dout = torch.zeros((P, d, m), dtype=torch.complex128, device="cuda")
    torch.cuda.synchronize()
    ctx.check(ctx._L.sqh_apply_derivate_to_dev(ctx.h, dps.data_ptr(), dbuf.data_ptr() + col0 * 16, d, m, d,
                                               dout.data_ptr()))
    ctx.synchronize()
    outd = dout.cpu().numpy()
    assert max(rel_fro(outd[k], ref[k]) for k in range(P)) < 1e-12


@pytest.mark.parametrize("n", [3, 5, 6])
def test_combined_unitary_derivatives_match_oracle(oracle, n):
    """Optimization_Problem_Combined_Unitary: C.Umtx and every dC/dp_k.Umtx."""
    say(f"combined unitary n={n}")
    gates = W.ansatz_layers(n, W.random_adjacent_pairs(n, 2 * n, 940 + n))
    P = W.param_count(gates)
    U = W.haar_unitary(1 << n, 941 + n).conj().T.copy()
    p = W.random_params(P, 942 + n)
    dec = N_Qubit_Decomposition_custom(U)
    dec.set_Gate_Structure(Circuit.from_gate_array(n, gates))
    Uc, D = dec.Optimization_Problem_Combined_Unitary(p)
    assert rel_fro(Uc, oracle.apply_to(gates, p, U)) < 1e-12
    ref = oracle.apply_derivate_to(gates, p, U)
    assert len(D) == P
    worst = max(rel_fro(D[k], ref[k]) for k in range(P))
    say(f"  worst relative error over {P} derivative matrices: {worst:.2e}")
    assert worst < 1e-12
