"""The tile-resident fused-block cost kernel (sqh_set_fused_cost: the circuit as dense
4-qubit blocks on the matrix cores, every column resident on the chip) against the
FP64 oracle: complex128 costs to 1e-12, complex64 to 1e-5, at 4-10 qubits, square and
rectangular targets, trace offsets, every cost variant it serves, the 8- and
10-qubit benchmark circuits, and the fallback of the correction variants."""
import sys

import numpy as np
import pytest

import oracle as orc
from squander_hip import _native as nat
from squander_hip import workloads as W

pytestmark = pytest.mark.gpu

VARIANTS = [(nat.COST_FROBENIUS, orc.COST_FROBENIUS), (nat.COST_HILBERT_SCHMIDT, orc.COST_HS),
            (nat.COST_INFIDELITY, orc.COST_INFIDELITY), (nat.COST_ABS_TRACE, orc.COST_ABS_TRACE)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _costs(ctx, ps, prec):
    c = np.empty(ps.shape[0])
    ctx.check(ctx._L.sqh_optimization_problem_batch(ctx.h, np.ascontiguousarray(ps).ctypes.data, ps.shape[0], prec,
                                                     c.ctypes.data))
    return c


@pytest.mark.parametrize("n", [4, 5, 6, 7, 8, 9, 10])
def test_fused_cost_matches_oracle(oracle, n):
    gates = W.mixed_gate_circuit(n, 4 * n + 8, 3100 + n)
    P = W.param_count(gates)
    d = 1 << n
    B = 5
    ps = W.random_params(P, 3200 + n, batch=B)
    for m, off in [(d, 0), (d // 2 + 3, 2), (3, 1)]:
        say(f"fused cost n={n} m={m} offset={off}")
        U = W.haar_unitary(d, 3300 + n)[:, :m].copy()
        ctx = nat.Context(0)
        ctx.set_circuit(n, gates)
        ctx.set_target(U)
        ctx.set_trace_offset(off)
        ctx.set_fused_cost(True)
        for nv, ov in VARIANTS:
            ctx.set_cost_function(nv)
            # the transposed chain (default) and the forward pass over the target's columns
            got = {}
            for chain in (True, False):
                ctx.set_fused_cost_chain(chain)
                got[chain] = (_costs(ctx, ps, nat.PREC_F64), _costs(ctx, ps, nat.PREC_F32))
            ctx.set_fused_cost_chain(True)
            toff = off if ov in (orc.COST_FROBENIUS, orc.COST_ABS_TRACE) else 0
            for b in range(B):
                fo = oracle.optimization_problem(gates, U, n, ps[b], variant=ov, trace_offset=toff)
                for chain, (f64, f32) in got.items():
                    assert abs(f64[b] - fo) < 1e-12, (nv, chain, b, f64[b], fo)
                    assert abs(f32[b] - fo) < 1e-5, (nv, chain, b, f32[b], fo)


@pytest.mark.parametrize("which", ["8q_40", "10q_60"])
def test_fused_cost_benchmark_circuits(oracle, which):
    """The 8-qubit / 40-layer and 10-qubit / 60-layer circuits: fused-block costs
    against the oracle and against the tile kernels."""
    n, gates = (8, W.ansatz_8q_40()) if which == "8q_40" else (10, W.ansatz_10q_60())
    say(f"fused cost {which}")
    P = W.param_count(gates)
    U = W.haar_unitary(1 << n, 2026 + n).conj().T.copy()
    B = 4
    ps = W.random_params(P, 41, batch=B)
    ctx = nat.Context(0)
    ctx.set_circuit(n, gates)
    ctx.set_target(U)
    tile64, tile32 = _costs(ctx, ps, nat.PREC_F64), _costs(ctx, ps, nat.PREC_F32)
    ctx.set_fused_cost(True)
    f64, f32 = _costs(ctx, ps, nat.PREC_F64), _costs(ctx, ps, nat.PREC_F32)
    ctx.set_fused_cost_chain(False)
    fw64, fw32 = _costs(ctx, ps, nat.PREC_F64), _costs(ctx, ps, nat.PREC_F32)
    for b in range(B):
        fo = oracle.optimization_problem(gates, U, n, ps[b])
        assert abs(f64[b] - fo) < 1e-12 and abs(tile64[b] - fo) < 1e-12 and abs(fw64[b] - fo) < 1e-12
        assert abs(f32[b] - fo) < 1e-5 and abs(tile32[b] - fo) < 1e-5 and abs(fw32[b] - fo) < 1e-5
    say(f"  max |fused - oracle| c128 {np.max(np.abs(f64 - tile64)):.2e} (vs tile), c64 "
        f"{max(abs(f32[b] - oracle.optimization_problem(gates, U, n, ps[b])) for b in range(B)):.2e}")


def test_fused_cost_batch_equals_single_evaluations():
    n = 8
    gates = W.ansatz_8q_40()
    P = W.param_count(gates)
    U = W.haar_unitary(256, 5).conj().T.copy()
    ps = W.random_params(P, 6, batch=37)
    ctx = nat.Context(0)
    ctx.set_circuit(n, gates)
    ctx.set_target(U)
    ctx.set_fused_cost(True)
    full = _costs(ctx, ps, nat.PREC_F64)
    for b in (0, 17, 36):
        assert _costs(ctx, ps[b:b + 1], nat.PREC_F64)[0] == full[b]


def test_correction_variants_fall_back_to_the_tile_kernels():
    n = 5
    gates = W.mixed_gate_circuit(n, 20, 77)
    P = W.param_count(gates)
    U = W.haar_unitary(32, 78).conj().T.copy()
    ps = W.random_params(P, 79, batch=3)
    out = {}
    for fused in (False, True):
        ctx = nat.Context(0)
        ctx.set_circuit(n, gates)
        ctx.set_target(U)
        ctx.set_cost_function(nat.COST_FROBENIUS_CORRECTION2)
        ctx.set_fused_cost(fused)
        out[fused] = _costs(ctx, ps, nat.PREC_F64)
    assert np.array_equal(out[False], out[True])


SHAPES = [(t, g, w, p) for t in (2, 4) for g in (1, 2) for w in (1, 2, 4) for p in (0, 1)
          if g == 1 or (t == 4 and p == 1)]


@pytest.mark.parametrize("chain", [True, False])
@pytest.mark.parametrize("n", [4, 6, 8, 9, 10])
def test_fused_cost_launch_shapes_are_bit_identical(n, chain):
    """Every launch shape the tuner may pick (tiles per wavefront, exchange groups,
    waves per workgroup, prefetch) gives bit-identical costs, in both precisions; the
    tuner's own pick equals them too.  Shapes that do not fit the qubit count (a column
    must lie in one exchange group) are rejected, and so are two exchange groups for the
    transposed chain."""
    gates = W.mixed_gate_circuit(n, 5 * n, 3400 + n)
    P = W.param_count(gates)
    d = 1 << n
    U = W.haar_unitary(d, 3500 + n)[:, :d - 3].copy()
    ps = W.random_params(P, 3600 + n, batch=9)
    ctx = nat.Context(0)
    ctx.set_circuit(n, gates)
    ctx.set_target(U)
    ctx.set_fused_cost(True)
    ctx.set_fused_cost_chain(chain)
    for prec in (nat.PREC_F64, nat.PREC_F32):
        ctx.set_fused_cost_config(0, 0, 0, 0)
        tuned = _costs(ctx, ps, prec)
        pick = ctx.fused_cost_config(prec == nat.PREC_F32, ps.shape[0])
        say(f"fused shapes n={n} prec={prec}: tuner picked {pick}")
        ran = 0
        for shape in SHAPES:
            ctx.set_fused_cost_config(*shape)
            try:
                got = _costs(ctx, ps, prec)
            except nat.SqhError:
                # too few vector columns per group for this qubit count
                O = d // 16
                assert 16 * shape[0] < O or (16 * shape[0] // shape[1]) % O != 0 or (chain and shape[1] != 1)
                continue
            ran += 1
            assert np.array_equal(got, tuned), shape
        assert ran >= 4
    ctx.set_fused_cost_config(0, 0, 0, 0)


@pytest.mark.parametrize("n", [4, 6, 7, 8, 10])
def test_fused_apply_to_matches_oracle(oracle, n):
    """Gates_block::apply_to through the fused blocks on the matrix cores (the store pass
    of the tile-resident kernel): a batch of matrices in place, per-element and shared
    parameters, a rectangular strided view inside a wider buffer; complex128 to 1e-12
    relative Frobenius norm, complex64 to 1e-5."""
    import torch
    gates = W.mixed_gate_circuit(n, 4 * n + 6, 3700 + n)
    P = W.param_count(gates)
    d = 1 << n
    m, stride = d // 2 + 3, d + 5
    B = 3
    rng = np.random.default_rng(3800 + n)
    X = rng.standard_normal((B, d, stride)) + 1j * rng.standard_normal((B, d, stride))
    ps = W.random_params(P, 3900 + n, batch=B)
    for f32 in (False, True):
        for shared in (False, True):
            say(f"fused apply_to n={n} f32={f32} shared={shared}")
            ctx = nat.Context(0)
            ctx.set_circuit(n, gates)
            ctx.set_fused_cost(True)
            dt = np.complex64 if f32 else np.complex128
            Xd = torch.from_numpy(X.astype(dt)).cuda()
            pd = torch.from_numpy(ps).cuda()
            torch.cuda.synchronize()
            ctx.check(ctx._L.sqh_apply_to_dev(ctx.h, pd.data_ptr(), 0 if shared else P, Xd.data_ptr(), d, m, stride,
                                              d * stride, B, int(f32)))
            ctx.synchronize()
            R = Xd.cpu().numpy()
            for b in range(B):
                p = ps[0] if shared else ps[b]
                ref = oracle.apply_to(gates, p, X[b][:, :m].astype(dt).astype(np.complex128))
                err = np.linalg.norm(R[b][:, :m] - ref) / np.linalg.norm(ref)
                assert err < (1e-5 if f32 else 1e-12), (b, err)
                # the columns beyond the view are untouched
                assert np.array_equal(R[b][:, m:], X[b][:, m:].astype(dt))


def _combined(ctx, ps, prec):
    B, P = ps.shape
    f = np.empty(B)
    g = np.empty((B, P))
    ctx.check(ctx._L.sqh_optimization_problem_combined_batch(ctx.h, np.ascontiguousarray(ps).ctypes.data, B, prec,
                                                              f.ctypes.data, g.ctypes.data))
    return f, g


@pytest.mark.parametrize("n,sub", [(8, 6), (10, 7)])
def test_fused_chain_with_qubits_outside_every_block(oracle, n, sub):
    """A circuit on the first `sub` qubits of an n-qubit register: the blocks never cover
    the register, so the transposed chain's support stays short up to its last block
    (the generic row path of its final product) and the gradient falls back to the
    backward sweep.  Costs (chain on and off) and gradients against the oracle, square
    and rectangular targets with offsets."""
    gates = W.mixed_gate_circuit(sub, 6 * sub, 4300 + n)
    P = W.param_count(gates)
    d = 1 << n
    B = 3
    ps = W.random_params(P, 4400 + n, batch=B)
    for m, off in [(d, 0), (d // 2 + 5, 3)]:
        say(f"fused chain, {sub} of {n} qubits, m={m} offset={off}")
        U = W.haar_unitary(d, 4500 + n)[:, :m].copy()
        ctx = nat.Context(0)
        ctx.set_circuit(n, gates)
        ctx.set_target(U)
        ctx.set_trace_offset(off)
        ctx.set_fused_cost(True)
        ctx.set_cost_function(nat.COST_FROBENIUS)
        got = {}
        for chain in (True, False):
            ctx.set_fused_cost_chain(chain)
            got[chain] = (_costs(ctx, ps, nat.PREC_F64), _costs(ctx, ps, nat.PREC_F32))
        ctx.set_fused_cost_chain(True)
        f64, g64 = _combined(ctx, ps, nat.PREC_F64)
        for b in range(B):
            fo, go = oracle.optimization_problem_combined(gates, U, n, ps[b], orc.COST_FROBENIUS, off)
            for chain, (c64, c32) in got.items():
                assert abs(c64[b] - fo) < 1e-12, (chain, b, c64[b], fo)
                assert abs(c32[b] - fo) < 1e-5, (chain, b, c32[b], fo)
            assert abs(f64[b] - fo) < 1e-12 and np.abs(g64[b] - go).max() < 1e-12


@pytest.mark.parametrize("n", [4, 5, 6, 7, 8, 9, 10])
def test_fused_gradient_matches_oracle(oracle, n):
    """Costs and gradients through the fused blocks (forward on the matrix cores, the
    adjoint sweep block by block, each block's correlations resolved to its gates)
    against the FP64 oracle: complex128 to 1e-12, complex64 to 1e-5, every cost variant,
    square and rectangular targets, trace offsets."""
    gates = W.mixed_gate_circuit(n, 4 * n + 8, 3700 + n)
    P = W.param_count(gates)
    d = 1 << n
    B = 3
    ps = W.random_params(P, 3800 + n, batch=B)
    for m, off in [(d, 0), (d // 2 + 3, 2), (3, 1)]:
        say(f"fused gradient n={n} m={m} offset={off}")
        U = W.haar_unitary(d, 3900 + n)[:, :m].copy()
        ctx = nat.Context(0)
        ctx.set_circuit(n, gates)
        ctx.set_target(U)
        ctx.set_trace_offset(off)
        ctx.set_fused_cost(True)
        for nv, ov in VARIANTS:
            ctx.set_cost_function(nv)
            # adjoint vectors from the transposed chain (default) and from the backward sweep
            got = {}
            for chain in (True, False):
                ctx.set_fused_cost_chain(chain)
                got[chain] = _combined(ctx, ps, nat.PREC_F64) + _combined(ctx, ps, nat.PREC_F32)
            ctx.set_fused_cost_chain(True)
            toff = off if ov in (orc.COST_FROBENIUS, orc.COST_ABS_TRACE) else 0
            for b in range(B):
                fo, go = oracle.optimization_problem_combined(gates, U, n, ps[b], ov, toff)
                for chain, (f64, g64, f32, g32) in got.items():
                    assert abs(f64[b] - fo) < 1e-12, (nv, chain, b, f64[b], fo)
                    assert np.abs(g64[b] - go).max() < 1e-12, (nv, chain, b, np.abs(g64[b] - go).max())
                    assert abs(f32[b] - fo) < 1e-5, (nv, chain, b, f32[b], fo)
                    assert np.abs(g32[b] - go).max() < 1e-5, (nv, chain, b, np.abs(g32[b] - go).max())


@pytest.mark.parametrize("which", ["8q_40", "10q_60"])
def test_fused_gradient_benchmark_circuits(oracle, which):
    """The 8-qubit / 40-layer and 10-qubit / 60-layer circuits: fused-block costs and
    gradients against the oracle and against the tile adjoint kernels."""
    n, gates = (8, W.ansatz_8q_40()) if which == "8q_40" else (10, W.ansatz_10q_60())
    say(f"fused gradient {which}")
    P = W.param_count(gates)
    U = W.haar_unitary(1 << n, 2126 + n).conj().T.copy()
    B = 3
    ps = W.random_params(P, 43, batch=B)
    ctx = nat.Context(0)
    ctx.set_circuit(n, gates)
    ctx.set_target(U)
    ctx.set_fused_cost(False)
    tf, tg = _combined(ctx, ps, nat.PREC_F64)
    ctx.set_fused_cost(True)
    f64, g64 = _combined(ctx, ps, nat.PREC_F64)
    f32, g32 = _combined(ctx, ps, nat.PREC_F32)
    ctx.set_fused_cost_chain(False)
    fb64, gb64 = _combined(ctx, ps, nat.PREC_F64)
    for b in range(B):
        fo, go = oracle.optimization_problem_combined(gates, U, n, ps[b])
        assert abs(f64[b] - fo) < 1e-12 and np.abs(g64[b] - go).max() < 1e-12
        assert abs(fb64[b] - fo) < 1e-12 and np.abs(gb64[b] - go).max() < 1e-12
        assert abs(f32[b] - fo) < 1e-5 and np.abs(g32[b] - go).max() < 1e-5
    say(f"  max |fused - tile| gradient c128 {np.abs(g64 - tg).max():.2e}, c64 vs c128 {np.abs(g32 - g64).max():.2e}")


def test_fused_gradient_batch_equals_single_evaluations():
    n = 8
    gates = W.ansatz_8q_40()
    P = W.param_count(gates)
    U = W.haar_unitary(256, 15).conj().T.copy()
    ps = W.random_params(P, 16, batch=21)
    ctx = nat.Context(0)
    ctx.set_circuit(n, gates)
    ctx.set_target(U)
    ctx.set_fused_cost(True)
    f, g = _combined(ctx, ps, nat.PREC_F64)
    for b in (0, 9, 20):
        fb, gb = _combined(ctx, ps[b:b + 1], nat.PREC_F64)
        assert fb[0] == f[b] and np.array_equal(gb[0], g[b])


@pytest.mark.parametrize("n", [4, 7, 9, 10])
def test_fused_gradient_launch_shapes(oracle, n):
    """Every launch shape of the cost + gradient kernel (tiles per wavefront, waves per
    workgroup) agrees with the oracle; shapes that do not fit fall back to the rule."""
    gates = W.mixed_gate_circuit(n, 4 * n, 4000 + n)
    P = W.param_count(gates)
    d = 1 << n
    U = W.haar_unitary(d, 4100 + n)[:, :d - 8].copy()
    ps = W.random_params(P, 4200 + n, batch=4)
    ctx = nat.Context(0)
    ctx.set_circuit(n, gates)
    ctx.set_target(U)
    ctx.set_fused_cost(True)
    ref = [oracle.optimization_problem_combined(gates, U, n, ps[b]) for b in range(4)]
    for t in (2, 4):
        for w in (1, 2, 4, 8):
            say(f"fused gradient shape n={n} tiles={t} waves={w}")
            ctx.set_fused_grad_config(t, w)
            f, g = _combined(ctx, ps, nat.PREC_F64)
            for b in range(4):
                assert abs(f[b] - ref[b][0]) < 1e-12 and np.abs(g[b] - ref[b][1]).max() < 1e-12
    with pytest.raises(nat.SqhError):
        ctx.set_fused_grad_config(3, 2)


@pytest.mark.parametrize("n", [4, 5, 7, 8, 10])
def test_fused_apply_derivate_to_matches_oracle(oracle, n):
    """apply_derivate_to through the fused blocks (one store pass over the P elements,
    each element's block matrices built from its kernel table with that parameter's gate
    differentiated; a controlled gate's control-0 block zeroed): every derivative matrix
    within 1e-12 (relative Frobenius) of the oracle's, square and rectangular inputs, a
    strided input view, every gate kind."""
    import torch
    gates = W.mixed_gate_circuit(n, 4 * n + 6, 4800 + n)
    P = W.param_count(gates)
    d = 1 << n
    p = W.random_params(P, 4900 + n)
    rng = np.random.default_rng(5000 + n)
    ctx = nat.Context(0)
    ctx.set_circuit(n, gates)
    ctx.set_fused_cost(True)
    for m in sorted({d, max(1, d // 2 + 3)}):
        say(f"fused apply_derivate_to n={n} m={m} P={P}")
        X = rng.standard_normal((d, m)) + 1j * rng.standard_normal((d, m))
        ref = oracle.apply_derivate_to(gates, p, X)
        out = np.empty((P, d, m), dtype=np.complex128)
        ctx.check(ctx._L.sqh_apply_derivate_to(ctx.h, p.ctypes.data, X.ctypes.data, d, m, m, out.ctypes.data))
        worst = max(np.linalg.norm(out[k] - ref[k]) / max(np.linalg.norm(ref[k]), 1e-300) for k in range(P))
        assert worst < 1e-12, worst
    # strided device input: the m columns from column 3 of a d x (m + 5) buffer
    m = max(1, d // 2)
    buf = rng.standard_normal((d, m + 5)) + 1j * rng.standard_normal((d, m + 5))
    ref = oracle.apply_derivate_to(gates, p, buf[:, 3:3 + m].copy())
    dbuf = torch.from_numpy(buf).cuda()
    dps = torch.from_numpy(p).cuda()
    dout = torch.zeros((P, d, m), dtype=torch.complex128, device="cuda")
    torch.cuda.synchronize()
    ctx.check(ctx._L.sqh_apply_derivate_to_dev(ctx.h, dps.data_ptr(), dbuf.data_ptr() + 3 * 16, d, m, m + 5,
                                                dout.data_ptr()))
    torch.cuda.synchronize()
    got = dout.cpu().numpy()
    worst = max(np.linalg.norm(got[k] - ref[k]) / max(np.linalg.norm(ref[k]), 1e-300) for k in range(P))
    assert worst < 1e-12, worst
