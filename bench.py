"""Benchmark: batched cost evaluations of the 8-qubit / 40-layer ansatz on one MI355X.

A step is one call of optimization_problem_batched on 512 parameter vectors
(8 qubits, U3 layer + 40 seeded-random adjacent CNOT layers: 128 gates, 264
angles; 256 x 256 complex128 Haar target), inputs resident in HBM.  The headline
path runs the circuit as its 6 fused dense 4-qubit blocks on the matrix cores with
every column resident on the chip (k_fused_cost); the JSON line also carries: that
kernel's roofline (live HIP-event timing inside the library), the gate-by-gate
register-tile kernels (circuit-specialised and generic) on the same batch, the
unfused streaming per-gate path's HBM bandwidth (512 MiB of matrices), the complex64
legs with their error against the FP64 oracle, the 10-qubit / 60-layer batch, the
8-qubit cost+gradient, and the CPU oracle timed on the host cores on a bounded
sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sequential-quantum-gate-decomposer_amd"))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

# HBM bytes per launch of the headline kernel from two rocprofv3 --pmc runs of this
# script (FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00, factors from a 256 MiB device copy
# in the same runs): profiles/k_tile_forward_k_stream_gate_k_block_mfma4_8q_b512.md
TILE_FORWARD_8Q_B512_TRAFFIC = (11922.6 * 2.0 + 2048.0) * 1024
# the same for the circuit-specialised kernel: profiles/sqh_jit_kernel_8q_b512.md
JIT_FORWARD_8Q_B512_TRAFFIC = (24042.5 * 2.0 + 2048.0) * 1024
# the same for the fused-block cost chain: profiles/k_fused_trace_8q_b512.md
FUSED_COST_8Q_B512_TRAFFIC = (10323.0 * 2.0 + 2048.0) * 1024
FP32_MFMA_PEAK_TFLOPS = 157.3


def _fusion_blocks(n, gates):
    from squander_hip import _native as nat
    return nat.inspect_fusion(n, gates)


def block_flops(gates, n, m):
    """Algorithmic flops of one pass of the fused dense 4-qubit blocks over a d x m
    matrix (the apply_to store pass): a complex 16 x 16 block times every 16-row group
    is 128 real flop per element, per block of the fusion plan."""
    from squander_hip import _native as nat
    return len(nat.inspect_fusion(n, gates)) * 128.0 * (1 << n) * m


def fused_cost_flops(gates, n, m, chain=True):
    """Algorithmic flops of one cost evaluation through the fused blocks.

    chain (the library's default cost pass): per column, the needed row of the circuit
    matrix built from a unit vector, the blocks transposed from the last; block b then
    acts on the 2^|S_b| rows inside the support S_b (the union of the qubits of blocks
    b .. nb-1): 128 flop per row for every block but the last (whose row is loaded),
    then 8 flop per row for the product with the target's column.

    forward: every block but the last over the whole matrix (128 flop per element per
    block), then of the last block only row c + offset of each column c (a 16-term
    complex dot product, 128 flop per column)."""
    from squander_hip import _native as nat
    plan = nat.inspect_fusion(n, gates)
    nb = len(plan)
    if not chain or nb < 2:
        return (nb - 1) * 128.0 * (1 << n) * m + 128.0 * m
    support, rows = set(plan[-1][0]), []
    for qs, _ in reversed(plan[:-1]):
        support |= set(qs)
        rows.append(1 << len(support))
    return (128.0 * sum(rows) + 8.0 * rows[-1]) * m


def fused_forward_leg(gates, U, params, costs, steps, n, d):
    """The fused cost pass with the target's columns run forward through the blocks
    (the transposed chain switched off): the comparison for the headline kernel."""
    import torch
    from squander_hip import _native as nat
    B = params.shape[0]
    fctx = nat.Context(0)
    fctx.set_circuit(n, gates)
    fctx.set_target(U)
    fctx.set_fused_cost(True)
    fctx.set_fused_cost_chain(False)
    step = lambda: fctx.check(fctx._L.sqh_optimization_problem_batch_dev(fctx.h, params.data_ptr(), B,
                                                                        nat.PREC_F64, costs.data_ptr()))
    step()
    fctx.synchronize()
    fctx.set_kernel_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    fctx.synchronize()
    dt = time.perf_counter() - t0
    kms, nl = fctx.kernel_timing()
    fctx.set_kernel_timing(False)
    k_s = kms / max(nl, 1) / 1e3
    fl = fused_cost_flops(gates, n, d, chain=False) * B
    return {"evals_per_s": B * steps / dt, "kernel_avg_ms": k_s * 1e3, "achieved_TFLOPs": fl / k_s / 1e12,
            "frac": fl / k_s / 1e12 / FP64_PEAK_TFLOPS, "algorithmic_flops_per_launch": fl,
            "launch_shape": dict(zip(("tiles_per_wave", "exchange_groups", "waves", "prefetch"),
                                     fctx.fused_cost_config(False, B))),
            "kernel": "k_fused_cost<double,8> (5 blocks over the target's columns, the 6th as one trace-row "
                      "dot product per column)"}


def fused_grad_flops(gates, n, m, chain=True):
    """Algorithmic flops of one cost + gradient evaluation through the fused blocks.

    chain (the library's default where the blocks' qubits cover the register): the
    adjoint vectors from the transposed chain (the cost pass's flops, fused_cost_flops),
    then a sweep up from the target: per block the correlation Z X^T, X <- F X (all but
    the last) and Z <- conj(F) Z (all but the last two) -- 3 nb - 3 products of 128 flop
    per element.

    backward sweep: nb block products forward, then per block X <- F^dagger X, the
    correlation Y X^T and (except for the first block) Y <- F^T Y -- 4 nb - 1 products."""
    from squander_hip import _native as nat
    plan = nat.inspect_fusion(n, gates)
    nb = len(plan)
    covered = set().union(*(set(qs) for qs, _ in plan)) == set(range(n))
    if not chain or nb < 2 or not covered:
        return (4 * nb - 1) * 128.0 * (1 << n) * m
    return fused_cost_flops(gates, n, m, chain=True) + (3 * nb - 3) * 128.0 * (1 << n) * m


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gate_work(gates, d, m, esz):
    """Algorithmic (flops, streamed bytes) of one pass of the gate list over a d x m
    matrix: 1-qubit parametric gate 14*d*m flop / 2*e*d*m B; CRY half of that;
    CNOT 0 flop / e*d*m B; CZ 0 flop / e*d*m/2 B."""
    from squander_hip import _native as nat
    fl, by = 0.0, 0.0
    for t, *_ in gates:
        t = int(t)
        if t in (nat.GATE_U3, nat.GATE_RX, nat.GATE_RY, nat.GATE_RZ):
            fl += 14.0 * d * m
            by += 2.0 * esz * d * m
        elif t == nat.GATE_CRY:
            fl += 7.0 * d * m
            by += esz * d * m
        elif t == nat.GATE_CNOT:
            by += esz * d * m
        elif t == nat.GATE_CZ:
            by += esz * d * m / 2
    return fl, by


def host_info():
    """Core count, CPU model and the SIMD extensions the oracle can use."""
    info = {"nproc": os.cpu_count()}
    try:
        with open("/proc/cpuinfo") as f:
            txt = f.read()
        model = [l.split(":", 1)[1].strip() for l in txt.splitlines() if l.startswith("model name")]
        flags = [l.split(":", 1)[1].split() for l in txt.splitlines() if l.startswith("flags")]
        info["model"] = model[0] if model else None
        info["simd"] = sorted(f for f in (flags[0] if flags else []) if f in ("avx2", "fma", "avx512f", "avx512dq"))
    except OSError:
        pass
    return info


def cpu_baseline(gates, U, n, P, o, seconds=10.0):
    """The CPU oracle (the restated C path, -O3 -mavx2 -mfma, OpenMP over the batch) on
    the host cores: the 8-qubit batched cost on all cores (the reported value) and on
    one core, the 8-qubit single-vector cost + gradient latency (P derivative
    matrices, as the reference's combined evaluation forms them) and the 10-qubit
    batched cost."""
    import numpy as np

    from squander_hip import workloads as W
    threads = min(os.cpu_count() or 1, 64)
    ps = W.random_params(P, 99, batch=64)
    done, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        o.optimization_problem_batched(gates, U, n, ps, nthreads=threads)
        done += ps.shape[0]
    dt = time.perf_counter() - t0
    res = {"value": done / dt, "unit": "cost evals/s", "cores": threads, "kind": "port",
           "sample": f"{done} cost evals of the 8q/40-layer ansatz (batches of 64, OpenMP over the batch, "
                     f"{dt:.1f} s)", "host": host_info()}
    done1, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < 3.0:
        o.optimization_problem_batched(gates, U, n, ps[:8], nthreads=1)
        done1 += 8
    res["one_core"] = {"value": done1 / (time.perf_counter() - t0), "unit": "cost evals/s", "cores": 1}
    t0 = time.perf_counter()
    o.optimization_problem_combined(gates, U, n, ps[0])
    res["c2_cost_grad_latency_s"] = {"value": time.perf_counter() - t0, "cores": 1,
                                     "sample": "one 264-parameter cost + gradient from the derivative matrices"}
    g10 = W.ansatz_10q_60(2027)
    U10 = W.haar_unitary(1024, 2027).conj().T.copy()
    p10 = W.random_params(W.param_count(g10), 8, batch=threads)
    t0 = time.perf_counter()
    o.optimization_problem_batched(g10, U10, 10, p10, nthreads=threads)
    res["c3_10q"] = {"value": threads / (time.perf_counter() - t0), "unit": "cost evals/s", "cores": threads,
                     "sample": f"one batch of {threads} 10-qubit evaluations"}
    return res


def tile_leg(gates, U, params, costs, steps, flops, jit):
    """The same batch with the gate-by-gate register-tile kernel: specialised to the
    circuit (jit) or the generic one (one program walked op by op)."""
    from squander_hip import PREC_F64, Context
    ctx = Context(0)
    ctx.set_circuit(8, gates)
    ctx.set_target(U)
    ctx.set_jit(jit)
    ctx.set_fused_cost(False)
    B = params.shape[0]
    step = lambda: ctx.check(ctx._L.sqh_optimization_problem_batch_dev(ctx.h, params.data_ptr(), B, PREC_F64,
                                                                      costs.data_ptr()))
    step()
    ctx.synchronize()
    ctx.set_kernel_timing(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    ctx.synchronize()
    dt = time.perf_counter() - t0
    kms, nl = ctx.kernel_timing()
    k_s = kms / max(nl, 1) / 1e3
    return {"evals_per_s": B * steps / dt, "kernel_avg_ms": k_s * 1e3, "achieved_TFLOPs": flops / k_s / 1e12,
            "frac_of_fp64_valu_peak": flops / k_s / 1e12 / FP64_PEAK_TFLOPS,
            "flops": "gate arithmetic (14 flop per element per 1-qubit gate)",
            "kernel": ("sqh_jit_kernel (k_tile_forward specialised to the circuit, diagonal mode)" if jit
                       else "k_tile_forward<8,double,1>")}


def ten_qubit_leg(steps, o):
    """10 qubits, U3 layer + 60 random adjacent CNOT layers (390 angles), 1024 x 1024
    complex128 Haar target, batch of 64: complex128 and complex64 cost evaluations."""
    import numpy as np
    import torch

    from squander_hip import PREC_F32, PREC_F64, Context
    from squander_hip import workloads as W
    n, d, B = 10, 1024, 64
    gates = W.ansatz_10q_60(2027)
    P = W.param_count(gates)
    U = W.haar_unitary(d, 2027).conj().T.copy()
    ctx = Context(0)
    ctx.set_circuit(n, gates)
    ctx.set_target(U)
    ctx.set_jit(True)
    ctx.set_fused_cost(False)  # the register-tile kernels here; the fused legs below
    params = torch.from_numpy(W.random_params(P, 8, batch=B)).cuda()
    costs = torch.empty(B, dtype=torch.float64, device="cuda")
    flops, _ = gate_work(gates, d, d, 16)
    res = {"workload": "10 qubits, 60 CNOT layers, 390 angles, batch 64", "algorithmic_flops_per_eval": flops}
    ref = o.optimization_problem_batched(gates, U, n, params[:2].cpu().numpy(), nthreads=2)
    for name, prec in (("c128", PREC_F64), ("c64", PREC_F32)):
        step = lambda: ctx.check(ctx._L.sqh_optimization_problem_batch_dev(ctx.h, params.data_ptr(), B, prec,
                                                                          costs.data_ptr()))
        step()
        ctx.synchronize()
        ctx.set_kernel_timing(True)
        t0 = time.perf_counter()
        k = max(2, steps // 4)
        for _ in range(k):
            step()
        ctx.synchronize()
        dt = time.perf_counter() - t0
        kms, nl = ctx.kernel_timing()
        ctx.set_kernel_timing(False)
        # FP32 without packed math (not generated here) peaks at the FP64 rate
        peak = FP64_PEAK_TFLOPS
        tf = flops * B / (kms / max(nl, 1) / 1e3) / 1e12
        res[name] = {"evals_per_s": B * k / dt, "kernel_avg_ms": kms / max(nl, 1), "achieved_TFLOPs": tf,
                     "peak_TFLOPs": peak, "frac": tf / peak,
                     "err_vs_fp64_oracle": float(np.abs(costs[:2].cpu().numpy() - ref).max())}
        log(f"10q {name}: {res[name]['evals_per_s']:.0f} evals/s, {tf:.1f} TFLOP/s")
    # the same through the fused dense 4-qubit blocks on the matrix cores
    fctx = Context(0)
    fctx.set_circuit(n, gates)
    fctx.set_target(U)
    fctx.set_fused_cost(True)
    bfl = fused_cost_flops(gates, n, d)
    for name, prec, peak in (("fused_c128", PREC_F64, FP64_PEAK_TFLOPS), ("fused_c64", PREC_F32, FP32_MFMA_PEAK_TFLOPS)):
        step = lambda: fctx.check(fctx._L.sqh_optimization_problem_batch_dev(fctx.h, params.data_ptr(), B, prec,
                                                                            costs.data_ptr()))
        step()
        fctx.synchronize()
        fctx.set_kernel_timing(True)
        t0 = time.perf_counter()
        k = max(2, steps // 4)
        for _ in range(k):
            step()
        fctx.synchronize()
        dt = time.perf_counter() - t0
        kms, nl = fctx.kernel_timing()
        fctx.set_kernel_timing(False)
        tf = bfl * B / (kms / max(nl, 1) / 1e3) / 1e12
        res[name] = {"evals_per_s": B * k / dt, "kernel_avg_ms": kms / max(nl, 1), "achieved_TFLOPs": tf,
                     "peak_TFLOPs": peak, "frac": tf / peak, "blocks": len(_fusion_blocks(n, gates)),
                     "algorithmic_flops_per_eval": bfl,
                     "err_vs_fp64_oracle": float(np.abs(costs[:2].cpu().numpy() - ref).max())}
        log(f"10q {name}: {res[name]['evals_per_s']:.0f} evals/s, {tf:.1f} TFLOP/s")
    # cost + gradient over the batch through the fused blocks (the transposed chain for
    # the adjoint vectors, then 3 nb - 3 block products up from the target), complex128
    # and complex64
    grad = torch.empty((B, P), dtype=torch.float64, device="cuda")
    _, gr = o.optimization_problem_combined(gates, U, n, params[0].cpu().numpy())
    for name, prec, peak in (("c128_gradient_fused", PREC_F64, FP64_PEAK_TFLOPS),
                             ("c64_gradient_fused", PREC_F32, FP32_MFMA_PEAK_TFLOPS)):
        fstep = lambda: fctx.check(fctx._L.sqh_optimization_problem_combined_batch_dev(
            fctx.h, params.data_ptr(), B, prec, costs.data_ptr(), grad.data_ptr()))
        fstep()
        fctx.synchronize()
        fctx.set_kernel_timing(True)
        t0 = time.perf_counter()
        for _ in range(3):
            fstep()
        fctx.synchronize()
        dt = (time.perf_counter() - t0) / 3
        kms, nl = fctx.kernel_timing()
        fctx.set_kernel_timing(False)
        k_s = kms / max(nl, 1) / 1e3
        gfl = fused_grad_flops(gates, n, d) * B
        res[name] = {"kernel_avg_ms": k_s * 1e3, "ms_per_call": dt * 1e3, "grads_per_s": B / dt,
                     "achieved_TFLOPs": gfl / k_s / 1e12, "peak_TFLOPs": peak, "frac": gfl / k_s / 1e12 / peak,
                     "err_grad_vs_oracle": float(np.abs(grad[0].cpu().numpy() - gr).max())}
        log(f"10q {name}: {dt * 1e3:.2f} ms per call ({k_s * 1e3:.2f} ms kernel)")
    # complex128 cost + gradient over the batch (specialised register-tile adjoint kernel)
    gstep = lambda: ctx.check(ctx._L.sqh_optimization_problem_combined_batch_dev(ctx.h, params.data_ptr(), B, PREC_F64,
                                                                                costs.data_ptr(), grad.data_ptr()))
    gstep()
    ctx.synchronize()
    ctx.set_kernel_timing(True)
    for _ in range(3):
        gstep()
    ctx.synchronize()
    kms, nl = ctx.kernel_timing()
    ctx.set_kernel_timing(False)
    k_s = kms / max(nl, 1) / 1e3
    gflops = flops * B * (3.0 + 8.0 / 7.0)
    _, gr = o.optimization_problem_combined(gates, U, n, params[0].cpu().numpy())
    res["c128_gradient"] = {"kernel_avg_ms": k_s * 1e3, "grads_per_s": B / k_s, "achieved_TFLOPs": gflops / k_s / 1e12,
                            "frac_of_fp64_peak": gflops / k_s / 1e12 / FP64_PEAK_TFLOPS,
                            "err_grad_vs_oracle": float(np.abs(grad[0].cpu().numpy() - gr).max())}
    log(f"10q c128 cost+gradient: {k_s * 1e3:.2f} ms")
    return res


def gradient_leg(ctx, params, gates, U, o, reps=5, fused=False):
    """Batched cost + gradient (optimization_problem_combined over the batch) of the
    8-qubit ansatz: through the fused 4-qubit blocks on the matrix cores (fused), or the
    circuit-specialised register-tile adjoint kernel."""
    import numpy as np
    import torch

    from squander_hip import PREC_F64
    n, d = 8, 256
    B, P = params.shape
    L = ctx._L
    G = torch.empty((B, P), dtype=torch.float64, device="cuda")
    fb = torch.empty(B, dtype=torch.float64, device="cuda")

    def grad_batch():
        ctx.check(L.sqh_optimization_problem_combined_batch_dev(ctx.h, params.data_ptr(), B, PREC_F64, fb.data_ptr(),
                                                                 G.data_ptr()))

    grad_batch()
    ctx.synchronize()
    ctx.set_kernel_timing(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        grad_batch()
    ctx.synchronize()
    t1 = time.perf_counter()
    gkms, gnl = ctx.kernel_timing()
    ctx.set_kernel_timing(False)
    gk_s = gkms / max(gnl, 1) / 1e3
    _, gr1 = o.optimization_problem_combined(gates, U, n, params[1].cpu().numpy())
    if fused:
        gflops = fused_grad_flops(gates, n, d) * B
        peak = FP64_PEAK_TFLOPS
        kernel = ("k_fused_grad_chain<double,8,T> (v_mfma_f64_16x16x4f64: the transposed chain for the "
                  "adjoint vectors, then the sweep up from the target)")
    else:
        # adjoint sweep: forward pass, G^dagger on the state, G^T on the adjoint and the
        # pair correlations (32 flop per pair against 28 for a gate): (3 + 8/7) x forward
        gflops = gate_work(gates, d, d, 16)[0] * B * (3.0 + 8.0 / 7.0)
        peak = FP64_PEAK_TFLOPS
        kernel = "sqh_jit_kernel (adjoint gradient specialised to the circuit)"
    log(f"8q B={B} cost+gradient ({'fused' if fused else 'tile'}): {gk_s * 1e3:.3f} ms kernel, "
        f"{(t1 - t0) / reps * 1e3:.3f} ms per call, {gflops / gk_s / 1e12:.1f} TFLOP/s")
    return {"ms": (t1 - t0) / reps * 1e3, "kernel_avg_ms": gk_s * 1e3, "grads_per_s": B / ((t1 - t0) / reps),
            "algorithmic_flops": gflops, "achieved_TFLOPs": gflops / gk_s / 1e12, "peak_TFLOPs": peak,
            "frac_of_peak": gflops / gk_s / 1e12 / peak,
            "err_grad_vs_oracle": float(np.abs(G[1].cpu().numpy() - gr1).max()), "kernel": kernel}


def derivate_leg(gates, P, U):
    """apply_derivate_to of the 8-qubit ansatz on a 256 x 256 matrix: 264 derivative
    matrices of 1 MiB written to HBM per call.  A forward store pass runs over 264
    elements sharing the input, element k with gate k's kernel replaced by its
    derivative: one circuit pass per parameter (the reference's replay count), through
    the fused blocks (the default path; 128 flop per element per block) or the generic
    and the circuit-specialised register-tile kernels (the gate arithmetic)."""
    import torch

    from squander_hip import Context
    from squander_hip import workloads as W
    d = 256
    p = torch.from_numpy(W.random_params(P, 3)).cuda()
    M = torch.from_numpy(U.copy()).cuda()
    out = torch.empty((P, d, d), dtype=torch.complex128, device="cuda")
    wbytes = P * d * d * 16
    flops_eval, _ = gate_work(gates, d, d, 16)
    # one pass of the circuit per parameter
    flops = P * flops_eval
    res = {"bound": "mfma (FP64)", "algorithmic_flops": flops, "bytes_written": wbytes}
    bflops = P * block_flops(gates, 8, d)
    for name, jit in (("fused", False), ("generic", False), ("specialised", True)):
        ctx = Context(0)
        ctx.set_circuit(8, gates)
        ctx.set_jit(jit)
        ctx.set_fused_cost(name == "fused")
        call = lambda: ctx.check(ctx._L.sqh_apply_derivate_to_dev(ctx.h, p.data_ptr(), M.data_ptr(), d, d, d,
                                                                  out.data_ptr()))
        call()
        ctx.synchronize()
        ctx.set_kernel_timing(True)
        t0 = time.perf_counter()
        for _ in range(5):
            call()
        ctx.synchronize()
        dt = (time.perf_counter() - t0) / 5
        kms, nl = ctx.kernel_timing()
        ctx.set_kernel_timing(False)
        k_s = kms / max(nl, 1) / 1e3
        fl = bflops if name == "fused" else flops
        res[name] = {"ms": dt * 1e3, "kernel_avg_ms": k_s * 1e3, "achieved_TFLOPs": fl / k_s / 1e12,
                     "frac_of_fp64_peak": fl / k_s / 1e12 / FP64_PEAK_TFLOPS, "write_GBs": wbytes / k_s / 1e9,
                     "kernel": {"fused": "k_fused_cost<double,8,...> store pass over 264 elements (per-parameter "
                                         "block matrices)",
                                "specialised": "sqh_jit_kernel (forward store, per-parameter kernel tables)",
                                "generic": "k_tile_forward<8,double,0> (per-parameter kernel tables)"}[name]}
        log(f"apply_derivate_to 8q {name}: {k_s * 1e3:.3f} ms")
    # error of the 264 derivative matrices (the last call's) against the FP64 oracle
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import oracle as orc
    ref = orc.Oracle().apply_derivate_to(gates, p.cpu().numpy(), U)
    got = out.cpu().numpy()
    res["err_rel_fro_vs_fp64_oracle"] = max(float(np.linalg.norm(got[k] - ref[k]) / np.linalg.norm(ref[k]))
                                            for k in range(P))
    return res


def mlp_train_leg(samples=65536, epochs=4, batch=1024):
    """Training throughput of the warm-start network at its configured size (512 ->
    3 x 1024 GELU -> 168, BF16 weights and activations, FP32 accumulation and master
    weights, Adam) on synthetic inputs and targets: every layer's forward and both
    backward products run on k_gemm_bf16 (v_mfma_f32_16x16x32_bf16).  Algorithmic
    flops per sample and step: 2 * sum(in * out) forward, twice that backward."""
    import numpy as np
    import torch

    from squander_hip import Context
    from squander_hip.warmstart import WarmStart
    ctx = Context(0)
    in_dim, P, hidden = 512, 84, 1024
    ws = WarmStart(ctx, in_dim, P, hidden=hidden, n_hidden=3, seed=1)
    rng = np.random.default_rng(2)
    X = rng.standard_normal((samples, in_dim)).astype(np.float32)
    Y = rng.uniform(-1, 1, (samples, 2 * P)).astype(np.float32)
    ws.train(X[:batch * 4], Y[:batch * 4], epochs=1, batch=batch, lr=1e-3, seed=0)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses = ws.train(X, Y, epochs=epochs, batch=batch, lr=1e-3, seed=0)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    dims = [in_dim, hidden, hidden, hidden, 2 * P]
    fwd = 2.0 * sum(a * b for a, b in zip(dims[:-1], dims[1:]))
    steps = epochs * (samples // batch)
    flops = 3.0 * fwd * batch * steps
    return {"workload": f"warm-start MLP 512 -> 3 x 1024 GELU -> 168, batch {batch}, {steps} Adam steps, synthetic data",
            "seconds": dt, "samples_per_s": batch * steps / dt, "wall_TFLOPs": flops / dt / 1e12,
            "peak_TFLOPs_bf16_dense": 2500.0, "algorithmic_flops_per_step": 3.0 * fwd * batch,
            "loss_first_last": [losses[0], losses[-1]]}


def agents_leg(o):
    """AGENTS (gradient-free population search, AGENTS.cpp:323-600) on the 8-qubit
    ansatz: 64 agents, each iteration's population costs in one batched launch.  The
    CPU oracle follows the identical trajectory (same generator consumption), timed on
    a bounded sample of iterations on one host thread."""
    import numpy as np

    from squander_hip import Context, agents, agents_options
    from squander_hip import workloads as W
    n = 8
    gates = W.ansatz_8q_40(2026)
    P = W.param_count(gates)
    theta = W.random_params(P, 31)
    T = o.apply_to(gates, theta, np.eye(256)).conj().T.copy()
    x0 = W.random_params(P, 32)
    ctx = Context(0)
    ctx.set_circuit(n, gates)
    ctx.set_target(T)
    ctx.set_jit(True)
    res = {"workload": "8-qubit ansatz (264 angles), reachable target, 64 agents, circuit-specialised kernels"}
    k_cpu, k_gpu = 12, 200
    agents(ctx, x0, 5, agents_options(max_iterations=2))  # warm-up
    t0 = time.perf_counter()
    fg, xg, itg, evg = agents(ctx, x0, 5, agents_options(max_iterations=k_gpu, optimization_tolerance=1e-8))
    tg = time.perf_counter() - t0
    fs, xs, its, evs = agents(ctx, x0, 5, agents_options(max_iterations=k_cpu, optimization_tolerance=1e-8))
    t0 = time.perf_counter()
    fo, xo, ito, evo = o.agents(gates, T, n, 5, x0, agent_num=64, max_iterations=k_cpu, tolerance=1e-8)
    tc = time.perf_counter() - t0
    res["gpu"] = {"iterations": itg, "evaluations": evg, "seconds": tg, "iterations_per_s": itg / tg,
                  "evals_per_s": evg / tg, "final_cost": fg}
    res["cpu"] = {"iterations": ito, "evaluations": evo, "seconds": tc, "iterations_per_s": ito / tc,
                  "evals_per_s": evo / tc, "cores": 1, "kind": "port",
                  "sample": f"the first {k_cpu} iterations of the same run"}
    res["same_trajectory"] = bool(its == ito and evs == evo and abs(fs - fo) < 1e-10)
    log(f"AGENTS 8q: GPU {itg / tg:.1f} it/s, CPU {ito / tc:.2f} it/s, same trajectory {res['same_trajectory']}")
    # K runs against K targets in lockstep: each stage's K x 64 costs in one launch
    from squander_hip import agents_batch
    K = 8
    Ts = np.array([o.apply_to(gates, W.random_params(P, 40 + k), np.eye(256)).conj().T for k in range(K)])
    x0s = np.array([W.random_params(P, 60 + k) for k in range(K)])
    seeds = np.arange(1, K + 1, dtype=np.uint32)
    ob = agents_options(max_iterations=k_gpu, optimization_tolerance=1e-8)
    agents_batch(ctx, Ts, x0s, seeds, agents_options(max_iterations=2))  # warm-up
    t0 = time.perf_counter()
    fb, xb, itb, evb = agents_batch(ctx, Ts, x0s, seeds, ob)
    tb = time.perf_counter() - t0
    ctx.set_target(Ts[K - 1])
    f1, x1, it1, ev1 = agents(ctx, x0s[K - 1], int(seeds[K - 1]), ob)
    same = bool(it1 == itb[K - 1] and ev1 == evb[K - 1] and f1 == fb[K - 1] and np.array_equal(x1, xb[K - 1]))
    res["gpu_batch"] = {"runs": K, "iterations": int(itb.sum()), "evaluations": int(evb.sum()), "seconds": tb,
                        "iterations_per_s": float(itb.sum()) / tb, "evals_per_s": float(evb.sum()) / tb,
                        "final_costs_max": float(fb.max()), "run_equals_single_call": same}
    log(f"AGENTS 8q x{K} targets: {itb.sum() / tb:.1f} it/s aggregate, run equals single call {same}")
    return res


def uneven_leg(o):
    """C4, with each reported final cost checked against the FP64 oracle's cost of the
    returned parameters (the acceptance is FP64 on the GPU; the descent is complex64)."""
    from squander_hip import uneven_bench
    res, probs, raw = uneven_bench.run(count=256, seed=2028, tol=1e-8, mixed=True, log=log)
    res["workload"] = ("256 decompositions, qubits uniform 3..6, U3 layer + 4..20 random adjacent CNOT layers, "
                       "half Haar / half ansatz-reachable targets, complex64 descent + FP64 acceptance")
    err = 0.0
    for i, (n, g, U, _, _) in enumerate(probs):
        err = max(err, abs(float(raw["f"][i]) - o.optimization_problem(g, U, n, raw["x"][i])))
    res["max_err_reported_cost_vs_fp64_oracle"] = err
    log(f"uneven: max |reported cost - FP64 oracle cost| = {err:.1e}")
    # the same batch in FP64 throughout, and through the lockstep rounds
    for key, kw in (("fp64", {"mixed": False}), ("rounds_mixed", {"mixed": True, "engine": "rounds"})):
        r2, _, _ = uneven_bench.run(count=256, seed=2028, tol=1e-8, log=log, **kw)
        res[key] = {k: r2[k] for k in ("seconds", "solves_per_s", "evals_p50", "evals_p99", "engine")}
        res[key]["success_reachable"] = r2["reachable"]["success"]
        res[key]["success_haar"] = r2["haar"]["success"]
    return res


def solves_leg(o):
    """C5, with every reported final cost of the batched and learned legs checked against
    the FP64 oracle's cost of the returned parameters."""
    from squander_hip import solve_bench
    from squander_hip.workloads import ansatz_4q_12
    keep = {}
    res = solve_bench.run(n_train=200_000, n_reach=4096, n_haar=1024, epochs=20, tol=1e-8, n_plain=128, log=log,
                          keep=keep)
    res["workload"] = ("4 qubits, U3 layer + 12 alternating CNOT layers (84 angles); warm-start network trained on "
                       "200k generated pairs for 20 epochs; evaluation on 4096 held-out reachable + 1024 Haar targets "
                       "(plain single-run leg on the first 128 of each)")
    gates = ansatz_4q_12()
    for name, kept in keep.items():
        for leg in ("batched", "learned"):
            f, x = kept[leg]
            err = max(abs(float(f[i]) - o.optimization_problem(gates, kept["Umtx"][i], 4, x[i])) for i in range(len(f)))
            res["results"][name][("batched_fp64" if leg == "batched" else "learned_mixed")][
                "max_err_reported_cost_vs_fp64_oracle"] = err
    log("4q solves: max |reported cost - FP64 oracle cost| = "
        + ", ".join(f"{k} {v['batched_fp64']['max_err_reported_cost_vs_fp64_oracle']:.1e} / "
                    f"{v['learned_mixed']['max_err_reported_cost_vs_fp64_oracle']:.1e}"
                    for k, v in res["results"].items()))
    return res


def c1_leg(o, layers=12, seconds=6.0):
    """The plain decomposition of 3-qubit targets on the host cores (the oracle's
    restated BFGS driver, one target per thread) beside the GPU's batched driver, with
    the phase-free cost 1 - |Tr(U^dagger V)|/8 and stop at 1e-8.  At 12 CNOT layers
    (the configured structure) neither finds exact solutions from random starts (a
    generic 3-qubit unitary needs >= 14 CNOTs, and the reachable targets sit in a
    landscape of local minima near 1e-3 that scipy's BFGS finds too); at 16 layers
    every target solves, which is where solves/s is measured."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor

    import oracle as orc
    from squander_hip import COST_ABS_TRACE, Context, decompose_batch, default_options, initial_guess
    from squander_hip import workloads as W
    n = 3
    gates = W.ansatz_3q(layers)
    P = W.param_count(gates)
    # half Haar targets, half reachable targets built from seeded angles
    Ts = [W.haar_unitary(8, 5000 + i).conj().T.copy() for i in range(128)]
    Ts += [o.apply_to(gates, W.random_params(P, 6000 + i), np.eye(8)).conj().T.copy() for i in range(128)]
    kind = np.array([0] * 128 + [1] * 128)
    threads = min(os.cpu_count() or 1, 64)
    order = np.random.default_rng(0).permutation(len(Ts))  # mix Haar and reachable in the time-bounded sample
    done, ok = 0, {0: 0, 1: 0}
    seen = {0: 0, 1: 0}
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        while time.perf_counter() - t0 < seconds and done < len(Ts):
            batch = [int(i) for i in order[done:min(done + threads, len(Ts))]]
            fs = list(ex.map(lambda i: o.decompose_custom(gates, Ts[i], n, 11 + i, optimization_tolerance=1e-8,
                                                          variant=orc.COST_ABS_TRACE)[0], batch))
            for i, f in zip(batch, fs):
                seen[kind[i]] += 1
                ok[kind[i]] += int(f <= 1e-8)
            done += len(batch)
    cpu_dt = time.perf_counter() - t0
    ctx = Context(0)
    ctx.set_circuit(n, gates)
    ctx.set_cost_function(COST_ABS_TRACE)
    x0 = np.array([initial_guess(P, 11 + i) for i in range(len(Ts))])
    t0 = time.perf_counter()
    res = decompose_batch(ctx, np.array(Ts), np.arange(len(Ts), dtype=np.int32), x0,
                          default_options(n, optimization_tolerance=1e-8))
    gpu_dt = time.perf_counter() - t0
    return {"workload": f"3 qubits, U3 layer + {layers} alternating CNOT layers ({P} angles), 128 Haar + 128 "
                        f"reachable targets, random starts, 1 - |Tr|/8 <= 1e-8",
            "cpu": {"targets": done, "seconds": cpu_dt, "decompositions_per_s": done / cpu_dt,
                    "solves_per_s": (ok[0] + ok[1]) / cpu_dt,
                    "success_reachable": ok[1] / max(seen[1], 1), "success_haar": ok[0] / max(seen[0], 1),
                    "cores": threads, "kind": "port"},
            "gpu_batched": {"targets": len(Ts), "seconds": gpu_dt, "decompositions_per_s": len(Ts) / gpu_dt,
                            "solves_per_s": float(np.sum(res["f"] <= 1e-8)) / gpu_dt,
                            "success_reachable": float(np.mean(res["f"][kind == 1] <= 1e-8)),
                            "success_haar": float(np.mean(res["f"][kind == 0] <= 1e-8)),
                            "max_fp64_cost_of_solves": float(max((o.optimization_problem(
                                gates, Ts[i], n, res["x"][i], variant=orc.COST_ABS_TRACE)
                                for i in range(len(Ts)) if res["f"][i] <= 1e-8), default=float("nan"))),
                            "same_start_points_as_cpu": True}}


def copy_calibration(dev):
    """A plain 256 MiB device copy, three times: the reference kernel whose PMC byte
    counters calibrate FETCH_SIZE / WRITE_SIZE."""
    import torch
    nbytes = 256 << 20
    src = torch.ones(nbytes, dtype=torch.uint8, device=dev)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    return {"bytes_read": nbytes, "bytes_written": nbytes, "launches": 3}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--only-main", action="store_true", help="time the headline step only (for profiling)")
    ap.add_argument("--only-gradient", action="store_true",
                    help="time the batched 8-qubit cost + gradient only (for profiling the adjoint kernel)")
    ap.add_argument("--skip-solves", action="store_true", help="skip the decomposition legs (4q solves, uneven batch)")
    ap.add_argument("--headline", choices=["fused", "jit", "generic"], default="fused",
                    help="headline kernel: fused 4-qubit blocks on the matrix cores (default), the "
                         "circuit-specialised register-tile kernel, or the generic one")
    ap.add_argument("--generic", action="store_true", help="same as --headline generic")
    ap.add_argument("--only-mlp", action="store_true",
                    help="time the warm-start network's training steps only (for profiling the BF16 GEMM kernel)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 256 MiB device copy (the reference kernel for PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the library has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from squander_hip import PREC_F32, PREC_F64, Context
    from squander_hip import workloads as W

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    dev = torch.device("cuda:0")
    n, d = 8, 256
    gates = W.ansatz_8q_40(2026)
    P = W.param_count(gates)
    B = args.batch
    U = W.haar_unitary(d, 2026).conj().T.copy()

    headline = "generic" if args.generic else args.headline
    # the context of the other legs: circuit-specialised register-tile kernels (compiled
    # on first use) for the cost-only legs that name them
    ctx = Context(0)
    ctx.set_circuit(n, gates)
    ctx.set_target(U)
    ctx.set_jit(True)
    ctx.set_fused_cost(False)
    L = ctx._L
    hctx = ctx
    if headline != "jit":
        hctx = Context(0)
        hctx.set_circuit(n, gates)
        hctx.set_target(U)
        hctx.set_fused_cost(headline == "fused")
    params = torch.from_numpy(W.random_params(P, 7, batch=B)).to(dev)
    costs = torch.empty(B, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    if args.only_mlp:
        res = mlp_train_leg()
        print(json.dumps({"metric": "warm-start MLP training samples/s (BF16 MFMA)", "value": res["samples_per_s"],
                          "unit": "samples/s", "higher_is_better": True, "dtype": "bf16", "mlp": res}), flush=True)
        return
    if args.only_gradient:
        gctx = Context(0)
        gctx.set_circuit(n, gates)
        gctx.set_target(U)
        gctx.set_fused_cost(True)
        res = gradient_leg(gctx, params, gates, U, orc.Oracle(), reps=args.steps, fused=True)
        line = {"metric": "cost+gradient evals/s @8q batch 512 (FP64)", "value": res["grads_per_s"],
                "unit": "gradients/s", "higher_is_better": True, "dtype": "f64", "gradient": res}
        if args.copy_calibration:
            line["copy_calibration"] = copy_calibration(dev)
        print(json.dumps(line), flush=True)
        return

    def step(prec=PREC_F64):
        hctx.check(L.sqh_optimization_problem_batch_dev(hctx.h, params.data_ptr(), B, prec, costs.data_ptr()))

    log(f"warmup {args.warmup} steps")
    t_w = time.perf_counter()
    for _ in range(max(args.warmup, 1)):
        step()
    hctx.synchronize()
    jit_ms, _ = hctx.jit_stats()
    log(f"warm-up done in {time.perf_counter() - t_w:.1f} s (run-time kernel compilation {jit_ms:.0f} ms)")
    hctx.kernel_timing()
    hctx.set_kernel_timing(True)
    log(f"timing {args.steps} steps of B={B} cost evaluations (FP64, headline kernel: {headline})")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    hctx.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kms, nl = hctx.kernel_timing()
    hctx.set_kernel_timing(False)
    wall = t1 - t0
    value = B * args.steps / wall
    k_avg_s = kms / max(nl, 1) / 1e3
    flops_eval, _ = gate_work(gates, d, d, 16)
    # the fused chain's algorithmic flops: per column 2 blocks on the 128 rows of a
    # 7-qubit support and 3 on all 256, then the product with the target's column (0.42x
    # the gate arithmetic)
    flops_launch = (fused_cost_flops(gates, n, d) if headline == "fused" else flops_eval) * B
    achieved = flops_launch / k_avg_s / 1e12
    log(f"headline: {value:.0f} evals/s, kernel {k_avg_s * 1e3:.3f} ms, {achieved:.1f} TFLOP/s")

    # error of the GPU costs against the FP64 oracle on a sample
    o = orc.Oracle()
    pc = params[:8].cpu().numpy()
    ref = o.optimization_problem_batched(gates, U, n, pc, nthreads=8)
    err64 = float(np.abs(costs[:8].cpu().numpy() - ref).max())

    out = {
        "metric": "cost evals/s @8q batch 512; 4q solves/s to cost\u22641e-8, learned+batched vs plain",
        "metric_parts": {"value": "FP64 (complex128) cost evals/s at 8 qubits, batch 512",
                         "solves": "4-qubit solves/s, learned + batched against plain: solves_4q"},
        "value": value,
        "unit": "cost evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Haar target, seeded uniform [0,2pi) angles)",
        "config": {"workload": "8-qubit U3 layer + 40 random adjacent CNOT layers (128 gates, 264 angles), "
                               "256x256 complex128 Haar target, batch of parameter vectors",
                   "qubits": n, "batch": B, "gates": int(len(gates)), "params": P},
        "err_vs_fp64_oracle": err64,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": (None if B != 512 else {"fused": FUSED_COST_8Q_B512_TRAFFIC,
                                                        "jit": JIT_FORWARD_8Q_B512_TRAFFIC,
                                                        "generic": TILE_FORWARD_8Q_B512_TRAFFIC}[headline]),
                     "kernel": {"fused": "k_fused_trace<double,8> (the transposed chain of the 6 fused 4-qubit "
                                         "blocks from unit vectors on v_mfma_f64_16x16x4f64: the last block's row "
                                         "loaded, 2 blocks on the 7-qubit support, 3 on all rows, then the product "
                                         "with the target's column)",
                                "jit": "sqh_jit_kernel (k_tile_forward specialised to the circuit, diagonal mode)",
                                "generic": "k_tile_forward<8,double,1>"}[headline],
                     "peak_note": ("FP64 dense matrix-core peak" if headline == "fused" else
                                   "FP64 vector peak (equal to the FP64 matrix peak on MI355X): a VALU kernel"),
                     "kernel_avg_ms": k_avg_s * 1e3, "algorithmic_flops_per_launch": flops_launch,
                     "launch_shape": (dict(zip(("tiles_per_wave", "exchange_groups", "waves", "prefetch"),
                                               hctx.fused_cost_config(False, B))) if headline == "fused" else None),
                     "flops": ("fused chain: per column 128 flop per support row for blocks 5..1 (128, "
                               "128, 256, 256, 256 rows), 8 flop per row for the product with the target's "
                               "column" if headline == "fused"
                               else "gate arithmetic: 14 flop per element per 1-qubit gate")},
        "kernel_compile_ms": jit_ms,
    }
    if headline == "fused":
        out["fused_forward_pass"] = fused_forward_leg(gates, U, params, costs, args.steps, n, d)
        log(f"fused forward pass (no chain): {out['fused_forward_pass']['evals_per_s']:.0f} evals/s")
    if not args.only_main:
        if headline != "jit":
            out["tile_kernel_specialised"] = tile_leg(gates, U, params, costs, args.steps, flops_eval * B, True)
            log(f"tile kernel (specialised): {out['tile_kernel_specialised']['evals_per_s']:.0f} evals/s")
        if headline != "generic":
            out["tile_kernel_generic"] = tile_leg(gates, U, params, costs, args.steps, flops_eval * B, False)
    if args.copy_calibration:
        out["copy_calibration"] = copy_calibration(dev)
    if args.only_main:
        print(json.dumps(out), flush=True)
        return

    # complex64 legs: the headline path and the specialised register-tile kernel
    def c64_leg(c, fl, fused):
        run = lambda: c.check(L.sqh_optimization_problem_batch_dev(c.h, params.data_ptr(), B, PREC_F32,
                                                                   costs.data_ptr()))
        for _ in range(2):
            run()
        c.synchronize()
        c.set_kernel_timing(True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            run()
        c.synchronize()
        t1 = time.perf_counter()
        kms32, nl32 = c.kernel_timing()
        c.set_kernel_timing(False)
        k32_s = kms32 / max(nl32, 1) / 1e3
        tf = fl / k32_s / 1e12
        r = {"value": B * args.steps / (t1 - t0), "unit": "cost evals/s",
             "err_vs_fp64_oracle": float(np.abs(costs[:8].cpu().numpy() - ref).max()),
             "kernel_avg_ms": k32_s * 1e3, "achieved_TFLOPs": tf}
        if fused:
            r.update({"kernel": "k_fused_cost<float,8> (v_mfma_f32_16x16x4f32, FP32 accumulation)",
                      "peak_TFLOPs": FP32_MFMA_PEAK_TFLOPS, "frac": tf / FP32_MFMA_PEAK_TFLOPS})
        else:
            # packed FP32 (two FMAs per lane per cycle, which the 157.3 TFLOP/s figure
            # assumes) is not generated here, so the attainable FP32 VALU rate is the FP64 one
            r.update({"kernel": "sqh_jit_kernel (complex64)", "peak_TFLOPs_unpacked": FP64_PEAK_TFLOPS,
                      "frac_of_unpacked_peak": tf / FP64_PEAK_TFLOPS})
        return r

    out["c64"] = c64_leg(hctx, flops_launch, headline == "fused")
    log(f"c64: {out['c64']['value']:.0f} evals/s, err {out['c64']['err_vs_fp64_oracle']:.2e}")
    if headline != "jit":
        out["c64_tile_kernel_specialised"] = c64_leg(ctx, flops_eval * B, False)
        log(f"c64 tile kernel: {out['c64_tile_kernel_specialised']['value']:.0f} evals/s")

    def apply_err(fn, f32=False):
        """Max relative Frobenius error of two matrices applied once against the oracle."""
        m2 = torch.from_numpy(np.array(np.broadcast_to(U, (2, d, d)), dtype=np.complex64 if f32 else np.complex128,
                                       order="C")).to(dev)
        ctx.check(fn(ctx.h, params.data_ptr(), P, m2.data_ptr(), d, d, d, d * d, 2, int(f32)))
        ctx.synchronize()
        got = m2.cpu().numpy()
        err = 0.0
        for b in range(2):
            ref = o.apply_to(gates, params[b].cpu().numpy(), U)
            err = max(err, float(np.linalg.norm(got[b] - ref) / np.linalg.norm(ref)))
        return err

    # unfused streaming path: one HBM pass per gate over B matrices of 1 MiB
    mats = torch.from_numpy(np.broadcast_to(U, (B, d, d)).copy()).to(dev)
    torch.cuda.synchronize()

    def stream_step():
        ctx.check(L.sqh_apply_to_stream_dev(ctx.h, params.data_ptr(), P, mats.data_ptr(), d, d, d, d * d, B, 0))

    stream_step()
    ctx.synchronize()
    ctx.set_kernel_timing(True)
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        stream_step()
    ctx.synchronize()
    t1 = time.perf_counter()
    skms, snl = ctx.kernel_timing()
    ctx.set_kernel_timing(False)
    _, bytes_pass = gate_work(gates, d, d, 16)
    gbs = bytes_pass * B * reps / (skms / 1e3) / 1e9
    out["unfused_stream"] = {"achieved_GBs": gbs, "peak_GBs": HBM_PEAK_GBS, "frac": gbs / HBM_PEAK_GBS,
                             "evals_per_s": B * reps / (t1 - t0), "kernel": "k_stream_gate<double>",
                             "launches": snl, "bytes_per_pass": bytes_pass * B,
                             "err_rel_fro_vs_fp64_oracle": apply_err(L.sqh_apply_to_stream_dev),
                             "note": "algorithmic bytes over summed per-gate kernel time"}
    log(f"unfused stream: {gbs:.0f} GB/s")

    # fused streaming path: runs of gates as 16x16 blocks on 4 qubits (FP64 matrix cores)
    def fused_step():
        ctx.check(L.sqh_apply_to_fused_dev(ctx.h, params.data_ptr(), P, mats.data_ptr(), d, d, d, d * d, B, 0))

    fused_step()
    ctx.synchronize()
    ctx.set_kernel_timing(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        fused_step()
    ctx.synchronize()
    t1 = time.perf_counter()
    fkms, fnl = ctx.kernel_timing()
    ctx.set_kernel_timing(False)
    blk_bytes = 2.0 * 16 * d * d * B     # read + write of every matrix per block pass
    out["fused_stream"] = {"evals_per_s": B * reps / (t1 - t0), "blocks_per_pass": fnl // reps,
                           "achieved_GBs": blk_bytes * fnl / (fkms / 1e3) / 1e9,
                           "algorithmic_TFLOPs": flops_eval * B * reps / (fkms / 1e3) / 1e12,
                           "err_rel_fro_vs_fp64_oracle": apply_err(L.sqh_apply_to_fused_dev),
                           "kernel": "k_block_mfma4<double> (v_mfma_f64_16x16x4f64)"}
    log(f"fused stream: {out['fused_stream']['evals_per_s']:.0f} evals/s")

    # apply_to of the same batch in place, every column resident on the chip: the fused
    # blocks' store pass on the matrix cores, and the specialised register-tile kernel
    def apply_leg(c, name, kernel, fl):
        run = lambda: c.check(L.sqh_apply_to_dev(c.h, params.data_ptr(), P, mats.data_ptr(), d, d, d, d * d, B, 0))
        run()
        c.synchronize()
        c.set_kernel_timing(True)
        t0 = time.perf_counter()
        for _ in range(reps):
            run()
        c.synchronize()
        t1 = time.perf_counter()
        akms, anl = c.kernel_timing()
        c.set_kernel_timing(False)
        k_s = akms / max(anl, 1) / 1e3
        r = {"evals_per_s": B * reps / (t1 - t0), "kernel_avg_ms": k_s * 1e3, "achieved_TFLOPs": fl * B / k_s / 1e12,
             "hbm_GBs": 2.0 * 16 * d * d * B / k_s / 1e9, "kernel": kernel,
             "err_rel_fro_vs_fp64_oracle": apply_err(lambda h, pp, Pp, mm, dd, cc, ss, bs, bb, f: L.sqh_apply_to_dev(
                 c.h, pp, Pp, mm, dd, cc, ss, bs, bb, f))}
        log(f"apply_to {name}: {r['evals_per_s']:.0f} matrices/s, kernel {k_s * 1e3:.3f} ms")
        return r

    fctx = Context(0)
    fctx.set_circuit(n, gates)
    fctx.set_fused_cost(True)
    out["apply_to_fused_resident"] = apply_leg(fctx, "fused resident",
                                               "k_fused_cost<double,8,...> store pass (v_mfma_f64_16x16x4f64)",
                                               block_flops(gates, n, d))
    out["apply_to_tile_specialised"] = apply_leg(ctx, "tile specialised",
                                                 "sqh_jit_kernel (k_tile_forward specialised, store mode)", flops_eval)
    del mats

    # the same in complex64: 16x16 blocks on v_mfma_f32_16x16x4f32, FP32 accumulation
    mats32 = torch.from_numpy(np.array(np.broadcast_to(U, (B, d, d)), dtype=np.complex64, order="C")).to(dev)
    assert mats32.is_contiguous() and mats32.stride(0) == d * d

    def fused_step32():
        ctx.check(L.sqh_apply_to_fused_dev(ctx.h, params.data_ptr(), P, mats32.data_ptr(), d, d, d, d * d, B, 1))

    fused_step32()
    ctx.synchronize()
    ctx.set_kernel_timing(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        fused_step32()
    ctx.synchronize()
    t1 = time.perf_counter()
    fkms32, fnl32 = ctx.kernel_timing()
    ctx.set_kernel_timing(False)
    out["fused_stream_c64"] = {"evals_per_s": B * reps / (t1 - t0), "blocks_per_pass": fnl32 // reps,
                               "achieved_GBs": blk_bytes / 2 * fnl32 / (fkms32 / 1e3) / 1e9,
                               "err_rel_fro_vs_fp64_oracle": apply_err(L.sqh_apply_to_fused_dev, f32=True),
                               "kernel": "k_block_mfma4<float> (v_mfma_f32_16x16x4f32)"}
    log(f"fused stream c64: {out['fused_stream_c64']['evals_per_s']:.0f} evals/s, "
        f"{out['fused_stream_c64']['achieved_GBs']:.0f} GB/s, err {out['fused_stream_c64']['err_rel_fro_vs_fp64_oracle']:.1e}")
    del mats32

    # single-vector cost + gradient latency (8 qubits, 264 angles): the default path
    # (fused blocks) and the register-tile adjoint
    f0 = torch.empty(1, dtype=torch.float64, device=dev)
    g0 = torch.empty(P, dtype=torch.float64, device=dev)
    p1 = params[:1].contiguous()
    fctx = Context(0)
    fctx.set_circuit(n, gates)
    fctx.set_target(U)
    fr, gr = o.optimization_problem_combined(gates, U, n, p1.cpu().numpy()[0])
    for name, c in (("grad_8q_single", fctx), ("grad_8q_single_tile", ctx)):
        def grad_step():
            c.check(L.sqh_optimization_problem_combined_batch_dev(c.h, p1.data_ptr(), 1, PREC_F64, f0.data_ptr(),
                                                                   g0.data_ptr()))

        for _ in range(3):
            grad_step()
        c.synchronize()
        t0 = time.perf_counter()
        for _ in range(20):
            grad_step()
        c.synchronize()
        t1 = time.perf_counter()
        out[name] = {"ms": (t1 - t0) / 20 * 1e3, "err_grad_vs_oracle": float(np.abs(g0.cpu().numpy() - gr).max()),
                     "err_cost_vs_oracle": float(abs(f0.item() - fr)),
                     "path": "fused blocks (default)" if c is fctx else "register-tile adjoint (specialised)"}
    log(f"8q single cost+gradient: fused {out['grad_8q_single']['ms']:.3f} ms, tile "
        f"{out['grad_8q_single_tile']['ms']:.3f} ms")

    fctx.set_fused_cost(True)
    out["grad_8q_b512"] = gradient_leg(fctx, params, gates, U, o, fused=True)
    out["grad_8q_b512_tile"] = gradient_leg(ctx, params, gates, U, o)
    out["ten_qubit_b64"] = ten_qubit_leg(args.steps, o)
    out["apply_derivate_8q"] = derivate_leg(gates, P, U)
    if not args.skip_solves:
        out["agents_8q"] = agents_leg(o)
        out["uneven_256"] = uneven_leg(o)
        out["solves_4q"] = solves_leg(o)
    if not args.skip_cpu:
        log("cpu baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(gates, U, n, P, o)
        out["c1_3q_plain_solves"] = c1_leg(o, 12)
        out["c1_3q_16_layers"] = c1_leg(o, 16)
        log(f"C1 3q: 12 layers GPU {out['c1_3q_plain_solves']['gpu_batched']['decompositions_per_s']:.0f} dec/s; "
            f"16 layers GPU {out['c1_3q_16_layers']['gpu_batched']['solves_per_s']:.0f} solves/s "
            f"(success {out['c1_3q_16_layers']['gpu_batched']['success_reachable']:.2f} reachable / "
            f"{out['c1_3q_16_layers']['gpu_batched']['success_haar']:.2f} Haar), CPU "
            f"{out['c1_3q_16_layers']['cpu']['solves_per_s']:.0f} solves/s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
