"""CPU checks of the host-side tile-program compiler and the kernels' algorithm.

The op/layout records come from libsqhip itself (sqh_inspect_program); the
emulator in tile_emulator.py executes them as the HIP kernels do and is compared
with the oracle: forward application (Gates_block::apply_to), the trace entries,
and the adjoint gradient.  No GPU is needed.
"""
import numpy as np
import pytest

import oracle as orc
import tile_emulator as te
from conftest import rel_fro
from squander_hip import _native as nat
from squander_hip import workloads as W


def kernels_for(oracle, n, gates, params):
    """Kernel table in the program's slot order (the compiler may run gates on
    disjoint qubits out of list order)."""
    ks = []
    for t, ps, npar, _ in nat.inspect_slots(n, gates):
        ks.append(oracle.gate_kernel(int(t), params[ps:ps + npar]).reshape(-1))
    return ks


@pytest.mark.parametrize("n", list(range(1, 11)))
def test_forward_program_matches_oracle(oracle, n):
    gates = W.mixed_gate_circuit(n, 6 * n + 10, seed=300 + n)
    P = W.param_count(gates)
    p = np.random.default_rng(n).uniform(0, 2 * np.pi, P)
    ops, lays = te.decode(*nat.inspect_program(n, gates))
    d = 1 << n
    m = min(d, 48)
    X = np.random.default_rng(7).standard_normal((d, m)) + 1j * np.random.default_rng(8).standard_normal((d, m))
    out, _ = te.run_forward(n, ops, lays, kernels_for(oracle, n, gates, p), X)
    assert rel_fro(out, oracle.apply_to(gates, p, X)) < 1e-13


@pytest.mark.parametrize("n,seed", [(5, 1), (6, 2), (8, 3), (10, 4)])
def test_schedule_keeps_each_qubits_gate_order(n, seed):
    """The compiler may run gates on disjoint qubits out of list order, never two gates
    that share a qubit: along every qubit, the parametric gates' slot order follows
    the list order."""
    gates = W.mixed_gate_circuit(n, 8 * n, seed=900 + seed)
    slots = nat.inspect_slots(n, gates)
    index_of = {}
    for i, (t, tg, c, ps) in enumerate(gates):
        if nat.PARAM_NUM[int(t)]:
            index_of[int(ps)] = i
    order = [index_of[int(ps)] for _, ps, _, _ in slots]
    assert sorted(order) == sorted(index_of.values())
    for q in range(n):
        on_q = [i for i in order if q in (int(gates[i][1]), int(gates[i][2]))]
        assert on_q == sorted(on_q)


def test_schedule_remap_counts_of_the_benchmark_circuits():
    """Dependency-order scheduling: 5 remaps for the 8-qubit 40-layer ansatz and 7 for
    the 10-qubit 60-layer one (gate order needs 14 and 24)."""
    for n, gates, limit in ((8, W.ansatz_8q_40(2026), 5), (10, W.ansatz_10q_60(2027), 7)):
        ops, _ = te.decode(*nat.inspect_program(n, gates))
        assert sum(1 for o in ops if o[0] == te.OP_REMAP) <= limit


def test_program_remaps_only_when_needed():
    # 8-qubit ansatz: every op target must be a register qubit; remaps are far fewer than gates.
    gates = W.ansatz_8q_40()
    ops, lays = te.decode(*nat.inspect_program(8, gates))
    nremap = sum(1 for o in ops if o[0] == te.OP_REMAP)
    assert len(ops) - nremap == len(gates)
    assert nremap < len(gates) // 3
    assert all(o[1] < 4 for o in ops if o[0] != te.OP_REMAP)


@pytest.mark.parametrize("n,off,variant", [(3, 0, orc.COST_FROBENIUS), (4, 0, orc.COST_HS), (5, 0, orc.COST_ABS_TRACE),
                                           (6, 8, orc.COST_FROBENIUS), (8, 0, orc.COST_INFIDELITY)])
def test_adjoint_gradient_matches_oracle(oracle, n, off, variant):
    gates = W.mixed_gate_circuit(n, 5 * n + 8, seed=400 + n)
    P = W.param_count(gates)
    p = np.random.default_rng(3 * n).uniform(0, 2 * np.pi, P)
    d = 1 << n
    m = d - off
    U = W.haar_unitary(d, 50 + n)[:, :m].copy()
    ops, lays = te.decode(*nat.inspect_program(n, gates))
    toff = off if variant in (orc.COST_FROBENIUS, orc.COST_ABS_TRACE) else 0
    trace, Ms = te.run_adjoint(n, ops, lays, kernels_for(oracle, n, gates, p), U, toff)
    # gradient from the correlations, as the finish kernel forms it
    grad = np.zeros(P)
    for slot, (t, ps, npar, _) in enumerate(nat.inspect_slots(n, gates)):
        for q in range(npar):
            D = oracle.deriv_kernel(int(t), p[ps:ps + npar], q).reshape(-1)
            tk = np.sum(D * Ms[slot])
            if variant == orc.COST_FROBENIUS:
                grad[ps + q] = -tk.real / m
            elif variant == orc.COST_HS:
                grad[ps + q] = -2.0 / m ** 2 * (trace.real * tk.real + trace.imag * tk.imag)
            elif variant == orc.COST_INFIDELITY:
                grad[ps + q] = -2.0 / m / (m + 1) * (trace.real * tk.real + trace.imag * tk.imag)
            else:
                grad[ps + q] = -(trace.real * tk.real + trace.imag * tk.imag) / (abs(trace) * m)
    f_ref, g_ref = oracle.optimization_problem_combined(gates, U, n, p, variant, off)
    assert abs(orc.brute_cost(oracle.apply_to(gates, p, U), variant, off) - f_ref) < 1e-13
    assert np.abs(grad - g_ref).max() < 1e-12


@pytest.mark.parametrize("n", [11, 12, 13, 16, 20, 24])
def test_compile_beyond_the_tile_range(n):
    """Circuits above the register tile (streaming paths) compile: one slot per
    parametric gate in gate order, one op per gate plus the remaps.  (A layout holds lane
    positions for the tile range only; 13 qubits and up once wrote past it.)"""
    gates = W.mixed_gate_circuit(n, 6 * n, seed=1700 + n)
    ops, lays = te.decode(*nat.inspect_program(n, gates))
    nremap = sum(1 for o in ops if o[0] == te.OP_REMAP)
    assert len(ops) - nremap == len(gates)
    slots = nat.inspect_slots(n, gates)
    par = [int(g[3]) for g in gates if nat.PARAM_NUM[int(g[0])]]
    assert [int(ps) for _, ps, _, _ in slots] == par


def _embed(n, qubits, F):
    """Dense 2^n matrix of the block F acting on the sorted qubits (local bit j = qubits[j])."""
    d = 1 << n
    full = np.zeros((d, d), dtype=complex)
    mask = sum(1 << q for q in qubits)
    for r in range(d):
        li = sum(((r >> q) & 1) << j for j, q in enumerate(qubits))
        base = r & ~mask
        for lo in range(1 << len(qubits)):
            ro = base | sum(((lo >> j) & 1) << q for j, q in enumerate(qubits))
            full[ro, r] += F[lo, li]
    return full


@pytest.mark.parametrize("n,which", [(4, "mixed"), (6, "mixed"), (8, "ansatz"), (9, "mixed")])
def test_fusion_plan_is_a_legal_reordering(oracle, n, which):
    """The streaming path's 4-qubit blocks: every gate once, each block holds its gates'
    qubits, every qubit's gates keep their list order, and the product of the blocks
    (each built by the oracle from its gates) is the circuit."""
    gates = W.ansatz_8q_40(2026) if which == "ansatz" else W.mixed_gate_circuit(n, 7 * n, seed=2700 + n)
    P = W.param_count(gates)
    p = np.random.default_rng(n).uniform(0, 2 * np.pi, P)
    plan = nat.inspect_fusion(n, gates)
    order = [i for _, src in plan for i in src]
    assert sorted(order) == list(range(len(gates)))
    for q in range(n):
        on_q = [i for i in order if q in (int(gates[i][1]), int(gates[i][2]))]
        assert on_q == sorted(on_q)
    C = np.eye(1 << n, dtype=complex)
    for qubits, src in plan:
        assert len(qubits) == 4
        local = {q: j for j, q in enumerate(qubits)}
        sub = []
        for i in src:
            t, tg, c, ps = (int(v) for v in gates[i])
            ctrl = t in (W.GATE_CNOT, W.GATE_CZ, W.GATE_CRY)
            assert tg in local and (not ctrl or c in local)
            sub.append((t, local[tg], local[c] if ctrl else -1, ps))
        F = oracle.apply_to(np.array(sub, dtype=np.int32).reshape(-1, 4), p, np.eye(16, dtype=complex))
        C = _embed(n, list(qubits), F) @ C
    assert rel_fro(C, oracle.apply_to(gates, p, np.eye(1 << n, dtype=complex))) < 1e-12


def test_fusion_plan_block_counts():
    """Dependency packing: 6 blocks for the 8-qubit 40-layer ansatz and 9 for the
    10-qubit 60-layer one (gate-list order needs 15 and 26)."""
    assert len(nat.inspect_fusion(8, W.ansatz_8q_40(2026))) <= 6
    assert len(nat.inspect_fusion(10, W.ansatz_10q_60(2027))) <= 9


@pytest.mark.parametrize("n", [17, 20])
def test_fusion_plan_in_list_order_above_sixteen_qubits(n):
    """Above 16 qubits blocks grow in gate-list order: every gate once and in list
    order, each block holding its gates' qubits."""
    gates = W.mixed_gate_circuit(n, 4 * n, seed=2900 + n)
    plan = nat.inspect_fusion(n, gates)
    order = [i for _, src in plan for i in src]
    assert order == list(range(len(gates)))
    for qubits, src in plan:
        assert len(qubits) == 4 and list(qubits) == sorted(qubits)
        for i in src:
            t, tg, c, _ = (int(v) for v in gates[i])
            assert tg in qubits and (t not in (W.GATE_CNOT, W.GATE_CZ, W.GATE_CRY) or c in qubits)


@pytest.mark.parametrize("which", ["8q_40", "10q_60", "6q_mixed"])
def test_chain_flop_accounting_counts_the_rows_the_chain_touches(which):
    """bench.py's flops for the transposed chain (fused_cost_flops) against a brute-force
    count: a unit row pushed through random dense 16x16 blocks of the fusion plan, last
    block first, transposed, counting the rows of every outer group whose input is not
    all zero (the last block's row is loaded, so it costs nothing), then 8 flop per
    nonzero row for the product with the target's column."""
    import bench
    n, gates = {"8q_40": (8, W.ansatz_8q_40()), "10q_60": (10, W.ansatz_10q_60()),
                "6q_mixed": (6, W.mixed_gate_circuit(6, 30, 5))}[which]
    qs = [q for q, _ in nat.inspect_fusion(n, gates)]
    rng = np.random.default_rng(7)
    D = 1 << n
    v = np.zeros(D, dtype=complex)
    v[3 % D] = 1.0
    rows = []
    for b in reversed(range(len(qs))):
        q = qs[b]
        off = [sum(1 << q[k] for k in range(4) if (i >> k) & 1) for i in range(16)]
        Q = sum(1 << x for x in q)
        F = rng.normal(size=(16, 16)) + 1j * rng.normal(size=(16, 16))
        out = np.zeros_like(v)
        touched = 0
        for base in range(D):
            if base & Q:
                continue
            idx = [base | o for o in off]
            x = v[idx]
            if np.any(x != 0):
                touched += 16
                out[idx] = F.T @ x
        if b < len(qs) - 1:
            rows.append(touched)
        v = out
    expected = 128.0 * sum(rows) + 8.0 * np.count_nonzero(v)
    assert bench.fused_cost_flops(gates, n, 1) == expected, (rows, np.count_nonzero(v))
    if which == "8q_40":
        assert rows == [128, 128, 256, 256, 256]
