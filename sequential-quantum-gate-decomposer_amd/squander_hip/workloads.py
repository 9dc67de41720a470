"""Seeded synthetic workloads: Haar-random and ansatz-reachable targets and the
fixed ansatz families the benchmarks use.

Gate lists are (G, 4) int32 arrays of (type, target, control, param_start) with
the reference's gate-type ids; parameter starts are assigned in list order as
Gates_block::add_gate does (Gates_block.cpp:2500-2525).
"""
from __future__ import annotations

import numpy as np

from ._native import GATE_CNOT, GATE_CRY, GATE_CZ, GATE_RX, GATE_RY, GATE_RZ, GATE_U3, PARAM_NUM


def haar_unitary(d: int, seed) -> np.ndarray:
    """Haar-random d x d unitary: QR of a complex Gaussian matrix with the
    R-diagonal phase fix (Mezzadri's algorithm; scipy.stats.unitary_group)."""
    rng = seed if isinstance(seed, np.random.Generator) else np.random.default_rng(seed)
    z = (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))) / np.sqrt(2.0)
    q, r = np.linalg.qr(z)
    ph = np.diag(r) / np.abs(np.diag(r))
    return (q * ph[None, :]).astype(np.complex128)


class GateListBuilder:
    """Appends gates and assigns parameter starts in order."""

    def __init__(self, qbit_num: int):
        self.qbit_num = qbit_num
        self.rows = []
        self.P = 0

    def add(self, gtype, target, control=-1):
        self.rows.append((gtype, target, control, self.P if PARAM_NUM[gtype] else 0))
        self.P += PARAM_NUM[gtype]
        return self

    def array(self) -> np.ndarray:
        return np.asarray(self.rows, dtype=np.int32).reshape(-1, 4)


def ansatz_layers(qbit_num: int, pairs) -> np.ndarray:
    """U3 on every qubit, then for each adjacent pair (a, a+1): CNOT(target a,
    control a+1), U3(a), U3(a+1)."""
    b = GateListBuilder(qbit_num)
    for q in range(qbit_num):
        b.add(GATE_U3, q)
    for a in pairs:
        b.add(GATE_CNOT, a, a + 1)
        b.add(GATE_U3, a)
        b.add(GATE_U3, a + 1)
    return b.array()


def random_adjacent_pairs(qbit_num: int, layers: int, seed) -> list:
    rng = np.random.default_rng(seed)
    return [int(a) for a in rng.integers(0, qbit_num - 1, size=layers)]


def alternating_pairs(qbit_num: int, layers: int) -> list:
    """0-1, 1-2, ..., wrapping over the adjacent pairs (the 3-qubit case alternates 0-1, 1-2)."""
    return [i % (qbit_num - 1) for i in range(layers)]


def ansatz_8q_40(seed: int = 2026) -> np.ndarray:
    """8 qubits, U3 layer + 40 random adjacent CNOT layers: 128 gates, 264 parameters."""
    return ansatz_layers(8, random_adjacent_pairs(8, 40, seed))


def ansatz_10q_60(seed: int = 2027) -> np.ndarray:
    """10 qubits, U3 layer + 60 random adjacent CNOT layers: 190 gates, 390 parameters."""
    return ansatz_layers(10, random_adjacent_pairs(10, 60, seed))


def ansatz_4q_12() -> np.ndarray:
    """4 qubits, U3 layer + 12 CNOT layers on alternating adjacent pairs: 84 parameters."""
    return ansatz_layers(4, alternating_pairs(4, 12))


def ansatz_3q(layers: int = 12) -> np.ndarray:
    return ansatz_layers(3, alternating_pairs(3, layers))


def param_count(gates) -> int:
    g = np.asarray(gates).reshape(-1, 4)
    return int(sum(PARAM_NUM[int(t)] for t in g[:, 0]))


def random_params(P: int, seed, batch: int | None = None) -> np.ndarray:
    """Raw parameters uniform in [0, 2pi) (the reference's RANDOM initial guess)."""
    rng = seed if isinstance(seed, np.random.Generator) else np.random.default_rng(seed)
    shape = (P,) if batch is None else (batch, P)
    return rng.uniform(0.0, 2.0 * np.pi, size=shape)


def mixed_gate_circuit(qbit_num: int, n_gates: int, seed) -> np.ndarray:
    """A random circuit over all seven supported gate types (test coverage)."""
    rng = np.random.default_rng(seed)
    b = GateListBuilder(qbit_num)
    types = [GATE_U3, GATE_RX, GATE_RY, GATE_RZ, GATE_CNOT, GATE_CZ, GATE_CRY]
    for _ in range(n_gates):
        t = types[int(rng.integers(len(types)))]
        tgt = int(rng.integers(qbit_num))
        if t in (GATE_CNOT, GATE_CZ, GATE_CRY):
            if qbit_num < 2:
                t = GATE_U3
                b.add(t, tgt)
                continue
            c = int(rng.integers(qbit_num - 1))
            c = c if c < tgt else c + 1
            b.add(t, tgt, c)
        else:
            b.add(t, tgt)
    return b.array()
