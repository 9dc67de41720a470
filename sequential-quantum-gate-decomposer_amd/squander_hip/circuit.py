"""``Circuit``: the reference's Python circuit object for the seven gates of the
hot path, executing on the GPU through libsqhip.

Method names, argument meaning and error behaviour follow
squander/gates/qgd_Circuit.py:73-600 and qgd_Circuit_Wrapper.cpp:861-998:
``apply_to`` works in place on a C-contiguous numpy matrix, complex128 with
float64 parameters or (``is_f32=True``) complex64 with float32 parameters, and
raises ``TypeError`` on a dtype mismatch; ``apply_derivate_to`` returns a list of
P new matrices.
"""
from __future__ import annotations

import threading

import numpy as np

from . import _native as nat


class Circuit:
    """Ordered gate list on ``qbit_num`` qubits (Gates_block)."""

    def __init__(self, qbit_num: int):
        if qbit_num < 1:
            raise ValueError("qbit_num must be positive")
        self.qbit_num = int(qbit_num)
        self._gates = []          # (type, target, control, param_start)
        self._P = 0
        self._ctx = None
        self._ctx_dirty = True
        self._ctx_lock = threading.Lock()   # first use from several threads at once

    # -- building --------------------------------------------------------
    def _add(self, gtype, target, control=-1):
        if not 0 <= target < self.qbit_num:
            raise Exception(f"{nat.GATE_NAMES[gtype]}: target qubit index out of range.")
        if gtype in (nat.GATE_CNOT, nat.GATE_CZ, nat.GATE_CRY):
            if not 0 <= control < self.qbit_num or control == target:
                raise Exception(f"{nat.GATE_NAMES[gtype]}: control qubit index invalid.")
        self._gates.append((gtype, int(target), int(control), self._P if nat.PARAM_NUM[gtype] else 0))
        self._P += nat.PARAM_NUM[gtype]
        self._ctx_dirty = True

    def add_U3(self, target_qbit):
        self._add(nat.GATE_U3, target_qbit)

    def add_RX(self, target_qbit):
        self._add(nat.GATE_RX, target_qbit)

    def add_RY(self, target_qbit):
        self._add(nat.GATE_RY, target_qbit)

    def add_RZ(self, target_qbit):
        self._add(nat.GATE_RZ, target_qbit)

    def add_CNOT(self, target_qbit, control_qbit):
        self._add(nat.GATE_CNOT, target_qbit, control_qbit)

    def add_CZ(self, target_qbit, control_qbit):
        self._add(nat.GATE_CZ, target_qbit, control_qbit)

    def add_CRY(self, target_qbit, control_qbit):
        self._add(nat.GATE_CRY, target_qbit, control_qbit)

    def add_Circuit(self, circuit: "Circuit"):
        """Append a sub-circuit; its parameters follow the current ones."""
        if circuit.qbit_num > self.qbit_num:
            raise Exception("add_Circuit: the sub-circuit acts on more qubits")
        for t, tg, c, ps in circuit._gates:
            self._gates.append((t, tg, c, self._P + ps if nat.PARAM_NUM[t] else 0))
        self._P += circuit._P
        self._ctx_dirty = True

    @classmethod
    def from_gate_array(cls, qbit_num, gates):
        c = cls(qbit_num)
        for t, tg, ctl, ps in np.asarray(gates).reshape(-1, 4):
            c._gates.append((int(t), int(tg), int(ctl), int(ps)))
        c._P = int(sum(nat.PARAM_NUM[int(t)] for t, *_ in c._gates))
        return c

    # -- queries -----------------------------------------------------------
    def get_Parameter_Num(self) -> int:
        return self._P

    def get_Qbit_Num(self) -> int:
        return self.qbit_num

    def get_Gate_Num(self) -> int:
        return len(self._gates)

    def get_Parameter_Start_Index(self):
        return [g[3] for g in self._gates]

    def get_Gates(self):
        return [{"type": nat.GATE_NAMES[t], "target_qbit": tg, "control_qbit": c, "parameter_start_idx": ps}
                for t, tg, c, ps in self._gates]

    def gate_array(self) -> np.ndarray:
        return np.asarray(self._gates, dtype=np.int32).reshape(-1, 4)

    # -- execution -----------------------------------------------------------
    def _context(self):
        with self._ctx_lock:
            if self._ctx is None:
                self._ctx = nat.Context(0)
                self._ctx_dirty = True
            if self._ctx_dirty:
                self._ctx.set_circuit(self.qbit_num, self.gate_array())
                self._ctx_dirty = False
            return self._ctx

    def _params(self, parameters_mtx, is_f32):
        p = np.asarray(parameters_mtx)
        want = np.float32 if is_f32 else np.float64
        if p.dtype != want:
            raise TypeError(f"Parameter vector should be {np.dtype(want).name} when is_f32={bool(is_f32)}")
        p = np.ascontiguousarray(p).reshape(-1)
        if p.size < self._P:
            raise Exception(f"not enough parameters: {p.size} given, {self._P} needed")
        return p

    def apply_to(self, parameters_mtx, unitary_mtx, parallel=1, is_f32=False):
        """In-place C . unitary_mtx (Gates_block::apply_to)."""
        if not isinstance(unitary_mtx, np.ndarray) or not unitary_mtx.flags.c_contiguous:
            raise Exception("input matrix is not memory contiguous")
        want = np.complex64 if is_f32 else np.complex128
        if unitary_mtx.dtype != want:
            raise TypeError(f"input matrix or state should be {np.dtype(want).name} when is_f32={bool(is_f32)}")
        p = self._params(parameters_mtx, is_f32)
        M = unitary_mtx if unitary_mtx.ndim == 2 else unitary_mtx.reshape(unitary_mtx.shape[0], -1)
        if M.shape[0] != 1 << self.qbit_num:
            raise Exception("Gates_block::apply_to: Wrong input size in Gates_block gate apply.")
        ctx = self._context()
        ctx.check(ctx._L.sqh_apply_to(ctx.h, p.ctypes.data, M.ctypes.data, M.shape[0], M.shape[1], M.shape[1],
                                      1 if is_f32 else 0))
        return 0

    def get_Matrix(self, parameters_mtx, is_f32=False):
        d = 1 << self.qbit_num
        M = np.eye(d, dtype=np.complex64 if is_f32 else np.complex128)
        self.apply_to(parameters_mtx, M, is_f32=is_f32)
        return M

    def apply_derivate_to(self, parameters_mtx, unitary_mtx, parallel=1, is_f32=False):
        """[dC/dp_k . unitary_mtx for k < P] (Gates_block::apply_derivate_to)."""
        if is_f32:
            raise TypeError("apply_derivate_to is implemented for float64 parameters only")
        p = self._params(parameters_mtx, False)
        M = np.ascontiguousarray(unitary_mtx)
        if M.dtype != np.complex128:
            raise TypeError("input matrix or state should be complex128 when is_f32=False")
        if M.ndim == 1:
            M = M.reshape(-1, 1)
        if M.shape[0] != 1 << self.qbit_num:
            raise Exception("Gates_block::apply_derivate_to: Wrong input size.")
        if self._P == 0:
            return []
        out = np.empty((self._P,) + M.shape, dtype=np.complex128)
        ctx = self._context()
        ctx.check(ctx._L.sqh_apply_derivate_to(ctx.h, p.ctypes.data, M.ctypes.data, M.shape[0], M.shape[1],
                                               M.shape[1], out.ctypes.data))
        shape = unitary_mtx.shape
        return [out[k].reshape(shape) for k in range(self._P)]

    def apply_to_combined(self, parameters_mtx, unitary_mtx, parallel=1, is_f32=False):
        """[C . U, dC/dp_0 . U, ...] (Gates_block::apply_to_combined)."""
        fwd = np.array(unitary_mtx, dtype=np.complex128, copy=True, order="C")
        derivs = self.apply_derivate_to(parameters_mtx, unitary_mtx, parallel)
        self.apply_to(parameters_mtx, fwd, parallel)
        return [fwd] + derivs
