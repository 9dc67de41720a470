"""Gate-list binary format of the reference (export_gate_list_to_binary /
import_gate_list_from_binary, Gates_block.cpp:4807-4970), for the seven gates of
the path.

Layout (native little-endian): int32 qbit_num, int32 parameter_num, int32 gates_num,
then per gate an int32 gate type followed by
  CNOT, CZ:          int32 target, int32 control
  U3, RX, RY, RZ:    int32 target, float64 x parameter_num
  CRY:               int32 target, int32 control, float64 x 1
  BLOCK (16):        a nested list with its own header (imported flattened).
Parameters are written in gate order (each gate's parameters follow those of the
gates before it, as Gates_block::add_gate numbers them).
"""
from __future__ import annotations

import struct

import numpy as np

from . import _native as nat
from .circuit import Circuit

BLOCK_OPERATION = 16


def export_gate_list_to_binary(parameters, circuit: Circuit, filename: str):
    p = np.ascontiguousarray(parameters, dtype=np.float64).reshape(-1)
    if p.size < circuit.get_Parameter_Num():
        raise Exception("export_gate_list_to_binary: not enough parameters")
    with open(filename, "wb") as f:
        f.write(struct.pack("<iii", circuit.qbit_num, circuit.get_Parameter_Num(), circuit.get_Gate_Num()))
        for t, tg, c, ps in circuit._gates:
            f.write(struct.pack("<i", t))
            npar = nat.PARAM_NUM[t]
            if t in (nat.GATE_CNOT, nat.GATE_CZ):
                f.write(struct.pack("<ii", tg, c))
            elif t in (nat.GATE_U3, nat.GATE_RX, nat.GATE_RY, nat.GATE_RZ):
                f.write(struct.pack("<i", tg))
                f.write(p[ps:ps + npar].astype("<f8").tobytes())
            elif t == nat.GATE_CRY:
                f.write(struct.pack("<ii", tg, c))
                f.write(p[ps:ps + npar].astype("<f8").tobytes())
            else:
                raise Exception("export_gate_list_to_binary: unimplemented gate")


def _read(f, fmt):
    n = struct.calcsize(fmt)
    b = f.read(n)
    if len(b) != n:
        raise Exception("import_gate_list_from_binary: truncated file")
    return struct.unpack(fmt, b)


def _import_block(f, circuit: Circuit, params: list):
    qbit_num, parameter_num, gates_num = _read(f, "<iii")
    if qbit_num > circuit.qbit_num:
        raise Exception("import_gate_list_from_binary: nested block acts on more qubits")
    for _ in range(gates_num):
        (t,) = _read(f, "<i")
        if t in (nat.GATE_CNOT, nat.GATE_CZ):
            tg, c = _read(f, "<ii")
            circuit._add(t, tg, c)
        elif t in (nat.GATE_U3, nat.GATE_RX, nat.GATE_RY, nat.GATE_RZ):
            (tg,) = _read(f, "<i")
            circuit._add(t, tg)
            params.extend(_read(f, "<" + "d" * nat.PARAM_NUM[t]))
        elif t == nat.GATE_CRY:
            tg, c = _read(f, "<ii")
            circuit._add(t, tg, c)
            params.extend(_read(f, "<d"))
        elif t == BLOCK_OPERATION:
            _import_block(f, circuit, params)
        else:
            raise Exception(f"import_gate_list_from_binary: unimplemented gate type {t}")
    return qbit_num, parameter_num


def import_gate_list_from_binary(filename: str):
    """Returns (Circuit, parameters) with nested blocks flattened."""
    with open(filename, "rb") as f:
        head = f.read(4)
        if len(head) != 4:
            raise Exception("import_gate_list_from_binary: truncated file")
        (qbit_num,) = struct.unpack("<i", head)
        f.seek(0)
        circ = Circuit(qbit_num)
        params = []
        _, parameter_num = _import_block(f, circ, params)
    if len(params) != parameter_num or circ.get_Parameter_Num() != parameter_num:
        raise Exception("import_gate_list_from_binary: parameter count mismatch")
    return circ, np.array(params, dtype=np.float64)
