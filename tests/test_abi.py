"""CPU checks of the C-ABI boundary: the library loads, exports every function
include/sqhip.h declares, compiles programs without a device, and the Python mirror
rejects bad input the way the reference does before any device work."""
import ctypes
import os
import re

import numpy as np
import pytest

from squander_hip import Circuit, _native as nat
from squander_hip import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            names += re.findall(r"^\s*(?:int|void|long long|const char\*|void\*)\s+\**(sqh_\w+)\s*\(", txt, re.M)
    return sorted(set(names))


def test_header_declares_the_python_binding_surface():
    decl = declared_functions()
    assert set(decl) == set(nat.EXPORTS), set(decl) ^ set(nat.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(nat.LIB_PATH)
    for name in declared_functions():
        assert hasattr(L, name), name


def test_version_string():
    assert b"gfx950" in nat.lib().sqh_version()


def test_gate_type_ids_match_header():
    txt = open(os.path.join(ROOT, "include", "sqhip.h")).read()
    ids = dict((k, int(v)) for k, v in re.findall(r"#define SQH_GATE_(\w+) (\d+)", txt))
    assert ids == {"CZ": nat.GATE_CZ, "CNOT": nat.GATE_CNOT, "U3": nat.GATE_U3, "RY": nat.GATE_RY,
                   "RX": nat.GATE_RX, "RZ": nat.GATE_RZ, "CRY": nat.GATE_CRY}


def test_program_compiler_rejects_bad_gates():
    with pytest.raises(nat.SqhError):
        nat.inspect_program(3, [(nat.GATE_U3, 3, -1, 0)])          # target out of range
    with pytest.raises(nat.SqhError):
        nat.inspect_program(3, [(nat.GATE_CNOT, 1, 1, 0)])         # control == target
    with pytest.raises(nat.SqhError):
        nat.inspect_program(3, [(99, 0, -1, 0)])                    # unknown gate
    with pytest.raises(nat.SqhError):
        nat.inspect_program(3, [(nat.GATE_U3, 0, -1, 5)])          # parameters out of range


def test_program_compiler_rejects_overlapping_parameter_ranges():
    """Parameter ranges must tile [0, P) as Gates_block assigns them
    (Gates_block.cpp:2500-2525): a shared index is rejected."""
    with pytest.raises(nat.SqhError):   # RY reads index 2, U3 reads 0..2: 4 parameters, index 3 unread
        nat.inspect_program(2, [(nat.GATE_U3, 0, -1, 0), (nat.GATE_RY, 1, -1, 2)])
    with pytest.raises(nat.SqhError):
        nat.inspect_program(2, [(nat.GATE_RX, 0, -1, 0), (nat.GATE_RY, 1, -1, 0)])
    ops, _ = nat.inspect_program(2, [(nat.GATE_RY, 1, -1, 3), (nat.GATE_U3, 0, -1, 0)])  # any order is fine
    assert len(ops) > 0


def test_program_for_small_circuits_has_no_remaps():
    ops, lays = nat.inspect_program(4, W.ansatz_4q_12())
    assert len(lays) == 1 and len(ops) == len(W.ansatz_4q_12())


def test_circuit_mirror_validates_before_device_work():
    c = Circuit(3)
    with pytest.raises(Exception):
        c.add_U3(3)
    with pytest.raises(Exception):
        c.add_CNOT(0, 0)
    c.add_U3(0)
    c.add_CNOT(1, 0)
    c.add_CRY(2, 1)
    assert c.get_Parameter_Num() == 4 and c.get_Parameter_Start_Index() == [0, 0, 3]
    with pytest.raises(TypeError):
        c.apply_to(np.zeros(4, dtype=np.float32), np.eye(8, dtype=np.complex128))
    with pytest.raises(TypeError):
        c.apply_to(np.zeros(4), np.eye(8, dtype=np.complex64))
    with pytest.raises(TypeError):
        c.apply_derivate_to(np.zeros(4, dtype=np.float32), np.eye(8, dtype=np.complex128), is_f32=True)


def test_sub_circuit_parameters_follow_in_order():
    a = Circuit(2)
    a.add_U3(0)
    b = Circuit(2)
    b.add_RY(1)
    b.add_CNOT(0, 1)
    b.add_RZ(0)
    a.add_Circuit(b)
    assert a.get_Parameter_Start_Index() == [0, 3, 0, 4] and a.get_Parameter_Num() == 5


def test_no_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(nat.SqhError):
        nat.Context(0)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    """Without the built HIP library the package raises instead of computing anything
    another way (there is no CPU fallback)."""
    monkeypatch.setattr(nat, "LIB_PATH", str(tmp_path / "libsqhip.so"))
    monkeypatch.setattr(nat, "_lib", None)
    with pytest.raises(ImportError, match="not built"):
        nat.lib()
    with pytest.raises(ImportError):
        Circuit.from_gate_array(2, W.ansatz_layers(2, [0])).get_Matrix(np.zeros(W.param_count(W.ansatz_layers(2, [0]))))


def test_initial_guess_matches_oracle_mt19937(oracle):
    from squander_hip import mt19937_uniform_2pi
    for seed in (1, 42, 2026):
        assert np.array_equal(mt19937_uniform_2pi(seed, 50), oracle.mt_stream(seed, 50, kind="uniform"))
