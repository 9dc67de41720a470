"""Gate-list binary format (Gates_block.cpp:4807-4970): byte layout known answer,
round trip, nested blocks flattened on import."""
import struct

import numpy as np
import pytest

from squander_hip import Circuit
from squander_hip.binary_io import export_gate_list_to_binary, import_gate_list_from_binary


def test_byte_layout_known_answer(tmp_path):
    c = Circuit(2)
    c.add_U3(0)
    c.add_CNOT(1, 0)
    c.add_CRY(0, 1)
    p = np.array([0.1, 0.2, 0.3, 0.4])
    fn = tmp_path / "c.bin"
    export_gate_list_to_binary(p, c, str(fn))
    want = (struct.pack("<iii", 2, 4, 3) + struct.pack("<ii", 7, 0) + struct.pack("<ddd", 0.1, 0.2, 0.3) +
            struct.pack("<iii", 5, 1, 0) + struct.pack("<iii", 14, 0, 1) + struct.pack("<d", 0.4))
    assert fn.read_bytes() == want


def test_round_trip_and_nested_block(tmp_path):
    c = Circuit(3)
    for q in range(3):
        c.add_U3(q)
    c.add_CNOT(0, 1)
    c.add_RX(2)
    c.add_CZ(2, 0)
    c.add_RZ(1)
    c.add_RY(0)
    p = np.random.default_rng(0).uniform(0, 6.28, c.get_Parameter_Num())
    fn = tmp_path / "c.bin"
    export_gate_list_to_binary(p, c, str(fn))
    c2, p2 = import_gate_list_from_binary(str(fn))
    assert c2._gates == c._gates and np.array_equal(p2, p)
    # a file with a nested block (type 16) imports flattened
    inner = struct.pack("<iii", 3, 1, 1) + struct.pack("<ii", 9, 2) + struct.pack("<d", 0.5)
    data = struct.pack("<iii", 3, 4, 2) + struct.pack("<ii", 7, 1) + struct.pack("<ddd", 1, 2, 3) + struct.pack("<i", 16) + inner
    fn2 = tmp_path / "n.bin"
    fn2.write_bytes(data)
    c3, p3 = import_gate_list_from_binary(str(fn2))
    assert [g[0] for g in c3._gates] == [7, 9] and np.array_equal(p3, [1, 2, 3, 0.5])


def test_truncated_file_raises(tmp_path):
    fn = tmp_path / "t.bin"
    fn.write_bytes(struct.pack("<iii", 2, 3, 1) + struct.pack("<ii", 7, 0))
    with pytest.raises(Exception):
        import_gate_list_from_binary(str(fn))
