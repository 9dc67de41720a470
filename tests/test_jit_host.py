"""Run-time specialised kernels, host side (no GPU): the generated straight-line
source compiles for gfx950 for every gate kind, both control placements (register
and lane bits), all three kernel kinds and both precisions."""
import pytest

from squander_hip import _native as nat
from squander_hip import workloads as W


@pytest.mark.parametrize("n,kind,f32", [(3, 1, False), (6, 1, False), (6, 0, True), (3, 2, False), (6, 2, True),
                                        (3, 0, False)])
def test_generated_kernel_compiles(n, kind, f32):
    gates = W.mixed_gate_circuit(n, 18, 40 + n)
    ok, log, code, src = nat.jit_compile_check(n, gates, kind=kind, is_f32=f32, uniform=n > 4)
    assert ok, log
    assert code > 0 and src > 0


def test_compile_check_rejects_bad_circuit():
    ok, log, _, _ = nat.jit_compile_check(2, [[nat.GATE_U3, 3, -1, 0]], kind=1)
    assert not ok and log


def test_persistent_solver_kernel_compiles():
    """The persistent solver with the evaluation specialised to the circuit (kind 3,
    up to 4 qubits) compiles for the 4-qubit ansatz; above 4 qubits it is refused."""
    ok, log, code, src = nat.jit_compile_check(4, W.ansatz_4q_12(), kind=3)
    assert ok, log
    assert code > 0 and src > 0
    ok, _, _, _ = nat.jit_compile_check(5, W.ansatz_layers(5, [0, 1]), kind=3)
    assert not ok



def test_jit_source_is_the_compiled_source():
    """sqh_jit_source returns the generated source whose length sqh_jit_compile_check
    reports: straight-line code naming every parametric slot's kernel, the kernel entry
    point, and the circuit's constants."""
    gates = W.mixed_gate_circuit(5, 20, 77)
    ok, log, code, nbytes = nat.jit_compile_check(5, gates, kind=1, uniform=True)
    assert ok, log
    src = nat.jit_source(5, gates, kind=1, uniform=True)
    assert len(src) == nbytes
    assert "sqh_jit" in src and "constexpr int N = 5;" in src
    solver = nat.jit_source(4, W.ansatz_4q_12(), kind=3)
    assert "sqh_jit_kernel" in solver and "solve_runs" in solver
    with pytest.raises(nat.SqhError):
        nat.jit_source(5, W.ansatz_layers(5, [0, 1]), kind=3)
