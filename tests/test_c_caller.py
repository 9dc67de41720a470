"""libsqhip.so from a plain C program (tests/c_caller/sqh_c_caller.c): no Python and
no PyTorch in the process, so the library's HIP dependency binds to the runtime it
was linked against (/opt/rocm), not to the copy PyTorch bundles.  The batched cost
and gradient it returns must equal the oracle's (1e-12)."""
import os
import subprocess

import numpy as np
import pytest

from squander_hip import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "c_caller", "sqh_c_caller")


def test_c_caller_is_built_and_loads_the_library():
    assert os.path.exists(EXE), "build() compiles tests/c_caller"
    r = subprocess.run([EXE, "--version"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("sqhip")


@pytest.mark.gpu
def test_c_caller_cost_and_gradient_match_oracle(oracle, tmp_path):
    n = 5
    gates = W.mixed_gate_circuit(n, 30, 4242)
    P = W.param_count(gates)
    d = 1 << n
    U = W.haar_unitary(d, 4243).conj().T.copy()
    B = 3
    ps = W.random_params(P, 4244, batch=B)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    g = np.ascontiguousarray(np.asarray(gates)[:, :4], dtype=np.int32)
    with open(fin, "wb") as fh:
        fh.write(np.array([n, len(gates), B], dtype=np.int32).tobytes())
        fh.write(g.tobytes())
        fh.write(np.ascontiguousarray(U, dtype=np.complex128).tobytes())
        fh.write(np.ascontiguousarray(ps, dtype=np.float64).tobytes())
    env = {k: v for k, v in os.environ.items() if k not in ("LD_LIBRARY_PATH",)}
    print(f"C caller n={n} gates={len(gates)} B={B}", flush=True)
    r = subprocess.run([EXE, str(fin), str(fout)], capture_output=True, text=True, timeout=120, env=env)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    runtime = [l for l in r.stdout.splitlines() if l.startswith("hip runtime:")]
    assert runtime and "torch" not in runtime[0]
    out = np.fromfile(fout, dtype=np.float64)
    f, grad = out[:B], out[B:].reshape(B, P)
    for b in range(B):
        fo, go = oracle.optimization_problem_combined(gates, U, n, ps[b])
        assert abs(f[b] - fo) < 1e-12
        assert np.max(np.abs(grad[b] - go)) < 1e-12
