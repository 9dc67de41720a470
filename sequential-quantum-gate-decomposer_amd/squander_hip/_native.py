"""ctypes binding of libsqhip.so (the C ABI declared in include/sqhip.h).

The library is the product: there is no CPU fallback.  Loading fails loudly when
libsqhip.so has not been built, and every compute call fails loudly when no GPU
is present.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsqhip.so")

SQH_OK = 0
SQH_ERR_ARG, SQH_ERR_GATE, SQH_ERR_STATE, SQH_ERR_DEVICE, SQH_ERR_NOMEM = -1, -2, -3, -4, -5

# Gate-type ids: the reference's gate_type enum (Gate.h:39-51).
GATE_CZ, GATE_CNOT, GATE_U3, GATE_RY, GATE_RX, GATE_RZ, GATE_CRY = 4, 5, 7, 8, 9, 10, 14
PARAM_NUM = {GATE_U3: 3, GATE_RX: 1, GATE_RY: 1, GATE_RZ: 1, GATE_CRY: 1, GATE_CNOT: 0, GATE_CZ: 0}
GATE_NAMES = {GATE_U3: "U3", GATE_RX: "RX", GATE_RY: "RY", GATE_RZ: "RZ", GATE_CRY: "CRY",
              GATE_CNOT: "CNOT", GATE_CZ: "CZ"}

# Cost variants: the reference's cost_function_type values + ABS_TRACE.
COST_FROBENIUS, COST_HILBERT_SCHMIDT, COST_INFIDELITY, COST_ABS_TRACE = 0, 3, 9, 100
COST_FROBENIUS_CORRECTION1, COST_FROBENIUS_CORRECTION2 = 1, 2
COST_HILBERT_SCHMIDT_CORRECTION1, COST_HILBERT_SCHMIDT_CORRECTION2 = 4, 5
PREC_F64, PREC_F32 = 0, 1
FUSED_COST_OFF, FUSED_COST_ON, FUSED_COST_AUTO = 0, 1, 2

# Exported symbols (checked by the CPU test suite against include/sqhip.h).
EXPORTS = [
    "sqh_create", "sqh_destroy", "sqh_last_error", "sqh_get_stream", "sqh_synchronize", "sqh_version",
    "sqh_set_circuit", "sqh_get_parameter_num", "sqh_set_target", "sqh_set_trace_offset",
    "sqh_set_cost_function", "sqh_set_correction_weights", "sqh_apply_to", "sqh_apply_to_dev", "sqh_apply_to_stream_dev",
    "sqh_apply_derivate_to", "sqh_apply_derivate_to_dev", "sqh_optimization_problem_batch",
    "sqh_optimization_problem_batch_dev", "sqh_optimization_problem_combined_batch",
    "sqh_optimization_problem_combined_batch_dev", "sqh_optimization_problem_combined_unitary",
    "sqh_optimization_problem_combined_batch_noalloc_dev", "sqh_inspect_program", "sqh_inspect_slots", "sqh_inspect_fusion", "sqh_set_kernel_timing", "sqh_get_kernel_timing", "sqh_decompose_batch",
    "sqh_mlp_create", "sqh_mlp_destroy", "sqh_mlp_param_count", "sqh_mlp_get_params", "sqh_mlp_set_params",
    "sqh_mlp_forward_dev", "sqh_mlp_train_step_dev", "sqh_mlp_gradients", "sqh_apply_to_fused_dev",
    "sqh_apply_large_kernel_dev", "sqh_decompose_problems", "sqh_agents", "sqh_agents_batch",
    "sqh_set_autotune", "sqh_set_block_config", "sqh_get_block_config",
    "sqh_set_jit", "sqh_get_jit_stats", "sqh_jit_compile_check", "sqh_jit_source", "sqh_set_fused_cost",
    "sqh_set_fused_cost_config", "sqh_get_fused_cost_config", "sqh_set_fused_cost_chain", "sqh_set_fused_grad_config",
    "sqh_set_tile_config", "sqh_get_tile_config",
    "sqh_optimization_problem_combined_targets_dev", "sqh_mlp_backward_step_dev",
]


class DecomposeOptions(ctypes.Structure):
    """sqh_decompose_options (include/sqhip.h)."""
    _fields_ = [("max_inner_iterations", ctypes.c_longlong), ("max_outer_iterations", ctypes.c_int),
                ("optimization_tolerance", ctypes.c_double), ("convergence_threshold", ctypes.c_double),
                ("mixed_precision", ctypes.c_int), ("switch_cost", ctypes.c_double),
                ("rounds_per_graph", ctypes.c_int), ("max_rounds", ctypes.c_longlong),
                ("iteration_loops", ctypes.c_int), ("restart_jitter_count", ctypes.c_int),
                ("restart_jitter", ctypes.c_void_p), ("engine", ctypes.c_int)]


class SqhError(Exception):
    """A failed library call (the reference raises a plain Exception with the message)."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class _Gate(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("target", ctypes.c_int32),
                ("control", ctypes.c_int32), ("param_start", ctypes.c_int32)]


class AgentsOptions(ctypes.Structure):
    """sqh_agents_options (include/sqhip.h)."""
    _fields_ = [("agent_num", ctypes.c_int), ("max_iterations", ctypes.c_longlong),
                ("optimization_tolerance", ctypes.c_double), ("agent_lifetime", ctypes.c_int),
                ("exploration_rate", ctypes.c_double), ("agent_randomization_rate", ctypes.c_double),
                ("radius", ctypes.c_double), ("randomization_rate", ctypes.c_double),
                ("convergence_length", ctypes.c_int)]


class Problem(ctypes.Structure):
    _fields_ = [("qbit_num", ctypes.c_int), ("n_gates", ctypes.c_int), ("gates", ctypes.POINTER(_Gate)),
                ("target", ctypes.c_void_p), ("x0", ctypes.c_void_p)]


_lib = None
_lock = threading.Lock()


def lib():
    """Load libsqhip.so once; raise if it is missing."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                              "(the HIP library has no CPU fallback)")
        # PyTorch-ROCm bundles its own libamdhip64 (soname libamdhip64.so.7) and its
        # HIP library names it "libamdhip64.so", which does not match a runtime already
        # loaded from /opt/rocm.  Loading torch first lets this library's
        # "libamdhip64.so.7" dependency bind to torch's copy, so the process holds one
        # HIP/HSA runtime and device pointers from either side are valid in both.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
        ll = ctypes.c_longlong
        sig = {
            "sqh_create": ([i, ctypes.POINTER(vp)], i),
            "sqh_destroy": ([vp], None),
            "sqh_last_error": ([vp], ctypes.c_char_p),
            "sqh_get_stream": ([vp], vp),
            "sqh_synchronize": ([vp], i),
            "sqh_version": ([], ctypes.c_char_p),
            "sqh_set_circuit": ([vp, i, ctypes.POINTER(_Gate), i], i),
            "sqh_get_parameter_num": ([vp], i),
            "sqh_set_target": ([vp, vp, i, i, i], i),
            "sqh_set_trace_offset": ([vp, i], i),
            "sqh_set_cost_function": ([vp, i], i),
            "sqh_set_correction_weights": ([vp, ctypes.c_double, ctypes.c_double, ctypes.c_double], i),
            "sqh_apply_to": ([vp, vp, vp, i, i, i, i], i),
            "sqh_apply_to_dev": ([vp, vp, i, vp, i, i, i, ll, i, i], i),
            "sqh_apply_to_stream_dev": ([vp, vp, i, vp, i, i, i, ll, i, i], i),
            "sqh_apply_to_fused_dev": ([vp, vp, i, vp, i, i, i, ll, i, i], i),
            "sqh_apply_large_kernel_dev": ([vp, vp, vp, i, vp, i, i, i, ll, i, i], i),
            "sqh_apply_derivate_to": ([vp, vp, vp, i, i, i, vp], i),
            "sqh_apply_derivate_to_dev": ([vp, vp, vp, i, i, i, vp], i),
            "sqh_optimization_problem_batch": ([vp, vp, i, i, vp], i),
            "sqh_optimization_problem_batch_dev": ([vp, vp, i, i, vp], i),
            "sqh_optimization_problem_combined_batch": ([vp, vp, i, i, vp, vp], i),
            "sqh_optimization_problem_combined_batch_dev": ([vp, vp, i, i, vp, vp], i),
            "sqh_optimization_problem_combined_unitary": ([vp, vp, vp, vp], i),
            "sqh_optimization_problem_combined_batch_noalloc_dev": ([vp, vp, i, i, vp, vp], i),
            "sqh_set_kernel_timing": ([vp, i], i),
            "sqh_mlp_create": ([vp, i, i, i, i, ctypes.c_ulonglong, ctypes.POINTER(vp)], i),
            "sqh_mlp_destroy": ([vp], None),
            "sqh_mlp_param_count": ([vp], ll),
            "sqh_mlp_get_params": ([vp, vp], i),
            "sqh_mlp_set_params": ([vp, vp], i),
            "sqh_mlp_forward_dev": ([vp, vp, i, vp], i),
            "sqh_mlp_train_step_dev": ([vp, vp, vp, i, ctypes.c_float, d], i),
            "sqh_mlp_gradients": ([vp, vp, vp, i, vp, d], i),
            "sqh_decompose_batch": ([vp, vp, i, i, i, vp, i, vp, ctypes.POINTER(DecomposeOptions), vp, vp, vp, vp,
                                     ctypes.POINTER(ll)], i),
            "sqh_decompose_problems": ([vp, ctypes.POINTER(Problem), i, ctypes.POINTER(DecomposeOptions), vp, vp, vp,
                                        vp, ctypes.POINTER(ll)], i),
            "sqh_agents": ([vp, vp, ctypes.POINTER(AgentsOptions), ctypes.c_uint, d, ctypes.POINTER(ll),
                            ctypes.POINTER(ll)], i),
            "sqh_agents_batch": ([vp, vp, i, i, i, vp, ctypes.POINTER(AgentsOptions), vp, vp, vp, vp], i),
            "sqh_set_autotune": ([vp, i], i),
            "sqh_set_block_config": ([vp, i, i], i),
            "sqh_get_block_config": ([vp, i, i, i, i, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                      d], i),
            "sqh_set_jit": ([vp, i], i),
            "sqh_set_fused_cost": ([vp, i], i),
            "sqh_set_fused_cost_config": ([vp, i, i, i, i], i),
            "sqh_set_fused_cost_chain": ([vp, i], i),
            "sqh_set_fused_grad_config": ([vp, i, i], i),
            "sqh_set_tile_config": ([vp, i, i], i),
            "sqh_get_tile_config": ([vp, i, i, i, ctypes.POINTER(i), ctypes.POINTER(i)], i),
            "sqh_get_fused_cost_config": ([vp, i, i, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i),
                                           ctypes.POINTER(i)], i),
            "sqh_optimization_problem_combined_targets_dev": ([vp, vp, i, vp, vp, i, i, vp, vp], i),
            "sqh_mlp_backward_step_dev": ([vp, vp, vp, i, ctypes.c_float], i),
            "sqh_jit_compile_check": ([i, ctypes.POINTER(_Gate), i, i, i, i, ctypes.c_char_p, i,
                                       ctypes.POINTER(ll), ctypes.POINTER(ll)], i),
            "sqh_jit_source": ([i, ctypes.POINTER(_Gate), i, i, i, i, ctypes.c_char_p, ll, ctypes.POINTER(ll)], i),
            "sqh_get_jit_stats": ([vp, d, ctypes.POINTER(ctypes.c_int)], i),
            "sqh_get_kernel_timing": ([vp, d, ctypes.POINTER(ctypes.c_int)], i),
            "sqh_inspect_slots": ([i, ctypes.POINTER(_Gate), i, vp, i, ctypes.POINTER(ctypes.c_int)], i),
            "sqh_inspect_fusion": ([i, ctypes.POINTER(_Gate), i, vp, vp, i, vp, i, ctypes.POINTER(ctypes.c_int)], i),
            "sqh_inspect_program": ([i, ctypes.POINTER(_Gate), i, vp, i, vp, i, ctypes.POINTER(ctypes.c_int),
                                     ctypes.POINTER(ctypes.c_int)], i),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
        return L


def gate_array(gates):
    g = np.ascontiguousarray(np.asarray(gates, dtype=np.int32).reshape(-1, 4))
    arr = (_Gate * max(len(g), 1))()
    for k, (t, tg, c, ps) in enumerate(g):
        arr[k].type, arr[k].target, arr[k].control, arr[k].param_start = int(t), int(tg), int(c), int(ps)
    return arr, len(g)


def inspect_program(qbit_num, gates):
    """The register-tile program the library compiles for a gate list:
    (ops[n,8] uint8 records, layouts[n,96] uint8 records)."""
    L = lib()
    garr, ng = gate_array(gates)
    nops, nl = ctypes.c_int(0), ctypes.c_int(0)
    rc = L.sqh_inspect_program(qbit_num, garr, ng, None, 0, None, 0, ctypes.byref(nops), ctypes.byref(nl))
    if rc:
        raise SqhError(rc, "program compilation failed")
    ops = np.zeros((max(nops.value, 1), 8), dtype=np.uint8)
    lays = np.zeros((max(nl.value, 1), 96), dtype=np.uint8)
    rc = L.sqh_inspect_program(qbit_num, garr, ng, ops.ctypes.data, nops.value, lays.ctypes.data, nl.value,
                               ctypes.byref(nops), ctypes.byref(nl))
    if rc:
        raise SqhError(rc, "program compilation failed")
    return ops[:nops.value], lays[:nl.value]


def inspect_slots(qbit_num, gates):
    """The program's parametric-gate records in slot order: int32 array [n, 4] of
    (gate type, first parameter, parameter count, op index)."""
    L = lib()
    garr, ng = gate_array(gates)
    n = ctypes.c_int(0)
    rc = L.sqh_inspect_slots(qbit_num, garr, ng, None, 0, ctypes.byref(n))
    if rc:
        raise SqhError(rc, "program compilation failed")
    sl = np.zeros((max(n.value, 1), 4), dtype=np.int32)
    rc = L.sqh_inspect_slots(qbit_num, garr, ng, sl.ctypes.data, n.value, ctypes.byref(n))
    if rc:
        raise SqhError(rc, "program compilation failed")
    return sl[:n.value]


def inspect_fusion(qbit_num, gates):
    """The streaming path's dense-block plan: a list of (sorted qubits, gate-list
    indices in run order), one entry per block."""
    L = lib()
    garr, ng = gate_array(gates)
    nb = ctypes.c_int(0)
    rc = L.sqh_inspect_fusion(qbit_num, garr, ng, None, None, 0, None, 0, ctypes.byref(nb))
    if rc:
        raise SqhError(rc, "program compilation failed")
    first = np.zeros(nb.value + 1, dtype=np.int32)
    qs = np.zeros((max(nb.value, 1), 4), dtype=np.int32)
    src = np.zeros(max(ng, 1), dtype=np.int32)
    rc = L.sqh_inspect_fusion(qbit_num, garr, ng, first.ctypes.data, qs.ctypes.data, nb.value, src.ctypes.data, ng,
                              ctypes.byref(nb))
    if rc:
        raise SqhError(rc, "program compilation failed")
    return [(tuple(int(q) for q in qs[b] if q >= 0), [int(x) for x in src[first[b]:first[b + 1]]])
            for b in range(nb.value)]


class Context:
    """One sqh_ctx: a device, its stream, and the circuit/target state."""

    def __init__(self, device: int = 0):
        self._L = lib()
        h = ctypes.c_void_p()
        rc = self._L.sqh_create(device, ctypes.byref(h))
        if rc != SQH_OK:
            raise SqhError(rc, f"sqh_create(device={device}) failed with status {rc}: no usable GPU")
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            self._L.sqh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        if rc != SQH_OK:
            msg = self._L.sqh_last_error(self.h)
            raise SqhError(rc, msg.decode() if msg else f"status {rc}")

    @property
    def stream(self) -> int:
        return self._L.sqh_get_stream(self.h) or 0

    def synchronize(self):
        self.check(self._L.sqh_synchronize(self.h))

    def set_circuit(self, qbit_num, gates):
        garr, ng = gate_array(gates)
        self.check(self._L.sqh_set_circuit(self.h, qbit_num, garr, ng))

    def parameter_num(self) -> int:
        return int(self._L.sqh_get_parameter_num(self.h))

    def set_target(self, U):
        U = np.ascontiguousarray(U, dtype=np.complex128)
        self.check(self._L.sqh_set_target(self.h, U.ctypes.data, U.shape[0], U.shape[1], U.shape[1]))

    def set_trace_offset(self, off):
        self.check(self._L.sqh_set_trace_offset(self.h, int(off)))

    def set_cost_function(self, variant):
        self.check(self._L.sqh_set_cost_function(self.h, int(variant)))

    def set_correction_weights(self, prev_cost, scale1, scale2):
        self.check(self._L.sqh_set_correction_weights(self.h, float(prev_cost), float(scale1), float(scale2)))

    def set_jit(self, enable=True):
        """Use run-time specialised kernels for this context's circuit."""
        self.check(self._L.sqh_set_jit(self.h, int(bool(enable))))

    def set_fused_cost(self, mode=True):
        """Batched evaluations through the fused 4-qubit blocks on the matrix cores
        (4-10 qubits, costs without correction channels): True / False, or "auto" (the
        default: costs from 7 qubits, or from 5 without specialised kernels; costs with
        gradients from 6 qubits)."""
        m = FUSED_COST_AUTO if mode == "auto" else (FUSED_COST_ON if mode else FUSED_COST_OFF)
        self.check(self._L.sqh_set_fused_cost(self.h, m))

    def set_fused_cost_config(self, tiles=0, groups=0, waves=0, prefetch=0):
        """Force the fused-block cost kernel's launch shape (all 0: run-time tuner)."""
        self.check(self._L.sqh_set_fused_cost_config(self.h, int(tiles), int(groups), int(waves), int(prefetch)))

    def set_fused_cost_chain(self, on=True):
        """Fused cost pass as the transposed chain from unit vectors (True, the default)
        or as the target's columns run forward through the blocks (False)."""
        self.check(self._L.sqh_set_fused_cost_chain(self.h, int(bool(on))))

    def set_fused_grad_config(self, tiles=0, waves=0):
        """Force the fused-block cost + gradient kernel's launch shape (both 0: the rule)."""
        self.check(self._L.sqh_set_fused_grad_config(self.h, int(tiles), int(waves)))

    def set_tile_config(self, waves_per_group=0, stage_program=-1):
        """Force the register-tile kernels' launch options (0, -1: run-time tuner)."""
        self.check(self._L.sqh_set_tile_config(self.h, int(waves_per_group), int(stage_program)))

    def tile_config(self, kind, f32, B):
        """(wavefronts per workgroup or -1 packed, program staged) in effect for kind 1
        (forward trace entries) or 2 (gradient)."""
        w, st = ctypes.c_int(0), ctypes.c_int(0)
        self.check(self._L.sqh_get_tile_config(self.h, int(kind), int(bool(f32)), int(B), ctypes.byref(w),
                                                ctypes.byref(st)))
        return w.value, st.value

    def fused_cost_config(self, f32, B):
        """(tiles per wavefront, exchange groups, waves per workgroup, prefetch) in use."""
        t, g, w, p = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        self.check(self._L.sqh_get_fused_cost_config(self.h, int(bool(f32)), int(B), ctypes.byref(t), ctypes.byref(g),
                                                      ctypes.byref(w), ctypes.byref(p)))
        return t.value, g.value, w.value, p.value

    def jit_stats(self):
        ms, n = ctypes.c_double(0.0), ctypes.c_int(0)
        self.check(self._L.sqh_get_jit_stats(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def set_autotune(self, enable=True):
        self.check(self._L.sqh_set_autotune(self.h, int(bool(enable))))

    def set_block_config(self, tiles_per_wave=0, threads=0):
        """Force the 4-qubit block kernel's launch shape; (0, 0) returns to automatic."""
        self.check(self._L.sqh_set_block_config(self.h, int(tiles_per_wave), int(threads)))

    def block_config(self, rows, cols, batch, f32=False):
        """(tiles per wavefront, threads per workgroup, tuning ms so far) for a call shape."""
        t, th, ms = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_double(0.0)
        self.check(self._L.sqh_get_block_config(self.h, int(bool(f32)), int(rows), int(cols), int(batch),
                                                ctypes.byref(t), ctypes.byref(th), ctypes.byref(ms)))
        return t.value, th.value, ms.value

    def set_kernel_timing(self, enable=True):
        self.check(self._L.sqh_set_kernel_timing(self.h, 1 if enable else 0))

    def kernel_timing(self):
        """(total kernel milliseconds, launches) since the last query; resets the timers."""
        ms = ctypes.c_double(0.0)
        n = ctypes.c_int(0)
        self.check(self._L.sqh_get_kernel_timing(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value


def jit_source(qbit_num, gates, kind=1, is_f32=False, uniform=True):
    """The generated source of the run-time specialised kernel of a gate list (no GPU
    needed)."""
    L = lib()
    garr, ng = gate_array(gates)
    sb = ctypes.c_longlong(0)
    rc = L.sqh_jit_source(qbit_num, garr, ng, kind, int(is_f32), int(uniform), None, 0, ctypes.byref(sb))
    if rc != 0:
        raise SqhError(rc, "sqh_jit_source failed")
    buf = ctypes.create_string_buffer(sb.value + 1)
    L.sqh_jit_source(qbit_num, garr, ng, kind, int(is_f32), int(uniform), buf, sb.value + 1, ctypes.byref(sb))
    return buf.value.decode()


def jit_compile_check(qbit_num, gates, kind=1, is_f32=False, uniform=True):
    """Generate and compile (for gfx950, no GPU needed) the run-time specialised kernel
    of a gate list; returns (ok, log, code bytes, source bytes)."""
    L = lib()
    garr, ng = gate_array(gates)
    buf = ctypes.create_string_buffer(4096)
    cb, sb = ctypes.c_longlong(0), ctypes.c_longlong(0)
    rc = L.sqh_jit_compile_check(qbit_num, garr, ng, kind, int(is_f32), int(uniform), buf, 4096, ctypes.byref(cb),
                                 ctypes.byref(sb))
    return rc == 0, buf.value.decode(errors="replace"), cb.value, sb.value
