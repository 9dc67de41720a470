"""CPU oracle for the gate-application / cost / gradient path.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` as the checker.  The product
library never imports it.

Two independent restatements live here:

* ``Oracle`` -- ctypes binding of ``liboracle.so`` (``sq_oracle.c``), the
  line-by-line restatement of the reference's pair-row kernels, gate loop,
  derivative matrices, cost variants and BFGS driver (each C function cites
  its reference file:line).
* ``brute_*`` -- a dense numpy model built from textbook definitions only:
  every gate is the Kronecker embedding of its 2x2 (or controlled 4x4) matrix,
  rotations come from ``scipy.linalg.expm`` of Pauli generators and U3 from the
  ZYZ identity U3(t,p,l) = e^{i(p+l)/2} RZ(p) RY(t) RZ(l) (Qiskit convention,
  the reference checks against it in tests/gates/test_gates.py:162-209).  It
  shares no formula with ``sq_oracle.c`` and pins it.

Conventions follow the reference: row-major complex128, qubit q = bit q of the
row index, gates applied in list order, raw parameters are half-angles for the
RX/RY/RZ/CRY angle and for U3's theta (U3.cpp:49-51, RX.cpp:24).
"""
from __future__ import annotations

import ctypes
import itertools
import os
import subprocess

import numpy as np
import scipy.linalg

HERE = os.path.dirname(os.path.abspath(__file__))

# Reference gate_type values (Gate.h:39-51).
CZ, CNOT, U3, RY, RX, RZ, CRY = 4, 5, 7, 8, 9, 10, 14
PARAM_NUM = {U3: 3, RX: 1, RY: 1, RZ: 1, CRY: 1, CNOT: 0, CZ: 0}

# Reference cost_function_type values (Optimization_Interface.h:43-45) + ABS_TRACE.
COST_FROBENIUS, COST_HS, COST_INFIDELITY, COST_ABS_TRACE = 0, 3, 9, 100
COST_FROBENIUS_C1, COST_FROBENIUS_C2, COST_HS_C1, COST_HS_C2 = 1, 2, 4, 5
# Optimization_Interface.cpp:74-76: prev_cost_fnv_val, correction1_scale, correction2_scale
DEFAULT_CORRECTION = (1.0, 1.0 / 1.7, 1.0 / 2.0)


class _C128(ctypes.Structure):
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


class _Gate(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("target", ctypes.c_int),
                ("control", ctypes.c_int), ("param_start", ctypes.c_int)]


class _Problem(ctypes.Structure):
    _fields_ = [("gates", ctypes.POINTER(_Gate)), ("ng", ctypes.c_int), ("P", ctypes.c_int),
                ("qbit_num", ctypes.c_int), ("target", ctypes.c_void_p),
                ("rows", ctypes.c_int), ("cols", ctypes.c_int), ("stride", ctypes.c_int),
                ("variant", ctypes.c_int), ("trace_offset", ctypes.c_int),
                ("prev_cost", ctypes.c_double), ("scale1", ctypes.c_double), ("scale2", ctypes.c_double)]


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    so = os.path.join(HERE, "liboracle.so")
    if force or not os.path.exists(so):
        subprocess.run(["make", "-s", "-C", HERE] + (["-B"] if force else []), check=True)
    return so


def gates_array(gates) -> np.ndarray:
    """(G,4) int32 array of (type, target, control, param_start)."""
    return np.ascontiguousarray(np.asarray(gates, dtype=np.int32).reshape(-1, 4))


def param_count(gates) -> int:
    g = gates_array(gates)
    return int(sum(PARAM_NUM[int(t)] for t in g[:, 0]))


class AgentsOpts(ctypes.Structure):
    """or_agents_opts (reference defaults)."""
    _fields_ = [("agent_num", ctypes.c_int), ("max_iterations", ctypes.c_long), ("tolerance", ctypes.c_double),
                ("lifetime", ctypes.c_int), ("exploration_rate", ctypes.c_double),
                ("agent_randomization_rate", ctypes.c_double), ("radius", ctypes.c_double),
                ("randomization_rate", ctypes.c_double), ("convergence_length", ctypes.c_int)]


def agents_opts(**kw):
    o = AgentsOpts()
    o.agent_num = int(kw.get("agent_num", 64))
    o.max_iterations = int(kw.get("max_iterations", 10000))
    o.tolerance = float(kw.get("tolerance", 1e-8))
    o.lifetime = int(kw.get("lifetime", 1000))
    o.exploration_rate = float(kw.get("exploration_rate", 0.2))
    o.agent_randomization_rate = float(kw.get("agent_randomization_rate", 0.2))
    o.radius = float(kw.get("radius", 1.0))
    o.randomization_rate = float(kw.get("randomization_rate", 0.3))
    o.convergence_length = int(kw.get("convergence_length", 20))
    return o


class BfgsEvent(ctypes.Structure):
    """or_bfgs_event (sq_oracle.h): one observed step of the optimiser."""
    _fields_ = [("kind", ctypes.c_int), ("n", ctypes.c_long), ("iter", ctypes.c_long),
                ("step", ctypes.c_double), ("f", ctypes.c_double), ("dg", ctypes.c_double),
                ("f0", ctypes.c_double), ("dg0", ctypes.c_double), ("f_lowest", ctypes.c_double),
                ("max_step", ctypes.c_double), ("exit_curvature", ctypes.c_int), ("exit_calls", ctypes.c_int),
                ("dTg", ctypes.c_double), ("g0_dot_delta", ctypes.c_double), ("norm_Z", ctypes.c_double)] + \
        [(name, ctypes.POINTER(ctypes.c_double)) for name in ("g", "d", "g0", "delta", "gamma", "Z", "Z_rot", "Ztg",
                                                              "x")]


EV_DIRECTION, EV_PROBE, EV_LS_END, EV_UPDATE, EV_SKIP, EV_START, EV_END = range(7)
_HOOK = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(BfgsEvent))
_FG = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                       ctypes.POINTER(ctypes.c_double))


def _event_dict(ev: BfgsEvent) -> dict:
    n = int(ev.n)
    out = {k: getattr(ev, k) for k in ("kind", "iter", "step", "f", "dg", "f0", "dg0", "f_lowest", "max_step",
                                       "exit_curvature", "exit_calls", "dTg", "g0_dot_delta", "norm_Z")}
    for name in ("g", "d", "g0", "delta", "gamma", "Ztg", "x"):
        ptr = getattr(ev, name)
        out[name] = np.ctypeslib.as_array(ptr, shape=(n,)).copy() if ptr else None
    for name in ("Z", "Z_rot"):
        ptr = getattr(ev, name)
        out[name] = np.ctypeslib.as_array(ptr, shape=(n, n)).copy() if ptr else None
    return out


class Oracle:
    """ctypes binding of sq_oracle.c."""

    def __init__(self):
        self.lib = ctypes.CDLL(build())
        self.correction = DEFAULT_CORRECTION
        L = self.lib
        dp = ctypes.POINTER(ctypes.c_double)
        L.or_apply_circuit.argtypes = [ctypes.POINTER(_Gate), ctypes.c_int, dp, ctypes.c_void_p,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_apply_derivate.argtypes = [ctypes.POINTER(_Gate), ctypes.c_int, dp, ctypes.c_void_p,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        L.or_gate_kernel.argtypes = [ctypes.c_int, dp, ctypes.c_void_p]
        L.or_deriv_kernel.argtypes = [ctypes.c_int, dp, ctypes.c_int, ctypes.c_void_p]
        L.or_cost.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_cost.restype = ctypes.c_double
        L.or_cost_corr.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.or_cost_corr.restype = ctypes.c_double
        L.or_trace_flips.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int]
        L.or_trace_flips.restype = _C128
        pp = ctypes.POINTER(_Problem)
        L.or_optimization_problem.argtypes = [pp, dp]
        L.or_optimization_problem.restype = ctypes.c_double
        L.or_optimization_problem_combined.argtypes = [pp, dp, dp, dp]
        L.or_optimization_problem_batched.argtypes = [pp, dp, ctypes.c_int, ctypes.c_int, dp]
        L.or_optimization_problem_combined_batched.argtypes = [pp, dp, ctypes.c_int, ctypes.c_int, dp, dp]
        L.or_bfgs_powell.argtypes = [pp, dp, ctypes.c_long, ctypes.POINTER(ctypes.c_long),
                                     ctypes.POINTER(ctypes.c_int)]
        L.or_bfgs_powell.restype = ctypes.c_double
        L.or_decompose_custom.argtypes = [pp, ctypes.c_uint, ctypes.c_int, ctypes.c_long, ctypes.c_long,
                                          ctypes.c_long, ctypes.c_double, ctypes.c_double, dp,
                                          ctypes.POINTER(ctypes.c_long)]
        L.or_decompose_custom.restype = ctypes.c_double
        L.or_mt_uniform_int.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_mt_uniform_int.restype = ctypes.c_int
        L.or_mt_canonical.argtypes = [ctypes.c_void_p]
        L.or_mt_canonical.restype = ctypes.c_double
        L.or_mt_uniform.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.or_mt_uniform.restype = ctypes.c_double
        L.or_agents.argtypes = [pp, ctypes.c_uint, ctypes.POINTER(AgentsOpts), dp, ctypes.POINTER(ctypes.c_long),
                                ctypes.POINTER(ctypes.c_long)]
        L.or_agents.restype = ctypes.c_double
        L.or_mt_seed.argtypes = [ctypes.c_void_p, ctypes.c_uint]
        L.or_mt_next.argtypes = [ctypes.c_void_p]
        L.or_mt_next.restype = ctypes.c_uint
        L.or_mt_uniform_2pi.argtypes = [ctypes.c_void_p]
        L.or_mt_uniform_2pi.restype = ctypes.c_double
        L.or_bfgs_set_hook.argtypes = [_HOOK, ctypes.c_void_p]
        L.or_bfgs_powell_fn.argtypes = [_FG, ctypes.c_void_p, ctypes.c_long, dp, ctypes.c_long,
                                        ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_int)]
        L.or_bfgs_powell_fn.restype = ctypes.c_double
        L.or_agents_hs_fit.argtypes = [dp, ctypes.c_double, dp, dp]
        L.or_agents_hs_fit.restype = None
        L.or_agents_hs_search.argtypes = [dp, ctypes.c_double, dp, ctypes.POINTER(ctypes.c_long)]
        L.or_agents_hs_search.restype = ctypes.c_double

    # -- helpers -----------------------------------------------------------
    @staticmethod
    def _gates(gates):
        g = gates_array(gates)
        arr = (_Gate * max(len(g), 1))()
        for i, (t, tg, c, ps) in enumerate(g):
            arr[i].type, arr[i].target, arr[i].control, arr[i].param_start = int(t), int(tg), int(c), int(ps)
        return arr, len(g)

    @staticmethod
    def _dp(a):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

    def _problem(self, gates, target, qbit_num, variant, trace_offset):
        garr, ng = self._gates(gates)
        target = np.ascontiguousarray(target, dtype=np.complex128)
        prev, s1, s2 = self.correction
        pr = _Problem(garr, ng, param_count(gates), qbit_num, target.ctypes.data,
                      target.shape[0], target.shape[1], target.shape[1], variant, trace_offset, prev, s1, s2)
        return pr, (garr, target)

    # -- gate level --------------------------------------------------------
    def gate_kernel(self, gtype, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        out = np.zeros(4, dtype=np.complex128)
        self.lib.or_gate_kernel(gtype, self._dp(p), out.ctypes.data)
        return out.reshape(2, 2)

    def deriv_kernel(self, gtype, params, pidx):
        p = np.ascontiguousarray(params, dtype=np.float64)
        out = np.zeros(4, dtype=np.complex128)
        self.lib.or_deriv_kernel(gtype, self._dp(p), pidx, out.ctypes.data)
        return out.reshape(2, 2)

    # -- circuit level -----------------------------------------------------
    def apply_to(self, gates, params, M):
        """Gates_block::apply_to on a copy of M (rows x cols)."""
        garr, ng = self._gates(gates)
        out = np.ascontiguousarray(M, dtype=np.complex128).copy()
        p = np.ascontiguousarray(params, dtype=np.float64)
        rc = self.lib.or_apply_circuit(garr, ng, self._dp(p), out.ctypes.data,
                                       out.shape[0], out.shape[1], out.shape[1])
        if rc:
            raise ValueError("bad gate in circuit")
        return out

    def apply_derivate_to(self, gates, params, M):
        garr, ng = self._gates(gates)
        M = np.ascontiguousarray(M, dtype=np.complex128)
        P = param_count(gates)
        out = np.zeros((P,) + M.shape, dtype=np.complex128)
        p = np.ascontiguousarray(params, dtype=np.float64)
        rc = self.lib.or_apply_derivate(garr, ng, self._dp(p), M.ctypes.data,
                                        M.shape[0], M.shape[1], M.shape[1], out.ctypes.data)
        if rc:
            raise ValueError("bad gate in circuit")
        return out

    def cost(self, M, variant=COST_FROBENIUS, trace_offset=0):
        """Cost of a transformed matrix (rows = 2^qbit_num); the correction variants use
        self.correction = (prev_cost, scale1, scale2)."""
        M = np.ascontiguousarray(M, dtype=np.complex128)
        n = int(M.shape[0]).bit_length() - 1
        prev, s1, s2 = self.correction
        r = np.sqrt(prev)
        return self.lib.or_cost_corr(M.ctypes.data, n, M.shape[1], M.shape[1], variant, trace_offset, r * s1, r * s2)

    def trace_flips(self, M, bits, trace_offset=0):
        """sum_c sum_{|mask| = bits} M[(c + trace_offset) ^ mask, c] as a complex number."""
        M = np.ascontiguousarray(M, dtype=np.complex128)
        n = int(M.shape[0]).bit_length() - 1
        t = self.lib.or_trace_flips(M.ctypes.data, n, M.shape[1], M.shape[1], trace_offset, bits)
        return complex(t.re, t.im)

    def optimization_problem(self, gates, target, qbit_num, params, variant=COST_FROBENIUS, trace_offset=0):
        pr, keep = self._problem(gates, target, qbit_num, variant, trace_offset)
        p = np.ascontiguousarray(params, dtype=np.float64)
        return self.lib.or_optimization_problem(ctypes.byref(pr), self._dp(p))

    def optimization_problem_combined(self, gates, target, qbit_num, params, variant=COST_FROBENIUS,
                                      trace_offset=0):
        pr, keep = self._problem(gates, target, qbit_num, variant, trace_offset)
        p = np.ascontiguousarray(params, dtype=np.float64)
        f0 = np.zeros(1)
        g = np.zeros(pr.P)
        self.lib.or_optimization_problem_combined(ctypes.byref(pr), self._dp(p), self._dp(f0), self._dp(g))
        return float(f0[0]), g

    def optimization_problem_batched(self, gates, target, qbit_num, params, variant=COST_FROBENIUS,
                                     trace_offset=0, nthreads=1):
        pr, keep = self._problem(gates, target, qbit_num, variant, trace_offset)
        p = np.ascontiguousarray(params, dtype=np.float64).reshape(-1, pr.P)
        out = np.zeros(p.shape[0])
        self.lib.or_optimization_problem_batched(ctypes.byref(pr), self._dp(p), p.shape[0], nthreads,
                                                 self._dp(out))
        return out

    def optimization_problem_combined_batched(self, gates, target, qbit_num, params, variant=COST_FROBENIUS,
                                              trace_offset=0, nthreads=1):
        pr, keep = self._problem(gates, target, qbit_num, variant, trace_offset)
        p = np.ascontiguousarray(params, dtype=np.float64).reshape(-1, pr.P)
        f = np.zeros(p.shape[0])
        g = np.zeros(p.shape)
        self.lib.or_optimization_problem_combined_batched(ctypes.byref(pr), self._dp(p), p.shape[0],
                                                          nthreads, self._dp(f), self._dp(g))
        return f, g

    def bfgs(self, gates, target, qbit_num, x0, max_calls=10000, variant=COST_FROBENIUS, trace_offset=0):
        pr, keep = self._problem(gates, target, qbit_num, variant, trace_offset)
        x = np.ascontiguousarray(x0, dtype=np.float64).copy()
        calls = ctypes.c_long(0)
        st = ctypes.c_int(0)
        f = self.lib.or_bfgs_powell(ctypes.byref(pr), self._dp(x), max_calls, ctypes.byref(calls),
                                    ctypes.byref(st))
        return f, x, calls.value, st.value

    def decompose_custom(self, gates, target, qbit_num, seed, initial_guess=1, max_outer_iterations=4,
                         iteration_loops=1, max_inner_iterations=10000, optimization_tolerance=1e-8,
                         convergence_threshold=1e-5, variant=COST_FROBENIUS, trace_offset=0):
        pr, keep = self._problem(gates, target, qbit_num, variant, trace_offset)
        x = np.zeros(pr.P)
        calls = ctypes.c_long(0)
        f = self.lib.or_decompose_custom(ctypes.byref(pr), seed, initial_guess, max_outer_iterations,
                                         iteration_loops, max_inner_iterations, optimization_tolerance,
                                         convergence_threshold, self._dp(x), ctypes.byref(calls))
        return f, x, calls.value

    def traced(self, call):
        """Runs call() with the optimiser hook set; returns (call's result, events)."""
        events = []
        hook = _HOOK(lambda user, ev: events.append(_event_dict(ev.contents)))
        self.lib.or_bfgs_set_hook(hook, None)
        try:
            res = call()
        finally:
            self.lib.or_bfgs_set_hook(_HOOK(), None)
        return res, events

    def bfgs_fn(self, fg, x0, max_calls=10000):
        """BFGS_Powell on a Python objective fg(x) -> (f, g); returns (f, x, calls, status)."""
        n = len(x0)

        def cb(user, xp, fp, gp):
            f, g = fg(np.ctypeslib.as_array(xp, shape=(n,)).copy())
            fp[0] = float(f)
            np.ctypeslib.as_array(gp, shape=(n,))[:] = g

        fn = _FG(cb)
        x = np.ascontiguousarray(x0, dtype=np.float64).copy()
        calls = ctypes.c_long(0)
        st = ctypes.c_int(0)
        f = self.lib.or_bfgs_powell_fn(fn, None, n, self._dp(x), max_calls, ctypes.byref(calls), ctypes.byref(st))
        return f, x, calls.value, st.value

    def agents(self, gates, target, qbit_num, seed, x0, variant=COST_FROBENIUS, **kw):
        """AGENTS from the solution guess x0 with mt19937(seed); returns
        (current_minimum, optimized parameters, iterations, evaluations)."""
        pr, keep = self._problem(gates, target, qbit_num, variant, 0)
        x = np.array(x0, dtype=np.float64, copy=True)
        it, ev = ctypes.c_long(0), ctypes.c_long(0)
        o = agents_opts(**kw)
        f = self.lib.or_agents(ctypes.byref(pr), seed, ctypes.byref(o), self._dp(x), ctypes.byref(it),
                               ctypes.byref(ev))
        return f, x, it.value, ev.value

    def agents_hs_fit(self, fv, p):
        """The five-point fit: (model[5], start shift) from fv = f at p, p + pi/4,
        p + pi/2, p + pi, p + 3pi/2."""
        fv = np.ascontiguousarray(fv, dtype=np.float64)
        model = np.zeros(5)
        s0 = np.zeros(1)
        self.lib.or_agents_hs_fit(self._dp(fv), float(p), self._dp(model), self._dp(s0))
        return model, float(s0[0])

    def agents_hs_search(self, fv, p):
        """The five-point line search: (model minimum, shift, BFGS calls)."""
        fv = np.ascontiguousarray(fv, dtype=np.float64)
        s = np.zeros(1)
        calls = ctypes.c_long(0)
        f = self.lib.or_agents_hs_search(self._dp(fv), float(p), self._dp(s), ctypes.byref(calls))
        return f, float(s[0]), calls.value

    def mt_draws(self, seed, plan):
        """Interleaved libstdc++ draws from one mt19937(seed): plan is a list of
        ("int", n) / ("canonical",) / ("uniform", a, b)."""
        state = ctypes.create_string_buffer(624 * 4 + 16)
        self.lib.or_mt_seed(state, seed)
        out = []
        for step in plan:
            if step[0] == "int":
                out.append(self.lib.or_mt_uniform_int(state, step[1]))
            elif step[0] == "canonical":
                out.append(self.lib.or_mt_canonical(state))
            else:
                out.append(self.lib.or_mt_uniform(state, step[1], step[2]))
        return out

    def mt_stream(self, seed, n, kind="u32"):
        state = ctypes.create_string_buffer(624 * 4 + 16)
        self.lib.or_mt_seed(state, seed)
        if kind == "u32":
            return np.array([self.lib.or_mt_next(state) for _ in range(n)], dtype=np.uint64)
        return np.array([self.lib.or_mt_uniform_2pi(state) for _ in range(n)])


# ---------------------------------------------------------------------------
# Dense brute-force model (textbook definitions, independent of sq_oracle.c)
# ---------------------------------------------------------------------------
_X = np.array([[0, 1], [1, 0]], dtype=complex)
_Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
_Z = np.array([[1, 0], [0, -1]], dtype=complex)


def textbook_1q(gtype, p):
    """2x2 matrix from the physical angles (2 x raw half-angle)."""
    if gtype == RX:
        return scipy.linalg.expm(-1j * (2 * p[0]) / 2 * _X)
    if gtype in (RY, CRY):
        return scipy.linalg.expm(-1j * (2 * p[0]) / 2 * _Y)
    if gtype == RZ:
        return scipy.linalg.expm(-1j * (2 * p[0]) / 2 * _Z)
    if gtype == U3:
        theta, phi, lam = 2 * p[0], p[1], p[2]
        rz = lambda a: scipy.linalg.expm(-1j * a / 2 * _Z)
        ry = lambda a: scipy.linalg.expm(-1j * a / 2 * _Y)
        return np.exp(1j * (phi + lam) / 2) * rz(phi) @ ry(theta) @ rz(lam)
    if gtype == CNOT:
        return _X.copy()
    if gtype == CZ:
        return _Z.copy()
    raise ValueError(gtype)


def embed(n, u2, target, control=-1):
    """Full 2^n x 2^n matrix of a (controlled) one-qubit gate; qubit q = bit q."""
    d = 1 << n
    full = np.zeros((d, d), dtype=complex)
    for r in range(d):
        if control >= 0 and not (r >> control) & 1:
            full[r, r] = 1.0
            continue
        b = (r >> target) & 1
        r0 = r & ~(1 << target)
        for bb in (0, 1):
            full[r0 | (bb << target), r] += u2[bb, b]
    return full


def brute_circuit_matrix(n, gates, params):
    g = gates_array(gates)
    C = np.eye(1 << n, dtype=complex)
    for t, tg, c, ps in g:
        t = int(t)
        u = textbook_1q(t, params[ps:ps + PARAM_NUM[t]])
        ctrl = int(c) if t in (CNOT, CZ, CRY) else -1
        C = embed(n, u, int(tg), ctrl) @ C
    return C


def brute_apply(n, gates, params, M):
    return brute_circuit_matrix(n, gates, params) @ M


def brute_trace(M, trace_offset=0):
    m = M.shape[1]
    return np.sum(M[np.arange(m) + trace_offset, np.arange(m)])


def brute_flip_trace(M, bits, trace_offset=0):
    """Bit-flip trace from operators instead of index arithmetic: the X string on a
    mask (a Kronecker product of Pauli X and identities, qubit 0 the least significant
    index bit) times the shift that puts column c on row c + trace_offset selects the
    entries M[(c + offset) ^ mask, c]; their sum is <S, M> summed over the masks."""
    d, m = M.shape
    n = d.bit_length() - 1
    X = np.array([[0, 1], [1, 0]], dtype=complex)
    shift = np.zeros((d, m), dtype=complex)
    shift[np.arange(m) + trace_offset, np.arange(m)] = 1.0
    total = 0.0j
    for qs in itertools.combinations(range(n), bits):
        op = np.ones((1, 1), dtype=complex)
        for q in reversed(range(n)):  # kron(A_{n-1}, ..., A_0): qubit 0 last
            op = np.kron(op, X if q in qs else np.eye(2))
        total += np.sum((op @ shift) * M)
    return total


def brute_cost(M, variant=COST_FROBENIUS, trace_offset=0, correction=DEFAULT_CORRECTION):
    m = M.shape[1]
    prev, s1, s2 = correction
    w1, w2 = np.sqrt(prev) * s1, np.sqrt(prev) * s2
    if variant in (COST_FROBENIUS_C1, COST_FROBENIUS_C2):
        c = 1.0 - (brute_trace(M, trace_offset).real + w1 * brute_flip_trace(M, 1, trace_offset).real) / m
        if variant == COST_FROBENIUS_C2:
            c -= w2 * brute_flip_trace(M, 2, trace_offset).real / m
        return c
    if variant in (COST_HS_C1, COST_HS_C2):
        s = abs(brute_trace(M, 0)) ** 2 + w1 * abs(brute_flip_trace(M, 1)) ** 2
        if variant == COST_HS_C2:
            s += w2 * abs(brute_flip_trace(M, 2)) ** 2
        return 1.0 - s / m ** 2
    if variant == COST_FROBENIUS:
        return 1.0 - brute_trace(M, trace_offset).real / m
    if variant == COST_HS:
        return 1.0 - abs(brute_trace(M, 0)) ** 2 / m ** 2
    if variant == COST_INFIDELITY:
        return 1.0 - (abs(brute_trace(M, 0)) ** 2 / m + 1.0) / (m + 1.0)
    if variant == COST_ABS_TRACE:
        return 1.0 - abs(brute_trace(M, trace_offset)) / m
    raise ValueError(variant)


def parameter_shift_derivatives(n, gates, params, M):
    """dC/dp_k . M by the exact parameter-shift identity for raw half-angles:
    dU/dp = [U(p + pi/2) - U(p - pi/2)] / 2 holds for every gate here because each
    raw parameter enters only through sin(p), cos(p) (or e^{ip})."""
    P = len(params)
    out = np.zeros((P,) + M.shape, dtype=complex)
    for k in range(P):
        pp = np.array(params, dtype=float).copy()
        pm = pp.copy()
        pp[k] += np.pi / 2
        pm[k] -= np.pi / 2
        out[k] = (brute_apply(n, gates, pp, M) - brute_apply(n, gates, pm, M)) / 2
    return out
