"""Batched decomposition on the GPU (sqh_decompose_batch) and the reference
driver's initial guesses.

Initial guesses follow Decomposition_Base::solve_optimization_problem
(Decomposition_Base.cpp:415-446): a std::mt19937 seeded with config
``random_seed`` (Decomposition_Base.cpp:32-42) feeds
std::uniform_real_distribution(0, 2*pi); RANDOM uses the draws, CLOSE_TO_ZERO
divides them by 100, ZEROS uses zeros.  Before that the decomposition object's
constructor draws its id from the same generator with
std::uniform_int_distribution<>(0, INT_MAX) (Optimization_Interface.cpp:179-180):
libstdc++ maps one 32-bit word onto that range of 2^31 values without rejection,
so the guess starts at the second word.  libstdc++ builds each double from two
32-bit draws (low word first), which is reproduced here on top of numpy's MT19937
with the standard's integer seeding.  The BFGS engine's restart jitter
(BFGS.cpp:121-133) continues on the same generator.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as nat

GUESS = {"ZEROS": 0, "RANDOM": 1, "CLOSE_TO_ZERO": 2}


# 32-bit words the decomposition object's constructor takes from its generator
# (the id draw) before the initial guess.
CONSTRUCTOR_WORDS = 1


def mt19937_uniform_2pi(seed: int, n: int, skip_words: int = 0) -> np.ndarray:
    """n draws of uniform_real_distribution<double>(0, 2pi) from std::mt19937(seed),
    after discarding skip_words 32-bit outputs."""
    bg = np.random.RandomState(int(seed) & 0xFFFFFFFF)._bit_generator
    raw = np.asarray(bg.random_raw(skip_words + 2 * n), dtype=np.uint64)[skip_words:]
    raw = raw.reshape(n, 2).astype(np.float64)
    u = (raw[:, 0] + raw[:, 1] * 4294967296.0) / (4294967296.0 * 4294967296.0)
    u = np.where(u >= 1.0, np.nextafter(1.0, 0.0), u)
    return u * (2.0 * np.pi)


def initial_guess(P: int, seed: int, kind: str = "RANDOM") -> np.ndarray:
    """The start point a custom decomposition with config random_seed = seed draws."""
    kind = kind.upper()
    if kind == "ZEROS":
        return np.zeros(P)
    draws = mt19937_uniform_2pi(seed, P, CONSTRUCTOR_WORDS)
    return draws / 100.0 if kind == "CLOSE_TO_ZERO" else draws


def restart_jitter(P: int, seed: int, kind: str, count: int) -> np.ndarray:
    """(count, P) jitter vectors U[0, 2pi)/100 that the BFGS engine adds after each
    improvement of an iteration loop (BFGS.cpp:121-133), in the order the generator
    yields them after the initial guess (which consumed no draws for ZEROS)."""
    if count <= 0:
        return np.zeros((0, P))
    used = 0 if kind.upper() == "ZEROS" else P
    draws = mt19937_uniform_2pi(seed, used + count * P, CONSTRUCTOR_WORDS)[used:]
    return draws.reshape(count, P) / 100.0


def iteration_loops_for(qbit_num: int, config: dict | None = None) -> int:
    """BFGS iteration loops of the reference (BFGS.cpp:65-82): config
    max_iteration_loops_bfgs, else max_iteration_loops, else 3 at 2 qubits and 1
    otherwise (Optimization_Interface.cpp:149 sets iteration_loops[2] = 3)."""
    cfg = config or {}
    for key in ("max_iteration_loops_bfgs", "max_iteration_loops"):
        if key in cfg:
            return int(cfg[key])
    return 3 if qbit_num == 2 else 1


def default_options(qbit_num: int, **kw) -> nat.DecomposeOptions:
    """Reference defaults of a custom decomposition (BFGS, 4 outer iterations at
    <= 5 qubits, 1 above; 10000 evaluations per BFGS run; tolerance 1e-7;
    convergence threshold 1e-5)."""
    o = nat.DecomposeOptions()
    o.max_inner_iterations = int(kw.get("max_inner_iterations", 10000))
    o.max_outer_iterations = int(kw.get("max_outer_iterations", 4 if qbit_num <= 5 else 1))
    o.optimization_tolerance = float(kw.get("optimization_tolerance", 1e-7))
    o.convergence_threshold = float(kw.get("convergence_threshold", 1e-5))
    o.mixed_precision = int(bool(kw.get("mixed_precision", False)))
    o.switch_cost = float(kw.get("switch_cost", 1e-4))
    o.rounds_per_graph = int(kw.get("rounds_per_graph", 16))
    o.max_rounds = int(kw.get("max_rounds", 0))
    o.iteration_loops = int(kw.get("iteration_loops", 1))
    o.engine = {"auto": 0, "rounds": 1, "persistent": 2}[str(kw.get("engine", "auto"))]
    o.restart_jitter_count = 0
    o.restart_jitter = None
    return o


# Restart-jitter vectors handed to the GPU per run: enough for one improvement per
# BFGS run of the first 64 outer iterations.
MAX_JITTER_OUTER = 64


def _attach_jitter(options: nat.DecomposeOptions, jitter):
    """Sets the options' jitter fields from a (runs, J, P) array (or a flat
    concatenation); returns the contiguous array that must stay alive for the call."""
    if jitter is None:
        options.restart_jitter_count = 0
        options.restart_jitter = None
        return None
    J = int(jitter[0].shape[0]) if isinstance(jitter, (list, tuple)) else int(np.shape(jitter)[1])
    if isinstance(jitter, (list, tuple)):
        flat = np.ascontiguousarray(np.concatenate([np.asarray(j, dtype=np.float64).ravel() for j in jitter]))
    else:
        flat = np.ascontiguousarray(jitter, dtype=np.float64).ravel()
    options.restart_jitter_count = J
    options.restart_jitter = flat.ctypes.data
    return flat


def decompose_batch(ctx: nat.Context, targets, run_target, x0, options: nat.DecomposeOptions, jitter=None):
    """Run len(run_target) decompositions of the context's circuit in lockstep.

    targets: (T, rows, cols) complex128 (each problem's Umtx); run_target: (R,)
    target index per run; x0: (R, P) start points; jitter: (R, J, P) restart-jitter
    vectors for options.iteration_loops > 1 (restart_jitter).  Returns a dict with x
    (R, P), f (R,), status (R,), evals (R,) and rounds."""
    T = np.ascontiguousarray(targets, dtype=np.complex128)
    if T.ndim == 2:
        T = T[None]
    rt = np.ascontiguousarray(run_target, dtype=np.int32)
    X0 = np.ascontiguousarray(x0, dtype=np.float64)
    R, P = X0.shape
    x = np.empty((R, P))
    f = np.empty(R)
    st = np.empty(R, dtype=np.int32)
    ev = np.empty(R, dtype=np.int64)
    rounds = ctypes.c_longlong(0)
    keep = _attach_jitter(options, jitter)  # noqa: F841 (alive for the call)
    ctx.check(ctx._L.sqh_decompose_batch(ctx.h, T.ctypes.data, T.shape[0], T.shape[1], T.shape[2], rt.ctypes.data, R,
                                         X0.ctypes.data, ctypes.byref(options), x.ctypes.data, f.ctypes.data,
                                         st.ctypes.data, ev.ctypes.data, ctypes.byref(rounds)))
    return {"x": x, "f": f, "status": st, "evals": ev, "rounds": int(rounds.value)}


def decompose_problems(ctx: nat.Context, problems, options: nat.DecomposeOptions, jitter=None):
    """Independent decompositions with their own circuits and qubit counts in one
    lockstep batch.  problems: sequence of (qbit_num, gates (G, 4) int array,
    Umtx (2^n, 2^n) complex128, x0 (P,)); jitter: per problem a (J, P_i) array of
    restart-jitter vectors (options.iteration_loops > 1).  Returns a dict with x (list
    of arrays), f, status, evals and rounds."""
    keep = []
    arr = (nat.Problem * len(problems))()
    Ps = []
    for i, (n, gates, U, x0) in enumerate(problems):
        g, ng = nat.gate_array(gates)
        Um = np.ascontiguousarray(U, dtype=np.complex128)
        if Um.shape != (1 << n, 1 << n):
            raise ValueError(f"problem {i}: target must be {1 << n} x {1 << n}")
        xs = np.ascontiguousarray(x0, dtype=np.float64)
        keep += [g, Um, xs]
        arr[i].qbit_num = n
        arr[i].n_gates = ng
        arr[i].gates = ctypes.cast(g, ctypes.POINTER(nat._Gate))
        arr[i].target = Um.ctypes.data
        arr[i].x0 = xs.ctypes.data
        Ps.append(xs.size)
    R = len(problems)
    xo = np.empty(int(sum(Ps)))
    f = np.empty(R)
    st = np.empty(R, dtype=np.int32)
    ev = np.empty(R, dtype=np.int64)
    rounds = ctypes.c_longlong(0)
    keep.append(_attach_jitter(options, jitter))
    ctx.check(ctx._L.sqh_decompose_problems(ctx.h, arr, R, ctypes.byref(options), xo.ctypes.data, f.ctypes.data,
                                            st.ctypes.data, ev.ctypes.data, ctypes.byref(rounds)))
    xs_out, o = [], 0
    for p in Ps:
        xs_out.append(xo[o:o + p].copy())
        o += p
    return {"x": xs_out, "f": f, "status": st, "evals": ev, "rounds": int(rounds.value)}


def agents_options(**kw) -> nat.AgentsOptions:
    """AGENTS options with the reference defaults (AGENTS.cpp:98-232)."""
    o = nat.AgentsOptions()
    o.agent_num = int(kw.get("agent_num", 64))
    o.max_iterations = int(kw.get("max_iterations", kw.get("max_inner_iterations", 10000)))
    o.optimization_tolerance = float(kw.get("optimization_tolerance", 1e-8))
    o.agent_lifetime = int(kw.get("agent_lifetime", 1000))
    o.exploration_rate = float(kw.get("agent_exploration_rate", 0.2))
    o.agent_randomization_rate = float(kw.get("agent_randomization_rate", 0.2))
    o.radius = float(kw.get("Randomized_Radius", 1.0))
    o.randomization_rate = float(kw.get("randomization_rate", 0.3))
    o.convergence_length = int(kw.get("convergence_length", 20))
    return o


def agents(ctx: nat.Context, x0, seed: int, options: nat.AgentsOptions | None = None):
    """AGENTS on the context's circuit and target from the guess x0; returns
    (current_minimum, optimized parameters, iterations, evaluations)."""
    x = np.array(x0, dtype=np.float64, copy=True)
    if x.shape != (ctx.parameter_num(),):
        raise ValueError("x0 must hold parameter_num values")
    f = ctypes.c_double(0.0)
    it, ev = ctypes.c_longlong(0), ctypes.c_longlong(0)
    o = options if options is not None else agents_options()
    ctx.check(ctx._L.sqh_agents(ctx.h, x.ctypes.data, ctypes.byref(o), seed, ctypes.byref(f), ctypes.byref(it),
                                ctypes.byref(ev)))
    return f.value, x, int(it.value), int(ev.value)


def agents_batch(ctx: nat.Context, targets, x0s, seeds, options: nat.AgentsOptions | None = None):
    """Independent AGENTS runs, run k against targets[k] from the guess x0s[k] with
    seed seeds[k], advanced in lockstep with one launch per stage; returns
    (current_minimum[K], parameters[K, P], iterations[K], evaluations[K]).  Each run
    equals agents() on a context holding that target."""
    T = np.ascontiguousarray(targets, dtype=np.complex128)
    if T.ndim != 3:
        raise ValueError("targets must be K x rows x cols")
    K, rows, cols = T.shape
    x = np.array(x0s, dtype=np.float64, copy=True)
    if x.ndim != 2 or x.shape != (K, ctx.parameter_num()):
        raise ValueError("x0s must be K x parameter_num")
    x = np.ascontiguousarray(x)
    sd = np.ascontiguousarray(seeds, dtype=np.uint32)
    if sd.shape != (K,):
        raise ValueError("seeds must hold one seed per target")
    f = np.empty(K)
    it = np.empty(K, dtype=np.int64)
    ev = np.empty(K, dtype=np.int64)
    o = options if options is not None else agents_options()
    ctx.check(ctx._L.sqh_agents_batch(ctx.h, T.ctypes.data, K, rows, cols, x.ctypes.data, ctypes.byref(o),
                                      sd.ctypes.data, f.ctypes.data, it.ctypes.data, ev.ctypes.data))
    return f, x, it, ev
