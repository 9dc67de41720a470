"""``N_Qubit_Decomposition_custom``: the reference decomposition object's
cost / gradient surface on the GPU.

Method names and semantics follow the reference's Python wrapper
(squander/decomposition/qgd_N_Qubit_Decompositions_Wrapper.cpp:1446-1714,
method table :3125-3211) and Optimization_Interface
(Optimization_Interface.cpp:617-1561):

* ``Optimization_Problem(p)``            -> cost (float)
* ``Optimization_Problem_Grad(p)``       -> gradient (P,)
* ``Optimization_Problem_Combined(p)``   -> (cost, gradient)
* ``Optimization_Problem_Batch(P2d)``    -> costs (B,)
* ``Optimization_Problem_Combined_Unitary(p)`` -> (C.Umtx, [dC/dp_k . Umtx])

The target passed to the constructor is ``Umtx``; the circuit is applied to it and
the cost is 1 - Re Tr(C.Umtx)/m by default (FROBENIUS_NORM), so callers pass the
adjoint of the unitary to decompose, as the reference's tests do
(tests/decomposition/test_decomposition.py:62).
"""
from __future__ import annotations

import numpy as np

from . import _native as nat
from .circuit import Circuit

_COST_NAMES = {"FROBENIUS_NORM": nat.COST_FROBENIUS,
               "FROBENIUS_NORM_CORRECTION1": nat.COST_FROBENIUS_CORRECTION1,
               "FROBENIUS_NORM_CORRECTION2": nat.COST_FROBENIUS_CORRECTION2,
               "HILBERT_SCHMIDT_TEST": nat.COST_HILBERT_SCHMIDT,
               "HILBERT_SCHMIDT_TEST_CORRECTION1": nat.COST_HILBERT_SCHMIDT_CORRECTION1,
               "HILBERT_SCHMIDT_TEST_CORRECTION2": nat.COST_HILBERT_SCHMIDT_CORRECTION2,
               "INFIDELITY": nat.COST_INFIDELITY, "ABS_TRACE": nat.COST_ABS_TRACE}


class N_Qubit_Decomposition_custom:
    def __init__(self, Umtx, qbit_num=None, initial_guess="RANDOM", config=None, accelerator_num=0):
        U = np.ascontiguousarray(Umtx, dtype=np.complex128)
        if U.ndim != 2:
            raise Exception("Umtx must be a matrix")
        rows = U.shape[0]
        n = int(round(np.log2(rows)))
        if (1 << n) != rows:
            raise Exception("Umtx must have 2^n rows")
        if qbit_num is not None and int(qbit_num) != n:
            raise Exception("qbit_num does not match the size of Umtx")
        self.qbit_num = n
        self.Umtx = U
        self.config = dict(config or {})
        self.initial_guess = str(initial_guess).upper() if initial_guess is not None else "RANDOM"
        self.trace_offset = 0
        self.cost_variant = nat.COST_FROBENIUS
        self.circuit = None
        self.optimized_parameters = None
        self.num_iters = 0
        self.decomposition_error = None
        self.optimization_tolerance = float(self.config.get("optimization_tolerance", 1e-7))
        self.max_iterations = None
        # N_Qubit_Decomposition_custom.cpp:63-75: BFGS with 4 outer iterations up to 5
        # qubits, ADAM with one above (ADAM is not built here: Start_Decomposition
        # then raises unless set_Optimizer("BFGS") was called).
        self.optimizer = "BFGS" if n <= 5 else "ADAM"
        self.max_outer_iterations = 4 if n <= 5 else 1
        self._ctx = nat.Context(0)
        self._ctx.set_target(U)
        # Circuit-specialised kernels (compiled on first use per circuit: about 0.6 s
        # for the 8-qubit cost kernel, 12 s for its gradient kernel); a config key this
        # package adds (the reference has none), off by default so single calls keep
        # their latency.
        self._ctx.set_jit(bool(self.config.get("specialised_kernels", False)))

    # -- configuration -------------------------------------------------------
    def set_Gate_Structure(self, circuit: Circuit):
        if not isinstance(circuit, Circuit):
            raise Exception("set_Gate_Structure expects a Circuit")
        if circuit.qbit_num != self.qbit_num:
            raise Exception("the gate structure acts on a different number of qubits")
        self.circuit = circuit
        self._ctx.set_circuit(self.qbit_num, circuit.gate_array())

    def get_Circuit(self):
        return self.circuit

    def get_Parameter_Num(self) -> int:
        return self.circuit.get_Parameter_Num() if self.circuit is not None else 0

    def get_Qbit_Num(self) -> int:
        return self.qbit_num

    def set_Trace_Offset(self, trace_offset):
        self._ctx.set_trace_offset(int(trace_offset))
        self.trace_offset = int(trace_offset)

    def get_Trace_Offset(self) -> int:
        return self.trace_offset

    def set_Cost_Function_Variant(self, costfnc=0):
        if isinstance(costfnc, str):
            costfnc = _COST_NAMES[costfnc.upper()]
        self._ctx.set_cost_function(int(costfnc))
        self.cost_variant = int(costfnc)

    def set_Correction_Weights(self, prev_cost=1.0, scale1=1.0 / 1.7, scale2=1.0 / 2.0):
        """Weights of the correction variants: Optimization_Interface's prev_cost_fnv_val,
        correction1_scale and correction2_scale (Optimization_Interface.cpp:74-76; the
        reference sets prev_cost_fnv_val only from its ADAM engines, ADAM.cpp:201)."""
        self._ctx.set_correction_weights(prev_cost, scale1, scale2)

    def set_Optimizer(self, optimizer="BFGS"):
        """Optimization_Interface::set_optimizer (Optimization_Interface.cpp:1620-1697,
        names as in the wrapper, qgd_N_Qubit_Decompositions_Wrapper.cpp:1197-1260).  BFGS
        sets the reference's BFGS limits (10000 evaluations per run, 1e8 outer
        iterations); the other engines are not built here and make Start_Decomposition
        raise."""
        name = str(optimizer).upper()
        known = ("BFGS", "ADAM", "GRAD_DESCEND", "ADAM_BATCHED", "COSINE", "GRAD_DESCEND_PARAMETER_SHIFT_RULE",
                 "AGENTS", "AGENTS_COMBINED", "BFGS2", "BAYES_OPT", "BAYES_AGENTS")
        if name not in known:
            raise Exception(f"set_Optimizer: unknown optimizer {optimizer}")
        self.optimizer = name
        if name == "BFGS":
            self.max_outer_iterations = 100000000
            self.max_iterations = 10000

    def set_Optimization_Tolerance(self, tol):
        self.optimization_tolerance = float(tol)

    def set_Max_Iterations(self, max_iters):
        self.max_iterations = int(max_iters)

    def get_Num_of_Iters(self) -> int:
        return self.num_iters

    def get_Unitary(self):
        return self.Umtx.copy()

    # -- the hot path -----------------------------------------------------------
    def _check_params(self, p, batch=False):
        if self.circuit is None:
            raise Exception("no gate structure set")
        p = np.asarray(p)
        if p.dtype == np.float32:
            p = p.astype(np.float64)
        p = np.ascontiguousarray(p, dtype=np.float64)
        P = self.get_Parameter_Num()
        if batch:
            if p.ndim != 2 or p.shape[1] != P:
                raise Exception(f"Optimization_Problem_Batch: parameters must be B x {P}")
        elif p.size != P:
            raise Exception(f"Optimization_Interface::optimization_problem: Number of free paramaters should be {P}, "
                            f"but got {p.size}")
        return p

    def Optimization_Problem(self, parameters, precision=nat.PREC_F64) -> float:
        p = self._check_params(parameters)
        out = np.zeros(1)
        self._ctx.check(self._ctx._L.sqh_optimization_problem_batch(self._ctx.h, p.ctypes.data, 1, precision,
                                                                     out.ctypes.data))
        self.num_iters += 1
        return float(out[0])

    def Optimization_Problem_Batch(self, parameters, precision=nat.PREC_F64):
        p = self._check_params(parameters, batch=True)
        out = np.zeros(p.shape[0])
        self._ctx.check(self._ctx._L.sqh_optimization_problem_batch(self._ctx.h, p.ctypes.data, p.shape[0],
                                                                     precision, out.ctypes.data))
        self.num_iters += p.shape[0]
        return out

    def Optimization_Problem_Combined_Batch(self, parameters, precision=nat.PREC_F64):
        p = self._check_params(parameters, batch=True)
        f = np.zeros(p.shape[0])
        g = np.zeros(p.shape)
        self._ctx.check(self._ctx._L.sqh_optimization_problem_combined_batch(
            self._ctx.h, p.ctypes.data, p.shape[0], precision, f.ctypes.data, g.ctypes.data))
        self.num_iters += p.shape[0] * (p.shape[1] + 1)
        return f, g

    def Optimization_Problem_Combined(self, parameters, precision=nat.PREC_F64):
        p = self._check_params(parameters)
        f, g = self.Optimization_Problem_Combined_Batch(p.reshape(1, -1), precision)
        return float(f[0]), g[0]

    def Optimization_Problem_Grad(self, parameters, precision=nat.PREC_F64):
        return self.Optimization_Problem_Combined(parameters, precision)[1]

    def Optimization_Problem_Combined_Unitary(self, parameters):
        p = self._check_params(parameters)
        P = p.size
        U = np.empty(self.Umtx.shape, dtype=np.complex128)
        D = np.empty((P,) + self.Umtx.shape, dtype=np.complex128)
        self._ctx.check(self._ctx._L.sqh_optimization_problem_combined_unitary(
            self._ctx.h, p.ctypes.data, U.ctypes.data, D.ctypes.data if P else None))
        return U, [D[k] for k in range(P)]

    def get_Matrix(self, parameters):
        return self.circuit.get_Matrix(np.ascontiguousarray(parameters, dtype=np.float64))

    # -- the optimiser ----------------------------------------------------------------
    def _options(self, **kw):
        from .optimizer import default_options
        cfg = dict(self.config)
        cfg.update(kw)
        if self.max_iterations is not None:
            cfg.setdefault("max_inner_iterations", self.max_iterations)
        cfg.setdefault("optimization_tolerance", self.optimization_tolerance)
        keys = ("max_inner_iterations", "max_outer_iterations", "optimization_tolerance", "convergence_threshold",
                "mixed_precision", "switch_cost", "rounds_per_graph", "max_rounds", "iteration_loops", "engine")
        if "max_inner_iterations_bfgs" in cfg:
            cfg["max_inner_iterations"] = cfg["max_inner_iterations_bfgs"]
        cfg.setdefault("max_outer_iterations", self.max_outer_iterations)
        return default_options(self.qbit_num, **{k: cfg[k] for k in keys if k in cfg})

    def Start_Decomposition(self, **kw):
        """Decompose Umtx with the gate structure (N_Qubit_Decomposition_custom::
        start_decomposition, N_Qubit_Decomposition_custom.cpp:116-190): initial guess from
        config ``random_seed`` (Decomposition_Base.cpp:415-446), BFGS restarted from the best
        point with the engine's iteration loops and restart jitter (BFGS.cpp:42-152: 3 loops
        at 2 qubits, config max_iteration_loops[_bfgs]), all on the GPU.  Extra keyword
        options: mixed_precision, switch_cost.  Above 5 qubits the reference runs ADAM
        (N_Qubit_Decomposition_custom.cpp:63-75), which is not built here: the call raises
        unless set_Optimizer("BFGS") chose BFGS."""
        from .optimizer import MAX_JITTER_OUTER, decompose_batch, initial_guess, iteration_loops_for, restart_jitter
        if self.circuit is None:
            raise Exception("no gate structure set")
        if self.optimizer != "BFGS":
            raise Exception(f"Start_Decomposition: the {self.optimizer} optimizer (the reference's choice above 5 "
                            f"qubits) is not built on the GPU; call set_Optimizer('BFGS') to decompose with BFGS")
        P = self.get_Parameter_Num()
        seed = int(self.config.get("random_seed", np.random.SeedSequence().entropy & 0xFFFFFFFF))
        if self.optimized_parameters is not None and self.optimized_parameters.size == P:
            x0 = self.optimized_parameters.copy()
        else:
            x0 = initial_guess(P, seed, self.initial_guess)
        opts = self._options(**kw)
        loops = iteration_loops_for(self.qbit_num, self.config)
        opts.iteration_loops = loops
        jitter = None
        if loops > 1:
            J = min(int(opts.max_outer_iterations), MAX_JITTER_OUTER) * loops
            jitter = restart_jitter(P, seed, self.initial_guess, J)[None]
        res = decompose_batch(self._ctx, self.Umtx, np.zeros(1, dtype=np.int32), x0[None, :], opts, jitter)
        self.optimized_parameters = res["x"][0].copy()
        self.num_iters += int(res["evals"][0]) * (P + 1)
        self.decomposition_error = float(res["f"][0])
        self.last_run = res
        return 0

    # -- results -------------------------------------------------------------------
    def get_Optimized_Parameters(self):
        return None if self.optimized_parameters is None else self.optimized_parameters.copy()

    def set_Optimized_Parameters(self, parameters):
        self.optimized_parameters = self._check_params(parameters).copy()

    def get_Decomposition_Error(self):
        """Cost of the optimised parameters (calc_decomposition_error)."""
        if self.optimized_parameters is None:
            raise Exception("no optimized parameters")
        return self.Optimization_Problem(self.optimized_parameters)
