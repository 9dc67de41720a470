"""squander_hip -- MI355X-native gate application, cost and gradient for quantum
gate decomposition, behind the SQUANDER reference's Python API.

The compute runs in libsqhip.so (hand-written HIP kernels for gfx950 behind the C
ABI in include/sqhip.h); this package mirrors the reference's ``Circuit`` and
decomposition objects on top of it.  There is no CPU fallback.
"""
from ._native import (COST_ABS_TRACE, COST_FROBENIUS, COST_HILBERT_SCHMIDT, COST_INFIDELITY, PREC_F32, PREC_F64,
                      Context, SqhError, inspect_program, lib)
from .circuit import Circuit
from .decomposition import N_Qubit_Decomposition_custom
from . import workloads
from .optimizer import agents, agents_batch, agents_options, decompose_batch, decompose_problems, default_options, initial_guess, mt19937_uniform_2pi

__all__ = ["Circuit", "N_Qubit_Decomposition_custom", "Context", "SqhError", "inspect_program", "lib", "workloads", "agents", "agents_batch", "agents_options", "decompose_batch", "decompose_problems", "default_options", "initial_guess",
           "mt19937_uniform_2pi",
           "COST_FROBENIUS", "COST_HILBERT_SCHMIDT", "COST_INFIDELITY", "COST_ABS_TRACE", "PREC_F64", "PREC_F32"]
