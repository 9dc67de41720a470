"""Seeded problems for the test suite (inputs only; built on squander_hip.workloads)."""
import numpy as np

from squander_hip import workloads as W


def ansatz(n, pairs):
    return W.ansatz_layers(n, pairs)


def problem(n, layers, seed):
    """(n, gates, Umtx, P): a Haar target's adjoint and the adjacent-pair ansatz."""
    gates = W.ansatz_layers(n, W.random_adjacent_pairs(n, layers, seed) if n > 1 else [])
    U = W.haar_unitary(1 << n, seed + 1000)
    return n, gates, U.conj().T.copy(), W.param_count(gates)


def reachable(n, layers, seed):
    """(n, gates, theta*, V = C(theta*)) for an ansatz-reachable target."""
    import oracle as orc
    gates = W.ansatz_layers(n, W.random_adjacent_pairs(n, layers, seed))
    P = W.param_count(gates)
    theta = np.random.default_rng(seed).uniform(0, 2 * np.pi, P)
    V = orc.Oracle().apply_to(gates, theta, np.eye(1 << n))
    return n, gates, theta, V
