"""The oracle's AGENTS engine and the libstdc++ distributions it consumes.

Pinned by: libstdc++ golden draws (tests/golden/mt19937_agents_draws.txt, written by
g++ from std::mt19937 / uniform_int_distribution / uniform_real_distribution); a
dense-grid brute force showing that every single-agent step moves exactly the
coordinate the generator picked to the global minimiser of the 1-D cost (the
sinusoid property the three-point line search relies on, AGENTS.cpp:425-442); and
the returned minimum being the FP64 cost of the returned parameters."""
import os

import numpy as np
import pytest

import oracle as orc
from squander_hip import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))


def test_libstdcxx_int_and_real_draws_match_golden(oracle):
    lines = open(os.path.join(HERE, "golden", "mt19937_agents_draws.txt")).read().split("\n")
    seed, n = map(int, lines[0].split())
    plan = []
    for _ in range(n):
        plan += [("int", 84), ("uniform", 0.0, 1.0), ("uniform", -2 * np.pi, 2 * np.pi)]
    got = oracle.mt_draws(seed, plan)
    for i in range(n):
        a, b, c = lines[1 + i].split()
        assert got[3 * i] == int(a)
        assert got[3 * i + 1] == float(b)
        assert got[3 * i + 2] == float(c)


def test_canonical_matches_uniform_unit_interval(oracle):
    a = oracle.mt_draws(3, [("canonical",)] * 10)
    b = oracle.mt_draws(3, [("uniform", 0.0, 1.0)] * 10)
    assert a == b


def _reachable(n, layers, seed):
    o = orc.Oracle()
    gates = W.ansatz_layers(n, W.alternating_pairs(n, layers))
    P = W.param_count(gates)
    th = np.random.default_rng(seed).uniform(0, 2 * np.pi, P)
    return gates, o.apply_to(gates, th, np.eye(1 << n)).conj().T


def test_single_agent_steps_are_exact_coordinate_minimisations(oracle):
    gates, U = _reachable(2, 2, 4)
    P = W.param_count(gates)
    x0 = np.random.default_rng(9).uniform(0, 2 * np.pi, P)
    seed = 21
    # generator replay for one agent: initial index, then per iteration one index
    # and two canonical draws (lifetime 1: every iteration publishes the agent's
    # state as optimized_parameters; a single agent is the most successful one, so
    # it never adopts or re-randomises and draws nothing more)
    prev = x0
    grid = np.linspace(0, 2 * np.pi, 2001)

    def cost(y):
        return orc.brute_cost(orc.brute_apply(2, gates, y, U))

    for k in range(1, 6):
        f, x, it, ev = oracle.agents(gates, U, 2, seed, x0, agent_num=1, max_iterations=k, lifetime=1,
                                     tolerance=-1.0)
        assert it == k
        plan = [("int", P)]
        for j in range(k):
            plan += [("int", P), ("canonical",), ("canonical",)]
        idx = oracle.mt_draws(seed, plan)[1 + 3 * (k - 1)]
        changed = np.flatnonzero(np.abs(np.mod(x - prev + np.pi, 2 * np.pi) - np.pi) > 1e-13)
        assert set(changed.tolist()) <= {idx}
        # the new value minimises the cost along that coordinate (brute-force grid)
        fs = []
        for v in grid:
            y = x.copy()
            y[idx] = v
            fs.append(cost(y))
        assert cost(x) <= min(fs) + 1e-9
        assert abs(f - oracle.optimization_problem(gates, U, 2, x)) < 1e-12
        prev = x


@pytest.mark.parametrize("n,layers,seed", [(2, 2, 1), (3, 4, 2)])
def test_agents_reaches_reachable_target(oracle, n, layers, seed):
    gates, U = _reachable(n, layers, seed)
    P = W.param_count(gates)
    x0 = np.random.default_rng(seed + 100).uniform(0, 2 * np.pi, P)
    f0 = oracle.optimization_problem(gates, U, n, x0)
    f, x, it, ev = oracle.agents(gates, U, n, seed, x0, agent_num=64, max_iterations=4000, lifetime=1000,
                                 tolerance=1e-8)
    assert f <= f0
    assert abs(f - oracle.optimization_problem(gates, U, n, x)) < 1e-9
    assert f < 1e-6
