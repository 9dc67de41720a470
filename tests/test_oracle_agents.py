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


def _hs_along(oracle, gates, U, n, x, j, shifts):
    out = []
    for s in shifts:
        y = x.copy()
        y[j] += s
        out.append(oracle.optimization_problem(gates, U, n, y, variant=orc.COST_HS))
    return np.array(out)


def test_hs_five_point_fit_reproduces_the_cost_along_a_parameter(oracle):
    """The HILBERT_SCHMIDT_TEST cost along one parameter is kappa sin(2s + a) +
    gamma sin(s + b) + offset (|Tr|^2 of a trace linear in cos/sin of the raw
    parameter): the fit from five values must give the cost at any other shift."""
    n = 3
    gates = W.mixed_gate_circuit(n, 20, 77)
    P = W.param_count(gates)
    U = W.haar_unitary(8, 78).conj().T.copy()
    rng = np.random.default_rng(79)
    x = rng.uniform(0, 2 * np.pi, P)
    for j in rng.choice(P, 8, replace=False):
        p = x[j]
        fv = [oracle.optimization_problem(gates, U, n, x, variant=orc.COST_HS)]
        y = x.copy()
        for step in (np.pi / 4, np.pi / 4, np.pi / 2, np.pi / 2):
            y[j] += step
            fv.append(oracle.optimization_problem(gates, U, n, y, variant=orc.COST_HS))
        m, s0 = oracle.agents_hs_fit(fv, p)
        s = rng.uniform(-2 * np.pi, 2 * np.pi, 16)
        model = m[0] * np.sin(2 * s + m[1]) + m[2] * np.sin(s + m[3]) + m[4]
        assert np.max(np.abs(model - _hs_along(oracle, gates, U, n, x, j, s))) < 1e-12
        # the start shift is the minimiser of the larger term, in (-2pi, 2pi); for the
        # double-frequency term the reference's start (AGENTS.cpp:658) subtracts p / 2
        # where the minimiser subtracts p, and that start is kept
        assert abs(s0) < 2 * np.pi
        if abs(m[2]) > abs(m[0]):
            assert abs(np.sin(s0 + m[3]) + 1) < 1e-12
        else:
            assert abs(np.sin(2 * s0 + m[1] - p) + 1) < 1e-12


def test_hs_search_descends_the_model_to_a_stationary_point(oracle):
    """BFGS_Powell on the 1-D model from the start shift (capped at 10 calls): the
    returned value is the model at the returned shift, no higher than at the start,
    and the shift is stationary to the search's step bound (or the cap was reached)."""
    rng = np.random.default_rng(5)
    hits = 0
    for _ in range(200):
        kappa, gamma = rng.uniform(0, 1, 2)
        a, b, off = rng.uniform(-np.pi, np.pi, 3)
        p = rng.uniform(0, 2 * np.pi)

        def model(s):
            return kappa * np.sin(2 * (p + s) + a) + gamma * np.sin(p + s + b) + off

        fv = [model(0.0), model(np.pi / 4), model(np.pi / 2), model(np.pi), model(3 * np.pi / 2)]
        m, s0 = oracle.agents_hs_fit(fv, p)
        f, s, calls = oracle.agents_hs_search(fv, p)
        assert abs(f - model(s)) < 1e-12
        assert f <= model(s0) + 1e-15
        dm = 2 * kappa * np.cos(2 * (p + s) + a) + gamma * np.cos(p + s + b)
        # the search stops once its step falls below 1e-5 (grad_descend.cpp:396-425),
        # i.e. at |f'| of about 1e-5 x the curvature (at most 4 kappa + gamma = 5)
        assert abs(dm) < 1e-4 or calls >= 10
        grid = model(np.linspace(0, 2 * np.pi, 4001))
        hits += f <= grid.min() + 1e-9
    # started near the larger term's minimum, the search mostly ends in the global
    # minimum of the two-frequency model (141 of these 200 draws)
    assert hits >= 120


def test_agents_hs_converges_and_reports_the_cost_of_its_parameters(oracle):
    gates, V = _reachable(2, 2, 11)
    P = W.param_count(gates)
    x0 = np.random.default_rng(12).uniform(0, 2 * np.pi, P)
    f, x, it, ev = oracle.agents(gates, V, 2, 3, x0, variant=orc.COST_HS, agent_num=16, max_iterations=2000,
                                 lifetime=10, tolerance=1e-10)
    # the 1-D search stops at steps of 1e-5, so the HS cost settles near 1e-8 and the
    # convergence window ends the run
    assert f < 1e-6
    assert abs(f - oracle.optimization_problem(gates, V, 2, x, variant=orc.COST_HS)) < 1e-12
    # four evaluations per agent and iteration besides the periodic re-evaluations
    assert ev >= 1 + 16 + 4 * 16 * it


def test_agents_rejects_costs_without_a_line_search(oracle):
    gates, V = _reachable(2, 1, 1)
    x0 = np.zeros(W.param_count(gates))
    f, _, it, ev = oracle.agents(gates, V, 2, 3, x0, variant=orc.COST_INFIDELITY, max_iterations=5)
    assert f == np.finfo(float).max and ev == 0
