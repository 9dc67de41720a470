"""Pins the oracle's optimiser (or_bfgs_powell / or_decompose_custom, the restatement
of BFGS_Powell.cpp:70-439, grad_descend.cpp:127-272, 396-425 and BFGS.cpp:42-152) by
what the mathematics of the method fixes, observed through the oracle's step hook:

* Powell's factored update (M. J. D. Powell, "Updating conjugate directions by the
  BFGS formula", Math. Prog. 38, 1987): with H = Z Z^T the inverse-Hessian estimate,
  every applied update satisfies the secant equation H gamma = delta, the new first
  column is delta / sqrt(delta.gamma) and the other columns are conjugate to gamma
  (gamma.z_c = 0); the Givens sweep before it leaves Z Z^T unchanged and folds Z^T g0
  into its first entry; no column other than the first ends shorter than norm_Z;
* the search direction is d = -Z Z^T g, a descent direction when the search runs;
* the inexact line search stops at the first probe meeting the Armijo condition with
  0.1 and the curvature condition with 0.7 (relative to the lowest accepted point), so
  an exit by these tests satisfies the Wolfe conditions against the start point, and
  every probe that does not exit fails one of them; the first probe is
  min(1, max_step) with max_step = min over |d_k| >= 1e-5 of 2 pi / |d_k|; the accepted
  point is no worse than the start;
* an update is applied exactly when delta.gamma >= 0.1 |g0.delta|;
* the driver's generator: start points and restart jitter follow libstdc++'s draws
  (tests/golden/mt19937_decomposition_draws.txt) after the constructor's id draw;
* on a strictly convex quadratic the method converges to A^-1 b until every
  component of the direction is below 1e-5, and gamma = A delta.

A swapped Wolfe constant, a dropped Givens rotation, a missing column rescale, a wrong
sign in the direction or a shifted random stream each fails one of these checks.
"""
import os

import numpy as np
import pytest

import oracle as orc
import workloads_for_tests as WT

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
TWO_PI = 2.0 * np.pi


def circuit_problem(n, layers, seed, variant):
    n, gates, theta, V = WT.reachable(n, layers, seed)
    x0 = np.random.default_rng(seed + 50).uniform(0, TWO_PI, theta.size)
    return gates, V.conj().T.copy(), n, x0, variant


def run_traced(oracle, prob):
    gates, U, n, x0, variant = prob
    return oracle.traced(lambda: oracle.bfgs(gates, U, n, x0, variant=variant))


def quadratic(n, seed, cond=50.0):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    A = Q @ np.diag(np.geomspace(1.0, cond, n)) @ Q.T
    b = rng.normal(size=n)
    return A, b


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


CASES = [("2q", (2, 3, 4, orc.COST_FROBENIUS)), ("3q", (3, 4, 9, orc.COST_FROBENIUS)),
         ("3q_hs", (3, 5, 12, orc.COST_HS)), ("4q", (4, 3, 21, orc.COST_FROBENIUS))]


@pytest.fixture(scope="module", params=[c[1] for c in CASES], ids=[c[0] for c in CASES])
def traced_run(request, oracle):
    (res, events) = run_traced(oracle, circuit_problem(*request.param))
    return res, events


def split_searches(events):
    """Groups the events into (direction, probes, ls_end, update-or-skip) per iteration."""
    out, cur = [], None
    for e in events:
        if e["kind"] == orc.EV_DIRECTION:
            cur = {"dir": e, "probes": [], "end": None, "upd": None}
            out.append(cur)
        elif e["kind"] == orc.EV_PROBE:
            cur["probes"].append(e)
        elif e["kind"] == orc.EV_LS_END:
            cur["end"] = e
        elif e["kind"] in (orc.EV_UPDATE, orc.EV_SKIP):
            cur["upd"] = e
    return out


def test_search_direction_is_minus_Z_ZT_g(traced_run):
    (f, x, calls, st), events = traced_run
    dirs = [e for e in events if e["kind"] == orc.EV_DIRECTION]
    assert len(dirs) >= 3
    for e in dirs:
        Z, g = e["Z"], e["g"]
        assert rel(e["Ztg"], Z.T @ g) < 1e-13
        assert rel(e["d"], -(Z @ (Z.T @ g))) < 1e-12
        assert abs(e["dg"] - e["d"] @ g) <= 1e-12 * np.linalg.norm(e["d"]) * np.linalg.norm(g)


def test_line_search_wolfe_exits_and_first_step(traced_run):
    (f, x, calls, st), events = traced_run
    n_exit = n_cont = 0
    for it in split_searches(events):
        probes = it["probes"]
        if not probes:
            continue
        d = it["dir"]["d"]
        dg0 = probes[0]["dg0"]
        assert dg0 < 0.0 and abs(dg0 - it["dir"]["dg"]) == 0.0
        big = np.abs(d) >= 1e-5
        ms = np.min(TWO_PI / np.abs(d[big])) if big.any() else 0.0
        assert abs(probes[0]["max_step"] - ms) <= 1e-15 * ms
        assert probes[0]["step"] == min(1.0, probes[0]["max_step"])
        for i, p in enumerate(probes):
            armijo = p["f"] < p["f_lowest"] + p["step"] * 0.1 * p["dg0"]
            curvature = p["dg"] >= 0.7 * p["dg0"]
            if p["exit_calls"]:
                continue
            if i + 1 < len(probes):
                # the search went on: this probe failed one of the tests
                n_cont += 1
                assert not (armijo and curvature)
            elif armijo and curvature:
                n_exit += 1
                # the Wolfe conditions against the search's start point
                assert p["f"] <= p["f0"] + 0.1 * p["step"] * p["dg0"]
                assert p["dg"] >= 0.7 * p["dg0"]
            else:
                # the last probe passed on without meeting both tests: the search ran
                # out of probes (50) or its step shrank below 1.42e-14
                assert len(probes) == 50 or p["step"] * 0.1 < 1.42e-14
        assert it["end"]["f"] <= it["end"]["f0"]
    assert n_exit > 0 and n_cont > 0


def test_powell_update_secant_and_conjugacy(traced_run):
    (f, x, calls, st), events = traced_run
    n_upd = n_exact = 0
    for it in split_searches(events):
        u = it["upd"]
        if u is None or it["end"] is None:
            continue
        delta, gamma, g0 = u["delta"], u["gamma"], u["g0"]
        # delta is the accepted step along d (formed as x - x0, so rounded at the
        # scale of x), gamma the gradient change
        assert np.max(np.abs(delta - it["end"]["step"] * it["dir"]["d"])) <= 1e-13
        assert rel(gamma, it["end"]["g"] - g0) < 1e-13
        dTg = delta @ gamma
        if u["kind"] == orc.EV_SKIP:
            assert dTg < 0.1 * abs(g0 @ delta)
            continue
        n_upd += 1
        assert dTg >= 0.1 * abs(g0 @ delta)
        Zb, Zr, Z = it["dir"]["Z"], u["Z_rot"], u["Z"]
        # Givens sweep: orthogonal, so Z Z^T is unchanged; folds Z^T g0 into entry 0
        assert rel(Zr @ Zr.T, Zb @ Zb.T) < 1e-12
        v = Zr.T @ g0
        nv = np.linalg.norm(Zb.T @ g0)
        assert abs(v[0] - nv) <= 1e-12 * nv and np.max(np.abs(v[1:])) <= 1e-12 * nv
        # new factor: first column delta / sqrt(delta.gamma) (delta.gamma clamped below
        # at 1.42e-14 and entries of |delta| <= 1.42e-14 dropped, grad_descend.cpp:40),
        # the other columns conjugate to gamma
        used = u["dTg"]
        assert abs(used - max(dTg, 1.42e-14)) <= 1e-13 * max(dTg, 1.42e-14)
        col0 = np.where(np.abs(delta) > 1.42e-14, delta / np.sqrt(used), 0.0)
        assert rel(Z[:, 0], col0) < 1e-13
        cn = np.linalg.norm(Z[:, 1:], axis=0)
        zg = Z.T @ gamma
        if dTg >= 1.42e-14 and np.all(np.abs(delta) > 1.42e-14):
            n_exact += 1
            assert np.all(np.abs(zg[1:]) <= 1e-10 * cn * np.linalg.norm(gamma))
            assert abs(zg[0] - np.sqrt(dTg)) <= 1e-12 * np.sqrt(dTg)
            # the secant equation of the inverse-Hessian estimate H = Z Z^T
            assert rel(Z @ (Z.T @ gamma), delta) < 1e-10
        # column rescale: no column beyond the first shorter than norm_Z
        sq = cn * cn
        assert np.all((sq >= u["norm_Z"] * (1 - 1e-12)) | (sq <= 1.42e-14))
    assert n_upd >= 2 and n_exact >= 2


def test_quadratic_converges_and_gamma_is_A_delta(oracle):
    for n, seed in ((6, 1), (12, 2), (20, 3)):
        A, b = quadratic(n, seed)
        xs = np.linalg.solve(A, b)
        fg = lambda x: (0.5 * x @ A @ x - b @ x, A @ x - b)  # noqa: E731
        (f, x, calls, st), events = oracle.traced(lambda: oracle.bfgs_fn(fg, np.zeros(n), max_calls=2000))
        # the method stops once every component of the direction is below 1e-5 (the
        # maximal step is then 0: grad_descend.cpp:396-425), else at a minimum
        last = [e for e in events if e["kind"] == orc.EV_DIRECTION][-1]
        assert st == 5 or (st == 2 and np.max(np.abs(last["d"])) < 1e-5)
        assert np.linalg.norm(x - xs) < 1e-4 * np.linalg.norm(xs)
        assert abs(f - fg(xs)[0]) < 1e-8
        ups = [e for e in events if e["kind"] == orc.EV_UPDATE]
        assert len(ups) >= n // 2
        # along a line a quadratic is a parabola, which the line search's fit through
        # (0, f0) and (s_h, f_h) with slope dg_h reproduces exactly: after a first probe
        # that fails the Armijo test the next lies at max(0.1 s_h, s*/2), s* the exact
        # line minimiser -dg0 / d.A.d (the fit's tip, halved: grad_descend.cpp:218-236)
        n_fit = 0
        for it in split_searches(events):
            pr = it["probes"]
            if len(pr) < 2 or pr[0]["step"] == 0.0 or pr[0]["f"] < pr[0]["f0"] + pr[0]["step"] * 0.1 * pr[0]["dg0"]:
                continue
            d = it["dir"]["d"]
            s_star = -pr[0]["dg0"] / (d @ A @ d)
            expect = max(0.1 * pr[0]["step"], 0.5 * s_star)
            assert abs(pr[1]["step"] - expect) <= 1e-9 * expect
            n_fit += 1
        assert n_fit >= 1
        for u in ups:
            assert rel(u["gamma"], A @ u["delta"]) < 1e-9
            Z = u["Z"]
            assert rel(Z @ (Z.T @ u["gamma"]), u["delta"]) < 1e-10


def golden_decomposition_draws():
    lines = open(os.path.join(GOLDEN, "mt19937_decomposition_draws.txt")).read().split("\n")
    out, i = {}, 0
    while i < len(lines) and lines[i].strip():
        seed, _id, n = (int(v) for v in lines[i].split())
        out[seed] = np.array([float(v) for v in lines[i + 1:i + 1 + n]])
        i += 1 + n
    return out


@pytest.mark.parametrize("seed", [1, 42])
def test_driver_start_and_restart_jitter_follow_libstdcxx(oracle, seed):
    """Two qubits, one CNOT layer (12 angles), three iteration loops (the reference's
    setting at 2 qubits): the first BFGS run starts at the RANDOM guess drawn after the
    constructor's id draw; after a run that improves the minimum, the next loop starts
    at its end point plus the next 12 draws / 100; after one that does not, at its end
    point unchanged; each outer iteration starts from the best point so far."""
    draws = golden_decomposition_draws()[seed]
    gates = WT.ansatz(2, [0])
    P = orc.param_count(gates)
    assert P == 12
    U = orc.Oracle().apply_to(gates, np.random.default_rng(seed).uniform(0, TWO_PI, P), np.eye(4)).conj().T
    (f, x, calls), events = oracle.traced(
        lambda: oracle.decompose_custom(gates, U, 2, seed, initial_guess=1, max_outer_iterations=2,
                                        iteration_loops=3, optimization_tolerance=1e-30))
    runs = []
    for e in events:
        if e["kind"] == orc.EV_START:
            runs.append({"x0": e["x"]})
        elif e["kind"] == orc.EV_END:
            runs[-1]["x1"], runs[-1]["f"] = e["x"], e["f"]
    assert len(runs) == 6
    assert np.array_equal(runs[0]["x0"], draws[:P])
    used, cur_min, best = P, np.inf, None
    for k, r in enumerate(runs[:-1]):
        improved = r["f"] < cur_min
        if improved:
            cur_min, best = r["f"], r["x1"].copy()
        nxt = runs[k + 1]["x0"]
        if (k + 1) % 3 == 0:  # a new outer iteration starts from the best point
            assert np.array_equal(nxt, best)
            if improved:
                used += P  # the draws were consumed even though no loop followed
        elif improved:
            assert np.array_equal(nxt, r["x1"] + draws[used:used + P] / 100.0)
            used += P
        else:
            assert np.array_equal(nxt, r["x1"])
    assert used <= draws.size


def test_python_start_points_follow_libstdcxx():
    """The Python mirror's initial guess and restart jitter reproduce the golden draws."""
    from squander_hip.optimizer import initial_guess, restart_jitter
    for seed, draws in golden_decomposition_draws().items():
        assert np.array_equal(initial_guess(12, seed, "RANDOM"), draws[:12])
        assert np.array_equal(initial_guess(12, seed, "CLOSE_TO_ZERO"), draws[:12] / 100.0)
        assert np.array_equal(restart_jitter(12, seed, "RANDOM", 3), draws[12:48].reshape(3, 12) / 100.0)
        assert np.array_equal(restart_jitter(16, seed, "ZEROS", 3), draws[:48].reshape(3, 16) / 100.0)
