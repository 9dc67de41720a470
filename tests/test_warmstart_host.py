"""Host-side pieces of the learned warm start: exact label canonicalisation (checked
against the oracle's circuit matrix, phase included), features and sin/cos coding."""
import numpy as np
import pytest

from squander_hip import workloads as W
from squander_hip.warmstart import (_rx, _u3_matrices, _u3_params, angles_to_sincos, canonical_angles,
                                    projective_canonical_angles, sincos_to_angles, target_features)


@pytest.mark.parametrize("n,layers,seed", [(4, 12, 0), (3, 8, 1), (5, 10, 2)])
def test_canonical_angles_keep_the_circuit_matrix(oracle, n, layers, seed):
    gates = W.ansatz_layers(n, W.alternating_pairs(n, layers))
    P = W.param_count(gates)
    rng = np.random.default_rng(seed)
    th = rng.uniform(-4 * np.pi, 4 * np.pi, (6, P))
    can = canonical_angles(gates, th)
    I = np.eye(1 << n)
    for b in range(th.shape[0]):
        A = oracle.apply_to(gates, th[b], I)
        C = oracle.apply_to(gates, can[b], I)
        assert np.linalg.norm(A - C) / np.linalg.norm(A) < 1e-12
    g = np.asarray(gates).reshape(-1, 4)
    u3 = [int(r[3]) for r in g if r[0] == W.GATE_U3]
    thetas = can[:, u3[:-1]]
    assert np.all(thetas >= 0) and np.all(thetas <= np.pi / 2 + 1e-12)
    # every U3 that is next the control of a CNOT followed by a U3 has phi = 0
    zeroed = 0
    for i, r in enumerate(g):
        if r[0] != W.GATE_U3:
            continue
        q = r[1]
        nxt = [j for j in range(i + 1, len(g)) if g[j, 1] == q or g[j, 2] == q]
        if len(nxt) >= 2 and g[nxt[0], 0] == W.GATE_CNOT and g[nxt[0], 2] == q and g[nxt[1], 0] == W.GATE_U3:
            assert np.all(np.abs(can[:, r[3] + 1]) < 1e-12)
            zeroed += 1
    assert zeroed > 0
    # idempotent
    assert np.allclose(canonical_angles(gates, can), can, atol=1e-12)


def test_sincos_round_trip():
    th = np.random.default_rng(3).uniform(0, 2 * np.pi, (5, 7))
    assert np.allclose(sincos_to_angles(angles_to_sincos(th)), th, atol=1e-5)


def test_target_features_remove_global_phase():
    V = W.haar_unitary(16, 5)[None]
    f1 = target_features(V)
    f2 = target_features(V * np.exp(1j * 0.7))
    assert f1.shape == (1, 512) and np.allclose(f1, f2, atol=1e-6)


def _random_gauge(gates, theta, rng):
    """An equivalent circuit (up to global phase): random diag(1, e^{i psi}) moved
    across every CNOT's control and random Rx(b) across its target."""
    g = np.asarray(gates).reshape(-1, 4)
    x = theta.copy()
    B = x.shape[0]
    mats = {i: _u3_matrices(x[:, g[i, 3]], x[:, g[i, 3] + 1], x[:, g[i, 3] + 2])
            for i in range(len(g)) if g[i, 0] == W.GATE_U3}

    def nxt(q, start):
        for j in range(start, len(g)):
            if g[j, 1] == q or (g[j, 0] == W.GATE_CNOT and g[j, 2] == q):
                return j if g[j, 0] == W.GATE_U3 else None

    pending = {}
    for i in range(len(g)):
        if g[i, 0] == W.GATE_U3:
            pending[g[i, 1]] = i
            continue
        t, c = g[i, 1], g[i, 2]
        ic, jc, it, jt = pending.pop(c, None), nxt(c, i + 1), pending.pop(t, None), nxt(t, i + 1)
        if ic is not None and jc is not None:
            D = np.zeros((B, 2, 2), complex)
            D[:, 0, 0] = 1
            D[:, 1, 1] = np.exp(1j * rng.uniform(-3, 3, B))
            mats[ic] = D @ mats[ic]
            mats[jc] = mats[jc] @ np.conj(D)
        if it is not None and jt is not None:
            b = rng.uniform(-3, 3, B)
            mats[it] = _rx(b) @ mats[it]
            mats[jt] = mats[jt] @ _rx(-b)
    for i, M in mats.items():
        a, p, l = _u3_params(M)
        x[:, g[i, 3]], x[:, g[i, 3] + 1], x[:, g[i, 3] + 2] = a, p, l
    return x


@pytest.mark.parametrize("n,layers", [(4, 12), (3, 6)])
def test_projective_canonical_form(oracle, n, layers):
    """Canonical angles give the same circuit up to a global phase, and every
    gauge-equivalent parameter set maps to the same canonical angles."""
    gates = W.ansatz_layers(n, W.alternating_pairs(n, layers))
    P = W.param_count(gates)
    rng = np.random.default_rng(n)
    th = rng.uniform(0, 2 * np.pi, (5, P))
    can = projective_canonical_angles(gates, th)
    other = _random_gauge(gates, th, rng)
    d = 1 << n
    for b in range(5):
        A = oracle.apply_to(gates, th[b], np.eye(d))
        C = oracle.apply_to(gates, can[b], np.eye(d))
        O = oracle.apply_to(gates, other[b], np.eye(d))
        assert abs(abs(np.trace(C.conj().T @ A)) / d - 1) < 1e-12
        assert abs(abs(np.trace(O.conj().T @ A)) / d - 1) < 1e-12
    assert np.max(np.abs(np.mod(other - th + np.pi, 2 * np.pi) - np.pi)) > 1.0
    c2 = projective_canonical_angles(gates, other)
    assert np.max(np.abs(np.mod(c2 - can + np.pi, 2 * np.pi) - np.pi)) < 1e-9
    u3 = [int(r[3]) for r in np.asarray(gates) if r[0] == W.GATE_U3]
    assert np.all(can[:, u3] >= 0) and np.all(can[:, u3] <= np.pi / 2 + 1e-12)
