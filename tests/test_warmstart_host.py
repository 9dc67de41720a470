"""Host-side pieces of the learned warm start: exact label canonicalisation (checked
against the oracle's circuit matrix, phase included), features and sin/cos coding."""
import numpy as np
import pytest

from squander_hip import workloads as W
from squander_hip.warmstart import angles_to_sincos, canonical_angles, sincos_to_angles, target_features


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
