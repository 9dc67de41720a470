"""Learned warm start for fixed-ansatz decompositions.

A network maps a target unitary V (the unitary to decompose; the decomposer's
Umtx is V^dagger) to start angles for the ansatz: the input is the real and imaginary
parts of V after a global-phase normalisation (the largest-magnitude entry of the
first column made real and positive), the output is (sin, cos) of every raw
parameter.  Training pairs come from seeded random angles theta*: V = C(theta*) is
built with the library's batched apply_to, and each pair is verified with the
decomposer's own cost on the GPU (C(theta*) applied to V^dagger must give trace d).
A prediction is only a start point: the batched optimiser refines it, and runs that
do not reach the tolerance fall back to random starts.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _native as nat
from .mlp import WarmStartMLP
from .workloads import param_count


def target_features(V: np.ndarray) -> np.ndarray:
    """(n, d, d) complex -> (n, 2 d^2) float32, global phase removed."""
    V = np.asarray(V)
    n, d, _ = V.shape
    col = V[:, :, 0]
    k = np.argmax(np.abs(col), axis=1)
    ph = col[np.arange(n), k]
    Vn = V * (np.conj(ph) / np.abs(ph))[:, None, None]
    return np.concatenate([Vn.real.reshape(n, -1), Vn.imag.reshape(n, -1)], axis=1).astype(np.float32)


def canonical_angles(gates, theta: np.ndarray) -> np.ndarray:
    """Representative of theta's class of exactly equal circuits (same matrix, same
    global phase), so that training labels are one function of the target.

    U3(t, p, l) = [[cos t, -sin t e^{il}], [sin t e^{ip}, cos t e^{i(p+l)}]]
    (gate_kernel_templates.h:509-521) satisfies
      U3(t, p, l) = diag(1, e^{ip}) U3(t, 0, l)  and  U3(t, p, l) diag(1, e^{im}) = U3(t, p, l + m),
      U3(t, p, l) = U3(-t, p + pi, l + pi),   U3(t + pi, p, l) = -U3(t, p, l).
    A diagonal on a CNOT's control commutes with the CNOT, so the phi of a U3 whose
    qubit is next the control of a CNOT that is followed (on that qubit) by a U3 moves
    to that U3's lambda.  Every theta is folded into [0, pi/2]; the product of the
    signs the folds take out goes back into the last U3 as theta + pi.
    Only U3 and CNOT/CZ gates are rewritten; other parametric gates are left alone."""
    g = np.asarray(gates, dtype=np.int64).reshape(-1, 4)
    x = np.array(theta, dtype=np.float64, copy=True)
    single = x.ndim == 1
    if single:
        x = x[None, :]
    u3 = [i for i in range(len(g)) if g[i, 0] == nat.GATE_U3]
    if not u3:
        return theta
    sign = np.ones(x.shape[0])
    for i in u3:
        ps = g[i, 3]
        t = x[:, ps]
        k = np.floor(t / np.pi)
        t0 = t - k * np.pi
        sign *= np.where(np.mod(k, 2) == 1, -1.0, 1.0)
        hi = t0 > np.pi / 2
        x[:, ps] = np.where(hi, np.pi - t0, t0)
        x[:, ps + 1] += np.where(hi, np.pi, 0.0)
        x[:, ps + 2] += np.where(hi, np.pi, 0.0)
        sign *= np.where(hi, -1.0, 1.0)

    def next_on(q, start):
        for j in range(start, len(g)):
            t, tg, c = g[j, 0], g[j, 1], g[j, 2]
            if tg == q or (c == q and t in (nat.GATE_CNOT, nat.GATE_CZ, nat.GATE_CRY)):
                return j
        return None

    for i in u3:
        q = g[i, 1]
        j = next_on(q, i + 1)
        if j is None or g[j, 0] not in (nat.GATE_CNOT, nat.GATE_CZ) or g[j, 2] != q:
            continue
        k = next_on(q, j + 1)
        if k is None or g[k, 0] != nat.GATE_U3:
            continue
        x[:, g[k, 3] + 2] += x[:, g[i, 3] + 1]
        x[:, g[i, 3] + 1] = 0.0
    last = g[u3[-1], 3]
    x[:, last] += np.where(sign < 0, np.pi, 0.0)
    x = np.mod(x, 2.0 * np.pi)
    return x[0] if single else x


def _u3_matrices(t, p, l):
    """(B, 2, 2) U3 kernels from raw parameters (gate_kernel_templates.h:509-521)."""
    c, s = np.cos(t), np.sin(t)
    M = np.empty(t.shape + (2, 2), dtype=np.complex128)
    M[..., 0, 0] = c
    M[..., 0, 1] = -np.exp(1j * l) * s
    M[..., 1, 0] = np.exp(1j * p) * s
    M[..., 1, 1] = np.exp(1j * (p + l)) * c
    return M


def _u3_params(M):
    """Raw U3 parameters (t in [0, pi/2]) of e^{-i g} M for the phase g that makes
    M[0, 0] real and non-negative: every U(2) matrix is a phase times such a U3."""
    a = M[..., 0, 0]
    ph = np.where(np.abs(a) > 1e-300, np.conj(a) / np.maximum(np.abs(a), 1e-300), 1.0)
    g = M * ph[..., None, None]
    c = np.clip(np.real(g[..., 0, 0]), 0.0, 1.0)
    s = np.abs(g[..., 1, 0])
    t = np.arctan2(s, c)
    p = np.where(s > 1e-12, np.angle(g[..., 1, 0]), 0.0)
    # lambda from M01 when sin t > 0, else from M11 = e^{i(p+l)} cos t
    l = np.where(s > 1e-12, np.angle(-g[..., 0, 1]), np.angle(g[..., 1, 1]) - p)
    return t, np.mod(p, 2 * np.pi), np.mod(l, 2 * np.pi)


def _rx(a):
    c, s = np.cos(a / 2), np.sin(a / 2)
    R = np.empty(a.shape + (2, 2), dtype=np.complex128)
    R[..., 0, 0] = c
    R[..., 0, 1] = -1j * s
    R[..., 1, 0] = -1j * s
    R[..., 1, 1] = c
    return R


def projective_canonical_angles(gates, theta: np.ndarray) -> np.ndarray:
    """Representative of theta's class of circuits equal up to a global phase, for
    U3 + CNOT circuits: a diagonal on a CNOT's control and an Rx on its target
    commute with the CNOT, so before every CNOT the control-side U3 is reduced to
    phi = 0 (its diag(1, e^{i phi}) moves into the next U3 on that qubit) and the
    target-side U3 M is replaced by Rx(-a) M with a chosen so that the Bloch vector
    M|0> lies in the x-z half plane z >= 0 (Rx(a) moves into the next U3 on that
    qubit); every U3 is then re-read with theta in [0, pi/2].  The circuit matrix
    changes by a global phase only.  Gates other than U3 and CNOT must not occur."""
    g = np.asarray(gates, dtype=np.int64).reshape(-1, 4)
    x = np.array(theta, dtype=np.float64, copy=True)
    single = x.ndim == 1
    if single:
        x = x[None, :]
    if not np.all(np.isin(g[:, 0], (nat.GATE_U3, nat.GATE_CNOT))):
        raise ValueError("projective canonical form needs a U3 + CNOT circuit")
    mats = {i: _u3_matrices(x[:, g[i, 3]], x[:, g[i, 3] + 1], x[:, g[i, 3] + 2])
            for i in range(len(g)) if g[i, 0] == nat.GATE_U3}

    def next_u3(q, start):
        for j in range(start, len(g)):
            if g[j, 1] == q or (g[j, 0] == nat.GATE_CNOT and g[j, 2] == q):
                return j if g[j, 0] == nat.GATE_U3 else None
        return None

    pending = {}
    for i in range(len(g)):
        if g[i, 0] == nat.GATE_U3:
            pending[int(g[i, 1])] = i
            continue
        t, c = int(g[i, 1]), int(g[i, 2])
        ic, jc = pending.pop(c, None), next_u3(c, i + 1)
        if ic is not None and jc is not None:
            tt, pp, ll = _u3_params(mats[ic])
            mats[ic] = _u3_matrices(tt, np.zeros_like(pp), ll)
            D = np.zeros(pp.shape + (2, 2), dtype=np.complex128)
            D[..., 0, 0] = 1.0
            D[..., 1, 1] = np.exp(1j * pp)
            mats[jc] = mats[jc] @ D
        it, jt = pending.pop(t, None), next_u3(t, i + 1)
        if it is not None and jt is not None:
            M = mats[it]
            v0, v1 = M[..., 0, 0], M[..., 1, 0]
            vy = 2 * np.imag(np.conj(v0) * v1)
            vz = np.abs(v0) ** 2 - np.abs(v1) ** 2
            a = np.arctan2(-vy, vz)  # Rx(-a) v has y = 0, z >= 0
            mats[it] = _rx(-a) @ M
            mats[jt] = mats[jt] @ _rx(a)
    for i, M in mats.items():
        tt, pp, ll = _u3_params(M)
        ps = g[i, 3]
        x[:, ps], x[:, ps + 1], x[:, ps + 2] = tt, pp, ll
    return x[0] if single else x


def angles_to_sincos(theta: np.ndarray) -> np.ndarray:
    return np.concatenate([np.sin(theta), np.cos(theta)], axis=1).astype(np.float32)


def sincos_to_angles(sc: np.ndarray) -> np.ndarray:
    P = sc.shape[1] // 2
    return np.mod(np.arctan2(sc[:, :P], sc[:, P:]), 2.0 * np.pi)


class CircuitBatch:
    """Batched apply_to of one circuit with per-matrix parameters, on the GPU."""

    def __init__(self, ctx: nat.Context, qbit_num: int, gates):
        self.ctx = ctx
        self.n = qbit_num
        self.gates = np.asarray(gates, dtype=np.int32).reshape(-1, 4)
        self.P = param_count(self.gates)
        ctx.set_circuit(qbit_num, self.gates)

    def apply(self, params: np.ndarray, mats: np.ndarray) -> np.ndarray:
        """C(params[b]) . mats[b] for every b (complex128)."""
        d = 1 << self.n
        B = params.shape[0]
        p = torch.from_numpy(np.ascontiguousarray(params, dtype=np.float64)).cuda()
        M = torch.from_numpy(np.ascontiguousarray(mats, dtype=np.complex128)).cuda()
        torch.cuda.synchronize()
        self.ctx.check(self.ctx._L.sqh_apply_to_dev(self.ctx.h, p.data_ptr(), self.P, M.data_ptr(), d, d, d, d * d, B,
                                                    0))
        self.ctx.synchronize()
        return M.cpu().numpy()

    def unitaries(self, params: np.ndarray, chunk: int = 16384) -> np.ndarray:
        d = 1 << self.n
        out = []
        for s in range(0, params.shape[0], chunk):
            p = params[s:s + chunk]
            out.append(self.apply(p, np.broadcast_to(np.eye(d, dtype=np.complex128), (p.shape[0], d, d))))
        return np.concatenate(out)

    def verify(self, params: np.ndarray, V: np.ndarray, chunk: int = 16384, phase_free: bool = False) -> np.ndarray:
        """The decomposer's FROBENIUS cost 1 - Re Tr(C(theta) V^dagger)/d of each pair
        (phase_free: 1 - |Tr(C(theta) V^dagger)|/d, blind to a global phase)."""
        d = 1 << self.n
        costs = []
        for s in range(0, params.shape[0], chunk):
            R = self.apply(params[s:s + chunk], np.conj(np.transpose(V[s:s + chunk], (0, 2, 1))))
            tr = np.trace(R, axis1=1, axis2=2)
            costs.append(1.0 - (np.abs(tr) if phase_free else np.real(tr)) / d)
        return np.concatenate(costs)


def generate_pairs(cb: CircuitBatch, count: int, seed: int):
    """(features, sin/cos labels, canonical angles, V, verification costs) of count
    seeded reachable targets V = C(theta*), theta* uniform in [0, 2pi).  For U3 +
    CNOT circuits the label is the projective canonical form (verified with the
    phase-blind cost, since the features are phase-normalised too); otherwise the
    exact canonical form (verified with the FROBENIUS cost)."""
    rng = np.random.default_rng(seed)
    theta = rng.uniform(0.0, 2.0 * np.pi, (count, cb.P))
    V = cb.unitaries(theta)
    g = np.asarray(cb.gates).reshape(-1, 4)
    if np.all(np.isin(g[:, 0], (nat.GATE_U3, nat.GATE_CNOT))):
        label = projective_canonical_angles(cb.gates, theta)
        costs = cb.verify(label, V, phase_free=True)
    else:
        label = canonical_angles(cb.gates, theta)
        costs = cb.verify(label, V)
    return target_features(V), angles_to_sincos(label), label, V, costs


class WarmStart:
    """The network plus its training loop."""

    def __init__(self, ctx: nat.Context, in_dim: int, P: int, hidden=1024, n_hidden=3, seed=0):
        self.ctx = ctx
        self.net = WarmStartMLP(ctx, in_dim, hidden, n_hidden, 2 * P, seed=seed)
        self.P = P

    def train(self, X: np.ndarray, Y: np.ndarray, epochs=10, batch=1024, lr=1e-3, seed=0, log=None):
        Xd = torch.from_numpy(np.ascontiguousarray(X, dtype=np.float32)).cuda()
        Yd = torch.from_numpy(np.ascontiguousarray(Y, dtype=np.float32)).cuda()
        torch.cuda.synchronize()
        n = X.shape[0]
        rng = np.random.default_rng(seed)
        losses = []
        for ep in range(epochs):
            perm = torch.from_numpy(rng.permutation(n)).cuda()
            Xs, Ys = Xd[perm].contiguous(), Yd[perm].contiguous()
            torch.cuda.synchronize()
            tot, nb = 0.0, 0
            for s in range(0, n - batch + 1, batch):
                want = (s // batch) % 50 == 0
                l = self.net.train_step_dev(Xs[s:].data_ptr(), Ys[s:].data_ptr(), batch, lr, want_loss=want)
                if want:
                    tot += l
                    nb += 1
            self.ctx.synchronize()
            losses.append(tot / max(nb, 1))
            if log:
                log(f"epoch {ep}: loss {losses[-1]:.4f}")
        return losses

    def finetune_cost(self, X: np.ndarray, Umtx: np.ndarray, gates, qbit_num: int, epochs=5, batch=1024, lr=1e-4,
                      seed=0, log=None, variant=None):
        """Train on the decomposition cost itself: for each target the loss is the
        phase-blind cost 1 - |Tr(C(theta) Umtx)|/d at the predicted angles
        theta = atan2(sin, cos), its gradient comes from the library's batched adjoint
        cost+gradient (one target per element), chained through the sin/cos output
        coding, and the network takes one Adam step per batch.  Returns the mean
        training cost per epoch."""
        import torch
        n = X.shape[0]
        P = self.P
        ctx = nat.Context(self.ctx.device)
        ctx.set_circuit(qbit_num, gates)
        ctx.set_cost_function(nat.COST_ABS_TRACE if variant is None else variant)
        Xd = torch.from_numpy(np.ascontiguousarray(X, dtype=np.float32)).cuda()
        Ud = torch.from_numpy(np.ascontiguousarray(Umtx, dtype=np.complex128)).cuda()
        Y = torch.empty((batch, 2 * P), dtype=torch.float32, device="cuda")
        f = torch.empty(batch, dtype=torch.float64, device="cuda")
        g = torch.empty((batch, P), dtype=torch.float64, device="cuda")
        rng = np.random.default_rng(seed)
        costs = []
        for ep in range(epochs):
            perm = torch.from_numpy(rng.permutation(n).astype(np.int32)).cuda()
            Xs = Xd[perm.long()].contiguous()
            torch.cuda.synchronize()
            tot, nb = 0.0, 0
            for s0 in range(0, n - batch + 1, batch):
                idx = perm[s0:s0 + batch].contiguous()
                xb = Xs[s0:s0 + batch]
                self.net.forward_dev(xb.data_ptr(), batch, Y.data_ptr())
                self.ctx.synchronize()
                sn, cs = Y[:, :P].double(), Y[:, P:].double()
                th = torch.atan2(sn, cs).contiguous()
                torch.cuda.synchronize()  # th is written on torch's stream, read on the context's
                ctx.check(ctx._L.sqh_optimization_problem_combined_targets_dev(
                    ctx.h, Ud.data_ptr(), Ud.shape[0], idx.data_ptr(), th.data_ptr(), batch, nat.PREC_F64,
                    f.data_ptr(), g.data_ptr()))
                ctx.synchronize()
                r2 = (sn * sn + cs * cs).clamp_min(1e-12)
                dth = g / batch
                dy = torch.cat([dth * cs / r2, -dth * sn / r2], dim=1).float().contiguous()
                torch.cuda.synchronize()
                self.net.backward_step_dev(xb.data_ptr(), dy.data_ptr(), batch, lr)
                tot += float(f.mean())
                nb += 1
            costs.append(tot / max(nb, 1))
            if log:
                log(f"cost epoch {ep}: mean phase-blind cost {costs[-1]:.4f}")
        return costs

    def finetune_cost_stream(self, gates, qbit_num: int, steps=2000, batch=1024, lr=1e-4, seed=0, log=None,
                             log_every=200):
        """Cost training on a stream of fresh targets: every step draws new angles
        theta* on the GPU, builds V = C(theta*) with the library's batched apply_to,
        takes the phase-normalised features (the rule of target_features) and the
        target Umtx = V^dagger, and makes one Adam step on the phase-blind cost at the
        predicted angles.  No target is seen twice.  Returns the mean cost per log
        window."""
        import torch
        P = self.P
        d = 1 << qbit_num
        ctx = nat.Context(self.ctx.device)
        ctx.set_circuit(qbit_num, gates)
        ctx.set_cost_function(nat.COST_ABS_TRACE)
        gen = torch.Generator(device="cuda")
        gen.manual_seed(int(seed))
        eye = torch.eye(d, dtype=torch.complex128, device="cuda").expand(batch, d, d).contiguous()
        idx = torch.arange(batch, dtype=torch.int32, device="cuda")
        Y = torch.empty((batch, 2 * P), dtype=torch.float32, device="cuda")
        f = torch.empty(batch, dtype=torch.float64, device="cuda")
        g = torch.empty((batch, P), dtype=torch.float64, device="cuda")
        out, win = [], []
        ar = torch.arange(batch, device="cuda")
        for step in range(steps):
            th = torch.rand((batch, P), generator=gen, dtype=torch.float64, device="cuda") * (2.0 * np.pi)
            V = eye.clone()
            torch.cuda.synchronize()
            ctx.check(ctx._L.sqh_apply_to_dev(ctx.h, th.data_ptr(), P, V.data_ptr(), d, d, d, d * d, batch, 0))
            ctx.synchronize()
            col = V[:, :, 0]
            k = torch.argmax(col.abs(), dim=1)
            ph = col[ar, k]
            Vn = V * (ph.conj() / ph.abs())[:, None, None]
            X = torch.cat([Vn.real.reshape(batch, -1), Vn.imag.reshape(batch, -1)], dim=1).float().contiguous()
            U = V.conj().transpose(1, 2).contiguous()
            torch.cuda.synchronize()
            self.net.forward_dev(X.data_ptr(), batch, Y.data_ptr())
            self.ctx.synchronize()
            sn, cs = Y[:, :P].double(), Y[:, P:].double()
            tp = torch.atan2(sn, cs).contiguous()
            torch.cuda.synchronize()
            ctx.check(ctx._L.sqh_optimization_problem_combined_targets_dev(
                ctx.h, U.data_ptr(), batch, idx.data_ptr(), tp.data_ptr(), batch, nat.PREC_F64, f.data_ptr(),
                g.data_ptr()))
            ctx.synchronize()
            r2 = (sn * sn + cs * cs).clamp_min(1e-12)
            dth = g / batch
            dy = torch.cat([dth * cs / r2, -dth * sn / r2], dim=1).float().contiguous()
            torch.cuda.synchronize()
            self.net.backward_step_dev(X.data_ptr(), dy.data_ptr(), batch, lr)
            win.append(float(f.mean()))
            if (step + 1) % log_every == 0 or step + 1 == steps:
                out.append(float(np.mean(win)))
                win = []
                if log:
                    log(f"stream step {step + 1}: mean phase-blind cost {out[-1]:.4f}")
        self.ctx.synchronize()
        return out

    def predict(self, X: np.ndarray) -> np.ndarray:
        Xd = torch.from_numpy(np.ascontiguousarray(X, dtype=np.float32)).cuda()
        Yd = torch.empty((X.shape[0], 2 * self.P), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        self.net.forward_dev(Xd.data_ptr(), X.shape[0], Yd.data_ptr())
        self.ctx.synchronize()
        return sincos_to_angles(Yd.cpu().numpy())
