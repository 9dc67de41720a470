"""CPU emulator of the register-tile program (test infrastructure).

Executes the op and layout records that libsqhip compiles for a circuit
(``sqh_inspect_program``) the way one wavefront of the HIP kernels does: each
column is spread over L lanes holding NR registers, ops act on register pairs
of a local target, controls come from register bits or lane bits, REMAP ops
exchange the slab through a (swizzled) LDS array.  It runs the forward pass and
the adjoint gradient pass, so the host-side compiler and the kernels' algorithm
are checked against the oracle without a GPU.
"""
from __future__ import annotations

import numpy as np

OP_U, OP_CU, OP_X, OP_Z, OP_REMAP = 0, 1, 2, 3, 4
CM_NONE, CM_REG, CM_LANE = 0, 1, 2


def decode(ops_raw, lays_raw):
    ops = []
    for r in ops_raw:
        idx = int(np.frombuffer(bytes(r[4:8]), dtype="<i4")[0])
        ops.append((int(r[0]), int(r[1]), int(r[2]), int(r[3]), idx))
    lays = []
    for r in lays_raw:
        reg_off = np.frombuffer(bytes(r[0:64]), dtype="<i4").astype(np.int64)
        sub_pos = np.frombuffer(bytes(r[64:96]), dtype="<i4").astype(np.int64)
        lays.append((reg_off, sub_pos))
    return ops, lays


class Tile:
    """All columns of one matrix as [L lanes, NR registers, m columns]."""

    def __init__(self, N, lays):
        self.N = N
        self.K = min(N, 4)
        self.NR = 1 << self.K
        self.L = 1 << (N - self.K)
        self.lays = lays

    def sub_off(self, lay):
        _, sub_pos = self.lays[lay]
        s = np.arange(self.L)
        off = np.zeros(self.L, dtype=np.int64)
        for j in range(self.N - self.K):
            off |= ((s >> j) & 1) << int(sub_pos[j])
        return off

    def rows(self, lay):
        reg_off, _ = self.lays[lay]
        return self.sub_off(lay)[:, None] + reg_off[None, : self.NR]   # [L, NR]

    def load(self, M, lay=0):
        return M[self.rows(lay), :].astype(np.complex128)             # [L, NR, m]

    def store(self, V, lay):
        d = 1 << self.N
        out = np.zeros((d, V.shape[2]), dtype=V.dtype)
        out[self.rows(lay), :] = V
        return out

    def remap(self, V, lo, ln):
        # Write under layout lo, read under ln: the value at row r moves with r.
        full = self.store(V, lo)
        return full[self.rows(ln), :]

    def ctrl_mask(self, cmode, c):
        """[L, NR] bool: rows whose control bit is set."""
        i = np.arange(self.NR)
        s = np.arange(self.L)
        if cmode == CM_NONE:
            return np.ones((self.L, self.NR), dtype=bool)
        if cmode == CM_LANE:
            return np.repeat((((s >> c) & 1) == 1)[:, None], self.NR, axis=1)
        return np.repeat((((i >> c) & 1) == 1)[None, :], self.L, axis=0)

    def pairs(self, t):
        i0 = [i for i in range(self.NR) if not (i >> t) & 1]
        return np.array(i0), np.array(i0) | (1 << t)

    def apply(self, V, op, k, deriv=False):
        kind, t, cmode, c, idx = op
        i0, i1 = self.pairs(t)
        mask = self.ctrl_mask(cmode, c)[:, i0][:, :, None]            # [L, npairs, 1]
        a, b = V[:, i0, :], V[:, i1, :]
        if kind in (OP_U, OP_CU):
            na = k[0] * a + k[1] * b
            nb = k[2] * a + k[3] * b
            if kind == OP_CU or deriv:
                zero = np.zeros_like(a)
                na = np.where(mask, na, zero if deriv else a)
                nb = np.where(mask, nb, zero if deriv else b)
        elif kind == OP_X:
            na, nb = np.where(mask, b, a), np.where(mask, a, b)
        elif kind == OP_Z:
            na, nb = a, np.where(mask, -b, b)
        else:
            raise ValueError(kind)
        V = V.copy()
        V[:, i0, :], V[:, i1, :] = na, nb
        return V

    def corr(self, Y, X, op):
        kind, t, cmode, c, idx = op
        i0, i1 = self.pairs(t)
        mask = self.ctrl_mask(cmode, c)[:, i0][:, :, None] if kind == OP_CU else True
        ya, yb, xa, xb = Y[:, i0, :], Y[:, i1, :], X[:, i0, :], X[:, i1, :]
        w = np.where(mask, 1.0, 0.0)
        return np.array([np.sum(w * ya * xa), np.sum(w * ya * xb), np.sum(w * yb * xa), np.sum(w * yb * xb)])


def run_forward(N, ops, lays, kernels, M):
    """C . M through the program; kernels[slot] is the 2x2 (flattened) kernel."""
    T = Tile(N, lays)
    V = T.load(M, 0)
    cur = 0
    for op in ops:
        if op[0] == OP_REMAP:
            o, n = op[4] >> 16, op[4] & 0xFFFF
            assert o == cur, "remap from a layout that is not current"
            V = T.remap(V, o, n)
            cur = n
            continue
        V = T.apply(V, op, kernels[op[4]] if op[0] in (OP_U, OP_CU) else None)
    return T.store(V, cur), cur


def run_adjoint(N, ops, lays, kernels, U, trace_off):
    """Forward to C.U, then the backward sweep; returns (trace, M[slot] 2x2 correlations)."""
    T = Tile(N, lays)
    X = T.load(U, 0)
    cur = 0
    for op in ops:
        if op[0] == OP_REMAP:
            o, n = op[4] >> 16, op[4] & 0xFFFF
            X = T.remap(X, o, n)
            cur = n
            continue
        X = T.apply(X, op, kernels[op[4]] if op[0] in (OP_U, OP_CU) else None)
    m = U.shape[1]
    rows = T.rows(cur)[:, :, None]
    cols = np.arange(m)[None, None, :]
    hit = rows == cols + trace_off
    trace = np.sum(np.where(hit, X, 0))
    Y = np.where(hit, 1.0 + 0j, 0j)
    Ms = {}
    for op in reversed(ops):
        if op[0] == OP_REMAP:
            o, n = op[4] >> 16, op[4] & 0xFFFF
            X = T.remap(X, n, o)
            Y = T.remap(Y, n, o)
            continue
        if op[0] in (OP_X, OP_Z):
            X = T.apply(X, op, None)
            Y = T.apply(Y, op, None)
            continue
        k = kernels[op[4]]
        kd = np.array([np.conj(k[0]), np.conj(k[2]), np.conj(k[1]), np.conj(k[3])])
        kt = np.array([k[0], k[2], k[1], k[3]])
        X = T.apply(X, op, kd)
        Ms[op[4]] = T.corr(Y, X, op)
        Y = T.apply(Y, op, kt)
    return trace, Ms
