// sqh_program.cpp -- circuit -> register-tile program compiler (host only).
//
// The gate loop of Gates_block::apply_to_inner (Gates_block.cpp:683-708) becomes a
// list of ops.  Each op's target must be one of the K register qubits; when it is
// not, a REMAP op moves the slab to a layout whose register qubits are the next K
// distinct targets of the circuit, so one LDS exchange serves as long a run of gates
// as possible.  Controls may sit in registers or in the lane index.
#include "sqh_program.h"

#include <algorithm>
#include <cstring>

namespace sqh {

Layout make_layout(int N, int K, const std::vector<int>& pos) {
    Layout l;
    std::memset(&l, 0, sizeof(l));
    for (int i = 0; i < (1 << K); ++i) {
        int off = 0;
        for (int j = 0; j < K; ++j)
            if ((i >> j) & 1) off |= 1 << pos[j];
        l.reg_off[i] = off;
    }
    for (int j = 0; j < N - K; ++j) l.sub_pos[j] = pos[K + j];
    return l;
}

static bool contains(const std::vector<int>& v, int x) { return std::find(v.begin(), v.end(), x) != v.end(); }

bool compile_program(int N, const std::vector<GateRec>& gates, Program& prog, std::string& err) {
    prog = Program();
    if (N < 1 || N > 30) {
        err = "qbit_num must be in [1, 30]";
        return false;
    }
    prog.N = N;
    const int K = N < 4 ? N : 4;
    prog.K = K;
    int P = 0;
    for (size_t gi = 0; gi < gates.size(); ++gi) {
        const GateRec& g = gates[gi];
        const int np = param_num_of(g.type);
        if (np < 0) {
            err = "unsupported gate type " + std::to_string(g.type) + " at gate " + std::to_string(gi);
            return false;
        }
        if (g.target < 0 || g.target >= N) {
            err = "target qubit index out of range at gate " + std::to_string(gi);
            return false;
        }
        const bool ctrl = g.type == GT_CNOT || g.type == GT_CZ || g.type == GT_CRY;
        if (ctrl && (g.control < 0 || g.control >= N || g.control == g.target)) {
            err = "control qubit index invalid at gate " + std::to_string(gi);
            return false;
        }
        if (np > 0 && g.param_start < 0) {
            err = "negative parameter start index at gate " + std::to_string(gi);
            return false;
        }
        P += np;
    }
    for (size_t gi = 0; gi < gates.size(); ++gi) {
        const int np = param_num_of(gates[gi].type);
        if (np > 0 && gates[gi].param_start + np > P) {
            err = "parameter start index beyond the parameter count at gate " + std::to_string(gi);
            return false;
        }
    }
    prog.P = P;

    std::vector<int> pos(N), where(N);
    for (int q = 0; q < N; ++q) pos[q] = where[q] = q;
    prog.layouts.push_back(make_layout(N, K, pos));
    int cur = 0;

    for (size_t gi = 0; gi < gates.size(); ++gi) {
        const GateRec& g = gates[gi];
        const bool ctrl = g.type == GT_CNOT || g.type == GT_CZ || g.type == GT_CRY;
        int t = g.target, c = ctrl ? g.control : -1;
        // CZ is symmetric in its two qubits: use whichever is already in registers.
        if (g.type == GT_CZ && where[t] >= K && where[c] < K) std::swap(t, c);
        if (where[t] >= K) {
            std::vector<int> want;
            for (size_t gj = gi; gj < gates.size() && (int)want.size() < K; ++gj) {
                int tt = gates[gj].target;
                if (gj == gi) tt = t;
                if (!contains(want, tt)) want.push_back(tt);
            }
            for (int j = 0; j < K && (int)want.size() < K; ++j)
                if (!contains(want, pos[j])) want.push_back(pos[j]);
            std::vector<int> incoming, free_reg;
            for (int q : want)
                if (where[q] >= K) incoming.push_back(q);
            for (int j = 0; j < K; ++j)
                if (!contains(want, pos[j])) free_reg.push_back(j);
            for (size_t i = 0; i < incoming.size(); ++i) {
                const int q_in = incoming[i], j_reg = free_reg[i];
                const int q_out = pos[j_reg], j_lane = where[q_in];
                pos[j_reg] = q_in;
                pos[j_lane] = q_out;
                where[q_in] = j_reg;
                where[q_out] = j_lane;
            }
            prog.layouts.push_back(make_layout(N, K, pos));
            const int nl = (int)prog.layouts.size() - 1;
            if (nl > 0xffff) {
                err = "circuit too long for the tile program";
                return false;
            }
            Op r;
            r.kind = OP_REMAP;
            r.t = r.cmode = r.c = 0;
            r.idx = (cur << 16) | nl;
            prog.ops.push_back(r);
            prog.remaps++;
            cur = nl;
        }
        Op op;
        op.t = (uint8_t)where[t];
        op.cmode = CM_NONE;
        op.c = 0;
        op.idx = 0;
        if (c >= 0) {
            if (where[c] < K) {
                op.cmode = CM_REG;
                op.c = (uint8_t)where[c];
            } else {
                op.cmode = CM_LANE;
                op.c = (uint8_t)(where[c] - K);
            }
        }
        switch (g.type) {
        case GT_U3: case GT_RX: case GT_RY: case GT_RZ: op.kind = OP_U; break;
        case GT_CRY: op.kind = OP_CU; break;
        case GT_CNOT: op.kind = OP_X; break;
        case GT_CZ: op.kind = OP_Z; break;
        default: err = "unsupported gate"; return false;
        }
        if (op.kind == OP_U || op.kind == OP_CU) {
            Slot s;
            s.type = g.type;
            s.param_start = g.param_start;
            s.nparam = param_num_of(g.type);
            s.op = (int)prog.ops.size();
            op.idx = (int)prog.slots.size();
            prog.slots.push_back(s);
        }
        prog.ops.push_back(op);
    }
    return true;
}

FusionPlan plan_fusion(int N, const std::vector<GateRec>& gates) {
    FusionPlan fp;
    const int kmax = N < 4 ? N : 4;
    fp.k = kmax;
    std::vector<int> cur;   // qubits of the open block
    std::vector<size_t> members;
    int slot = 0;
    std::vector<int> slot_of(gates.size(), -1);
    for (size_t i = 0; i < gates.size(); ++i)
        if (gates[i].type != GT_CNOT && gates[i].type != GT_CZ) slot_of[i] = slot++;
    auto close = [&]() {
        if (members.empty()) return;
        // pad to kmax qubits with the lowest unused ones
        for (int q = 0; q < N && (int)cur.size() < kmax; ++q)
            if (std::find(cur.begin(), cur.end(), q) == cur.end()) cur.push_back(q);
        std::sort(cur.begin(), cur.end());
        std::array<int, 4> qs = {0, 0, 0, 0};
        for (int j = 0; j < kmax; ++j) qs[j] = cur[j];
        fp.first.push_back((int)fp.gates.size());
        fp.qubits.push_back(qs);
        auto local = [&](int q) { return (int)(std::find(cur.begin(), cur.end(), q) - cur.begin()); };
        for (size_t gi : members) {
            const GateRec& g = gates[gi];
            FusedRec r;
            r.pad = 0;
            r.t = (int8_t)local(g.target);
            r.c = -1;
            r.slot = slot_of[gi];
            switch (g.type) {
            case GT_CNOT: r.kind = OP_X; r.c = (int8_t)local(g.control); break;
            case GT_CZ: r.kind = OP_Z; r.c = (int8_t)local(g.control); break;
            case GT_CRY: r.kind = OP_CU; r.c = (int8_t)local(g.control); break;
            default: r.kind = OP_U; break;
            }
            fp.gates.push_back(r);
        }
        cur.clear();
        members.clear();
    };
    for (size_t i = 0; i < gates.size(); ++i) {
        const GateRec& g = gates[i];
        std::vector<int> q = {g.target};
        if (g.type == GT_CNOT || g.type == GT_CZ || g.type == GT_CRY) q.push_back(g.control);
        std::vector<int> uni = cur;
        for (int x : q)
            if (std::find(uni.begin(), uni.end(), x) == uni.end()) uni.push_back(x);
        if ((int)uni.size() > kmax) {
            close();
            uni = q;
        }
        cur = uni;
        members.push_back(i);
    }
    close();
    fp.first.push_back((int)fp.gates.size());
    return fp;
}

}  // namespace sqh
