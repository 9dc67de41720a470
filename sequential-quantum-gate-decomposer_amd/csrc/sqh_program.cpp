// sqh_program.cpp -- circuit -> register-tile program compiler (host only).
//
// The gate loop of Gates_block::apply_to_inner (Gates_block.cpp:683-708) becomes a
// list of ops.  Each op's target must be one of the K register qubits; when it is
// not, a REMAP op moves the slab through LDS to a layout with other register qubits.
// Controls may sit in registers or in the lane index.
//
// Gates on disjoint qubits commute (they act on different row-index bits), so the
// compiler runs them in any order that keeps, for every qubit, the order of the gates
// touching it.  Up to 10 qubits it schedules greedily: with the current register set
// it runs every gate that becomes runnable (lowest gate index first), and when none
// is left it remaps to the register set, among all C(N, K), under which the most
// gates run next.  On the 8-qubit 40-layer ansatz this takes 5 remaps where keeping
// the gate order takes 14.
#include "sqh_program.h"

#include <algorithm>
#include <array>
#include <cstring>
#include <functional>
#include <queue>

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
    // Lane-bit positions exist for the tile kernels' range (N - K <= 8 entries); above
    // it the layouts are never read (the streaming paths apply the gates by slot).
    const int nsub = std::min(N - K, (int)(sizeof(l.sub_pos) / sizeof(l.sub_pos[0])));
    for (int j = 0; j < nsub; ++j) l.sub_pos[j] = pos[K + j];
    return l;
}

static bool contains(const std::vector<int>& v, int x) { return std::find(v.begin(), v.end(), x) != v.end(); }

namespace {

bool controlled(int type) { return type == GT_CNOT || type == GT_CZ || type == GT_CRY; }

// Register set R (bit q = qubit q) serves gate g: its target is in R (CZ is
// symmetric, either of its qubits will do).
bool serves(const GateRec& g, uint32_t R) {
    if ((R >> g.target) & 1u) return true;
    return g.type == GT_CZ && ((R >> g.control) & 1u);
}

// Dependency graph: each gate waits for the previous gate on each of its qubits.
struct Dag {
    std::vector<int> npred;
    std::vector<std::array<int, 2>> succ;  // next gate on each of the gate's qubits, -1 none
};

Dag build_dag(const std::vector<GateRec>& g) {
    const int G = (int)g.size();
    Dag d;
    d.npred.assign(G, 0);
    d.succ.assign(G, {-1, -1});
    std::vector<int> last(32, -1);
    for (int i = 0; i < G; ++i) {
        const int qs[2] = {g[i].target, controlled(g[i].type) ? g[i].control : -1};
        for (int k = 0; k < 2; ++k) {
            const int q = qs[k];
            if (q < 0) continue;
            const int p = last[q];
            if (p >= 0) {
                d.npred[i]++;
                auto& sp = d.succ[p];
                (sp[0] < 0 ? sp[0] : sp[1]) = i;
            }
            last[q] = i;
        }
    }
    return d;
}

// Register set R holds every qubit of gate g (a dense block on R can absorb it).
bool holds(const GateRec& g, uint32_t R) {
    if (!((R >> g.target) & 1u)) return false;
    return !controlled(g.type) || ((R >> g.control) & 1u);
}

// Runs, with register set R, every gate that is or becomes runnable (lowest index
// first) and that R takes (fits: serves for the tile program, holds for dense blocks);
// frontier holds the runnable-by-dependencies gates R does not take.  With order null
// the state is restored afterwards (a trial); returns the gates run.
template <typename Fits>
int closure(const std::vector<GateRec>& g, const Dag& d, std::vector<int>& pend, std::vector<int>& frontier,
            uint32_t R, std::vector<int>* order, Fits fits) {
    std::priority_queue<int, std::vector<int>, std::greater<int>> heap;
    std::vector<int> rest, touched;
    for (int i : frontier) (fits(g[i], R) ? (void)heap.push(i) : rest.push_back(i));
    int n = 0;
    while (!heap.empty()) {
        const int i = heap.top();
        heap.pop();
        ++n;
        if (order) order->push_back(i);
        for (int s : d.succ[i]) {
            if (s < 0) continue;
            touched.push_back(s);
            if (--pend[s] == 0) (fits(g[s], R) ? (void)heap.push(s) : rest.push_back(s));
        }
    }
    if (order) {
        frontier = rest;
    } else {
        for (int s : touched) pend[s]++;
    }
    return n;
}

// Execution order and the register set of each segment of it (segments[k] covers
// order[seg_start[k] .. seg_start[k+1])).
void schedule(int N, int K, const std::vector<GateRec>& g, std::vector<int>& order, std::vector<uint32_t>& sets,
              std::vector<int>& seg_start) {
    const int G = (int)g.size();
    Dag d = build_dag(g);
    std::vector<int> pend = d.npred, frontier;
    for (int i = 0; i < G; ++i)
        if (pend[i] == 0) frontier.push_back(i);
    order.clear();
    sets.clear();
    seg_start.clear();
    while ((int)order.size() < G) {
        uint32_t best = 0;
        int best_n = -1;
        for (uint32_t R = 0; R < (1u << N); ++R) {
            if (__builtin_popcount(R) != K) continue;
            const int n = closure(g, d, pend, frontier, R, nullptr, serves);
            if (n > best_n) {
                best_n = n;
                best = R;
            }
        }
        seg_start.push_back((int)order.size());
        sets.push_back(best);
        closure(g, d, pend, frontier, best, &order, serves);
    }
}

// Dense-block partition: repeatedly the 4-qubit set under which the most gates run
// (each gate with all its qubits in the set, in dependency order), as the tile
// schedule picks register sets.  blocks[k] lists the gates of the k-th block in run
// order; sets[k] is its qubit set.
void schedule_blocks(int N, const std::vector<GateRec>& g, std::vector<std::vector<int>>& blocks,
                     std::vector<uint32_t>& sets) {
    const int G = (int)g.size();
    Dag d = build_dag(g);
    std::vector<int> pend = d.npred, frontier;
    for (int i = 0; i < G; ++i)
        if (pend[i] == 0) frontier.push_back(i);
    blocks.clear();
    sets.clear();
    int placed = 0;
    while (placed < G) {
        uint32_t best = 0;
        int best_n = -1;
        for (uint32_t R = 0; R < (1u << N); ++R) {
            if (__builtin_popcount(R) != 4) continue;
            const int n = closure(g, d, pend, frontier, R, nullptr, holds);
            if (n > best_n) {
                best_n = n;
                best = R;
            }
        }
        blocks.emplace_back();
        sets.push_back(best);
        closure(g, d, pend, frontier, best, &blocks.back(), holds);
        placed += (int)blocks.back().size();
    }
}

}  // namespace

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
    // The gates' parameter ranges must tile [0, P) without overlap, as
    // Gates_block::add_gate assigns them (Gates_block.cpp:2500-2525): a shared index
    // would take only one gate's gradient term, an uncovered one none.
    std::vector<uint8_t> cover((size_t)P, 0);
    for (size_t gi = 0; gi < gates.size(); ++gi) {
        const int np = param_num_of(gates[gi].type);
        if (np > 0 && gates[gi].param_start + np > P) {
            err = "parameter start index beyond the parameter count at gate " + std::to_string(gi);
            return false;
        }
        for (int k = 0; k < np; ++k) {
            if (cover[(size_t)gates[gi].param_start + k]++) {
                err = "parameter index " + std::to_string(gates[gi].param_start + k) + " shared by two gates (gate " +
                      std::to_string(gi) + ")";
                return false;
            }
        }
    }
    prog.P = P;

    // Execution order and register sets: scheduled up to 10 qubits (the tile
    // kernels' range), gate order with the next K distinct targets above.
    std::vector<int> order;
    std::vector<uint32_t> sets;
    std::vector<int> seg_start;
    if (N <= 10 && N > K) {
        schedule(N, K, gates, order, sets, seg_start);
    } else {
        for (int i = 0; i < (int)gates.size(); ++i) order.push_back(i);
        uint32_t R = (1u << K) - 1u;
        for (int i = 0; i < (int)gates.size(); ++i) {
            if (serves(gates[i], R) && i > 0) continue;
            if (!serves(gates[i], R)) {
                std::vector<int> want;
                for (int j = i; j < (int)gates.size() && (int)want.size() < K; ++j)
                    if (!contains(want, gates[j].target)) want.push_back(gates[j].target);
                for (int q = 0; q < N && (int)want.size() < K; ++q)
                    if (((R >> q) & 1u) && !contains(want, q)) want.push_back(q);
                R = 0;
                for (int q : want) R |= 1u << q;
            }
            sets.push_back(R);
            seg_start.push_back(i);
        }
        if (sets.empty()) {
            sets.push_back(R);
            seg_start.push_back(0);
        }
    }

    // Layout 0 holds the first segment's register set (ascending qubits in register
    // positions, the rest in lane positions).
    std::vector<int> pos, where(N);
    for (int q = 0; q < N; ++q)
        if ((sets[0] >> q) & 1u) pos.push_back(q);
    for (int q = 0; q < N; ++q)
        if (!((sets[0] >> q) & 1u)) pos.push_back(q);
    for (int j = 0; j < N; ++j) where[pos[j]] = j;
    prog.layouts.push_back(make_layout(N, K, pos));
    int cur = 0;
    prog.gate_slot.assign(gates.size(), -1);

    size_t seg = 0;
    for (size_t oi = 0; oi < order.size(); ++oi) {
        const GateRec& g = gates[order[oi]];
        const bool ctrl = controlled(g.type);
        int t = g.target, c = ctrl ? g.control : -1;
        if (seg + 1 < seg_start.size() && (int)oi == seg_start[seg + 1]) {
            ++seg;
            // remap: the segment's incoming qubits take the register positions of the
            // outgoing ones, which move to the incoming qubits' lane positions
            std::vector<int> incoming, free_reg;
            for (int q = 0; q < N; ++q)
                if (((sets[seg] >> q) & 1u) && where[q] >= K) incoming.push_back(q);
            for (int j = 0; j < K; ++j)
                if (!((sets[seg] >> pos[j]) & 1u)) free_reg.push_back(j);
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
        // CZ is symmetric in its two qubits: use whichever is in registers.
        if (g.type == GT_CZ && where[t] >= K && where[c] < K) std::swap(t, c);
        if (where[t] >= K) {
            err = "internal error: gate target outside the register set";
            return false;
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
            prog.gate_slot[order[oi]] = op.idx;
            prog.slots.push_back(s);
        }
        prog.ops.push_back(op);
    }
    return true;
}

FusionPlan plan_fusion(int N, const std::vector<GateRec>& gates, const std::vector<int>& gate_slot) {
    FusionPlan fp;
    const int kmax = N < 4 ? N : 4;
    fp.k = kmax;
    // one block: its qubits (padded to kmax with the lowest unused ones) and gates in order
    auto emit = [&](std::vector<int> cur, const std::vector<int>& members) {
        for (int q = 0; q < N && (int)cur.size() < kmax; ++q)
            if (std::find(cur.begin(), cur.end(), q) == cur.end()) cur.push_back(q);
        std::sort(cur.begin(), cur.end());
        std::array<int, 4> qs = {0, 0, 0, 0};
        for (int j = 0; j < kmax; ++j) qs[j] = cur[j];
        fp.first.push_back((int)fp.gates.size());
        fp.qubits.push_back(qs);
        auto local = [&](int q) { return (int)(std::find(cur.begin(), cur.end(), q) - cur.begin()); };
        for (int gi : members) {
            const GateRec& g = gates[gi];
            FusedRec r;
            r.pad = 0;
            r.t = (int8_t)local(g.target);
            r.c = -1;
            r.slot = gate_slot[gi];
            switch (g.type) {
            case GT_CNOT: r.kind = OP_X; r.c = (int8_t)local(g.control); break;
            case GT_CZ: r.kind = OP_Z; r.c = (int8_t)local(g.control); break;
            case GT_CRY: r.kind = OP_CU; r.c = (int8_t)local(g.control); break;
            default: r.kind = OP_U; break;
            }
            fp.gates.push_back(r);
            fp.src.push_back(gi);
        }
    };
    if (N >= 4 && N <= kFuseScheduleMaxQubits) {
        // gates on disjoint qubits commute: pack each block with every gate that can
        // run inside its qubit set (8-qubit 40-layer ansatz: 6 blocks where list order
        // needs 15; 10-qubit 60-layer: 9 where it needs 26)
        std::vector<std::vector<int>> blocks;
        std::vector<uint32_t> sets;
        schedule_blocks(N, gates, blocks, sets);
        for (size_t k = 0; k < blocks.size(); ++k) {
            std::vector<int> cur;
            for (int q = 0; q < N; ++q)
                if ((sets[k] >> q) & 1u) cur.push_back(q);
            emit(cur, blocks[k]);
        }
    } else {
        // list order: grow a block while the union of touched qubits stays <= kmax
        std::vector<int> cur, members;
        for (size_t i = 0; i < gates.size(); ++i) {
            const GateRec& g = gates[i];
            std::vector<int> q = {g.target};
            if (controlled(g.type)) q.push_back(g.control);
            std::vector<int> uni = cur;
            for (int x : q)
                if (std::find(uni.begin(), uni.end(), x) == uni.end()) uni.push_back(x);
            if ((int)uni.size() > kmax) {
                if (!members.empty()) emit(cur, members);
                members.clear();
                uni = q;
            }
            cur = uni;
            members.push_back((int)i);
        }
        if (!members.empty()) emit(cur, members);
    }
    fp.first.push_back((int)fp.gates.size());
    return fp;
}

}  // namespace sqh
