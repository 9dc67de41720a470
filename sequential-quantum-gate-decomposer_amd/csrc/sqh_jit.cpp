// sqh_jit.cpp -- run-time specialised tile kernels.
//
// The generic tile kernels walk the compiled program op by op: every gate goes
// through a dispatch on (kind, target), its control mask is a run-time value, and the
// switch arms merge back into one register assignment, which costs register copies
// and per-pair selects on top of the 2x2 arithmetic.  For a circuit that is evaluated
// many times (optimiser loops, batched costs) the program is fixed, so this file
// emits it as straight-line HIP: each op becomes a call of the same device helpers
// with its target, control mask and layout offsets as literals, hiprtc compiles it
// for the device (with the library's own code-generation flags), and the module is
// cached per context and program.  Register-controlled CNOTs become register
// renaming, kernel-table loads get constant offsets, and nothing is dispatched.
#include <hip/hiprtc.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "sqh_context.h"
#include "sqh_solve.h"

namespace sqh {

extern const char* const kJitHeaderInternal;  // sqh_internal.h (generated at build time)
extern const char* const kJitHeaderDevice;    // sqh_device.h
extern const char* const kJitHeaderBfgsDevice;  // sqh_bfgs_device.h
extern const char* const kJitHeaderBfgsCore;    // sqh_bfgs_core.h
extern const char* const kJitHeaderSolve;       // sqh_solve.h
extern const char* const kJitHeaderSolveLoop;   // sqh_solve_loop.h

namespace {

int geo_k(int N) { return N < 4 ? N : 4; }

// Register mask of a control (same rule as ctrl_mask on the device), as source text.
std::string mask_expr(int N, const Op& op) {
    const int NR = 1 << geo_k(N);
    const unsigned all = NR >= 32 ? 0xffffffffu : ((1u << NR) - 1u);
    std::ostringstream o;
    if (op.cmode == CM_NONE) {
        o << all << "u";
    } else if (op.cmode == CM_LANE) {
        o << "(((s >> " << int(op.c) << ") & 1) ? " << all << "u : 0u)";
    } else {
        unsigned m = 0;
        for (int i = 0; i < NR; ++i) m |= (unsigned)((i >> op.c) & 1) << i;
        o << m << "u";
    }
    return o.str();
}

std::string sub_expr(int N, const Layout& l, const char* s = "s") {
    const int nsub = N - geo_k(N);
    if (nsub == 0) return "0";
    std::ostringstream o;
    for (int j = 0; j < nsub; ++j) {
        if (j) o << " | ";
        o << "(((" << s << " >> " << j << ") & 1) << " << l.sub_pos[j] << ")";
    }
    return o.str();
}

// Launch shape compiled into a kernel as constants (so lane and batch indexing
// divide by literals and hold no scalar registers for the sizes).
struct JitShape {
    int m = 0, m_pad = 0, in_stride = 0, out_stride = 0, trace_off = 0, ns = 0;
    long long in_bstride = 0, out_bstride = 0, kern_bstride = 0;
    bool operator==(const JitShape& b) const {
        return m == b.m && m_pad == b.m_pad && in_stride == b.in_stride && out_stride == b.out_stride &&
               trace_off == b.trace_off && ns == b.ns && in_bstride == b.in_bstride && out_bstride == b.out_bstride &&
               kern_bstride == b.kern_bstride;
    }
};

JitShape shape_of(const TileLaunch& t) {
    JitShape s;
    s.m = t.m;
    s.m_pad = t.m_pad;
    s.in_stride = t.in_stride;
    s.out_stride = t.out_stride;
    s.trace_off = t.trace_off;
    s.ns = t.ns;
    s.in_bstride = t.in_bstride;
    s.out_bstride = t.out_bstride;
    s.kern_bstride = t.kern_bstride;
    return s;
}

struct Gen {
    const Program& pr;
    int N, NR;
    JitShape sh;
    std::ostringstream o;
    Gen(const Program& p, const JitShape& s) : pr(p), N(p.N), NR(1 << geo_k(p.N)), sh(s) {}

    // Kernel-table loads run one parametric op ahead of their use: op p's kernel is
    // loaded into k<p> while the previous parametric op computes, and a scheduling
    // barrier after every op keeps the compiler from hoisting all loads to the top
    // (which runs out of scalar registers and spills them).
    std::vector<int> order_fwd() const {
        std::vector<int> v;
        for (int p = 0; p < (int)pr.ops.size(); ++p)
            if (pr.ops[p].kind == OP_U || pr.ops[p].kind == OP_CU) v.push_back(p);
        return v;
    }
    void declare_k(int p) { o << "  cpx<R> k" << p << "[4]; load_k(kb + " << 4 * pr.ops[p].idx << ", k" << p << ");\n"; }
    void declare_b(int p) {
        o << "  cpx<R> kb" << p << "[4]; load_k(kbb + " << 4 * pr.ops[p].idx << ", kb" << p << ");\n";
    }
    void barrier() { o << "  __builtin_amdgcn_sched_barrier(0);\n"; }

    // LDS exchange between two layouts; sv / cv name the lane's sub index and column.
    // A remap permutes rows within each column, so with rp_ > 1 the wavefront's columns
    // go through the buffer in rp_ groups (an LDS buffer 1/rp_ the slab's size).
    int rp_ = 1;
    // LDS slot of (row, column group cq): row * LW + (row / 16) * LPAD + cq (LW = CW / RP
    // columns, one padding slot per 16 rows to spread rows over the banks).  A row is a
    // lane part (sub-index bits) plus a register offset on disjoint bits, so the slot is
    // the lane's base plus a constant the compiler folds into the LDS instruction.
    int lw() const { return (64 >> (N - geo_k(N))) / rp_; }
    int slot_const(int off) const { return off * lw() + (off >> 4); }
    void remap(const char* v, int from, int to, const char* sv = "s", const char* cv = "cc") {
        if (N <= geo_k(N)) return;
        const Layout& a = pr.layouts[from];
        const Layout& b = pr.layouts[to];
        o << "  { const int so = " << sub_expr(N, a, sv) << ", sn = " << sub_expr(N, b, sv) << ";\n"
          << "    const int cq = " << cv << " % LW, ch = " << cv << " / LW; (void)ch;\n"
          << "    const int bo = ldsb(so, cq), bn = ldsb(sn, cq);\n";
        for (int pass = 0; pass < rp_; ++pass) {
            if (rp_ > 1) o << "    if (ch == " << pass << ") {\n";
            for (int i = 0; i < NR; ++i)
                o << "      lds[bo + " << slot_const(a.reg_off[i]) << "] = " << v << "[" << i << "];\n";
            if (rp_ > 1) o << "    }\n";
            o << "    __syncthreads();\n";
            if (rp_ > 1) o << "    if (ch == " << pass << ") {\n";
            for (int i = 0; i < NR; ++i)
                o << "      " << v << "[" << i << "] = lds[bn + " << slot_const(b.reg_off[i]) << "];\n";
            if (rp_ > 1) o << "    }\n";
            o << "    __syncthreads();\n";
        }
        o << "  }\n";
    }

    // Form of a parametric op's kernel (which entries are exact zeros).
    int form(const Op& op) const {
        switch (pr.slots[op.idx].type) {
        case GT_U3: return 1;                 // KF_U3
        case GT_RY: case GT_CRY: return 2;    // KF_REAL
        case GT_RZ: return 3;                 // KF_DIAG
        case GT_RX: return 4;                 // KF_RX
        default: return 0;
        }
    }

    // Forward execution of op p (not a remap) on tile v; a parametric op uses k<p>.
    void gate(const char* v, int p) {
        const Op& op = pr.ops[p];
        const int t = op.t;
        switch (op.kind) {
        case OP_U:
            o << "  pairs_apply_f<" << form(op) << ", " << t << ", false>(" << v << ", k" << p << ", 0u);\n";
            break;
        case OP_CU:
            o << "  pairs_apply_f<" << form(op) << ", " << t << ", true>(" << v << ", k" << p << ", "
              << mask_expr(N, op) << ");\n";
            if (zero_cu_)  // apply_derivate_to tables: this element differentiates this gate
                o << "  pairs_zero_off<" << t << ">(" << v << ", " << mask_expr(N, op) << ", zs == " << op.idx
                  << ");\n";
            break;
        case OP_X:
            o << "  pairs_swap<" << t << ">(" << v << ", " << mask_expr(N, op) << ");\n";
            break;
        case OP_Z:
            o << "  pairs_negate<" << t << ">(" << v << ", " << mask_expr(N, op) << ");\n";
            break;
        default:
            break;
        }
    }

    int forward(const char* v) {
        int cur = 0;
        const std::vector<int> par = order_fwd();
        size_t next = 0;  // next parametric op whose kernel is not loaded yet
        if (next < par.size()) declare_k(par[next++]);
        for (int p = 0; p < (int)pr.ops.size(); ++p) {
            const Op& op = pr.ops[p];
            if (op.kind == OP_REMAP) {
                const int a = op.idx >> 16, b = op.idx & 0xffff;
                remap(v, a, b);
                cur = b;
            } else {
                if ((op.kind == OP_U || op.kind == OP_CU) && next < par.size()) declare_k(par[next++]);
                gate(v, p);
            }
            barrier();
        }
        return cur;
    }

    void load(const char* v, const char* src, const char* stride) {
        const Layout& l0 = pr.layouts[0];
        o << "  { const int so = " << sub_expr(N, l0) << ";\n";
        for (int i = 0; i < NR; ++i)
            o << "    " << v << "[" << i << "] = li.valid ? " << src << "[(long long)(so + " << l0.reg_off[i] << ") * "
              << stride << " + li.col] : mk<R>(0, 0);\n";
        o << "  }\n";
    }

    void prologue(bool uni, bool f32, const char* name) {
        o << "#include \"sqh_device.h\"\nusing namespace sqh;\ntypedef " << (f32 ? "float" : "double")
          << " R;\nconstexpr int N = " << N << ";\ntypedef Geo<N> G;\n"
          << "constexpr int SM = " << sh.m << ", SMPAD = " << sh.m_pad << ", SINS = " << sh.in_stride
          << ", SOUTS = " << sh.out_stride << ", STR = " << sh.trace_off << ", SNS = " << sh.ns << ";\n"
          << "constexpr long long SINB = " << sh.in_bstride << "LL, SOUTB = " << sh.out_bstride
          << "LL, SKB = " << sh.kern_bstride << "LL;\n"
          << "constexpr int RP = " << rp_ << ", LW = G::CW / RP, LPAD = 1;\n"
          << "constexpr int LDSN = (N > G::K) ? G::D * LW + (G::D / 16) * LPAD : 1;\n"
          << "__device__ __forceinline__ int ldsb(int row, int cq) { return row * LW + (row >> 4) * LPAD + cq; }\n"
          << "extern \"C\" __global__ __launch_bounds__(64) void " << name << "(JitArgs a) {\n"
          << "  __shared__ cpx<R> lds[LDSN];\n"
          << "  int B = a.B;\n"
          << "  if (a.count) { B = *a.count; if (B <= 0) return; }\n"
          << "  if ((long long)blockIdx.x * G::CW / SMPAD >= B) return;\n"
          << "  const LaneInfo li = lane_info<N, " << (uni ? "true" : "false")
          << ">(SM, SMPAD, B, threadIdx.x, blockIdx.x);\n"
          << "  const int s = li.s, cc = li.cc;\n"
          << "  (void)s; (void)cc; (void)lds;\n"
          << "  const cpx<R>* in = (const cpx<R>*)a.in + (long long)li.b * SINB;\n"
          << "  if (a.count) in = (const cpx<R>*)a.in + (long long)a.run_target[a.elem_run[li.b]] * a.tgt_stride;\n"
          << "  if (a.in_index) in = (const cpx<R>*)a.in + (long long)a.in_index[li.b] * SINB;\n"
          << "  const cpx<R>* kb = (const cpx<R>*)a.kern + (long long)li.b * SKB;\n";
    }

    // Straight-line k_tile_forward (mode 0: store; 1: diagonal entries).  The store
    // kernel also serves apply_derivate_to (per-element kernel tables, a.zslot set).
    bool zero_cu_ = false;

    std::string forward_kernel(bool uni, bool f32, int mode) {
        {
            // two column groups per remap: half the LDS per wavefront (8 KB at complex128),
            // so LDS no longer caps the resident wavefronts below the register limit;
            // measured 1-2 % faster than one group and 6 % faster than four at 8 qubits
            const int cw = N > geo_k(N) ? 64 >> (N - geo_k(N)) : 64;
            rp_ = (N > geo_k(N) && cw % 2 == 0) ? 2 : 1;
        }
        prologue(uni, f32, "sqh_jit_kernel");
        if (mode == 0)
            for (const Op& op : pr.ops)
                if (op.kind == OP_CU) zero_cu_ = true;
        if (zero_cu_) o << "  const int zs = a.zslot ? a.zslot[li.b] : -1;\n";
        o << "  cpx<R> v[G::NR];\n";
        load("v", "in", "SINS");
        const int cur = forward("v");
        const Layout& lf = pr.layouts[cur];
        o << "  const int sf = " << sub_expr(N, lf) << ";\n";
        if (mode == 0) {
            o << "  cpx<R>* dst = (cpx<R>*)a.out + (long long)li.b * SOUTB;\n";
            for (int i = 0; i < NR; ++i)
                o << "  if (li.valid) dst[(long long)(sf + " << lf.reg_off[i] << ") * SOUTS + li.col] = v[" << i
                  << "];\n";
        } else {
            // the lane's diagonal entry (at most one register matches) picked with selects
            // and stored once: conditional stores per register let the compiler sink the
            // last gates into sixteen branches and spill scalar registers there
            o << "  const int tr = li.col + STR;\n  R dre = 0, dim = 0; bool hit = false;\n";
            for (int i = 0; i < NR; ++i)
                o << "  { const bool m = sf + " << lf.reg_off[i] << " == tr; dre = m ? v[" << i << "].re : dre; dim = m ? v["
                  << i << "].im : dim; hit = hit || m; }\n";
            o << "  if (li.valid && hit) { a.diag[2 * ((long long)li.b * SM + li.col)] = (double)dre; "
                 "a.diag[2 * ((long long)li.b * SM + li.col) + 1] = (double)dim; }\n";
        }
        o << "}\n";
        return o.str();
    }

    // Straight-line k_tile_grad: forward, diagonal and selector, adjoint sweep.
    bool uni_ = false;

    // The kernel holds two tiles (state and adjoint) and asks for two wavefronts per
    // SIMD (at most 256 registers): measured against one wavefront, the 8-qubit
    // batch-512 gradient runs 19 % faster and the 10-qubit one 33 % faster.
    std::string grad_kernel(bool uni, bool f32) {
        uni_ = uni;
        prologue(uni, f32, "__attribute__((amdgpu_waves_per_eu(2))) sqh_jit_kernel");
        o << "  cpx<R> x[G::NR], y[G::NR];\n";
        load("x", "in", "SINS");
        const int cur = forward("x");
        const Layout& lf = pr.layouts[cur];
        // diagonal entry by selects and one store (as in the forward kernel)
        o << "  { const int sf = " << sub_expr(N, lf) << ";\n    const int tr = li.col + STR;\n"
          << "    R dre = 0, dim = 0; bool any = false;\n";
        for (int i = 0; i < NR; ++i) {
            o << "    { const bool hit = li.valid && sf + " << lf.reg_off[i] << " == tr;\n"
              << "      dre = hit ? x[" << i << "].re : dre; dim = hit ? x[" << i << "].im : dim; any = any || hit;\n"
              << "      y[" << i << "] = mk<R>(hit ? (R)1 : (R)0, (R)0); }\n";
        }
        o << "    if (any) { a.diag[2 * ((long long)li.b * SM + li.col)] = (double)dre; "
             "a.diag[2 * ((long long)li.b * SM + li.col) + 1] = (double)dim; }\n";
        o << "  }\n"
          << "  const int lim = SMPAD < G::CW ? SMPAD : G::CW;\n"
          << "  const bool leader = li.bvalid && (li.s == 0) && (li.cc % lim == 0);\n"
          << "  const int wpb = SMPAD >= G::CW ? SMPAD / G::CW : 1;\n"
          << "  const long long item0 = (long long)blockIdx.x * G::CW + li.cc;\n"
          << "  const int wib = SMPAD >= G::CW ? (int)((item0 % SMPAD) / G::CW) : 0;\n"
          << "  double* mp = a.mpart + ((long long)li.b * wpb + wib) * (long long)SNS * 8;\n"
          // The adjoint sweep's remaps use the forward's LDS addresses in reverse; fresh
          // copies of the lane indices (through an empty asm statement) make the sweep
          // recompute them instead of the compiler keeping the forward's addresses live
          // across the whole sweep (about 200 registers at 8 qubits).
          << "  int sb = s, cb = cc;\n  __asm__ volatile(\"\" : \"+v\"(sb), \"+v\"(cb));\n"
          // ... and the sweep loads its kernel tables through an offset the compiler
          // cannot see is 0 (a.pad0), so it reloads them instead of keeping the
          // forward's loads live
          << "  const cpx<R>* kbb = kb + a.pad0;\n";
        std::vector<int> par = order_fwd();
        std::reverse(par.begin(), par.end());
        size_t next = 0;
        if (next < par.size()) {
            o << "  {\n";
            declare_b(par[next++]);
        }
        for (int p = (int)pr.ops.size() - 1; p >= 0; --p) {
            const Op& op = pr.ops[p];
            if (op.kind == OP_REMAP) {
                const int oa = op.idx >> 16, nb = op.idx & 0xffff;
                remap("x", nb, oa, "sb", "cb");
                remap("y", nb, oa, "sb", "cb");
            } else if (op.kind == OP_X || op.kind == OP_Z) {
                gate("x", p);
                gate("y", p);
            } else if (op.kind == OP_U || op.kind == OP_CU) {
                if (next < par.size()) declare_b(par[next++]);
                const bool masked = op.kind == OP_CU;
                const std::string mk = masked ? mask_expr(N, op) : std::string("0u");
                const char* mb = masked ? "true" : "false";
                o << "  { const cpx<R>* k = kb" << p << ";\n"
                  << "    cpx<R> kd[4] = {mk(k[0].re, -k[0].im), mk(k[2].re, -k[2].im), mk(k[1].re, -k[1].im), "
                     "mk(k[3].re, -k[3].im)};\n"
                  << "    cpx<R> kt[4] = {k[0], k[2], k[1], k[3]};\n"
                  << "    cpx<R> acc[4] = {mk<R>(0, 0), mk<R>(0, 0), mk<R>(0, 0), mk<R>(0, 0)};\n"
                  << "    const unsigned msk = " << mk << ";\n"
                  << "    pairs_apply_f<" << form(op) << ", " << int(op.t) << ", " << mb << ">(x, kd, msk);\n"
                  << "    pairs_corr<" << int(op.t) << ">(y, x, acc, msk, " << mb << ");\n"
                  << "    pairs_apply_f<" << form(op) << ", " << int(op.t) << ", " << mb << ">(y, kt, msk);\n"
                  << "    double r[8];\n"
                  << (uni_ ? "    for (int e = 0; e < 4; ++e) { r[2 * e] = wave_sum((double)acc[e].re); "
                             "r[2 * e + 1] = wave_sum((double)acc[e].im); }\n"
                           : "    for (int e = 0; e < 4; ++e) { r[2 * e] = seg_sum<G::CW>((double)acc[e].re, lim); "
                             "r[2 * e + 1] = seg_sum<G::CW>((double)acc[e].im, lim); }\n")
                  << "    if (leader) { double* d = mp + " << 8 * op.idx << "; for (int e = 0; e < 8; ++e) d[e] = r[e]; }\n"
                  << "  }\n";
            }
            barrier();
        }
        if (!par.empty()) o << "  }\n";
        o << "}\n";
        return o.str();
    }

    // The persistent solver (sqh_solve_loop.h) with the evaluation specialised to the
    // circuit: N <= 4, so one wavefront holds the whole matrix (one column per lane)
    // and no remaps occur.  jit_eval<R> is the straight-line forward pass, trace,
    // adjoint sweep and per-gate correlations of k_tile_grad for one run, reduced in
    // the wavefront into LDS, followed by the gradient from the derivative kernels.
    std::string solve_kernel() {
        o << "#include \"sqh_solve_loop.h\"\nusing namespace sqh;\nconstexpr int N = " << N
          << ";\ntypedef Geo<N> G;\nconstexpr int SM = " << sh.m << ", STR = " << sh.trace_off
          << ", SNS = " << sh.ns << ";\n"
          << "struct JitLane { bool valid; int col; };\n";
        o << "template <typename R>\n__device__ double jit_eval(const cpx<R>* in, const double* xr, double* gr, "
             "cpx<double>* kt_raw, double* mc, const Slot* slots, int variant) {\n"
          << "  const int lane = threadIdx.x;\n"
          << "  cpx<R>* kw = reinterpret_cast<cpx<R>*>(kt_raw);\n"
          << "  for (int sl = lane; sl < SNS; sl += 64) { const Slot q = slots[sl]; cpx<double> k[4]; "
             "dev_gate_kernel(q.type, xr + q.param_start, k); for (int e = 0; e < 4; ++e) kw[4 * sl + e] = cvt<R>(k[e]); }\n"
          << "  __syncthreads();\n"
          << "  JitLane li; li.col = lane; li.valid = lane < SM;\n"
          << "  const int s = 0, cc = lane; (void)s; (void)cc;\n"
          << "  const cpx<R>* kb = kw;\n"
          << "  cpx<R> x[G::NR], y[G::NR];\n";
        load("x", "in", "SM");
        const int cur = forward("x");
        const Layout& lf = pr.layouts[cur];
        o << "  double tre = 0.0, tim = 0.0;\n"
          << "  { const int sf = " << sub_expr(N, lf) << ";\n    const int tr = li.col + STR;\n";
        for (int i = 0; i < NR; ++i) {
            o << "    { const bool hit = li.valid && sf + " << lf.reg_off[i] << " == tr;\n"
              << "      if (hit) { tre += (double)x[" << i << "].re; tim += (double)x[" << i << "].im; }\n"
              << "      y[" << i << "] = mk<R>(hit ? (R)1 : (R)0, (R)0); }\n";
        }
        o << "  }\n  const cpx<R>* kbb = kb;\n";
        std::vector<int> par = order_fwd();
        std::reverse(par.begin(), par.end());
        size_t next = 0;
        if (next < par.size()) {
            o << "  {\n";
            declare_b(par[next++]);
        }
        for (int p = (int)pr.ops.size() - 1; p >= 0; --p) {
            const Op& op = pr.ops[p];
            if (op.kind == OP_X || op.kind == OP_Z) {
                gate("x", p);
                gate("y", p);
            } else if (op.kind == OP_U || op.kind == OP_CU) {
                if (next < par.size()) declare_b(par[next++]);
                const bool masked = op.kind == OP_CU;
                const std::string mk = masked ? mask_expr(N, op) : std::string("0u");
                const char* mb = masked ? "true" : "false";
                o << "  { const cpx<R>* k = kb" << p << ";\n"
                  << "    cpx<R> kd[4] = {mk(k[0].re, -k[0].im), mk(k[2].re, -k[2].im), mk(k[1].re, -k[1].im), "
                     "mk(k[3].re, -k[3].im)};\n"
                  << "    cpx<R> kt[4] = {k[0], k[2], k[1], k[3]};\n"
                  << "    cpx<R> acc[4] = {mk<R>(0, 0), mk<R>(0, 0), mk<R>(0, 0), mk<R>(0, 0)};\n"
                  << "    const unsigned msk = " << mk << ";\n"
                  << "    pairs_apply_f<" << form(op) << ", " << int(op.t) << ", " << mb << ">(x, kd, msk);\n"
                  << "    pairs_corr<" << int(op.t) << ">(y, x, acc, msk, " << mb << ");\n"
                  << "    pairs_apply_f<" << form(op) << ", " << int(op.t) << ", " << mb << ">(y, kt, msk);\n"
                  << "    double r[8];\n"
                  << "    for (int e = 0; e < 4; ++e) { r[2 * e] = wave_sum((double)acc[e].re); "
                     "r[2 * e + 1] = wave_sum((double)acc[e].im); }\n"
                  << "    if (lane == 0) { double* d = mc + " << 8 * op.idx << "; for (int e = 0; e < 8; ++e) d[e] = r[e]; }\n"
                  << "  }\n";
            }
            barrier();
        }
        if (!par.empty()) o << "  }\n";
        o << "  tre = wave_sum(tre);\n  tim = wave_sum(tim);\n  __syncthreads();\n"
          << "  for (int sl = lane; sl < SNS; sl += 64) {\n"
          << "    const Slot q = slots[sl];\n    const double* M = mc + 8 * sl;\n"
          << "    for (int j = 0; j < q.nparam; ++j) {\n"
          << "      cpx<double> dk[4]; dev_deriv_kernel(q.type, xr + q.param_start, j, dk);\n"
          << "      double kre = 0.0, kim = 0.0;\n"
          << "      for (int e = 0; e < 4; ++e) { kre += dk[e].re * M[2 * e] - dk[e].im * M[2 * e + 1]; "
             "kim += dk[e].re * M[2 * e + 1] + dk[e].im * M[2 * e]; }\n"
          << "      gr[q.param_start + j] = grad_of(variant, tre, tim, kre, kim, (double)SM);\n"
          << "    }\n  }\n  __syncthreads();\n"
          << "  return cost_of(variant, tre, tim, (double)SM);\n}\n";
        solve_entry();
        return o.str();
    }
    void solve_entry() {
        o << "extern \"C\" __global__ __launch_bounds__(64) void sqh_jit_kernel(SolveArgs a) {\n"
          << "  const int var = a.variant;\n"
          << "  solve_runs<1>(a,\n"
          << "    [=](const cpx<double>* U, const Op*, int, const Layout*, const Slot* sl, int, const double* xr, "
             "double* gr, cpx<double>* kt, double* mc, cpx<double>*) { return jit_eval<double>(U, xr, gr, kt, mc, sl, var); },\n"
          << "    [=](const cpx<float>* U, const Op*, int, const Layout*, const Slot* sl, int, const double* xr, "
             "double* gr, cpx<double>* kt, double* mc, cpx<float>*) { return jit_eval<float>(U, xr, gr, kt, mc, sl, var); });\n"
          << "}\n";
    }
};

}  // namespace

// One compiled kernel of the cache.
struct JitEntry {
    long long prog_version;
    int kind;  // 0: forward store, 1: forward diagonal, 2: gradient
    bool f32, uni;
    JitShape shape;
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;
};

struct JitCache {
    std::vector<JitEntry> entries;
    double compile_ms = 0.0;
    int compiles = 0;
    ~JitCache() {
        for (auto& e : entries)
            if (e.mod) (void)hipModuleUnload(e.mod);
    }
};

void jit_release(sqh_ctx* c) {
    delete c->jit_cache;
    c->jit_cache = nullptr;
}

namespace {

std::string jit_source(const Program& pr, int kind, bool f32, bool uni, const JitShape& sh) {
    Gen g(pr, sh);
    if (kind == kJitSolve) return g.solve_kernel();
    return kind == 2 ? g.grad_kernel(uni, f32) : g.forward_kernel(uni, f32, kind);
}

// hiprtc compile of a generated source for arch; code object in code, log on failure.
bool jit_compile(const std::string& src, const std::string& arch, std::vector<char>& code, std::string& log) {
    hiprtcProgram p = nullptr;
    const char* hs[] = {kJitHeaderInternal, kJitHeaderDevice, kJitHeaderBfgsDevice, kJitHeaderBfgsCore,
                        kJitHeaderSolve, kJitHeaderSolveLoop};
    const char* hn[] = {"sqh_internal.h", "sqh_device.h", "sqh_bfgs_device.h", "sqh_bfgs_core.h", "sqh_solve.h",
                        "sqh_solve_loop.h"};
    if (hiprtcCreateProgram(&p, src.c_str(), "sqh_jit.hip", 6, hs, hn) != HIPRTC_SUCCESS) {
        log = "hiprtcCreateProgram failed";
        return false;
    }
    const std::string a = "--offload-arch=" + arch;
    // the library's own code-generation flags (build_lib.py)
    const char* opts[] = {a.c_str(), "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                          "-packed-fp32-ops", "-mllvm", "-simplifycfg-sink-common=false"};
    const hiprtcResult rr = hiprtcCompileProgram(p, 9, opts);
    if (rr != HIPRTC_SUCCESS) {
        size_t ls = 0;
        (void)hiprtcGetProgramLogSize(p, &ls);
        log.assign(ls + 1, '\0');
        (void)hiprtcGetProgramLog(p, &log[0]);
        (void)hiprtcDestroyProgram(&p);
        return false;
    }
    size_t cs = 0;
    (void)hiprtcGetCodeSize(p, &cs);
    code.resize(cs);
    (void)hiprtcGetCode(p, code.data());
    (void)hiprtcDestroyProgram(&p);
    return true;
}

}  // namespace

// Compiles (once) and returns the kernel for the current program; nullptr with the
// context's error set on failure.
static hipFunction_t jit_get(sqh_ctx* c, int kind, bool f32, bool uni, const JitShape& sh, bool compile = true) {
    if (!compile) {  // cache lookup only (allocates nothing)
        if (!c->jit_cache) return nullptr;
        for (auto& e : c->jit_cache->entries)
            if (e.prog_version == c->prog_version && e.kind == kind && e.f32 == f32 && e.uni == uni && e.shape == sh)
                return e.fn;
        return nullptr;
    }
    if (!c->jit_cache) c->jit_cache = new JitCache();
    JitCache& jc = *c->jit_cache;
    for (auto& e : jc.entries)
        if (e.prog_version == c->prog_version && e.kind == kind && e.f32 == f32 && e.uni == uni && e.shape == sh)
            return e.fn;
    // Kernels of circuits the context no longer holds are unloaded once the stream
    // that ran them is idle.
    bool stale = false;
    for (auto& e : jc.entries) stale = stale || e.prog_version != c->prog_version;
    if (stale && hipStreamSynchronize(c->stream) == hipSuccess) {
        std::vector<JitEntry> keep;
        for (auto& e : jc.entries) {
            if (e.prog_version == c->prog_version) {
                keep.push_back(e);
            } else if (e.mod) {
                (void)hipModuleUnload(e.mod);
            }
        }
        jc.entries.swap(keep);
    }
    const auto t0 = std::chrono::steady_clock::now();
    hipDeviceProp_t prop;
    std::string arch = "gfx950";
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) {
        std::string name = prop.gcnArchName;
        arch = name.substr(0, name.find(':'));
    }
    std::vector<char> code;
    std::string log;
    if (!jit_compile(jit_source(c->prog, kind, f32, uni, sh), arch, code, log)) {
        fail(c, SQH_ERR_DEVICE, "run-time kernel compilation failed: %.400s", log.c_str());
        return nullptr;
    }
    JitEntry e;
    e.prog_version = c->prog_version;
    e.kind = kind;
    e.f32 = f32;
    e.uni = uni;
    e.shape = sh;
    hipError_t he = hipModuleLoadData(&e.mod, code.data());
    if (he == hipSuccess) he = hipModuleGetFunction(&e.fn, e.mod, "sqh_jit_kernel");
    if (he != hipSuccess) {
        if (e.mod) (void)hipModuleUnload(e.mod);
        fail(c, SQH_ERR_DEVICE, "loading the run-time kernel: %s", hipGetErrorString(he));
        return nullptr;
    }
    jc.entries.push_back(e);
    jc.compiles++;
    jc.compile_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return e.fn;
}

int jit_launch(sqh_ctx* c, const TileLaunch& t, int kind, hipStream_t s, bool compile) {
    const int cw = tile_columns_per_wave(t.N);
    const bool uni = t.N > 4 && t.m_pad % cw == 0;
    if (t.bi.run_prog) return fail(c, SQH_ERR_ARG, "run-time kernels serve one circuit");
    hipFunction_t fn = jit_get(c, kind, t.f32, uni, shape_of(t), compile);
    if (!fn) return compile ? SQH_ERR_DEVICE : kJitNotReady;
    JitArgs a;
    std::memset(&a, 0, sizeof(a));  // pad0 = 0: the gradient kernel's table offset
    a.in = t.in;
    a.in_bstride = t.in_bstride;
    a.in_stride = t.in_stride;
    a.m = t.m;
    a.m_pad = t.m_pad;
    a.B = t.B;
    a.out = t.out;
    a.out_bstride = t.out_bstride;
    a.out_stride = t.out_stride;
    a.trace_off = t.trace_off;
    a.ns = t.ns;
    a.kern = t.kern;
    a.kern_bstride = t.kern_bstride;
    a.diag = t.diag;
    a.mpart = t.mpart;
    a.count = t.bi.count;
    a.elem_run = t.bi.elem_run;
    a.run_target = t.bi.run_target;
    a.tgt_stride = t.bi.tgt_stride;
    a.zslot = t.zslot;
    a.in_index = t.in_index;
    const long long items = (long long)t.B * t.m_pad;
    const unsigned grid = (unsigned)((items + cw - 1) / cw);
    if (grid == 0) return SQH_OK;
    void* params[] = {&a};
    HIPTRY(c, hipModuleLaunchKernel(fn, grid, 1, 1, 64, 1, 1, 0, s, params, nullptr));
    return SQH_OK;
}

// Gradient kernels of long programs compile slowly (the straight-line adjoint sweep
// of the 142-op 8-qubit circuit takes about 12 s on the host, the 214-op 10-qubit one
// about 17 s); longer programs keep the generic gradient kernel.
constexpr size_t kJitMaxGradOps = 256;

bool jit_wants(sqh_ctx* c, int kind) {
    if (!c->jit || c->prog.N > kMaxTileQubits) return false;
    if (kind == kJitSolve) return c->prog.N <= 4 && correction_channels(c->variant) == 0;
    // the generated epilogues read the diagonal only
    if ((kind == 1 || kind == 2) && correction_channels(c->variant) > 0) return false;
    return kind != 2 || c->prog.ops.size() <= kJitMaxGradOps;
}

int jit_launch_solve(sqh_ctx* c, const SolveArgs& a, int grid, size_t shm, hipStream_t s) {
    JitShape sh;
    sh.m = a.m;
    sh.trace_off = a.trace_off;
    sh.ns = a.ns;
    hipFunction_t fn = jit_get(c, kJitSolve, false, false, sh);
    if (!fn) return SQH_ERR_DEVICE;
    SolveArgs arg = a;
    void* params[] = {&arg};
    HIPTRY(c, hipModuleLaunchKernel(fn, grid, 1, 1, 64, 1, 1, (unsigned)shm, s, params, nullptr));
    return SQH_OK;
}

int jit_prepare(sqh_ctx* c, const TileLaunch& t, int kind) {
    const int cw = tile_columns_per_wave(t.N);
    const bool uni = t.N > 4 && t.m_pad % cw == 0;
    return jit_get(c, kind, t.f32, uni, shape_of(t)) ? SQH_OK : SQH_ERR_DEVICE;
}

double jit_compile_ms(sqh_ctx* c, int* compiles) {
    if (!c->jit_cache) {
        if (compiles) *compiles = 0;
        return 0.0;
    }
    if (compiles) *compiles = c->jit_cache->compiles;
    return c->jit_cache->compile_ms;
}

}  // namespace sqh

using namespace sqh;

extern "C" int sqh_jit_source(int qbit_num, const sqh_gate* gates, int n_gates, int kind, int is_f32, int uniform,
                              char* buf, long long cap, long long* source_bytes) {
    if (n_gates < 0 || (n_gates > 0 && !gates) || kind < 0 || kind > kJitSolve || cap < 0) return SQH_ERR_ARG;
    std::vector<GateRec> g(n_gates);
    for (int i = 0; i < n_gates; ++i) g[i] = {gates[i].type, gates[i].target, gates[i].control, gates[i].param_start};
    Program pr;
    std::string err;
    if (!compile_program(qbit_num, g, pr, err) || pr.N > kMaxTileQubits || (kind == kJitSolve && pr.N > 4))
        return SQH_ERR_GATE;
    JitShape sh;  // the representative launch of sqh_jit_compile_check
    const int d = 1 << pr.N;
    sh.m = d;
    sh.m_pad = d;
    sh.in_stride = sh.out_stride = d;
    sh.ns = (int)pr.slots.size();
    sh.kern_bstride = 4LL * sh.ns;
    sh.out_bstride = (long long)d * d;
    const std::string src = jit_source(pr, kind, is_f32 != 0, uniform != 0, sh);
    if (source_bytes) *source_bytes = (long long)src.size();
    if (buf && cap > 0) {
        const size_t n = std::min<size_t>(src.size(), (size_t)cap - 1);
        std::memcpy(buf, src.data(), n);
        buf[n] = '\0';
    }
    return SQH_OK;
}

extern "C" int sqh_jit_compile_check(int qbit_num, const sqh_gate* gates, int n_gates, int kind, int is_f32,
                                     int uniform, char* log, int log_cap, long long* code_bytes,
                                     long long* source_bytes) {
    if (n_gates < 0 || (n_gates > 0 && !gates) || kind < 0 || kind > kJitSolve) return SQH_ERR_ARG;
    std::vector<GateRec> g(n_gates);
    for (int i = 0; i < n_gates; ++i) g[i] = {gates[i].type, gates[i].target, gates[i].control, gates[i].param_start};
    Program pr;
    std::string err;
    if (!compile_program(qbit_num, g, pr, err) || pr.N > kMaxTileQubits || (kind == kJitSolve && pr.N > 4)) {
        if (log && log_cap > 0) std::snprintf(log, (size_t)log_cap, "%s", err.c_str());
        return SQH_ERR_GATE;
    }
    JitShape sh;  // a representative launch: square target, one parameter vector per element
    const int d = 1 << pr.N;
    sh.m = d;
    sh.m_pad = d;
    sh.in_stride = sh.out_stride = d;
    sh.ns = (int)pr.slots.size();
    sh.kern_bstride = 4LL * sh.ns;
    sh.out_bstride = (long long)d * d;
    const std::string src = jit_source(pr, kind, is_f32 != 0, uniform != 0, sh);
    if (source_bytes) *source_bytes = (long long)src.size();
    std::vector<char> code;
    std::string lg;
    const bool ok = jit_compile(src, "gfx950", code, lg);
    if (code_bytes) *code_bytes = (long long)code.size();
    if (log && log_cap > 0) std::snprintf(log, (size_t)log_cap, "%s", lg.c_str());
    return ok ? SQH_OK : SQH_ERR_DEVICE;
}
