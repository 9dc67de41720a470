// sqh_internal.h -- types shared by the host-side circuit compiler and the HIP kernels.
//
// A circuit (an ordered gate list, Gates_block.h:44-1025) is compiled on the host
// into a "tile program": a list of ops executed by every wavefront on the column
// slab of the matrix it holds in registers.  Each lane holds NR = 2^K complex
// elements of one column; the K "register qubits" are the row-index bits that
// vary across those registers, the remaining N-K row bits select the lane
// (the "sub" index) inside the column's lane group.  A gate whose target is not a
// register qubit is preceded by a REMAP op that exchanges the slab through LDS into
// a layout where it is.
#pragma once
#ifndef __HIPCC_RTC__
#include <cstdint>
#else
using __hip_internal::int32_t;
using __hip_internal::uint32_t;
using __hip_internal::uint8_t;
using __hip_internal::int8_t;
#endif

namespace sqh {

// Reference gate_type values (Gate.h:39-51).
enum GateType : int32_t {
    GT_CZ = 4, GT_CNOT = 5, GT_U3 = 7, GT_RY = 8, GT_RX = 9, GT_RZ = 10, GT_CRY = 14
};

// Reference cost_function_type values (Optimization_Interface.h:43-45); ABS_TRACE
// (1 - |Tr|/m) has no reference counterpart.
enum CostVariant : int32_t {
    CV_FROBENIUS = 0, CV_FROBENIUS_C1 = 1, CV_FROBENIUS_C2 = 2, CV_HS = 3, CV_HS_C1 = 4, CV_HS_C2 = 5,
    CV_INFIDELITY = 9, CV_ABS_TRACE = 100
};

// Extra trace channels a cost variant reads beyond the diagonal: the single (1) and
// double (2) bit-flip sums of the arXiv:2210.09191 corrections.
constexpr int correction_channels(int variant) {
    return (variant == CV_FROBENIUS_C1 || variant == CV_HS_C1) ? 1
           : (variant == CV_FROBENIUS_C2 || variant == CV_HS_C2) ? 2 : 0;
}

enum OpKind : uint8_t {
    OP_U = 0,      // uncontrolled 2x2 from the kernel table (slot = idx)
    OP_CU = 1,     // controlled 2x2 (CRY), slot = idx
    OP_X = 2,      // CNOT: controlled row swap
    OP_Z = 3,      // CZ: controlled sign flip of the target-bit-set row
    OP_REMAP = 4,  // idx = (old_layout << 16) | new_layout
    OP_NOP = 5,    // padding around the program (never executed, only prefetched)
};

// Padding records on each side of the device copy of a program, so the kernels'
// look-ahead loads need no bounds checks.
constexpr int kOpPad = 2;

// Control modes: none, control qubit is a register bit (c = register position),
// control qubit is a lane bit (c = sub-index bit).
enum CtrlMode : uint8_t { CM_NONE = 0, CM_REG = 1, CM_LANE = 2 };

struct alignas(8) Op {
    uint8_t kind, t, cmode, c;
    int32_t idx;
};

// A layout: reg_off[i] is the row-index contribution of register i, sub_pos[j] the
// qubit carried by bit j of the lane's sub index.
struct alignas(16) Layout {
    int32_t reg_off[16];
    int32_t sub_pos[8];
};

// One parametric gate ("slot"): its type, first raw parameter, parameter count,
// and the index of its op in the program.
struct alignas(16) Slot {
    int32_t type, param_start, nparam, op;
};

// Program of one run when the runs of one launch carry different circuits on the
// same number of qubits: offsets into the launch's concatenated op, layout and slot
// arrays (each program's ops padded by kOpPad no-ops on both sides).
struct ProgRef {
    int32_t op_off, nops, lay_off, slot_off, ns, nlays, pad1, pad2;
};

// Optional indirection for batched optimisation: launch element j is run
// elem_run[j], whose target matrix is run_target[run] (targets tgt_stride complex
// entries apart); only the first *count elements are live.  All null: element j
// is batch element j on the shared input.
struct BatchIndex {
    const int* count = nullptr;
    const int* elem_run = nullptr;
    const int* run_target = nullptr;
    long long tgt_stride = 0;
    const ProgRef* run_prog = nullptr;  // per run; requires one element per wavefront
};

// Trace channels and gradient seed of a launch.  Channel k (0 <= k <= nch) of
// column c sums M[(c + off) ^ mask, c] over the masks of k set bits below 2^N
// (k = 0: the trace entry itself); diag holds channel k at offset k * chstride
// complex entries.  The gradient kernel's adjoint seed weights row r of column c
// by w[popcount(r ^ (c + off))]: w = (1, w1, w2) when seed is null, else the
// element's three complex weights seed[6 b ... 6 b + 5].
struct TraceSpec {
    int nch = 0;
    int pad = 0;
    long long chstride = 0;
    const double* seed = nullptr;
    double w1 = 0.0, w2 = 0.0;
};

// Arguments of a run-time compiled tile kernel (one struct, passed by value).
struct JitArgs {
    const void* in;
    long long in_bstride;
    int in_stride, m, m_pad, B;
    void* out;
    long long out_bstride;
    int out_stride, trace_off, ns, pad0;  // pad0 = 0 (the gradient kernel's opaque table offset)
    const void* kern;
    long long kern_bstride;
    double* diag;
    double* mpart;
    const int* count;
    const int* elem_run;
    const int* run_target;
    long long tgt_stride;
    const int* zslot;     // per element: the controlled slot whose control-0 block is zeroed, or -1
    const int* in_index;  // forward kernels: element b reads input in_index[b] (null: input b)
};

constexpr int kMaxTileQubits = 10;   // one column fits one wavefront up to 2^10 rows

inline int param_num_of(int type) {
    switch (type) {
    case GT_U3: return 3;
    case GT_RX: case GT_RY: case GT_RZ: case GT_CRY: return 1;
    case GT_CNOT: case GT_CZ: return 0;
    default: return -1;
    }
}

}  // namespace sqh
