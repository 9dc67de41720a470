// sqh_program.h -- host-side compilation of a gate list into a register-tile program.
#pragma once
#include <array>
#include <string>
#include <vector>

#include "sqh_internal.h"

namespace sqh {

struct GateRec {
    int type, target, control, param_start;
};

struct Program {
    int N = 0;   // qubits
    int K = 0;   // register qubits
    int P = 0;   // parameters
    std::vector<Op> ops;
    std::vector<Layout> layouts;
    std::vector<Slot> slots;
    std::vector<int> gate_slot;  // per gate of the input list: its slot (kernel-table index), or -1
    int remaps = 0;
};

// Fusion of consecutive gates into dense blocks on at most 4 qubits (padded to 4
// qubits when the circuit has them, so every block runs on the matrix cores).
struct FusedRec {
    int8_t kind, t, c, pad;  // OpKind, local target, local control (-1 none)
    int32_t slot;            // parametric slot or -1
};

struct FusionPlan {
    std::vector<FusedRec> gates;
    std::vector<int> src;               // gate-list index of each entry of gates
    std::vector<int> first;             // block b covers gates[first[b] .. first[b+1])
    std::vector<std::array<int, 4>> qubits;  // sorted qubits of each block
    int k = 0;                          // qubits per block
};

// Up to this many qubits the blocks are packed by dependency scheduling (a search over
// the C(N, 4) qubit sets per block); above it, in gate-list order.
constexpr int kFuseScheduleMaxQubits = 16;

// gate_slot: the kernel-table index of each parametric gate (Program::gate_slot).
FusionPlan plan_fusion(int N, const std::vector<GateRec>& gates, const std::vector<int>& gate_slot);

// Validates the gate list (types, qubit indices, parameter ranges) the way the
// reference's gate constructors and Gates_block::add_gate do (U3.cpp:41-46,
// CRY.cpp:60-66, Gates_block.cpp:2500-2525) and compiles it.  Returns false with
// err set on a bad gate.
bool compile_program(int N, const std::vector<GateRec>& gates, Program& prog, std::string& err);

// Layout of a qubit placement: pos[j] is the qubit at position j (registers first).
Layout make_layout(int N, int K, const std::vector<int>& pos);

}  // namespace sqh
