// Host-side circuit planner: lowers a gate list to device passes.
//
// Two lowering modes:
//   * exact (fuse = 0): one full-state sweep per gate, in circuit order, with
//     the reference's unfused multiply/add sequence (src/lib.rs:1037-1039), so
//     complex128 results match the reference CPU path bit for bit;
//   * fused (fuse = 1): tile passes.  A tile is the set of 2^M amplitudes whose
//     indices differ only in M chosen qubits (the `low` contiguous qubits
//     0..low-1 plus up to M-low further qubits).  One workgroup loads a tile
//     into LDS, applies every scheduled gate whose qubits lie in the tile, and
//     writes it back: one HBM sweep for many gates.  Diagonal gates (Rz, Z, S,
//     T, CZ, controlled phase, ZZ rotations) are applied from the element's
//     global index, so they never need their qubits inside the tile.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace qf {

struct JitProgram;

struct cd {
    double re, im;
};

// Device op kinds (shared with the tile kernel).
enum TileOpKind : int32_t {
    TOP_MAT1 = 0,   // dense 2x2 on tile bit q0
    TOP_DIAG = 1,   // diagonal on up to 2 global qubits (g0, g1): d[bits(g1,g0)]
    TOP_CNOT = 2,   // permutation: control tile bit q0, target tile bit q1
    TOP_SWAP = 3,   // permutation on tile bits q0, q1
    TOP_MAT2 = 4,   // dense 4x4 on tile bits (q0 = low bit of the 2-bit index, q1 = high bit)
    TOP_ROT = 5,    // exp(-i theta P / 2), P = X/Y on tile bit q0, or Z on global qubit q0; axis in q1
    TOP_ROTZZ = 6,  // exp(-i theta Z_q0 Z_q1 / 2) on global qubits q0, q1
    TOP_PHASE = 7,  // register kernel: switch the register-bit set (q0 = packed tile bits, 4 bits each)
};

// Register-resident tile kernel ("phases"): each thread keeps 2^RB amplitudes
// in registers; the ops of a phase address register bits:
//   TOP_MAT1, TOP_ROT (X/Y), TOP_SWAP: q0 (and q1) are register indices;
//   TOP_MAT2: register indices q0 < q1;
//   TOP_CNOT: q0 = control as a global qubit, q1 = target register index;
//   TOP_DIAG, TOP_ROT (Z), TOP_ROTZZ: global qubits (bits taken from the
//   element's global index).

// Structure of a TOP_MAT1 matrix in the register kernel (TileOp::flags).
enum : int32_t {
    MAT_GENERAL = 0,  // dense complex 2x2
    MAT_REAL = 1,     // all entries real
    MAT_CROSS = 2,    // m00, m11 real; m01, m10 purely imaginary (Rx-like)
    MAT_SWAP = 3,     // exactly [[0, 1], [1, 0]]
};

// TileOp flags (rotations).
enum : int32_t {
    TF_NEGATE = 1,   // apply the inverse rotation (angle -theta)
    TF_RECORD = 2,   // before applying, record Im <lambda| G |psi> for this occurrence (reverse sweep)
};

// One op as the tile kernel reads it (uniform across the workgroup).
// Rotation angles are per instance: the kernel reads (cos(theta/2), sin(theta/2))
// of parameter `param` from the instance's row of the angle table.
struct TileOp {
    int32_t kind;
    int32_t q0, q1;   // tile bit positions (MAT1/CNOT/SWAP/MAT2, ROT X/Y) or global qubits (DIAG, ROT Z, ROTZZ)
    int32_t flags;
    int32_t param;    // rotation: column of the per-instance angle table
    int32_t slot;     // rotation: occurrence index (gradient partial column)
    int32_t pad[2];
    double m[32];     // MAT1: 4 complex; DIAG: 4 complex (index g1bit*2+g0bit); MAT2: 16 complex row major
};

// A diagonal gate applied before a dense block (DENSE passes): on the index
// bits a and b (a == b for a one-qubit diagonal) with factor d[bit(b) * 2 +
// bit(a)].  In `col` lists both qubits lie outside the block (one factor per
// column); in `blk` lists a is a position in the block and b a qubit outside it.
struct DiagPre {
    uint32_t a, b;
    cd d[4];
};

// One full-state pass.
struct Pass {
    enum Kind : int32_t { SWEEP_1Q = 0, SWEEP_2Q = 1, TILE = 2, SWEEP_DIAG = 3, DENSE = 4 } kind;
    // SWEEP_1Q: q0 target, m; SWEEP_2Q: op2 (0 CNOT, 1 CZ, 2 SWAP), q0, q1; SWEEP_DIAG: q0, q1, m as diag
    int32_t op2 = 0;
    uint32_t q0 = 0, q1 = 0;
    cd m[4]{};
    // TILE
    uint32_t tile_bits = 0;            // M
    uint32_t reg_bits = 0;             // > 0: register kernel with 2^reg_bits amplitudes per thread
    uint32_t low = 0;                  // contiguous low qubits
    std::vector<uint32_t> high;        // global qubit of tile bit low+k (ascending)
    std::vector<TileOp> ops;
    // Compiled passes only: the index bit each tile bit is stored to (empty:
    // the bit it was loaded from).  A permutation of the tile's bits; the
    // qubits stored to bits 0..low-1 are the next pass's contiguous ones.
    std::vector<uint32_t> store_phys;
    // DENSE: a fused k-qubit block (matrix-core pass).  block: its qubits in
    // ascending order (bit i of a block index is qubit block[i]); U: the
    // 2^k x 2^k block matrix, row major, composed in FP64 on the host;
    // pre_col / pre_blk: diagonal gates applied before U (see DiagPre).
    std::vector<uint32_t> block;
    std::vector<cd> U;
    std::vector<DiagPre> pre_col, pre_blk;
};

struct PlanOptions {
    int fuse = 1;
    uint32_t tile_qubits = 0;  // 0 = default for the precision
    uint32_t low_qubits = 0;   // 0 = default
    uint32_t max_block = 0;
    int kernel = 0;            // 0 = default, 1 = LDS op-by-op tile kernel, 2 = register kernel,
                               // 3 = compiled passes (register-kernel plan, one generated kernel per pass),
                               // 4 = dense k-qubit blocks on the matrix cores (k = max_block, 4 or 5)
    uint32_t reg_bits = 0;     // register kernel: amplitudes per thread = 2^reg_bits (0 = default)
    // Compiled passes may store qubits to other index bits than they were
    // loaded from (the plan's last pass restores the layout); such a plan
    // runs only as compiled kernels.  Off for plans that also run through
    // the interpreting kernels (the adjoint sweep's reverse plans).
    int relabel = 1;
    // 1: compiled passes below kCompileMinQubits as well (a large batch of
    // small states, where one launch still streams much memory).
    int compile = 0;
    // 1: among compiled plans with the fewest passes, prefer the one whose
    // heaviest pass carries the least arithmetic (passes capped at a share
    // of the circuit's arithmetic, so that work moves from issue-bound
    // passes into streaming-bound ones).
    int balance = 1;
    // 1: the run-time X and Y rotations of a compiled plan read angle rows of
    // (scale, ratio) instead of (cos, sin) (PassGen and fill_row): one FMA per
    // output component.  Set only where every such rotation lands in a
    // compiled pass (Plan::scaled_rotations).
    int scaled_rotations = 0;
};

// Circuits of at least this many qubits run compiled passes by default
// (smaller states are launch-bound and not worth a compile).
constexpr uint32_t kCompileMinQubits = 18;

struct PlanStats {
    uint64_t passes = 0, ops_in = 0, ops_applied = 0, dense_blocks = 0, max_ops_per_pass = 0, phases = 0;
};

// Kernel families a plan launch used (bit mask, PlanStats::kernels_run).
enum : uint32_t {
    KR_SWEEP = 1,     // one full-state sweep per gate (exact lowering)
    KR_LDS_TILE = 2,  // k_tile / k_tile_adjoint (LDS, op by op)
    KR_REG_TILE = 4,  // k_tile_reg (interpreting register kernel)
    KR_COMPILED = 8,  // generated per-pass kernels
    KR_MFMA = 16,     // dense-block matrix-core passes
};

// Where a plan came from; a run entry point accepts only its own kind.
enum PlanOrigin : int32_t { PLAN_CIRCUIT = 0, PLAN_PARAMETERIZED = 1, PLAN_DENSITY = 2 };

// One circuit gate in planner form.
struct GateIn {
    int32_t kind;      // 0 1q matrix, 1 CNOT, 2 CZ, 3 SWAP, 4 diag2 (controlled phase etc.), 5 rotation,
                       // 6 dense 4x4 on (a, b) (a density matrix's channel superoperator)
    uint32_t a, b;
    cd m[4];           // 1q: matrix; diag2: d[bits(b,a)]
    cd m4[16]{};       // kind 6: row-major 4x4, basis index bit(b) * 2 + bit(a)
    int32_t axis = 0;  // rotation: 0 X, 1 Y, 2 Z
    int32_t param = 0; // rotation: angle-table column
    int32_t slot = 0;  // rotation: occurrence index
    int32_t flags = 0; // rotation: TF_* flags
};

struct Plan {
    uint32_t qubits = 0;
    int precision = 0;
    std::vector<Pass> passes;
    PlanStats stats;
    // Hash of the planner gate list the plan was built from (circuit_key):
    // a run entry point lowers the ops it is given and refuses a plan whose
    // key differs, so a plan never runs in place of another circuit.
    std::string circuit_key;
    // Rotations whose angles come from a per-instance table at launch
    // (GateIn kind 5): such a plan needs that table and runs only through
    // the parameterized entry points.
    uint32_t runtime_rotations = 0;
    int32_t origin = PLAN_CIRCUIT;
    // KR_* mask of the kernels the last launch of this plan object used.
    mutable uint32_t kernels_run = 0;
    int compiled = 0;          // 1: register passes run as compiled kernels (2: explicitly requested)
    int relabeled = 0;         // passes with a store permutation (Pass::store_phys): compiled only
    int scaled_rotations = 0;  // X/Y rotations read (scale, ratio) rows: compiled only
    // Compiled programs per form (0: one state, 1: adjoint, psi and lambda)
    // and precision, attached on first launch (jit_state: 0 not tried, 1
    // compiled, -1 the compile failed and the register kernel runs).
    mutable std::shared_ptr<JitProgram> jit[2][2];
    mutable int jit_state[2][2] = {{0, 0}, {0, 0}};
};

// Builds a plan; returns an empty string on success or an error message.
std::string build_plan(const std::vector<GateIn>& gates, uint32_t qubits, int precision, const PlanOptions& opt,
                       Plan& out);

// Hash of a planner gate list and width (every field that changes what the
// gates do).
std::string circuit_key(const std::vector<GateIn>& gates, uint32_t qubits);

// JSON text of a plan (passes, tile sets, ops with exact matrices).
std::string describe_plan(const Plan& plan);

// Default tile geometry for a precision (amplitude bytes 16 or 8).
uint32_t default_tile_bits(int precision);
uint32_t default_low_bits(int precision);

}  // namespace qf
