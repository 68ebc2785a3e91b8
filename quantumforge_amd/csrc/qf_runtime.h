// Internal helpers shared by the runtime translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <string>
#include <vector>

#include "../../include/qf_hip.h"
#include "qf_kernels.h"
#include "qf_plan.h"

#define QF_HIP_TRY(expr)                                            \
    do {                                                            \
        hipError_t e_ = (expr);                                     \
        if (e_ != hipSuccess) return ::qfrt::hip_fail(e_, #expr);   \
    } while (0)

struct qf_plan {
    qf::Plan plan;
};

struct qf_backend {
    int32_t ordinal = 0;
    hipStream_t stream = nullptr;
    std::atomic<int> refs{1};
    qf::PlanOptions opts;
    // Stream-ordered scratch; reused after a stream synchronize.
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    double* d_out = nullptr;
    size_t d_out_count = 0;
    void* d_ops = nullptr;
    size_t d_ops_bytes = 0;
    void* d_dense = nullptr;
    size_t d_dense_bytes = 0;
    // Per-instance measurement outcomes / masks and collapse scales.
    void* meas_a = nullptr;
    size_t meas_a_bytes = 0;
    void* meas_b = nullptr;
    size_t meas_b_bytes = 0;
    // Measurement record of the last batched gradient / expectation call.
    int32_t profile = 0;
    qf_gradient_profile grad_prof{};
    // Device arena of the batch engines (QAOA waves): kept for the next call,
    // released by qf_backend_release_cached, at destroy, or when a device
    // allocation fails.
    void* arena = nullptr;
    size_t arena_bytes = 0;
    // Pinned host staging of the batch engines' tables and results (DMA at
    // full link rate; pageable copies are staged by the runtime in chunks).
    void* pinned = nullptr;
    size_t pinned_bytes = 0;
};

struct qf_state {
    qf_backend* be = nullptr;
    void* d = nullptr;
    uint32_t qubits = 0;
    int32_t precision = QF_C128;
    uint64_t dim = 0;
};

// A density matrix of `qubits` qubits: dim x dim complex128 entries, row
// major (a state vector of 2 * qubits qubits).
struct qf_density {
    qf_backend* be = nullptr;
    void* d = nullptr;
    uint32_t qubits = 0;
    uint64_t dim = 0;
};

// A batch of `count` states stored back to back.  Equal-width batches have
// one segment; a batch of different widths (qf_batch_upload /
// qf_batch_zero_uneven) is cut into segments of consecutive states of one
// width, and every batch operation runs segment by segment.  qubits is the
// widest state, dim its dimension (equal-width batches: every state's).
struct qf_batch {
    qf_backend* be = nullptr;
    void* d = nullptr;
    uint32_t qubits = 0;
    uint32_t count = 0;
    int32_t precision = QF_C128;
    uint64_t dim = 0;
    struct Seg {
        uint32_t qubits, first, count;
        uint64_t off;  // amplitude offset of the segment's first state
    };
    std::vector<Seg> segs;
    uint64_t total = 0;  // amplitudes of all states
    bool uneven() const { return segs.size() > 1; }
};

namespace qfrt {

// Extra inputs of a plan launch: the per-instance angle table (cos, sin of
// half angles, [count][nslots]) and, for the reverse sweep, the costate and
// the gradient partial buffer.
struct LaunchCtx {
    const double* cs = nullptr;
    uint32_t nslots = 0;
    void* lam = nullptr;
    double* partial = nullptr;
    // Dense-block passes: the plan's uploaded operand blob (dense_blob).
    const unsigned char* dense = nullptr;
    // Passes [pass_begin, pass_end) of the plan are launched.
    size_t pass_begin = 0, pass_end = SIZE_MAX;
};

qf_status fail(qf_status st, const std::string& msg);
qf::PlanOptions plan_options_from(const qf_plan_options* o);
qf_status hip_fail(hipError_t e, const char* what);
// cos and sin as two separate libm calls (the reference's f64::cos, f64::sin),
// out of line so that no compiler merges a pair into sincos, whose last bit
// can differ from cos's.
double host_cos(double x);
double host_sin(double x);
void gate_matrix(int gate, double theta, qf::cd m[4]);
bool is_rotation(int gate);
qf_status check_pair(uint32_t qubits, uint32_t a, uint32_t b);
qf_status check_circuit_plan(const qf::Plan& plan, const std::vector<qf::GateIn>& g, uint32_t qubits);
qf_status check_qubits(uint32_t qubits);
qf_status lower_ops(const qf_op* ops, size_t count, uint32_t qubits, std::vector<qf::GateIn>& out);
qf_status activate(qf_backend* be);
qf_status device_alloc(void** out, size_t bytes);
qf_status ensure(qf_backend* be, void** buf, size_t* cap, size_t bytes);
// The backend's arena of at least `bytes` (kept between calls; grown by a
// free and a new allocation).
qf_status arena(qf_backend* be, size_t bytes, void** out);
// The backend's pinned host buffer of at least `bytes` (kept between calls).
qf_status pinned_host(qf_backend* be, size_t bytes, void** out);
qf_status launch_check(const char* what);
size_t amp_bytes(int precision);
unsigned grid_for(uint64_t work);
qf_status run_plan(qf_backend* be, void* d, int precision, uint32_t qubits, uint32_t count, const qf::Plan& plan);
qf_status upload_plans(qf_backend* be, const std::vector<const qf::Plan*>& plans,
                       std::vector<const qf::TileOp*>& bases);
qf_status launch_plan(qf_backend* be, void* d, int precision, uint32_t qubits, uint32_t count, const qf::Plan& plan,
                      const qf::TileOp* d_ops, const LaunchCtx& ctx);
qf_status apply_gates(qf_backend* be, void* d, int precision, uint32_t qubits, uint32_t count,
                      const std::vector<qf::GateIn>& gates, const qf::PlanOptions& opt);
qf_status reduce(qf_backend* be, const void* d, const void* d2, int precision, uint64_t dim, uint32_t count, int kind,
                 uint32_t q, const std::vector<qf::PauliDev>& terms, std::vector<double>& out);
qf_backend* state_backend(const qf_state* st);
void backend_release(qf_backend* be);
// Noise-channel parameters (QF_CHANNEL_*): InvalidArgument, InvalidProbability.
qf_status check_channel(int32_t kind, double p0, double p1);
// One Kraus-trajectory step on `count` contiguous states, one uniform each.
qf_status trajectory_states(qf_backend* be, void* d, int precision, uint32_t qubits, uint32_t count, int32_t channel,
                            double p0, double p1, uint32_t q, const double* uniforms);

// A parameterized circuit as the gradient code sees it (qf_pop records).
struct PCirc {
    const qf_pop* ops = nullptr;
    size_t nops = 0;
    uint32_t nparams = 0, ninputs = 0;
    std::vector<int32_t> slot_param;  // rotation occurrence -> parameter
    std::vector<size_t> slot_op;      // rotation occurrence -> op index
};
qf_status validate_values(const PCirc& pc, const double* inputs, size_t inputs_len, const double* params,
                          size_t params_len);
qf_status validate_ops(PCirc& pc, uint32_t qubits);
std::vector<qf::GateIn> param_gates(const PCirc& pc, const double* inputs, bool reverse);
void fill_row(const PCirc& pc, const double* params, int64_t shift_slot, double delta, double* row,
              bool scaled = false);
std::vector<qf::PauliDev> sorted_terms(const std::vector<qf::PauliDev>& t);
void adjoint_gate_id(int gate, double theta, int* g, double* th);
qf_status pauli_terms(const qf_pauli_term* terms, size_t count, uint32_t qubits, std::vector<qf::PauliDev>& out);

}  // namespace qfrt
