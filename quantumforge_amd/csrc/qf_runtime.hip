// Host runtime of the state-vector engine: implements the C ABI of
// include/qf_hip.h over HIP.  One HIP stream per backend (default priority),
// 64-bit indexing throughout, validation on the host before any launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/qf_hip.h"
#include "qf_dense.h"
#include "qf_kernels.h"
#include "qf_regtile.h"
#include "qf_jit.h"
#include "qf_plan.h"
#include "qf_runtime.h"

using qf::cd;
using qf::cx;

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------
namespace {
thread_local std::string g_last_error;
}

namespace qfrt {

qf_status fail(qf_status st, const std::string& msg) {
    g_last_error = msg;
    return st;
}

qf_status hip_fail(hipError_t e, const char* what) {
    return fail(QF_DEVICE_ERROR, std::string("HIP error ") + hipGetErrorName(e) + " (" + std::to_string(int(e)) +
                                     "): " + hipGetErrorString(e) + " in " + what);
}

}  // namespace qfrt

using namespace qfrt;

extern "C" const char* qf_last_error_message(void) { return g_last_error.c_str(); }

extern "C" const char* qf_status_name(qf_status s) {
    switch (s) {
        case QF_OK: return "Ok";
        case QF_ZERO_QUBITS: return "ZeroQubits";
        case QF_QUBIT_OUT_OF_RANGE: return "QubitOutOfRange";
        case QF_INVALID_QUBIT_PAIR: return "InvalidQubitPair";
        case QF_NON_FINITE_ANGLE: return "NonFiniteAngle";
        case QF_STATE_TOO_LARGE: return "StateTooLarge";
        case QF_EMPTY_BATCH: return "EmptyBatch";
        case QF_INCOMPATIBLE_BATCH: return "IncompatibleBatch";
        case QF_INVALID_PARAMETER_COUNT: return "InvalidParameterCount";
        case QF_INVALID_INPUT_COUNT: return "InvalidInputCount";
        case QF_NON_FINITE_PARAMETER: return "NonFiniteParameter";
        case QF_OUT_OF_MEMORY: return "MemoryAllocationFailed";
        case QF_DEVICE_ERROR: return "Driver";
        case QF_REQUIRES_DENSITY_MATRIX: return "OperationRequiresDensityMatrix";
        case QF_REQUIRES_SHOT_EXECUTION: return "OperationRequiresShotExecution";
        case QF_INVALID_ARGUMENT: return "InvalidArgument";
        case QF_INCOMPATIBLE_DEVICE_STATE: return "IncompatibleDeviceState";
        case QF_STATE_SIZE_NOT_POWER_OF_TWO: return "StateSizeNotPowerOfTwo";
        case QF_INVALID_BATCH_SAMPLES: return "InvalidBatchSamples";
        case QF_MEMORY_BUDGET_EXCEEDED: return "MemoryBudgetExceeded";
        case QF_INVALID_FD_STEP: return "InvalidFiniteDifferenceStep";
        case QF_SIMULATION_SIZE_OVERFLOW: return "SimulationSizeOverflow";
        case QF_INVALID_PROBABILITY: return "InvalidProbability";
        case QF_INVALID_SHOT_COUNT: return "InvalidShotCount";
        case QF_CLASSICAL_BIT_OUT_OF_RANGE: return "ClassicalBitOutOfRange";
        case QF_NON_HERMITIAN_DENSITY_MATRIX: return "NonHermitianDensityMatrix";
        case QF_INVALID_SUBSYSTEM_SELECTION: return "InvalidSubsystemSelection";
        default: return "Unknown";
    }
}

#define HIP_TRY(expr)                                  \
    do {                                               \
        hipError_t e_ = (expr);                        \
        if (e_ != hipSuccess) return hip_fail(e_, #expr); \
    } while (0)

// ---------------------------------------------------------------------------
// Gate matrices on the host: Gate::matrix, src/lib.rs:2267-2333.
// ---------------------------------------------------------------------------
namespace qfrt {

__attribute__((noinline)) double host_cos(double x) { return std::cos(x); }
__attribute__((noinline)) double host_sin(double x) { return std::sin(x); }

void gate_matrix(int gate, double theta, cd m[4]) {
    const cd z{0, 0}, o{1, 0}, i{0, 1};
    const double s = 0.70710678118654752440;  // FRAC_1_SQRT_2
    switch (gate) {
        case QF_GATE_I: m[0] = o; m[1] = z; m[2] = z; m[3] = o; return;
        case QF_GATE_X: m[0] = z; m[1] = o; m[2] = o; m[3] = z; return;
        case QF_GATE_Y: m[0] = z; m[1] = cd{0, -1}; m[2] = i; m[3] = z; return;
        case QF_GATE_Z: m[0] = o; m[1] = z; m[2] = z; m[3] = cd{-1, 0}; return;
        case QF_GATE_H: m[0] = cd{s, 0}; m[1] = cd{s, 0}; m[2] = cd{s, 0}; m[3] = cd{-s, 0}; return;
        case QF_GATE_S: m[0] = o; m[1] = z; m[2] = z; m[3] = i; return;
        case QF_GATE_SDG: m[0] = o; m[1] = z; m[2] = z; m[3] = cd{0, -1}; return;
        case QF_GATE_T: m[0] = o; m[1] = z; m[2] = z; m[3] = cd{s, s}; return;
        case QF_GATE_TDG: m[0] = o; m[1] = z; m[2] = z; m[3] = cd{s, -s}; return;
        case QF_GATE_RX: {
            const double c = host_cos(theta * 0.5), si = host_sin(theta * 0.5);
            m[0] = cd{c, 0}; m[1] = cd{0, -si}; m[2] = cd{0, -si}; m[3] = cd{c, 0};
            return;
        }
        case QF_GATE_RY: {
            const double c = host_cos(theta * 0.5), si = host_sin(theta * 0.5);
            m[0] = cd{c, 0}; m[1] = cd{-si, 0}; m[2] = cd{si, 0}; m[3] = cd{c, 0};
            return;
        }
        default: {  // RZ
            const double half = theta * 0.5;
            m[0] = cd{host_cos(half), -host_sin(half)}; m[1] = z; m[2] = z;
            m[3] = cd{host_cos(half), host_sin(half)};
            return;
        }
    }
}

bool is_rotation(int gate) { return gate == QF_GATE_RX || gate == QF_GATE_RY || gate == QF_GATE_RZ; }

qf::PlanOptions plan_options_from(const qf_plan_options* o) {
    qf::PlanOptions po;
    if (o) {
        po.fuse = o->fuse;
        po.tile_qubits = o->tile_qubits;
        po.low_qubits = o->low_qubits;
        po.max_block = o->max_block;
        po.kernel = o->kernel;
        po.reg_bits = o->reg_bits;
    }
    return po;
}

}  // namespace qfrt

// ---------------------------------------------------------------------------
// Handles
// ---------------------------------------------------------------------------
namespace qfrt {

qf_backend* state_backend(const qf_state* st) { return st->be; }

// Backends holding an arena, so that a failed allocation can release them.
std::mutex g_arena_mu;
std::vector<qf_backend*> g_arena_owners;

void arena_free(qf_backend* be) {
    if (!be->arena) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(be->ordinal);
    (void)hipStreamSynchronize(be->stream);
    (void)hipFree(be->arena);
    (void)hipSetDevice(cur);
    be->arena = nullptr;
    be->arena_bytes = 0;
}

// Releases every backend's arena; true when one was held.
bool release_arenas() {
    std::lock_guard<std::mutex> lk(g_arena_mu);
    const bool any = !g_arena_owners.empty();
    for (qf_backend* be : g_arena_owners) arena_free(be);
    g_arena_owners.clear();
    return any;
}

qf_status arena(qf_backend* be, size_t bytes, void** out) {
    std::lock_guard<std::mutex> lk(g_arena_mu);
    if (be->arena && be->arena_bytes >= bytes) {
        *out = be->arena;
        return QF_OK;
    }
    arena_free(be);
    g_arena_owners.erase(std::remove(g_arena_owners.begin(), g_arena_owners.end(), be), g_arena_owners.end());
    hipError_t e = hipMalloc(&be->arena, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        be->arena = nullptr;
        // Other backends' arenas may hold the memory.
        for (qf_backend* o : g_arena_owners) arena_free(o);
        g_arena_owners.clear();
        e = hipMalloc(&be->arena, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            be->arena = nullptr;
            return fail(QF_OUT_OF_MEMORY, "could not reserve " + std::to_string(bytes) + " bytes of device memory");
        }
    }
    be->arena_bytes = bytes;
    g_arena_owners.push_back(be);
    *out = be->arena;
    return QF_OK;
}

qf_status pinned_host(qf_backend* be, size_t bytes, void** out) {
    if (be->pinned && be->pinned_bytes >= bytes) {
        *out = be->pinned;
        return QF_OK;
    }
    if (be->pinned) {
        HIP_TRY(hipStreamSynchronize(be->stream));
        (void)hipHostFree(be->pinned);
        be->pinned = nullptr;
        be->pinned_bytes = 0;
    }
    if (hipHostMalloc(&be->pinned, bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        be->pinned = nullptr;
        return fail(QF_OUT_OF_MEMORY, "could not allocate " + std::to_string(bytes) + " bytes of pinned host memory");
    }
    be->pinned_bytes = bytes;
    *out = be->pinned;
    return QF_OK;
}

void backend_release(qf_backend* be) {
    if (!be) return;
    if (be->refs.fetch_sub(1) == 1) {
        {
            std::lock_guard<std::mutex> lk(g_arena_mu);
            arena_free(be);
            g_arena_owners.erase(std::remove(g_arena_owners.begin(), g_arena_owners.end(), be),
                                 g_arena_owners.end());
        }
        (void)hipSetDevice(be->ordinal);
        (void)hipStreamSynchronize(be->stream);
        if (be->scratch) (void)hipFree(be->scratch);
        if (be->d_out) (void)hipFree(be->d_out);
        if (be->d_ops) (void)hipFree(be->d_ops);
        if (be->d_dense) (void)hipFree(be->d_dense);
        if (be->meas_a) (void)hipFree(be->meas_a);
        if (be->meas_b) (void)hipFree(be->meas_b);
        if (be->pinned) (void)hipHostFree(be->pinned);
        (void)hipStreamDestroy(be->stream);
        delete be;
    }
}

qf_status activate(qf_backend* be) {
    HIP_TRY(hipSetDevice(be->ordinal));
    return QF_OK;
}

// Grows a backend-owned device buffer; waits for the stream first so no
// in-flight kernel still reads the old contents.
qf_status ensure(qf_backend* be, void** buf, size_t* cap, size_t bytes) {
    if (*cap >= bytes && *buf) return QF_OK;
    HIP_TRY(hipStreamSynchronize(be->stream));
    if (*buf) {
        HIP_TRY(hipFree(*buf));
        *buf = nullptr;
        *cap = 0;
    }
    size_t want = bytes < 4096 ? 4096 : bytes;
    hipError_t e = hipMalloc(buf, want);
    if (e != hipSuccess) {
        *buf = nullptr;
        (void)hipGetLastError();
        return fail(QF_OUT_OF_MEMORY, "could not allocate " + std::to_string(want) + " bytes of device scratch");
    }
    *cap = want;
    return QF_OK;
}

size_t amp_bytes(int precision) { return precision == QF_C128 ? 16 : 8; }

// Size checks of StateVector::zero_with_memory_budget, src/lib.rs:940-966,
// with 64-bit device indexing (states up to 2^40 amplitudes).
qf_status check_qubits(uint32_t qubits) {
    if (qubits == 0) return fail(QF_ZERO_QUBITS, "a state vector must contain at least one qubit");
    if (qubits > 40) return fail(QF_STATE_TOO_LARGE, "state vector is too large for the HIP kernels (more than 40 qubits)");
    return QF_OK;
}

qf_status device_alloc(void** out, size_t bytes) {
    hipError_t e = hipMalloc(out, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        // A batch engine's arena kept from an earlier call may hold the memory.
        if (release_arenas()) {
            e = hipMalloc(out, bytes);
            if (e == hipSuccess) return QF_OK;
            (void)hipGetLastError();
        }
        *out = nullptr;
        return fail(QF_OUT_OF_MEMORY, "could not reserve " + std::to_string(bytes) +
                                          " bytes of device memory for the dense quantum state");
    }
    return QF_OK;
}

unsigned grid_for(uint64_t work) {
    uint64_t g = (work + qf::kThreads - 1) / qf::kThreads;
    if (g > (uint64_t(1) << 20)) g = uint64_t(1) << 20;  // grid-stride beyond 2^28 threads
    return unsigned(g ? g : 1);
}

qf_status launch_check(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, what);
    return QF_OK;
}

// Validates a circuit op list; converts unitary ops to planner gates.
// Errors follow the first offending op in circuit order, with the variants of
// Circuit::run / CudaBackend::run_circuit (src/lib.rs:2786-2802, src/cuda.rs:1050-1083).
qf_status lower_ops(const qf_op* ops, size_t count, uint32_t qubits, std::vector<qf::GateIn>& out) {
    out.clear();
    out.reserve(count);
    for (size_t i = 0; i < count; ++i) {
        const qf_op& op = ops[i];
        qf::GateIn g{};
        switch (op.kind) {
            case 0:
                if (op.gate < QF_GATE_I || op.gate > QF_GATE_RZ)
                    return fail(QF_INVALID_ARGUMENT, "unknown gate id " + std::to_string(op.gate));
                if (op.a >= qubits)
                    return fail(QF_QUBIT_OUT_OF_RANGE, "qubit index " + std::to_string(op.a) +
                                                           " is out of range for " + std::to_string(qubits) + " qubits");
                if (is_rotation(op.gate) && !std::isfinite(op.theta))
                    return fail(QF_NON_FINITE_ANGLE, "rotation angle must be finite");
                g.kind = 0;
                g.a = op.a;
                gate_matrix(op.gate, op.theta, g.m);
                break;
            case 1: case 2: case 3: {
                qf_status st = check_pair(qubits, op.a, op.b);
                if (st) return st;
                g.kind = op.kind;
                g.a = op.a;
                g.b = op.b;
                break;
            }
            case 7: {
                qf_status st = check_pair(qubits, op.a, op.b);
                if (st) return st;
                if (!std::isfinite(op.theta)) return fail(QF_NON_FINITE_ANGLE, "rotation angle must be finite");
                g.kind = 4;
                g.a = op.a;
                g.b = op.b;
                g.m[0] = g.m[1] = g.m[2] = cd{1, 0};
                g.m[3] = cd{host_cos(op.theta), host_sin(op.theta)};
                break;
            }
            case 4: return fail(QF_REQUIRES_DENSITY_MATRIX, "this circuit operation requires density-matrix simulation");
            case 5: case 6:
                return fail(QF_REQUIRES_SHOT_EXECUTION,
                            "measurement and classically conditioned circuit operations require shot execution");
            default: return fail(QF_INVALID_ARGUMENT, "unknown op kind " + std::to_string(op.kind));
        }
        out.push_back(g);
    }
    return QF_OK;
}

// check_pair, src/lib.rs:1124-1141 (the CPU's variants; see the header).
qf_status check_pair(uint32_t qubits, uint32_t a, uint32_t b) {
    if (a >= qubits)
        return fail(QF_QUBIT_OUT_OF_RANGE,
                    "qubit index " + std::to_string(a) + " is out of range for " + std::to_string(qubits) + " qubits");
    if (b >= qubits)
        return fail(QF_QUBIT_OUT_OF_RANGE,
                    "qubit index " + std::to_string(b) + " is out of range for " + std::to_string(qubits) + " qubits");
    if (a == b)
        return fail(QF_INVALID_QUBIT_PAIR, "two-qubit operation requires distinct qubits, got " + std::to_string(a) +
                                               " and " + std::to_string(b));
    return QF_OK;
}

template <class R>
qf_status launch_pass(qf_backend* be, cx<R>* d, uint32_t qubits, uint32_t count, const qf::Pass& p,
                      const qf::TileOp* d_ops, const LaunchCtx& ctx) {
    const uint64_t dim = uint64_t(1) << qubits;
    const uint64_t total = dim * count;
    hipStream_t st = be->stream;
    switch (p.kind) {
        case qf::Pass::DENSE: return fail(QF_INVALID_ARGUMENT, "dense pass without uploaded operands");
        case qf::Pass::SWEEP_1Q:
            hipLaunchKernelGGL(qf::k_sweep_1q<R>, dim3(grid_for(total / 2)), dim3(qf::kThreads), 0, st, d, total / 2,
                               p.q0, p.m[0], p.m[1], p.m[2], p.m[3]);
            return launch_check("k_sweep_1q");
        case qf::Pass::SWEEP_2Q:
            hipLaunchKernelGGL(qf::k_sweep_2q<R>, dim3(grid_for(total / 4)), dim3(qf::kThreads), 0, st, d, total / 4,
                               p.q0, p.q1, p.op2);
            return launch_check("k_sweep_2q");
        case qf::Pass::SWEEP_DIAG:
            hipLaunchKernelGGL(qf::k_sweep_diag<R>, dim3(grid_for(total)), dim3(qf::kThreads), 0, st, d, total,
                               int(p.q0), int(p.q1), p.m[0], p.m[1], p.m[2], p.m[3]);
            return launch_check("k_sweep_diag");
        case qf::Pass::TILE: {
            qf::TileArgs a{};
            a.n = qubits;
            a.M = p.tile_bits;
            a.low = p.low;
            a.log_blocks = qubits - p.tile_bits;
            a.nops = uint32_t(p.ops.size());
            a.nslots = ctx.nslots;
            for (size_t k = 0; k < p.high.size(); ++k) a.high[k] = p.high[k];
            const uint32_t gx_log0 = a.log_blocks < 16 ? a.log_blocks : 16;
            const dim3 grid0(1u << gx_log0, count << (a.log_blocks - gx_log0));
            if (p.reg_bits) {
                const int ns = ctx.lam ? 2 : 1;
                const size_t lds = size_t(ns) << p.tile_bits;
                const size_t bytes = lds * sizeof(cx<R>);
                const dim3 block(1u << (p.tile_bits - p.reg_bits));
                const void* fn = nullptr;
#define QF_REG_CASE(RBV, NSV)                                                                          \
    if (p.reg_bits == RBV && ns == NSV) {                                                              \
        fn = reinterpret_cast<const void*>(&qf::k_tile_reg<R, RBV, NSV>);                              \
        if (bytes > 64 * 1024)                                                                         \
            HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)));  \
        hipLaunchKernelGGL((qf::k_tile_reg<R, RBV, NSV>), grid0, block, bytes, st, d,                  \
                           static_cast<cx<R>*>(ctx.lam), a, d_ops, ctx.cs, ctx.partial);               \
    }
                QF_REG_CASE(3, 1)
                QF_REG_CASE(3, 2)
                QF_REG_CASE(4, 1)
                QF_REG_CASE(4, 2)
#undef QF_REG_CASE
                if (!fn) return fail(QF_INVALID_ARGUMENT, "unsupported register-kernel shape");
                return launch_check("k_tile_reg");
            }
            const uint32_t nh = p.tile_bits - p.low;
            const int tiles = ctx.lam ? 2 : 1;
            const size_t lds =
                tiles * (size_t(1) << p.tile_bits) * sizeof(cx<R>) + (size_t(1) << nh) * sizeof(uint64_t);
            const void* fn = ctx.lam ? reinterpret_cast<const void*>(&qf::k_tile_adjoint<R>)
                                     : reinterpret_cast<const void*>(&qf::k_tile<R>);
            if (lds > 64 * 1024)
                HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
            const uint32_t gx_log = a.log_blocks < 16 ? a.log_blocks : 16;
            const dim3 grid(1u << gx_log, count << (a.log_blocks - gx_log));
            if (ctx.lam) {
                hipLaunchKernelGGL(qf::k_tile_adjoint<R>, grid, dim3(qf::kThreads), lds, st, d,
                                   static_cast<cx<R>*>(ctx.lam), a, d_ops, ctx.cs, ctx.partial);
                return launch_check("k_tile_adjoint");
            }
            hipLaunchKernelGGL(qf::k_tile<R>, grid, dim3(qf::kThreads), lds, st, d, a, d_ops, ctx.cs);
            return launch_check("k_tile");
        }
    }
    return fail(QF_INVALID_ARGUMENT, "unknown pass kind");
}

// Copies the tile ops of several plans into the backend's op buffer (one
// upload); bases[i] is the device address of plan i's first op.
qf_status upload_plans(qf_backend* be, const std::vector<const qf::Plan*>& plans,
                       std::vector<const qf::TileOp*>& bases) {
    std::vector<qf::TileOp> host;
    std::vector<size_t> start;
    for (const qf::Plan* pl : plans) {
        start.push_back(host.size());
        for (const auto& p : pl->passes) host.insert(host.end(), p.ops.begin(), p.ops.end());
    }
    bases.assign(plans.size(), nullptr);
    if (host.empty()) return QF_OK;
    // The buffer may still be read by earlier launches: wait before reuse.
    HIP_TRY(hipStreamSynchronize(be->stream));
    qf_status s = ensure(be, &be->d_ops, &be->d_ops_bytes, host.size() * sizeof(qf::TileOp));
    if (s) return s;
    HIP_TRY(hipMemcpy(be->d_ops, host.data(), host.size() * sizeof(qf::TileOp), hipMemcpyHostToDevice));
    for (size_t i = 0; i < plans.size(); ++i) bases[i] = static_cast<const qf::TileOp*>(be->d_ops) + start[i];
    return QF_OK;
}

// ---------------------------------------------------------------------------
// Compiled passes (qf_jit.h): one program per plan and precision, loaded as
// modules once per device.
// ---------------------------------------------------------------------------
namespace {
std::mutex g_jit_mu;
// Loaded programs per (program key, device).  Beyond kLoadedPrograms the
// oldest is unloaded (after the device drains, since a launch may still run
// it); a plan that needs it again reloads its code objects.
struct Loaded {
    std::vector<hipModule_t> mods;
    std::vector<hipFunction_t> fns;
};
std::map<std::pair<std::string, int>, Loaded> g_jit_fns;
std::deque<std::pair<std::string, int>> g_jit_order;
constexpr size_t kLoadedPrograms = 256;
}  // namespace

// The plan's compiled program for `precision` (compiled on first use), or
// nullptr when the plan runs the register kernel.  A failed compile is an
// error when compiled passes were requested explicitly or the plan runs only
// compiled (relabeled passes, rotations read as (scale, ratio)); otherwise it
// is reported once and the register kernel runs the same plan.
qf_status plan_program(const qf::Plan& plan, int precision, int ns, const qf::JitProgram** out) {
    *out = nullptr;
    if (!plan.compiled) return QF_OK;
    std::lock_guard<std::mutex> lk(g_jit_mu);
    const int pr = precision == QF_C128 ? 0 : 1, f = ns == 2 ? 1 : 0;
    if (plan.jit_state[f][pr] == 0) {
        std::string err;
        plan.jit[f][pr] = qf::jit_compile_plan(plan, pr, &err, ns);
        plan.jit_state[f][pr] = plan.jit[f][pr] ? 1 : -1;
        if (!plan.jit[f][pr]) {
            if (plan.compiled == 2 || plan.relabeled || plan.scaled_rotations) {
                plan.jit_state[f][pr] = 0;
                return fail(QF_INVALID_ARGUMENT, "pass compiler: " + err);
            }
            std::fprintf(stderr, "quantumforge_amd: pass compiler failed, running the register kernel: %s\n",
                         err.c_str());
        }
    }
    *out = plan.jit[f][pr].get();
    return QF_OK;
}

qf_status jit_functions(const qf::JitProgram& prog, const std::vector<hipFunction_t>** out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_jit_mu);
    const auto key = std::make_pair(prog.key, dev);
    auto it = g_jit_fns.find(key);
    if (it == g_jit_fns.end()) {
        if (g_jit_order.size() >= kLoadedPrograms) {
            HIP_TRY(hipDeviceSynchronize());
            while (g_jit_order.size() >= kLoadedPrograms) {
                auto old = g_jit_fns.find(g_jit_order.front());
                if (old != g_jit_fns.end()) {
                    for (hipModule_t m : old->second.mods) (void)hipModuleUnload(m);
                    g_jit_fns.erase(old);
                }
                g_jit_order.pop_front();
            }
        }
        Loaded L;
        L.mods.assign(prog.codes.size(), nullptr);
        for (size_t c = 0; c < prog.codes.size(); ++c) {
            hipError_t e = hipModuleLoadData(&L.mods[c], prog.codes[c].data());
            if (e != hipSuccess) {
                for (size_t k = 0; k < c; ++k) (void)hipModuleUnload(L.mods[k]);
                return hip_fail(e, "hipModuleLoadData");
            }
        }
        L.fns.assign(prog.names.size(), nullptr);
        for (size_t i = 0; i < prog.names.size(); ++i) {
            hipError_t e = hipModuleGetFunction(&L.fns[i], L.mods[prog.code_of_kernel[i]], prog.names[i].c_str());
            if (e != hipSuccess) {
                for (hipModule_t m : L.mods) (void)hipModuleUnload(m);
                return hip_fail(e, "hipModuleGetFunction");
            }
        }
        it = g_jit_fns.emplace(key, std::move(L)).first;
        g_jit_order.push_back(key);
    }
    *out = &it->second.fns;
    return QF_OK;
}

qf_status launch_compiled(qf_backend* be, void* d, uint32_t qubits, uint32_t count, const qf::Pass& p,
                          hipFunction_t fn, uint32_t threads, int ns, const LaunchCtx& ctx) {
    const uint32_t log_blocks = qubits - p.tile_bits;
    const uint32_t gx_log = log_blocks < 16 ? log_blocks : 16;
    const uint64_t gy = uint64_t(count) << (log_blocks - gx_log);
    if (gy > 65535) return fail(QF_INVALID_ARGUMENT, "too many states for one compiled-pass launch");
    void* sp = d;
    const double* cs = ctx.cs;
    unsigned nslots = ctx.nslots;
    void* lam = ctx.lam;
    double* partial = ctx.partial;
    void* args1[] = {&sp, &cs, &nslots};
    void* args2[] = {&sp, &cs, &nslots, &lam, &partial};  // the adjoint form
    HIP_TRY(hipModuleLaunchKernel(fn, 1u << gx_log, uint32_t(gy), 1, threads, 1, 1, 0, be->stream,
                                  ns == 2 ? args2 : args1, nullptr));
    return QF_OK;
}

// ---------------------------------------------------------------------------
// Dense-block passes: per pass, the A operand of the real embedding of U in
// lane order ([tile][k-step][lane], the amplitude precision) and the two
// diagonal lists, each section 256-byte aligned.
// ---------------------------------------------------------------------------
size_t align256(size_t v) { return (v + 255) / 256 * 256; }

size_t dense_pass_bytes(const qf::Pass& p, int precision) {
    const size_t k = p.block.size(), nt = size_t(1) << (k - 3), nks = size_t(1) << (k - 1);
    const size_t rb = precision == QF_C128 ? 8 : 4;
    return align256(nt * nks * 64 * rb) + align256(p.pre_col.size() * sizeof(qf::DenseDiag)) +
           align256(p.pre_blk.size() * sizeof(qf::DenseDiag));
}

template <class R>
void fill_dense_pass(const qf::Pass& p, unsigned char* out) {
    const size_t k = p.block.size(), D = size_t(1) << k, nt = size_t(1) << (k - 3), nks = size_t(1) << (k - 1);
    R* aop = reinterpret_cast<R*>(out);
    for (size_t t = 0; t < nt; ++t)
        for (size_t ks = 0; ks < nks; ++ks)
            for (size_t l = 0; l < 64; ++l) {
                const size_t i = 16 * t + (l % 16), kk = 4 * ks + l / 16;
                const size_t row = i % D, colx = kk % D;
                const qf::cd u = p.U[row * D + colx];
                double v;
                if (i < D) v = kk < D ? u.re : -u.im;  // [Ur -Ui]
                else v = kk < D ? u.im : u.re;          // [Ui  Ur]
                aop[(t * nks + ks) * 64 + l] = R(v);
            }
    unsigned char* q = out + align256(nt * nks * 64 * sizeof(R));
    for (int list = 0; list < 2; ++list) {
        const auto& L = list ? p.pre_blk : p.pre_col;
        qf::DenseDiag* d = reinterpret_cast<qf::DenseDiag*>(q);
        for (size_t e = 0; e < L.size(); ++e) {
            d[e] = qf::DenseDiag{};
            d[e].a = L[e].a;
            d[e].b = L[e].b;
            for (int c = 0; c < 4; ++c) {
                d[e].d[2 * c] = L[e].d[c].re;
                d[e].d[2 * c + 1] = L[e].d[c].im;
            }
        }
        q += align256(L.size() * sizeof(qf::DenseDiag));
    }
}

bool plan_has_dense(const qf::Plan& plan) {
    for (const auto& p : plan.passes)
        if (p.kind == qf::Pass::DENSE) return true;
    return false;
}

// Uploads the dense passes' operands of a plan into the backend's buffer.
qf_status upload_dense(qf_backend* be, const qf::Plan& plan, int precision, const unsigned char** out) {
    *out = nullptr;
    size_t total = 0;
    for (const auto& p : plan.passes)
        if (p.kind == qf::Pass::DENSE) total += dense_pass_bytes(p, precision);
    if (!total) return QF_OK;
    std::vector<unsigned char> host(total, 0);
    size_t off = 0;
    for (const auto& p : plan.passes)
        if (p.kind == qf::Pass::DENSE) {
            if (precision == QF_C128) fill_dense_pass<double>(p, host.data() + off);
            else fill_dense_pass<float>(p, host.data() + off);
            off += dense_pass_bytes(p, precision);
        }
    HIP_TRY(hipStreamSynchronize(be->stream));  // earlier launches may still read the buffer
    qf_status s = ensure(be, &be->d_dense, &be->d_dense_bytes, total);
    if (s) return s;
    HIP_TRY(hipMemcpy(be->d_dense, host.data(), total, hipMemcpyHostToDevice));
    *out = static_cast<const unsigned char*>(be->d_dense);
    return QF_OK;
}

template <class R>
qf_status launch_dense(qf_backend* be, void* d, uint32_t qubits, uint32_t count, const qf::Pass& p,
                       const unsigned char* blob) {
    const uint32_t k = uint32_t(p.block.size());
    if ((k != 4 && k != 5) || qubits < k + 4) return fail(QF_INVALID_ARGUMENT, "dense pass shape");
    qf::DenseArgs a{};
    a.n = qubits;
    a.ncol = uint32_t(p.pre_col.size());
    a.nblk = uint32_t(p.pre_blk.size());
    for (uint32_t i = 0; i < k; ++i) a.S[i] = p.block[i];
    a.groups = uint64_t(1) << (qubits - k - 4);
    a.states = count;
    const size_t nt = size_t(1) << (k - 3), nks = size_t(1) << (k - 1);
    const R* aop = reinterpret_cast<const R*>(blob);
    const qf::DenseDiag* dcol = reinterpret_cast<const qf::DenseDiag*>(blob + align256(nt * nks * 64 * sizeof(R)));
    const qf::DenseDiag* dblk = reinterpret_cast<const qf::DenseDiag*>(
        blob + align256(nt * nks * 64 * sizeof(R)) + align256(p.pre_col.size() * sizeof(qf::DenseDiag)));
    const uint64_t waves = a.groups * a.states;
    const unsigned blocks = unsigned(std::min<uint64_t>((waves + 3) / 4, 65536));
    if (k == 5)
        hipLaunchKernelGGL((qf::k_dense_block<R, 5>), dim3(blocks), dim3(256), 0, be->stream, static_cast<cx<R>*>(d), a,
                           aop, dcol, dblk);
    else
        hipLaunchKernelGGL((qf::k_dense_block<R, 4>), dim3(blocks), dim3(256), 0, be->stream, static_cast<cx<R>*>(d), a,
                           aop, dcol, dblk);
    return launch_check("k_dense_block");
}

// Launches every pass of a plan whose ops were uploaded at d_ops (d_ops may
// be null when every tile pass runs compiled).
qf_status launch_plan(qf_backend* be, void* d, int precision, uint32_t qubits, uint32_t count, const qf::Plan& plan,
                      const qf::TileOp* d_ops, const LaunchCtx& ctx) {
    if (plan.qubits != qubits) return fail(QF_INCOMPATIBLE_DEVICE_STATE, "plan was built for a different qubit count");
    const qf::JitProgram* prog = nullptr;
    const std::vector<hipFunction_t>* fns = nullptr;
    const int ns = ctx.lam ? 2 : 1;
    if (plan.compiled) {
        qf_status s = plan_program(plan, precision, ns, &prog);
        if (s) return s;
        if (prog && (s = jit_functions(*prog, &fns))) return s;
    }
    if (plan.relabeled && !prog)
        return fail(QF_INVALID_ARGUMENT, "a plan that relabels qubits between passes runs only as compiled kernels");
    if (plan.scaled_rotations && !prog)
        return fail(QF_INVALID_ARGUMENT, "a plan whose rotations read (scale, ratio) rows runs only as compiled kernels");
    const qf::TileOp* cur = d_ops;
    uint32_t kinds = 0;
    size_t dense_off = 0;
    for (size_t i = 0; i < std::min(plan.passes.size(), ctx.pass_end); ++i) {
        const qf::Pass& p = plan.passes[i];
        if (i < ctx.pass_begin) {
            if (p.kind == qf::Pass::DENSE) dense_off += dense_pass_bytes(p, precision);
            if (cur) cur += p.ops.size();
            continue;
        }
        qf_status s;
        const int kern = prog ? prog->kernel_of_pass[i] : -1;
        if (p.kind == qf::Pass::DENSE) {
            if (!ctx.dense || ctx.lam) return fail(QF_INVALID_ARGUMENT, "dense pass without uploaded operands");
            kinds |= qf::KR_MFMA;
            s = precision == QF_C128 ? launch_dense<double>(be, d, qubits, count, p, ctx.dense + dense_off)
                                     : launch_dense<float>(be, d, qubits, count, p, ctx.dense + dense_off);
            dense_off += dense_pass_bytes(p, precision);
        } else if (kern >= 0) {
            kinds |= qf::KR_COMPILED;
            s = launch_compiled(be, d, qubits, count, p, (*fns)[kern], prog->threads[kern], ns, ctx);
        } else {
            kinds |= p.kind != qf::Pass::TILE ? qf::KR_SWEEP : (p.reg_bits ? qf::KR_REG_TILE : qf::KR_LDS_TILE);
            if (p.kind == qf::Pass::TILE && !cur) return fail(QF_INVALID_ARGUMENT, "tile pass without uploaded ops");
            s = precision == QF_C128 ? launch_pass<double>(be, static_cast<cx<double>*>(d), qubits, count, p, cur, ctx)
                                     : launch_pass<float>(be, static_cast<cx<float>*>(d), qubits, count, p, cur, ctx);
        }
        if (s) return s;
        if (cur) cur += p.ops.size();
    }
    plan.kernels_run = kinds;
    return QF_OK;
}

// Runs a plan on `count` contiguous states of `qubits` qubits.
qf_status run_plan(qf_backend* be, void* d, int precision, uint32_t qubits, uint32_t count, const qf::Plan& plan) {
    if (plan_has_dense(plan)) {
        LaunchCtx ctx;
        qf_status s = upload_dense(be, plan, precision, &ctx.dense);
        if (s) return s;
        return launch_plan(be, d, precision, qubits, count, plan, nullptr, ctx);
    }
    if (plan.compiled) {
        const qf::JitProgram* prog = nullptr;
        qf_status s = plan_program(plan, precision, 1, &prog);
        if (s) return s;
        if (prog) return launch_plan(be, d, precision, qubits, count, plan, nullptr, LaunchCtx{});
    }
    std::vector<const qf::TileOp*> bases;
    qf_status s = upload_plans(be, {&plan}, bases);
    if (s) return s;
    return launch_plan(be, d, precision, qubits, count, plan, bases[0], LaunchCtx{});
}

qf_status apply_gates(qf_backend* be, void* d, int precision, uint32_t qubits, uint32_t count,
                      const std::vector<qf::GateIn>& gates, const qf::PlanOptions& opt) {
    qf::Plan plan;
    std::string err = qf::build_plan(gates, qubits, precision, opt, plan);
    if (!err.empty()) return fail(QF_INVALID_ARGUMENT, err);
    return run_plan(be, d, precision, qubits, count, plan);
}

// Compensated reduction over `count` contiguous states; out gets count (re, im) pairs.
qf_status reduce(qf_backend* be, const void* d, const void* d2, int precision, uint64_t dim, uint32_t count, int kind,
                 uint32_t q, const std::vector<qf::PauliDev>& terms, std::vector<double>& out) {
    if (count > 65535) return fail(QF_INVALID_ARGUMENT, "reductions take at most 65535 states per call");
    uint64_t blocks = (dim + qf::kThreads * 4 - 1) / (qf::kThreads * 4);
    if (blocks > 1024) blocks = 1024;
    if (blocks == 0) blocks = 1;
    const size_t partial_bytes = size_t(count) * blocks * 2 * sizeof(qf::dd);
    const size_t terms_bytes = terms.size() * sizeof(qf::PauliDev);
    qf_status s = ensure(be, &be->scratch, &be->scratch_bytes, partial_bytes + terms_bytes + 256);
    if (s) return s;
    size_t cap = be->d_out_count * sizeof(double);
    void* dout = be->d_out;
    s = ensure(be, &dout, &cap, size_t(count) * 2 * sizeof(double));
    if (s) return s;
    be->d_out = static_cast<double*>(dout);
    be->d_out_count = cap / sizeof(double);
    qf::dd* partial = static_cast<qf::dd*>(be->scratch);
    qf::PauliDev* d_terms = nullptr;
    if (!terms.empty()) {
        d_terms = reinterpret_cast<qf::PauliDev*>(static_cast<char*>(be->scratch) + ((partial_bytes + 255) / 256) * 256);
        HIP_TRY(hipMemcpyAsync(d_terms, terms.data(), terms_bytes, hipMemcpyHostToDevice, be->stream));
    }
    dim3 grid(unsigned(blocks), count);
    if (precision == QF_C128)
        hipLaunchKernelGGL(qf::k_reduce<double>, grid, dim3(qf::kThreads), 0, be->stream,
                           static_cast<const cx<double>*>(d), static_cast<const cx<double>*>(d2), dim, kind, q,
                           d_terms, uint32_t(terms.size()), partial);
    else
        hipLaunchKernelGGL(qf::k_reduce<float>, grid, dim3(qf::kThreads), 0, be->stream,
                           static_cast<const cx<float>*>(d), static_cast<const cx<float>*>(d2), dim, kind, q, d_terms,
                           uint32_t(terms.size()), partial);
    s = launch_check("k_reduce");
    if (s) return s;
    hipLaunchKernelGGL(qf::k_reduce_finish<>, dim3(count), dim3(qf::kThreads), 0, be->stream, partial,
                       uint32_t(blocks), be->d_out);
    s = launch_check("k_reduce_finish");
    if (s) return s;
    out.resize(size_t(count) * 2);
    HIP_TRY(hipMemcpyAsync(out.data(), be->d_out, out.size() * sizeof(double), hipMemcpyDeviceToHost, be->stream));
    HIP_TRY(hipStreamSynchronize(be->stream));
    return QF_OK;
}

// Device form of a Pauli sum: c' = c * i^(#Y) (see qf::PauliDev).
qf_status pauli_terms(const qf_pauli_term* terms, size_t count, uint32_t qubits, std::vector<qf::PauliDev>& out) {
    out.clear();
    for (size_t t = 0; t < count; ++t) {
        const uint64_t all = terms[t].x_mask | terms[t].z_mask;
        if (qubits < 64 && (all >> qubits))
            return fail(QF_QUBIT_OUT_OF_RANGE, "Pauli term acts on a qubit out of range for " +
                                                   std::to_string(qubits) + " qubits");
        const int ny = __builtin_popcountll(terms[t].x_mask & terms[t].z_mask) & 3;
        double cr = terms[t].re, ci = terms[t].im;
        for (int k = 0; k < ny; ++k) {  // multiply by i
            const double r = -ci, i = cr;
            cr = r;
            ci = i;
        }
        out.push_back(qf::PauliDev{cr, ci, terms[t].x_mask, terms[t].z_mask});
    }
    return QF_OK;
}

}  // namespace qfrt

// ---------------------------------------------------------------------------
// Fixed device scratch for the entry points that allocate nothing
// (qf_state_expectation_pauli_sum_noalloc).  One array per device, held until
// the process ends; every call on a device shares it, so calls that overlap on
// one device (two backends' streams at once) race.  Layout: the reduction's
// per-block partials, then the two result doubles, then the term table.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kScratchPartialBytes = 1024 * 2 * sizeof(qf::dd);  // k_reduce: at most 1024 blocks
constexpr size_t kScratchOutOff = kScratchPartialBytes;
constexpr size_t kScratchTermsOff = kScratchOutOff + 256;
constexpr size_t kNoAllocMaxTerms = 4096;
constexpr size_t kScratchBytes = kScratchTermsOff + kNoAllocMaxTerms * sizeof(qf::PauliDev);
static_assert(sizeof(qf::PauliDev) == sizeof(qf_pauli_term), "term records are converted in place");
static_assert(kScratchBytes % 256 == 0, "scratch size");

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {};

// The device address of `scratch` on device `ordinal` (which must be
// current): hipGetSymbolAddress once per device, then the cached address.
qf_status scratch_address(int ordinal, void** out) {
    if (ordinal < 0 || ordinal >= 64) return fail(QF_DEVICE_ERROR, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[ordinal]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
        }
        g_scratch_ptr[ordinal] = p;
    }
    *out = g_scratch_ptr[ordinal];
    return QF_OK;
}

// qf_pauli_term -> qf::PauliDev in place: c' = c * i^(#Y), the same exact
// swaps and negations as pauli_terms on the host.
__global__ void k_pauli_terms_in_place(qf::PauliDev* __restrict__ t, uint32_t count) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    double cr = t[k].cr, ci = t[k].ci;
    const int ny = __popcll(t[k].x & t[k].z) & 3;
    for (int r = 0; r < ny; ++r) {
        const double nr = -ci, ni = cr;
        cr = nr;
        ci = ni;
    }
    t[k].cr = cr;
    t[k].ci = ci;
}
}  // namespace

// ---------------------------------------------------------------------------
// Backend
// ---------------------------------------------------------------------------
extern "C" qf_status qf_backend_create(int32_t ordinal, qf_backend** out) {
    if (!out) return fail(QF_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (ordinal < 0 || ordinal >= n)
        return fail(QF_DEVICE_ERROR, "HIP device ordinal " + std::to_string(ordinal) + " not present (" +
                                         std::to_string(n) + " devices)");
    HIP_TRY(hipSetDevice(ordinal));
    // The no-allocation entry points' scratch address, once per device with
    // that device current (a failure here is reported by those entry points).
    {
        void* sp = nullptr;
        if (scratch_address(ordinal, &sp) != QF_OK) g_last_error.clear();
    }
    qf_backend* be = new qf_backend;
    be->ordinal = ordinal;
    hipError_t e = hipStreamCreateWithFlags(&be->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete be;
        return hip_fail(e, "hipStreamCreateWithFlags");
    }
    *out = be;
    return QF_OK;
}

extern "C" qf_status qf_backend_destroy(qf_backend* be) {
    if (!be) return fail(QF_INVALID_ARGUMENT, "null backend");
    backend_release(be);
    return QF_OK;
}

extern "C" qf_status qf_backend_release_cached(qf_backend* be) {
    if (!be) return fail(QF_INVALID_ARGUMENT, "null backend");
    std::lock_guard<std::mutex> lk(g_arena_mu);
    arena_free(be);
    g_arena_owners.erase(std::remove(g_arena_owners.begin(), g_arena_owners.end(), be), g_arena_owners.end());
    return QF_OK;
}

extern "C" qf_status qf_backend_synchronize(qf_backend* be) {
    if (!be) return fail(QF_INVALID_ARGUMENT, "null backend");
    qf_status s = activate(be);
    if (s) return s;
    HIP_TRY(hipStreamSynchronize(be->stream));
    return QF_OK;
}

extern "C" qf_status qf_backend_device_ordinal(const qf_backend* be, int32_t* out) {
    if (!be || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    *out = be->ordinal;
    return QF_OK;
}

extern "C" qf_status qf_backend_set_plan_options(qf_backend* be, const qf_plan_options* o) {
    if (!be || !o) return fail(QF_INVALID_ARGUMENT, "null argument");
    if (o->tile_qubits > 14) return fail(QF_INVALID_ARGUMENT, "tile_qubits must be at most 14");
    be->opts.fuse = o->fuse;
    be->opts.tile_qubits = o->tile_qubits;
    be->opts.low_qubits = o->low_qubits;
    be->opts.max_block = o->max_block;
    be->opts.kernel = o->kernel;
    be->opts.reg_bits = o->reg_bits;
    return QF_OK;
}

extern "C" qf_status qf_backend_stream(const qf_backend* be, void** out) {
    if (!be || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    *out = be->stream;
    return QF_OK;
}

// ---------------------------------------------------------------------------
// States
// ---------------------------------------------------------------------------
namespace {

qf_status new_state(qf_backend* be, uint32_t qubits, int32_t precision, qf_state** out) {
    if (!be || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    qf_status s = check_qubits(qubits);
    if (s) return s;
    if (precision != QF_C128 && precision != QF_C64) return fail(QF_INVALID_ARGUMENT, "unknown precision");
    s = activate(be);
    if (s) return s;
    const uint64_t dim = uint64_t(1) << qubits;
    void* d = nullptr;
    s = device_alloc(&d, dim * amp_bytes(precision));
    if (s) return s;
    qf_state* st = new qf_state;
    st->be = be;
    be->refs.fetch_add(1);
    st->d = d;
    st->qubits = qubits;
    st->precision = precision;
    st->dim = dim;
    *out = st;
    return QF_OK;
}

template <class R>
void fill_zero(qf_backend* be, void* d, uint64_t total, uint32_t qubits) {
    hipLaunchKernelGGL(qf::k_fill_zero_state<R>, dim3(grid_for(total)), dim3(qf::kThreads), 0, be->stream,
                       static_cast<cx<R>*>(d), total, qubits);
}

}  // namespace

extern "C" qf_status qf_state_zero(qf_backend* be, uint32_t qubits, int32_t precision, qf_state** out) {
    qf_status s = new_state(be, qubits, precision, out);
    if (s) return s;
    qf_state* st = *out;
    if (precision == QF_C128)
        fill_zero<double>(be, st->d, st->dim, qubits);
    else
        fill_zero<float>(be, st->d, st->dim, qubits);
    s = launch_check("k_fill_zero_state");
    if (s) {
        qf_state_destroy(st);
        *out = nullptr;
    }
    return s;
}

// StateVector::zero_with_memory_budget, src/lib.rs:940-966: the budget covers
// the amplitude payload (2^qubits amplitudes of the state's precision).
extern "C" qf_status qf_state_zero_with_memory_budget(qf_backend* be, uint32_t qubits, int32_t precision,
                                                      uint64_t budget_bytes, qf_state** out) {
    if (!be || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (qubits == 0) return fail(QF_ZERO_QUBITS, "a state vector must contain at least one qubit");
    if (precision != QF_C128 && precision != QF_C64) return fail(QF_INVALID_ARGUMENT, "unknown precision");
    if (qubits >= 64) return fail(QF_SIMULATION_SIZE_OVERFLOW, "dense simulation size exceeds the supported addressable range");
    const unsigned __int128 required = (unsigned __int128)(uint64_t(1) << qubits) * amp_bytes(precision);
    if (required > budget_bytes) {
        char msg[160];
        std::snprintf(msg, sizeof msg, "dense state needs at least %llu bytes, exceeding the %llu-byte memory budget",
                      (unsigned long long)required, (unsigned long long)budget_bytes);
        return fail(QF_MEMORY_BUDGET_EXCEEDED, msg);
    }
    return qf_state_zero(be, qubits, precision, out);
}

extern "C" qf_status qf_state_upload(qf_backend* be, const qf_c128* amps, uint32_t qubits, qf_state** out) {
    if (!amps) return fail(QF_INVALID_ARGUMENT, "null amplitudes");
    qf_status s = new_state(be, qubits, QF_C128, out);
    if (s) return s;
    hipError_t e = hipMemcpy((*out)->d, amps, (*out)->dim * 16, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        qf_state_destroy(*out);
        *out = nullptr;
        return hip_fail(e, "hipMemcpy upload");
    }
    return QF_OK;
}

extern "C" qf_status qf_state_upload_c64(qf_backend* be, const qf_c64* amps, uint32_t qubits, qf_state** out) {
    if (!amps) return fail(QF_INVALID_ARGUMENT, "null amplitudes");
    qf_status s = new_state(be, qubits, QF_C64, out);
    if (s) return s;
    hipError_t e = hipMemcpy((*out)->d, amps, (*out)->dim * 8, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        qf_state_destroy(*out);
        *out = nullptr;
        return hip_fail(e, "hipMemcpy upload");
    }
    return QF_OK;
}

extern "C" qf_status qf_state_download(const qf_state* st, qf_c128* out) {
    if (!st || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    qf_backend* be = st->be;
    qf_status s = activate(be);
    if (s) return s;
    if (st->precision == QF_C128) {
        HIP_TRY(hipMemcpyAsync(out, st->d, st->dim * 16, hipMemcpyDeviceToHost, be->stream));
        HIP_TRY(hipStreamSynchronize(be->stream));
        return QF_OK;
    }
    void* tmp = nullptr;
    s = device_alloc(&tmp, st->dim * 16);
    if (s) return s;
    hipLaunchKernelGGL((qf::k_convert<float, double>), dim3(grid_for(st->dim)), dim3(qf::kThreads), 0, be->stream,
                       static_cast<const cx<float>*>(st->d), static_cast<cx<double>*>(tmp), st->dim);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, tmp, st->dim * 16, hipMemcpyDeviceToHost, be->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(be->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return hip_fail(e, "download complex64");
    return QF_OK;
}

extern "C" qf_status qf_state_download_c64(const qf_state* st, qf_c64* out) {
    if (!st || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    if (st->precision != QF_C64) return fail(QF_INCOMPATIBLE_DEVICE_STATE, "state is complex128");
    qf_status s = activate(st->be);
    if (s) return s;
    HIP_TRY(hipMemcpyAsync(out, st->d, st->dim * 8, hipMemcpyDeviceToHost, st->be->stream));
    HIP_TRY(hipStreamSynchronize(st->be->stream));
    return QF_OK;
}

extern "C" qf_status qf_state_destroy(qf_state* st) {
    if (!st) return fail(QF_INVALID_ARGUMENT, "null state");
    qf_backend* be = st->be;
    (void)hipSetDevice(be->ordinal);
    (void)hipStreamSynchronize(be->stream);
    (void)hipFree(st->d);
    delete st;
    backend_release(be);
    return QF_OK;
}

extern "C" qf_status qf_state_qubits(const qf_state* st, uint32_t* out) {
    if (!st || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    *out = st->qubits;
    return QF_OK;
}

extern "C" qf_status qf_state_precision(const qf_state* st, int32_t* out) {
    if (!st || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    *out = st->precision;
    return QF_OK;
}

extern "C" qf_status qf_state_copy_from(qf_state* dst, const qf_state* src) {
    if (!dst || !src) return fail(QF_INVALID_ARGUMENT, "null argument");
    if (dst->dim != src->dim || dst->be != src->be)
        return fail(QF_INCOMPATIBLE_DEVICE_STATE, "device states must share a backend and have equal dimensions");
    qf_status s = activate(dst->be);
    if (s) return s;
    if (dst->precision == src->precision) {
        HIP_TRY(hipMemcpyAsync(dst->d, src->d, dst->dim * amp_bytes(dst->precision), hipMemcpyDeviceToDevice,
                               dst->be->stream));
        return QF_OK;
    }
    // Precision conversion on the device (complex64 <-> complex128).
    if (src->precision == QF_C64)
        hipLaunchKernelGGL((qf::k_convert<float, double>), dim3(grid_for(dst->dim)), dim3(qf::kThreads), 0,
                           dst->be->stream, static_cast<const cx<float>*>(src->d), static_cast<cx<double>*>(dst->d),
                           dst->dim);
    else
        hipLaunchKernelGGL((qf::k_convert<double, float>), dim3(grid_for(dst->dim)), dim3(qf::kThreads), 0,
                           dst->be->stream, static_cast<const cx<double>*>(src->d), static_cast<cx<float>*>(dst->d),
                           dst->dim);
    return launch_check("k_convert");
}

extern "C" qf_status qf_state_device_pointer(const qf_state* st, void** out) {
    if (!st || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    *out = st->d;
    return QF_OK;
}

namespace {
qf_status exact_opts(qf::PlanOptions& o) {
    o = qf::PlanOptions{};
    o.fuse = 0;
    return QF_OK;
}
}  // namespace

extern "C" qf_status qf_state_apply_gate(qf_state* st, int32_t gate, uint32_t target, double theta) {
    if (!st) return fail(QF_INVALID_ARGUMENT, "null state");
    qf_op op{0, gate, target, 0, theta};
    std::vector<qf::GateIn> g;
    qf_status s = lower_ops(&op, 1, st->qubits, g);
    if (s) return s;
    if ((s = activate(st->be))) return s;
    qf::PlanOptions o;
    exact_opts(o);
    return apply_gates(st->be, st->d, st->precision, st->qubits, 1, g, o);
}

extern "C" qf_status qf_state_apply_matrix(qf_state* st, uint32_t target, const double* m8) {
    if (!st || !m8) return fail(QF_INVALID_ARGUMENT, "null argument");
    if (target >= st->qubits)
        return fail(QF_QUBIT_OUT_OF_RANGE, "qubit index " + std::to_string(target) + " is out of range for " +
                                               std::to_string(st->qubits) + " qubits");
    for (int k = 0; k < 8; ++k)
        if (!std::isfinite(m8[k])) return fail(QF_INVALID_ARGUMENT, "matrix entries must be finite");
    qf::GateIn g{};
    g.kind = 0;
    g.a = target;
    std::memcpy(g.m, m8, sizeof(g.m));
    qf_status s = activate(st->be);
    if (s) return s;
    qf::PlanOptions o;
    exact_opts(o);
    return apply_gates(st->be, st->d, st->precision, st->qubits, 1, {g}, o);
}

extern "C" qf_status qf_state_apply_rotation_chain(qf_state* st, const qf_rotation* r, size_t count) {
    if (!st || (!r && count)) return fail(QF_INVALID_ARGUMENT, "null argument");
    // validate_gates, src/cuda.rs:2256-2272: every gate before any launch.
    std::vector<qf_op> ops(count);
    for (size_t i = 0; i < count; ++i) {
        if (r[i].axis < 0 || r[i].axis > 2) return fail(QF_INVALID_ARGUMENT, "unknown rotation axis");
        ops[i] = qf_op{0, QF_GATE_RX + r[i].axis, r[i].qubit, 0, r[i].theta};
    }
    std::vector<qf::GateIn> g;
    qf_status s = lower_ops(ops.data(), count, st->qubits, g);
    if (s) return s;
    if ((s = activate(st->be))) return s;
    qf::PlanOptions o;
    exact_opts(o);
    return apply_gates(st->be, st->d, st->precision, st->qubits, 1, g, o);
}

extern "C" qf_status qf_state_apply_2q(qf_state* st, int32_t op, uint32_t a, uint32_t b) {
    if (!st) return fail(QF_INVALID_ARGUMENT, "null state");
    if (op < 0 || op > 2) return fail(QF_INVALID_ARGUMENT, "unknown two-qubit op");
    qf_op o{op + 1, 0, a, b, 0.0};
    std::vector<qf::GateIn> g;
    qf_status s = lower_ops(&o, 1, st->qubits, g);
    if (s) return s;
    if ((s = activate(st->be))) return s;
    qf::PlanOptions po;
    exact_opts(po);
    return apply_gates(st->be, st->d, st->precision, st->qubits, 1, g, po);
}

extern "C" qf_status qf_state_apply_cp(qf_state* st, uint32_t control, uint32_t target, double phi) {
    if (!st) return fail(QF_INVALID_ARGUMENT, "null state");
    qf_status s = check_pair(st->qubits, control, target);
    if (s) return s;
    if (!std::isfinite(phi)) return fail(QF_NON_FINITE_ANGLE, "rotation angle must be finite");
    qf::GateIn g{};
    g.kind = 4;
    g.a = control;
    g.b = target;
    g.m[0] = g.m[1] = g.m[2] = cd{1, 0};
    g.m[3] = cd{host_cos(phi), host_sin(phi)};
    if ((s = activate(st->be))) return s;
    qf::PlanOptions po;
    exact_opts(po);
    return apply_gates(st->be, st->d, st->precision, st->qubits, 1, {g}, po);
}

// A plan passed to a circuit run must have been built from this very
// circuit (same lowered gates and width) by qf_plan_build: a plan of another
// circuit, a parameterized plan (its rotations read an angle table the
// circuit runs do not supply) or a density plan is refused before any launch.
qf_status qfrt::check_circuit_plan(const qf::Plan& plan, const std::vector<qf::GateIn>& g, uint32_t qubits) {
    if (plan.origin != qf::PLAN_CIRCUIT || plan.runtime_rotations)
        return fail(QF_INVALID_ARGUMENT, "plan was not built by qf_plan_build (parameterized or density plan)");
    if (plan.qubits != qubits) return fail(QF_INVALID_ARGUMENT, "plan was built for a different qubit count");
    if (plan.circuit_key != qf::circuit_key(g, qubits))
        return fail(QF_INVALID_ARGUMENT, "plan was built for a different circuit");
    return QF_OK;
}

extern "C" qf_status qf_state_run_circuit(qf_state* st, const qf_op* ops, size_t count, const qf_plan* plan) {
    if (!st || (!ops && count)) return fail(QF_INVALID_ARGUMENT, "null argument");
    std::vector<qf::GateIn> g;
    qf_status s = lower_ops(ops, count, st->qubits, g);
    if (s) return s;
    if (plan && (s = check_circuit_plan(plan->plan, g, st->qubits))) return s;
    if ((s = activate(st->be))) return s;
    if (plan) return run_plan(st->be, st->d, st->precision, st->qubits, 1, plan->plan);
    return apply_gates(st->be, st->d, st->precision, st->qubits, 1, g, st->be->opts);
}

// Amplitudes at `count` basis indices, as complex128 (StateVector::amplitudes()[i],
// src/lib.rs:992-994, without downloading the state).
extern "C" qf_status qf_state_gather(const qf_state* st, const uint64_t* indices, size_t count, qf_c128* out) {
    if (!st || ((!indices || !out) && count)) return fail(QF_INVALID_ARGUMENT, "null argument");
    for (size_t k = 0; k < count; ++k)
        if (indices[k] >= st->dim)
            return fail(QF_INVALID_ARGUMENT, "basis index " + std::to_string(indices[k]) + " is out of range for " +
                                                 std::to_string(st->qubits) + " qubits");
    if (count == 0) return QF_OK;
    qf_backend* be = st->be;
    qf_status s = activate(be);
    if (s) return s;
    const size_t ib = count * sizeof(uint64_t), ob = count * sizeof(qf_c128);
    void* tmp = nullptr;
    if ((s = device_alloc(&tmp, ((ib + 255) / 256) * 256 + ob))) return s;
    uint64_t* d_idx = static_cast<uint64_t*>(tmp);
    cx<double>* d_out = reinterpret_cast<cx<double>*>(static_cast<char*>(tmp) + ((ib + 255) / 256) * 256);
    hipError_t e = hipMemcpyAsync(d_idx, indices, ib, hipMemcpyHostToDevice, be->stream);
    if (e == hipSuccess) {
        if (st->precision == QF_C128)
            hipLaunchKernelGGL(qf::k_gather<double>, dim3(grid_for(count)), dim3(qf::kThreads), 0, be->stream,
                               static_cast<const cx<double>*>(st->d), d_idx, d_out, uint64_t(count));
        else
            hipLaunchKernelGGL(qf::k_gather<float>, dim3(grid_for(count)), dim3(qf::kThreads), 0, be->stream,
                               static_cast<const cx<float>*>(st->d), d_idx, d_out, uint64_t(count));
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, ob, hipMemcpyDeviceToHost, be->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(be->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return hip_fail(e, "qf_state_gather");
    return QF_OK;
}

// The device's in-place streaming rate over this state's bytes (the memory
// pattern of one tile pass: read every amplitude once, write it back).  The
// amplitudes are multiplied by 1.0, so the state is unchanged.
extern "C" qf_status qf_state_stream_rate(qf_state* st, int32_t variant, uint32_t launches, double* gbs,
                                          double* avg_ms) {
    if (!st || !gbs) return fail(QF_INVALID_ARGUMENT, "null argument");
    if (variant < 0 || variant > 4) return fail(QF_INVALID_ARGUMENT, "unknown streaming variant");
    if (launches == 0) launches = 1;
    qf_backend* be = st->be;
    qf_status s = activate(be);
    if (s) return s;
    const uint64_t bytes = st->dim * amp_bytes(st->precision);
    const uint64_t words = bytes / 16;
    if (words == 0) return fail(QF_INVALID_ARGUMENT, "state too small to stream");
    double2* p = static_cast<double2*>(st->d);
    const uint32_t chunk_log = 12;  // 64 KiB per workgroup, the bytes of a 12-qubit complex128 tile
    auto launch = [&]() {
        switch (variant) {
            case 0:
                hipLaunchKernelGGL(qf::k_stream_probe<0>, dim3(grid_for(words)), dim3(qf::kThreads), 0, be->stream,
                                   p, words, 1.0, chunk_log);
                break;
            case 1:
                hipLaunchKernelGGL(qf::k_stream_probe<1>, dim3(grid_for(words)), dim3(qf::kThreads), 0, be->stream,
                                   p, words, 1.0, chunk_log);
                break;
            case 2:
                hipLaunchKernelGGL(qf::k_stream_probe<2>, dim3(grid_for((words + 3) / 4)), dim3(qf::kThreads), 0,
                                   be->stream, p, words, 1.0, chunk_log);
                break;
            case 4:
                hipLaunchKernelGGL(qf::k_stream_probe<4>, dim3(grid_for((words + 7) / 8)), dim3(qf::kThreads), 0,
                                   be->stream, p, words, 1.0, chunk_log);
                break;
            default:
                hipLaunchKernelGGL(qf::k_stream_probe<3>,
                                   dim3(unsigned((words + (uint64_t(1) << chunk_log) - 1) >> chunk_log)),
                                   dim3(qf::kThreads), 0, be->stream, p, words, 1.0, chunk_log);
        }
    };
    launch();  // warm
    hipEvent_t a = nullptr, b = nullptr;
    HIP_TRY(hipEventCreate(&a));
    hipError_t e = hipEventCreate(&b);
    if (e != hipSuccess) {
        (void)hipEventDestroy(a);
        return hip_fail(e, "hipEventCreate");
    }
    e = hipEventRecord(a, be->stream);
    for (uint32_t k = 0; k < launches && e == hipSuccess; ++k) {
        launch();
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(b, be->stream);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (e != hipSuccess) return hip_fail(e, "qf_state_stream_rate");
    const double per = double(ms) / launches;
    *gbs = 2.0 * double(bytes) / (per * 1e-3) / 1e9;
    if (avg_ms) *avg_ms = per;
    return QF_OK;
}

extern "C" qf_status qf_state_norm_sqr(const qf_state* st, double* out) {
    if (!st || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    qf_status s = activate(st->be);
    if (s) return s;
    std::vector<double> r;
    s = reduce(st->be, st->d, nullptr, st->precision, st->dim, 1, qf::RED_NORM, 0, {}, r);
    if (s) return s;
    *out = r[0];
    return QF_OK;
}

extern "C" qf_status qf_state_probabilities(const qf_state* st, double* out) {
    if (!st || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    qf_backend* be = st->be;
    qf_status s = activate(be);
    if (s) return s;
    void* tmp = nullptr;
    s = device_alloc(&tmp, st->dim * sizeof(double));
    if (s) return s;
    if (st->precision == QF_C128)
        hipLaunchKernelGGL(qf::k_probabilities<double>, dim3(grid_for(st->dim)), dim3(qf::kThreads), 0, be->stream,
                           static_cast<const cx<double>*>(st->d), static_cast<double*>(tmp), st->dim);
    else
        hipLaunchKernelGGL(qf::k_probabilities<float>, dim3(grid_for(st->dim)), dim3(qf::kThreads), 0, be->stream,
                           static_cast<const cx<float>*>(st->d), static_cast<double*>(tmp), st->dim);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, tmp, st->dim * sizeof(double), hipMemcpyDeviceToHost, be->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(be->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return hip_fail(e, "probabilities");
    return QF_OK;
}

extern "C" qf_status qf_state_expectation(const qf_state* st, int32_t obs, uint32_t qubit, double* out) {
    if (!st || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    if (obs < 0 || obs > 2) return fail(QF_INVALID_ARGUMENT, "unknown observable");
    if (qubit >= st->qubits)
        return fail(QF_QUBIT_OUT_OF_RANGE, "qubit index " + std::to_string(qubit) + " is out of range for " +
                                               std::to_string(st->qubits) + " qubits");
    qf_status s = activate(st->be);
    if (s) return s;
    std::vector<double> r;
    s = reduce(st->be, st->d, nullptr, st->precision, st->dim, 1, obs == QF_OBS_Z ? qf::RED_Z : qf::RED_XY, qubit, {},
               r);
    if (s) return s;
    // expectation_x = sum 2 Re(conj(a0) a1), expectation_y = sum 2 Im(conj(a0) a1) (src/lib.rs:1272-1303).
    *out = obs == QF_OBS_Z ? r[0] : (obs == QF_OBS_X ? 2.0 * r[0] : 2.0 * r[1]);
    return QF_OK;
}

extern "C" qf_status qf_state_expectation_pauli_sum(const qf_state* st, const qf_pauli_term* terms, size_t count,
                                                    double* out_re, double* out_im) {
    if (!st || (!terms && count) || !out_re || !out_im) return fail(QF_INVALID_ARGUMENT, "null argument");
    std::vector<qf::PauliDev> t;
    qf_status s = pauli_terms(terms, count, st->qubits, t);
    if (s) return s;
    if (count == 0) {
        *out_re = *out_im = 0.0;
        return QF_OK;
    }
    if ((s = activate(st->be))) return s;
    std::vector<double> r;
    s = reduce(st->be, st->d, nullptr, st->precision, st->dim, 1, qf::RED_PAULI, 0, t, r);
    if (s) return s;
    *out_re = r[0];
    *out_im = r[1];
    return QF_OK;
}

// qf_state_expectation_pauli_sum without any allocation: the reduction
// partials, the term table and the result live in the device's fixed scratch
// array (kScratchBytes); the same kernels, grid and summation order as the
// allocating entry point, so the results are identical.  Nothing in the
// scratch needs clearing: every partial, term and result slot a call reads
// was written by that call.
extern "C" qf_status qf_state_expectation_pauli_sum_noalloc(const qf_state* st, const qf_pauli_term* terms,
                                                            size_t count, double* out_re, double* out_im) {
    if (!st || (!terms && count) || !out_re || !out_im) return fail(QF_INVALID_ARGUMENT, "null argument");
    if (count > kNoAllocMaxTerms)
        return fail(QF_INVALID_ARGUMENT, "the no-allocation expectation takes at most " +
                                             std::to_string(kNoAllocMaxTerms) + " Pauli terms, got " +
                                             std::to_string(count));
    for (size_t t = 0; t < count; ++t) {
        const uint64_t all = terms[t].x_mask | terms[t].z_mask;
        if (st->qubits < 64 && (all >> st->qubits))
            return fail(QF_QUBIT_OUT_OF_RANGE, "Pauli term acts on a qubit out of range for " +
                                                   std::to_string(st->qubits) + " qubits");
    }
    if (count == 0) {
        *out_re = *out_im = 0.0;
        return QF_OK;
    }
    qf_backend* be = st->be;
    qf_status s = activate(be);
    if (s) return s;
    void* base = nullptr;
    if ((s = scratch_address(be->ordinal, &base))) return s;
    char* sc = static_cast<char*>(base);
    qf::dd* partial = reinterpret_cast<qf::dd*>(sc);
    double* d_out = reinterpret_cast<double*>(sc + kScratchOutOff);
    qf::PauliDev* d_terms = reinterpret_cast<qf::PauliDev*>(sc + kScratchTermsOff);
    uint64_t blocks = (st->dim + qf::kThreads * 4 - 1) / (qf::kThreads * 4);
    if (blocks > 1024) blocks = 1024;
    if (blocks == 0) blocks = 1;
    HIP_TRY(hipMemcpyAsync(d_terms, terms, count * sizeof(qf_pauli_term), hipMemcpyHostToDevice, be->stream));
    hipLaunchKernelGGL(k_pauli_terms_in_place, dim3(unsigned((count + 255) / 256)), dim3(256), 0, be->stream, d_terms,
                       uint32_t(count));
    if ((s = launch_check("k_pauli_terms_in_place"))) return s;
    if (st->precision == QF_C128)
        hipLaunchKernelGGL(qf::k_reduce<double>, dim3(unsigned(blocks), 1), dim3(qf::kThreads), 0, be->stream,
                           static_cast<const cx<double>*>(st->d), static_cast<const cx<double>*>(nullptr), st->dim,
                           int(qf::RED_PAULI), 0u, d_terms, uint32_t(count), partial);
    else
        hipLaunchKernelGGL(qf::k_reduce<float>, dim3(unsigned(blocks), 1), dim3(qf::kThreads), 0, be->stream,
                           static_cast<const cx<float>*>(st->d), static_cast<const cx<float>*>(nullptr), st->dim,
                           int(qf::RED_PAULI), 0u, d_terms, uint32_t(count), partial);
    if ((s = launch_check("k_reduce"))) return s;
    hipLaunchKernelGGL(qf::k_reduce_finish<>, dim3(1), dim3(qf::kThreads), 0, be->stream, partial, uint32_t(blocks),
                       d_out);
    if ((s = launch_check("k_reduce_finish"))) return s;
    double r[2] = {0.0, 0.0};
    HIP_TRY(hipMemcpyAsync(r, d_out, sizeof r, hipMemcpyDeviceToHost, be->stream));
    HIP_TRY(hipStreamSynchronize(be->stream));
    *out_re = r[0];
    *out_im = r[1];
    return QF_OK;
}

extern "C" qf_status qf_state_inner(const qf_state* a, const qf_state* b, double* out_re, double* out_im) {
    if (!a || !b || !out_re || !out_im) return fail(QF_INVALID_ARGUMENT, "null argument");
    if (a->dim != b->dim || a->precision != b->precision || a->be != b->be)
        return fail(QF_INCOMPATIBLE_DEVICE_STATE, "device states must share a backend and have equal dimensions");
    qf_status s = activate(a->be);
    if (s) return s;
    std::vector<double> r;
    s = reduce(a->be, a->d, b->d, a->precision, a->dim, 1, qf::RED_INNER, 0, {}, r);
    if (s) return s;
    *out_re = r[0];
    *out_im = r[1];
    return QF_OK;
}

// ---------------------------------------------------------------------------
// Batches
// ---------------------------------------------------------------------------
namespace {
// A batch of count states of widths qubits[0..count) (uniform: all equal to
// qubits[0]); segments of consecutive equal widths.
qf_status new_batch_widths(qf_backend* be, const uint32_t* qubits, uint32_t count, int32_t precision,
                           qf_batch** out) {
    if (!be || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (count == 0) return fail(QF_EMPTY_BATCH, "a HIP state batch cannot be empty");
    if (precision != QF_C128 && precision != QF_C64) return fail(QF_INVALID_ARGUMENT, "unknown precision");
    std::vector<qf_batch::Seg> segs;
    uint64_t total = 0;
    uint32_t widest = 0;
    for (uint32_t i = 0; i < count; ++i) {
        qf_status s = check_qubits(qubits[i]);
        if (s) return s;
        if (segs.empty() || segs.back().qubits != qubits[i]) segs.push_back(qf_batch::Seg{qubits[i], i, 0, total});
        ++segs.back().count;
        const uint64_t dim = uint64_t(1) << qubits[i];
        if (total + dim < total || (total + dim) > (uint64_t(1) << 40))
            return fail(QF_STATE_TOO_LARGE, "batch too large");
        total += dim;
        widest = std::max(widest, qubits[i]);
    }
    qf_status s = activate(be);
    if (s) return s;
    void* d = nullptr;
    s = device_alloc(&d, total * amp_bytes(precision));
    if (s) return s;
    qf_batch* b = new qf_batch;
    b->be = be;
    be->refs.fetch_add(1);
    b->d = d;
    b->qubits = widest;
    b->count = count;
    b->precision = precision;
    b->dim = uint64_t(1) << widest;
    b->segs = std::move(segs);
    b->total = total;
    *out = b;
    return QF_OK;
}

qf_status new_batch(qf_backend* be, uint32_t qubits, uint32_t count, int32_t precision, qf_batch** out) {
    if (count == 0) {
        if (out) *out = nullptr;
        return fail(QF_EMPTY_BATCH, "a HIP state batch cannot be empty");
    }
    std::vector<uint32_t> w(count, qubits);
    return new_batch_widths(be, w.data(), count, precision, out);
}

qf_status upload_into(qf_batch** out, const qf_c128* amps) {
    hipError_t e = hipMemcpy((*out)->d, amps, (*out)->total * 16, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        qf_batch_destroy(*out);
        *out = nullptr;
        return hip_fail(e, "hipMemcpy batch upload");
    }
    return QF_OK;
}

void* seg_base(const qf_batch* b, const qf_batch::Seg& g) {
    return static_cast<char*>(b->d) + g.off * amp_bytes(b->precision);
}
}  // namespace

// Uneven widths allowed: the states are stored back to back, each with its
// own width (no reference counterpart; CudaBackend::upload_batch accepts
// equal widths only, src/cuda.rs:871-954, kept as qf_batch_upload_equal).
extern "C" qf_status qf_batch_upload(qf_backend* be, const qf_c128* amps, const uint32_t* qubits, uint32_t count,
                                     qf_batch** out) {
    if (!out) return fail(QF_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    if (count == 0) return fail(QF_EMPTY_BATCH, "a HIP state batch cannot be empty");
    if (!amps || !qubits) return fail(QF_INVALID_ARGUMENT, "null argument");
    qf_status s = new_batch_widths(be, qubits, count, QF_C128, out);
    if (s) return s;
    return upload_into(out, amps);
}

// CudaBackend::upload_batch, src/cuda.rs:871-954: EmptyBatch, IncompatibleBatch
// (src/cuda.rs:875-877) for states of different widths.
extern "C" qf_status qf_batch_upload_equal(qf_backend* be, const qf_c128* amps, const uint32_t* qubits,
                                           uint32_t count, qf_batch** out) {
    if (!out) return fail(QF_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    if (count == 0) return fail(QF_EMPTY_BATCH, "a HIP state batch cannot be empty");
    if (!amps || !qubits) return fail(QF_INVALID_ARGUMENT, "null argument");
    for (uint32_t i = 1; i < count; ++i)
        if (qubits[i] != qubits[0])
            return fail(QF_INCOMPATIBLE_BATCH, "all states in a HIP batch must have equal dimensions");
    qf_status s = new_batch(be, qubits[0], count, QF_C128, out);
    if (s) return s;
    return upload_into(out, amps);
}

extern "C" qf_status qf_batch_zero(qf_backend* be, uint32_t qubits, uint32_t count, int32_t precision,
                                   qf_batch** out) {
    if (count == 0) {
        if (out) *out = nullptr;
        return fail(QF_EMPTY_BATCH, "a HIP state batch cannot be empty");
    }
    std::vector<uint32_t> w(count, qubits);
    return qf_batch_zero_uneven(be, w.data(), count, precision, out);
}

extern "C" qf_status qf_batch_zero_uneven(qf_backend* be, const uint32_t* qubits, uint32_t count, int32_t precision,
                                          qf_batch** out) {
    if (!qubits && count) return fail(QF_INVALID_ARGUMENT, "null argument");
    qf_status s = new_batch_widths(be, qubits, count, precision, out);
    if (s) return s;
    qf_batch* b = *out;
    for (const auto& g : b->segs) {
        const uint64_t n = (uint64_t(1) << g.qubits) * g.count;
        if (precision == QF_C128)
            fill_zero<double>(be, seg_base(b, g), n, g.qubits);
        else
            fill_zero<float>(be, seg_base(b, g), n, g.qubits);
    }
    s = launch_check("k_fill_zero_state");
    if (s) {
        qf_batch_destroy(b);
        *out = nullptr;
    }
    return s;
}

extern "C" qf_status qf_batch_widths(const qf_batch* b, uint32_t* out) {
    if (!b || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    for (const auto& g : b->segs)
        for (uint32_t i = 0; i < g.count; ++i) out[g.first + i] = g.qubits;
    return QF_OK;
}

extern "C" qf_status qf_batch_destroy(qf_batch* b) {
    if (!b) return fail(QF_INVALID_ARGUMENT, "null batch");
    qf_backend* be = b->be;
    (void)hipSetDevice(be->ordinal);
    (void)hipStreamSynchronize(be->stream);
    (void)hipFree(b->d);
    delete b;
    backend_release(be);
    return QF_OK;
}

extern "C" qf_status qf_batch_shape(const qf_batch* b, uint32_t* qubits, uint32_t* count) {
    if (!b) return fail(QF_INVALID_ARGUMENT, "null batch");
    if (qubits) *qubits = b->qubits;
    if (count) *count = b->count;
    return QF_OK;
}

namespace {
// Lowers `ops` for every segment width first (errors follow the narrowest
// state that rejects an op; nothing is launched before all are valid), then
// runs them segment by segment.
qf_status batch_ops(qf_batch* b, const qf_op* ops, size_t count, const qf::PlanOptions& po) {
    std::vector<std::vector<qf::GateIn>> g(b->segs.size());
    for (size_t i = 0; i < b->segs.size(); ++i) {
        qf_status s = lower_ops(ops, count, b->segs[i].qubits, g[i]);
        if (s) return s;
    }
    qf_status s = activate(b->be);
    if (s) return s;
    for (size_t i = 0; i < b->segs.size(); ++i) {
        const auto& sg = b->segs[i];
        if ((s = apply_gates(b->be, seg_base(b, sg), b->precision, sg.qubits, sg.count, g[i], po))) return s;
    }
    return QF_OK;
}
}  // namespace

extern "C" qf_status qf_batch_apply_gate(qf_batch* b, int32_t gate, uint32_t target, double theta) {
    if (!b) return fail(QF_INVALID_ARGUMENT, "null batch");
    qf_op op{0, gate, target, 0, theta};
    qf::PlanOptions o;
    exact_opts(o);
    return batch_ops(b, &op, 1, o);
}

extern "C" qf_status qf_batch_apply_2q(qf_batch* b, int32_t op, uint32_t a, uint32_t c) {
    if (!b) return fail(QF_INVALID_ARGUMENT, "null batch");
    if (op < 0 || op > 2) return fail(QF_INVALID_ARGUMENT, "unknown two-qubit op");
    qf_op o{op + 1, 0, a, c, 0.0};
    qf::PlanOptions po;
    exact_opts(po);
    return batch_ops(b, &o, 1, po);
}

extern "C" qf_status qf_batch_run_circuit(qf_batch* b, const qf_op* ops, size_t count, const qf_plan* plan) {
    if (!b || (!ops && count)) return fail(QF_INVALID_ARGUMENT, "null argument");
    if (!plan) return batch_ops(b, ops, count, b->be->opts);
    if (b->uneven())
        return fail(QF_INCOMPATIBLE_BATCH, "a plan is built for one width: run a batch of different widths without one");
    std::vector<qf::GateIn> g;
    qf_status s = lower_ops(ops, count, b->qubits, g);
    if (s) return s;
    if ((s = check_circuit_plan(plan->plan, g, b->qubits))) return s;
    if ((s = activate(b->be))) return s;
    return run_plan(b->be, b->d, b->precision, b->qubits, b->count, plan->plan);
}

extern "C" qf_status qf_batch_download(const qf_batch* b, qf_c128* out) {
    if (!b || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    qf_backend* be = b->be;
    qf_status s = activate(be);
    if (s) return s;
    const uint64_t total = b->total;
    if (b->precision == QF_C128) {
        HIP_TRY(hipMemcpyAsync(out, b->d, total * 16, hipMemcpyDeviceToHost, be->stream));
        HIP_TRY(hipStreamSynchronize(be->stream));
        return QF_OK;
    }
    void* tmp = nullptr;
    s = device_alloc(&tmp, total * 16);
    if (s) return s;
    hipLaunchKernelGGL((qf::k_convert<float, double>), dim3(grid_for(total)), dim3(qf::kThreads), 0, be->stream,
                       static_cast<const cx<float>*>(b->d), static_cast<cx<double>*>(tmp), total);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, tmp, total * 16, hipMemcpyDeviceToHost, be->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(be->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return hip_fail(e, "batch download complex64");
    return QF_OK;
}

extern "C" qf_status qf_batch_expectation_pauli_sum(const qf_batch* b, const qf_pauli_term* terms, size_t nterms,
                                                    double* out) {
    if (!b || (!terms && nterms) || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    std::vector<std::vector<qf::PauliDev>> t(b->segs.size());
    for (size_t k = 0; k < b->segs.size(); ++k) {
        qf_status s = pauli_terms(terms, nterms, b->segs[k].qubits, t[k]);
        if (s) return s;
    }
    qf_status s = activate(b->be);
    if (s) return s;
    std::vector<double> r;
    for (size_t k = 0; k < b->segs.size(); ++k) {
        const auto& sg = b->segs[k];
        const uint64_t dim = uint64_t(1) << sg.qubits;
        for (uint32_t off = 0; off < sg.count; off += 65535) {
            const uint32_t n = std::min<uint32_t>(65535, sg.count - off);
            const char* base = static_cast<const char*>(seg_base(b, sg)) + size_t(off) * dim * amp_bytes(b->precision);
            s = reduce(b->be, base, nullptr, b->precision, dim, n, qf::RED_PAULI, 0, t[k], r);
            if (s) return s;
            for (uint32_t i = 0; i < n; ++i) out[sg.first + off + i] = r[2 * i];
        }
    }
    return QF_OK;
}

// ---------------------------------------------------------------------------
// Plans
// ---------------------------------------------------------------------------
extern "C" qf_status qf_plan_build(const qf_op* ops, size_t count, uint32_t qubits, const qf_plan_options* o,
                                   qf_plan** out) {
    if (!out || (!ops && count)) return fail(QF_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    qf_status s = check_qubits(qubits);
    if (s) return s;
    std::vector<qf::GateIn> g;
    s = lower_ops(ops, count, qubits, g);
    if (s) return s;
    const qf::PlanOptions po = plan_options_from(o);
    const int precision = o ? o->precision : QF_C128;
    if (precision != QF_C128 && precision != QF_C64) return fail(QF_INVALID_ARGUMENT, "unknown precision");
    qf_plan* p = new qf_plan;
    std::string err = qf::build_plan(g, qubits, precision, po, p->plan);
    if (!err.empty()) {
        delete p;
        return fail(QF_INVALID_ARGUMENT, err);
    }
    *out = p;
    return QF_OK;
}

extern "C" qf_status qf_plan_stats_get(const qf_plan* p, qf_plan_stats* out) {
    if (!p || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    out->passes = p->plan.stats.passes;
    out->ops_in = p->plan.stats.ops_in;
    out->ops_applied = p->plan.stats.ops_applied;
    out->dense_blocks = p->plan.stats.dense_blocks;
    out->max_ops_per_pass = p->plan.stats.max_ops_per_pass;
    out->phases = p->plan.stats.phases;
    out->kernels_run = p->plan.kernels_run;
    out->runtime_rotations = p->plan.runtime_rotations;
    return QF_OK;
}

extern "C" qf_status qf_plan_describe(const qf_plan* p, char* buf, size_t cap, size_t* needed) {
    if (!p) return fail(QF_INVALID_ARGUMENT, "null plan");
    const std::string s = qf::describe_plan(p->plan);
    if (needed) *needed = s.size() + 1;
    if (buf && cap) {
        const size_t n = std::min(cap - 1, s.size());
        std::memcpy(buf, s.data(), n);
        buf[n] = '\0';
    }
    return QF_OK;
}

namespace {
qf_status plan_source(const qf_plan* p, int32_t precision, int ns, char* buf, size_t cap, size_t* needed) {
    if (!p) return fail(QF_INVALID_ARGUMENT, "null plan");
    if (precision != QF_C128 && precision != QF_C64) return fail(QF_INVALID_ARGUMENT, "unknown precision");
    qf::JitProgram prog;
    std::string err;
    const std::string s = qf::jit_plan_source(p->plan, precision == QF_C128 ? 0 : 1, &prog, &err, ns);
    if (s.empty()) return fail(QF_INVALID_ARGUMENT, "pass compiler: " + err);
    if (needed) *needed = s.size() + 1;
    if (buf && cap) {
        const size_t n = std::min(cap - 1, s.size());
        std::memcpy(buf, s.data(), n);
        buf[n] = '\0';
    }
    return QF_OK;
}
}  // namespace

extern "C" qf_status qf_plan_source(const qf_plan* p, int32_t precision, char* buf, size_t cap, size_t* needed) {
    return plan_source(p, precision, 1, buf, cap, needed);
}

extern "C" qf_status qf_plan_adjoint_source(const qf_plan* p, int32_t precision, char* buf, size_t cap,
                                            size_t* needed) {
    return plan_source(p, precision, 2, buf, cap, needed);
}

extern "C" qf_status qf_plan_compile(const qf_plan* p, int32_t precision, double* compile_ms, int32_t* from_cache) {
    if (!p) return fail(QF_INVALID_ARGUMENT, "null plan");
    if (precision != QF_C128 && precision != QF_C64) return fail(QF_INVALID_ARGUMENT, "unknown precision");
    if (!p->plan.compiled) return fail(QF_INVALID_ARGUMENT, "the plan does not run compiled passes");
    std::string err;
    auto prog = qf::jit_compile_plan(p->plan, precision == QF_C128 ? 0 : 1, &err);
    if (!prog) return fail(QF_INVALID_ARGUMENT, "pass compiler: " + err);
    if (compile_ms) *compile_ms = prog->compile_ms;
    if (from_cache) *from_cache = prog->from_disk ? 1 : 0;
    return QF_OK;
}

extern "C" qf_status qf_plan_destroy(qf_plan* p) {
    if (!p) return fail(QF_INVALID_ARGUMENT, "null plan");
    delete p;
    return QF_OK;
}
