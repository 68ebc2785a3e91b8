// Density matrices on the device (CudaDeviceDensityMatrix and
// CudaBackend::upload_density / run_density_circuit, src/cuda.rs:969-1047,
// 2007-2253; CPU semantics DensityMatrix, src/lib.rs:1553-2227).
//
// A density matrix of n qubits is stored as its row-major entries, which is a
// state vector of 2n qubits: entry (row, col) sits at index row * 2^n + col,
// so row qubit q is vector bit n + q and column qubit q is vector bit q.  In
// that view rho -> U rho U^dagger is U on bit n + q followed by conj(U) on bit
// q, CNOT/CZ/SWAP act on the row bits and on the column bits, and a noise
// channel is one 4x4 superoperator on the bit pair (q, n + q).  Every circuit
// therefore goes through the same fused tile planner (and, from 9 qubits, the
// same compiled passes) as state vectors, in place: a whole layer of gates and
// channels costs one or a few HBM sweeps, where the reference runs two
// out-of-place sweeps per gate (src/cuda.rs:2031-2063).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "qf_runtime.h"

using qf::cd;
using qf::cx;
using namespace qfrt;

namespace {

constexpr uint32_t kMaxDensityQubits = 15;  // validate_density_size, src/cuda.rs:2239-2253

__device__ __forceinline__ uint64_t dthread() { return uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; }

// Density-matrix collapse (DensityMatrix::measure, src/lib.rs:1939-1948):
// entries whose row or column bit q differs from the outcome become zero, the
// others are divided by the outcome's probability.  Row bit q is index bit
// n + q, column bit q is index bit q.
__global__ __launch_bounds__(qf::kThreads) void k_density_collapse(cx<double>* __restrict__ s, uint64_t total,
                                                                      uint32_t n, uint32_t q, uint32_t outcome,
                                                                      double probability) {
#pragma clang fp contract(off)
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = dthread(); i < total; i += stride) {
        const uint32_t rb = uint32_t((i >> (n + q)) & 1), cb = uint32_t((i >> q) & 1);
        const cx<double> v = s[i];
        s[i] = (rb != outcome || cb != outcome) ? cx<double>{0.0, 0.0}
                                                : cx<double>{v.re / probability, v.im / probability};
    }
}

// rho = |psi><psi| (DensityMatrix::from_state, src/lib.rs:1642-1656):
// entry (row, col) = psi[row] * conj(psi[col]), the unfused complex product.
__global__ __launch_bounds__(qf::kThreads) void k_density_outer(const cx<double>* __restrict__ psi,
                                                                   cx<double>* __restrict__ rho, uint32_t n) {
#pragma clang fp contract(off)
    const uint64_t total = uint64_t(1) << (2 * n), mask = (uint64_t(1) << n) - 1;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = dthread(); i < total; i += stride) {
        const cx<double> a = psi[i >> n], b = psi[i & mask];
        const double br = b.re, bi = -b.im;
        rho[i] = cx<double>{a.re * br - a.im * bi, a.re * bi + a.im * br};
    }
}

// I / 2^n (DensityMatrix::maximally_mixed, src/lib.rs:1697-1699).
__global__ __launch_bounds__(qf::kThreads) void k_density_mixed(cx<double>* __restrict__ rho, uint32_t n) {
    const uint64_t total = uint64_t(1) << (2 * n), mask = (uint64_t(1) << n) - 1;
    const double v = 1.0 / double(uint64_t(1) << n);
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = dthread(); i < total; i += stride)
        rho[i] = cx<double>{(i >> n) == (i & mask) ? v : 0.0, 0.0};
}


std::string qstr(uint32_t q, uint32_t n) {
    return "qubit index " + std::to_string(q) + " is out of range for " + std::to_string(n) + " qubits";
}

// validate_density_size, src/cuda.rs:2239-2253.
qf_status check_density_qubits(uint32_t qubits) {
    if (qubits == 0) return fail(QF_ZERO_QUBITS, "a density matrix must contain at least one qubit");
    if (qubits > kMaxDensityQubits) return fail(QF_STATE_TOO_LARGE, "density matrix is too large for the HIP kernels");
    return QF_OK;
}

// Superoperator of a channel on (row bit r, column bit c), index 2 r + c:
// S[2r+c][2r'+c'] = sum_k w_k K_k[r][r'] conj(K_k[c][c']) over the Kraus
// operators of apply_channel / depolarize (src/lib.rs:2057-2129, 2200-2226).
void channel_superop(int32_t kind, double p0, double p1, cd S[16]) {
    struct Kraus {
        double w;
        cd k[4];
    };
    const cd z{0, 0}, o{1, 0};
    std::vector<Kraus> ks;
    switch (kind) {
        case QF_CHANNEL_DEPOLARIZING:
            ks = {{1.0 - p0, {o, z, z, o}},
                  {p0 / 3.0, {z, o, o, z}},
                  {p0 / 3.0, {z, cd{0, -1}, cd{0, 1}, z}},
                  {p0 / 3.0, {o, z, z, cd{-1, 0}}}};
            break;
        case QF_CHANNEL_AMPLITUDE_DAMPING:
            ks = {{1.0, {o, z, z, cd{std::sqrt(1.0 - p0), 0}}}, {p0, {z, o, z, z}}};
            break;
        case QF_CHANNEL_PHASE_DAMPING:
            ks = {{1.0, {o, z, z, cd{std::sqrt(1.0 - p0), 0}}}, {p0, {z, z, z, o}}};
            break;
        default: {  // thermal damping: gamma_up = gamma_down * exp(-beta_delta)
            const double up = p0 * std::exp(-p1), down = p0;
            ks = {{1.0, {cd{std::sqrt(1.0 - up), 0}, z, z, cd{std::sqrt(1.0 - down), 0}}},
                  {down, {z, o, z, z}},
                  {up, {z, z, o, z}}};
            break;
        }
    }
    for (int i = 0; i < 16; ++i) S[i] = z;
    for (const Kraus& K : ks)
        for (int r = 0; r < 2; ++r)
            for (int c = 0; c < 2; ++c)
                for (int r2 = 0; r2 < 2; ++r2)
                    for (int c2 = 0; c2 < 2; ++c2) {
                        const cd a = K.k[r * 2 + r2], b = K.k[c * 2 + c2];
                        // a * conj(b), weighted
                        const cd v{K.w * (a.re * b.re + a.im * b.im), K.w * (a.im * b.re - a.re * b.im)};
                        cd& t = S[(2 * r + c) * 4 + (2 * r2 + c2)];
                        t.re += v.re;
                        t.im += v.im;
                    }
}

// Lowers a density circuit to planner gates on the 2n-qubit view, validating
// the ops in order with the CPU path's errors (Circuit::run_density,
// src/lib.rs:2804-2819; gate angles as CudaDeviceDensityMatrix::apply_gate).
qf_status lower_density(const qf_op* ops, size_t nops, uint32_t n, std::vector<qf::GateIn>& out) {
    out.clear();
    for (size_t i = 0; i < nops; ++i) {
        const qf_op& op = ops[i];
        switch (op.kind) {
            case 0: {
                if (op.gate < QF_GATE_I || op.gate > QF_GATE_RZ)
                    return fail(QF_INVALID_ARGUMENT, "unknown gate id " + std::to_string(op.gate));
                if (op.a >= n) return fail(QF_QUBIT_OUT_OF_RANGE, qstr(op.a, n));
                if (is_rotation(op.gate) && !std::isfinite(op.theta))
                    return fail(QF_NON_FINITE_ANGLE, "rotation angle must be finite");
                qf::GateIn row{}, col{};
                row.kind = col.kind = 0;
                gate_matrix(op.gate, op.theta, row.m);
                row.a = n + op.a;
                col.a = op.a;
                for (int k = 0; k < 4; ++k) col.m[k] = cd{row.m[k].re, -row.m[k].im};
                out.push_back(row);
                out.push_back(col);
                break;
            }
            case 1: case 2: case 3: case 7: {
                qf_status s = check_pair(n, op.a, op.b);
                if (s) return s;
                qf::GateIn row{}, col{};
                if (op.kind == 7) {
                    // Exact controlled phase: diag(1,1,1,e^{i phi}) on rows, its conjugate on columns.
                    row.kind = col.kind = 4;
                    row.m[0] = row.m[1] = row.m[2] = col.m[0] = col.m[1] = col.m[2] = cd{1, 0};
                    row.m[3] = cd{host_cos(op.theta), host_sin(op.theta)};
                    col.m[3] = cd{row.m[3].re, -row.m[3].im};
                } else {
                    row.kind = col.kind = op.kind;
                }
                row.a = n + op.a;
                row.b = n + op.b;
                col.a = op.a;
                col.b = op.b;
                out.push_back(row);
                out.push_back(col);
                break;
            }
            case 4: {
                if (op.a >= n) return fail(QF_QUBIT_OUT_OF_RANGE, qstr(op.a, n));
                qf_status s = check_channel(op.gate, op.theta, op.theta2);
                if (s) return s;
                qf::GateIn g{};
                g.kind = 6;
                g.a = op.a;      // column bit: low bit of the superoperator index
                g.b = n + op.a;  // row bit: high bit
                channel_superop(op.gate, op.theta, op.theta2, g.m4);
                // A diagonal superoperator (phase damping) is a two-qubit
                // diagonal: applied from the element index, no register bits.
                bool diagonal = true;
                for (int r = 0; r < 4; ++r)
                    for (int c = 0; c < 4; ++c)
                        if (r != c && (g.m4[r * 4 + c].re != 0.0 || g.m4[r * 4 + c].im != 0.0)) diagonal = false;
                if (diagonal) {
                    g.kind = 4;
                    for (int k = 0; k < 4; ++k) g.m[k] = g.m4[k * 5];
                }
                out.push_back(g);
                break;
            }
            case 5: case 6:
                return fail(QF_REQUIRES_SHOT_EXECUTION, "this circuit operation requires shot execution");
            default:
                return fail(QF_INVALID_ARGUMENT, "unknown op kind " + std::to_string(op.kind));
        }
    }
    return QF_OK;
}

// Plan options for density circuits: the backend's, always fused (the 4x4
// superoperators exist only in tile passes).
qf::PlanOptions density_options(const qf_backend* be) {
    qf::PlanOptions o = be->opts;
    o.fuse = 1;
    return o;
}

// Single-op calls use the LDS tile kernel: one HBM sweep, and no pass compile
// per distinct gate angle.
qf_status apply_one(qf_density* d, const std::vector<qf::GateIn>& g) {
    qf_status s = activate(d->be);
    if (s) return s;
    qf::PlanOptions o = density_options(d->be);
    o.kernel = 1;
    o.reg_bits = 0;
    return apply_gates(d->be, d->d, QF_C128, 2 * d->qubits, 1, g, o);
}

qf_status new_density(qf_backend* be, uint32_t qubits, qf_density** out) {
    if (!be || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    qf_status s = check_density_qubits(qubits);
    if (s) return s;
    if ((s = activate(be))) return s;
    const uint64_t dim = uint64_t(1) << qubits;
    void* p = nullptr;
    if ((s = device_alloc(&p, dim * dim * sizeof(qf_c128)))) return s;
    qf_density* d = new qf_density;
    d->be = be;
    be->refs.fetch_add(1);
    d->d = p;
    d->qubits = qubits;
    d->dim = dim;
    *out = d;
    return QF_OK;
}

qf_status density_reduce(const qf_density* d, int kind, uint32_t q, double* re, double* im) {
    qf_status s = activate(d->be);
    if (s) return s;
    std::vector<double> r;
    const uint64_t n = kind == qf::RED_NORM ? d->dim * d->dim : d->dim;
    if ((s = reduce(d->be, d->d, nullptr, QF_C128, n, 1, kind, q, {}, r))) return s;
    if (re) *re = r[0];
    if (im) *im = r[1];
    return QF_OK;
}

}  // namespace

namespace qfrt {

// validate_probability (src/lib.rs:2229-2235) and NoiseChannel::thermal_rates
// (src/lib.rs:1611-1627).
qf_status check_channel(int32_t kind, double p0, double p1) {
    auto prob = [](double p) { return std::isfinite(p) && p >= 0.0 && p <= 1.0; };
    if (kind < QF_CHANNEL_DEPOLARIZING || kind > QF_CHANNEL_THERMAL_DAMPING)
        return fail(QF_INVALID_ARGUMENT, "unknown noise channel " + std::to_string(kind));
    if (!prob(p0)) return fail(QF_INVALID_PROBABILITY, "probability must be in [0, 1]");
    if (kind == QF_CHANNEL_THERMAL_DAMPING && (!std::isfinite(p1) || p1 < 0.0))
        return fail(QF_INVALID_PROBABILITY, "beta_delta must be finite and non-negative");
    return QF_OK;
}

}  // namespace qfrt

extern "C" qf_status qf_density_upload(qf_backend* be, const qf_c128* entries, size_t len, uint32_t qubits,
                                      qf_density** out) {
    if (!be || !entries || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    qf_status s = check_density_qubits(qubits);
    if (s) return s;
    const uint64_t dim = uint64_t(1) << qubits;
    if (len != dim * dim) return fail(QF_STATE_TOO_LARGE, "entry count does not match 4^qubits");
    if ((s = new_density(be, qubits, out))) return s;
    hipError_t e = hipMemcpy((*out)->d, entries, len * sizeof(qf_c128), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        qf_density_destroy(*out);
        *out = nullptr;
        return hip_fail(e, "hipMemcpy density upload");
    }
    return QF_OK;
}

extern "C" qf_status qf_density_from_state(const qf_state* st, qf_density** out) {
    if (!st || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    if (st->precision != QF_C128) return fail(QF_INCOMPATIBLE_DEVICE_STATE, "density matrices are complex128");
    qf_status s = new_density(st->be, st->qubits, out);
    if (s) return s;
    qf_density* d = *out;
    hipLaunchKernelGGL(k_density_outer, dim3(grid_for(d->dim * d->dim)), dim3(qf::kThreads), 0, d->be->stream,
                       static_cast<const cx<double>*>(st->d), static_cast<cx<double>*>(d->d), d->qubits);
    if ((s = launch_check("k_density_outer"))) {
        qf_density_destroy(d);
        *out = nullptr;
    }
    return s;
}

extern "C" qf_status qf_density_maximally_mixed(qf_backend* be, uint32_t qubits, qf_density** out) {
    qf_status s = new_density(be, qubits, out);
    if (s) return s;
    qf_density* d = *out;
    hipLaunchKernelGGL(k_density_mixed, dim3(grid_for(d->dim * d->dim)), dim3(qf::kThreads), 0, be->stream,
                       static_cast<cx<double>*>(d->d), d->qubits);
    if ((s = launch_check("k_density_mixed"))) {
        qf_density_destroy(d);
        *out = nullptr;
    }
    return s;
}

// DensityMatrix::maximally_mixed_with_memory_budget, src/lib.rs:1668-1700:
// the budget covers the 4^qubits complex128 entries.
extern "C" qf_status qf_density_maximally_mixed_with_memory_budget(qf_backend* be, uint32_t qubits,
                                                                   uint64_t budget_bytes, qf_density** out) {
    if (!be || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (qubits == 0) return fail(QF_ZERO_QUBITS, "a state vector must contain at least one qubit");
    if (qubits >= 32) return fail(QF_SIMULATION_SIZE_OVERFLOW, "dense simulation size exceeds the supported addressable range");
    const unsigned __int128 required = (unsigned __int128)(uint64_t(1) << (2 * qubits)) * 16;
    if (required > budget_bytes) {
        char msg[160];
        std::snprintf(msg, sizeof msg, "dense state needs at least %llu bytes, exceeding the %llu-byte memory budget",
                      (unsigned long long)required, (unsigned long long)budget_bytes);
        return fail(QF_MEMORY_BUDGET_EXCEEDED, msg);
    }
    return qf_density_maximally_mixed(be, qubits, out);
}

extern "C" qf_status qf_density_destroy(qf_density* d) {
    if (!d) return fail(QF_INVALID_ARGUMENT, "null density matrix");
    qf_backend* be = d->be;
    (void)hipSetDevice(be->ordinal);
    (void)hipStreamSynchronize(be->stream);
    (void)hipFree(d->d);
    delete d;
    backend_release(be);
    return QF_OK;
}

extern "C" qf_status qf_density_qubits(const qf_density* d, uint32_t* out) {
    if (!d || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    *out = d->qubits;
    return QF_OK;
}

extern "C" qf_status qf_density_download(const qf_density* d, qf_c128* out, size_t len) {
    if (!d || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    if (len != d->dim * d->dim) return fail(QF_INVALID_ARGUMENT, "output needs 4^qubits entries");
    qf_status s = activate(d->be);
    if (s) return s;
    QF_HIP_TRY(hipMemcpyAsync(out, d->d, len * sizeof(qf_c128), hipMemcpyDeviceToHost, d->be->stream));
    QF_HIP_TRY(hipStreamSynchronize(d->be->stream));
    return QF_OK;
}

extern "C" qf_status qf_density_apply_gate(qf_density* d, int32_t gate, uint32_t target, double theta) {
    if (!d) return fail(QF_INVALID_ARGUMENT, "null density matrix");
    qf_op op{};
    op.kind = 0;
    op.gate = gate;
    op.a = target;
    op.theta = theta;
    std::vector<qf::GateIn> g;
    qf_status s = lower_density(&op, 1, d->qubits, g);
    return s ? s : apply_one(d, g);
}

namespace {
qf_status apply_pair(qf_density* d, int32_t kind, uint32_t a, uint32_t b) {
    if (!d) return fail(QF_INVALID_ARGUMENT, "null density matrix");
    qf_op op{};
    op.kind = kind;
    op.a = a;
    op.b = b;
    std::vector<qf::GateIn> g;
    qf_status s = lower_density(&op, 1, d->qubits, g);
    return s ? s : apply_one(d, g);
}
}  // namespace

extern "C" qf_status qf_density_apply_cnot(qf_density* d, uint32_t control, uint32_t target) {
    return apply_pair(d, 1, control, target);
}
extern "C" qf_status qf_density_apply_cz(qf_density* d, uint32_t first, uint32_t second) {
    return apply_pair(d, 2, first, second);
}
extern "C" qf_status qf_density_apply_swap(qf_density* d, uint32_t first, uint32_t second) {
    return apply_pair(d, 3, first, second);
}

extern "C" qf_status qf_density_apply_channel(qf_density* d, int32_t channel, uint32_t target, double p0, double p1) {
    if (!d) return fail(QF_INVALID_ARGUMENT, "null density matrix");
    qf_op op{};
    op.kind = 4;
    op.gate = channel;
    op.a = target;
    op.theta = p0;
    op.theta2 = p1;
    std::vector<qf::GateIn> g;
    qf_status s = lower_density(&op, 1, d->qubits, g);
    return s ? s : apply_one(d, g);
}

extern "C" qf_status qf_density_measure(qf_density* d, uint32_t qubit, double uniform, uint8_t* outcome) {
    if (!d || !outcome) return fail(QF_INVALID_ARGUMENT, "null argument");
    if (qubit >= d->qubits) return fail(QF_QUBIT_OUT_OF_RANGE, qstr(qubit, d->qubits));
    if (!std::isfinite(uniform) || !(uniform >= 0.0 && uniform < 1.0))
        return fail(QF_INVALID_BATCH_SAMPLES, "measurement needs a uniform sample in [0, 1)");
    double p = 0.0;
    qf_status s = density_reduce(d, qf::RED_DP1, qubit, &p, nullptr);
    if (s) return s;
    const double p_one = std::min(std::max(p, 0.0), 1.0);
    const uint8_t o = uint8_t(uniform < p_one);
    const double probability = o ? p_one : 1.0 - p_one;
    if (probability <= DBL_EPSILON)
        return fail(QF_INVALID_PROBABILITY, "measurement probability must be in [0, 1]");
    const uint64_t total = d->dim * d->dim;
    hipLaunchKernelGGL(k_density_collapse, dim3(grid_for(total)), dim3(qf::kThreads), 0, d->be->stream,
                       static_cast<cx<double>*>(d->d), total, d->qubits, qubit, uint32_t(o), probability);
    if ((s = launch_check("k_density_collapse"))) return s;
    QF_HIP_TRY(hipStreamSynchronize(d->be->stream));
    *outcome = o;
    return QF_OK;
}

extern "C" qf_status qf_density_trace(const qf_density* d, double* re, double* im) {
    if (!d || !re || !im) return fail(QF_INVALID_ARGUMENT, "null argument");
    return density_reduce(d, qf::RED_DTRACE, 0, re, im);
}

extern "C" qf_status qf_density_purity(const qf_density* d, double* out) {
    if (!d || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    return density_reduce(d, qf::RED_NORM, 0, out, nullptr);
}

extern "C" qf_status qf_density_expectation(const qf_density* d, int32_t obs, uint32_t qubit, double* out) {
    if (!d || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    if (obs < QF_OBS_X || obs > QF_OBS_Z) return fail(QF_INVALID_ARGUMENT, "unknown observable");
    if (qubit >= d->qubits) return fail(QF_QUBIT_OUT_OF_RANGE, qstr(qubit, d->qubits));
    double re = 0.0, im = 0.0;
    if (obs == QF_OBS_Z) {
        qf_status s = density_reduce(d, qf::RED_DP1, qubit, nullptr, &re);
        *out = re;
        return s;
    }
    qf_status s = density_reduce(d, qf::RED_DXY, qubit, &re, &im);
    *out = obs == QF_OBS_X ? 2.0 * re : -2.0 * im;
    return s;
}

extern "C" qf_status qf_density_probability_one(const qf_density* d, uint32_t qubit, double* out) {
    if (!d || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    if (qubit >= d->qubits) return fail(QF_QUBIT_OUT_OF_RANGE, qstr(qubit, d->qubits));
    double p = 0.0;
    qf_status s = density_reduce(d, qf::RED_DP1, qubit, &p, nullptr);
    *out = std::min(std::max(p, 0.0), 1.0);
    return s;
}

extern "C" qf_status qf_density_run_circuit(qf_density* d, const qf_op* ops, size_t nops) {
    if (!d || (!ops && nops)) return fail(QF_INVALID_ARGUMENT, "null argument");
    std::vector<qf::GateIn> g;
    qf_status s = lower_density(ops, nops, d->qubits, g);
    if (s || g.empty()) return s;
    if ((s = activate(d->be))) return s;
    return apply_gates(d->be, d->d, QF_C128, 2 * d->qubits, 1, g, density_options(d->be));
}

extern "C" qf_status qf_run_density_circuit(qf_backend* be, const qf_c128* initial, size_t len, uint32_t qubits,
                                           const qf_op* ops, size_t nops, qf_c128* out, size_t out_len) {
    if (!be || !initial || !out || (!ops && nops)) return fail(QF_INVALID_ARGUMENT, "null argument");
    qf_density* d = nullptr;
    qf_status s = qf_density_upload(be, initial, len, qubits, &d);
    if (s) return s;
    s = qf_density_run_circuit(d, ops, nops);
    if (!s) s = qf_density_download(d, out, out_len);
    qf_density_destroy(d);
    return s;
}

extern "C" qf_status qf_density_plan_build(const qf_op* ops, size_t nops, uint32_t qubits,
                                          const qf_plan_options* options, qf_plan** out) {
    if ((!ops && nops) || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    qf_status s = check_density_qubits(qubits);
    if (s) return s;
    std::vector<qf::GateIn> g;
    if ((s = lower_density(ops, nops, qubits, g))) return s;
    qf::PlanOptions o = plan_options_from(options);
    o.fuse = 1;
    qf_plan* p = new qf_plan;
    const std::string e = qf::build_plan(g, 2 * qubits, QF_C128, o, p->plan);
    if (!e.empty()) {
        delete p;
        return fail(QF_INVALID_ARGUMENT, e);
    }
    p->plan.origin = qf::PLAN_DENSITY;
    *out = p;
    return QF_OK;
}

extern "C" qf_status qf_density_run_plan(qf_density* d, const qf_plan* plan) {
    if (!d || !plan) return fail(QF_INVALID_ARGUMENT, "null argument");
    if (plan->plan.origin != qf::PLAN_DENSITY || plan->plan.runtime_rotations)
        return fail(QF_INVALID_ARGUMENT, "plan was not built by qf_density_plan_build");
    if (plan->plan.qubits != 2 * d->qubits || plan->plan.precision != QF_C128)
        return fail(QF_INVALID_ARGUMENT, "plan was built for another density size");
    qf_status s = activate(d->be);
    if (s) return s;
    return run_plan(d->be, d->d, QF_C128, 2 * d->qubits, 1, plan->plan);
}
