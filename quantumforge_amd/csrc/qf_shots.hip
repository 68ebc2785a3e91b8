// Projective measurement, classically conditioned gates, Kraus trajectories
// and seeded shots on resident states (CudaDeviceState::measure,
// CudaDeviceStateBatch::measure_with_uniforms / apply_gate_masked /
// apply_channel_trajectory_with_uniforms, CudaBackend::run_circuit_shots and
// run_circuit_shots_batched, src/cuda.rs:1085-1269, 1386-1442, 1500-1790).
//
// A measurement is one FP64 reduction of P(1) per state (RED_P1), the outcome
// and its probability decided on the host exactly as the CPU path does
// (StateVector::measure, src/lib.rs:1321-1350), and one collapse kernel that
// zeroes the other branch and divides the kept one by sqrt(probability).
// Shots draw their uniforms from SmallRng in the CPU's shot-major order, so
// seeded results equal the CPU's (Circuit::run_shots, src/lib.rs:2843-2911).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "qf_runtime.h"

using qf::cd;
using qf::cx;
using namespace qfrt;

#define HIP_TRY_RET(expr)                                  \
    do {                                                   \
        hipError_t e_ = (expr);                            \
        if (e_ != hipSuccess) return hip_fail(e_, #expr);  \
    } while (0)

namespace {

// SmallRng, src/lib.rs:4881-4912: splitmix64 finalizer seeding, xorshift64
// (13, 7, 17), next_f64 = (x >> 11) * 2^-53.
struct SmallRng {
    uint64_t s;
    explicit SmallRng(uint64_t seed) {
        uint64_t z = seed + 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        s = z ? z : 0x9E3779B97F4A7C15ull;
    }
    uint64_t next_u64() {
        uint64_t x = s;
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        s = x;
        return x;
    }
    double next_f64() { return double(next_u64() >> 11) * (1.0 / double(uint64_t(1) << 53)); }
};

constexpr uint32_t kClassicalBits = 4096;  // run_shots_impl, src/lib.rs:2852-2861

std::string num(double v) {
    char b[64];
    std::snprintf(b, sizeof b, "%.17g", v);
    return b;
}

template <class T>
qf_status upload_small(qf_backend* be, void** buf, size_t* cap, const std::vector<T>& host, T** dev) {
    qf_status s = ensure(be, buf, cap, host.size() * sizeof(T) + 16);
    if (s) return s;
    HIP_TRY_RET(hipMemcpyAsync(*buf, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice, be->stream));
    *dev = static_cast<T*>(*buf);
    return QF_OK;
}

}  // namespace

namespace qfrt {

// Measures `qubit` on `count` contiguous states with one uniform each.
// Validates every outcome before collapsing any state (measure_with_uniforms,
// src/cuda.rs:1567-1585).
qf_status measure_states(qf_backend* be, void* d, int precision, uint32_t qubits, uint32_t count, uint32_t qubit,
                         const double* uniforms, uint8_t* outcomes) {
    const uint64_t dim = uint64_t(1) << qubits;
    std::vector<double> r;
    std::vector<double> p1(count);
    for (uint32_t c0 = 0; c0 < count; c0 += 65535) {
        const uint32_t n = std::min<uint32_t>(65535, count - c0);
        const char* base = static_cast<const char*>(d) + size_t(c0) * dim * amp_bytes(precision);
        qf_status s = reduce(be, base, nullptr, precision, dim, n, qf::RED_P1, qubit, {}, r);
        if (s) return s;
        for (uint32_t i = 0; i < n; ++i) p1[c0 + i] = r[2 * i];
    }
    std::vector<uint8_t> out(count);
    std::vector<double> scale(count);
    for (uint32_t i = 0; i < count; ++i) {
        const double p_one = std::min(std::max(p1[i], 0.0), 1.0);
        const uint8_t o = uint8_t(uniforms[i] < p_one);
        const double prob = o ? p_one : 1.0 - p_one;
        if (prob <= DBL_EPSILON)
            return fail(QF_INVALID_PROBABILITY, "measurement probability must be in [0, 1], got " + num(prob));
        out[i] = o;
        scale[i] = std::sqrt(prob);
    }
    uint8_t* d_out = nullptr;
    double* d_scale = nullptr;
    qf_status s = upload_small(be, &be->meas_a, &be->meas_a_bytes, out, &d_out);
    if (s) return s;
    if ((s = upload_small(be, &be->meas_b, &be->meas_b_bytes, scale, &d_scale))) return s;
    const uint64_t total = dim * count;
    if (precision == QF_C128)
        hipLaunchKernelGGL(qf::k_collapse<double>, dim3(grid_for(total)), dim3(qf::kThreads), 0, be->stream,
                           static_cast<cx<double>*>(d), total, qubits, qubit, d_out, d_scale);
    else
        hipLaunchKernelGGL(qf::k_collapse<float>, dim3(grid_for(total)), dim3(qf::kThreads), 0, be->stream,
                           static_cast<cx<float>*>(d), total, qubits, qubit, d_out, d_scale);
    if ((s = launch_check("k_collapse"))) return s;
    HIP_TRY_RET(hipStreamSynchronize(be->stream));
    std::memcpy(outcomes, out.data(), count);
    return QF_OK;
}

// One-qubit gate on the states with enabled[i] != 0.
qf_status masked_gate(qf_backend* be, void* d, int precision, uint32_t qubits, uint32_t count, int gate,
                      uint32_t target, double theta, const uint8_t* enabled) {
    cd m[4];
    gate_matrix(gate, theta, m);
    std::vector<uint8_t> en(enabled, enabled + count);
    uint8_t* d_en = nullptr;
    qf_status s = upload_small(be, &be->meas_a, &be->meas_a_bytes, en, &d_en);
    if (s) return s;
    const uint64_t npairs = (uint64_t(count) << qubits) / 2;
    if (precision == QF_C128)
        hipLaunchKernelGGL(qf::k_sweep_1q_masked<double>, dim3(grid_for(npairs)), dim3(qf::kThreads), 0, be->stream,
                           static_cast<cx<double>*>(d), npairs, qubits, target, d_en, m[0], m[1], m[2], m[3]);
    else
        hipLaunchKernelGGL(qf::k_sweep_1q_masked<float>, dim3(grid_for(npairs)), dim3(qf::kThreads), 0, be->stream,
                           static_cast<cx<float>*>(d), npairs, qubits, target, d_en, m[0], m[1], m[2], m[3]);
    if ((s = launch_check("k_sweep_1q_masked"))) return s;
    // The host copy of `enabled` must outlive the upload.
    HIP_TRY_RET(hipStreamSynchronize(be->stream));
    return QF_OK;
}

// Rust's saturating float -> u8 cast (`as u8`): NaN and negatives give 0.
static int as_u8(double x) { return !(x > 0.0) ? 0 : (x >= 255.0 ? 255 : int(x)); }

// One Kraus-trajectory step of `channel` on qubit `q` of `count` contiguous
// states, one uniform each, with the branch rules of
// StateVector::apply_channel_trajectory (src/lib.rs:1356-1505): the norm and
// P(1) of each state come from one FP64 reduction, the branch and its scales
// are decided on the host, and one kernel applies every state's branch.  All
// states are checked before any changes.
qf_status trajectory_states(qf_backend* be, void* d, int precision, uint32_t qubits, uint32_t count, int32_t channel,
                            double p0, double p1, uint32_t q, const double* uniforms) {
#pragma clang fp contract(off)
    std::vector<uint8_t> code(count, 0);
    std::vector<double> sc(2 * size_t(count), 0.0);
    if (channel == QF_CHANNEL_DEPOLARIZING) {
        for (uint32_t i = 0; i < count; ++i) {
            const double u = uniforms[i];
            if (u >= 1.0 - p0 && p0 > 0.0)
                code[i] = uint8_t(1 + std::min(as_u8(std::floor((u - (1.0 - p0)) * (3.0 / p0))), 2));
        }
    } else {
        const uint64_t dim = uint64_t(1) << qubits;
        std::vector<double> nrm(count), pone(count), r;
        for (uint32_t c0 = 0; c0 < count; c0 += 65535) {
            const uint32_t n = std::min<uint32_t>(65535, count - c0);
            const char* base = static_cast<const char*>(d) + size_t(c0) * dim * amp_bytes(precision);
            qf_status s = reduce(be, base, nullptr, precision, dim, n, qf::RED_NORM_P1, q, {}, r);
            if (s) return s;
            for (uint32_t i = 0; i < n; ++i) {
                nrm[c0 + i] = r[2 * i];
                pone[c0 + i] = r[2 * i + 1];
            }
        }
        const double up = channel == QF_CHANNEL_THERMAL_DAMPING ? p0 * std::exp(-p1) : 0.0, down = p0;
        for (uint32_t i = 0; i < count; ++i) {
            const double norm = nrm[i], p_one = pone[i], u = uniforms[i];
            if (!std::isfinite(norm) || norm <= DBL_EPSILON)
                return fail(QF_INVALID_PROBABILITY, "state norm must be positive, got " + num(norm));
            double branch_norm;
            if (channel != QF_CHANNEL_THERMAL_DAMPING) {
                const double gamma = p0;
                const double jump_probability = std::min(std::max(gamma * p_one / norm, 0.0), 1.0);
                const bool jump = u < jump_probability;
                branch_norm = jump ? gamma * p_one : norm - gamma * p_one;
                if (!std::isfinite(branch_norm) || branch_norm <= DBL_EPSILON)
                    return fail(QF_INVALID_PROBABILITY, "trajectory branch probability " + num(branch_norm));
                const double inverse = 1.0 / std::sqrt(branch_norm);
                if (jump) {
                    code[i] = channel == QF_CHANNEL_AMPLITUDE_DAMPING ? 5 : 7;
                    sc[2 * i] = std::sqrt(gamma) * inverse;
                } else {
                    code[i] = 4;
                    sc[2 * i] = inverse;
                    sc[2 * i + 1] = std::sqrt(1.0 - gamma) * inverse;
                }
                continue;
            }
            const double pr1 = std::min(std::max(p_one / norm, 0.0), 1.0), pr0 = 1.0 - pr1;
            const double go_down = pr1 * down, go_up = pr0 * up;
            const double stay = pr0 * (1.0 - up) + pr1 * (1.0 - down);
            const int branch = u < go_down ? 1 : (u < go_down + go_up ? 2 : 0);
            branch_norm = branch == 1 ? go_down : (branch == 2 ? go_up : stay);
            if (!std::isfinite(branch_norm) || branch_norm <= DBL_EPSILON)
                return fail(QF_INVALID_PROBABILITY, "trajectory branch probability " + num(branch_norm));
            const double inverse = 1.0 / std::sqrt(branch_norm);
            if (branch == 1) {
                code[i] = 5;
                sc[2 * i] = std::sqrt(down) * inverse;
            } else if (branch == 2) {
                code[i] = 6;
                sc[2 * i] = std::sqrt(up) * inverse;
            } else {
                code[i] = 4;
                sc[2 * i] = std::sqrt(1.0 - up) * inverse;
                sc[2 * i + 1] = std::sqrt(1.0 - down) * inverse;
            }
        }
    }
    if (std::all_of(code.begin(), code.end(), [](uint8_t c) { return c == 0; })) return QF_OK;
    uint8_t* d_code = nullptr;
    double* d_sc = nullptr;
    qf_status s = upload_small(be, &be->meas_a, &be->meas_a_bytes, code, &d_code);
    if (s) return s;
    if ((s = upload_small(be, &be->meas_b, &be->meas_b_bytes, sc, &d_sc))) return s;
    const uint64_t npairs = (uint64_t(count) << qubits) / 2;
    if (precision == QF_C128)
        hipLaunchKernelGGL(qf::k_trajectory<double>, dim3(grid_for(npairs)), dim3(qf::kThreads), 0, be->stream,
                           static_cast<cx<double>*>(d), npairs, qubits, q, d_code, d_sc);
    else
        hipLaunchKernelGGL(qf::k_trajectory<float>, dim3(grid_for(npairs)), dim3(qf::kThreads), 0, be->stream,
                           static_cast<cx<float>*>(d), npairs, qubits, q, d_code, d_sc);
    if ((s = launch_check("k_trajectory"))) return s;
    // The host copies must outlive the uploads.
    HIP_TRY_RET(hipStreamSynchronize(be->stream));
    return QF_OK;
}

}  // namespace qfrt

namespace {

qf_status check_gate(int gate, uint32_t target, double theta, uint32_t qubits) {
    if (gate < QF_GATE_I || gate > QF_GATE_RZ) return fail(QF_INVALID_ARGUMENT, "unknown gate id");
    if (target >= qubits)
        return fail(QF_QUBIT_OUT_OF_RANGE, "qubit index " + std::to_string(target) + " is out of range for " +
                                               std::to_string(qubits) + " qubits");
    if (is_rotation(gate) && !std::isfinite(theta)) return fail(QF_NON_FINITE_ANGLE, "rotation angle must be finite");
    return QF_OK;
}

// One step of a shot circuit: a run of unitary ops (exact plan), a
// measurement, a classically conditioned gate or a noise channel.
struct Step {
    enum Kind { UNITARY, MEASURE, CONDITIONAL, CHANNEL } kind;
    qf::Plan plan;
    uint32_t target = 0, cbit = 0;
    uint8_t value = 0;
    int gate = 0;  // CONDITIONAL: gate id; CHANNEL: QF_CHANNEL_*
    double theta = 0, theta2 = 0;
};

// nrandom counts the ops that draw one uniform per shot: measurements and
// (with noise) channels.
qf_status parse_shot_circuit(const qf_op* ops, size_t nops, uint32_t qubits, bool allow_noise,
                             std::vector<Step>& steps, uint32_t& width, uint32_t& nrandom) {
    width = 0;
    nrandom = 0;
    // Classical-bit range and channels in one pass in op order
    // (run_circuit_shots, src/cuda.rs:1097-1119).
    for (size_t i = 0; i < nops; ++i) {
        const qf_op& op = ops[i];
        if (op.kind == 5 || op.kind == 6) {
            if (op.kind == 6 && (op.b >> 17))
                return fail(QF_INVALID_ARGUMENT, "conditional op: b must be classical_bit | value << 16");
            const uint32_t cbit = op.kind == 5 ? op.b : (op.b & 0xFFFFu);
            if (cbit >= kClassicalBits)
                return fail(QF_CLASSICAL_BIT_OUT_OF_RANGE, "classical bit index " + std::to_string(cbit) +
                                                               " exceeds the supported maximum of 4095");
            width = std::max(width, cbit + 1);
        } else if (op.kind == 4 && !allow_noise) {
            return fail(QF_REQUIRES_DENSITY_MATRIX, "this circuit operation requires density-matrix simulation");
        }
    }
    qf::PlanOptions exact;
    exact.fuse = 0;
    size_t i = 0;
    while (i < nops) {
        const qf_op& op = ops[i];
        if (op.kind == 5) {
            if (op.a >= qubits)
                return fail(QF_QUBIT_OUT_OF_RANGE, "qubit index " + std::to_string(op.a) + " is out of range for " +
                                                       std::to_string(qubits) + " qubits");
            Step st{Step::MEASURE};
            st.target = op.a;
            st.cbit = op.b;
            steps.push_back(std::move(st));
            ++nrandom;
            ++i;
            continue;
        }
        if (op.kind == 4) {
            if (op.a >= qubits)
                return fail(QF_QUBIT_OUT_OF_RANGE, "qubit index " + std::to_string(op.a) + " is out of range for " +
                                                       std::to_string(qubits) + " qubits");
            qf_status s = check_channel(op.gate, op.theta, op.theta2);
            if (s) return s;
            Step st{Step::CHANNEL};
            st.gate = op.gate;
            st.target = op.a;
            st.theta = op.theta;
            st.theta2 = op.theta2;
            steps.push_back(std::move(st));
            ++nrandom;
            ++i;
            continue;
        }
        if (op.kind == 6) {
            qf_status s = check_gate(op.gate, op.a, op.theta, qubits);
            if (s) return s;
            Step st{Step::CONDITIONAL};
            st.gate = op.gate;
            st.target = op.a;
            st.theta = op.theta;
            st.cbit = op.b & 0xFFFFu;
            st.value = uint8_t((op.b >> 16) & 1u);
            steps.push_back(std::move(st));
            ++i;
            continue;
        }
        size_t j = i;
        while (j < nops && ops[j].kind != 4 && ops[j].kind != 5 && ops[j].kind != 6) ++j;
        std::vector<qf::GateIn> g;
        qf_status s = lower_ops(ops + i, j - i, qubits, g);
        if (s) return s;
        Step st{Step::UNITARY};
        const std::string e = qf::build_plan(g, qubits, QF_C128, exact, st.plan);
        if (!e.empty()) return fail(QF_INVALID_ARGUMENT, e);
        steps.push_back(std::move(st));
        i = j;
    }
    return QF_OK;
}

// Shots one after another on one state, restored from a device copy of the
// initial state before each shot (run_circuit_shots, src/cuda.rs:1085-1159).
qf_status shots_sequential(qf_backend* be, const qf_c128* initial, uint32_t qubits, const std::vector<Step>& steps,
                           uint64_t shots, uint64_t seed, uint32_t width, uint8_t* registers) {
    const size_t bytes = (size_t(1) << qubits) * sizeof(qf_c128);
    void* d0 = nullptr;
    void* d = nullptr;
    qf_status s = device_alloc(&d0, bytes);
    if (s) return s;
    if ((s = device_alloc(&d, bytes))) {
        (void)hipFree(d0);
        return s;
    }
    auto done = [&](qf_status r) {
        (void)hipStreamSynchronize(be->stream);
        (void)hipFree(d0);
        (void)hipFree(d);
        return r;
    };
    if (hipMemcpy(d0, initial, bytes, hipMemcpyHostToDevice) != hipSuccess)
        return done(fail(QF_DEVICE_ERROR, "copy of the initial state failed"));
    SmallRng rng(seed);
    std::vector<uint8_t> classical(width);
    for (uint64_t shot = 0; shot < shots; ++shot) {
        if (hipMemcpyAsync(d, d0, bytes, hipMemcpyDeviceToDevice, be->stream) != hipSuccess)
            return done(fail(QF_DEVICE_ERROR, "device copy of the initial state failed"));
        std::fill(classical.begin(), classical.end(), 0);
        for (const Step& st : steps) {
            switch (st.kind) {
                case Step::UNITARY: s = run_plan(be, d, QF_C128, qubits, 1, st.plan); break;
                case Step::MEASURE: {
                    const double u = rng.next_f64();
                    uint8_t o = 0;
                    s = measure_states(be, d, QF_C128, qubits, 1, st.target, &u, &o);
                    classical[st.cbit] = o;
                    break;
                }
                case Step::CONDITIONAL: {
                    const uint8_t en = classical[st.cbit] == st.value;
                    s = en ? masked_gate(be, d, QF_C128, qubits, 1, st.gate, st.target, st.theta, &en) : QF_OK;
                    break;
                }
                case Step::CHANNEL:
                    s = fail(QF_REQUIRES_DENSITY_MATRIX, "this circuit operation requires density-matrix simulation");
                    break;
            }
            if (s) return done(s);
        }
        if (width) std::memcpy(registers + shot * width, classical.data(), width);
    }
    return done(QF_OK);
}

__global__ void k_broadcast(cx<double>* __restrict__ dst, const cx<double>* __restrict__ src, uint64_t dim,
                            uint64_t total) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += stride)
        dst[i] = src[i & (dim - 1)];
}

// All shots as batches of states (run_circuit_shots_batched,
// src/cuda.rs:1161-1269): the uniforms are drawn shot-major up front, then
// every step runs on the whole batch.  Batches are sized to free device
// memory; shots are independent, so the split does not change any result.
// With noise channels the initial state is first normalized on the host as
// StateVector::normalize does (src/lib.rs:1005-1014).
qf_status shots_batched(qf_backend* be, const qf_c128* initial_in, uint32_t qubits, const std::vector<Step>& steps,
                        uint64_t shots, uint64_t seed, uint32_t width, uint32_t nrandom, uint8_t* registers) {
    SmallRng rng(seed);
    std::vector<double> draws(size_t(shots) * nrandom);
    for (double& x : draws) x = rng.next_f64();
    const uint64_t dim = uint64_t(1) << qubits;
    const size_t per = size_t(dim) * sizeof(qf_c128);
    const qf_c128* initial = initial_in;
    std::vector<qf_c128> normalized;
    if (std::any_of(steps.begin(), steps.end(), [](const Step& st) { return st.kind == Step::CHANNEL; })) {
#pragma clang fp contract(off)
        double acc = 0.0;
        for (uint64_t i = 0; i < dim; ++i) acc += initial_in[i].re * initial_in[i].re + initial_in[i].im * initial_in[i].im;
        const double norm = std::sqrt(acc);
        if (norm == 0.0 || !std::isfinite(norm))
            return fail(QF_INVALID_PROBABILITY, "initial state norm must be positive, got " + num(norm));
        normalized.resize(dim);
        for (uint64_t i = 0; i < dim; ++i) normalized[i] = qf_c128{initial_in[i].re / norm, initial_in[i].im / norm};
        initial = normalized.data();
    }
    size_t freeb = 0, totalb = 0;
    if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) freeb = size_t(8) << 30;
    uint64_t chunk = uint64_t(double(freeb) * 0.7) / (per ? per : 1);
    chunk = std::max<uint64_t>(1, std::min<uint64_t>({chunk - (chunk > 1 ? 1 : 0), shots, uint64_t(1) << 20}));
    void* d0 = nullptr;
    qf_status s = device_alloc(&d0, per);
    if (s) return s;
    if (hipMemcpy(d0, initial, per, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d0);
        return fail(QF_DEVICE_ERROR, "copy of the initial state failed");
    }
    void* d = nullptr;
    if ((s = device_alloc(&d, per * chunk))) {
        (void)hipFree(d0);
        return s;
    }
    auto done = [&](qf_status r) {
        (void)hipStreamSynchronize(be->stream);
        (void)hipFree(d0);
        (void)hipFree(d);
        return r;
    };
    std::vector<uint8_t> classical;
    std::vector<double> u;
    std::vector<uint8_t> o, en;
    for (uint64_t first = 0; first < shots; first += chunk) {
        const uint32_t n = uint32_t(std::min<uint64_t>(chunk, shots - first));
        hipLaunchKernelGGL(k_broadcast, dim3(grid_for(dim * n)), dim3(qf::kThreads), 0, be->stream,
                           static_cast<cx<double>*>(d), static_cast<const cx<double>*>(d0), dim, dim * n);
        if ((s = launch_check("k_broadcast"))) return done(s);
        classical.assign(size_t(n) * width, 0);
        uint32_t mi = 0;
        for (const Step& st : steps) {
            switch (st.kind) {
                case Step::UNITARY: s = run_plan(be, d, QF_C128, qubits, n, st.plan); break;
                case Step::MEASURE: {
                    u.resize(n);
                    o.resize(n);
                    for (uint32_t k = 0; k < n; ++k) u[k] = draws[(first + k) * nrandom + mi];
                    s = measure_states(be, d, QF_C128, qubits, n, st.target, u.data(), o.data());
                    for (uint32_t k = 0; k < n; ++k) classical[size_t(k) * width + st.cbit] = o[k];
                    ++mi;
                    break;
                }
                case Step::CHANNEL: {
                    u.resize(n);
                    for (uint32_t k = 0; k < n; ++k) u[k] = draws[(first + k) * nrandom + mi];
                    s = trajectory_states(be, d, QF_C128, qubits, n, st.gate, st.theta, st.theta2, st.target, u.data());
                    ++mi;
                    break;
                }
                case Step::CONDITIONAL: {
                    en.resize(n);
                    bool any = false;
                    for (uint32_t k = 0; k < n; ++k) {
                        en[k] = classical[size_t(k) * width + st.cbit] == st.value;
                        any = any || en[k];
                    }
                    s = any ? masked_gate(be, d, QF_C128, qubits, n, st.gate, st.target, st.theta, en.data()) : QF_OK;
                    break;
                }
            }
            if (s) return done(s);
        }
        if (width) std::memcpy(registers + first * width, classical.data(), classical.size());
    }
    return done(QF_OK);
}

}  // namespace

extern "C" qf_status qf_state_measure(qf_state* st, uint32_t qubit, double uniform, uint8_t* outcome) {
    if (!st || !outcome) return fail(QF_INVALID_ARGUMENT, "null argument");
    if (qubit >= st->qubits)
        return fail(QF_QUBIT_OUT_OF_RANGE, "qubit index " + std::to_string(qubit) + " is out of range for " +
                                               std::to_string(st->qubits) + " qubits");
    if (!std::isfinite(uniform) || !(uniform >= 0.0 && uniform < 1.0))
        return fail(QF_INVALID_BATCH_SAMPLES, "measurement needs a uniform sample in [0, 1)");
    qf_status s = activate(st->be);
    if (s) return s;
    return measure_states(st->be, st->d, st->precision, st->qubits, 1, qubit, &uniform, outcome);
}

extern "C" qf_status qf_batch_measure_with_uniforms(qf_batch* b, uint32_t qubit, const double* uniforms,
                                                   size_t count, uint8_t* outcomes) {
    if (!b || !outcomes || (!uniforms && count)) return fail(QF_INVALID_ARGUMENT, "null argument");
    if (b->uneven()) return fail(QF_INCOMPATIBLE_BATCH, "this operation needs a batch of equal widths");
    if (qubit >= b->qubits)
        return fail(QF_QUBIT_OUT_OF_RANGE, "qubit index " + std::to_string(qubit) + " is out of range for " +
                                               std::to_string(b->qubits) + " qubits");
    bool ok = count == b->count;
    for (size_t i = 0; ok && i < count; ++i)
        ok = std::isfinite(uniforms[i]) && uniforms[i] >= 0.0 && uniforms[i] < 1.0;
    if (!ok) return fail(QF_INVALID_BATCH_SAMPLES, "batch measurement needs one uniform sample per state");
    qf_status s = activate(b->be);
    if (s) return s;
    return measure_states(b->be, b->d, b->precision, b->qubits, b->count, qubit, uniforms, outcomes);
}

extern "C" qf_status qf_batch_apply_gate_masked(qf_batch* b, int32_t gate, uint32_t target, double theta,
                                               const uint8_t* enabled, size_t count) {
    if (!b || (!enabled && count)) return fail(QF_INVALID_ARGUMENT, "null argument");
    if (b->uneven()) return fail(QF_INCOMPATIBLE_BATCH, "this operation needs a batch of equal widths");
    qf_status s = check_gate(gate, target, theta, b->qubits);
    if (s) return s;
    if (count != b->count) return fail(QF_INVALID_BATCH_SAMPLES, "the mask needs one entry per state");
    if ((s = activate(b->be))) return s;
    return masked_gate(b->be, b->d, b->precision, b->qubits, b->count, gate, target, theta, enabled);
}

extern "C" qf_status qf_run_circuit_shots(qf_backend* be, const qf_c128* initial, uint32_t qubits, const qf_op* ops,
                                         size_t nops, uint64_t shots, uint64_t seed, int32_t batched,
                                         uint8_t* registers, uint32_t* classical_width) {
    if (!be || !initial || (!ops && nops) || !classical_width) return fail(QF_INVALID_ARGUMENT, "null argument");
    if (shots == 0) return fail(QF_INVALID_SHOT_COUNT, "shot count must be greater than zero");
    qf_status s = check_qubits(qubits);
    if (s) return s;
    std::vector<Step> steps;
    uint32_t width = 0, nrandom = 0;
    if ((s = parse_shot_circuit(ops, nops, qubits, batched != 0, steps, width, nrandom))) return s;
    *classical_width = width;
    if (!registers) return QF_OK;
    if ((s = activate(be))) return s;
    return batched ? shots_batched(be, initial, qubits, steps, shots, seed, width, nrandom, registers)
                   : shots_sequential(be, initial, qubits, steps, shots, seed, width, registers);
}

extern "C" qf_status qf_batch_apply_channel_trajectory(qf_batch* b, int32_t channel, uint32_t target, double p0,
                                                      double p1, const double* uniforms, size_t count) {
    if (!b || (!uniforms && count)) return fail(QF_INVALID_ARGUMENT, "null argument");
    if (b->uneven()) return fail(QF_INCOMPATIBLE_BATCH, "this operation needs a batch of equal widths");
    if (target >= b->qubits)
        return fail(QF_QUBIT_OUT_OF_RANGE, "qubit index " + std::to_string(target) + " is out of range for " +
                                               std::to_string(b->qubits) + " qubits");
    bool ok = count == b->count;
    for (size_t i = 0; ok && i < count; ++i)
        ok = std::isfinite(uniforms[i]) && uniforms[i] >= 0.0 && uniforms[i] < 1.0;
    if (!ok) return fail(QF_INVALID_BATCH_SAMPLES, "a trajectory step needs one uniform sample per state");
    qf_status s = check_channel(channel, p0, p1);
    if (s) return s;
    if ((s = activate(b->be))) return s;
    return trajectory_states(b->be, b->d, b->precision, b->qubits, b->count, channel, p0, p1, target, uniforms);
}
