// Gradients of parameterized circuits on the device.
//
// Adjoint (two-state reverse sweep):
//   forward:  psi_b = U(theta_b)|0..0> for every parameter set b (tile passes;
//             rotation angles read per instance from an angle table);
//   costate:  lambda_b = O psi_b and E_b = <psi_b|lambda_b> in one kernel;
//   reverse:  over the inverted ops in reverse order, each rotation
//             occurrence first records Im <lambda|G|psi> (G its generator)
//             per tile, then psi and lambda are both multiplied by U_k^dagger.
// This replaces the reference's store-every-state adjoint
// (src/lib.rs:3564-3704, O(ops * 2^n) memory) with O(2 * 2^n) per set; the
// accumulated quantity is the same: dE/dtheta = 2 Re <lambda|(-i/2) G|psi_after>
// = Im <lambda|G|psi_after> (src/lib.rs:3649-3664).
//
// Parameter shift / finite differences: the 2 x (rotation occurrences)
// shifted circuits of every set run as one batch of instances whose angle
// table rows differ in the shifted occurrence (src/lib.rs:3516-3546,
// src/gradients.rs:344-375).
//
// Backprop: the reference's own store-every-state reverse mode
// (src/gradients.rs:392-506) with exact per-gate sweeps, kept as an
// independent implementation to cross-check the adjoint.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "qf_runtime.h"

using qf::cd;
using qf::cx;
using namespace qfrt;

struct qf_backend;
struct qf_state;

namespace qfrt {

// Host-side helpers shared with the uneven-batch engine.
qf_status qubit_error(uint32_t q, uint32_t qubits) {
    return fail(QF_QUBIT_OUT_OF_RANGE,
                "qubit index " + std::to_string(q) + " is out of range for " + std::to_string(qubits) + " qubits");
}

// validate_inputs / validate_parameters (src/lib.rs:4651-4675) for one set.
qf_status validate_values(const PCirc& pc, const double* inputs, size_t inputs_len, const double* params,
                          size_t params_len) {
    if (inputs_len != pc.ninputs)
        return fail(QF_INVALID_INPUT_COUNT, "expected " + std::to_string(pc.ninputs) + " encoded inputs, got " +
                                                std::to_string(inputs_len));
    for (size_t i = 0; i < inputs_len; ++i)
        if (!std::isfinite(inputs[i])) return fail(QF_NON_FINITE_PARAMETER, "circuit parameters must be finite");
    if (params_len != pc.nparams)
        return fail(QF_INVALID_PARAMETER_COUNT, "expected " + std::to_string(pc.nparams) + " circuit parameters, got " +
                                                    std::to_string(params_len));
    for (size_t i = 0; i < params_len; ++i)
        if (!std::isfinite(params[i])) return fail(QF_NON_FINITE_PARAMETER, "circuit parameters must be finite");
    return QF_OK;
}

// Structural checks in op order (run_with_shift, src/lib.rs:4686-4736) and the
// rotation-occurrence table.
qf_status validate_ops(PCirc& pc, uint32_t qubits) {
    pc.slot_param.clear();
    pc.slot_op.clear();
    for (size_t i = 0; i < pc.nops; ++i) {
        const qf_pop& op = pc.ops[i];
        switch (op.kind) {
            case 0:
                if (op.gate < QF_GATE_I || op.gate > QF_GATE_RZ) return fail(QF_INVALID_ARGUMENT, "unknown gate id");
                if (op.a >= qubits) return qubit_error(op.a, qubits);
                if (is_rotation(op.gate) && !std::isfinite(op.theta))
                    return fail(QF_NON_FINITE_ANGLE, "rotation angle must be finite");
                break;
            case 1:
            case 2:
                if (op.gate < 0 || op.gate > 2) return fail(QF_INVALID_ARGUMENT, "unknown rotation axis");
                if (op.kind == 1 && op.index >= pc.nparams)
                    return fail(QF_INVALID_PARAMETER_COUNT, "expected " + std::to_string(pc.nparams) +
                                                                " circuit parameters, got " +
                                                                std::to_string(op.index + 1));
                if (op.kind == 2 && op.index >= pc.ninputs)
                    return fail(QF_INVALID_INPUT_COUNT, "expected " + std::to_string(pc.ninputs) +
                                                            " encoded inputs, got " + std::to_string(op.index + 1));
                if (op.a >= qubits) return qubit_error(op.a, qubits);
                if (op.kind == 1) {
                    pc.slot_param.push_back(int32_t(op.index));
                    pc.slot_op.push_back(i);
                }
                break;
            case 3:
            case 4:
            case 5: {
                qf_status s = check_pair(qubits, op.a, op.b);
                if (s) return s;
                break;
            }
            case 6: return fail(QF_REQUIRES_DENSITY_MATRIX, "this circuit operation requires density-matrix simulation");
            default: return fail(QF_INVALID_ARGUMENT, "unknown op kind " + std::to_string(op.kind));
        }
    }
    return QF_OK;
}

void adjoint_gate_id(int gate, double theta, int* g, double* th) {  // adjoint_gate, src/lib.rs:2370-2382
    *th = theta;
    switch (gate) {
        case QF_GATE_RX: case QF_GATE_RY: case QF_GATE_RZ: *g = gate; *th = -theta; break;
        case QF_GATE_S: *g = QF_GATE_SDG; break;
        case QF_GATE_SDG: *g = QF_GATE_S; break;
        case QF_GATE_T: *g = QF_GATE_TDG; break;
        case QF_GATE_TDG: *g = QF_GATE_T; break;
        default: *g = gate; break;
    }
}

// Planner gates of the circuit (forward) or of its inverse in reverse order
// with every rotation flagged TF_NEGATE | TF_RECORD (reverse).
std::vector<qf::GateIn> param_gates(const PCirc& pc, const double* inputs, bool reverse) {
    std::vector<qf::GateIn> out;
    out.reserve(pc.nops);
    std::vector<int32_t> slot_of(pc.nops, -1);
    for (size_t s = 0; s < pc.slot_op.size(); ++s) slot_of[pc.slot_op[s]] = int32_t(s);
    for (size_t k = 0; k < pc.nops; ++k) {
        const size_t i = reverse ? pc.nops - 1 - k : k;
        const qf_pop& op = pc.ops[i];
        qf::GateIn g{};
        g.a = op.a;
        g.b = op.b;
        switch (op.kind) {
            case 0: {
                int gid = op.gate;
                double th = op.theta;
                if (reverse) adjoint_gate_id(op.gate, op.theta, &gid, &th);
                g.kind = 0;
                gate_matrix(gid, th, g.m);
                break;
            }
            case 1:
                g.kind = 5;
                g.axis = op.gate;
                g.param = int32_t(op.index);
                g.slot = slot_of[i];
                g.flags = reverse ? (qf::TF_NEGATE | qf::TF_RECORD) : 0;
                break;
            case 2: {
                const double th = reverse ? -inputs[op.index] : inputs[op.index];
                g.kind = 0;
                gate_matrix(QF_GATE_RX + op.gate, th, g.m);
                break;
            }
            case 3: g.kind = 1; break;
            case 4: g.kind = 2; break;
            default: g.kind = 3; break;
        }
        out.push_back(g);
    }
    return out;
}

// Angle table row: (cos(theta/2), sin(theta/2)) per rotation occurrence with
// theta = params[param] (+ delta on occurrence `shift_slot`), computed on the
// host as Gate::matrix does (src/lib.rs:2303-2331).  scaled (the rows of a
// plan with Plan::scaled_rotations): X and Y occurrences hold instead
// (f, u) with f the larger of c = cos, s = sin in magnitude and u the other
// over f, the lowest mantissa bit of u set when f = s; the rotation is then
// f [[1, -u], [u, 1]] (Ry, f = c) or f [[u, -1], [1, u]] (f = s), and Rx
// likewise with -i u, -i.
void fill_row(const PCirc& pc, const double* params, int64_t shift_slot, double delta, double* row, bool scaled) {
    for (size_t s = 0; s < pc.slot_param.size(); ++s) {
        double th = params[pc.slot_param[s]];
        if (int64_t(s) == shift_slot) th = th + delta;
        const double c = host_cos(th * 0.5), sn = host_sin(th * 0.5);
        if (!scaled || pc.ops[pc.slot_op[s]].gate == 2) {
            row[2 * s] = c;
            row[2 * s + 1] = sn;
            continue;
        }
        const bool by_s = std::fabs(sn) > std::fabs(c);
        uint64_t ub;
        const double u = by_s ? c / sn : sn / c;
        std::memcpy(&ub, &u, 8);
        ub = (ub & ~uint64_t(1)) | uint64_t(by_s);
        row[2 * s] = by_s ? sn : c;
        std::memcpy(&row[2 * s + 1], &ub, 8);
    }
}


// Runs f(b) for b in [0, count) on up to 16 host threads (the angle tables
// of large batches: one cos and one sin per occurrence and row), about one
// thread per 8192 occurrences (row_work occurrences a row).
template <class F>
void parallel_rows(uint64_t count, F f, uint64_t row_work = 32) {
    const uint64_t hw = std::max<uint64_t>(1, std::min<uint64_t>(16, std::thread::hardware_concurrency()));
    const uint64_t nt = std::min<uint64_t>({hw, count * std::max<uint64_t>(row_work, 1) / 8192 + 1, count});
    if (nt <= 1) {
        for (uint64_t b = 0; b < count; ++b) f(b);
        return;
    }
    std::vector<std::thread> pool;
    for (uint64_t w = 0; w < nt; ++w)
        pool.emplace_back([&, w] {
            for (uint64_t b = count * w / nt; b < count * (w + 1) / nt; ++b) f(b);
        });
    for (auto& t : pool) t.join();
}

// Observable terms sorted by x mask (k_pauli_apply reads psi[i ^ x] once per x).
std::vector<qf::PauliDev> sorted_terms(const std::vector<qf::PauliDev>& t) {
    std::vector<qf::PauliDev> s(t);
    std::stable_sort(s.begin(), s.end(), [](const qf::PauliDev& a, const qf::PauliDev& b) { return a.x < b.x; });
    return s;
}

// The sorted terms' runs of one x mask: terms [t0, t1) share psi[i ^ x].
struct PauliGroup {
    uint64_t x;
    uint32_t t0, t1;
};

// Device image of a sorted term list: the terms, then their x-mask groups.
std::vector<unsigned char> pack_terms(const std::vector<qf::PauliDev>& sorted, uint32_t* ngroups) {
    std::vector<PauliGroup> g;
    for (uint32_t t = 0; t < sorted.size(); ++t)
        if (g.empty() || g.back().x != sorted[t].x) g.push_back(PauliGroup{sorted[t].x, t, t + 1});
        else g.back().t1 = t + 1;
    *ngroups = uint32_t(g.size());
    std::vector<unsigned char> out(std::max<size_t>(sorted.size() * sizeof(qf::PauliDev) + g.size() * sizeof(PauliGroup), 16));
    if (!sorted.empty()) std::memcpy(out.data(), sorted.data(), sorted.size() * sizeof(qf::PauliDev));
    if (!g.empty()) std::memcpy(out.data() + sorted.size() * sizeof(qf::PauliDev), g.data(), g.size() * sizeof(PauliGroup));
    return out;
}

// Windowed form of a sorted term list for k_pauli_windows.  The terms of one
// x mask act on psi[i ^ x] with the coefficient
//   C_x(i) = sum_t c'_t (-1)^popcount((i ^ x) & z_t),
// which depends only on the bits of i in the union of the z masks.  Terms are
// packed into windows of at most kPauliWindowBits consecutive index bits that
// hold their whole z mask; a window's 2^w coefficients (every sign pattern,
// x folded in) are tabulated on the host in FP64, so the device takes one
// table lookup per window instead of one popcount and sign per term.  Terms
// whose z mask spans more bits stay per-term ("wide").
struct PWGroup {
    uint64_t x;
    uint32_t w0, w1;  // windows [w0, w1)
    uint32_t t0, t1;  // wide terms [t0, t1)
};
struct PWWin {
    uint32_t lo, mask, off, pad;  // key = (i >> lo) & mask, coefficient at off + key
};
struct PWHeader {
    uint32_t ngroups, nwin, ncoef, nwide;
    uint64_t groups_off, wins_off, coef_off, wide_off;  // byte offsets in the pack
};
constexpr uint32_t kPauliWindowBits = 6;
constexpr uint32_t kPauliMaxCoef = 2048;  // 32 KiB of LDS

// Returns an empty vector when the windowed tables would exceed kPauliMaxCoef
// entries (k_pauli_apply takes such term lists).
std::vector<unsigned char> pack_pauli_windows(const std::vector<qf::PauliDev>& sorted) {
    std::vector<PWGroup> groups;
    std::vector<PWWin> wins;
    std::vector<double> coef;  // (re, im) pairs
    std::vector<qf::PauliDev> wide;
    size_t t = 0;
    while (t < sorted.size()) {
        size_t e = t;
        while (e < sorted.size() && sorted[e].x == sorted[t].x) ++e;
        PWGroup g{sorted[t].x, uint32_t(wins.size()), 0, uint32_t(wide.size()), 0};
        // terms of this x mask by the lowest bit of their z mask (z = 0 first)
        std::vector<size_t> idx;
        for (size_t k = t; k < e; ++k) idx.push_back(k);
        std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
            const uint64_t za = sorted[a].z, zb = sorted[b].z;
            const int la = za ? __builtin_ctzll(za) : -1, lb = zb ? __builtin_ctzll(zb) : -1;
            return la < lb;
        });
        std::vector<bool> used(idx.size(), false);
        for (size_t a = 0; a < idx.size(); ++a) {
            if (used[a]) continue;
            const uint64_t za = sorted[idx[a]].z;
            if (za && 64 - __builtin_clzll(za) - __builtin_ctzll(za) > int(kPauliWindowBits)) {
                wide.push_back(sorted[idx[a]]);
                used[a] = true;
                continue;
            }
            const uint32_t lo = za ? uint32_t(__builtin_ctzll(za)) : 0;
            const uint64_t wmask = ((uint64_t(1) << kPauliWindowBits) - 1) << lo;
            std::vector<size_t> members;
            uint64_t span = 0;
            for (size_t b = a; b < idx.size(); ++b) {
                const uint64_t zb = sorted[idx[b]].z;
                if (!used[b] && (zb & ~wmask) == 0) {
                    used[b] = true;
                    members.push_back(idx[b]);
                    span |= zb;
                }
            }
            const uint32_t width = span ? uint32_t(64 - __builtin_clzll(span)) - lo : 0;
            const uint32_t size = 1u << width;
            PWWin w{lo, size - 1, uint32_t(coef.size() / 2), 0};
            for (uint32_t key = 0; key < size; ++key) {
                const uint64_t bits = (uint64_t(key) << lo) ^ g.x;
                double cr = 0.0, ci = 0.0;
                for (size_t m : members) {
                    const double sg = (__builtin_popcountll(bits & sorted[m].z) & 1) ? -1.0 : 1.0;
                    cr += sg * sorted[m].cr;
                    ci += sg * sorted[m].ci;
                }
                coef.push_back(cr);
                coef.push_back(ci);
            }
            wins.push_back(w);
        }
        g.w1 = uint32_t(wins.size());
        g.t1 = uint32_t(wide.size());
        groups.push_back(g);
        t = e;
    }
    if (coef.size() / 2 > kPauliMaxCoef) return {};
    auto al = [](size_t v) { return (v + 15) / 16 * 16; };
    PWHeader h{uint32_t(groups.size()), uint32_t(wins.size()), uint32_t(coef.size() / 2), uint32_t(wide.size()),
               0, 0, 0, 0};
    h.groups_off = al(sizeof(PWHeader));
    h.wins_off = al(h.groups_off + groups.size() * sizeof(PWGroup));
    h.coef_off = al(h.wins_off + wins.size() * sizeof(PWWin));
    h.wide_off = al(h.coef_off + coef.size() * sizeof(double));
    std::vector<unsigned char> out(al(h.wide_off + wide.size() * sizeof(qf::PauliDev)) + 16, 0);
    std::memcpy(out.data(), &h, sizeof h);
    if (!groups.empty()) std::memcpy(out.data() + h.groups_off, groups.data(), groups.size() * sizeof(PWGroup));
    if (!wins.empty()) std::memcpy(out.data() + h.wins_off, wins.data(), wins.size() * sizeof(PWWin));
    if (!coef.empty()) std::memcpy(out.data() + h.coef_off, coef.data(), coef.size() * sizeof(double));
    if (!wide.empty()) std::memcpy(out.data() + h.wide_off, wide.data(), wide.size() * sizeof(qf::PauliDev));
    return out;
}

}  // namespace qfrt

namespace {

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// lambda = O psi and the energy partials conj(psi_i) lambda_i (FP64,
// compensated) from the windowed term tables (pack_pauli_windows).  The grid
// is 8 x (blocks / 8): workgroup b runs on XCD b mod 8 (round-robin dispatch)
// and takes logical block (b mod 8) * (blocks / 8) + b / 8, so every XCD walks
// its own contiguous run of instances and the partner amplitudes
// psi[i ^ x] of an instance are read from that XCD's L2 rather than from HBM
// once per x mask.  Logical block L is (instance L / eblk, chunk L % eblk);
// its partial goes to the same slot as k_pauli_apply's, so k_reduce_finish is
// shared.
template <class R>
__global__ __launch_bounds__(qf::kThreads) void k_pauli_windows(const cx<R>* __restrict__ psi,
                                                                cx<R>* __restrict__ lam, uint32_t n,
                                                                const unsigned char* __restrict__ pack,
                                                                uint32_t cnt, uint32_t eblk,
                                                                qf::dd* __restrict__ partial) {
    extern __shared__ double2 coef_s[];
    const PWHeader* __restrict__ h = reinterpret_cast<const PWHeader*>(pack);
    const PWGroup* __restrict__ groups = reinterpret_cast<const PWGroup*>(pack + h->groups_off);
    const PWWin* __restrict__ wins = reinterpret_cast<const PWWin*>(pack + h->wins_off);
    const double2* __restrict__ coef = reinterpret_cast<const double2*>(pack + h->coef_off);
    const qf::PauliDev* __restrict__ wide = reinterpret_cast<const qf::PauliDev*>(pack + h->wide_off);
    const uint32_t per = gridDim.x >> 3;
    const uint64_t L = uint64_t(blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (L >= uint64_t(cnt) * eblk) return;  // the whole workgroup leaves together
    for (uint32_t k = threadIdx.x; k < h->ncoef; k += blockDim.x) coef_s[k] = coef[k];
    __syncthreads();
    const uint64_t inst = L / eblk, chunk = L % eblk;
    const uint64_t dim = uint64_t(1) << n;
    const cx<R>* __restrict__ p = psi + inst * dim;
    cx<R>* __restrict__ l = lam ? lam + inst * dim : nullptr;
    const uint32_t ngroups = h->ngroups;
    qf::dd er{0, 0}, ei{0, 0};
    // Two amplitudes per thread and step, i and i + stride (the thread's
    // next one, so its additions keep their order): each group's window and
    // term records are read once for both, and the gathers of 4 groups for
    // both amplitudes are in flight together.
    const uint64_t stride = uint64_t(eblk) * blockDim.x;
    for (uint64_t i = chunk * blockDim.x + threadIdx.x; i < dim; i += 2 * stride) {
        const uint64_t j = i + stride;
        const bool two = j < dim;
        double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
        for (uint32_t g0 = 0; g0 < ngroups; g0 += 4) {
            cx<R> v[4], w2[4];
#pragma unroll
            for (uint32_t u = 0; u < 4; ++u)
                if (g0 + u < ngroups) {
                    const uint64_t x = groups[g0 + u].x;
                    v[u] = p[i ^ x];
                    if (two) w2[u] = p[j ^ x];
                }
#pragma unroll
            for (uint32_t u = 0; u < 4; ++u) {
                if (g0 + u >= ngroups) break;
                const PWGroup G = groups[g0 + u];
                double cr = 0.0, ci = 0.0, dr = 0.0, di = 0.0;
                for (uint32_t w = G.w0; w < G.w1; ++w) {
                    const PWWin W = wins[w];
                    const double2 c = coef_s[W.off + (uint32_t(i >> W.lo) & W.mask)];
                    const double2 d = coef_s[W.off + (uint32_t(j >> W.lo) & W.mask)];
                    cr += c.x;
                    ci += c.y;
                    dr += d.x;
                    di += d.y;
                }
                for (uint32_t t = G.t0; t < G.t1; ++t) {
                    const qf::PauliDev& T = wide[t];
                    const double sg = (__popcll((i ^ G.x) & T.z) & 1) ? -1.0 : 1.0;
                    const double sh = (__popcll((j ^ G.x) & T.z) & 1) ? -1.0 : 1.0;
                    cr += sg * T.cr;
                    ci += sg * T.ci;
                    dr += sh * T.cr;
                    di += sh * T.ci;
                }
                const double pr = v[u].re, pi = v[u].im;
                ar += cr * pr - ci * pi;
                ai += cr * pi + ci * pr;
                if (two) {
                    const double qr = w2[u].re, qi = w2[u].im;
                    br += dr * qr - di * qi;
                    bi += dr * qi + di * qr;
                }
            }
        }
        if (l) l[i] = cx<R>{R(ar), R(ai)};
        const cx<R> v = p[i];
        const double vr = v.re, vi = v.im;
        qf::neumaier(er, vr * ar + vi * ai);  // Re(conj(psi) lambda)
        qf::neumaier(ei, vr * ai - vi * ar);  // Im(conj(psi) lambda)
        if (two) {
            if (l) l[j] = cx<R>{R(br), R(bi)};
            const cx<R> q = p[j];
            const double qr = q.re, qi = q.im;
            qf::neumaier(er, qr * br + qi * bi);
            qf::neumaier(ei, qr * bi - qi * br);
        }
    }
    qf::block_reduce(er, ei);
    if (threadIdx.x == 0) {
        partial[2 * (inst * eblk + chunk)] = er;
        partial[2 * (inst * eblk + chunk) + 1] = ei;
    }
}

// lambda = O psi (Pauli sum, terms sorted by x mask, groups after them) and
// the energy partials conj(psi_i) lambda_i, FP64 compensated, one row per
// instance (grid.y).  The gathers psi[i ^ x] of up to 8 groups are issued
// together before their terms are added, so their latencies overlap.
template <class R>
__global__ __launch_bounds__(qf::kThreads) void k_pauli_apply(const cx<R>* __restrict__ psi, cx<R>* __restrict__ lam,
                                                              uint32_t n, const qf::PauliDev* __restrict__ terms,
                                                              uint32_t nterms, uint32_t ngroups,
                                                              qf::dd* __restrict__ partial) {
    const PauliGroup* __restrict__ groups = reinterpret_cast<const PauliGroup*>(terms + nterms);
    const uint64_t dim = uint64_t(1) << n;
    const uint64_t inst = blockIdx.y;
    const cx<R>* __restrict__ p = psi + inst * dim;
    cx<R>* __restrict__ l = lam ? lam + inst * dim : nullptr;
    qf::dd er{0, 0}, ei{0, 0};
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < dim; i += stride) {
        double ar = 0.0, ai = 0.0;
        for (uint32_t g0 = 0; g0 < ngroups; g0 += 8) {
            cx<R> v[8];
#pragma unroll
            for (uint32_t u = 0; u < 8; ++u)
                if (g0 + u < ngroups) v[u] = p[i ^ groups[g0 + u].x];
#pragma unroll
            for (uint32_t u = 0; u < 8; ++u) {
                if (g0 + u >= ngroups) break;
                const PauliGroup G = groups[g0 + u];
                const double pr = v[u].re, pi = v[u].im;
                for (uint32_t t = G.t0; t < G.t1; ++t) {
                    const qf::PauliDev& T = terms[t];
                    // P|col> = c' (-1)^{popcount(col & z)} |col ^ x>, col = i ^ x.
                    const double sg = (__popcll((i ^ G.x) & T.z) & 1) ? -1.0 : 1.0;
                    const double cr = sg * T.cr, ci = sg * T.ci;
                    ar += cr * pr - ci * pi;
                    ai += cr * pi + ci * pr;
                }
            }
        }
        if (l) l[i] = cx<R>{R(ar), R(ai)};
        const cx<R> v = p[i];
        const double vr = v.re, vi = v.im;
        qf::neumaier(er, vr * ar + vi * ai);  // Re(conj(psi) lambda)
        qf::neumaier(ei, vr * ai - vi * ar);  // Im(conj(psi) lambda)
    }
    qf::block_reduce(er, ei);
    if (threadIdx.x == 0) {
        partial[2 * (inst * gridDim.x + blockIdx.x)] = er;
        partial[2 * (inst * gridDim.x + blockIdx.x) + 1] = ei;
    }
}

// grads[inst][p] = sum over the occurrences of p, in reverse circuit order as
// the reference's reverse loop adds them, of the per-block overlaps summed
// in block order (Neumaier).
__global__ __launch_bounds__(qf::kThreads) void k_grad_finish(const double* __restrict__ partial, uint32_t nblk,
                                                              uint32_t nslots, const uint32_t* __restrict__ pstart,
                                                              const uint32_t* __restrict__ pslots, uint32_t nparams,
                                                              uint32_t count, double* __restrict__ grads) {
    const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= uint64_t(count) * nparams) return;
    const uint64_t inst = t / nparams;
    const uint32_t p = uint32_t(t % nparams);
    double g = 0.0;
    for (uint32_t k = pstart[p]; k < pstart[p + 1]; ++k) {
        const uint32_t slot = pslots[k];
        qf::dd acc{0, 0};
        for (uint32_t b = 0; b < nblk; ++b) qf::neumaier(acc, partial[(inst * nblk + b) * nslots + slot]);
        g += acc.s + acc.c;
    }
    grads[t] = g;
}

// Im <lambda| G |psi> for G = X/Y/Z on qubit q (backprop's per-gate overlap).
template <class R>
__global__ __launch_bounds__(qf::kThreads) void k_gen_overlap(const cx<R>* __restrict__ lam,
                                                              const cx<R>* __restrict__ psi, uint64_t dim, int axis,
                                                              uint32_t q, qf::dd* __restrict__ partial) {
    qf::dd acc{0, 0}, unused{0, 0};
    const uint64_t m = uint64_t(1) << q;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < dim; i += stride) {
        const cx<R> l = lam[i];
        const cx<R> p = axis == 2 ? psi[i] : psi[i ^ m];
        const double lr = l.re, li = l.im, pr = p.re, pi = p.im;
        const double re = lr * pr + li * pi, im = lr * pi - li * pr;
        double v;
        if (axis == 0)
            v = im;
        else if (axis == 1)
            v = (i & m) ? re : -re;
        else
            v = (i & m) ? -im : im;
        qf::neumaier(acc, v);
    }
    qf::block_reduce(acc, unused);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = acc;
        partial[2 * blockIdx.x + 1] = qf::dd{0, 0};
    }
}

struct DevBuf {
    void* p = nullptr;
    bool owned = true;
    ~DevBuf() {
        if (p && owned) (void)hipFree(p);
    }
    qf_status alloc(size_t bytes) {
        if (p && owned) (void)hipFree(p);
        p = nullptr;
        owned = true;
        return device_alloc(&p, bytes ? bytes : 16);
    }
};

// Buffers of one call carved from the backend's arena (kept between calls:
// an optimizer loop pays for its device allocations and frees once).
struct ArenaCarve {
    std::vector<std::pair<DevBuf*, size_t>> v;
    void add(DevBuf& b, size_t bytes) { v.push_back({&b, ((bytes ? bytes : 16) + 255) / 256 * 256}); }
    qf_status commit(qf_backend* be) {
        size_t total = 0;
        for (const auto& x : v) total += x.second;
        void* base = nullptr;
        qf_status s = arena(be, total, &base);
        if (s) return s;
        size_t off = 0;
        for (auto& x : v) {
            x.first->p = static_cast<unsigned char*>(base) + off;
            x.first->owned = false;
            off += x.second;
        }
        return QF_OK;
    }
};

template <class R>
void fill_zero_states(hipStream_t st, void* d, uint64_t total, uint32_t qubits) {
    hipLaunchKernelGGL(qf::k_fill_zero_state<R>, dim3(grid_for(total)), dim3(qf::kThreads), 0, st,
                       static_cast<cx<R>*>(d), total, qubits);
}


struct GradRun {
    qf_backend* be;
    PCirc pc;
    uint32_t qubits;
    int precision;
    std::vector<qf::PauliDev> terms;
    qf::Plan fwd, rev;
    uint32_t fwd_tile = 0, rev_tile = 0;
};

// Compiled gradient plans read X/Y rotations as (scale, ratio) rows unless
// QF_SCALED_ROTATIONS=0 (the direct (cos, sin) form).
bool scaled_rotations_on() {
    const char* e = std::getenv("QF_SCALED_ROTATIONS");
    return !(e && e[0] == '0');
}

qf_status plan_for(GradRun& R, const double* inputs, uint32_t tile, bool need_rev, uint64_t sets) {
    qf::PlanOptions po;
    po.fuse = 1;
    // A batch of many small states streams as much memory per launch as one
    // large state: compiled passes (the reverse sweep in their adjoint form)
    // pay for their compile there too.
    if ((sets << R.qubits) >= (uint64_t(1) << 22)) po.compile = 1;
    const bool compiled = po.compile || R.qubits >= qf::kCompileMinQubits;
    // Tiles of compiled gradient plans, measured on the 16-qubit VQE over 4096
    // sets (forward / reverse): complex128 12 / 12 (51.8 ms; 12 / 11: 54.0 ms,
    // 13 / 12: 54.4 ms; 11 / 11 took 57.3 ms with direct-form rotations),
    // complex64 12 / 11 (38.0-38.5 ms; 11 / 11: 39.0-39.3 ms, 13 / 12: 40.3
    // ms, 12 / 12 and 13 / 13 were slower than 11 / 11); complex64 states of
    // fewer than 12 qubits keep 11 / 11.
    const bool c64_split = compiled && !tile && R.precision != QF_C128 && R.qubits >= 12;
    po.tile_qubits = tile ? tile
                     : compiled ? (R.precision == QF_C128 || c64_split ? 12u : 11u)
                                : qf::default_tile_bits(R.precision);
    po.scaled_rotations = compiled && scaled_rotations_on() ? 1 : 0;
    std::string e = qf::build_plan(param_gates(R.pc, inputs, false), R.qubits, R.precision, po, R.fwd);
    if (!e.empty()) return fail(QF_INVALID_ARGUMENT, e);
    if (need_rev) {
        // A reverse plan that runs compiled (its adjoint form) may relabel
        // and keeps the forward tile (its half-buffer exchanges fit both
        // states), except complex64 from 12 qubits (one qubit smaller, as
        // measured above); one that runs through the interpreting adjoint
        // kernel may not relabel and takes a tile one qubit smaller (two
        // states in LDS).
        if (!compiled || tile || c64_split)
            po.tile_qubits = po.tile_qubits > 2 ? po.tile_qubits - 1 : po.tile_qubits;
        po.relabel = compiled ? 1 : 0;
        e = qf::build_plan(param_gates(R.pc, inputs, true), R.qubits, R.precision, po, R.rev);
        if (!e.empty()) return fail(QF_INVALID_ARGUMENT, e);
        // One angle table serves both plans: both read (scale, ratio) rows
        // or neither does.
        if (R.rev.scaled_rotations != R.fwd.scaled_rotations && R.fwd.runtime_rotations) {
            po.scaled_rotations = 0;
            e = qf::build_plan(param_gates(R.pc, inputs, false), R.qubits, R.precision, po, R.fwd);
            if (e.empty()) e = qf::build_plan(param_gates(R.pc, inputs, true), R.qubits, R.precision, po, R.rev);
            if (!e.empty()) return fail(QF_INVALID_ARGUMENT, e);
        }
    }
    return QF_OK;
}

uint32_t reduce_blocks(uint64_t dim) {
    uint64_t b = (dim + qf::kThreads * 4 - 1) / (qf::kThreads * 4);
    if (b > 1024) b = 1024;
    return uint32_t(b ? b : 1);
}

// Largest wave of instances that fits `per_inst` bytes each in the free device memory.
uint32_t wave_size(uint64_t per_inst, uint32_t want, uint32_t cap_chunk, uint32_t log_blocks_max,
                   size_t reusable = 0) {
    size_t freeb = 0, total = 0;
    if (hipMemGetInfo(&freeb, &total) != hipSuccess) freeb = size_t(8) << 30;
    freeb += reusable;  // the backend's arena, reused by this call
    const double budget = double(freeb) * 0.8;
    uint64_t n = per_inst ? uint64_t(budget / double(per_inst)) : want;
    uint32_t ylim = 65535u >> (log_blocks_max > 16 ? std::min<uint32_t>(log_blocks_max - 16, 16) : 0);
    n = std::min<uint64_t>({n, want, ylim ? ylim : 1, 65535});
    if (cap_chunk) n = std::min<uint64_t>(n, cap_chunk);
    return uint32_t(n ? n : 0);
}

// Core: runs `cnt` instances whose angle-table rows are given (table is
// [cnt][nslots][2] host doubles): forward, energy, and with grads != nullptr
// the reverse sweep (grads [cnt][nparams]).  Waves of instances sized to the
// free device memory.
// Fills the angle-table rows [a, b) of a call; table_run calls it for each
// chunk of sets just before the chunk's rows are needed.
using RowFill = std::function<void(uint32_t, uint32_t)>;

qf_status table_run(GradRun& R, const double* table, uint32_t cnt_total, uint32_t chunk_opt, double* energies,
                    double* grads, double* stage = nullptr, const RowFill* fill = nullptr, double* estage = nullptr) {
    qf_backend* be = R.be;
    hipStream_t st = nullptr;
    {
        void* sp = nullptr;
        qf_backend_stream(be, &sp);
        st = static_cast<hipStream_t>(sp);
    }
    const bool want_grad = grads != nullptr;
    const uint32_t n = R.qubits;
    const uint64_t dim = uint64_t(1) << n;
    const size_t ab = amp_bytes(R.precision);
    const uint32_t nslots = uint32_t(R.pc.slot_param.size());
    const uint32_t rev_M = want_grad && !R.rev.passes.empty() ? R.rev.passes[0].tile_bits : n;
    const uint32_t nblk = uint32_t(1) << (n - std::min(rev_M, n));
    const uint32_t eblk = reduce_blocks(dim);
    const uint64_t per_inst = dim * ab * (want_grad ? 2 : 1) + uint64_t(nslots) * 16 +
                              (want_grad ? uint64_t(nblk) * nslots * 8 + uint64_t(R.pc.nparams) * 8 : 0) +
                              eblk * 32 + 64;
    const uint32_t fwd_M = R.fwd.passes.empty() ? n : R.fwd.passes[0].tile_bits;
    const uint32_t logmax = n - std::min(std::min(fwd_M, rev_M), n);
    uint32_t wave = wave_size(per_inst, cnt_total, chunk_opt, logmax, be->arena_bytes);
    if (wave == 0) return fail(QF_OUT_OF_MEMORY, "device memory cannot hold one parameter set's states");

    // parameter -> occurrences in reverse circuit order (CSR)
    std::vector<uint32_t> ps(R.pc.nparams + 1, 0), pl(nslots);
    for (uint32_t k = 0; k < nslots; ++k) ps[R.pc.slot_param[k] + 1]++;
    for (uint32_t p = 0; p < R.pc.nparams; ++p) ps[p + 1] += ps[p];
    {
        std::vector<uint32_t> fill(ps.begin(), ps.end() - 1);
        for (uint32_t k = nslots; k-- > 0;) pl[fill[R.pc.slot_param[k]]++] = k;
    }
    const std::vector<qf::PauliDev> tsorted = sorted_terms(R.terms);
    uint32_t ngroups = 0;
    const std::vector<unsigned char> tpack = pack_terms(tsorted, &ngroups);
    const std::vector<unsigned char> wpack = pack_pauli_windows(tsorted);
    // Arena bytes of a wave of w sets (the carve below, 256-byte aligned).
    auto carve_bytes = [&](uint64_t w) {
        auto a = [](uint64_t b) { return ((b ? b : 16) + 255) / 256 * 256; };
        return a(w * dim * ab) + a(want_grad ? w * dim * ab : 16) + a(w * nslots * 16) +
               a(want_grad ? w * nblk * nslots * 8 : 16) + a(w * eblk * 2 * sizeof(qf::dd)) + a(w * 2 * 8) +
               a(want_grad ? w * R.pc.nparams * 8 : 16) + a(ps.size() * 4) + a(std::max<size_t>(pl.size(), 1) * 4) +
               a(tpack.size()) + a(wpack.size());
    };
    // An arena kept from an earlier call is used as it is when the call
    // needs no more waves within it: growing it means freeing most of HBM
    // and allocating it again, which has taken seconds.
    if (be->arena_bytes > 0 && carve_bytes(wave) > be->arena_bytes) {
        uint64_t lo = 0, hi = wave;  // largest w with carve_bytes(w) <= arena_bytes
        while (lo < hi) {
            const uint64_t mid = (lo + hi + 1) / 2;
            if (carve_bytes(mid) <= be->arena_bytes) lo = mid;
            else hi = mid - 1;
        }
        if (lo > 0 && (cnt_total + lo - 1) / lo == (cnt_total + wave - 1) / wave) wave = uint32_t(lo);
    }
    DevBuf psi, lam, cs, part, epart, eout, gout, pstart, pslots, terms, wterms;
    qf_status s;
    {
        ArenaCarve ac;
        ac.add(psi, size_t(wave) * dim * ab);
        ac.add(lam, want_grad ? size_t(wave) * dim * ab : 16);
        ac.add(cs, size_t(wave) * nslots * 16);
        ac.add(part, want_grad ? size_t(wave) * nblk * nslots * 8 : 16);
        ac.add(epart, size_t(wave) * eblk * 2 * sizeof(qf::dd));
        ac.add(eout, size_t(wave) * 2 * sizeof(double));
        ac.add(gout, want_grad ? size_t(wave) * R.pc.nparams * sizeof(double) : 16);
        ac.add(pstart, ps.size() * 4);
        ac.add(pslots, std::max<size_t>(pl.size(), 1) * 4);
        ac.add(terms, tpack.size());
        ac.add(wterms, wpack.size());
        if ((s = ac.commit(be))) return s;
    }
    uint32_t wcoef = 0;
    if (!wpack.empty()) {
        wcoef = reinterpret_cast<const PWHeader*>(wpack.data())->ncoef;
        QF_HIP_TRY(hipMemcpyAsync(wterms.p, wpack.data(), wpack.size(), hipMemcpyHostToDevice, st));
    }
    QF_HIP_TRY(hipMemcpyAsync(pstart.p, ps.data(), ps.size() * 4, hipMemcpyHostToDevice, st));
    if (!pl.empty()) QF_HIP_TRY(hipMemcpyAsync(pslots.p, pl.data(), pl.size() * 4, hipMemcpyHostToDevice, st));
    QF_HIP_TRY(hipMemcpyAsync(terms.p, tpack.data(), tpack.size(), hipMemcpyHostToDevice, st));
    QF_HIP_TRY(hipStreamSynchronize(st));  // host vectors above go out of scope with this function only

    std::vector<const qf::TileOp*> bases;
    std::vector<const qf::Plan*> plans{&R.fwd};
    if (want_grad) plans.push_back(&R.rev);
    if ((s = upload_plans(be, plans, bases))) return s;

    qf_gradient_profile& P = be->grad_prof;
    P = qf_gradient_profile{};
    P.sets = cnt_total;
    P.forward_passes = uint32_t(R.fwd.passes.size());
    P.reverse_passes = want_grad ? uint32_t(R.rev.passes.size()) : 0;
    P.costate_kernel = wpack.empty() ? 2 : 1;
    P.profiled = be->profile ? 1 : 0;
    const double sweep = 2.0 * double(dim) * double(ab);
    P.forward_bytes = sweep * P.forward_passes * cnt_total;
    P.costate_bytes = sweep * cnt_total;
    P.reverse_bytes = 2.0 * sweep * P.reverse_passes * cnt_total;
    struct Ev {
        hipEvent_t e[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
        ~Ev() {
            for (auto& x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } ev;
    if (be->profile)
        for (auto& x : ev.e) QF_HIP_TRY(hipEventCreate(&x));
    auto mark = [&](int k) { return be->profile ? hipEventRecord(ev.e[k], st) : hipSuccess; };

    // The energies come back through pinned staging when the caller has
    // some (a copy to pageable memory waits on a bounce buffer).
    std::vector<double> eo_pageable(estage ? 0 : size_t(wave) * 2);
    double* eo = estage ? estage : eo_pageable.data();
    for (uint32_t off = 0; off < cnt_total; off += wave) {
        const uint32_t cnt = std::min(wave, cnt_total - off);
        ++P.waves;
        // Synchronous copy: the previous wave finished at its download below.
        if (nslots && !fill)
            QF_HIP_TRY(hipMemcpy(cs.p, table + size_t(off) * nslots * 2, size_t(cnt) * nslots * 16,
                                 hipMemcpyHostToDevice));
        if (R.precision == QF_C128)
            fill_zero_states<double>(st, psi.p, dim * cnt, n);
        else
            fill_zero_states<float>(st, psi.p, dim * cnt, n);
        if ((s = launch_check("k_fill_zero_state"))) return s;
        LaunchCtx fctx;
        fctx.cs = static_cast<const double*>(cs.p);
        fctx.nslots = nslots;
        QF_HIP_TRY(mark(0));
        if (fill && nslots) {
            // The host fills the angle rows of the next chunk of sets while
            // the GPU runs the first forward pass over this one (and, before
            // the first chunk, the zero fill); the other passes then run
            // over the whole wave.  Chunks of 1/16, 2/16, 4/16 and 9/16 of
            // the sets: only the first, small one waits for the host, and
            // each later fill is shorter than the GPU's run of the chunk
            // before it.
            const uint32_t K = cnt >= 1024 && R.fwd.passes.size() > 1 ? 4 : 1;
            static const uint32_t edge16[5] = {0, 1, 3, 7, 16};
            auto edge = [&](uint32_t j) {
                return K == 1 ? (j ? cnt : 0u) : uint32_t(uint64_t(cnt) * edge16[j] / 16);
            };
            for (uint32_t j = 0; j < K; ++j) {
                const uint32_t a = edge(j), b = edge(j + 1);
                if (j == 0) (*fill)(off + a, off + b);
                double* csj = static_cast<double*>(cs.p) + size_t(a) * nslots * 2;
                QF_HIP_TRY(hipMemcpyAsync(csj, table + size_t(off + a) * nslots * 2, size_t(b - a) * nslots * 16,
                                          hipMemcpyHostToDevice, st));
                LaunchCtx c1 = fctx;
                c1.cs = csj;
                if (K > 1) c1.pass_end = 1;
                if ((s = launch_plan(be, static_cast<char*>(psi.p) + size_t(a) * dim * ab, R.precision, n, b - a,
                                     R.fwd, bases[0], c1)))
                    return s;
                P.forward_kernels |= R.fwd.kernels_run;
                if (j + 1 < K) (*fill)(off + b, off + edge(j + 2));
            }
            if (K > 1) {
                LaunchCtx c2 = fctx;
                c2.pass_begin = 1;
                if ((s = launch_plan(be, psi.p, R.precision, n, cnt, R.fwd, bases[0], c2))) return s;
                P.forward_kernels |= R.fwd.kernels_run;
            }
        } else {
            if (fill) (*fill)(off, off + cnt);
            if ((s = launch_plan(be, psi.p, R.precision, n, cnt, R.fwd, bases[0], fctx))) return s;
            P.forward_kernels |= R.fwd.kernels_run;
        }
        QF_HIP_TRY(mark(1));
        void* lam_out = want_grad ? lam.p : nullptr;
        if (!wpack.empty()) {
            const uint64_t blocks = (uint64_t(cnt) * eblk + 7) / 8 * 8;
            const size_t lds = size_t(std::max<uint32_t>(wcoef, 1)) * 16;
            if (R.precision == QF_C128)
                hipLaunchKernelGGL(k_pauli_windows<double>, dim3(unsigned(blocks)), dim3(qf::kThreads), lds, st,
                                   static_cast<const cx<double>*>(psi.p), static_cast<cx<double>*>(lam_out), n,
                                   static_cast<const unsigned char*>(wterms.p), cnt, eblk,
                                   static_cast<qf::dd*>(epart.p));
            else
                hipLaunchKernelGGL(k_pauli_windows<float>, dim3(unsigned(blocks)), dim3(qf::kThreads), lds, st,
                                   static_cast<const cx<float>*>(psi.p), static_cast<cx<float>*>(lam_out), n,
                                   static_cast<const unsigned char*>(wterms.p), cnt, eblk,
                                   static_cast<qf::dd*>(epart.p));
            if ((s = launch_check("k_pauli_windows"))) return s;
        } else {
            dim3 eg(eblk, cnt);
            if (R.precision == QF_C128)
                hipLaunchKernelGGL(k_pauli_apply<double>, eg, dim3(qf::kThreads), 0, st,
                                   static_cast<const cx<double>*>(psi.p), static_cast<cx<double>*>(lam_out), n,
                                   static_cast<const qf::PauliDev*>(terms.p), uint32_t(tsorted.size()), ngroups,
                                   static_cast<qf::dd*>(epart.p));
            else
                hipLaunchKernelGGL(k_pauli_apply<float>, eg, dim3(qf::kThreads), 0, st,
                                   static_cast<const cx<float>*>(psi.p), static_cast<cx<float>*>(lam_out), n,
                                   static_cast<const qf::PauliDev*>(terms.p), uint32_t(tsorted.size()), ngroups,
                                   static_cast<qf::dd*>(epart.p));
            if ((s = launch_check("k_pauli_apply"))) return s;
        }
        QF_HIP_TRY(mark(2));
        hipLaunchKernelGGL(qf::k_reduce_finish<>, dim3(cnt), dim3(qf::kThreads), 0, st,
                           static_cast<const qf::dd*>(epart.p), eblk, static_cast<double*>(eout.p));
        if ((s = launch_check("k_reduce_finish"))) return s;
        QF_HIP_TRY(mark(3));
        if (want_grad) {
            LaunchCtx rctx;
            rctx.cs = static_cast<const double*>(cs.p);
            rctx.nslots = nslots;
            rctx.lam = lam.p;
            rctx.partial = static_cast<double*>(part.p);
            if ((s = launch_plan(be, psi.p, R.precision, n, cnt, R.rev, bases[1], rctx))) return s;
            P.reverse_kernels |= R.rev.kernels_run;
            const uint64_t work = uint64_t(cnt) * R.pc.nparams;
            if (work) {
                hipLaunchKernelGGL(k_grad_finish, dim3(unsigned((work + 255) / 256)), dim3(256), 0, st,
                                   static_cast<const double*>(part.p), nblk, nslots,
                                   static_cast<const uint32_t*>(pstart.p), static_cast<const uint32_t*>(pslots.p),
                                   R.pc.nparams, cnt, static_cast<double*>(gout.p));
                if ((s = launch_check("k_grad_finish"))) return s;
                // Into pinned staging when the caller has one (one DMA), else
                // straight to the caller's memory.
                QF_HIP_TRY(hipMemcpyAsync((stage ? stage : grads) + size_t(off) * R.pc.nparams, gout.p,
                                          work * sizeof(double), hipMemcpyDeviceToHost, st));
            }
        }
        QF_HIP_TRY(mark(4));
        QF_HIP_TRY(hipMemcpyAsync(eo, eout.p, size_t(cnt) * 2 * sizeof(double), hipMemcpyDeviceToHost, st));
        QF_HIP_TRY(hipStreamSynchronize(st));
        for (uint32_t b = 0; b < cnt; ++b) energies[off + b] = eo[2 * b];
        if (stage && want_grad && R.pc.nparams) {
            // Out of the pinned staging in 1 MiB pieces on host threads.
            const size_t bytes = size_t(cnt) * R.pc.nparams * sizeof(double);
            char* dst = reinterpret_cast<char*>(grads + size_t(off) * R.pc.nparams);
            const char* src = reinterpret_cast<const char*>(stage + size_t(off) * R.pc.nparams);
            constexpr size_t piece = size_t(1) << 20;
            parallel_rows(
                (bytes + piece - 1) / piece,
                [&](uint64_t k) {
                    const size_t a = size_t(k) * piece;
                    std::memcpy(dst + a, src + a, std::min(piece, bytes - a));
                },
                8192);
        }
        if (be->profile) {
            // ev 0..1 forward, 1..2 costate, 2..3 energy finish, 3..4 reverse + gradient finish
            float t[4] = {0, 0, 0, 0};
            for (int k = 0; k < 4; ++k) QF_HIP_TRY(hipEventElapsedTime(&t[k], ev.e[k], ev.e[k + 1]));
            P.forward_ms += t[0];
            P.costate_ms += t[1];
            P.finish_ms += t[2];
            P.reverse_ms += t[3];
        }
    }
    return QF_OK;
}

// Energies (and with grads, adjoint gradients) at `sets` parameter rows.
qf_status adjoint_sets(GradRun& R, const double* params, uint32_t sets, uint32_t chunk_opt, double* energies,
                       double* grads) {
    const uint32_t nslots = uint32_t(R.pc.slot_param.size());
    // The angle table and the gradient staging in the backend's pinned buffer.
    const size_t tdoubles = std::max<size_t>(size_t(sets) * nslots * 2, 1);
    const size_t gdoubles = grads ? size_t(sets) * R.pc.nparams : 0;
    // Energies of one memory wave (at most all sets).
    const size_t edoubles = size_t(sets) * 2;
    void* hp = nullptr;
    std::vector<double> fallback;
    double* table = nullptr;
    double* stage = nullptr;
    double* estage = nullptr;
    if (pinned_host(R.be, (tdoubles + gdoubles + edoubles) * sizeof(double), &hp) == QF_OK) {
        table = static_cast<double*>(hp);
        stage = gdoubles ? table + tdoubles : nullptr;
        estage = table + tdoubles + gdoubles;
    } else {
        fallback.assign(tdoubles, 0.0);
        table = fallback.data();
    }
    const RowFill rows = [&](uint32_t a, uint32_t b) {
        parallel_rows(
            b - a,
            [&](uint64_t i) {
                fill_row(R.pc, params + size_t(a + i) * R.pc.nparams, -1, 0.0, table + size_t(a + i) * nslots * 2,
                         R.fwd.scaled_rotations != 0);
            },
            nslots);
    };
    return table_run(R, table, sets, chunk_opt, energies, grads, stage, &rows, estage);
}

// Shift rule (fd = false, delta = pi/2) or central differences (fd = true,
// delta = step): 2 x occurrences shifted circuits per set, run as instances
// whose angle rows differ in the shifted occurrence only.
qf_status shift_sets(GradRun& R, const double* params, uint32_t sets, uint32_t chunk_opt, bool fd, double step,
                     double* energies, double* grads) {
    const uint32_t nslots = uint32_t(R.pc.slot_param.size());
    qf_status s = adjoint_sets(R, params, sets, chunk_opt, energies, nullptr);
    if (s) return s;
    std::memset(grads, 0, size_t(sets) * R.pc.nparams * sizeof(double));
    if (nslots == 0) return QF_OK;
    const double delta = fd ? step : 1.57079632679489661923;  // FRAC_PI_2
    // Groups of sets keep the host tables bounded (about 64 MiB).
    const uint64_t row_bytes = uint64_t(nslots) * 16;
    const uint64_t per_set = row_bytes * nslots * 2;
    const uint32_t per_group = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(sets, (uint64_t(64) << 20) / per_set)));
    std::vector<double> table;
    std::vector<double> eg;
    for (uint32_t g0 = 0; g0 < sets; g0 += per_group) {
        const uint32_t gs = std::min(per_group, sets - g0);
        const uint32_t cnt = gs * nslots * 2;
        table.assign(size_t(cnt) * nslots * 2, 0.0);
        parallel_rows(
            uint64_t(gs) * nslots * 2,
            [&](uint64_t row) {
                const uint64_t b = row / (uint64_t(nslots) * 2), k = (row / 2) % nslots;
                fill_row(R.pc, params + size_t(g0 + b) * R.pc.nparams, int64_t(k), row % 2 == 0 ? delta : -delta,
                         table.data() + row * nslots * 2, R.fwd.scaled_rotations != 0);
            },
            nslots);
        eg.assign(cnt, 0.0);
        if ((s = table_run(R, table.data(), cnt, chunk_opt, eg.data(), nullptr))) return s;
        for (uint32_t b = 0; b < gs; ++b) {
            double* gr = grads + size_t(g0 + b) * R.pc.nparams;
            for (uint32_t k = 0; k < nslots; ++k) {  // occurrence order, as src/lib.rs:3534-3543
                const double plus = eg[(size_t(b) * nslots + k) * 2], minus = eg[(size_t(b) * nslots + k) * 2 + 1];
                if (fd)
                    gr[R.pc.slot_param[k]] += (plus - minus) / (2.0 * step);
                else
                    gr[R.pc.slot_param[k]] += 0.5 * (plus - minus);
            }
        }
    }
    return QF_OK;
}

}  // namespace

namespace {

// The reference's backprop (src/gradients.rs:392-506): forward storing every
// intermediate state, costate O psi_final, reverse accumulation of
// Im <lambda|G|psi_after> with exact per-gate sweeps.  One set at a time.
qf_status backprop_set(GradRun& R, const double* inputs, const double* params, double* energy, double* grad) {
    qf_backend* be = R.be;
    void* sp = nullptr;
    qf_backend_stream(be, &sp);
    hipStream_t st = static_cast<hipStream_t>(sp);
    const uint32_t n = R.qubits;
    const uint64_t dim = uint64_t(1) << n;
    const size_t ab = amp_bytes(R.precision);
    const size_t nstates = R.pc.nops + 1;
    size_t freeb = 0, total = 0;
    QF_HIP_TRY(hipMemGetInfo(&freeb, &total));
    const double need = double(nstates + 1) * double(dim) * double(ab);
    if (need > 0.5 * double(freeb))
        return fail(QF_MEMORY_BUDGET_EXCEEDED, "backprop stores every intermediate state: needs " +
                                                   std::to_string(uint64_t(need)) + " bytes, " +
                                                   std::to_string(uint64_t(freeb / 2)) + " allowed");
    DevBuf states, lam, part, eout, terms;
    qf_status s;
    if ((s = states.alloc(nstates * dim * ab))) return s;
    if ((s = lam.alloc(dim * ab))) return s;
    const uint32_t eblk = reduce_blocks(dim);
    if ((s = part.alloc(size_t(eblk) * 2 * sizeof(qf::dd)))) return s;
    if ((s = eout.alloc(2 * sizeof(double)))) return s;
    const std::vector<qf::PauliDev> tsorted = sorted_terms(R.terms);
    uint32_t ngroups = 0;
    const std::vector<unsigned char> tpack = pack_terms(tsorted, &ngroups);
    if ((s = terms.alloc(tpack.size()))) return s;
    QF_HIP_TRY(hipMemcpy(terms.p, tpack.data(), tpack.size(), hipMemcpyHostToDevice));
    char* base = static_cast<char*>(states.p);
    auto sp_at = [&](size_t k) { return static_cast<void*>(base + k * dim * ab); };
    if (R.precision == QF_C128)
        fill_zero_states<double>(st, sp_at(0), dim, n);
    else
        fill_zero_states<float>(st, sp_at(0), dim, n);
    if ((s = launch_check("k_fill_zero_state"))) return s;
    qf::PlanOptions exact;
    exact.fuse = 0;
    auto gate_of = [&](size_t i, bool inverse) {
        const qf_pop& op = R.pc.ops[i];
        qf::GateIn g{};
        g.a = op.a;
        g.b = op.b;
        if (op.kind == 0) {
            int gid = op.gate;
            double th = op.theta;
            if (inverse) adjoint_gate_id(op.gate, op.theta, &gid, &th);
            g.kind = 0;
            gate_matrix(gid, th, g.m);
        } else if (op.kind == 1 || op.kind == 2) {
            const double v = op.kind == 1 ? params[op.index] : inputs[op.index];
            g.kind = 0;
            gate_matrix(QF_GATE_RX + op.gate, inverse ? -v : v, g.m);
        } else {
            g.kind = op.kind - 2;  // 3 CNOT -> 1, 4 CZ -> 2, 5 SWAP -> 3
        }
        return g;
    };
    for (size_t i = 0; i < R.pc.nops; ++i) {
        QF_HIP_TRY(hipMemcpyAsync(sp_at(i + 1), sp_at(i), dim * ab, hipMemcpyDeviceToDevice, st));
        if ((s = apply_gates(be, sp_at(i + 1), R.precision, n, 1, {gate_of(i, false)}, exact))) return s;
    }
    if (R.precision == QF_C128)
        hipLaunchKernelGGL(k_pauli_apply<double>, dim3(eblk, 1), dim3(qf::kThreads), 0, st,
                           static_cast<const cx<double>*>(sp_at(R.pc.nops)), static_cast<cx<double>*>(lam.p), n,
                           static_cast<const qf::PauliDev*>(terms.p), uint32_t(tsorted.size()), ngroups,
                           static_cast<qf::dd*>(part.p));
    else
        hipLaunchKernelGGL(k_pauli_apply<float>, dim3(eblk, 1), dim3(qf::kThreads), 0, st,
                           static_cast<const cx<float>*>(sp_at(R.pc.nops)), static_cast<cx<float>*>(lam.p), n,
                           static_cast<const qf::PauliDev*>(terms.p), uint32_t(tsorted.size()), ngroups,
                           static_cast<qf::dd*>(part.p));
    if ((s = launch_check("k_pauli_apply"))) return s;
    double e2[2];
    hipLaunchKernelGGL(qf::k_reduce_finish<>, dim3(1), dim3(qf::kThreads), 0, st, static_cast<const qf::dd*>(part.p),
                       eblk, static_cast<double*>(eout.p));
    QF_HIP_TRY(hipMemcpyAsync(e2, eout.p, sizeof(e2), hipMemcpyDeviceToHost, st));
    QF_HIP_TRY(hipStreamSynchronize(st));
    *energy = e2[0];
    for (uint32_t p = 0; p < R.pc.nparams; ++p) grad[p] = 0.0;
    for (size_t i = R.pc.nops; i-- > 0;) {
        const qf_pop& op = R.pc.ops[i];
        if (op.kind == 1) {
            if (R.precision == QF_C128)
                hipLaunchKernelGGL(k_gen_overlap<double>, dim3(eblk), dim3(qf::kThreads), 0, st,
                                   static_cast<const cx<double>*>(lam.p), static_cast<const cx<double>*>(sp_at(i + 1)),
                                   dim, op.gate, op.a, static_cast<qf::dd*>(part.p));
            else
                hipLaunchKernelGGL(k_gen_overlap<float>, dim3(eblk), dim3(qf::kThreads), 0, st,
                                   static_cast<const cx<float>*>(lam.p), static_cast<const cx<float>*>(sp_at(i + 1)),
                                   dim, op.gate, op.a, static_cast<qf::dd*>(part.p));
            if ((s = launch_check("k_gen_overlap"))) return s;
            hipLaunchKernelGGL(qf::k_reduce_finish<>, dim3(1), dim3(qf::kThreads), 0, st,
                               static_cast<const qf::dd*>(part.p), eblk, static_cast<double*>(eout.p));
            QF_HIP_TRY(hipMemcpyAsync(e2, eout.p, sizeof(e2), hipMemcpyDeviceToHost, st));
            QF_HIP_TRY(hipStreamSynchronize(st));
            grad[op.index] += e2[0];
        }
        if ((s = apply_gates(be, lam.p, R.precision, n, 1, {gate_of(i, true)}, exact))) return s;
    }
    QF_HIP_TRY(hipStreamSynchronize(st));
    return QF_OK;
}

qf_gradient_options default_options() {
    qf_gradient_options o{};
    o.strategy = QF_GRAD_AUTO;
    o.precision = QF_C128;
    o.fd_step = 1e-3;  // GradientOptions::default, src/gradients.rs:161
    return o;
}

// Shared validation of the gradient entry points, in the reference's order.
qf_status prepare(qf_backend* be, const qf_pop* ops, size_t nops, uint32_t nparams, uint32_t ninputs,
                  const double* inputs, size_t inputs_len, const double* params, uint32_t sets, uint32_t qubits,
                  const qf_pauli_term* obs, size_t nterms, const qf_gradient_options& o, int32_t strategy,
                  GradRun& R) {
    if (!be || (!ops && nops) || (!params && nparams && sets) || (!inputs && inputs_len) || (!obs && nterms))
        return fail(QF_INVALID_ARGUMENT, "null argument");
    qf_status s = check_qubits(qubits);
    if (s) return s;
    if (o.precision != QF_C128 && o.precision != QF_C64) return fail(QF_INVALID_ARGUMENT, "unknown precision");
    R.be = be;
    R.pc.ops = ops;
    R.pc.nops = nops;
    R.pc.nparams = nparams;
    R.pc.ninputs = ninputs;
    R.qubits = qubits;
    R.precision = o.precision;
    if (strategy == QF_GRAD_FINITE_DIFFERENCE && (!std::isfinite(o.fd_step) || o.fd_step <= 0.0))
        return fail(QF_INVALID_FD_STEP, "finite-difference step must be positive and finite, got " +
                                            std::to_string(o.fd_step));
    for (uint32_t b = 0; b < sets; ++b)
        if ((s = validate_values(R.pc, inputs, inputs_len, params + size_t(b) * nparams, nparams))) return s;
    if ((s = pauli_terms(obs, nterms, qubits, R.terms))) return s;
    if ((s = validate_ops(R.pc, qubits))) return s;
    return activate(be);
}

int32_t resolve(int32_t strategy, const qf_pop* ops, size_t nops, uint64_t shots, bool* channel) {
    *channel = false;
    for (size_t i = 0; i < nops; ++i)
        if (ops[i].kind == 6) *channel = true;
    if (strategy != QF_GRAD_AUTO) return strategy;
    // GradientStrategy::select, src/gradients.rs:81-89.
    if (*channel) return -1;
    return shots > 0 ? QF_GRAD_PARAMETER_SHIFT : QF_GRAD_ADJOINT;
}

qf_status run_strategy(GradRun& R, int32_t strategy, const double* inputs, const double* params, uint32_t sets,
                       const qf_gradient_options& o, double* energies, double* grads) {
    qf_status s;
    switch (strategy) {
        case QF_GRAD_ADJOINT:
            if ((s = plan_for(R, inputs, o.tile_qubits, true, sets))) return s;
            return adjoint_sets(R, params, sets, o.chunk, energies, grads);
        case QF_GRAD_PARAMETER_SHIFT:
        case QF_GRAD_FINITE_DIFFERENCE:
            // The shifted circuits run as one batch of 2 x occurrences per set.
            if ((s = plan_for(R, inputs, o.tile_qubits, false,
                              uint64_t(sets) * 2 * std::max<size_t>(R.pc.slot_param.size(), 1))))
                return s;
            return shift_sets(R, params, sets, o.chunk, strategy == QF_GRAD_FINITE_DIFFERENCE, o.fd_step, energies,
                              grads);
        case QF_GRAD_BACKPROP:
            for (uint32_t b = 0; b < sets; ++b)
                if ((s = backprop_set(R, inputs, params + size_t(b) * R.pc.nparams, energies + b,
                                      grads + size_t(b) * R.pc.nparams)))
                    return s;
            return QF_OK;
        default: return fail(QF_INVALID_ARGUMENT, "unknown gradient strategy " + std::to_string(strategy));
    }
}

}  // namespace

extern "C" qf_status qf_batch_gradient(qf_backend* be, const qf_pop* ops, size_t nops, uint32_t nparams,
                                       uint32_t ninputs, const double* inputs, uint32_t qubits, const double* params,
                                       uint32_t sets, const qf_pauli_term* obs, size_t nterms,
                                       const qf_gradient_options* options, double* energies, double* grads) {
    if (sets == 0) return fail(QF_EMPTY_BATCH, "a gradient batch cannot be empty");
    if (!energies || (!grads && nparams)) return fail(QF_INVALID_ARGUMENT, "null output");
    const qf_gradient_options o = options ? *options : default_options();
    bool channel = false;
    const int32_t strategy = resolve(o.strategy, ops, nops, o.shots, &channel);
    if (strategy < 0) return fail(QF_REQUIRES_DENSITY_MATRIX, "this circuit operation requires density-matrix simulation");
    GradRun R;
    qf_status s = prepare(be, ops, nops, nparams, ninputs, inputs, ninputs, params, sets, qubits, obs, nterms, o,
                          strategy, R);
    if (s) return s;
    std::vector<double> gtmp;
    if (!grads) {
        gtmp.resize(1);
        grads = gtmp.data();
    }
    return run_strategy(R, strategy, inputs, params, sets, o, energies, grads);
}

extern "C" qf_status qf_gradient(qf_backend* be, const qf_pop* ops, size_t nops, uint32_t nparams, uint32_t ninputs,
                                 const double* inputs, size_t inputs_len, const double* params, size_t params_len,
                                 uint32_t qubits, const qf_pauli_term* obs, size_t nterms,
                                 const qf_gradient_options* options, double* grad_out, double* value_out,
                                 int32_t* strategy_used, uint64_t* evaluations) {
    const qf_gradient_options o = options ? *options : default_options();
    bool channel = false;
    const int32_t strategy = resolve(o.strategy, ops, nops, o.shots, &channel);
    if (strategy < 0) return fail(QF_REQUIRES_DENSITY_MATRIX, "this circuit operation requires density-matrix simulation");
    if (!grad_out && nparams) return fail(QF_INVALID_ARGUMENT, "null output");
    GradRun R;
    // validate_parameters checks the supplied length before anything else on it.
    if (strategy == QF_GRAD_FINITE_DIFFERENCE && (!std::isfinite(o.fd_step) || o.fd_step <= 0.0))
        return fail(QF_INVALID_FD_STEP, "finite-difference step must be positive and finite, got " +
                                            std::to_string(o.fd_step));
    if (inputs_len != ninputs)
        return fail(QF_INVALID_INPUT_COUNT, "expected " + std::to_string(ninputs) + " encoded inputs, got " +
                                                std::to_string(inputs_len));
    for (size_t i = 0; i < inputs_len; ++i)
        if (!std::isfinite(inputs[i])) return fail(QF_NON_FINITE_PARAMETER, "circuit parameters must be finite");
    if (params_len != nparams)
        return fail(QF_INVALID_PARAMETER_COUNT, "expected " + std::to_string(nparams) + " circuit parameters, got " +
                                                    std::to_string(params_len));
    qf_status s = prepare(be, ops, nops, nparams, ninputs, inputs, inputs_len, params, 1, qubits, obs, nterms, o,
                          strategy, R);
    if (s) return s;
    double energy = 0.0;
    std::vector<double> g(std::max<size_t>(nparams, 1));
    if ((s = run_strategy(R, strategy, inputs, params, 1, o, &energy, g.data()))) return s;
    if (nparams) std::memcpy(grad_out, g.data(), nparams * sizeof(double));
    if (value_out) *value_out = energy;
    if (strategy_used) *strategy_used = strategy;
    if (evaluations) {  // GradientEstimate::evaluations, src/gradients.rs:256-306
        const uint64_t occ = R.pc.slot_param.size();
        *evaluations = (strategy == QF_GRAD_ADJOINT || strategy == QF_GRAD_BACKPROP) ? 2 : 2 * occ;
    }
    return QF_OK;
}

extern "C" qf_status qf_batch_expectation(qf_backend* be, const qf_pop* ops, size_t nops, uint32_t nparams,
                                          uint32_t ninputs, const double* inputs, uint32_t qubits,
                                          const double* params, uint32_t sets, const qf_pauli_term* obs,
                                          size_t nterms, const qf_gradient_options* options, double* energies) {
    if (sets == 0) return fail(QF_EMPTY_BATCH, "a batch cannot be empty");
    if (!energies) return fail(QF_INVALID_ARGUMENT, "null output");
    const qf_gradient_options o = options ? *options : default_options();
    GradRun R;
    qf_status s = prepare(be, ops, nops, nparams, ninputs, inputs, ninputs, params, sets, qubits, obs, nterms, o,
                          QF_GRAD_ADJOINT, R);
    if (s) return s;
    for (size_t i = 0; i < nops; ++i)
        if (ops[i].kind == 6)
            return fail(QF_REQUIRES_DENSITY_MATRIX, "this circuit operation requires density-matrix simulation");
    if ((s = plan_for(R, inputs, o.tile_qubits, false, sets))) return s;
    return adjoint_sets(R, params, sets, o.chunk, energies, nullptr);
}

extern "C" qf_status qf_state_run_parameterized(qf_state* st, const qf_pop* ops, size_t nops, uint32_t nparams,
                                                uint32_t ninputs, const double* inputs, size_t inputs_len,
                                                const double* params, size_t params_len) {
    if (!st) return fail(QF_INVALID_ARGUMENT, "null state");
    uint32_t qubits = 0;
    int32_t precision = 0;
    qf_state_qubits(st, &qubits);
    qf_state_precision(st, &precision);
    void* d = nullptr;
    qf_state_device_pointer(st, &d);
    qf_backend* be = state_backend(st);
    GradRun R;
    R.be = be;
    R.pc.ops = ops;
    R.pc.nops = nops;
    R.pc.nparams = nparams;
    R.pc.ninputs = ninputs;
    R.qubits = qubits;
    R.precision = precision;
    if ((!ops && nops) || (!params && params_len) || (!inputs && inputs_len))
        return fail(QF_INVALID_ARGUMENT, "null argument");
    qf_status s = validate_values(R.pc, inputs, inputs_len, params, params_len);
    if (s) return s;
    if ((s = validate_ops(R.pc, qubits))) return s;
    if ((s = activate(be))) return s;
    if ((s = plan_for(R, inputs, 0, false, 1))) return s;
    const uint32_t nslots = uint32_t(R.pc.slot_param.size());
    std::vector<double> row(std::max<size_t>(size_t(nslots) * 2, 2));
    fill_row(R.pc, params, -1, 0.0, row.data(), R.fwd.scaled_rotations != 0);
    void* sp = nullptr;
    qf_backend_stream(be, &sp);
    hipStream_t strm = static_cast<hipStream_t>(sp);
    DevBuf cs;
    if ((s = cs.alloc(row.size() * sizeof(double)))) return s;
    QF_HIP_TRY(hipMemcpy(cs.p, row.data(), row.size() * sizeof(double), hipMemcpyHostToDevice));
    const uint64_t dim = uint64_t(1) << qubits;
    if (precision == QF_C128)
        fill_zero_states<double>(strm, d, dim, qubits);
    else
        fill_zero_states<float>(strm, d, dim, qubits);
    if ((s = launch_check("k_fill_zero_state"))) return s;
    std::vector<const qf::TileOp*> bases;
    if ((s = upload_plans(be, {&R.fwd}, bases))) return s;
    LaunchCtx ctx;
    ctx.cs = static_cast<const double*>(cs.p);
    ctx.nslots = nslots;
    if ((s = launch_plan(be, d, precision, qubits, 1, R.fwd, bases[0], ctx))) return s;
    QF_HIP_TRY(hipStreamSynchronize(strm));  // cs is freed on return
    return QF_OK;
}

// Plan of a parameterized circuit as the gradient engine builds it: rotation
// occurrences read their angles from the per-instance table (forward), or run
// inverted and recorded in reverse order (reverse != 0).
extern "C" qf_status qf_plan_build_parameterized(const qf_pop* ops, size_t nops, uint32_t nparams, uint32_t ninputs,
                                                 const double* inputs, size_t inputs_len, uint32_t qubits,
                                                 int32_t reverse, const qf_plan_options* options, qf_plan** out) {
    if (!out || (!ops && nops) || (!inputs && inputs_len)) return fail(QF_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    qf_status s = check_qubits(qubits);
    if (s) return s;
    PCirc pc;
    pc.ops = ops;
    pc.nops = nops;
    pc.nparams = nparams;
    pc.ninputs = ninputs;
    if ((s = validate_ops(pc, qubits))) return s;
    if (inputs_len != ninputs)
        return fail(QF_INVALID_INPUT_COUNT, "expected " + std::to_string(ninputs) + " encoded inputs, got " +
                                                std::to_string(inputs_len));
    qf::PlanOptions po = plan_options_from(options);
    po.fuse = 1;
    const bool compiled = po.kernel == 3 || (po.kernel == 0 && (po.compile || qubits >= qf::kCompileMinQubits));
    // A reverse plan relabels only when it runs compiled (its adjoint form);
    // compiled plans take the gradient engine's (scale, ratio) rows.
    if (reverse && !compiled) po.relabel = 0;
    po.scaled_rotations = compiled && scaled_rotations_on() ? 1 : 0;
    qf_plan* p = new qf_plan;
    const int precision = options ? options->precision : QF_C128;
    if (precision != QF_C128 && precision != QF_C64) {
        delete p;
        return fail(QF_INVALID_ARGUMENT, "unknown precision");
    }
    const std::string e = qf::build_plan(param_gates(pc, inputs, reverse != 0), qubits, precision, po, p->plan);
    if (!e.empty()) {
        delete p;
        return fail(QF_INVALID_ARGUMENT, e);
    }
    p->plan.origin = qf::PLAN_PARAMETERIZED;
    *out = p;
    return QF_OK;
}

extern "C" qf_status qf_backend_set_profiling(qf_backend* be, int32_t on) {
    if (!be) return fail(QF_INVALID_ARGUMENT, "null backend");
    be->profile = on ? 1 : 0;
    return QF_OK;
}

extern "C" qf_status qf_backend_last_gradient_profile(const qf_backend* be, qf_gradient_profile* out) {
    if (!be || !out) return fail(QF_INVALID_ARGUMENT, "null argument");
    *out = be->grad_prof;
    return QF_OK;
}
