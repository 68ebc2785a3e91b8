// FP64 CPU oracle for the state-vector apply path.
//
// TEST INFRASTRUCTURE ONLY.  This file restates the reference crate's CPU
// algorithm (crate `quantumforge` 0.2.10, Rust) so that the HIP engine can be
// checked against it.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker / the timed CPU
// baseline; the product library never links or calls it.
//
// Compiled with -ffp-contract=off: Rust never contracts a*b+c into an FMA,
// so neither may the restatement (the complex product below is the exact
// operation sequence of `impl Mul for Complex64`, src/lib.rs:486-494).
//
// Every function names the reference lines it follows.  Pinning of each
// function (golden vectors, closed forms, brute force) lives in
// tests/test_oracle_*.py.
#include <cmath>
#include <limits>
#include <cstdint>
#include <cstring>
#include <map>
#include <vector>
#include <algorithm>
#include <array>
#include <chrono>
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" double or_rng_next_f64(uint64_t* s);

namespace {

struct C {  // src/lib.rs:437-445  repr(C) {re, im}
    double re, im;
};
inline C cadd(C a, C b) { return {a.re + b.re, a.im + b.im}; }           // lib.rs:472-477
inline C csub(C a, C b) { return {a.re - b.re, a.im - b.im}; }           // lib.rs:478-483
inline C cmul(C a, C b) {                                                // lib.rs:486-494
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
inline C cmulf(C a, double s) { return {a.re * s, a.im * s}; }           // lib.rs:496-500
inline C cconj(C a) { return {a.re, -a.im}; }                            // lib.rs:467-469
inline double cnorm(C a) { return a.re * a.re + a.im * a.im; }           // lib.rs:459-461

// Status codes shared with include/qf_hip.h.
enum {
    OK = 0,
    ZERO_QUBITS = 1,
    QUBIT_OUT_OF_RANGE = 2,
    INVALID_QUBIT_PAIR = 3,
    NON_FINITE_ANGLE = 4,
    STATE_TOO_LARGE = 5,
    INVALID_PARAMETER_COUNT = 8,
    INVALID_INPUT_COUNT = 9,
    NON_FINITE_PARAMETER = 10,
    REQUIRES_DENSITY_MATRIX = 13,
    REQUIRES_SHOT_EXECUTION = 14,
    INVALID_ARGUMENT = 15,
    INVALID_FD_STEP = 20,
    INVALID_PROBABILITY = 22,
    INVALID_SHOT_COUNT = 23,
    CLASSICAL_BIT_OUT_OF_RANGE = 24,
};

// Gate ids (qf_gate in include/qf_hip.h) in the order of `enum Gate`, src/lib.rs:2239-2265.
enum { G_I, G_X, G_Y, G_Z, G_H, G_S, G_SDG, G_T, G_TDG, G_RX, G_RY, G_RZ };

// cos and sin as two separate libm calls, like the reference's f64::cos and
// f64::sin.  Kept out of line so that the compiler cannot merge a cos/sin pair
// of one argument into sincos, whose last bit can differ from cos's.
__attribute__((noinline)) static double lcos(double x) { return std::cos(x); }
__attribute__((noinline)) static double lsin(double x) { return std::sin(x); }

// Gate::matrix, src/lib.rs:2267-2333.  (m00, m01, m10, m11) row major.
void gate_matrix(int gate, double theta, C m[4]) {
    const C z{0, 0}, o{1, 0}, i{0, 1};
    const double s = 0.70710678118654752440;  // std::f64::consts::FRAC_1_SQRT_2
    switch (gate) {
        case G_I: m[0] = o; m[1] = z; m[2] = z; m[3] = o; break;
        case G_X: m[0] = z; m[1] = o; m[2] = o; m[3] = z; break;
        case G_Y: m[0] = z; m[1] = C{0, -1}; m[2] = i; m[3] = z; break;
        case G_Z: m[0] = o; m[1] = z; m[2] = z; m[3] = C{-1, 0}; break;
        case G_H: m[0] = C{s, 0}; m[1] = C{s, 0}; m[2] = C{s, 0}; m[3] = C{-s, 0}; break;
        case G_S: m[0] = o; m[1] = z; m[2] = z; m[3] = i; break;
        case G_SDG: m[0] = o; m[1] = z; m[2] = z; m[3] = C{0, -1}; break;
        case G_T: m[0] = o; m[1] = z; m[2] = z; m[3] = C{s, s}; break;
        case G_TDG: m[0] = o; m[1] = z; m[2] = z; m[3] = C{s, -s}; break;
        case G_RX: {
            double c = lcos(theta * 0.5), si = lsin(theta * 0.5);
            m[0] = C{c, 0}; m[1] = C{0, -si}; m[2] = C{0, -si}; m[3] = C{c, 0};
            break;
        }
        case G_RY: {
            double c = lcos(theta * 0.5), si = lsin(theta * 0.5);
            m[0] = C{c, 0}; m[1] = C{-si, 0}; m[2] = C{si, 0}; m[3] = C{c, 0};
            break;
        }
        default: {  // G_RZ
            double half = theta * 0.5;
            m[0] = C{lcos(half), -lsin(half)}; m[1] = z; m[2] = z;
            m[3] = C{lcos(half), lsin(half)};
            break;
        }
    }
}

// adjoint_gate, src/lib.rs:2370-2382.
void adjoint_gate(int gate, double theta, int* out_gate, double* out_theta) {
    *out_theta = theta;
    switch (gate) {
        case G_RX: case G_RY: case G_RZ: *out_gate = gate; *out_theta = -theta; break;
        case G_S: *out_gate = G_SDG; break;
        case G_SDG: *out_gate = G_S; break;
        case G_T: *out_gate = G_TDG; break;
        case G_TDG: *out_gate = G_T; break;
        default: *out_gate = gate; break;
    }
}

struct State {
    std::vector<C> amps;
    uint32_t qubits;
};

// StateVector::apply, rayon branch, src/lib.rs:1017-1044: clone, parallel map, collect.
int apply_1q_m(State& st, const C m[4], uint32_t q) {
    if (q >= st.qubits) return QUBIT_OUT_OF_RANGE;
    const size_t mask = size_t(1) << q;
    const size_t n = st.amps.size();
    std::vector<C> old(st.amps);          // self.amplitudes.clone()
    std::vector<C> next(n);               // collect() target
    const C* o = old.data();
    C* w = next.data();
#pragma omp parallel for schedule(static)
    for (int64_t sidx = 0; sidx < int64_t(n); ++sidx) {
        size_t idx = size_t(sidx);
        if ((idx & mask) == 0) {
            w[idx] = cadd(cmul(m[0], o[idx]), cmul(m[1], o[idx | mask]));
        } else {
            w[idx] = cadd(cmul(m[2], o[idx & ~mask]), cmul(m[3], o[idx]));
        }
    }
    st.amps.swap(next);                   // self.amplitudes = new
    return OK;
}

int apply_1q(State& st, int gate, double theta, uint32_t q) {
    if (q >= st.qubits) return QUBIT_OUT_OF_RANGE;
    C m[4];
    gate_matrix(gate, theta, m);
    return apply_1q_m(st, m, q);
}

// check_pair, src/lib.rs:1124-1141.
int check_pair(const State& st, uint32_t a, uint32_t b) {
    if (a >= st.qubits) return QUBIT_OUT_OF_RANGE;
    if (b >= st.qubits) return QUBIT_OUT_OF_RANGE;
    if (a == b) return INVALID_QUBIT_PAIR;
    return OK;
}

// Shared free-qubit enumeration of apply_cnot / apply_cz / apply_swap
// (src/lib.rs:1062-1122): sequential, index rebuilt bit by bit.
template <class F>
void for_free(const State& st, uint32_t a, uint32_t b, size_t seed, F f) {
    std::vector<uint32_t> free_q;
    for (uint32_t q = 0; q < st.qubits; ++q)
        if (q != a && q != b) free_q.push_back(q);
    const size_t n_free = free_q.size();
    for (size_t fv = 0; fv < (size_t(1) << n_free); ++fv) {
        size_t base = seed;
        for (size_t i = 0; i < n_free; ++i)
            if (fv & (size_t(1) << i)) base |= size_t(1) << free_q[i];
        f(base);
    }
}

// StateVector::apply_cnot, src/lib.rs:1062-1080.
int apply_cnot(State& st, uint32_t c, uint32_t t) {
    int e = check_pair(st, c, t);
    if (e) return e;
    const size_t tm = size_t(1) << t;
    for_free(st, c, t, size_t(1) << c, [&](size_t base) { std::swap(st.amps[base], st.amps[base | tm]); });
    return OK;
}

// StateVector::apply_cz, src/lib.rs:1083-1101.
int apply_cz(State& st, uint32_t a, uint32_t b) {
    int e = check_pair(st, a, b);
    if (e) return e;
    for_free(st, a, b, (size_t(1) << a) | (size_t(1) << b),
             [&](size_t idx) { st.amps[idx] = cmulf(st.amps[idx], -1.0); });
    return OK;
}

// StateVector::apply_swap, src/lib.rs:1104-1122.
int apply_swap(State& st, uint32_t a, uint32_t b) {
    int e = check_pair(st, a, b);
    if (e) return e;
    const size_t flip = (size_t(1) << a) ^ (size_t(1) << b);
    for_free(st, a, b, size_t(1) << b, [&](size_t idx) { std::swap(st.amps[idx], st.amps[idx ^ flip]); });
    return OK;
}

// Exact controlled phase diag(1,1,1,e^{i phi}) (no reference counterpart: the
// crate lowers it to Rz/CNOT, src/algorithms.rs:139-145, = this x e^{-i phi/4}).
int apply_cp(State& st, uint32_t a, uint32_t b, double phi) {
    int e = check_pair(st, a, b);
    if (e) return e;
    const C ph{lcos(phi), lsin(phi)};
    for_free(st, a, b, (size_t(1) << a) | (size_t(1) << b), [&](size_t idx) { st.amps[idx] = cmul(ph, st.amps[idx]); });
    return OK;
}

constexpr double kEps = std::numeric_limits<double>::epsilon();

// StateVector::norm_sqr, src/lib.rs:1000-1002: sequential left-to-right sum.
double norm_sqr(const State& st) {
    double acc = 0.0;
    for (const C& a : st.amps) acc += cnorm(a);
    return acc;
}

// StateVector::normalize, src/lib.rs:1005-1014.
int normalize(State& st) {
    const double norm = std::sqrt(norm_sqr(st));
    if (norm == 0.0 || !std::isfinite(norm)) return INVALID_PROBABILITY;
    for (C& a : st.amps) a = C{a.re / norm, a.im / norm};
    return OK;
}

// Noise channels (NoiseChannel, src/lib.rs:1561-1627), numbered as their
// declaration order; p0 / p1 are NoiseChannel::params().
enum { CH_DEPOLARIZING = 0, CH_AMPLITUDE_DAMPING = 1, CH_PHASE_DAMPING = 2, CH_THERMAL_DAMPING = 3 };

// validate_probability, src/lib.rs:2229-2235.
bool valid_probability(double p) { return std::isfinite(p) && p >= 0.0 && p <= 1.0; }

// NoiseChannel::thermal_rates, src/lib.rs:1611-1627: (gamma_up, gamma_down).
int thermal_rates(double gamma_down, double beta_delta, double* up, double* down) {
    if (!valid_probability(gamma_down)) return INVALID_PROBABILITY;
    if (!std::isfinite(beta_delta) || beta_delta < 0.0) return INVALID_PROBABILITY;
    *up = gamma_down * std::exp(-beta_delta);
    *down = gamma_down;
    return OK;
}

int check_channel(int kind, double p0, double p1) {
    switch (kind) {
        case CH_DEPOLARIZING:
        case CH_AMPLITUDE_DAMPING:
        case CH_PHASE_DAMPING: return valid_probability(p0) ? OK : INVALID_PROBABILITY;
        case CH_THERMAL_DAMPING: {
            double u, d;
            return thermal_rates(p0, p1, &u, &d);
        }
        default: return INVALID_ARGUMENT;
    }
}

// Rust's saturating float -> u8 cast (`as u8`): NaN and negatives give 0.
int as_u8(double x) { return !(x > 0.0) ? 0 : (x >= 255.0 ? 255 : int(x)); }

// StateVector::apply_channel_trajectory, src/lib.rs:1356-1505: one uniform
// per channel, drawn from `rng` after validation (and, for the damping
// channels, after the norm and P(1) sums).
int channel_trajectory(State& st, int kind, double p0, double p1, uint32_t q, uint64_t* rng) {
    if (q >= st.qubits) return QUBIT_OUT_OF_RANGE;
    const size_t mask = size_t(1) << q;
    const size_t n = st.amps.size();
    if (kind == CH_DEPOLARIZING) {
        if (!valid_probability(p0)) return INVALID_PROBABILITY;
        const double sample = or_rng_next_f64(rng);
        if (sample >= 1.0 - p0 && p0 > 0.0) {
            const int branch = std::min(as_u8(std::floor((sample - (1.0 - p0)) * (3.0 / p0))), 2);
            return apply_1q(st, branch == 0 ? G_X : (branch == 1 ? G_Y : G_Z), 0.0, q);
        }
        return OK;
    }
    double up = 0.0, down = 0.0;
    if (kind == CH_THERMAL_DAMPING) {
        const int e = thermal_rates(p0, p1, &up, &down);
        if (e) return e;
    } else if (kind == CH_AMPLITUDE_DAMPING || kind == CH_PHASE_DAMPING) {
        if (!valid_probability(p0)) return INVALID_PROBABILITY;
    } else {
        return INVALID_ARGUMENT;
    }
    const double norm = norm_sqr(st);
    if (!std::isfinite(norm) || norm <= kEps) return INVALID_PROBABILITY;
    double p_one = 0.0;
    for (size_t i = 0; i < n; ++i)
        if (i & mask) p_one += cnorm(st.amps[i]);
    if (kind != CH_THERMAL_DAMPING) {
        const double gamma = p0;
        const double jump_probability = std::min(std::max(gamma * p_one / norm, 0.0), 1.0);
        const bool jump = or_rng_next_f64(rng) < jump_probability;
        const double branch_norm = jump ? gamma * p_one : norm - gamma * p_one;
        if (!std::isfinite(branch_norm) || branch_norm <= kEps) return INVALID_PROBABILITY;
        const double inverse = 1.0 / std::sqrt(branch_norm);
        if (jump) {
            const double scale = std::sqrt(gamma) * inverse;
            for (size_t base = 0; base < n; ++base) {
                if (base & mask) continue;
                const C value = cmulf(st.amps[base | mask], scale);
                if (kind == CH_AMPLITUDE_DAMPING) {
                    st.amps[base] = value;
                    st.amps[base | mask] = C{0.0, 0.0};
                } else {
                    st.amps[base] = C{0.0, 0.0};
                    st.amps[base | mask] = value;
                }
            }
        } else {
            const double excited = std::sqrt(1.0 - gamma) * inverse;
            for (size_t i = 0; i < n; ++i) st.amps[i] = cmulf(st.amps[i], (i & mask) ? excited : inverse);
        }
        return OK;
    }
    const double pr1 = std::min(std::max(p_one / norm, 0.0), 1.0);
    const double pr0 = 1.0 - pr1;
    const double go_down = pr1 * down, go_up = pr0 * up;
    const double stay = pr0 * (1.0 - up) + pr1 * (1.0 - down);
    const double sample = or_rng_next_f64(rng);
    double branch_norm;
    int branch;
    if (sample < go_down) {
        branch_norm = go_down;
        branch = 1;
    } else if (sample < go_down + go_up) {
        branch_norm = go_up;
        branch = 2;
    } else {
        branch_norm = stay;
        branch = 0;
    }
    if (!std::isfinite(branch_norm) || branch_norm <= kEps) return INVALID_PROBABILITY;
    const double inverse = 1.0 / std::sqrt(branch_norm);
    if (branch == 1) {
        const double scale = std::sqrt(down) * inverse;
        for (size_t base = 0; base < n; ++base) {
            if (base & mask) continue;
            st.amps[base] = cmulf(st.amps[base | mask], scale);
            st.amps[base | mask] = C{0.0, 0.0};
        }
    } else if (branch == 2) {
        const double scale = std::sqrt(up) * inverse;
        for (size_t base = 0; base < n; ++base) {
            if (base & mask) continue;
            st.amps[base | mask] = cmulf(st.amps[base], scale);
            st.amps[base] = C{0.0, 0.0};
        }
    } else {
        const double ground = std::sqrt(1.0 - up) * inverse, excited = std::sqrt(1.0 - down) * inverse;
        for (size_t i = 0; i < n; ++i) st.amps[i] = cmulf(st.amps[i], (i & mask) ? excited : ground);
    }
    return OK;
}

// ---- DensityMatrix, src/lib.rs:1553-2227: row-major entries[row * dim + col].
struct Density {
    std::vector<C> e;
    size_t dim;
    uint32_t qubits;
};

// DensityMatrix::local_transform, src/lib.rs:2157-2198: rows with m, then
// columns with conj(m), through a full temporary.
std::vector<C> local_transform(const Density& d, const C m[4], uint32_t q) {
    const size_t mask = size_t(1) << q, D = d.dim;
    std::vector<C> left(d.e.size());
    for (size_t row = 0; row < D; ++row) {
        const size_t row0 = row & ~mask, row1 = row0 | mask;
        for (size_t col = 0; col < D; ++col) {
            const C a = d.e[row0 * D + col], b = d.e[row1 * D + col];
            left[row * D + col] = (row & mask) == 0 ? cadd(cmul(m[0], a), cmul(m[1], b))
                                                    : cadd(cmul(m[2], a), cmul(m[3], b));
        }
    }
    std::vector<C> out(d.e.size());
    for (size_t row = 0; row < D; ++row)
        for (size_t col = 0; col < D; ++col) {
            const size_t col0 = col & ~mask, col1 = col0 | mask;
            const C a = left[row * D + col0], b = left[row * D + col1];
            out[row * D + col] = (col & mask) == 0 ? cadd(cmul(a, cconj(m[0])), cmul(b, cconj(m[1])))
                                                   : cadd(cmul(a, cconj(m[2])), cmul(b, cconj(m[3])));
        }
    return out;
}

// DensityMatrix::apply, src/lib.rs:2044-2054.
int density_apply(Density& d, int gate, double theta, uint32_t q) {
    if (q >= d.qubits) return QUBIT_OUT_OF_RANGE;
    C m[4];
    gate_matrix(gate, theta, m);
    d.e = local_transform(d, m, q);
    return OK;
}

// check_qubit_pair, src/lib.rs:1532-1550.
int density_pair(const Density& d, uint32_t a, uint32_t b) {
    if (a >= d.qubits || b >= d.qubits) return QUBIT_OUT_OF_RANGE;
    return a == b ? INVALID_QUBIT_PAIR : OK;
}

// apply_basis_permutation, src/lib.rs:2033-2042.
template <class F>
void density_permute(Density& d, F perm) {
    const std::vector<C> original(d.e);
    const size_t D = d.dim;
    for (size_t row = 0; row < D; ++row)
        for (size_t col = 0; col < D; ++col) d.e[row * D + col] = original[perm(row) * D + perm(col)];
}

// DensityMatrix::apply_cnot / apply_cz / apply_swap, src/lib.rs:1984-2031.
int density_2q(Density& d, int kind, uint32_t a, uint32_t b, double phi) {
    const int e = density_pair(d, a, b);
    if (e) return e;
    const size_t ma = size_t(1) << a, mb = size_t(1) << b, D = d.dim;
    if (kind == 1) {
        density_permute(d, [&](size_t i) { return (i & ma) ? i ^ mb : i; });
    } else if (kind == 3) {
        density_permute(d, [&](size_t i) { return ((i & ma) == 0) != ((i & mb) == 0) ? i ^ ma ^ mb : i; });
    } else if (kind == 2) {
        for (size_t row = 0; row < D; ++row)
            for (size_t col = 0; col < D; ++col) {
                C& x = d.e[row * D + col];
                if ((row & ma) && (row & mb)) x = cmulf(x, -1.0);
                if ((col & ma) && (col & mb)) x = cmulf(x, -1.0);
            }
    } else {  // exact controlled phase (no reference counterpart), rho -> CP rho CP^dagger
        const C ph{lcos(phi), lsin(phi)};
        for (size_t row = 0; row < D; ++row)
            for (size_t col = 0; col < D; ++col) {
                C& x = d.e[row * D + col];
                if ((row & ma) && (row & mb)) x = cmul(ph, x);
                if ((col & ma) && (col & mb)) x = cmul(x, cconj(ph));
            }
    }
    return OK;
}

// DensityMatrix::apply_kraus, src/lib.rs:2131-2144.
void apply_kraus(Density& d, const std::vector<std::array<C, 4>>& ks, uint32_t q) {
    std::vector<C> sum(d.e.size(), C{0.0, 0.0});
    for (const auto& k : ks) {
        const std::vector<C> branch = local_transform(d, k.data(), q);
        for (size_t i = 0; i < sum.size(); ++i) sum[i] = cadd(sum[i], branch[i]);
    }
    d.e.swap(sum);
}

// DensityMatrix::apply_channel, src/lib.rs:2057-2129, and depolarize,
// src/lib.rs:2200-2226.
int density_channel(Density& d, int kind, double p0, double p1, uint32_t q) {
    if (q >= d.qubits) return QUBIT_OUT_OF_RANGE;
    const C z{0.0, 0.0}, one{1.0, 0.0};
    switch (kind) {
        case CH_DEPOLARIZING: {
            if (!valid_probability(p0)) return INVALID_PROBABILITY;
            const Density original = d;
            std::vector<C> mixed(d.e.size(), z);
            for (int gate : {int(G_X), int(G_Y), int(G_Z)}) {
                C m[4];
                gate_matrix(gate, 0.0, m);
                const std::vector<C> branch = local_transform(original, m, q);
                for (size_t i = 0; i < mixed.size(); ++i) mixed[i] = cadd(mixed[i], cmulf(branch[i], p0 / 3.0));
            }
            for (size_t i = 0; i < mixed.size(); ++i) mixed[i] = cadd(mixed[i], cmulf(original.e[i], 1.0 - p0));
            d.e.swap(mixed);
            return OK;
        }
        case CH_AMPLITUDE_DAMPING:
        case CH_PHASE_DAMPING: {
            if (!valid_probability(p0)) return INVALID_PROBABILITY;
            const std::array<C, 4> k0{one, z, z, C{std::sqrt(1.0 - p0), 0.0}};
            const std::array<C, 4> k1 = kind == CH_AMPLITUDE_DAMPING ? std::array<C, 4>{z, C{std::sqrt(p0), 0.0}, z, z}
                                                                     : std::array<C, 4>{z, z, z, C{std::sqrt(p0), 0.0}};
            apply_kraus(d, {k0, k1}, q);
            return OK;
        }
        case CH_THERMAL_DAMPING: {
            double up, down;
            const int e = thermal_rates(p0, p1, &up, &down);
            if (e) return e;
            const std::array<C, 4> k0{C{std::sqrt(1.0 - up), 0.0}, z, z, C{std::sqrt(1.0 - down), 0.0}};
            const std::array<C, 4> k1{z, C{std::sqrt(down), 0.0}, z, z};
            const std::array<C, 4> k2{z, z, C{std::sqrt(up), 0.0}, z};
            apply_kraus(d, {k0, k1, k2}, q);
            return OK;
        }
        default: return INVALID_ARGUMENT;
    }
}

// DensityMatrix::probability_one, src/lib.rs:1908-1921.
double density_p1(const Density& d, uint32_t q) {
    const size_t mask = size_t(1) << q;
    double acc = 0.0;
    for (size_t i = 0; i < d.dim; ++i)
        if (i & mask) acc += d.e[i * d.dim + i].re;
    return std::min(std::max(acc, 0.0), 1.0);
}

// DensityMatrix::measure, src/lib.rs:1924-1950.
int density_measure(Density& d, uint32_t q, double u, uint8_t* outcome) {
    if (q >= d.qubits) return QUBIT_OUT_OF_RANGE;
    const double p_one = density_p1(d, q);
    const uint8_t o = uint8_t(u < p_one);
    const double probability = o ? p_one : 1.0 - p_one;
    if (probability <= kEps) return INVALID_PROBABILITY;
    const size_t mask = size_t(1) << q, D = d.dim;
    for (size_t row = 0; row < D; ++row)
        for (size_t col = 0; col < D; ++col) {
            C& x = d.e[row * D + col];
            if (uint8_t((row & mask) != 0) != o || uint8_t((col & mask) != 0) != o) x = C{0.0, 0.0};
            else x = C{x.re / probability, x.im / probability};
        }
    *outcome = o;
    return OK;
}

}  // namespace

extern "C" {

// ---- op records (layout identical to qf_op / qf_pop / qf_pauli_term in include/qf_hip.h)
struct or_op {
    int32_t kind;   // 0 single-qubit, 1 CNOT, 2 CZ, 3 SWAP, 4 channel {gate = channel kind, a, theta, theta2},
                    // 5 measure, 6 conditional, 7 controlled phase
    int32_t gate;
    uint32_t a, b;
    double theta;
    double theta2;  // second channel parameter (ThermalDamping beta_delta)
};
struct or_pop {
    int32_t kind;   // 0 fixed, 1 rotation, 2 encoding rotation, 3 CNOT, 4 CZ, 5 SWAP, 6 channel
    int32_t gate;   // gate id (fixed) or axis 0=X 1=Y 2=Z (rotations)
    uint32_t a, b;
    uint32_t index; // parameter / input index
    uint32_t pad;
    double theta;   // angle of a fixed rotation gate
};
struct or_pauli_term {
    double re, im;
    uint64_t x_mask, z_mask;  // Y on q <=> both bits set
};

// ---- SmallRng, src/lib.rs:4881-4912
uint64_t or_rng_seeded(uint64_t seed) {
    uint64_t z = seed + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return z == 0 ? 0x9E3779B97F4A7C15ull : z;
}
uint64_t or_rng_next_u64(uint64_t* s) {
    uint64_t x = *s;
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    *s = x;
    return x;
}
double or_rng_next_f64(uint64_t* s) {
    return double(or_rng_next_u64(s) >> 11) * (1.0 / double(uint64_t(1) << 53));
}

void or_gate_matrix(int gate, double theta, double* out8) {
    C m[4];
    gate_matrix(gate, theta, m);
    std::memcpy(out8, m, sizeof(m));
}

// ---- state handle
void* or_state_zero(uint32_t qubits) {  // StateVector::zero, src/lib.rs:930-966
    if (qubits == 0 || qubits > 40) return nullptr;
    State* st = new State;
    st->qubits = qubits;
    st->amps.assign(size_t(1) << qubits, C{0, 0});
    st->amps[0] = C{1, 0};
    return st;
}
void* or_state_from(const double* amps, uint32_t qubits) {  // from_amplitudes, src/lib.rs:969-978
    State* st = new State;
    st->qubits = qubits;
    st->amps.resize(size_t(1) << qubits);
    std::memcpy(st->amps.data(), amps, st->amps.size() * sizeof(C));
    return st;
}
void or_state_free(void* s) { delete static_cast<State*>(s); }
void or_state_get(const void* s, double* out) {
    const State* st = static_cast<const State*>(s);
    std::memcpy(out, st->amps.data(), st->amps.size() * sizeof(C));
}
int or_apply_gate(void* s, int gate, double theta, uint32_t q) {
    return apply_1q(*static_cast<State*>(s), gate, theta, q);
}
int or_apply_matrix(void* s, const double* m8, uint32_t q) {
    C m[4];
    std::memcpy(m, m8, sizeof(m));
    return apply_1q_m(*static_cast<State*>(s), m, q);
}
int or_apply_cnot(void* s, uint32_t c, uint32_t t) { return apply_cnot(*static_cast<State*>(s), c, t); }
int or_apply_cz(void* s, uint32_t a, uint32_t b) { return apply_cz(*static_cast<State*>(s), a, b); }
int or_apply_swap(void* s, uint32_t a, uint32_t b) { return apply_swap(*static_cast<State*>(s), a, b); }

// Circuit::run, src/lib.rs:2786-2802 (ops validated one at a time, as they run).
int or_run(void* s, const or_op* ops, size_t nops) {
    State& st = *static_cast<State*>(s);
    for (size_t i = 0; i < nops; ++i) {
        const or_op& op = ops[i];
        int e = OK;
        switch (op.kind) {
            case 0: e = apply_1q(st, op.gate, op.theta, op.a); break;
            case 1: e = apply_cnot(st, op.a, op.b); break;
            case 2: e = apply_cz(st, op.a, op.b); break;
            case 3: e = apply_swap(st, op.a, op.b); break;
            case 4: return REQUIRES_DENSITY_MATRIX;
            case 5: case 6: return REQUIRES_SHOT_EXECUTION;
            case 7: e = apply_cp(st, op.a, op.b, op.theta); break;
            default: return INVALID_ARGUMENT;
        }
        if (e) return e;
    }
    return OK;
}

// Wall-clock of or_run over `reps` fresh |0..0> states (bench CPU leg).
double or_time_run(uint32_t qubits, const or_op* ops, size_t nops, int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
    State* st = static_cast<State*>(or_state_zero(qubits));
    auto t0 = std::chrono::steady_clock::now();
    int e = or_run(st, ops, nops);
    auto t1 = std::chrono::steady_clock::now();
    or_state_free(st);
    if (e) return -double(e);
    return std::chrono::duration<double>(t1 - t0).count();
}
int or_max_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// ---- measurement and shots
// StateVector::measure, src/lib.rs:1321-1350: P(1) summed left to right over
// the indices with bit q set, outcome = u < P(1), branch probability checked
// against f64::EPSILON, kept amplitudes divided by sqrt(probability).
int measure_state(State& st, uint32_t q, double u, uint8_t* outcome) {
    if (q >= st.qubits) return QUBIT_OUT_OF_RANGE;
    const size_t mask = size_t(1) << q;
    double p_one = 0.0;
    for (size_t i = 0; i < st.amps.size(); ++i)
        if (i & mask) p_one += cnorm(st.amps[i]);
    p_one = std::min(std::max(p_one, 0.0), 1.0);
    const uint8_t o = uint8_t(u < p_one);
    const double probability = o ? p_one : 1.0 - p_one;
    if (probability <= std::numeric_limits<double>::epsilon()) return INVALID_PROBABILITY;
    const double scale = std::sqrt(probability);
    for (size_t i = 0; i < st.amps.size(); ++i) {
        if (uint8_t((i & mask) != 0) != o) st.amps[i] = C{0.0, 0.0};
        else st.amps[i] = C{st.amps[i].re / scale, st.amps[i].im / scale};  // Complex64 / f64, lib.rs:502-505
    }
    *outcome = o;
    return OK;
}

int or_measure(void* s, uint32_t q, double u, uint8_t* outcome) {
    return measure_state(*static_cast<State*>(s), q, u, outcome);
}

// Circuit::run_shots / run_shots_with_noise (allow_noise = 1),
// src/lib.rs:2822-2911: op kinds as or_op, 5 = measure {a target, b classical
// bit}, 6 = conditional gate {gate, a target, theta, b = classical bit | value
// << 16}, 4 = channel (a Kraus trajectory with noise, else
// OperationRequiresDensityMatrix).  With noise each shot's state is normalized
// first.  Writes every shot's classical register ([shots][width],
// least-significant bit first).
int or_run_shots(const double* initial, uint32_t qubits, const or_op* ops, size_t nops, uint64_t shots,
                 uint64_t seed, uint8_t* registers, uint32_t* width_out, int allow_noise) {
    if (shots == 0) return INVALID_SHOT_COUNT;
    uint32_t width = 0;
    for (size_t i = 0; i < nops; ++i) {
        if (ops[i].kind != 5 && ops[i].kind != 6) continue;
        const uint32_t bit = ops[i].kind == 5 ? ops[i].b : (ops[i].b & 0xFFFFu);
        if (bit >= 4096) return CLASSICAL_BIT_OUT_OF_RANGE;
        width = std::max(width, bit + 1);
    }
    *width_out = width;
    if (!registers) return OK;
    uint64_t rng = or_rng_seeded(seed);
    const size_t dim = size_t(1) << qubits;
    std::vector<uint8_t> classical(width);
    for (uint64_t shot = 0; shot < shots; ++shot) {
        State st;
        st.qubits = qubits;
        st.amps.resize(dim);
        std::memcpy(st.amps.data(), initial, dim * sizeof(C));
        if (allow_noise) {
            const int e = normalize(st);
            if (e) return e;
        }
        std::fill(classical.begin(), classical.end(), 0);
        for (size_t i = 0; i < nops; ++i) {
            const or_op& op = ops[i];
            int e = OK;
            switch (op.kind) {
                case 0: e = apply_1q(st, op.gate, op.theta, op.a); break;
                case 1: e = apply_cnot(st, op.a, op.b); break;
                case 2: e = apply_cz(st, op.a, op.b); break;
                case 3: e = apply_swap(st, op.a, op.b); break;
                case 4:
                    if (!allow_noise) return REQUIRES_DENSITY_MATRIX;
                    e = channel_trajectory(st, op.gate, op.theta, op.theta2, op.a, &rng);
                    break;
                case 5: {
                    const double u = or_rng_next_f64(&rng);
                    e = measure_state(st, op.a, u, &classical[op.b]);
                    break;
                }
                case 6:
                    if (classical[op.b & 0xFFFFu] == ((op.b >> 16) & 1u)) e = apply_1q(st, op.gate, op.theta, op.a);
                    break;
                case 7: e = apply_cp(st, op.a, op.b, op.theta); break;
                default: return INVALID_ARGUMENT;
            }
            if (e) return e;
        }
        if (width) std::memcpy(registers + shot * width, classical.data(), width);
    }
    return OK;
}

int or_channel_trajectory(void* s, int kind, double p0, double p1, uint32_t q, uint64_t* rng) {
    return channel_trajectory(*static_cast<State*>(s), kind, p0, p1, q, rng);
}
int or_normalize(void* s) { return normalize(*static_cast<State*>(s)); }

// ---- density matrices (DensityMatrix, src/lib.rs:1553-2227)
void* or_density_from(const double* entries, uint32_t qubits) {  // DensityMatrix::from_entries, src/lib.rs:1707-1715
    if (qubits == 0 || qubits > 14) return nullptr;
    Density* d = new Density;
    d->qubits = qubits;
    d->dim = size_t(1) << qubits;
    d->e.resize(d->dim * d->dim);
    std::memcpy(d->e.data(), entries, d->e.size() * sizeof(C));
    return d;
}
void* or_density_from_state(const double* amps, uint32_t qubits) {  // DensityMatrix::from_state, src/lib.rs:1642-1656
    if (qubits == 0 || qubits > 14) return nullptr;
    Density* d = new Density;
    d->qubits = qubits;
    d->dim = size_t(1) << qubits;
    d->e.resize(d->dim * d->dim);
    const C* a = reinterpret_cast<const C*>(amps);
    for (size_t row = 0; row < d->dim; ++row)
        for (size_t col = 0; col < d->dim; ++col) d->e[row * d->dim + col] = cmul(a[row], cconj(a[col]));
    return d;
}
void or_density_free(void* d) { delete static_cast<Density*>(d); }
void or_density_get(const void* d, double* out) {
    const Density* x = static_cast<const Density*>(d);
    std::memcpy(out, x->e.data(), x->e.size() * sizeof(C));
}
int or_density_apply_gate(void* d, int gate, double theta, uint32_t q) {
    return density_apply(*static_cast<Density*>(d), gate, theta, q);
}
// kind: 1 CNOT {a control, b target}, 2 CZ, 3 SWAP, 7 controlled phase.
int or_density_apply_2q(void* d, int kind, uint32_t a, uint32_t b, double phi) {
    return density_2q(*static_cast<Density*>(d), kind, a, b, phi);
}
int or_density_apply_channel(void* d, int kind, double p0, double p1, uint32_t q) {
    return density_channel(*static_cast<Density*>(d), kind, p0, p1, q);
}
int or_density_measure(void* d, uint32_t q, double u, uint8_t* outcome) {
    return density_measure(*static_cast<Density*>(d), q, u, outcome);
}
int or_density_probability_one(const void* d, uint32_t q, double* out) {
    const Density& x = *static_cast<const Density*>(d);
    if (q >= x.qubits) return QUBIT_OUT_OF_RANGE;
    *out = density_p1(x, q);
    return OK;
}
// Circuit::run_density, src/lib.rs:2804-2819 (ops run and validated in order).
int or_density_run(void* d, const or_op* ops, size_t nops) {
    Density& x = *static_cast<Density*>(d);
    for (size_t i = 0; i < nops; ++i) {
        const or_op& op = ops[i];
        int e = OK;
        switch (op.kind) {
            case 0: e = density_apply(x, op.gate, op.theta, op.a); break;
            case 1: case 2: case 3: e = density_2q(x, op.kind, op.a, op.b, 0.0); break;
            case 7: e = density_2q(x, 7, op.a, op.b, op.theta); break;
            case 4: e = density_channel(x, op.gate, op.theta, op.theta2, op.a); break;
            case 5: case 6: return REQUIRES_SHOT_EXECUTION;
            default: return INVALID_ARGUMENT;
        }
        if (e) return e;
    }
    return OK;
}

// ---- reductions
double or_norm_sqr(const void* s) {  // src/lib.rs:1000-1002, sequential left-to-right
    const State& st = *static_cast<const State*>(s);
    double acc = 0.0;
    for (const C& a : st.amps) acc += cnorm(a);
    return acc;
}
void or_probabilities(const void* s, double* out) {  // src/lib.rs:1144-1146
    const State& st = *static_cast<const State*>(s);
    for (size_t i = 0; i < st.amps.size(); ++i) out[i] = cnorm(st.amps[i]);
}
// expectation_z / _x / _y, src/lib.rs:1149-1318.  obs: 0=X 1=Y 2=Z.
int or_expectation(const void* s, int obs, uint32_t q, double* out) {
    const State& st = *static_cast<const State*>(s);
    if (q >= st.qubits) return QUBIT_OUT_OF_RANGE;
    const size_t mask = size_t(1) << q;
    double acc = 0.0;
    if (obs == 2) {
        for (size_t i = 0; i < st.amps.size(); ++i) {
            double sign = (i & mask) == 0 ? 1.0 : -1.0;
            acc += sign * cnorm(st.amps[i]);
        }
    } else {
        for (size_t i = 0; i < st.amps.size(); ++i) {
            if (i & mask) continue;
            C p = cmul(cconj(st.amps[i]), st.amps[i | mask]);
            acc += (obs == 0 ? p.re : p.im) * 2.0;
        }
    }
    *out = acc;
    return OK;
}

// PauliWord::apply_basis_index, src/analytic.rs:2629-2652.
static void pauli_basis(const or_pauli_term& t, uint32_t qubits, size_t index, size_t* dest, C* phase) {
    C ph{1, 0};
    for (uint32_t q = 0; q < qubits; ++q) {
        const uint64_t m = uint64_t(1) << q;
        const bool x = t.x_mask & m, z = t.z_mask & m;
        const bool one = index & m;
        if (x && z) {  // Y
            ph = cmul(ph, one ? C{0, -1} : C{0, 1});
            index ^= m;
        } else if (x) {
            index ^= m;
        } else if (z && one) {
            ph = cmulf(ph, -1.0);
        }
    }
    *dest = index;
    *phase = ph;
}

// PauliSum::expectation_ket, src/analytic.rs:2751-2778 -> SparseOperator::new
// (entries merged per (row, col), src/quantum.rs:797-837) -> apply_amplitudes
// (src/quantum.rs:1340-1354) -> Bra::inner (src/quantum.rs:327-337).
int or_pauli_expectation(const void* s, const or_pauli_term* terms, size_t nterms, double* out_re,
                         double* out_im) {
    const State& st = *static_cast<const State*>(s);
    const size_t dim = st.amps.size();
    // Group terms by x_mask: every term of one group maps column -> row = column ^ x.
    std::map<uint64_t, std::vector<size_t>> groups;
    for (size_t t = 0; t < nterms; ++t) groups[terms[t].x_mask].push_back(t);
    C acc{0, 0};
    std::vector<std::pair<size_t, C>> row_entries;
    for (size_t row = 0; row < dim; ++row) {
        row_entries.clear();
        for (auto& g : groups) {
            const size_t col = row ^ size_t(g.first);
            C v{0, 0};
            bool any = false;
            for (size_t t : g.second) {
                size_t dest;
                C ph;
                pauli_basis(terms[t], st.qubits, col, &dest, &ph);
                C coeff{terms[t].re, terms[t].im};
                v = any ? cadd(v, cmul(coeff, ph)) : cmul(coeff, ph);
                any = true;
            }
            row_entries.emplace_back(col, v);
        }
        std::sort(row_entries.begin(), row_entries.end(),
                  [](const std::pair<size_t, C>& x, const std::pair<size_t, C>& y) { return x.first < y.first; });
        C hv{0, 0};
        for (auto& e : row_entries) hv = cadd(hv, cmul(e.second, st.amps[e.first]));
        acc = cadd(acc, cmul(cconj(st.amps[row]), hv));
    }
    *out_re = acc.re;
    *out_im = acc.im;
    return OK;
}

// H|psi> (sparse-operator product, same merge as above) written into out.
static void pauli_apply(const State& st, const or_pauli_term* terms, size_t nterms, std::vector<C>& out) {
    const size_t dim = st.amps.size();
    out.assign(dim, C{0, 0});
    for (size_t t = 0; t < nterms; ++t) {
        C coeff{terms[t].re, terms[t].im};
        for (size_t col = 0; col < dim; ++col) {
            size_t row;
            C ph;
            pauli_basis(terms[t], st.qubits, col, &row, &ph);
            out[row] = cadd(out[row], cmul(cmul(coeff, ph), st.amps[col]));
        }
    }
}

// ---- parameterized circuits
struct PCirc {
    const or_pop* ops;
    size_t nops;
    uint32_t nparams, ninputs;
};

// validate_parameters / validate_inputs, src/lib.rs:4651-4675.
static int validate(const PCirc& pc, const double* params, size_t np, const double* inputs, size_t ni) {
    if (ni != pc.ninputs) return INVALID_INPUT_COUNT;
    for (size_t i = 0; i < ni; ++i)
        if (!std::isfinite(inputs[i])) return NON_FINITE_PARAMETER;
    if (np != pc.nparams) return INVALID_PARAMETER_COUNT;
    for (size_t i = 0; i < np; ++i)
        if (!std::isfinite(params[i])) return NON_FINITE_PARAMETER;
    return OK;
}

static int axis_gate(int axis) { return axis == 0 ? G_RX : (axis == 1 ? G_RY : G_RZ); }

// run_with_shift, src/lib.rs:4677-4738.
static int run_with_shift(const PCirc& pc, const double* params, const double* inputs, State& st,
                          int64_t shift_op, double delta) {
    for (size_t i = 0; i < pc.nops; ++i) {
        const or_pop& op = pc.ops[i];
        int e = OK;
        switch (op.kind) {
            case 0: e = apply_1q(st, op.gate, op.theta, op.a); break;
            case 1: {
                if (op.index >= pc.nparams) return INVALID_PARAMETER_COUNT;
                double th = params[op.index] + (int64_t(i) == shift_op ? delta : 0.0);
                e = apply_1q(st, axis_gate(op.gate), th, op.a);
                break;
            }
            case 2:
                if (op.index >= pc.ninputs) return INVALID_INPUT_COUNT;
                e = apply_1q(st, axis_gate(op.gate), inputs[op.index], op.a);
                break;
            case 3: e = apply_cnot(st, op.a, op.b); break;
            case 4: e = apply_cz(st, op.a, op.b); break;
            case 5: e = apply_swap(st, op.a, op.b); break;
            case 6: return REQUIRES_DENSITY_MATRIX;
            default: return INVALID_ARGUMENT;
        }
        if (e) return e;
    }
    return OK;
}

// Observable = Pauli sum; a single-qubit observable is one term with coefficient 1.
static double energy(const State& st, const or_pauli_term* terms, size_t nterms) {
    double re, im;
    or_pauli_expectation(&st, terms, nterms, &re, &im);
    return re;
}
static bool single_qubit_obs(const or_pauli_term* terms, size_t nterms, int* obs, uint32_t* q) {
    if (nterms != 1 || terms[0].re != 1.0 || terms[0].im != 0.0) return false;
    uint64_t x = terms[0].x_mask, z = terms[0].z_mask, all = x | z;
    if (all == 0 || (all & (all - 1))) return false;
    *q = uint32_t(__builtin_ctzll(all));
    *obs = (x && z) ? 1 : (x ? 0 : 2);
    return true;
}
static double observe(const State& st, const or_pauli_term* terms, size_t nterms) {
    int obs;
    uint32_t q;
    if (single_qubit_obs(terms, nterms, &obs, &q)) {  // State::expectation, src/lib.rs:1306-1318
        double v;
        or_expectation(&st, obs, q, &v);
        return v;
    }
    return energy(st, terms, nterms);
}
static int check_obs(const or_pauli_term* terms, size_t nterms, uint32_t qubits) {
    for (size_t t = 0; t < nterms; ++t)
        if (qubits < 64 && ((terms[t].x_mask | terms[t].z_mask) >> qubits)) return QUBIT_OUT_OF_RANGE;
    return OK;
}

int or_pc_run(const or_pop* ops, size_t nops, uint32_t nparams, uint32_t ninputs, const double* params,
              size_t np, const double* inputs, size_t ni, uint32_t qubits, int64_t shift_op, double delta,
              double* out_amps) {
    PCirc pc{ops, nops, nparams, ninputs};
    int e = validate(pc, params, np, inputs, ni);
    if (e) return e;
    if (qubits == 0) return ZERO_QUBITS;
    State* st = static_cast<State*>(or_state_zero(qubits));
    e = run_with_shift(pc, params, inputs, *st, shift_op, delta);
    if (!e && out_amps) std::memcpy(out_amps, st->amps.data(), st->amps.size() * sizeof(C));
    or_state_free(st);
    return e;
}

int or_pc_expectation(const or_pop* ops, size_t nops, uint32_t nparams, uint32_t ninputs, const double* params,
                      const double* inputs, uint32_t qubits, const or_pauli_term* terms, size_t nterms,
                      double* out) {
    PCirc pc{ops, nops, nparams, ninputs};
    int e = validate(pc, params, nparams, inputs, ninputs);
    if (e) return e;
    if ((e = check_obs(terms, nterms, qubits))) return e;
    State* st = static_cast<State*>(or_state_zero(qubits));
    e = run_with_shift(pc, params, inputs, *st, -1, 0.0);
    if (!e) *out = observe(*st, terms, nterms);
    or_state_free(st);
    return e;
}

// parameter_shift_gradient_with_inputs, src/lib.rs:3516-3546 (shift ±pi/2 per
// rotation occurrence), and finite_difference_gradient_with_inputs,
// src/gradients.rs:344-375 (step h, divided by 2h).  fd_step <= 0 selects the shift rule.
static int shift_like(const or_pop* ops, size_t nops, uint32_t nparams, uint32_t ninputs, const double* params,
                      const double* inputs, uint32_t qubits, const or_pauli_term* terms, size_t nterms,
                      bool fd, double step, double* grad) {
    PCirc pc{ops, nops, nparams, ninputs};
    if (fd && (!std::isfinite(step) || step <= 0.0)) return INVALID_FD_STEP;
    int e = validate(pc, params, nparams, inputs, ninputs);
    if (e) return e;
    if ((e = check_obs(terms, nterms, qubits))) return e;
    for (uint32_t p = 0; p < nparams; ++p) grad[p] = 0.0;
    const double shift = fd ? step : 1.57079632679489661923;  // FRAC_PI_2
    for (size_t i = 0; i < nops; ++i) {
        if (ops[i].kind != 1) continue;
        State* a = static_cast<State*>(or_state_zero(qubits));
        State* b = static_cast<State*>(or_state_zero(qubits));
        e = run_with_shift(pc, params, inputs, *a, int64_t(i), shift);
        if (!e) e = run_with_shift(pc, params, inputs, *b, int64_t(i), -shift);
        if (!e) {
            double plus = observe(*a, terms, nterms), minus = observe(*b, terms, nterms);
            if (fd)
                grad[ops[i].index] += (plus - minus) / (2.0 * step);
            else
                grad[ops[i].index] += 0.5 * (plus - minus);
        }
        or_state_free(a);
        or_state_free(b);
        if (e) return e;
    }
    return OK;
}
int or_pc_shift_gradient(const or_pop* ops, size_t nops, uint32_t nparams, uint32_t ninputs, const double* params,
                         const double* inputs, uint32_t qubits, const or_pauli_term* terms, size_t nterms,
                         double* grad) {
    return shift_like(ops, nops, nparams, ninputs, params, inputs, qubits, terms, nterms, false, 0.0, grad);
}
int or_pc_fd_gradient(const or_pop* ops, size_t nops, uint32_t nparams, uint32_t ninputs, const double* params,
                      const double* inputs, uint32_t qubits, const or_pauli_term* terms, size_t nterms, double step,
                      double* grad) {
    return shift_like(ops, nops, nparams, ninputs, params, inputs, qubits, terms, nterms, true, step, grad);
}

// adjoint_gradient_with_inputs, src/lib.rs:3564-3704: forward sweep storing every
// state, costate = O psi_final, reverse sweep accumulating
// 2 Re sum conj(lambda) * ((P psi_after) * (-i/2)) then lambda <- U^dagger lambda.
// With a Pauli-sum observable the costate is H psi_final (the sum's sparse product).
int or_pc_adjoint_gradient(const or_pop* ops, size_t nops, uint32_t nparams, uint32_t ninputs,
                           const double* params, const double* inputs, uint32_t qubits,
                           const or_pauli_term* terms, size_t nterms, double* grad) {
    PCirc pc{ops, nops, nparams, ninputs};
    int e = validate(pc, params, nparams, inputs, ninputs);
    if (e) return e;
    if ((e = check_obs(terms, nterms, qubits))) return e;
    std::vector<State> states;
    states.reserve(nops + 1);
    State* z = static_cast<State*>(or_state_zero(qubits));
    states.push_back(*z);
    or_state_free(z);
    for (size_t i = 0; i < nops; ++i) {
        State next = states[i];
        const or_pop& op = ops[i];
        switch (op.kind) {
            case 0: e = apply_1q(next, op.gate, op.theta, op.a); break;
            case 1:
                if (op.index >= nparams) return INVALID_PARAMETER_COUNT;
                e = apply_1q(next, axis_gate(op.gate), params[op.index], op.a);
                break;
            case 2:
                if (op.index >= ninputs) return INVALID_INPUT_COUNT;
                e = apply_1q(next, axis_gate(op.gate), inputs[op.index], op.a);
                break;
            case 3: e = apply_cnot(next, op.a, op.b); break;
            case 4: e = apply_cz(next, op.a, op.b); break;
            case 5: e = apply_swap(next, op.a, op.b); break;
            case 6: return REQUIRES_DENSITY_MATRIX;
            default: return INVALID_ARGUMENT;
        }
        if (e) return e;
        states.push_back(std::move(next));
    }
    State costate = states.back();
    int obs;
    uint32_t oq;
    if (single_qubit_obs(terms, nterms, &obs, &oq)) {
        e = apply_1q(costate, obs == 0 ? G_X : (obs == 1 ? G_Y : G_Z), 0.0, oq);
        if (e) return e;
    } else {
        pauli_apply(states.back(), terms, nterms, costate.amps);
    }
    for (uint32_t p = 0; p < nparams; ++p) grad[p] = 0.0;
    const C minus_i_half{0.0, -0.5};
    for (size_t ii = nops; ii-- > 0;) {
        const or_pop& op = ops[ii];
        const State& after = states[ii + 1];
        if (op.kind == 1) {
            State deriv = after;
            e = apply_1q(deriv, op.gate == 0 ? G_X : (op.gate == 1 ? G_Y : G_Z), 0.0, op.a);
            if (e) return e;
            double overlap = 0.0;
            for (size_t k = 0; k < deriv.amps.size(); ++k)
                overlap += cmul(cconj(costate.amps[k]), cmul(deriv.amps[k], minus_i_half)).re;
            grad[op.index] += 2.0 * overlap;
        }
        switch (op.kind) {
            case 0: {
                int g;
                double th;
                adjoint_gate(op.gate, op.theta, &g, &th);
                e = apply_1q(costate, g, th, op.a);
                break;
            }
            case 1: e = apply_1q(costate, axis_gate(op.gate), -params[op.index], op.a); break;
            case 2: e = apply_1q(costate, axis_gate(op.gate), -inputs[op.index], op.a); break;
            case 3: e = apply_cnot(costate, op.a, op.b); break;
            case 4: e = apply_cz(costate, op.a, op.b); break;
            case 5: e = apply_swap(costate, op.a, op.b); break;
            default: break;
        }
        if (e) return e;
    }
    return OK;
}

// ---- in-place CPU variant (reported separately from the restatement in the bench):
// same arithmetic per pair, no clone, OpenMP over pairs and over permutation quads.
int or_fast_run(double* amps_d, uint32_t qubits, const or_op* ops, size_t nops) {
    C* amps = reinterpret_cast<C*>(amps_d);
    const size_t half = size_t(1) << (qubits - 1), quarter = size_t(1) << (qubits >= 2 ? qubits - 2 : 0);
    for (size_t i = 0; i < nops; ++i) {
        const or_op& op = ops[i];
        if (op.kind == 0) {
            if (op.a >= qubits) return QUBIT_OUT_OF_RANGE;
            C m[4];
            gate_matrix(op.gate, op.theta, m);
            const size_t mask = size_t(1) << op.a, lo = mask - 1;
#pragma omp parallel for schedule(static)
            for (int64_t sp = 0; sp < int64_t(half); ++sp) {
                size_t p = size_t(sp);
                size_t i0 = ((p & ~lo) << 1) | (p & lo), i1 = i0 | mask;
                C a0 = amps[i0], a1 = amps[i1];
                amps[i0] = cadd(cmul(m[0], a0), cmul(m[1], a1));
                amps[i1] = cadd(cmul(m[2], a0), cmul(m[3], a1));
            }
        } else if (op.kind >= 1 && op.kind <= 3) {
            if (op.a >= qubits || op.b >= qubits) return QUBIT_OUT_OF_RANGE;
            if (op.a == op.b) return INVALID_QUBIT_PAIR;
            const uint32_t lo_q = std::min(op.a, op.b), hi_q = std::max(op.a, op.b);
            const size_t ma = size_t(1) << op.a, mb = size_t(1) << op.b;
#pragma omp parallel for schedule(static)
            for (int64_t sp = 0; sp < int64_t(quarter); ++sp) {
                size_t p = size_t(sp);
                size_t lm = (size_t(1) << lo_q) - 1;
                size_t x = ((p & ~lm) << 1) | (p & lm);
                size_t hm = (size_t(1) << hi_q) - 1;
                size_t base = ((x & ~hm) << 1) | (x & hm);
                if (op.kind == 1) std::swap(amps[base | ma], amps[base | ma | mb]);
                else if (op.kind == 2) amps[base | ma | mb] = cmulf(amps[base | ma | mb], -1.0);
                else std::swap(amps[base | ma], amps[base | mb]);
            }
        } else if (op.kind == 4) {
            return REQUIRES_DENSITY_MATRIX;
        } else {
            return REQUIRES_SHOT_EXECUTION;
        }
    }
    return OK;
}

}  // extern "C"

// ---- Lanczos (sparse_lanczos_gpu, src/quantum.rs:1853-2023; DMRG::run_sparse_lanczos,
// src/tensor_network.rs:1535-1689, whose GPU form run_sparse_lanczos_gpu,
// src/tensor_network.rs:1692-2025, runs the same iteration)
namespace {

// Start vector of both: seed = seed * 6364136223846793005 + 1 (wrapping, seed
// 42 first), x = seed / u64::MAX - 0.5, then divided by its 2-norm.
void lanczos_start(std::vector<double>& v) {
    uint64_t seed = 42;
    for (double& x : v) {
        seed = seed * 6364136223846793005ull + 1ull;
        x = double(seed) / double(UINT64_MAX) - 0.5;
    }
    double acc = 0.0;
    for (double x : v) acc += x * x;
    const double norm = std::sqrt(acc);
    for (double& x : v) x /= norm;
}

double dot(const std::vector<double>& a, const std::vector<double>& b) {
    double acc = 0.0;
    for (size_t i = 0; i < a.size(); ++i) acc += a[i] * b[i];
    return acc;
}

// tfim_matvec of src/tensor_network.rs:1702-1718 (H = -sum Z_i Z_{i+1} - sum
// X_i on an open chain, site i = bit n-1-i of the index).
void tfim_apply(uint32_t n, const std::vector<double>& v, std::vector<double>& w) {
    const size_t dim = v.size();
    for (size_t a = 0; a < dim; ++a) {
        double r = 0.0;
        for (uint32_t i = 0; i + 1 < n; ++i) {
            const double zi = ((a >> (n - 1 - i)) & 1) ? -1.0 : 1.0;
            const double zj = ((a >> (n - 2 - i)) & 1) ? -1.0 : 1.0;
            r += -zi * zj * v[a];
        }
        for (uint32_t i = 0; i < n; ++i) r += -v[a ^ (size_t(1) << (n - 1 - i))];
        w[a] = r;
    }
}

}  // namespace

extern "C" {

// The Lanczos start vector (for pinning the LCG against exact integer arithmetic).
void or_lanczos_start(double* out, uint64_t dim) {
    std::vector<double> v(dim);
    lanczos_start(v);
    std::memcpy(out, v.data(), dim * sizeof(double));
}

// sparse_lanczos_gpu on a CSR operator with real values (the reference keeps
// the real part of each entry): m = min(max_iterations, dim, 100); per step
// w = A v, w -= beta_{j-1} v_prev, modified Gram-Schmidt against v_0..v_{j-1},
// alpha_j = v_j . w, w -= alpha_j v_j, beta_j = |w|; stop when beta_j <
// tolerance or after m steps.  Writes alpha[0..m), beta[0..m) and the size of
// the tridiagonal model (the first index with alpha == 0.0, else m).
int or_sparse_lanczos(const uint64_t* rows, const uint64_t* cols, const double* vals, uint64_t dim,
                      uint64_t max_iterations, double tolerance, double* alpha, double* beta, uint64_t* size) {
    const size_t m = std::min<uint64_t>(std::min<uint64_t>(max_iterations, dim), 100);
    std::vector<double> v(dim), vprev(dim, 0.0), w(dim);
    lanczos_start(v);
    std::vector<std::vector<double>> vall{v};
    std::fill(alpha, alpha + m, 0.0);
    std::fill(beta, beta + m, 0.0);
    for (size_t j = 0; j < m; ++j) {
        for (size_t i = 0; i < dim; ++i) {
            double s = 0.0;
            for (uint64_t k = rows[i]; k < rows[i + 1]; ++k) s += vals[k] * v[cols[k]];
            w[i] = s;
        }
        if (j > 0)
            for (size_t i = 0; i < dim; ++i) w[i] -= beta[j - 1] * vprev[i];
        for (size_t k = 0; k < j; ++k) {
            const double p = dot(vall[k], w);
            for (size_t i = 0; i < dim; ++i) w[i] -= p * vall[k][i];
        }
        const double a = dot(v, w);
        for (size_t i = 0; i < dim; ++i) w[i] -= a * v[i];
        const double b = std::sqrt(dot(w, w));
        alpha[j] = a;
        if (j + 1 < m) beta[j] = b;
        if (b < tolerance || j + 1 == m) break;
        vprev = v;
        for (size_t i = 0; i < dim; ++i) v[i] = w[i] / b;
        vall.push_back(v);
    }
    size_t actual = m;
    for (size_t j = 0; j < m; ++j)
        if (alpha[j] == 0.0) {
            actual = j;
            break;
        }
    *size = actual;
    return OK;
}

// DMRG::run_sparse_lanczos on the TFIM chain of n sites: m = min(n >= 24 ? 30
// : 100, 2^n); per step w = H v, alpha_j = v . w, w -= alpha_j v + beta_{j-1}
// v_prev, modified Gram-Schmidt against every stored vector (the current one
// included), beta_j = |w|, stop when beta_j < 1e-12.  The tridiagonal model
// always has m rows (zero beyond the last step).  Writes alpha, beta (m
// entries) and m.
int or_tfim_lanczos(uint32_t n, double* alpha, double* beta, uint64_t* size) {
    if (n < 2 || n > 26) return INVALID_ARGUMENT;
    const size_t dim = size_t(1) << n;
    const size_t m = std::min<size_t>(n >= 24 ? 30 : 100, dim);
    std::vector<double> v(dim), vprev(dim, 0.0), w(dim);
    lanczos_start(v);
    std::vector<std::vector<double>> vall{v};
    std::fill(alpha, alpha + m, 0.0);
    std::fill(beta, beta + m, 0.0);
    for (size_t j = 0; j < m; ++j) {
        tfim_apply(n, v, w);
        alpha[j] = dot(v, w);
        for (size_t i = 0; i < dim; ++i) {
            w[i] -= alpha[j] * v[i];
            if (j > 0) w[i] -= beta[j - 1] * vprev[i];
        }
        for (const auto& u : vall) {
            const double p = dot(w, u);
            for (size_t i = 0; i < dim; ++i) w[i] -= p * u[i];
        }
        beta[j] = std::sqrt(dot(w, w));
        if (beta[j] < 1e-12) break;
        vprev = v;
        for (size_t i = 0; i < dim; ++i) v[i] = w[i] / beta[j];
        vall.push_back(v);
    }
    *size = m;
    return OK;
}

}  // extern "C"
