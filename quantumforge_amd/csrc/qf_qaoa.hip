// QAOA-form instances of an uneven batch on a tile engine of their own.
//
// Circuits in QAOA form (qf_qaoa.h) differ from one another only in their
// graph (which qubit pairs the cost rotations join), their angles and their
// width, so one set of kernels serves every instance with no compile per
// graph.  A tile holds 2^12 amplitudes of one instance: index bits 0..2 (the
// contiguous 128-byte runs every load and store moves) and nine more qubits
// chosen per pass; the other n - 12 bits index the tiles.  512 threads hold 8
// amplitudes each in registers.  In phase g the registers hold tile bits
// 3g..3g+2 and the thread index the other nine (ascending), so the X rotations
// of one tile bit group are register butterflies; phases change through a
// 64 KiB LDS tile (XOR-swizzled 16-byte slots).
//
// A pass applies, in order: the rest of a layer's X rotations (sweep A:
// phases 0, 1, 2, 3), one cost layer D (phase 3 after sweep A, else 0), and
// the next layer's X rotations on the tile's qubits (sweep B: phases 3..0
// after sweep A, else 0..3 and back to 0).  The cost layer is the diagonal
// exp(-i sum_e theta_e/2 z_u z_v) with the instance's own edge table: per
// thread, the edges with no register endpoint multiply into one factor, the
// few edges with one register endpoint into a factor per register
// bit, and edges inside the registers come from a host table of 8 phases; so
// the cost layer rides along in a pass that also moves the mixer.  With the
// tile's qubit sets cycling A, B (two passes per layer pair, n <= 21) or A, B,
// C (a cost layer every second pass, n <= 30), every pass carries a whole
// layer's worth of rotations.  The last forward pass also writes the costate
// lambda = O psi and the energy partials (the observable's Z and Z Z words,
// evaluated like the cost layer).
//
// The reverse sweep runs the same passes backwards on psi and lambda: before
// undoing an X rotation, each thread adds Im(conj(lambda) X psi) over its
// registers, and four such shares are summed over the wave at once (transpose
// steps, then lane exchanges); a cost layer's edge overlaps sum_x z_u z_v
// Im(conj(lambda_x) psi_x) come from one Walsh transform over registers and
// lanes of the thread's Im(conj(lambda) psi) (the sign of an edge over a wave
// is a Walsh coefficient, read from one lane), so the per-edge cost is scalar.
// Per-tile sums go to a [slot][tile] partial table; a finishing kernel adds
// every parameter's occurrences in a fixed order (Neumaier), and the energy
// partials.
//
// Each memory wave of instances is one captured hipGraph of launches (forward
// pass j of every instance, reverse pass j, finish) over device tables; the
// graph is replayed per wave with new tables, its grids sized for the largest
// wave.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "qf_qaoa.h"

using qf::cx;
using qf::dd;

namespace qfq {
namespace {

constexpr int kQM = 12;                // tile qubits
constexpr int kQT = 512;               // threads per workgroup, 8 amplitudes each
constexpr int kQW = kQT / 64;          // waves per workgroup
constexpr uint32_t kMaxTerms = 120;    // diagonal terms of one pass
constexpr uint32_t kRows = 24 + kMaxTerms;

enum : uint32_t { F_LOAD = 1, F_SWA = 2, F_DIAG = 4, F_SWB = 8, F_COST = 16, F_LAMBDA = 32 };

// One term of a pass's diagonal (a cost edge or an observable word),
// classified for the phase the diagonal is applied in.  Its sign at index x
// is (-1)^(popcount(x & gm) + parity of the register bits in regmask).
struct QTerm {
    uint32_t gm;        // index bits outside the registers
    uint32_t lanemask;  // records: lane bits among the term's qubits
    uint32_t wbmask;    // records: index bits of the term's wave and block qubits
    uint32_t slot;      // angle slot (cost layer) or coefficient index (observable)
};

// Terms of a pass, sorted by register class: counts of regmask 0, 1, 2, 4 (one
// register bit), 3, 5, 6 (two register bits).
constexpr int kClasses = 7;
__host__ __device__ constexpr uint32_t class_mask(int c) {
    return c == 0 ? 0u : c == 1 ? 1u : c == 2 ? 2u : c == 3 ? 4u : c == 4 ? 3u : c == 5 ? 5u : 6u;
}

struct QDesc {
    uint64_t state_off;   // amplitude offset of the instance in the wave's psi / lambda
    uint64_t part_off;    // gradient partials (doubles), [slot][tile]
    uint64_t epart_off;   // energy partials (dd), [tile]
    uint32_t n, logb, flags, nslots1;
    uint32_t cs_off, coef_off, t_off, pad0;
    uint32_t cnt[kClasses];  // cost layer terms by register class (phase 3 after sweep A, else 0)
    uint32_t t_n;
    uint32_t o_off;
    uint32_t ocnt[4];     // observable terms of register class 0, 1, 2, 4 (phase 0)
    uint32_t hi[9];       // tile bits 3..11 -> index bits
    uint32_t nout;
    uint32_t outpos[20];  // tile index bits -> index bits outside the tile
    uint32_t slotA[12], slotB[12];
    double w2[16];        // cost layer: the phases of its register-pair edges, per register
    double ow2[8];        // observable: its register-pair words' sum, per register
    double c0;            // observable constant
};

struct QHeader {
    uint64_t total_blocks;
    uint32_t nentries, pad;
};
struct QEntry {
    uint64_t block_start;
    uint32_t desc, pad;
};

// Finishing work: one parameter (its occurrence slots, CSR) or, with
// nslots == ~0u, one instance's energy.
struct QFin {
    uint64_t part_off;
    uint64_t out;         // gradient index, or energy index
    uint32_t nblk, first, count, is_energy;
};

// ---------------------------------------------------------------------------
// Lane exchanges of doubles
// ---------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, int(b), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, int(b >> 32), CTRL, 0xf, 0xf, false);
    return __builtin_bit_cast(double, (long long)(unsigned)lo | ((long long)hi << 32));
}
template <int PAT>
__device__ __forceinline__ double swz_f64(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_ds_swizzle(int(b), PAT);
    const int hi = __builtin_amdgcn_ds_swizzle(int(b >> 32), PAT);
    return __builtin_bit_cast(double, (long long)(unsigned)lo | ((long long)hi << 32));
}
__device__ __forceinline__ double bperm_f64(double v, unsigned lane) {
    const long long b = __builtin_bit_cast(long long, v);
    const int a = int(lane << 2);
    const int lo = __builtin_amdgcn_ds_bpermute(a, int(b));
    const int hi = __builtin_amdgcn_ds_bpermute(a, int(b >> 32));
    return __builtin_bit_cast(double, (long long)(unsigned)lo | ((long long)hi << 32));
}
// The value of lane ln ^ 2^s.
template <int S>
__device__ __forceinline__ double lane_xor(double v, unsigned ln) {
    if constexpr (S == 0) return dpp_f64<0xB1>(v);
    else if constexpr (S == 1) return dpp_f64<0x4E>(v);
    else if constexpr (S < 5) return swz_f64<((1 << S) << 10) | 0x1F>(v);
    else return bperm_f64(v, ln ^ 32u);
}
__device__ __forceinline__ double readlane_f64(double v, uint32_t lane) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane(int(b), lane);
    const int hi = __builtin_amdgcn_readlane(int(b >> 32), lane);
    return __builtin_bit_cast(double, (long long)(unsigned)lo | ((long long)hi << 32));
}

// Walsh transform over the wave's lanes: lane m ends with
// sum_l (-1)^popcount(l & m) v_l.
template <int S = 0>
__device__ __forceinline__ double lane_walsh(double v, unsigned ln) {
    if constexpr (S == 6) {
        return v;
    } else {
        const double p = lane_xor<S>(v, ln);
        return lane_walsh<S + 1>(((ln >> S) & 1u) ? p - v : v + p, ln);
    }
}

// Wave sums of four values: transpose steps on lane bits 0 and 1 leave lane l
// with value l & 3 summed over its quad, lane exchanges across bits 2..5
// complete it; every lane of a class holds its value's sum.
__device__ __forceinline__ double wave_sum4(double a, double b, double c, double d, unsigned ln) {
    const bool h0 = ln & 1u, h1 = (ln >> 1) & 1u;
    const double x0 = (h0 ? b : a) + dpp_f64<0xB1>(h0 ? a : b);
    const double x1 = (h0 ? d : c) + dpp_f64<0xB1>(h0 ? c : d);
    double v = (h1 ? x1 : x0) + dpp_f64<0x4E>(h1 ? x0 : x1);
    v += swz_f64<(4 << 10) | 0x1F>(v);
    v += swz_f64<(8 << 10) | 0x1F>(v);
    v += swz_f64<(16 << 10) | 0x1F>(v);
    v += bperm_f64(v, ln ^ 32u);
    return v;
}

template <class R>
struct Regs {
    R re[8], im[8];
};

// 16-byte (complex128) or 8-byte slot of tile index x in the LDS tile.
__device__ __forceinline__ uint32_t swz(uint32_t x) { return x ^ ((x >> 3) & 15u); }
// Tile index of thread t's register 0 in phase G.
template <int G>
__device__ __forceinline__ uint32_t tile_base(uint32_t t) {
    constexpr uint32_t lo = (1u << (3 * G)) - 1u;
    return (t & lo) | ((t >> (3 * G)) << (3 * G + 3));
}

// Moves the registers of every state from phase G0 to phase G1 through the
// LDS tile (one state at a time).
template <class R, int NS, int G0, int G1>
__device__ __forceinline__ void exchange(Regs<R> (&s)[NS], cx<R>* xb, uint32_t t) {
    uint32_t bw = tile_base<G0>(t), br = tile_base<G1>(t);
    // Opaque per exchange, so that the compiler does not keep every
    // exchange's 16 slot addresses live from the kernel's start.
    asm volatile("" : "+v"(bw), "+v"(br));
#pragma unroll
    for (int st = 0; st < NS; ++st) {
#pragma unroll
        for (int k = 0; k < 8; ++k) xb[swz(bw | (uint32_t(k) << (3 * G0)))] = cx<R>{s[st].re[k], s[st].im[k]};
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const cx<R> v = xb[swz(br | (uint32_t(k) << (3 * G1)))];
            s[st].re[k] = v.re;
            s[st].im[k] = v.im;
        }
        __syncthreads();
    }
}

// Stores the thread's 8 amplitudes (phase 0: tile bits 0..2 in registers, one
// contiguous run) so that every store instruction writes whole runs: the
// wave's threads pass their registers through the wave's own 512 slots of
// the LDS tile, and lane l then writes amplitude l % 8 of the wave's thread
// 8k + l / 8 (k = 0..7).  Stored straight from the registers, each
// instruction writes 16 bytes into each of 64 runs, and the L2 wrote the
// lines of the two-state reverse passes back 1.4 to 1.9 times over.  The LDS tile
// must be free (every exchange ends with a barrier).
template <class R>
__device__ __forceinline__ void store_runs(cx<R>* __restrict__ dst, const Regs<R>& s, uint32_t g0, cx<R>* xb,
                                           uint32_t t) {
    const uint32_t ln = t & 63u, e = ln & 7u, hi = ln >> 3;
    cx<R>* wb = xb + (t & ~63u) * 8u;
#pragma unroll
    for (int k = 0; k < 8; ++k) wb[ln * 8u + (uint32_t(k) ^ e)] = cx<R>{s.re[k], s.im[k]};
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t sl = 8u * uint32_t(k) + hi;
        const cx<R> v = wb[sl * 8u + (e ^ hi)];
        const uint32_t ga = uint32_t(__shfl(int(g0), int(sl), 64));
        dst[ga + e] = v;
    }
    __syncthreads();
}

// X rotation exp(-i theta/2 X) (s = sin(theta/2); the reverse sweep passes
// -s) on register bit J of every state.
template <class R, int NS, int J>
__device__ __forceinline__ void rx_reg(Regs<R> (&s)[NS], double cd, double sd) {
    const R c = R(cd), sn = R(sd);
#pragma unroll
    for (int st = 0; st < NS; ++st)
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (k & (1 << J)) continue;
            const int k1 = k | (1 << J);
            const R a0r = s[st].re[k], a0i = s[st].im[k], a1r = s[st].re[k1], a1i = s[st].im[k1];
            s[st].re[k] = c * a0r + sn * a1i;
            s[st].im[k] = c * a0i - sn * a1r;
            s[st].re[k1] = c * a1r + sn * a0i;
            s[st].im[k1] = c * a1i - sn * a0r;
        }
}

// Im <lambda| X_J |psi> over the thread's registers.
template <class R, int J>
__device__ __forceinline__ double rx_share(const Regs<R>& p, const Regs<R>& l) {
    double q = 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int kp = k ^ (1 << J);
        q += double(l.re[k]) * double(p.im[kp]) - double(l.im[k]) * double(p.re[kp]);
    }
    return q;
}

// The X rotations of tile bit group G (register bits 0..2) with the slots
// slots[3G..3G+2].  Reverse (NS == 2): each rotation's overlap is recorded
// first (rows row0 + j of the shared table), then the rotation is undone.
template <class R, int NS, int G>
__device__ __forceinline__ void rx_group(Regs<R> (&s)[NS], const double2* rxc, double* qred, uint32_t row0,
                                         uint32_t t) {
    const double2 a0 = rxc[row0 + 3 * G], a1 = rxc[row0 + 3 * G + 1], a2 = rxc[row0 + 3 * G + 2];
    double c0 = a0.x, n0 = a0.y, c1 = a1.x, n1 = a1.y, c2 = a2.x, n2 = a2.y;
    if constexpr (NS == 2) {
        const unsigned ln = t & 63u;
        // X on different bits commute with one another's rotations: the three
        // shares are taken before any of the three is undone.
        const double q = wave_sum4(rx_share<R, 0>(s[0], s[1]), rx_share<R, 1>(s[0], s[1]),
                                   rx_share<R, 2>(s[0], s[1]), 0.0, ln);
        if (ln < 3u) qred[(row0 + 3 * G + ln) * kQW + (t >> 6)] = q;
        n0 = -n0;
        n1 = -n1;
        n2 = -n2;
    }
    rx_reg<R, NS, 0>(s, c0, n0);
    rx_reg<R, NS, 1>(s, c1, n1);
    rx_reg<R, NS, 2>(s, c2, n2);
}

// Index bits of thread t's register 0 in phase P (0 or 3), without the tile's
// own block bits (bb).
template <int P>
__device__ __forceinline__ uint32_t thread_bits(const QDesc& d, uint32_t t) {
    uint32_t g = 0;
    if constexpr (P == 0) {
#pragma unroll
        for (int i = 0; i < 9; ++i) g |= ((t >> i) & 1u) << d.hi[i];
    } else {
        g = t & 7u;
#pragma unroll
        for (int i = 0; i < 6; ++i) g |= ((t >> (3 + i)) & 1u) << d.hi[i];
    }
    return g;
}

__device__ __forceinline__ void cmul(double& ar, double& ai, double br, double bi) {
    const double r = ar * br - ai * bi;
    ai = ar * bi + ai * br;
    ar = r;
}

// The cost layer of the pass in phase P: forward, psi_r *= f_r; reverse
// (NS == 2), the edge overlaps are recorded and both states take conj(f_r).
template <class R, int NS, int P>
__device__ __forceinline__ void diag_layer(Regs<R> (&s)[NS], const QDesc& d, const QTerm* __restrict__ T,
                                           const double* __restrict__ cs, double* qred,
                                           uint32_t bb, uint32_t t) {
    const uint32_t g = bb | thread_bits<P>(d, t);
    const QTerm* tp = T + d.t_off;
    // Edges with no register endpoint: one factor per thread, the product of
    // their c - i z_u z_v s (products keep the rounding of each factor's
    // cos and sin, as the reference's one Rz per edge does).
    double ar0 = 1.0, ai0 = 0.0;
    for (uint32_t e = 0; e < d.cnt[0]; ++e) {
        const QTerm q = tp[e];
        const double c = cs[2 * q.slot], sn = cs[2 * q.slot + 1];
        cmul(ar0, ai0, c, (__popc(g & q.gm) & 1) ? sn : -sn);
    }
    // Edges with one register endpoint j: a factor per register bit value
    // (bit 0: P_j, bit 1: conj(P_j)).
    double pr[3] = {1.0, 1.0, 1.0}, pi[3] = {0.0, 0.0, 0.0};
    uint32_t e0 = d.cnt[0];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        for (uint32_t e = e0; e < e0 + d.cnt[1 + j]; ++e) {
            const QTerm q = tp[e];
            const double c = cs[2 * q.slot], sn = cs[2 * q.slot + 1];
            cmul(pr[j], pi[j], c, (__popc(g & q.gm) & 1) ? sn : -sn);
        }
        e0 += d.cnt[1 + j];
    }
    // f_r = F P_0^(r0) P_1^(r1) P_2^(r2) W_r, where P^(1) = conj(P): the bit
    // 0 and 1 products first, then one register at a time.
    double br[4], bi[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        br[k] = ar0;
        bi[k] = ai0;
        cmul(br[k], bi[k], pr[0], (k & 1) ? -pi[0] : pi[0]);
        cmul(br[k], bi[k], pr[1], (k & 2) ? -pi[1] : pi[1]);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        double fr = br[k & 3], fi = bi[k & 3];
        cmul(fr, fi, pr[2], (k & 4) ? -pi[2] : pi[2]);
        cmul(fr, fi, d.w2[2 * k], d.w2[2 * k + 1]);
#pragma unroll
        for (int st = 0; st < NS; ++st) {
            double ar = double(s[st].re[k]), ai = double(s[st].im[k]);
            cmul(ar, ai, fr, NS == 2 ? -fi : fi);
            s[st].re[k] = R(ar);
            s[st].im[k] = R(ai);
        }
    }
    if constexpr (NS == 2) {
        // Records: q_r = Im(conj(lambda_r) psi_r) (the diagonal leaves it
        // unchanged, so it is taken after the states are updated).
        const unsigned ln = t & 63u;
        double q[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            q[k] = double(s[1].re[k]) * double(s[0].im[k]) - double(s[1].im[k]) * double(s[0].re[k]);
        // Walsh transform over the registers: Q[m] = sum_r (-1)^popcount(r & m) q_r.
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (!(k & (1 << b))) {
                    const double x = q[k], y = q[k | (1 << b)];
                    q[k] = x + y;
                    q[k | (1 << b)] = x - y;
                }
        // The wave's index bits outside the lanes and registers (lane 0 of the
        // wave has lane bits 0).
        const uint32_t gw = __builtin_amdgcn_readfirstlane(g);
        const uint32_t wv = t >> 6;
        uint32_t e = 0;
#pragma unroll
        for (int c = 0; c < kClasses; ++c) {
            const uint32_t m = class_mask(c);
            const uint32_t ne = d.cnt[c];
            if (ne == 0) continue;
            __builtin_amdgcn_sched_barrier(0);
            const double W = lane_walsh(q[m], ln);
            for (uint32_t i = 0; i < ne; ++i, ++e) {
                const QTerm qt = tp[e];
                double v = readlane_f64(W, qt.lanemask);
                if (__popc(gw & qt.wbmask) & 1) v = -v;
                if (ln == 0) qred[(24 + e) * kQW + wv] = v;
            }
        }
    }
}

// lambda = O psi and the thread's share of <psi|O|psi>, in phase 0.
template <class R>
__device__ __forceinline__ double costate(Regs<R>& p, Regs<R>& l, const QDesc& d, const QTerm* __restrict__ T,
                                          const double* __restrict__ coef, uint32_t g) {
    const QTerm* tp = T + d.o_off;
    double B[4] = {d.c0, 0.0, 0.0, 0.0};
    uint32_t e = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        for (uint32_t i = 0; i < d.ocnt[c]; ++i, ++e) {
            const QTerm q = tp[e];
            const double v = coef[q.slot];
            B[c] += (__popc(g & q.gm) & 1) ? -v : v;
        }
    }
    double en = 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const double Ck = B[0] + ((k & 1) ? -B[1] : B[1]) + ((k & 2) ? -B[2] : B[2]) + ((k & 4) ? -B[3] : B[3]) +
                          d.ow2[k];
        const double pr = double(p.re[k]), pi = double(p.im[k]);
        en += Ck * (pr * pr + pi * pi);
        l.re[k] = R(Ck * pr);
        l.im[k] = R(Ck * pi);
    }
    return en;
}

template <class E>
__device__ __forceinline__ uint32_t find_entry(const E* __restrict__ e, uint32_t n, uint64_t b) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (e[mid].block_start <= b)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Pass shapes: the schedule makes every pass one of three (the cost layer
// after sweep A in phase 3, or in phase 0 without it).
enum Shape : int { SH_DB = 0, SH_ADB = 1, SH_B = 2 };
constexpr int kShapes = 3;

// One pass of shape SH of every instance in the launch table: forward
// (NS == 1; COST: the instance's last pass, with the costate and the energy)
// or its reverse (NS == 2: psi and lambda, overlap records).  Each shape is
// its own straight-line kernel; the reverse form is held to 128 VGPRs, so two
// workgroups (16 waves) share a CU.
template <class R, int NS, int SH, bool COST>
__global__ __launch_bounds__(kQT, NS == 2 ? 4 : 2) void k_qaoa_pass(cx<R>* __restrict__ psi, cx<R>* __restrict__ lam,
                                                      const unsigned char* __restrict__ launch,
                                                      const QDesc* __restrict__ descs, const QTerm* __restrict__ terms,
                                                      const double* __restrict__ cs,
                                                      const double* __restrict__ coef, double* __restrict__ partial,
                                                      dd* __restrict__ epart) {
    constexpr bool A = SH == SH_ADB, D = SH != SH_B;
    __shared__ cx<R> xb[1 << kQM];
    __shared__ double qred[NS == 2 ? kRows * kQW : 1];
    __shared__ dd ered[kQW];
    // (cos, sin) of the pass's 24 X rotations (sweep A tile bits, then sweep B).
    __shared__ double2 rxc[24];
    const QHeader* h = reinterpret_cast<const QHeader*>(launch);
    const QEntry* E = reinterpret_cast<const QEntry*>(launch + sizeof(QHeader));
    const uint64_t total = h->total_blocks;
    const uint32_t ne = h->nentries;
    const uint32_t t = threadIdx.x;
    // One workgroup per tile; the grid (a multiple of 8) covers the largest
    // wave's launch and the workgroups past this wave's tiles leave at once.
    // Workgroup x runs on XCD x mod 8 and takes tile (x mod 8) * per + x / 8
    // (per = this launch's tiles / 8, rounded up), so each XCD walks a
    // contiguous run of tiles and every XCD gets an eighth of them.
    const uint64_t per = (total + 7) / 8, m = blockIdx.x >> 3;
    if (m >= per) return;
    const uint64_t b = uint64_t(blockIdx.x & 7u) * per + m;
    if (b >= total) return;
    const QEntry& en = E[find_entry(E, ne, b)];
    const QDesc& d = descs[en.desc];
    const uint32_t lb = uint32_t(b - en.block_start);
    const uint32_t fl = d.flags;
    uint32_t bb = 0;
    for (uint32_t i = 0; i < d.nout; ++i) bb |= ((lb >> i) & 1u) << d.outpos[i];
    const uint32_t g0 = bb | thread_bits<0>(d, t);
    const double* csi = cs + d.cs_off;
    if (t < 24) {
        const uint32_t sl = t < 12 ? d.slotA[t] : d.slotB[t - 12];
        rxc[t] = make_double2(csi[2 * sl], csi[2 * sl + 1]);
    }
    __syncthreads();
    Regs<R> s[NS];
    cx<R>* sp = psi + d.state_off;
    cx<R>* lp = lam + d.state_off;
    if (NS == 2 || (fl & F_LOAD)) {
#pragma unroll
        for (int st = 0; st < NS; ++st) {
            const cx<R>* src = (st ? lp : sp) + g0;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const cx<R> v = src[k];
                s[st].re[k] = v.re;
                s[st].im[k] = v.im;
            }
        }
    } else {
        // |+>^n: the Hadamard layer on |0..0>.
        const R a = R(1.0 / std::sqrt(double(1ull << d.n)));
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            s[0].re[k] = a;
            s[0].im[k] = R(0);
        }
    }
    if constexpr (NS == 1) {
        if constexpr (A) {
            rx_group<R, 1, 0>(s, rxc, qred, 0, t);
            exchange<R, 1, 0, 1>(s, xb, t);
            rx_group<R, 1, 1>(s, rxc, qred, 0, t);
            exchange<R, 1, 1, 2>(s, xb, t);
            rx_group<R, 1, 2>(s, rxc, qred, 0, t);
            exchange<R, 1, 2, 3>(s, xb, t);
            rx_group<R, 1, 3>(s, rxc, qred, 0, t);
            diag_layer<R, 1, 3>(s, d, terms, csi, qred, bb, t);
            rx_group<R, 1, 3>(s, rxc, qred, 12, t);
            exchange<R, 1, 3, 2>(s, xb, t);
            rx_group<R, 1, 2>(s, rxc, qred, 12, t);
            exchange<R, 1, 2, 1>(s, xb, t);
            rx_group<R, 1, 1>(s, rxc, qred, 12, t);
            exchange<R, 1, 1, 0>(s, xb, t);
            rx_group<R, 1, 0>(s, rxc, qred, 12, t);
        } else {
            if constexpr (D) diag_layer<R, 1, 0>(s, d, terms, csi, qred, bb, t);
            rx_group<R, 1, 0>(s, rxc, qred, 12, t);
            exchange<R, 1, 0, 1>(s, xb, t);
            rx_group<R, 1, 1>(s, rxc, qred, 12, t);
            exchange<R, 1, 1, 2>(s, xb, t);
            rx_group<R, 1, 2>(s, rxc, qred, 12, t);
            exchange<R, 1, 2, 3>(s, xb, t);
            rx_group<R, 1, 3>(s, rxc, qred, 12, t);
            exchange<R, 1, 3, 0>(s, xb, t);
        }
        if constexpr (COST) {
            Regs<R> l;
            const double en_t = costate<R>(s[0], l, d, terms, coef + d.coef_off, g0);
            if (fl & F_LAMBDA) store_runs<R>(lp, l, g0, xb, t);
            // The tile's energy partial: wave sums, then the waves in order.
            dd acc{en_t, 0.0};
            for (int m = 32; m >= 1; m >>= 1)
                acc = qf::dd_add(acc, dd{__shfl_xor(acc.s, m, 64), __shfl_xor(acc.c, m, 64)});
            if ((t & 63u) == 0) ered[t >> 6] = acc;
            __syncthreads();
            if (t == 0) {
                dd tot = ered[0];
                for (int w = 1; w < kQW; ++w) tot = qf::dd_add(tot, ered[w]);
                epart[d.epart_off + lb] = tot;
            }
        }
        store_runs<R>(sp, s[0], g0, xb, t);
    } else {
        // The forward pass backwards.
        if constexpr (A) {  // forward: A on 0..3, D in phase 3, B on 3..0
            rx_group<R, 2, 0>(s, rxc, qred, 12, t);
            exchange<R, 2, 0, 1>(s, xb, t);
            rx_group<R, 2, 1>(s, rxc, qred, 12, t);
            exchange<R, 2, 1, 2>(s, xb, t);
            rx_group<R, 2, 2>(s, rxc, qred, 12, t);
            exchange<R, 2, 2, 3>(s, xb, t);
            rx_group<R, 2, 3>(s, rxc, qred, 12, t);
            diag_layer<R, 2, 3>(s, d, terms, csi, qred, bb, t);
            rx_group<R, 2, 3>(s, rxc, qred, 0, t);
            exchange<R, 2, 3, 2>(s, xb, t);
            rx_group<R, 2, 2>(s, rxc, qred, 0, t);
            exchange<R, 2, 2, 1>(s, xb, t);
            rx_group<R, 2, 1>(s, rxc, qred, 0, t);
            exchange<R, 2, 1, 0>(s, xb, t);
            rx_group<R, 2, 0>(s, rxc, qred, 0, t);
        } else {  // forward: D in phase 0 (SH_DB), B on 0..3 and back to 0
            exchange<R, 2, 0, 3>(s, xb, t);
            rx_group<R, 2, 3>(s, rxc, qred, 12, t);
            exchange<R, 2, 3, 2>(s, xb, t);
            rx_group<R, 2, 2>(s, rxc, qred, 12, t);
            exchange<R, 2, 2, 1>(s, xb, t);
            rx_group<R, 2, 1>(s, rxc, qred, 12, t);
            exchange<R, 2, 1, 0>(s, xb, t);
            rx_group<R, 2, 0>(s, rxc, qred, 12, t);
            if constexpr (D) diag_layer<R, 2, 0>(s, d, terms, csi, qred, bb, t);
        }
        store_runs<R>(sp, s[0], g0, xb, t);
        store_runs<R>(lp, s[1], g0, xb, t);
        // The tile's overlap of every recorded rotation: its waves' sums in
        // wave order, into the [slot][tile] partial table.
        __syncthreads();
        const uint32_t nblk = 1u << d.logb;
        const uint32_t rows = 24 + (D ? d.t_n : 0);
        for (uint32_t r = t; r < rows; r += kQT) {
            uint32_t slot;
            if (r < 12) {
                if (!A) continue;
                slot = d.slotA[r];
            } else if (r < 24) {
                slot = d.slotB[r - 12];
            } else {
                slot = terms[d.t_off + (r - 24)].slot;
            }
            double q = 0.0;
            for (int w = 0; w < kQW; ++w) q += qred[r * kQW + w];
            partial[d.part_off + uint64_t(slot) * nblk + lb] = q;
        }
    }
}

// Gradients (one workgroup per parameter: its occurrences' partial columns,
// tiles in order, Neumaier) and energies (one per instance).
__global__ __launch_bounds__(256) void k_qaoa_finish(const QFin* __restrict__ fin, const uint32_t* __restrict__ nfin_p,
                                                     const uint32_t* __restrict__ slots,
                                                     const double* __restrict__ partial, const dd* __restrict__ epart,
                                                     double* __restrict__ grads, double* __restrict__ energies) {
    __shared__ dd red[4];
    const uint32_t nfin = *nfin_p;
    for (uint32_t w = blockIdx.x; w < nfin; w += gridDim.x) {
        const QFin f = fin[w];
        dd acc{0.0, 0.0};
        if (f.is_energy) {
            for (uint32_t b = threadIdx.x; b < f.nblk; b += blockDim.x) acc = qf::dd_add(acc, epart[f.part_off + b]);
        } else {
            for (uint32_t o = f.first; o < f.first + f.count; ++o) {
                const double* col = partial + f.part_off + uint64_t(slots[o]) * f.nblk;
                for (uint32_t b = threadIdx.x; b < f.nblk; b += blockDim.x) qf::neumaier(acc, col[b]);
            }
        }
        for (int m = 32; m >= 1; m >>= 1) acc = qf::dd_add(acc, dd{__shfl_xor(acc.s, m, 64), __shfl_xor(acc.c, m, 64)});
        if ((threadIdx.x & 63u) == 0) red[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            dd tot = red[0];
            for (int i = 1; i < 4; ++i) tot = qf::dd_add(tot, red[i]);
            if (f.is_energy) energies[f.out] = tot.s + tot.c;
            else grads[f.out] = tot.s + tot.c;
        }
        __syncthreads();
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host
// ---------------------------------------------------------------------------
namespace {

using qfrt::fail;

// One pass of an instance's schedule.
struct PassPlan {
    uint32_t hi[9];
    uint32_t flags = 0;
    int diag = -1;                    // cost layer applied in this pass
    int layer_a = -1, layer_b = -1;   // layers of the sweep A and sweep B rotations
    bool rot_a[12] = {}, rot_b[12] = {};  // tile bits rotated in each sweep
};

// Passes between two cost layers: 1 while two tile sets cover the qubits
// (n <= 21), else 2 (three sets, n <= 30).  QF_QAOA_LAYER_PASSES = 2 asks for
// the second form on narrower instances too (tests).
int layer_stride(uint32_t n) {
    if (const char* e = std::getenv("QF_QAOA_LAYER_PASSES"))
        if (std::atoi(e) == 2) return 2;
    return n <= 21 ? 1 : 2;
}

// The tile sets: qubits 0..2 in every set, the others in k + 1 chunks of nine
// (a short chunk is filled from the first qubits), so any k + 1 consecutive
// sets hold every qubit.
std::vector<std::vector<uint32_t>> tile_sets(uint32_t n, int k) {
    std::vector<uint32_t> rest;
    for (uint32_t q = 3; q < n; ++q) rest.push_back(q);
    std::vector<std::vector<uint32_t>> sets(size_t(k) + 1);
    size_t pos = 0;
    for (auto& s : sets) {
        while (s.size() < 9 && pos < rest.size()) s.push_back(rest[pos++]);
        for (size_t i = 0; s.size() < 9; ++i)
            if (std::find(s.begin(), s.end(), rest[i]) == s.end()) s.push_back(rest[i]);
        std::sort(s.begin(), s.end());
    }
    return sets;
}

int tile_bit_of(const uint32_t hi[9], uint32_t q) {
    if (q < 3) return int(q);
    for (int i = 0; i < 9; ++i)
        if (hi[i] == q) return 3 + i;
    return -1;
}

// The pass shape of a pass's flags (-1: none of the three).  Qubits 0..2 sit
// in every tile, so a pass with a cost layer always rotates after it.
int shape_of(uint32_t fl) {
    const bool a = fl & F_SWA, dg = fl & F_DIAG, b = fl & F_SWB;
    if (a && dg && b) return SH_ADB;
    if (!a && dg && b) return SH_DB;
    if (!a && !dg && b) return SH_B;
    return -1;
}

// Rotations of cost layer l at pass l * k; X rotation (l, q) at the first pass
// from there whose tile holds q: before the next cost layer, which the tile
// sets guarantee.  A pass with a cost layer rotates the previous layer's
// leftovers before it (sweep A) and its own layer after it (sweep B); a pass
// without one rotates its layer in sweep B.
bool schedule(const Form& f, std::vector<PassPlan>& out) {
    const uint32_t n = f.n;
    const int k = layer_stride(n);
    if (n < 12 || n > 32 || (k == 1 && n > 21) || n > 30) return false;
    const auto sets = tile_sets(n, k);
    const int p = int(f.layers.size());
    auto set_of = [&](int j) -> const std::vector<uint32_t>& { return sets[size_t(j) % sets.size()]; };
    std::vector<std::vector<int>> at(size_t(p), std::vector<int>(n, -1));
    int last = (p - 1) * k;
    for (int l = 0; l < p; ++l)
        for (uint32_t q = 0; q < n; ++q) {
            int j = l * k;
            while (!(q < 3 || std::binary_search(set_of(j).begin(), set_of(j).end(), q))) ++j;
            if (l + 1 < p && j > (l + 1) * k) return false;
            at[size_t(l)][q] = j;
            last = std::max(last, j);
        }
    out.assign(size_t(last) + 1, PassPlan{});
    for (int j = 0; j <= last; ++j) {
        PassPlan& P = out[size_t(j)];
        std::copy(set_of(j).begin(), set_of(j).end(), P.hi);
        if (j % k == 0 && j / k < p) P.diag = j / k;
    }
    for (int l = 0; l < p; ++l)
        for (uint32_t q = 0; q < n; ++q) {
            PassPlan& P = out[size_t(at[size_t(l)][q])];
            const int b = tile_bit_of(P.hi, q);
            if (b < 0) return false;
            if (P.diag >= 0 && l < P.diag) {
                if (P.layer_a >= 0 && P.layer_a != l) return false;
                P.layer_a = l;
                P.rot_a[b] = true;
            } else {
                if (P.layer_b >= 0 && P.layer_b != l) return false;
                P.layer_b = l;
                P.rot_b[b] = true;
            }
        }
    // A pass whose tile holds no rotation still owed (three sets on a narrow
    // instance) does nothing and is dropped.
    std::vector<PassPlan> kept;
    for (const PassPlan& P : out)
        if (P.layer_a >= 0 || P.diag >= 0 || P.layer_b >= 0) kept.push_back(P);
    out.swap(kept);
    for (size_t j = 0; j < out.size(); ++j) {
        PassPlan& P = out[j];
        P.flags = (j > 0 ? F_LOAD : 0u) | (P.layer_a >= 0 ? F_SWA : 0u) | (P.diag >= 0 ? F_DIAG : 0u) |
                  (P.layer_b >= 0 ? F_SWB : 0u) | (j + 1 == out.size() ? F_COST : 0u);
        if (shape_of(P.flags) < 0) return false;
    }
    return true;
}

// The qubits of each role in phase P (0 or 3) of a pass with tile hi[].
struct Roles {
    int reg[3];      // register bit j -> qubit
    int lane[6];     // lane bit -> qubit
    int wave[3];     // wave bit -> qubit
};
Roles roles(const uint32_t hi[9], int P) {
    Roles r{};
    auto qubit_of_tile_bit = [&](int b) { return b < 3 ? b : int(hi[b - 3]); };
    for (int j = 0; j < 3; ++j) r.reg[j] = qubit_of_tile_bit(3 * P + j);
    int tb[9], m = 0;
    for (int b = 0; b < 12; ++b)
        if (b / 3 != P) tb[m++] = b;
    for (int i = 0; i < 6; ++i) r.lane[i] = qubit_of_tile_bit(tb[i]);
    for (int i = 0; i < 3; ++i) r.wave[i] = qubit_of_tile_bit(tb[6 + i]);
    return r;
}

// Classifies the words (a, b) (b == a: one Z) of a diagonal for phase P into
// QTerms sorted by register class (the first nclasses classes).  cls_out[c]
// counts class c.  pairs also gets the words inside the registers (register
// mask 3, 5 or 6), with their slot.
void classify(const std::vector<std::array<uint32_t, 3>>& words, const uint32_t hi[9], int P, int nclasses,
              std::vector<QTerm>& terms, uint32_t* cls_out, std::vector<std::pair<uint32_t, uint32_t>>* pairs) {
    const Roles R = roles(hi, P);
    std::vector<std::vector<QTerm>> by(static_cast<size_t>(kClasses));
    for (const auto& w : words) {
        uint32_t qs[2] = {w[0], w[1]};
        const int nq = w[0] == w[1] ? 1 : 2;
        QTerm t{0, 0, 0, w[2]};
        uint32_t regmask = 0;
        for (int i = 0; i < nq; ++i) {
            const int q = int(qs[i]);
            int role = 0;  // 0 block, 1 register, 2 lane, 3 wave
            for (int j = 0; j < 3; ++j)
                if (R.reg[j] == q) { role = 1; regmask |= 1u << j; }
            for (int j = 0; j < 6 && role == 0; ++j)
                if (R.lane[j] == q) { role = 2; t.lanemask |= 1u << j; }
            for (int j = 0; j < 3 && role == 0; ++j)
                if (R.wave[j] == q) role = 3;
            if (role != 1) t.gm |= 1u << q;
            if (role == 0 || role == 3) t.wbmask |= 1u << q;
        }
        int c = 0;
        for (; c < kClasses; ++c)
            if (class_mask(c) == regmask) break;
        if (c >= 4 && pairs) pairs->push_back({regmask, w[2]});
        if (c >= nclasses) continue;
        by[size_t(c)].push_back(t);
    }
    for (int c = 0; c < nclasses; ++c) {
        cls_out[c] = uint32_t(by[size_t(c)].size());
        terms.insert(terms.end(), by[size_t(c)].begin(), by[size_t(c)].end());
    }
}

struct Dev {
    void* p = nullptr;
    ~Dev() {
        if (p) (void)hipFree(p);
    }
};

// The launches of one memory wave: forward pass j of every instance (j < F),
// reverse pass j (j < Rv), the finishing kernel; captured once, replayed per
// wave over tables rewritten in place.
struct Buffers {
    void* psi;
    void* lam;
    const unsigned char* launch;
    size_t stride;
    const QDesc* descs;
    const QTerm* terms;
    const double *cs, *coef;
    double* part;
    dd* epart;
    const QFin* fin;
    const uint32_t* nfin;
    const uint32_t* slots;
    double* grads;
    double* energies;
};

// Launch slots of a wave: forward pass j, shape, last pass or not
// (F x kShapes x 2), then reverse pass j (from each instance's last), shape.
struct Slot {
    bool fwd;
    uint32_t j;
    int shape;
    bool cost;
};
std::vector<Slot> launch_slots(uint32_t F, uint32_t Rv) {
    std::vector<Slot> v;
    for (uint32_t j = 0; j < F; ++j)
        for (int sh = 0; sh < kShapes; ++sh)
            for (int c = 0; c < 2; ++c) v.push_back({true, j, sh, c == 1});
    for (uint32_t j = 0; j < Rv; ++j)
        for (int sh = 0; sh < kShapes; ++sh) v.push_back({false, j, sh, false});
    return v;
}

template <class R, int NS, int SH, bool COST>
void launch_pass(hipStream_t st, dim3 g, const unsigned char* L, const Buffers& B) {
    hipLaunchKernelGGL((k_qaoa_pass<R, NS, SH, COST>), g, dim3(kQT), 0, st, static_cast<cx<R>*>(B.psi),
                       static_cast<cx<R>*>(B.lam), L, B.descs, B.terms, B.cs, B.coef, B.part, B.epart);
}

template <class R>
void launch_slot(hipStream_t st, const Slot& sl, dim3 g, const unsigned char* L, const Buffers& B) {
    if (sl.fwd) {
        if (sl.shape == SH_DB) sl.cost ? launch_pass<R, 1, SH_DB, true>(st, g, L, B) : launch_pass<R, 1, SH_DB, false>(st, g, L, B);
        else if (sl.shape == SH_ADB) sl.cost ? launch_pass<R, 1, SH_ADB, true>(st, g, L, B) : launch_pass<R, 1, SH_ADB, false>(st, g, L, B);
        else sl.cost ? launch_pass<R, 1, SH_B, true>(st, g, L, B) : launch_pass<R, 1, SH_B, false>(st, g, L, B);
    } else {
        if (sl.shape == SH_DB) launch_pass<R, 2, SH_DB, false>(st, g, L, B);
        else if (sl.shape == SH_ADB) launch_pass<R, 2, SH_ADB, false>(st, g, L, B);
        else launch_pass<R, 2, SH_B, false>(st, g, L, B);
    }
}

template <class R>
qf_status capture(hipStream_t st, const std::vector<Slot>& slots, const std::vector<uint64_t>& grid,
                  const Buffers& B, hipGraphExec_t* exec, uint32_t* nodes) {
    hipGraph_t graph = nullptr;
    uint32_t launched = 0;
    QF_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (size_t k = 0; k < slots.size(); ++k) {
        if (grid[k] == 0) continue;  // no instance of this shape at this pass in any wave
        launch_slot<R>(st, slots[k], dim3(unsigned(grid[k])), B.launch + k * B.stride, B);
        ++launched;
    }
    hipLaunchKernelGGL(k_qaoa_finish, dim3(4096), dim3(256), 0, st, B.fin, B.nfin, B.slots, B.part, B.epart, B.grads,
                       B.energies);
    hipError_t ce = hipStreamEndCapture(st, &graph);
    if (ce != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return qfrt::hip_fail(ce, "hipStreamEndCapture");
    }
    hipError_t ie = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) return qfrt::hip_fail(ie, "hipGraphInstantiate");
    *nodes = launched + 1;
    return QF_OK;
}

}  // namespace

bool match(const qf_uneven_instance& q, const qfrt::PCirc& pc, Form& f) {
    const uint32_t n = q.qubits;
    if (n < 12 || n > 30) return false;
    f = Form{};
    f.n = n;
    size_t i = 0;
    uint32_t slot = 0;
    uint64_t seen = 0;
    const uint64_t all = (uint64_t(1) << n) - 1;
    while (i < q.nops && q.ops[i].kind == 0 && q.ops[i].gate == QF_GATE_H) {
        if (seen >> q.ops[i].a & 1) return false;
        seen |= uint64_t(1) << q.ops[i].a;
        ++i;
    }
    if (seen != all) return false;
    while (i < q.nops) {
        Form::Layer L;
        while (i + 2 < q.nops && q.ops[i].kind == 3 && q.ops[i + 1].kind == 1 && q.ops[i + 1].gate == 2 &&
               q.ops[i + 1].a == q.ops[i].b && q.ops[i + 2].kind == 3 && q.ops[i + 2].a == q.ops[i].a &&
               q.ops[i + 2].b == q.ops[i].b) {
            L.edges.push_back({q.ops[i].a, q.ops[i].b, slot++});
            i += 3;
        }
        L.mix.assign(n, ~0u);
        uint64_t mixed = 0;
        while (i < q.nops && q.ops[i].kind == 1 && q.ops[i].gate == 0 && !(mixed >> q.ops[i].a & 1)) {
            mixed |= uint64_t(1) << q.ops[i].a;
            L.mix[q.ops[i].a] = slot++;
            ++i;
        }
        if (mixed != all || L.edges.size() > kMaxTerms) return false;
        f.layers.push_back(std::move(L));
    }
    if (f.layers.empty() || slot != pc.slot_param.size()) return false;
    // Observable: real Z and Z Z words and a constant.
    for (size_t t = 0; t < q.nterms; ++t) {
        const qf_pauli_term& w = q.obs[t];
        if (w.x_mask != 0 || w.im != 0.0 || (w.z_mask >> n) != 0) return false;
        const int pc_ = __builtin_popcountll(w.z_mask);
        if (pc_ == 0) {
            f.c0 += w.re;
        } else if (pc_ <= 2) {
            const uint32_t a = uint32_t(__builtin_ctzll(w.z_mask));
            const uint32_t b = pc_ == 2 ? uint32_t(63 - __builtin_clzll(w.z_mask)) : a;
            f.obs.push_back({a, b, w.re});
        } else {
            return false;
        }
    }
    if (f.obs.size() > kMaxTerms) return false;
    std::vector<PassPlan> plan;
    return schedule(f, plan);
}

namespace {

template <class R>
qf_status run_t(qf_backend* be, const qf_uneven_instance* inst, const std::vector<uint32_t>& ids,
                const std::vector<qfrt::PCirc>& pcs, const std::vector<Form>& forms,
                const std::vector<uint64_t>& grad_off, int32_t precision, bool want_grad, double* energies,
                double* grads, qf_uneven_stats& S) {
    qf_status s;
    void* sp = nullptr;
    qf_backend_stream(be, &sp);
    hipStream_t st = static_cast<hipStream_t>(sp);
    const size_t ab = qfrt::amp_bytes(precision);
    struct Inst {
        uint32_t k = 0, n = 0, nslots = 0;
        std::vector<PassPlan> passes;
        std::vector<double> cs;
        uint64_t grad_local = 0;  // first gradient of the instance in the call's device array
    };
    std::vector<Inst> X(ids.size());
    uint64_t ngrad = 0;
    uint32_t F = 0;
    for (size_t i = 0; i < ids.size(); ++i) {
        Inst& I = X[i];
        I.k = ids[i];
        I.n = inst[I.k].qubits;
        const qfrt::PCirc& pc = pcs[I.k];
        I.nslots = uint32_t(pc.slot_param.size());
        if (!schedule(forms[I.k], I.passes)) return fail(QF_INVALID_ARGUMENT, "QAOA schedule failed");
        I.cs.assign(size_t(I.nslots + 1) * 2, 0.0);
        qfrt::fill_row(pc, inst[I.k].params, -1, 0.0, I.cs.data());
        I.cs[2 * I.nslots] = 1.0;  // the identity slot of tile bits a sweep does not rotate
        I.grad_local = ngrad;
        ngrad += inst[I.k].nparams;
        F = std::max<uint32_t>(F, uint32_t(I.passes.size()));
    }
    const uint32_t Rv = want_grad ? F : 0;

    // Memory waves: first fit, widest first.
    auto nblk_of = [](uint32_t n) { return uint64_t(1) << (n - kQM); };
    auto footprint = [&](const Inst& I) {
        const uint64_t nb = nblk_of(I.n);
        return (want_grad ? 2 : 1) * (uint64_t(1) << I.n) * ab + (want_grad ? uint64_t(I.nslots + 1) * nb * 8 : 0) +
               nb * sizeof(dd) + 4096;
    };
    std::vector<uint32_t> order(X.size());
    for (uint32_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return X[a].n > X[b].n; });
    size_t freeb = 0, totalb = 0;
    QF_HIP_TRY(hipMemGetInfo(&freeb, &totalb));
    // The backend's arena from an earlier call is reused, so it counts as free.
    uint64_t budget = uint64_t(double(freeb + be->arena_bytes) * 0.8);
    if (const char* cap = std::getenv("QF_UNEVEN_WAVE_BYTES")) {
        const unsigned long long v = std::strtoull(cap, nullptr, 10);
        if (v > 0 && v < budget) budget = v;
    }
    // First fit, widest first; false (too_big = the instance) when an
    // instance alone exceeds cap.
    uint32_t too_big = 0;
    auto pack = [&](uint64_t cap, std::vector<std::vector<uint32_t>>& W, std::vector<uint64_t>& WB) {
        W.clear();
        WB.clear();
        for (uint32_t i : order) {
            const uint64_t f = footprint(X[i]);
            if (f > cap) {
                too_big = X[i].k;
                return false;
            }
            bool placed = false;
            for (size_t w = 0; w < W.size() && !placed; ++w)
                if (WB[w] + f <= cap) {
                    W[w].push_back(i);
                    WB[w] += f;
                    placed = true;
                }
            if (!placed) {
                W.push_back({i});
                WB.push_back(f);
            }
        }
        return true;
    };
    auto up = [](uint64_t x) { return (x + 255) / 256 * 256; };
    // Arena bytes of a wave set: psi, lambda, partial table and energy
    // partials, each sized for its largest wave.
    auto arena_need = [&](const std::vector<std::vector<uint32_t>>& W) {
        uint64_t a_max = 0, p_max = 0, e_max = 0;
        for (const auto& w : W) {
            uint64_t a = 0, p = 0, e = 0;
            for (uint32_t i : w) {
                a += uint64_t(1) << X[i].n;
                if (want_grad) p += uint64_t(X[i].nslots + 1) * nblk_of(X[i].n);
                e += nblk_of(X[i].n);
            }
            a_max = std::max(a_max, a);
            p_max = std::max(p_max, p);
            e_max = std::max(e_max, e);
        }
        return up(a_max * ab) * (want_grad ? 2 : 1) + up(std::max<uint64_t>(p_max, 1) * 8) + up(e_max * sizeof(dd));
    };
    std::vector<std::vector<uint32_t>> waves;
    std::vector<uint64_t> wave_bytes;
    if (!pack(budget, waves, wave_bytes))
        return fail(QF_OUT_OF_MEMORY, "instance " + std::to_string(too_big) + " does not fit in device memory");
    // An arena kept from an earlier call is used as it is when the batch
    // packs into as few waves within it: growing it means freeing most of
    // HBM and allocating it again, and such an allocation right after the
    // free has taken 6 s.
    if (be->arena_bytes > 0 && arena_need(waves) > be->arena_bytes) {
        std::vector<std::vector<uint32_t>> w2;
        std::vector<uint64_t> b2;
        if (pack(std::min<uint64_t>(be->arena_bytes, budget), w2, b2) && w2.size() <= waves.size() &&
            arena_need(w2) <= be->arena_bytes) {
            waves.swap(w2);
            wave_bytes.swap(b2);
        }
    }
    // Buffers sized for the largest wave.
    uint64_t max_amps = 0, max_part = 0, max_epart = 0;
    size_t max_inst = 1, max_desc = 1, max_terms = 1, max_cs = 2, max_coef = 1, max_fin = 1, max_slots = 1;
    for (const auto& w : waves) {
        uint64_t a = 0, p = 0, e = 0;
        size_t nd = 0, nt = 0, nc = 0, no = 0, nf = 0, nsl = 0;
        for (uint32_t i : w) {
            const Inst& I = X[i];
            a += uint64_t(1) << I.n;
            if (want_grad) p += uint64_t(I.nslots + 1) * nblk_of(I.n);
            e += nblk_of(I.n);
            nd += I.passes.size();
            for (const auto& f : forms[I.k].layers) nt += f.edges.size();
            nt += forms[I.k].obs.size();
            nc += I.cs.size();
            no += forms[I.k].obs.size();
            nf += 1 + (want_grad ? inst[I.k].nparams : 0);
            nsl += I.nslots;
        }
        max_amps = std::max(max_amps, a);
        max_part = std::max(max_part, p);
        max_epart = std::max(max_epart, e);
        max_inst = std::max(max_inst, w.size());
        max_desc = std::max(max_desc, nd);
        max_terms = std::max(max_terms, nt);
        max_cs = std::max(max_cs, nc);
        max_coef = std::max(max_coef, no);
        max_fin = std::max(max_fin, nf);
        max_slots = std::max(max_slots, nsl);
    }
    const size_t stride = (sizeof(QHeader) + max_inst * sizeof(QEntry) + 255) / 256 * 256;
    // psi, lambda, the partial tables and the energy partials live in the
    // backend's arena, kept for the next call (freeing a wave's worth of
    // device memory costs the host most of a second).
    Dev d_launch, d_desc, d_terms, d_cs, d_coef, d_fin, d_nfin, d_slots, d_grads, d_energy;
    struct {
        void* p;
    } d_psi{}, d_lam{}, d_part{}, d_epart{};
    {
        const uint64_t b_psi = up(max_amps * ab), b_lam = want_grad ? up(max_amps * ab) : 0,
                       b_part = up(std::max<uint64_t>(max_part, 1) * 8), b_ep = up(max_epart * sizeof(dd));
        void* base = nullptr;
        if ((s = qfrt::arena(be, b_psi + b_lam + b_part + b_ep, &base))) return s;
        unsigned char* a = static_cast<unsigned char*>(base);
        d_psi.p = a;
        d_lam.p = want_grad ? a + b_psi : nullptr;
        d_part.p = a + b_psi + b_lam;
        d_epart.p = a + b_psi + b_lam + b_part;
    }
    if ((s = qfrt::device_alloc(&d_launch.p, launch_slots(F, Rv).size() * stride))) return s;
    if ((s = qfrt::device_alloc(&d_desc.p, max_desc * sizeof(QDesc)))) return s;
    if ((s = qfrt::device_alloc(&d_terms.p, max_terms * sizeof(QTerm)))) return s;
    if ((s = qfrt::device_alloc(&d_cs.p, max_cs * 8))) return s;
    if ((s = qfrt::device_alloc(&d_coef.p, max_coef * 8))) return s;
    if ((s = qfrt::device_alloc(&d_fin.p, max_fin * sizeof(QFin)))) return s;
    if ((s = qfrt::device_alloc(&d_nfin.p, 16))) return s;
    if ((s = qfrt::device_alloc(&d_slots.p, max_slots * 4))) return s;
    if ((s = qfrt::device_alloc(&d_grads.p, std::max<uint64_t>(ngrad, 1) * 8))) return s;
    if ((s = qfrt::device_alloc(&d_energy.p, ids.size() * 8))) return s;

    Buffers B{d_psi.p, d_lam.p, static_cast<const unsigned char*>(d_launch.p), stride,
              static_cast<const QDesc*>(d_desc.p), static_cast<const QTerm*>(d_terms.p),
              static_cast<const double*>(d_cs.p), static_cast<const double*>(d_coef.p), static_cast<double*>(d_part.p), static_cast<dd*>(d_epart.p),
              static_cast<const QFin*>(d_fin.p), static_cast<const uint32_t*>(d_nfin.p),
              static_cast<const uint32_t*>(d_slots.p), static_cast<double*>(d_grads.p),
              static_cast<double*>(d_energy.p)};
    // Instance i's pass for launch slot sl, or -1.
    const std::vector<Slot> slots_l = launch_slots(F, Rv);
    auto pass_of = [&](const Inst& I, const Slot& sl) -> int {
        const uint32_t np = uint32_t(I.passes.size());
        if (sl.j >= np) return -1;
        const uint32_t j = sl.fwd ? sl.j : np - 1 - sl.j;
        const PassPlan& P = I.passes[j];
        if (shape_of(P.flags) != sl.shape) return -1;
        if (sl.fwd && bool(P.flags & F_COST) != sl.cost) return -1;
        return int(j);
    };
    // Workgroups of each slot: the most tiles it has in any wave.
    std::vector<uint64_t> grid(slots_l.size(), 0);
    for (const auto& w : waves)
        for (size_t k = 0; k < slots_l.size(); ++k) {
            uint64_t blocks = 0;
            for (uint32_t i : w)
                if (pass_of(X[i], slots_l[k]) >= 0) blocks += nblk_of(X[i].n);
            grid[k] = std::max(grid[k], (blocks + 7) / 8 * 8);
        }
    hipGraphExec_t exec = nullptr;
    uint32_t nodes = 0;
    if ((s = capture<R>(st, slots_l, grid, B, &exec, &nodes))) return s;
    struct ExecGuard {
        hipGraphExec_t e;
        ~ExecGuard() {
            if (e) (void)hipGraphExecDestroy(e);
        }
    } eg{exec};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    QF_HIP_TRY(hipEventCreate(&ev0));
    QF_HIP_TRY(hipEventCreate(&ev1));
    struct EvGuard {
        hipEvent_t a, b;
        ~EvGuard() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } evg{ev0, ev1};

    std::vector<unsigned char> launch(slots_l.size() * stride);
    std::vector<QDesc> descs;
    std::vector<QTerm> terms;
    std::vector<double> csbuf, coef;
    std::vector<QFin> fin;
    std::vector<uint32_t> slots;
    for (const auto& w : waves) {
        descs.clear();
        terms.clear();
        csbuf.clear();
        coef.clear();
        fin.clear();
        slots.clear();
        std::fill(launch.begin(), launch.end(), 0);
        std::vector<uint32_t> desc0(w.size());
        uint64_t amp_off = 0, part_off = 0, ep_off = 0;
        for (size_t wi = 0; wi < w.size(); ++wi) {
            const Inst& I = X[w[wi]];
            const Form& f = forms[I.k];
            const uint32_t cs_off = uint32_t(csbuf.size());
            csbuf.insert(csbuf.end(), I.cs.begin(), I.cs.end());
            const uint32_t coef_off = uint32_t(coef.size());
            std::vector<std::array<uint32_t, 3>> owords;
            for (size_t t = 0; t < f.obs.size(); ++t) {
                coef.push_back(f.obs[t].coef);
                owords.push_back({f.obs[t].a, f.obs[t].b, uint32_t(t)});
            }
            const uint64_t nb = nblk_of(I.n);
            desc0[wi] = uint32_t(descs.size());
            for (const PassPlan& P : I.passes) {
                QDesc d{};
                d.state_off = amp_off;
                d.part_off = part_off;
                d.epart_off = ep_off;
                d.n = I.n;
                d.logb = I.n - kQM;
                d.flags = P.flags | ((P.flags & F_COST) && want_grad ? F_LAMBDA : 0u);
                d.nslots1 = I.nslots + 1;
                d.cs_off = cs_off;
                d.coef_off = coef_off;
                std::copy(P.hi, P.hi + 9, d.hi);
                bool in_tile[32] = {};
                for (uint32_t q = 0; q < 3; ++q) in_tile[q] = true;
                for (int q = 0; q < 9; ++q) in_tile[P.hi[q]] = true;
                for (uint32_t q = 0; q < I.n; ++q)
                    if (!in_tile[q]) d.outpos[d.nout++] = q;
                for (int b = 0; b < 12; ++b) {
                    const uint32_t q = b < 3 ? uint32_t(b) : P.hi[b - 3];
                    d.slotA[b] = P.rot_a[b] ? f.layers[size_t(P.layer_a)].mix[q] : I.nslots;
                    d.slotB[b] = P.rot_b[b] ? f.layers[size_t(P.layer_b)].mix[q] : I.nslots;
                }
                for (int k = 0; k < 8; ++k) {
                    d.w2[2 * k] = 1.0;
                    d.w2[2 * k + 1] = 0.0;
                }
                if (P.diag >= 0) {
                    std::vector<std::array<uint32_t, 3>> words;
                    for (const auto& e : f.layers[size_t(P.diag)].edges) words.push_back({e.u, e.v, e.slot});
                    std::vector<std::pair<uint32_t, uint32_t>> pairs;
                    d.t_off = uint32_t(terms.size());
                    classify(words, P.hi, (P.flags & F_SWA) ? 3 : 0, kClasses, terms, d.cnt, &pairs);
                    for (int c = 0; c < kClasses; ++c) d.t_n += d.cnt[c];
                    // exp(-i theta/2 z z) of the register-pair edges, per register.
                    for (const auto& pr : pairs)
                        for (int k = 0; k < 8; ++k) {
                            const bool odd = __builtin_popcount(uint32_t(k) & pr.first) & 1;
                            const double c = I.cs[2 * pr.second], sn = odd ? I.cs[2 * pr.second + 1]
                                                                           : -I.cs[2 * pr.second + 1];
                            const double r = d.w2[2 * k] * c - d.w2[2 * k + 1] * sn;
                            d.w2[2 * k + 1] = d.w2[2 * k] * sn + d.w2[2 * k + 1] * c;
                            d.w2[2 * k] = r;
                        }
                }
                if (P.flags & F_COST) {
                    std::vector<std::pair<uint32_t, uint32_t>> pairs;
                    d.o_off = uint32_t(terms.size());
                    classify(owords, P.hi, 0, 4, terms, d.ocnt, &pairs);
                    for (const auto& pr : pairs)
                        for (int k = 0; k < 8; ++k) {
                            const bool odd = __builtin_popcount(uint32_t(k) & pr.first) & 1;
                            d.ow2[k] += odd ? -f.obs[pr.second].coef : f.obs[pr.second].coef;
                        }
                    d.c0 = f.c0;
                }
                descs.push_back(d);
            }
            // Finishing work: the energy, then every parameter's occurrences.
            fin.push_back(QFin{ep_off, uint64_t(&I - X.data()), uint32_t(nb), 0, 0, 1});
            if (want_grad) {
                const qfrt::PCirc& pc = pcs[I.k];
                const uint32_t np = inst[I.k].nparams;
                std::vector<uint32_t> ps(np + 1, 0);
                for (uint32_t q = 0; q < I.nslots; ++q) ps[pc.slot_param[q] + 1]++;
                for (uint32_t p = 0; p < np; ++p) ps[p + 1] += ps[p];
                const uint32_t base = uint32_t(slots.size());
                slots.resize(slots.size() + I.nslots);
                std::vector<uint32_t> fill(ps.begin(), ps.end() - 1);
                for (uint32_t q = 0; q < I.nslots; ++q) slots[base + fill[pc.slot_param[q]]++] = q;
                for (uint32_t p = 0; p < np; ++p)
                    fin.push_back(QFin{part_off, I.grad_local + p, uint32_t(nb), base + ps[p], ps[p + 1] - ps[p], 0});
            }
            amp_off += uint64_t(1) << I.n;
            if (want_grad) part_off += uint64_t(I.nslots + 1) * nb;
            ep_off += nb;
        }
        // Launch tables, one per slot.
        for (size_t k = 0; k < slots_l.size(); ++k) {
            QHeader* h = reinterpret_cast<QHeader*>(launch.data() + k * stride);
            QEntry* E = reinterpret_cast<QEntry*>(launch.data() + k * stride + sizeof(QHeader));
            uint64_t blocks = 0;
            uint32_t ne = 0;
            for (size_t wi = 0; wi < w.size(); ++wi) {
                const Inst& I = X[w[wi]];
                const int j = pass_of(I, slots_l[k]);
                if (j < 0) continue;
                E[ne].block_start = blocks;
                E[ne].desc = desc0[wi] + uint32_t(j);
                ++ne;
                blocks += nblk_of(I.n);
            }
            h->total_blocks = blocks;
            h->nentries = ne;
        }
        const uint32_t nfin = uint32_t(fin.size());
        QF_HIP_TRY(hipMemcpy(d_launch.p, launch.data(), launch.size(), hipMemcpyHostToDevice));
        QF_HIP_TRY(hipMemcpy(d_desc.p, descs.data(), descs.size() * sizeof(QDesc), hipMemcpyHostToDevice));
        if (!terms.empty())
            QF_HIP_TRY(hipMemcpy(d_terms.p, terms.data(), terms.size() * sizeof(QTerm), hipMemcpyHostToDevice));
        QF_HIP_TRY(hipMemcpy(d_cs.p, csbuf.data(), csbuf.size() * 8, hipMemcpyHostToDevice));
        if (!coef.empty()) QF_HIP_TRY(hipMemcpy(d_coef.p, coef.data(), coef.size() * 8, hipMemcpyHostToDevice));
        QF_HIP_TRY(hipMemcpy(d_fin.p, fin.data(), fin.size() * sizeof(QFin), hipMemcpyHostToDevice));
        QF_HIP_TRY(hipMemcpy(d_nfin.p, &nfin, 4, hipMemcpyHostToDevice));
        if (!slots.empty()) QF_HIP_TRY(hipMemcpy(d_slots.p, slots.data(), slots.size() * 4, hipMemcpyHostToDevice));
        QF_HIP_TRY(hipEventRecord(ev0, st));
        QF_HIP_TRY(hipGraphLaunch(exec, st));
        QF_HIP_TRY(hipEventRecord(ev1, st));
        QF_HIP_TRY(hipStreamSynchronize(st));
        float ms = 0.0f;
        QF_HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
        S.device_ms += ms;
        for (uint32_t i : w) {
            const Inst& I = X[i];
            const double np = double(I.passes.size()), dim = double(uint64_t(1) << I.n);
            S.alg_bytes += np * 2 * dim * double(ab) + (want_grad ? np * 4 * dim * double(ab) + dim * double(ab) : 0.0);
        }
    }
    std::vector<double> g(ngrad), e(ids.size());
    if (ngrad) QF_HIP_TRY(hipMemcpy(g.data(), d_grads.p, ngrad * 8, hipMemcpyDeviceToHost));
    QF_HIP_TRY(hipMemcpy(e.data(), d_energy.p, ids.size() * 8, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < ids.size(); ++i) {
        const Inst& I = X[i];
        energies[I.k] = e[i];
        if (want_grad)
            for (uint32_t p = 0; p < inst[I.k].nparams; ++p) grads[grad_off[I.k] + p] = g[I.grad_local + p];
    }
    S.waves += uint32_t(waves.size());
    S.graph_replays += uint32_t(waves.size());
    S.graph_nodes = std::max(S.graph_nodes, nodes);
    S.qaoa_instances += uint32_t(ids.size());
    S.qaoa_forward_passes = std::max(S.qaoa_forward_passes, F);
    return QF_OK;
}

}  // namespace

qf_status run(qf_backend* be, const qf_uneven_instance* inst, const std::vector<uint32_t>& ids,
              const std::vector<qfrt::PCirc>& pcs, const std::vector<Form>& forms,
              const std::vector<uint64_t>& grad_off, int32_t precision, bool want_grad, double* energies,
              double* grads, qf_uneven_stats& S) {
    if (ids.empty()) return QF_OK;
    return precision == QF_C128
               ? run_t<double>(be, inst, ids, pcs, forms, grad_off, precision, want_grad, energies, grads, S)
               : run_t<float>(be, inst, ids, pcs, forms, grad_off, precision, want_grad, energies, grads, S);
}

}  // namespace qfq
