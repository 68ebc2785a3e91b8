// Register-resident tile kernel for gfx950.
//
// A workgroup of 2^(M-RB) threads owns one 2^M-amplitude tile (same tile
// geometry as k_tile).  Each thread keeps 2^RB amplitudes of the tile in
// registers: the RB "register bits" of the current phase index them, the
// remaining M-RB tile bits index the thread.  Gates on register bits are
// pure register arithmetic; diagonal gates and CNOT controls read their bits
// from the element's global index, so they run on any qubit.  Between phases
// the tile is re-distributed through LDS (one write + one read), instead of
// one LDS round trip per gate.  Loads and stores go straight between HBM and
// registers with the first / last phase's mapping.
//
// NS = 2 carries a second state (the adjoint costate) through the same ops and
// records Im <lambda|G|psi> for TF_RECORD rotations before applying them.
#pragma once
#include "qf_kernels.h"

namespace qf {

// Global offset (without the workgroup base) of tile index j.
__device__ __forceinline__ uint64_t tile_to_global(uint32_t j, const TileArgs& a) {
    uint64_t g = j & ((1u << a.low) - 1);
    const uint32_t nh = a.M - a.low;
    for (uint32_t k = 0; k < nh; ++k) g |= uint64_t((j >> (a.low + k)) & 1) << a.high[k];
    return g;
}

// Tile index of the thread's register-0 element: thread bits go, in order,
// into the tile bits that are not register bits.
__device__ __forceinline__ uint32_t thread_tile_index(uint32_t t, uint32_t rbmask, uint32_t M) {
    uint32_t j = 0, bit = 0;
    for (uint32_t b = 0; b < M; ++b) {
        if ((rbmask >> b) & 1) continue;
        j |= ((t >> bit) & 1) << b;
        ++bit;
    }
    return j;
}

template <int RB>
struct RegMap {
    uint32_t jk[1 << RB];  // tile index offset of register k
    uint64_t gk[1 << RB];  // global offset of register k
    uint32_t jt;           // tile index of the thread's register 0
    uint64_t gt;           // global index of the thread's register 0 (with base)
};

template <int RB>
__device__ __forceinline__ void decode_map(uint32_t packed, const TileArgs& a, uint64_t base, RegMap<RB>& m) {
    uint32_t rb[RB];
    uint32_t mask = 0;
#pragma unroll
    for (int i = 0; i < RB; ++i) {
        rb[i] = (packed >> (4 * i)) & 15u;
        mask |= 1u << rb[i];
    }
#pragma unroll
    for (int k = 0; k < (1 << RB); ++k) {
        uint32_t j = 0;
#pragma unroll
        for (int i = 0; i < RB; ++i) j |= uint32_t((k >> i) & 1) << rb[i];
        m.jk[k] = j;
        m.gk[k] = tile_to_global(j, a);
    }
    m.jt = thread_tile_index(threadIdx.x, mask, a.M);
    m.gt = base + tile_to_global(m.jt, a);
}

// LDS slot of tile element j during a phase exchange: XOR-folds the higher
// index bits into the low bits that select the LDS bank row (256 B), so lanes
// whose tile indices differ only above the register bits do not collide.
template <class R>
__device__ __forceinline__ uint32_t lds_slot(uint32_t j) {
    constexpr uint32_t B = sizeof(cx<R>) == 16 ? 4 : 5;  // log2(elements per 256 B)
    constexpr uint32_t MASK = (1u << B) - 1;
    return j ^ (((j >> B) ^ (j >> (2 * B))) & MASK);
}

template <class R>
__device__ __forceinline__ cx<R> cm(R mr, R mi, cx<R> v) {
    return cx<R>{mr * v.re - mi * v.im, mr * v.im + mi * v.re};
}

// 2x2 on register bit B.  FORM selects the arithmetic by the matrix's
// structure (MAT_* in qf_plan.h): general complex (16 flop per amplitude),
// real (8), "cross" = real diagonal + imaginary off-diagonal such as Rx (8),
// or the X permutation (0).  m holds (m00, m01, m10, m11) as re, im pairs.
template <class R, int RB, int B, int FORM>
__device__ __forceinline__ void reg_mat1(cx<R> (&a)[1 << RB], const R (&m)[8]) {
    if constexpr (B < RB) {
#pragma unroll
        for (int k = 0; k < (1 << RB); ++k) {
            if (k & (1 << B)) continue;
            const int k1 = k | (1 << B);
            const cx<R> a0 = a[k], a1 = a[k1];
            if constexpr (FORM == MAT_SWAP) {
                a[k] = a1;
                a[k1] = a0;
            } else if constexpr (FORM == MAT_REAL) {
                a[k] = cx<R>{m[0] * a0.re + m[2] * a1.re, m[0] * a0.im + m[2] * a1.im};
                a[k1] = cx<R>{m[4] * a0.re + m[6] * a1.re, m[4] * a0.im + m[6] * a1.im};
            } else if constexpr (FORM == MAT_CROSS) {  // m00, m11 real; m01 = i*m[3], m10 = i*m[5]
                a[k] = cx<R>{m[0] * a0.re - m[3] * a1.im, m[0] * a0.im + m[3] * a1.re};
                a[k1] = cx<R>{m[6] * a1.re - m[5] * a0.im, m[6] * a1.im + m[5] * a0.re};
            } else {
                const cx<R> x0 = cm(m[0], m[1], a0), x1 = cm(m[2], m[3], a1);
                const cx<R> y0 = cm(m[4], m[5], a0), y1 = cm(m[6], m[7], a1);
                a[k] = cx<R>{x0.re + x1.re, x0.im + x1.im};
                a[k1] = cx<R>{y0.re + y1.re, y0.im + y1.im};
            }
        }
    }
}

template <class R, int RB, int FORM>
__device__ __forceinline__ void reg_mat1_bit(cx<R> (&a)[1 << RB], int bit, const R (&m)[8]) {
    switch (bit) {
        case 0: reg_mat1<R, RB, 0, FORM>(a, m); break;
        case 1: reg_mat1<R, RB, 1, FORM>(a, m); break;
        case 2: reg_mat1<R, RB, 2, FORM>(a, m); break;
        case 3: reg_mat1<R, RB, 3, FORM>(a, m); break;
        default: reg_mat1<R, RB, 4, FORM>(a, m); break;
    }
}

template <class R, int RB>
__device__ __forceinline__ void reg_mat1_any(cx<R> (&a)[1 << RB], int bit, const R (&m)[8], int form = MAT_GENERAL) {
    switch (form) {
        case MAT_REAL: reg_mat1_bit<R, RB, MAT_REAL>(a, bit, m); break;
        case MAT_CROSS: reg_mat1_bit<R, RB, MAT_CROSS>(a, bit, m); break;
        case MAT_SWAP: reg_mat1_bit<R, RB, MAT_SWAP>(a, bit, m); break;
        default: reg_mat1_bit<R, RB, MAT_GENERAL>(a, bit, m); break;
    }
}

// CNOT with target register bit B and a control bit read from the global index.
template <class R, int RB, int B>
__device__ __forceinline__ void reg_cnot(cx<R> (&a)[1 << RB], const RegMap<RB>& mp, int ctrl) {
    if constexpr (B < RB) {
#pragma unroll
        for (int k = 0; k < (1 << RB); ++k) {
            if (k & (1 << B)) continue;
            const int k1 = k | (1 << B);
            const bool c = ((mp.gt | mp.gk[k]) >> ctrl) & 1;
            const cx<R> u = a[k], v = a[k1];
            a[k] = c ? v : u;
            a[k1] = c ? u : v;
        }
    }
}

template <class R, int RB>
__device__ __forceinline__ void reg_cnot_any(cx<R> (&a)[1 << RB], const RegMap<RB>& mp, int ctrl, int bit) {
    switch (bit) {
        case 0: reg_cnot<R, RB, 0>(a, mp, ctrl); break;
        case 1: reg_cnot<R, RB, 1>(a, mp, ctrl); break;
        case 2: reg_cnot<R, RB, 2>(a, mp, ctrl); break;
        case 3: reg_cnot<R, RB, 3>(a, mp, ctrl); break;
        default: reg_cnot<R, RB, 4>(a, mp, ctrl); break;
    }
}

template <class R, int RB, int I, int J>
__device__ __forceinline__ void reg_swap(cx<R> (&a)[1 << RB]) {
    if constexpr (I < RB && J < RB && I != J) {
#pragma unroll
        for (int k = 0; k < (1 << RB); ++k) {
            if (!((k >> I) & 1) || ((k >> J) & 1)) continue;
            const int k2 = k ^ (1 << I) ^ (1 << J);
            const cx<R> t = a[k];
            a[k] = a[k2];
            a[k2] = t;
        }
    }
}

template <class R, int RB>
__device__ __forceinline__ void reg_swap_any(cx<R> (&a)[1 << RB], int i, int j) {
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    switch (lo * 8 + hi) {
        case 0 * 8 + 1: reg_swap<R, RB, 0, 1>(a); break;
        case 0 * 8 + 2: reg_swap<R, RB, 0, 2>(a); break;
        case 0 * 8 + 3: reg_swap<R, RB, 0, 3>(a); break;
        case 0 * 8 + 4: reg_swap<R, RB, 0, 4>(a); break;
        case 1 * 8 + 2: reg_swap<R, RB, 1, 2>(a); break;
        case 1 * 8 + 3: reg_swap<R, RB, 1, 3>(a); break;
        case 1 * 8 + 4: reg_swap<R, RB, 1, 4>(a); break;
        case 2 * 8 + 3: reg_swap<R, RB, 2, 3>(a); break;
        case 2 * 8 + 4: reg_swap<R, RB, 2, 4>(a); break;
        default: reg_swap<R, RB, 3, 4>(a); break;
    }
}

// Dense 4x4 on register bits I < J: a_r' = sum_c m[r][c] a_c over the index
// r = bit(I) + 2 bit(J); m holds 16 complex entries (re, im) row major.
template <class R, int RB, int I, int J>
__device__ __forceinline__ void reg_mat2(cx<R> (&a)[1 << RB], const double* __restrict__ m) {
    if constexpr (I < RB && J < RB && I < J) {
        R mr[16], mi[16];
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            mr[e] = R(m[2 * e]);
            mi[e] = R(m[2 * e + 1]);
        }
#pragma unroll
        for (int k = 0; k < (1 << RB); ++k) {
            if ((k >> I) & 1 || (k >> J) & 1) continue;
            const int idx[4] = {k, k | (1 << I), k | (1 << J), k | (1 << I) | (1 << J)};
            cx<R> v[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) v[c] = a[idx[c]];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                cx<R> acc = cm(mr[4 * r], mi[4 * r], v[0]);
#pragma unroll
                for (int c = 1; c < 4; ++c) {
                    const cx<R> t = cm(mr[4 * r + c], mi[4 * r + c], v[c]);
                    acc.re += t.re;
                    acc.im += t.im;
                }
                a[idx[r]] = acc;
            }
        }
    }
}

template <class R, int RB>
__device__ __forceinline__ void reg_mat2_any(cx<R> (&a)[1 << RB], int i, int j, const double* __restrict__ m) {
    switch (i * 8 + j) {
        case 0 * 8 + 1: reg_mat2<R, RB, 0, 1>(a, m); break;
        case 0 * 8 + 2: reg_mat2<R, RB, 0, 2>(a, m); break;
        case 0 * 8 + 3: reg_mat2<R, RB, 0, 3>(a, m); break;
        case 0 * 8 + 4: reg_mat2<R, RB, 0, 4>(a, m); break;
        case 1 * 8 + 2: reg_mat2<R, RB, 1, 2>(a, m); break;
        case 1 * 8 + 3: reg_mat2<R, RB, 1, 3>(a, m); break;
        case 1 * 8 + 4: reg_mat2<R, RB, 1, 4>(a, m); break;
        case 2 * 8 + 3: reg_mat2<R, RB, 2, 3>(a, m); break;
        case 2 * 8 + 4: reg_mat2<R, RB, 2, 4>(a, m); break;
        default: reg_mat2<R, RB, 3, 4>(a, m); break;
    }
}

// Diagonal on global qubits q0 (, q1): a_k *= d[bit(q1)*2 + bit(q0)].
template <class R, int RB>
__device__ __forceinline__ void reg_diag(cx<R> (&a)[1 << RB], const RegMap<RB>& mp, const TileOp& op) {
    const R d0r = R(op.m[0]), d0i = R(op.m[1]), d1r = R(op.m[2]), d1i = R(op.m[3]);
    const R d2r = R(op.m[4]), d2i = R(op.m[5]), d3r = R(op.m[6]), d3i = R(op.m[7]);
    const int q0 = op.q0, q1 = op.q1;
#pragma unroll
    for (int k = 0; k < (1 << RB); ++k) {
        const uint64_t gi = mp.gt | mp.gk[k];
        const bool b0 = (gi >> q0) & 1;
        const bool b1 = q1 >= 0 ? ((gi >> q1) & 1) : false;
        const R dr = b1 ? (b0 ? d3r : d2r) : (b0 ? d1r : d0r);
        const R di = b1 ? (b0 ? d3i : d2i) : (b0 ? d1i : d0i);
        a[k] = cm(dr, di, a[k]);
    }
}

// Bit k: parity of global qubits q0 (and q1 if >= 0) in register k's index.
// The register offsets gk[k] are ORs of the single-bit offsets gk[1 << b]
// and share no bit with the thread's gt, so one parity of gt and one per
// register bit (uniform across the workgroup) give all 2^RB parities.
template <int RB>
__device__ __forceinline__ uint32_t reg_parities(const RegMap<RB>& mp, int q0, int q1) {
    const uint64_t mask = (uint64_t(1) << q0) | (q1 >= 0 ? uint64_t(1) << q1 : 0);
    uint32_t p = (__popcll(mp.gt & mask) & 1) ? (1u << (1 << RB)) - 1 : 0u;
#pragma unroll
    for (int b = 0; b < RB; ++b) {
        uint32_t pattern = 0;  // registers whose index has bit b set
#pragma unroll
        for (int k = 0; k < (1 << RB); ++k) pattern |= uint32_t((k >> b) & 1) << k;
        if (__popcll(mp.gk[1 << b] & mask) & 1) p ^= pattern;
    }
    return p;
}

// exp(-i theta Z/2) or exp(-i theta Z Z/2) from (c, s): c - is on even parity, c + is on odd.
template <class R, int RB>
__device__ __forceinline__ void reg_rotz(cx<R> (&a)[1 << RB], const RegMap<RB>& mp, int q0, int q1, R c, R s) {
    const uint32_t par = reg_parities<RB>(mp, q0, q1);
#pragma unroll
    for (int k = 0; k < (1 << RB); ++k) {
        const bool odd = (par >> k) & 1;
        a[k] = cm(c, odd ? s : -s, a[k]);
    }
}

// Thread-local Im <lambda| G |psi> for a rotation generator.
template <class R, int RB, int B>
__device__ __forceinline__ double reg_overlap_xy(const cx<R> (&psi)[1 << RB], const cx<R> (&lam)[1 << RB],
                                                 int axis) {
    double acc = 0.0;
    if constexpr (B < RB) {
#pragma unroll
        for (int k = 0; k < (1 << RB); ++k) {
            const int kp = k ^ (1 << B);
            const double lr = lam[k].re, li = lam[k].im, pr = psi[kp].re, pi = psi[kp].im;
            const double re = lr * pr + li * pi, im = lr * pi - li * pr;
            acc += axis == 0 ? im : ((k >> B) & 1 ? re : -re);
        }
    }
    return acc;
}

template <class R, int RB>
__device__ __forceinline__ double reg_overlap(const cx<R> (&psi)[1 << RB], const cx<R> (&lam)[1 << RB],
                                              const RegMap<RB>& mp, const TileOp& op) {
    if (op.kind == TOP_ROT && op.q1 != 2) {
        switch (op.q0) {
            case 0: return reg_overlap_xy<R, RB, 0>(psi, lam, op.q1);
            case 1: return reg_overlap_xy<R, RB, 1>(psi, lam, op.q1);
            case 2: return reg_overlap_xy<R, RB, 2>(psi, lam, op.q1);
            case 3: return reg_overlap_xy<R, RB, 3>(psi, lam, op.q1);
            default: return reg_overlap_xy<R, RB, 4>(psi, lam, op.q1);
        }
    }
    const int q0 = op.q0, q1 = op.kind == TOP_ROTZZ ? op.q1 : -1;
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < (1 << RB); ++k) {
        const uint64_t gi = mp.gt | mp.gk[k];
        const bool odd = (((gi >> q0) ^ (q1 >= 0 ? (gi >> q1) : 0)) & 1) != 0;
        const double im = double(lam[k].re) * double(psi[k].im) - double(lam[k].im) * double(psi[k].re);
        acc += odd ? -im : im;
    }
    return acc;
}

template <class R, int RB>
__device__ __forceinline__ void reg_apply(cx<R> (&a)[1 << RB], const RegMap<RB>& mp, const TileOp& op,
                                          const double* __restrict__ cs_row) {
    switch (op.kind) {
        case TOP_MAT1: {
            const R m[8] = {R(op.m[0]), R(op.m[1]), R(op.m[2]), R(op.m[3]),
                            R(op.m[4]), R(op.m[5]), R(op.m[6]), R(op.m[7])};
            reg_mat1_any<R, RB>(a, op.q0, m, op.flags);
            break;
        }
        case TOP_DIAG: reg_diag<R, RB>(a, mp, op); break;
        case TOP_CNOT: reg_cnot_any<R, RB>(a, mp, op.q0, op.q1); break;
        case TOP_SWAP: reg_swap_any<R, RB>(a, op.q0, op.q1); break;
        case TOP_MAT2: reg_mat2_any<R, RB>(a, op.q0, op.q1, op.m); break;
        case TOP_ROT:
        case TOP_ROTZZ: {
            R c, s;
            rot_cs<R>(op, cs_row, c, s);
            if (op.kind == TOP_ROT && op.q1 == 0) {  // Rx = [[c, -is], [-is, c]]
                const R m[8] = {c, R(0), R(0), -s, R(0), -s, c, R(0)};
                reg_mat1_any<R, RB>(a, op.q0, m, MAT_CROSS);
            } else if (op.kind == TOP_ROT && op.q1 == 1) {  // Ry = [[c, -s], [s, c]]
                const R m[8] = {c, R(0), -s, R(0), s, R(0), c, R(0)};
                reg_mat1_any<R, RB>(a, op.q0, m, MAT_REAL);
            } else {
                reg_rotz<R, RB>(a, mp, op.q0, op.kind == TOP_ROTZZ ? op.q1 : -1, c, s);
            }
            break;
        }
        default: break;
    }
}

// Recorded overlaps wait in LDS in batches of kRecBatch: each wave adds its
// share (DPP row operations, no LDS traffic, no barrier) into a row of red,
// and one pass over the batch with a single barrier writes the tile sums
// (waves added in order).
constexpr uint32_t kRecBatch = 32;
constexpr uint32_t kRedDoubles = kRecBatch * (kThreads / 64) + kRecBatch;

// One DPP move of a double (two 32-bit halves); lanes of rows outside
// ROWS (and lanes whose source is out of range) read 0.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_move(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, int(b), CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, int(b >> 32), CTRL, ROWS, 0xf, false);
    return __builtin_bit_cast(double, (long long)(unsigned)lo | ((long long)hi << 32));
}

// The wave's sum, in lane 63: quad permutes (swap neighbours, swap pairs),
// row half-mirror and mirror (every lane of a row holds its row's sum),
// then row_bcast:15 and row_bcast:31 carry rows into rows 1 and 3, 2 and 3.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_move<0xB1, 0xf>(v);
    v += dpp_move<0x4E, 0xf>(v);
    v += dpp_move<0x141, 0xf>(v);
    v += dpp_move<0x140, 0xf>(v);
    v += dpp_move<0x142, 0xa>(v);
    v += dpp_move<0x143, 0xc>(v);
    return v;
}

// Writes the pending records' tile sums (row r: waves' values, then the
// slot at red[kRecBatch * W + r]) and empties the batch.
__device__ __forceinline__ void flush_records(double* red, uint32_t& nrec, double* __restrict__ prow) {
    const uint32_t W = blockDim.x / 64;
    __syncthreads();
    for (uint32_t r = threadIdx.x; r < nrec; r += blockDim.x) {
        double t = 0.0;
        for (uint32_t w = 0; w < W; ++w) t += red[r * W + w];
        prow[uint32_t(red[kRecBatch * W + r])] = t;
    }
    __syncthreads();
    nrec = 0;
}

// One workgroup's tile: block index b (instance = b >> log_blocks).  ops[0]
// is always a TOP_PHASE header.  NS = 1: s0 is the state; NS = 2: s0 = psi,
// s1 = lambda, partial as k_tile_adjoint.  `xb` is the LDS exchange buffer
// (NS x 2^M amplitudes), `red` kRedDoubles doubles of LDS.
template <class R, int RB, int NS>
__device__ __forceinline__ void tile_reg_body(cx<R>* __restrict__ s0, cx<R>* __restrict__ s1, const TileArgs& args,
                                              const TileOp* __restrict__ ops, const double* __restrict__ cs,
                                              double* __restrict__ partial, uint64_t b, cx<R>* xb, double* red) {
    constexpr int NR = 1 << RB;
    const uint32_t T = 1u << args.M;
    const uint64_t inst = b >> args.log_blocks;
    const uint64_t bl = b & ((uint64_t(1) << args.log_blocks) - 1);
    uint64_t base = 0;
    {
        const uint32_t nh = args.M - args.low;
        uint32_t k = 0, bit = 0;
        for (uint32_t g = args.low; g < args.n; ++g) {
            if (k < nh && args.high[k] == g) {
                ++k;
                continue;
            }
            base |= ((bl >> bit) & 1) << g;
            ++bit;
        }
    }
    base += inst << args.n;
    const double* cs_row = cs ? cs + 2 * inst * args.nslots : nullptr;
    double* prow = (NS == 2) ? partial + (inst * (uint64_t(1) << args.log_blocks) + bl) * args.nslots : nullptr;

    RegMap<RB> mp;
    decode_map<RB>(uint32_t(ops[0].q0), args, base, mp);
    uint32_t nrec = 0;
    cx<R> a[NS][NR];
    double t[NR];       // NS = 2: Im(conj(lambda) psi) per register while t_ok
    bool t_ok = false;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        a[0][k] = s0[mp.gt | mp.gk[k]];
        if constexpr (NS == 2) a[1][k] = s1[mp.gt | mp.gk[k]];
    }
    for (uint32_t o = 1; o < args.nops; ++o) {
        const TileOp& op = ops[o];
        if (op.kind == TOP_PHASE) {
            __syncthreads();  // earlier readers of the exchange buffer are done
#pragma unroll
            for (int s = 0; s < NS; ++s)
#pragma unroll
                for (int k = 0; k < NR; ++k) xb[s * T + lds_slot<R>(mp.jt | mp.jk[k])] = a[s][k];
            decode_map<RB>(uint32_t(op.q0), args, base, mp);
            __syncthreads();
#pragma unroll
            for (int s = 0; s < NS; ++s)
#pragma unroll
                for (int k = 0; k < NR; ++k) a[s][k] = xb[s * T + lds_slot<R>(mp.jt | mp.jk[k])];
            t_ok = false;
            continue;
        }
        if constexpr (NS == 2) {
            if (op.kind == TOP_ROTZZ || (op.kind == TOP_ROT && op.q1 == 2)) {
                // Z / ZZ rotation of both states: one parity per register for
                // the record and both products.  Im(conj(lambda) psi) per
                // register is unchanged by diagonal unitaries applied to both
                // states, so a run of recorded Z rotations computes it once.
                R c, sn;
                rot_cs<R>(op, cs_row, c, sn);
                const bool rec = (op.flags & TF_RECORD) != 0;
                if (rec && !t_ok) {
#pragma unroll
                    for (int k = 0; k < NR; ++k)
                        t[k] = double(a[1][k].re) * double(a[0][k].im) - double(a[1][k].im) * double(a[0][k].re);
                    t_ok = true;
                }
                const uint32_t par = reg_parities<RB>(mp, op.q0, op.kind == TOP_ROTZZ ? op.q1 : -1);
                double acc = 0.0;
#pragma unroll
                for (int k = 0; k < NR; ++k) {
                    const bool odd = (par >> k) & 1;
                    acc += odd ? -t[k] : t[k];
                    a[0][k] = cm(c, odd ? sn : -sn, a[0][k]);
                    a[1][k] = cm(c, odd ? sn : -sn, a[1][k]);
                }
                if (rec) {
                    const double v = wave_sum(acc);
                    const uint32_t W = blockDim.x / 64;
                    if (threadIdx.x % 64 == 63) red[nrec * W + threadIdx.x / 64] = v;
                    if (threadIdx.x == 0) red[kRecBatch * W + nrec] = double(op.slot);
                    if (++nrec == kRecBatch) flush_records(red, nrec, prow);
                }
                continue;
            }
            t_ok = false;  // any other op changes the registers' overlaps
            if ((op.kind == TOP_ROT || op.kind == TOP_ROTZZ) && (op.flags & TF_RECORD)) {
                const double v = wave_sum(reg_overlap<R, RB>(a[0], a[1], mp, op));
                const uint32_t W = blockDim.x / 64;
                if (threadIdx.x % 64 == 63) red[nrec * W + threadIdx.x / 64] = v;
                if (threadIdx.x == 0) red[kRecBatch * W + nrec] = double(op.slot);
                if (++nrec == kRecBatch) flush_records(red, nrec, prow);
            }
        }
#pragma unroll
        for (int s = 0; s < NS; ++s) reg_apply<R, RB>(a[s], mp, op, cs_row);
    }
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        s0[mp.gt | mp.gk[k]] = a[0][k];
        if constexpr (NS == 2) s1[mp.gt | mp.gk[k]] = a[1][k];
    }
    if constexpr (NS == 2) {
        if (nrec) flush_records(red, nrec, prow);
    }
}

template <class R, int RB, int NS>
__global__ __launch_bounds__(kThreads) void k_tile_reg(cx<R>* __restrict__ s0, cx<R>* __restrict__ s1, TileArgs args,
                                                       const TileOp* __restrict__ ops, const double* __restrict__ cs,
                                                       double* __restrict__ partial) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double red[kRedDoubles];
    const uint64_t b = uint64_t(blockIdx.x) + uint64_t(blockIdx.y) * gridDim.x;
    tile_reg_body<R, RB, NS>(s0, s1, args, ops, cs, partial, b, reinterpret_cast<cx<R>*>(smem_raw), red);
}

}  // namespace qf
