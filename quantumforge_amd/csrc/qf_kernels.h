// HIP kernels of the state-vector engine, written for gfx950 (CDNA4):
// 64-lane wavefronts, 256-thread workgroups (4 waves), LDS tiles.
//
// Amplitudes are interleaved {re, im} (the reference's repr(C) Complex64,
// src/lib.rs:437-445), indexed with 64-bit integers (the reference kernels
// use 32-bit `int` and stop at 30 qubits, src/cuda.rs:2256-2259).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "qf_plan.h"

namespace qf {

template <class R>
struct alignas(2 * sizeof(R)) cx {
    R re, im;
};

constexpr int kThreads = 256;

// Index of the p-th pair whose bit q is zero.
__device__ __forceinline__ uint64_t insert_zero(uint64_t p, uint32_t q) {
    const uint64_t lo = p & ((uint64_t(1) << q) - 1);
    return ((p ^ lo) << 1) | lo;
}

__device__ __forceinline__ uint64_t global_thread() {
    return uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
}

// ---------------------------------------------------------------------------
// Exact sweeps: one kernel per gate, reference arithmetic order, no FMA.
// ---------------------------------------------------------------------------

// new0 = m00*a0 + m01*a1, new1 = m10*a0 + m11*a1 with the complex product of
// src/lib.rs:486-494 and no contraction (StateVector::apply, src/lib.rs:1037-1039).
template <class R>
__global__ __launch_bounds__(kThreads) void k_sweep_1q(cx<R>* __restrict__ s, uint64_t npairs, uint32_t q,
                                                       cd m00, cd m01, cd m10, cd m11) {
#pragma clang fp contract(off)
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t p = global_thread(); p < npairs; p += stride) {
    const uint64_t i0 = insert_zero(p, q);
    const uint64_t i1 = i0 | (uint64_t(1) << q);
    const cx<R> a0 = s[i0], a1 = s[i1];
    const R ar0 = a0.re, ai0 = a0.im, ar1 = a1.re, ai1 = a1.im;
    const R b00r = R(m00.re), b00i = R(m00.im), b01r = R(m01.re), b01i = R(m01.im);
    const R b10r = R(m10.re), b10i = R(m10.im), b11r = R(m11.re), b11i = R(m11.im);
    R p0r = b00r * ar0 - b00i * ai0;
    R p0i = b00r * ai0 + b00i * ar0;
    R p1r = b01r * ar1 - b01i * ai1;
    R p1i = b01r * ai1 + b01i * ar1;
    cx<R> o0{p0r + p1r, p0i + p1i};
    R q0r = b10r * ar0 - b10i * ai0;
    R q0i = b10r * ai0 + b10i * ar0;
    R q1r = b11r * ar1 - b11i * ai1;
    R q1i = b11r * ai1 + b11i * ar1;
    cx<R> o1{q0r + q1r, q0i + q1i};
    s[i0] = o0;
    s[i1] = o1;
  }
}

// CNOT / CZ / SWAP over the 2^(n-2) quads (src/lib.rs:1062-1122).  Bit-exact:
// swaps move amplitudes, CZ multiplies by -1.0 exactly as src/lib.rs:1099.
template <class R>
__global__ __launch_bounds__(kThreads) void k_sweep_2q(cx<R>* __restrict__ s, uint64_t nquads, uint32_t a,
                                                       uint32_t b, int op) {
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t p = global_thread(); p < nquads; p += stride) {
    const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    const uint64_t base = insert_zero(insert_zero(p, lo), hi);
    const uint64_t ma = uint64_t(1) << a, mb = uint64_t(1) << b;
    if (op == 0) {  // CNOT(control a, target b)
        const cx<R> x = s[base | ma], y = s[base | ma | mb];
        s[base | ma] = y;
        s[base | ma | mb] = x;
    } else if (op == 1) {  // CZ
        cx<R> x = s[base | ma | mb];
        x.re = x.re * R(-1.0);
        x.im = x.im * R(-1.0);
        s[base | ma | mb] = x;
    } else {  // SWAP
        const cx<R> x = s[base | ma], y = s[base | mb];
        s[base | ma] = y;
        s[base | mb] = x;
    }
  }
}

// Diagonal on one or two qubits: a *= d[bit(q1)*2 + bit(q0)] (q1 < 0: one qubit).
template <class R>
__global__ __launch_bounds__(kThreads) void k_sweep_diag(cx<R>* __restrict__ s, uint64_t n, int q0, int q1,
                                                         cd d0, cd d1, cd d2, cd d3) {
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t i = global_thread(); i < n; i += stride) {
    const uint32_t k = uint32_t((i >> q0) & 1) | (q1 >= 0 ? uint32_t((i >> q1) & 1) << 1 : 0u);
    const cd d = k == 0 ? d0 : (k == 1 ? d1 : (k == 2 ? d2 : d3));
    const cx<R> a = s[i];
    const R dr = R(d.re), di = R(d.im);
    s[i] = cx<R>{dr * a.re - di * a.im, dr * a.im + di * a.re};
  }
}

// Projective collapse after measuring qubit q of each of the `total >>
// qubits` states: amplitudes whose bit q differs from the instance's outcome
// become 0, the others are divided by the instance's sqrt(probability)
// (StateVector::measure, src/lib.rs:1342-1348, which divides too).
template <class R>
__global__ __launch_bounds__(kThreads) void k_collapse(cx<R>* __restrict__ s, uint64_t total, uint32_t qubits,
                                                       uint32_t q, const uint8_t* __restrict__ outcomes,
                                                       const double* __restrict__ scales) {
#pragma clang fp contract(off)
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = global_thread(); i < total; i += stride) {
        const uint64_t inst = i >> qubits;
        const uint32_t bit = uint32_t((i >> q) & 1);
        cx<R> v = s[i];
        if (bit != outcomes[inst]) {
            v = cx<R>{R(0), R(0)};
        } else {
            const R sc = R(scales[inst]);
            v = cx<R>{v.re / sc, v.im / sc};
        }
        s[i] = v;
    }
}

// One-qubit gate on the states whose enabled[inst] is set (classically
// conditioned gates of a shot batch), with k_sweep_1q's unfused arithmetic.
template <class R>
__global__ __launch_bounds__(kThreads) void k_sweep_1q_masked(cx<R>* __restrict__ s, uint64_t npairs,
                                                              uint32_t qubits, uint32_t q,
                                                              const uint8_t* __restrict__ enabled, cd m00, cd m01,
                                                              cd m10, cd m11) {
#pragma clang fp contract(off)
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t p = global_thread(); p < npairs; p += stride) {
        const uint64_t i0 = insert_zero(p, q);
        if (!enabled[i0 >> qubits]) continue;
        const uint64_t i1 = i0 | (uint64_t(1) << q);
        const cx<R> a0 = s[i0], a1 = s[i1];
        const R ar0 = a0.re, ai0 = a0.im, ar1 = a1.re, ai1 = a1.im;
        const R b00r = R(m00.re), b00i = R(m00.im), b01r = R(m01.re), b01i = R(m01.im);
        const R b10r = R(m10.re), b10i = R(m10.im), b11r = R(m11.re), b11i = R(m11.im);
        R p0r = b00r * ar0 - b00i * ai0;
        R p0i = b00r * ai0 + b00i * ar0;
        R p1r = b01r * ar1 - b01i * ai1;
        R p1i = b01r * ai1 + b01i * ar1;
        R q0r = b10r * ar0 - b10i * ai0;
        R q0i = b10r * ai0 + b10i * ar0;
        R q1r = b11r * ar1 - b11i * ai1;
        R q1i = b11r * ai1 + b11i * ar1;
        s[i0] = cx<R>{p0r + p1r, p0i + p1i};
        s[i1] = cx<R>{q0r + q1r, q0i + q1i};
    }
}

// |0..0> for `count` states of 2^qubits amplitudes each (a fill kernel, so it
// may sit inside a captured graph).
template <class R>
__global__ __launch_bounds__(kThreads) void k_fill_zero_state(cx<R>* __restrict__ s, uint64_t total, uint32_t qubits) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = global_thread(); i < total; i += stride) {
        const bool first = (i & ((uint64_t(1) << qubits) - 1)) == 0;
        s[i] = cx<R>{first ? R(1) : R(0), R(0)};
    }
}

// In-place streaming probe over n 16-byte words (a state's amplitudes viewed as
// double2): every word is read once and written back multiplied by `scale`,
// the memory pattern of a tile pass without its arithmetic.  Variant 0: one
// word per thread per grid-stride step; 1: the same with nontemporal loads and
// stores; 2: four words per thread per step (64 bytes in flight per lane);
// 3: each workgroup streams one contiguous chunk of 2^chunk_log words (the
// grid covers the buffer once); 4: eight words per thread per step (128 bytes
// in flight per lane, the load depth of a 5-qubit dense-block pass).
template <int V>
__global__ __launch_bounds__(kThreads) void k_stream_probe(double2* __restrict__ s, uint64_t n, double scale,
                                                           uint32_t chunk_log) {
    if constexpr (V == 3) {
        const uint64_t base = uint64_t(blockIdx.x) << chunk_log, end = min(base + (uint64_t(1) << chunk_log), n);
        for (uint64_t i = base + threadIdx.x; i < end; i += blockDim.x) {
            double2 v = s[i];
            v.x *= scale;
            v.y *= scale;
            s[i] = v;
        }
    } else if constexpr (V == 2 || V == 4) {
        constexpr int W = V == 2 ? 4 : 8;
        const uint64_t stride = uint64_t(gridDim.x) * blockDim.x * W;
        for (uint64_t i = (uint64_t(blockIdx.x) * blockDim.x) * W + threadIdx.x; i < n; i += stride) {
            double2 v[W];
#pragma unroll
            for (int k = 0; k < W; ++k)
                if (i + uint64_t(k) * blockDim.x < n) v[k] = s[i + uint64_t(k) * blockDim.x];
#pragma unroll
            for (int k = 0; k < W; ++k)
                if (i + uint64_t(k) * blockDim.x < n) {
                    v[k].x *= scale;
                    v[k].y *= scale;
                    s[i + uint64_t(k) * blockDim.x] = v[k];
                }
        }
    } else {
        const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
        for (uint64_t i = global_thread(); i < n; i += stride) {
            double2 v;
            if constexpr (V == 1) {
                v.x = __builtin_nontemporal_load(&s[i].x);
                v.y = __builtin_nontemporal_load(&s[i].y);
            } else {
                v = s[i];
            }
            v.x *= scale;
            v.y *= scale;
            if constexpr (V == 1) {
                __builtin_nontemporal_store(v.x, &s[i].x);
                __builtin_nontemporal_store(v.y, &s[i].y);
            } else {
                s[i] = v;
            }
        }
    }
}

// out[k] = s[idx[k]] as complex128 (indices checked on the host).
template <class R>
__global__ __launch_bounds__(kThreads) void k_gather(const cx<R>* __restrict__ s, const uint64_t* __restrict__ idx,
                                                     cx<double>* __restrict__ out, uint64_t n) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t k = global_thread(); k < n; k += stride) {
        const cx<R> v = s[idx[k]];
        out[k] = cx<double>{double(v.re), double(v.im)};
    }
}

template <class A, class B>
__global__ __launch_bounds__(kThreads) void k_convert(const cx<A>* __restrict__ in, cx<B>* __restrict__ out,
                                                      uint64_t n) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = global_thread(); i < n; i += stride) out[i] = cx<B>{B(in[i].re), B(in[i].im)};
}

// ---------------------------------------------------------------------------
// Tile pass: load a 2^M-amplitude tile into LDS, apply a list of ops, store.
// ---------------------------------------------------------------------------
// Tile bit k < low is global qubit k (a contiguous run of 2^low amplitudes,
// so loads and stores are 2^low * sizeof(amplitude)-byte coalesced runs); tile
// bit low+k is global qubit high[k].  Workgroup b of instance `inst` takes the
// remaining n-M qubits from the bits of b.  Instances of a batch are
// contiguous 2^n-amplitude slabs.
struct TileArgs {
    uint32_t n;          // qubits per state
    uint32_t M;          // tile bits
    uint32_t low;        // contiguous low qubits
    uint32_t log_blocks; // n - M
    uint32_t nops;
    uint32_t nslots;     // columns of the per-instance angle table (rotation occurrences)
    uint32_t high[40];   // high[k] for k < M - low
};

template <class R>
__device__ __forceinline__ cx<R> cmul_d(const double* m, int k, cx<R> a) {
    const R mr = R(m[2 * k]), mi = R(m[2 * k + 1]);
    return cx<R>{mr * a.re - mi * a.im, mr * a.im + mi * a.re};
}

template <class R>
__device__ __forceinline__ cx<R> cmul_rr(R mr, R mi, cx<R> a) {
    return cx<R>{mr * a.re - mi * a.im, mr * a.im + mi * a.re};
}

// Workgroup geometry of a tile: global index of tile element j is
// base + hoff[j >> low] + (j & lowmask).
struct TileGeo {
    uint64_t inst, bl, base;
    uint32_t T, low, lowmask, nh;
};

__device__ __forceinline__ TileGeo tile_geometry(const TileArgs& args, uint64_t* hoff) {
    TileGeo t;
    t.nh = args.M - args.low;
    t.T = 1u << args.M;
    t.low = args.low;
    t.lowmask = (1u << args.low) - 1;
    const uint64_t b = uint64_t(blockIdx.x) + uint64_t(blockIdx.y) * gridDim.x;
    t.inst = b >> args.log_blocks;
    t.bl = b & ((uint64_t(1) << args.log_blocks) - 1);
    // Deposit the workgroup bits into the qubits outside the tile.
    uint64_t base = 0;
    uint32_t k = 0, bit = 0;
    for (uint32_t g = args.low; g < args.n; ++g) {
        if (k < t.nh && args.high[k] == g) {
            ++k;
            continue;
        }
        base |= ((t.bl >> bit) & 1) << g;
        ++bit;
    }
    t.base = base + (t.inst << args.n);
    for (uint32_t h = threadIdx.x; h < (1u << t.nh); h += blockDim.x) {
        uint64_t off = 0;
        for (uint32_t kk = 0; kk < t.nh; ++kk) off |= uint64_t((h >> kk) & 1) << args.high[kk];
        hoff[h] = off;
    }
    return t;
}

// Rotation exp(-i theta P/2) from the instance's (cos(theta/2), sin(theta/2));
// the inverse (TF_NEGATE) uses sin -> -sin.  Matrices as Gate::matrix,
// src/lib.rs:2303-2331.
template <class R>
__device__ __forceinline__ void rot_cs(const TileOp& op, const double* __restrict__ cs_row, R& c, R& s) {
    const double cc = cs_row[2 * op.slot], ss = cs_row[2 * op.slot + 1];
    c = R(cc);
    s = R((op.flags & TF_NEGATE) ? -ss : ss);
}

// Applies one op to a tile held in LDS (all threads participate; caller syncs).
template <class R>
__device__ __forceinline__ void tile_apply(cx<R>* __restrict__ tile, const TileOp& op, const TileGeo& g,
                                           const uint64_t* __restrict__ hoff, const double* __restrict__ cs_row) {
    const uint32_t T = g.T;
    const int kind = op.kind;
    if (kind == TOP_MAT1 || (kind == TOP_ROT && op.q1 != 2)) {
        const uint32_t q = uint32_t(op.q0);
        R m0r, m0i, m1r, m1i, m2r, m2i, m3r, m3i;
        if (kind == TOP_MAT1) {
            m0r = R(op.m[0]); m0i = R(op.m[1]); m1r = R(op.m[2]); m1i = R(op.m[3]);
            m2r = R(op.m[4]); m2i = R(op.m[5]); m3r = R(op.m[6]); m3i = R(op.m[7]);
        } else {
            R c, s;
            rot_cs<R>(op, cs_row, c, s);
            if (op.q1 == 0) {  // Rx = [[c, -is], [-is, c]]
                m0r = c; m0i = 0; m1r = 0; m1i = -s; m2r = 0; m2i = -s; m3r = c; m3i = 0;
            } else {           // Ry = [[c, -s], [s, c]]
                m0r = c; m0i = 0; m1r = -s; m1i = 0; m2r = s; m2i = 0; m3r = c; m3i = 0;
            }
        }
        for (uint32_t p = threadIdx.x; p < T / 2; p += blockDim.x) {
            const uint32_t lo = p & ((1u << q) - 1);
            const uint32_t i0 = ((p ^ lo) << 1) | lo, i1 = i0 | (1u << q);
            const cx<R> a0 = tile[i0], a1 = tile[i1];
            const cx<R> x0 = cmul_rr(m0r, m0i, a0), x1 = cmul_rr(m1r, m1i, a1);
            const cx<R> y0 = cmul_rr(m2r, m2i, a0), y1 = cmul_rr(m3r, m3i, a1);
            tile[i0] = cx<R>{x0.re + x1.re, x0.im + x1.im};
            tile[i1] = cx<R>{y0.re + y1.re, y0.im + y1.im};
        }
    } else if (kind == TOP_DIAG) {
        const int q0 = op.q0, q1 = op.q1;
        for (uint32_t j = threadIdx.x; j < T; j += blockDim.x) {
            const uint64_t gi = g.base + hoff[j >> g.low] + (j & g.lowmask);
            const int k = int((gi >> q0) & 1) | (q1 >= 0 ? int((gi >> q1) & 1) << 1 : 0);
            tile[j] = cmul_d(op.m, k, tile[j]);
        }
    } else if (kind == TOP_ROT || kind == TOP_ROTZZ) {  // Rz (diagonal) or ZZ rotation
        R c, s;
        rot_cs<R>(op, cs_row, c, s);
        const int q0 = op.q0, q1 = kind == TOP_ROTZZ ? op.q1 : -1;
        for (uint32_t j = threadIdx.x; j < T; j += blockDim.x) {
            const uint64_t gi = g.base + hoff[j >> g.low] + (j & g.lowmask);
            const uint32_t odd = uint32_t((gi >> q0) & 1) ^ (q1 >= 0 ? uint32_t((gi >> q1) & 1) : 0u);
            tile[j] = cmul_rr(c, odd ? s : -s, tile[j]);  // e^{-i theta/2} on +1, e^{+i theta/2} on -1
        }
    } else if (kind == TOP_CNOT || kind == TOP_SWAP) {
        const uint32_t a = uint32_t(op.q0), c = uint32_t(op.q1);
        const uint32_t lo = a < c ? a : c, hi = a < c ? c : a;
        for (uint32_t p = threadIdx.x; p < T / 4; p += blockDim.x) {
            uint32_t x = p;
            uint32_t l = x & ((1u << lo) - 1);
            x = ((x ^ l) << 1) | l;
            l = x & ((1u << hi) - 1);
            x = ((x ^ l) << 1) | l;
            const uint32_t ia = x | (1u << a), ic = x | (1u << c);
            if (kind == TOP_CNOT) {
                const cx<R> u = tile[ia], v = tile[ia | (1u << c)];
                tile[ia] = v;
                tile[ia | (1u << c)] = u;
            } else {
                const cx<R> u = tile[ia], v = tile[ic];
                tile[ia] = v;
                tile[ic] = u;
            }
        }
    } else {  // TOP_MAT2
        const uint32_t a = uint32_t(op.q0), c = uint32_t(op.q1);
        const uint32_t lo = a < c ? a : c, hi = a < c ? c : a;
        for (uint32_t p = threadIdx.x; p < T / 4; p += blockDim.x) {
            uint32_t x = p;
            uint32_t l = x & ((1u << lo) - 1);
            x = ((x ^ l) << 1) | l;
            l = x & ((1u << hi) - 1);
            x = ((x ^ l) << 1) | l;
            const uint32_t idx[4] = {x, x | (1u << a), x | (1u << c), x | (1u << a) | (1u << c)};
            cx<R> v[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) v[r] = tile[idx[r]];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                cx<R> acc{R(0), R(0)};
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const cx<R> t = cmul_d(op.m, r * 4 + k, v[k]);
                    acc.re += t.re;
                    acc.im += t.im;
                }
                tile[idx[r]] = acc;
            }
        }
    }
}

template <class R>
__global__ __launch_bounds__(kThreads) void k_tile(cx<R>* __restrict__ s, TileArgs args,
                                                   const TileOp* __restrict__ ops, const double* __restrict__ cs) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const uint32_t T = 1u << args.M;
    cx<R>* tile = reinterpret_cast<cx<R>*>(smem_raw);
    uint64_t* hoff = reinterpret_cast<uint64_t*>(smem_raw + sizeof(cx<R>) * T);
    const TileGeo g = tile_geometry(args, hoff);
    const double* cs_row = cs ? cs + 2 * g.inst * args.nslots : nullptr;
    __syncthreads();
    cx<R>* __restrict__ gp = s + g.base;
    for (uint32_t j = threadIdx.x; j < T; j += blockDim.x) tile[j] = gp[hoff[j >> g.low] + (j & g.lowmask)];
    __syncthreads();
    for (uint32_t o = 0; o < args.nops; ++o) {
        tile_apply<R>(tile, ops[o], g, hoff, cs_row);
        __syncthreads();
    }
    for (uint32_t j = threadIdx.x; j < T; j += blockDim.x) gp[hoff[j >> g.low] + (j & g.lowmask)] = tile[j];
}

// Sum of one double over the workgroup; every thread gets the result.
__device__ __forceinline__ double block_sum(double v, double* red) {
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < int(blockDim.x / 64); ++w) t += red[w];
    __syncthreads();
    return t;
}

// Im <lambda| G |psi> over one tile for the generator G of a rotation op
// (X / Y on a tile bit, Z on a global qubit, Z_u Z_v for TOP_ROTZZ).  FP64.
template <class R>
__device__ __forceinline__ double tile_overlap(const cx<R>* __restrict__ psi, const cx<R>* __restrict__ lam,
                                               const TileOp& op, const TileGeo& g, const uint64_t* __restrict__ hoff) {
    double acc = 0.0;
    const uint32_t T = g.T;
    if (op.kind == TOP_ROT && op.q1 != 2) {
        const uint32_t m = 1u << op.q0;
        for (uint32_t j = threadIdx.x; j < T; j += blockDim.x) {
            const cx<R> l = lam[j], p = psi[j ^ m];
            const double lr = l.re, li = l.im, pr = p.re, pi = p.im;
            const double re = lr * pr + li * pi;      // Re(conj(l) p)
            const double im = lr * pi - li * pr;      // Im(conj(l) p)
            if (op.q1 == 0)
                acc += im;                            // X: (X psi)_j = psi_{j^m}
            else
                acc += (j & m) ? re : -re;            // Y: (Y psi)_j = -i psi_{j|m} (bit 0), +i psi_{j&~m} (bit 1)
        }
    } else {
        const int q0 = op.q0, q1 = op.kind == TOP_ROTZZ ? op.q1 : -1;
        for (uint32_t j = threadIdx.x; j < T; j += blockDim.x) {
            const uint64_t gi = g.base + hoff[j >> g.low] + (j & g.lowmask);
            const uint32_t odd = uint32_t((gi >> q0) & 1) ^ (q1 >= 0 ? uint32_t((gi >> q1) & 1) : 0u);
            const cx<R> l = lam[j], p = psi[j];
            const double im = double(l.re) * double(p.im) - double(l.im) * double(p.re);
            acc += odd ? -im : im;
        }
    }
    return acc;
}

// Reverse sweep over (psi, lambda): for each op (already inverted and in
// reverse order), record Im <lambda|G|psi> for rotations flagged TF_RECORD
// into partial[(inst * 2^(n-M) + block) * nslots + slot], then apply the op
// to both states (psi is uncomputed, lambda is carried back).
template <class R>
__global__ __launch_bounds__(kThreads) void k_tile_adjoint(cx<R>* __restrict__ psi, cx<R>* __restrict__ lam,
                                                           TileArgs args, const TileOp* __restrict__ ops,
                                                           const double* __restrict__ cs,
                                                           double* __restrict__ partial) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double red[kThreads / 64];
    const uint32_t T = 1u << args.M;
    cx<R>* tp = reinterpret_cast<cx<R>*>(smem_raw);
    cx<R>* tl = tp + T;
    uint64_t* hoff = reinterpret_cast<uint64_t*>(smem_raw + 2 * sizeof(cx<R>) * T);
    const TileGeo g = tile_geometry(args, hoff);
    const double* cs_row = cs + 2 * g.inst * args.nslots;
    double* prow = partial + (g.inst * (uint64_t(1) << args.log_blocks) + g.bl) * args.nslots;
    __syncthreads();
    cx<R>* __restrict__ gp = psi + g.base;
    cx<R>* __restrict__ gl = lam + g.base;
    for (uint32_t j = threadIdx.x; j < T; j += blockDim.x) {
        const uint64_t off = hoff[j >> g.low] + (j & g.lowmask);
        tp[j] = gp[off];
        tl[j] = gl[off];
    }
    __syncthreads();
    for (uint32_t o = 0; o < args.nops; ++o) {
        const TileOp& op = ops[o];
        if ((op.kind == TOP_ROT || op.kind == TOP_ROTZZ) && (op.flags & TF_RECORD)) {
            const double v = block_sum(tile_overlap<R>(tp, tl, op, g, hoff), red);
            if (threadIdx.x == 0) prow[op.slot] = v;
        }
        tile_apply<R>(tp, op, g, hoff, cs_row);
        tile_apply<R>(tl, op, g, hoff, cs_row);
        __syncthreads();
    }
    for (uint32_t j = threadIdx.x; j < T; j += blockDim.x) {
        const uint64_t off = hoff[j >> g.low] + (j & g.lowmask);
        gp[off] = tp[j];
        gl[off] = tl[j];
    }
}

// ---------------------------------------------------------------------------
// FP64 compensated reductions (Neumaier per lane, two-sum tree across lanes).
// ---------------------------------------------------------------------------
struct dd {
    double s, c;  // value = s + c
};

__device__ __forceinline__ void neumaier(dd& a, double x) {
    const double t = a.s + x;
    if (fabs(a.s) >= fabs(x))
        a.c += (a.s - t) + x;
    else
        a.c += (x - t) + a.s;
    a.s = t;
}

__device__ __forceinline__ dd dd_add(dd a, dd b) {
    const double s = a.s + b.s;
    const double bb = s - a.s;
    const double e = (a.s - (s - bb)) + (b.s - bb);  // two-sum error
    return dd{s, a.c + b.c + e};
}

__device__ __forceinline__ dd shfl_xor_dd(dd v, int mask) {
    return dd{__shfl_xor(v.s, mask, 64), __shfl_xor(v.c, mask, 64)};
}

// Reduces (re, im) accumulators of the workgroup into lane 0 of wave 0.
__device__ __forceinline__ void block_reduce(dd& re, dd& im) {
    __shared__ dd red[2][kThreads / 64];
    for (int m = 32; m >= 1; m >>= 1) {
        re = dd_add(re, shfl_xor_dd(re, m));
        im = dd_add(im, shfl_xor_dd(im, m));
    }
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) {
        red[0][wave] = re;
        red[1][wave] = im;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < int(blockDim.x / 64); ++w) {
            re = dd_add(re, red[0][w]);
            im = dd_add(im, red[1][w]);
        }
    }
}

// RED_P1: sum of |a_i|^2 over the indices with bit q set (the probability of
// measuring 1 on qubit q, StateVector::measure, src/lib.rs:1328-1335).
// RED_NORM_P1: (sum |a_i|^2, the same over bit q set) in one pass (Kraus
// trajectories, src/lib.rs:1384-1395).
// Density matrices (dim = D = 2^n, entries row-major, one instance):
// RED_DTRACE (sum rho_rr), RED_DP1 (sum of Re rho_rr over rows with bit q
// set, and sum of +-Re rho_rr = <Z_q>), RED_DXY (sum of rho[r0][r0 | 2^q]
// over rows r0 with bit q clear: <X> = 2 Re, <Y> = -2 Im).
enum RedKind : int {
    RED_NORM = 0, RED_Z = 1, RED_XY = 2, RED_PAULI = 3, RED_INNER = 4, RED_P1 = 5,
    RED_NORM_P1 = 6, RED_DTRACE = 7, RED_DP1 = 8, RED_DXY = 9
};

// Device form of a Pauli term: c' = coefficient * i^(number of Y), sign from
// parity(index & z_mask); contribution conj(psi[i ^ x]) * c' * sign * psi[i].
struct PauliDev {
    double cr, ci;
    uint64_t x, z;
};

// partial[(inst * gridDim.x + block)] = (re, im) compensated partial sums.
template <class R>
__global__ __launch_bounds__(kThreads) void k_reduce(const cx<R>* __restrict__ s, const cx<R>* __restrict__ s2,
                                                     uint64_t dim, int kind, uint32_t q,
                                                     const PauliDev* __restrict__ terms, uint32_t nterms,
                                                     dd* __restrict__ partial) {
    const uint64_t inst = blockIdx.y;
    const cx<R>* __restrict__ a = s + inst * dim;
    dd re{0, 0}, im{0, 0};
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    if (kind == RED_NORM || kind == RED_Z || kind == RED_P1) {
        for (uint64_t i = global_thread(); i < dim; i += stride) {
            if (kind == RED_P1 && !((i >> q) & 1)) continue;
            const cx<R> v = a[i];
            const double x = double(v.re) * double(v.re) + double(v.im) * double(v.im);
            neumaier(re, (kind == RED_Z && ((i >> q) & 1)) ? -x : x);
        }
    } else if (kind == RED_NORM_P1) {
        for (uint64_t i = global_thread(); i < dim; i += stride) {
            const cx<R> v = a[i];
            const double x = double(v.re) * double(v.re) + double(v.im) * double(v.im);
            neumaier(re, x);
            if ((i >> q) & 1) neumaier(im, x);
        }
    } else if (kind == RED_DTRACE || kind == RED_DP1) {
        for (uint64_t r = global_thread(); r < dim; r += stride) {
            const cx<R> v = a[r * (dim + 1)];
            if (kind == RED_DTRACE) {
                neumaier(re, double(v.re));
                neumaier(im, double(v.im));
            } else {
                const bool one = (r >> q) & 1;
                if (one) neumaier(re, double(v.re));
                neumaier(im, one ? -double(v.re) : double(v.re));
            }
        }
    } else if (kind == RED_DXY) {
        for (uint64_t p = global_thread(); p < dim / 2; p += stride) {
            const uint64_t r0 = insert_zero(p, q);
            const cx<R> v = a[r0 * dim + (r0 | (uint64_t(1) << q))];
            neumaier(re, double(v.re));
            neumaier(im, double(v.im));
        }
    } else if (kind == RED_XY) {  // sum over bit-q-clear i of conj(a_i) a_{i|m}; x2 done on host
        for (uint64_t p = global_thread(); p < dim / 2; p += stride) {
            const uint64_t i0 = insert_zero(p, q), i1 = i0 | (uint64_t(1) << q);
            const cx<R> u = a[i0], v = a[i1];
            const double ur = u.re, ui = u.im, vr = v.re, vi = v.im;
            neumaier(re, ur * vr + ui * vi);
            neumaier(im, ur * vi - ui * vr);
        }
    } else if (kind == RED_PAULI) {
        for (uint64_t i = global_thread(); i < dim; i += stride) {
            const cx<R> v = a[i];
            const double vr = v.re, vi = v.im;
            for (uint32_t t = 0; t < nterms; ++t) {
                const PauliDev& P = terms[t];
                const cx<R> u = P.x ? a[i ^ P.x] : v;
                const double ur = u.re, ui = -double(u.im);  // conj(psi[i ^ x])
                const double sign = (__popcll(i & P.z) & 1) ? -1.0 : 1.0;
                const double wr = ur * vr - ui * vi, wi = ur * vi + ui * vr;
                neumaier(re, sign * (P.cr * wr - P.ci * wi));
                neumaier(im, sign * (P.cr * wi + P.ci * wr));
            }
        }
    } else {  // RED_INNER: sum conj(a_i) b_i
        const cx<R>* __restrict__ bvec = s2 + inst * dim;
        for (uint64_t i = global_thread(); i < dim; i += stride) {
            const cx<R> u = a[i], v = bvec[i];
            const double ur = u.re, ui = u.im, vr = v.re, vi = v.im;
            neumaier(re, ur * vr + ui * vi);
            neumaier(im, ur * vi - ui * vr);
        }
    }
    block_reduce(re, im);
    if (threadIdx.x == 0) {
        partial[2 * (inst * gridDim.x + blockIdx.x)] = re;
        partial[2 * (inst * gridDim.x + blockIdx.x) + 1] = im;
    }
}

// One workgroup per instance folds its `nblocks` partials; out[inst] = (re, im).
// (A template so the header can be included by several translation units.)
template <int Unused = 0>
__global__ __launch_bounds__(kThreads) void k_reduce_finish(const dd* __restrict__ partial, uint32_t nblocks,
                                                            double* __restrict__ out) {
    const uint64_t inst = blockIdx.x;
    dd re{0, 0}, im{0, 0};
    for (uint32_t k = threadIdx.x; k < nblocks; k += blockDim.x) {
        re = dd_add(re, partial[2 * (inst * nblocks + k)]);
        im = dd_add(im, partial[2 * (inst * nblocks + k) + 1]);
    }
    block_reduce(re, im);
    if (threadIdx.x == 0) {
        out[2 * inst] = re.s + re.c;
        out[2 * inst + 1] = im.s + im.c;
    }
}

// One Kraus-trajectory branch per state (StateVector::apply_channel_trajectory,
// src/lib.rs:1356-1505) on the pair (a0, a1) of qubit q; code per state:
//   0 nothing, 1 X, 2 Y, 3 Z (depolarizing branches, exact permutations and
//   negations of Gate::X/Y/Z applied as matrices),
//   4 a0 *= s0, a1 *= s1 (no jump),
//   5 a0 = a1 * s0, a1 = 0 (decay |1> -> |0>),
//   6 a1 = a0 * s0, a0 = 0 (excitation |0> -> |1>),
//   7 a0 = 0, a1 *= s0 (phase-damping jump).
// scales holds (s0, s1) per state.  Complex64 * f64 products as the CPU path.
template <class R>
__global__ __launch_bounds__(kThreads) void k_trajectory(cx<R>* __restrict__ s, uint64_t npairs, uint32_t qubits,
                                                         uint32_t q, const uint8_t* __restrict__ code,
                                                         const double* __restrict__ scales) {
#pragma clang fp contract(off)
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t p = global_thread(); p < npairs; p += stride) {
        const uint64_t inst = p >> (qubits - 1);
        const int c = code[inst];
        if (c == 0) continue;
        const uint64_t i0 = (inst << qubits) | insert_zero(p & ((uint64_t(1) << (qubits - 1)) - 1), q);
        const uint64_t i1 = i0 | (uint64_t(1) << q);
        const cx<R> a0 = s[i0], a1 = s[i1];
        const R s0 = R(scales[2 * inst]), s1 = R(scales[2 * inst + 1]);
        cx<R> b0 = a0, b1 = a1;
        switch (c) {
            case 1: b0 = a1; b1 = a0; break;
            case 2: b0 = cx<R>{a1.im, -a1.re}; b1 = cx<R>{-a0.im, a0.re}; break;
            case 3: b1 = cx<R>{-a1.re, -a1.im}; break;
            case 4: b0 = cx<R>{a0.re * s0, a0.im * s0}; b1 = cx<R>{a1.re * s1, a1.im * s1}; break;
            case 5: b0 = cx<R>{a1.re * s0, a1.im * s0}; b1 = cx<R>{R(0), R(0)}; break;
            case 6: b1 = cx<R>{a0.re * s0, a0.im * s0}; b0 = cx<R>{R(0), R(0)}; break;
            default: b0 = cx<R>{R(0), R(0)}; b1 = cx<R>{a1.re * s0, a1.im * s0}; break;
        }
        s[i0] = b0;
        s[i1] = b1;
    }
}

template <class R>
__global__ __launch_bounds__(kThreads) void k_probabilities(const cx<R>* __restrict__ s, double* __restrict__ out,
                                                            uint64_t n) {
#pragma clang fp contract(off)
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = global_thread(); i < n; i += stride) {
        const cx<R> v = s[i];
        // norm_sqr as src/lib.rs:459-461: re*re + im*im, unfused.
        out[i] = double(v.re) * double(v.re) + double(v.im) * double(v.im);
    }
}

}  // namespace qf
