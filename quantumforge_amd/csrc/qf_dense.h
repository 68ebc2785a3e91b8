// Dense-block passes on the matrix cores (plan kernel = 4).
//
// A pass applies one k-qubit block matrix U (2^k x 2^k complex, k = 4 or 5)
// to every column of 2^k amplitudes whose indices differ only in the block's
// qubits.  The complex product is one real product with the real embedding
//
//     [out_re]   [Ur  -Ui] [in_re]
//     [out_im] = [Ui   Ur] [in_im]        (2^(k+1) x 2^(k+1)),
//
// tiled into 16 x 16 x 4 matrix-core steps (v_mfma_f64_16x16x4f64 for
// complex128, v_mfma_f32_16x16x4f32 for complex64).  One wave takes 16
// columns at a time: lane l = 16 r + j holds column j (j = l mod 16) and the
// block indices b = r + 4 s of it (s = 0 .. 2^(k-2) - 1), which are exactly
// the B-operand entries (rows 4 ks + r of [in_re; in_im]) it supplies at
// step ks; accumulator i of its output tile t holds output row 16 t + r + 4 i
// (the FP64 instruction interleaves rows across lanes) or 16 t + 4 r + i (the
// FP32 one keeps blocks of four; both measured on gfx950), which it stores.
// The A operand (the embedding in lane order, 2^(k-3) x 2^(k-1) values per
// lane) is precomputed on the host and read from LDS.
//
// Diagonal gates ready at the start of the pass are applied to the loaded
// amplitudes first: factors of gates outside the block multiply the whole
// column (from the column's index bits), gates with one qubit in the block
// give a two-valued factor per block bit; the lane multiplies its
// amplitudes by the product, built by doubling over the block bits.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "qf_kernels.h"

namespace qf {

struct DenseDiag {
    uint32_t a, b;  // column list: two qubits outside the block; block list: a = block position, b = qubit
    uint32_t pad[2];
    double d[8];    // 4 complex factors, index bit(b) * 2 + bit(a)
};

struct DenseArgs {
    uint32_t n;          // qubits of the state
    uint32_t ncol, nblk; // diagonal gates per column / per block bit
    uint32_t pad;
    uint32_t S[5];       // block qubits, ascending
    uint32_t pad2[3];
    uint64_t groups;     // column groups of 16 (per state)
    uint64_t states;     // contiguous states (batches)
};

typedef double qf_double4 __attribute__((ext_vector_type(4)));
typedef float qf_float4 __attribute__((ext_vector_type(4)));

// row(r, i): the output row (within a 16-row tile) of accumulator i in the
// lanes 16 r .. 16 r + 15.
template <class R>
struct MfmaAcc;
template <>
struct MfmaAcc<double> {
    typedef qf_double4 T;
    __device__ static T step(double a, double b, T c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static uint32_t row(uint32_t r, uint32_t i) { return r + 4 * i; }
};
template <>
struct MfmaAcc<float> {
    typedef qf_float4 T;
    __device__ static T step(float a, float b, T c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static uint32_t row(uint32_t r, uint32_t i) { return 4 * r + i; }
};

__device__ __forceinline__ cx<double> dmul(cx<double> x, cx<double> y) {
    return cx<double>{x.re * y.re - x.im * y.im, x.re * y.im + x.im * y.re};
}

template <int K>
__device__ __forceinline__ uint64_t block_offset(uint32_t b, const uint32_t* S) {
    uint64_t o = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) o |= uint64_t((b >> i) & 1) << S[i];
    return o;
}

// aop: [NT][NKS][64] values of the A operand in lane order, copied to LDS
// (shared by the workgroup's waves).
template <class R, int K>
__global__ __launch_bounds__(256) void k_dense_block(cx<R>* __restrict__ state, DenseArgs args,
                                                     const R* __restrict__ aop, const DenseDiag* __restrict__ dcol,
                                                     const DenseDiag* __restrict__ dblk) {
    constexpr int NT = 1 << (K - 3);   // 16-row output tiles (half real, half imaginary)
    constexpr int NKS = 1 << (K - 1);  // K-steps of 4
    constexpr int NIN = 1 << (K - 2);  // amplitudes per lane
    using Acc = MfmaAcc<R>;
    __shared__ R Ash[NT * NKS * 64];
    for (int t = threadIdx.x; t < NT * NKS * 64; t += blockDim.x) Ash[t] = aop[t];
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63, j = lane & 15, r = lane >> 4;
    uint32_t S[K];
#pragma unroll
    for (int i = 0; i < K; ++i) S[i] = args.S[i];
    // Offsets of block indices: the lane's part (two block bits from r) plus a
    // part that is the same in every lane.
    const uint64_t in_lane = block_offset<K>(r, S);
    const uint64_t out_lane = block_offset<K>(Acc::row(r, 0), S);
    const uint64_t total = args.groups * args.states;
    const uint64_t waves = uint64_t(gridDim.x) * (blockDim.x >> 6);
    const uint64_t dim = uint64_t(1) << args.n;
    const bool diag = args.ncol + args.nblk > 0;
    for (uint64_t g = uint64_t(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6); g < total; g += waves) {
        const uint64_t inst = g / args.groups, gi = g % args.groups;
        // The A operand is read from LDS at every use: an opaque per-iteration
        // copy of the lane index keeps the compiler from hoisting the reads
        // into 2^(2k-4) registers (which would leave one wave per SIMD).
        uint32_t alane = lane;
        asm volatile("" : "+v"(alane));
        // column index -> base address: insert a zero bit at every block qubit
        uint64_t base = gi * 16 + j;
#pragma unroll
        for (int i = 0; i < K; ++i) base = ((base >> S[i]) << (S[i] + 1)) | (base & ((uint64_t(1) << S[i]) - 1));
        cx<R>* __restrict__ col = state + inst * dim + base;
        cx<R> a[NIN];
#pragma unroll
        for (int u = 0; u < NIN; ++u) a[u] = col[in_lane + block_offset<K>(4 * u, S)];
        if (diag) {
            cx<double> c{1.0, 0.0};
            for (uint32_t e = 0; e < args.ncol; ++e) {
                const uint32_t k2 = uint32_t(((base >> dcol[e].b) & 1) << 1) | uint32_t((base >> dcol[e].a) & 1);
                c = dmul(c, cx<double>{dcol[e].d[2 * k2], dcol[e].d[2 * k2 + 1]});
            }
            cx<double> f0[K], f1[K];
#pragma unroll
            for (int p = 0; p < K; ++p) f0[p] = f1[p] = cx<double>{1.0, 0.0};
            for (uint32_t e = 0; e < args.nblk; ++e) {
                const uint32_t pa = dblk[e].a;
                const uint32_t qb = uint32_t((base >> dblk[e].b) & 1) << 1;
                const cx<double> x0{dblk[e].d[2 * qb], dblk[e].d[2 * qb + 1]};
                const cx<double> x1{dblk[e].d[2 * qb + 2], dblk[e].d[2 * qb + 3]};
#pragma unroll
                for (int p = 0; p < K; ++p)
                    if (pa == uint32_t(p)) {
                        f0[p] = dmul(f0[p], x0);
                        f1[p] = dmul(f1[p], x1);
                    }
            }
            // lane bits 0, 1 of the block index are r; bits 2.. are u
            c = dmul(c, (r & 1) ? f1[0] : f0[0]);
            c = dmul(c, (r & 2) ? f1[1] : f0[1]);
            cx<double> P[NIN];
            P[0] = c;
#pragma unroll
            for (int bit = 0; bit < K - 2; ++bit)
#pragma unroll
                for (int u = 0; u < (1 << bit); ++u) {
                    P[u | (1 << bit)] = dmul(P[u], f1[bit + 2]);
                    P[u] = dmul(P[u], f0[bit + 2]);
                }
#pragma unroll
            for (int u = 0; u < NIN; ++u) {
                const cx<double> v = dmul(P[u], cx<double>{double(a[u].re), double(a[u].im)});
                a[u] = cx<R>{R(v.re), R(v.im)};
            }
        }
        typename Acc::T acc[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t] = typename Acc::T{R(0), R(0), R(0), R(0)};
#pragma unroll
        for (int ks = 0; ks < NKS; ++ks) {
            const R bv = ks < NKS / 2 ? a[ks].re : a[ks - NKS / 2].im;
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[t] = Acc::step(Ash[(t * NKS + ks) * 64 + alane], bv, acc[t]);
        }
        // output u: tile u / 4 (real) and NT / 2 + u / 4 (imaginary), accumulator u % 4
#pragma unroll
        for (int u = 0; u < NIN; ++u)
            col[out_lane + block_offset<K>(16 * (u >> 2) + Acc::row(0, u & 3), S)] =
                cx<R>{acc[u >> 2][u & 3], acc[NT / 2 + (u >> 2)][u & 3]};
    }
}

}  // namespace qf
