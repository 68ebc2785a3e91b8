// Sparse Lanczos on the device (sparse_lanczos_gpu, src/quantum.rs:1853-2023,
// and the TFIM Lanczos of DMRG, src/tensor_network.rs:1535-1689 and
// 1692-2025).
//
// The reference keeps its Krylov vectors on the host and runs the
// reorthogonalization as one dot and one axpy per stored vector, each a
// separate pass (and, in the TFIM form, a separate launch plus a device sync).
// Here the whole basis stays in HBM, (m + 1) vectors of dim doubles, and one
// Lanczos step is:
//   1. the operator: w = A v_j - beta_{j-1} v_{j-1}, written straight into the
//      next basis slot (CSR SpMV with 1/4/16/64 lanes per row, or the
//      matrix-free TFIM kernel);
//   2. one multi-dot pass: c_k = v_k . w for every stored vector and |w|^2,
//      eight vectors per workgroup row, compensated FP64 partial sums reduced
//      in a fixed order (deterministic);
//   3. one fused pass: w -= sum_k c_k v_k and |w|^2.  If that norm fell below
//      1/sqrt(2) of the three-term remainder's (w - c_j v_j), steps 2-3 run
//      once more (classical Gram-Schmidt twice, as accurate as the
//      reference's modified Gram-Schmidt);
//   4. alpha_j = c_j, beta_j = |w|, v_{j+1} = w / beta_j.
// The tridiagonal model's lowest eigenvalue is found on the host by Sturm
// bisection.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "qf_runtime.h"

using namespace qfrt;
using qf::dd;

namespace {

constexpr uint32_t kMaxSteps = 100;
constexpr uint32_t kGroup = 8;          // vectors per multi-dot workgroup row
constexpr uint32_t kDotBlocks = 512;    // workgroups per vector in a multi-dot
constexpr uint32_t kNormBlocks = 1024;  // workgroups of the update pass

__device__ __forceinline__ uint64_t tid() { return uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; }

// Sum of kGroup compensated accumulators over the workgroup into thread 0.
__device__ __forceinline__ void reduce_group(dd (&acc)[kGroup]) {
    __shared__ dd red[kGroup][qf::kThreads / 64];
#pragma unroll
    for (uint32_t t = 0; t < kGroup; ++t)
        for (int m = 32; m >= 1; m >>= 1)
            acc[t] = qf::dd_add(acc[t], qf::dd{__shfl_xor(acc[t].s, m, 64), __shfl_xor(acc[t].c, m, 64)});
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0)
        for (uint32_t t = 0; t < kGroup; ++t) red[t][wave] = acc[t];
    __syncthreads();
    if (threadIdx.x == 0)
        for (uint32_t t = 0; t < kGroup; ++t)
            for (int w = 1; w < int(blockDim.x / 64); ++w) acc[t] = qf::dd_add(acc[t], red[t][w]);
}

// Lanczos start vector before normalization: the affine LCG seed_{k+1} =
// seed_k * 6364136223846793005 + 1 from seed 42, x_k = seed_{k+1} / 2^64 - 0.5.
// Each thread handles 64 consecutive entries and jumps to its first seed by
// composing the affine map (square and multiply over the entry index).
__global__ __launch_bounds__(qf::kThreads) void k_lanczos_start(double* __restrict__ v, uint64_t dim) {
    const uint64_t first = tid() * 64;
    if (first >= dim) return;
    const uint64_t A = 6364136223846793005ull, Cc = 1ull;
    // (a, c) of the map x -> a x + c applied `first` times.
    uint64_t ra = 1, rc = 0, pa = A, pc = Cc, k = first;
    while (k) {
        if (k & 1) {
            rc = pa * rc + pc;
            ra = pa * ra;
        }
        pc = pa * pc + pc;
        pa = pa * pa;
        k >>= 1;
    }
    uint64_t seed = ra * 42ull + rc;
    const double inv = double(UINT64_MAX);
    const uint64_t end = first + 64 < dim ? first + 64 : dim;
    for (uint64_t i = first; i < end; ++i) {
        seed = seed * A + Cc;
        v[i] = double(seed) / inv - 0.5;
    }
}

// w = A v - beta * vprev over a CSR operator, LANES lanes per row.
template <int LANES>
__global__ __launch_bounds__(qf::kThreads) void k_csr_spmv(const uint32_t* __restrict__ rows,
                                                           const uint32_t* __restrict__ cols,
                                                           const double* __restrict__ vals,
                                                           const double* __restrict__ v, double* __restrict__ w,
                                                           uint32_t dim, const double* __restrict__ vprev,
                                                           double beta) {
#pragma clang fp contract(off)
    const uint64_t t = tid();
    const uint64_t row = t / LANES;
    const int lane = int(t % LANES);
    double s = 0.0;
    if (row < dim)
        for (uint32_t k = rows[row] + uint32_t(lane); k < rows[row + 1]; k += LANES) s += vals[k] * v[cols[k]];
    if constexpr (LANES > 1)
        for (int off = LANES / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, LANES);
    if (row < dim && lane == 0) w[row] = vprev ? s - beta * vprev[row] : s;
}

// w = H v - beta * vprev for the open TFIM chain H = -sum Z_i Z_{i+1} - sum
// X_i (site i = bit n-1-i; tfim_matvec, src/tensor_network.rs:1702-1718): the
// ZZ part is (2 * #unequal neighbour bits - (n - 1)) v[a], the X part
// subtracts v[a ^ 2^(n-1-i)] for i = 0..n-1 in that order.
__global__ __launch_bounds__(qf::kThreads) void k_tfim(const double* __restrict__ v, double* __restrict__ w,
                                                       uint32_t n, uint64_t dim, const double* __restrict__ vprev,
                                                       double beta) {
#pragma clang fp contract(off)
    const uint64_t mask = (uint64_t(1) << (n - 1)) - 1;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t a = tid(); a < dim; a += stride) {
        const int differ = __popcll((a ^ (a >> 1)) & mask);
        double r = double(2 * differ - int(n - 1)) * v[a];
        for (uint32_t i = 0; i < n; ++i) r -= v[a ^ (uint64_t(1) << (n - 1 - i))];
        w[a] = vprev ? r - beta * vprev[a] : r;
    }
}

// The same operator in two passes for chains of kTfimSplit sites or more,
// whose high-bit flip partners miss the caches in the one-pass kernel.
// k_tfim_high stages 2^(kTfimHigh + kTfimRun) entries in LDS (runs of
// 2^kTfimRun contiguous entries, every combination of the top kTfimHigh bits)
// and writes the diagonal term minus the partners across those bits;
// k_tfim_low stages 2^kTfimLow contiguous entries and subtracts the partners
// across the bits between (read through the XCD's L2: each XCD walks its own
// contiguous share of the vector), then those across the low kTfimLow bits
// from LDS, then beta * vprev.  The subtractions run in the one-pass kernel's
// order (bits n-1 down to 0), so w is the same to the bit.
constexpr uint32_t kTfimRun = 4, kTfimHigh = 9, kTfimLow = 13, kTfimSplit = 14;
// 1024-thread workgroups: two share a CU beside their 64 KiB tiles, 32 waves.
constexpr uint32_t kTfimThreads = 1024;

__global__ __launch_bounds__(kTfimThreads) void k_tfim_high(const double* __restrict__ v, double* __restrict__ w,
                                                            uint32_t n) {
#pragma clang fp contract(off)
    constexpr uint32_t T = kTfimHigh + kTfimRun, E = 1u << T;
    __shared__ double xs[E];
    const uint64_t mid = blockIdx.x;  // index bits kTfimRun .. n - kTfimHigh - 1
    const uint32_t hs = n - kTfimHigh;
    auto gidx = [&](uint32_t e) {
        return (uint64_t(e >> kTfimRun) << hs) | (mid << kTfimRun) | uint64_t(e & ((1u << kTfimRun) - 1));
    };
    for (uint32_t e = threadIdx.x; e < E; e += blockDim.x) xs[e] = v[gidx(e)];
    __syncthreads();
    const uint64_t mask = (uint64_t(1) << (n - 1)) - 1;
    for (uint32_t e = threadIdx.x; e < E; e += blockDim.x) {
        const uint64_t a = gidx(e);
        const int differ = __popcll((a ^ (a >> 1)) & mask);
        double r = double(2 * differ - int(n - 1)) * xs[e];
        for (uint32_t i = 0; i < kTfimHigh; ++i) r -= xs[e ^ (1u << (T - 1 - i))];
        w[a] = r;
    }
}

// lo = min(kTfimLow, n - kTfimHigh) staged bits; the grid (2^(n - lo)
// workgroups) is a multiple of 8.
__global__ __launch_bounds__(kTfimThreads) void k_tfim_low(const double* __restrict__ v, double* __restrict__ w,
                                                           uint32_t n, uint32_t lo, const double* __restrict__ vprev,
                                                           double beta) {
#pragma clang fp contract(off)
    __shared__ double xs[1u << kTfimLow];
    const uint32_t E = 1u << lo;
    const uint64_t per = gridDim.x >> 3;
    const uint64_t base = (uint64_t(blockIdx.x & 7) * per + (blockIdx.x >> 3)) << lo;
    for (uint32_t e = threadIdx.x; e < E; e += blockDim.x) xs[e] = v[base + e];
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < E; e += blockDim.x) {
        const uint64_t a = base + e;
        double r = w[a];
        for (uint32_t b = n - kTfimHigh; b-- > lo;) r -= v[a ^ (uint64_t(1) << b)];
        for (uint32_t b = lo; b-- > 0;) r -= xs[e ^ (1u << b)];
        w[a] = vprev ? r - beta * vprev[a] : r;
    }
}

// partial[k * gridDim.x + blockIdx.x] = sum_i V[k][i] * w[i] for the vectors
// k of workgroup row blockIdx.y (k == nv: sum_i w[i]^2).
__global__ __launch_bounds__(qf::kThreads) void k_multidot(const double* __restrict__ V, uint64_t ld, uint32_t nv,
                                                           const double* __restrict__ w, uint64_t dim,
                                                           dd* __restrict__ partial) {
    const uint32_t k0 = blockIdx.y * kGroup;
    dd acc[kGroup];
#pragma unroll
    for (uint32_t t = 0; t < kGroup; ++t) acc[t] = dd{0.0, 0.0};
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = tid(); i < dim; i += stride) {
        const double x = w[i];
#pragma unroll
        for (uint32_t t = 0; t < kGroup; ++t) {
            const uint32_t k = k0 + t;
            if (k < nv) qf::neumaier(acc[t], V[k * ld + i] * x);
            else if (k == nv) qf::neumaier(acc[t], x * x);
        }
    }
    reduce_group(acc);
    if (threadIdx.x == 0)
        for (uint32_t t = 0; t < kGroup; ++t)
            if (k0 + t <= nv) partial[uint64_t(k0 + t) * gridDim.x + blockIdx.x] = acc[t];
}

// out[k] = the fixed-order sum of vector k's nblocks partials (one workgroup per vector).
__global__ __launch_bounds__(qf::kThreads) void k_finish(const dd* __restrict__ partial, uint32_t nblocks,
                                                         double* __restrict__ out) {
    dd acc[kGroup];
    for (uint32_t t = 0; t < kGroup; ++t) acc[t] = dd{0.0, 0.0};
    for (uint32_t j = threadIdx.x; j < nblocks; j += blockDim.x)
        acc[0] = qf::dd_add(acc[0], partial[uint64_t(blockIdx.x) * nblocks + j]);
    reduce_group(acc);
    if (threadIdx.x == 0) out[blockIdx.x] = acc[0].s + acc[0].c;
}

// w -= sum_k c[k] V[k] (k in order), partial[blockIdx.x] = the workgroup's sum of w^2.
__global__ __launch_bounds__(qf::kThreads) void k_update(const double* __restrict__ V, uint64_t ld, uint32_t nv,
                                                         const double* __restrict__ c, double* __restrict__ w,
                                                         uint64_t dim, dd* __restrict__ partial) {
#pragma clang fp contract(off)
    __shared__ double cs[kMaxSteps + 1];
    for (uint32_t k = threadIdx.x; k < nv; k += blockDim.x) cs[k] = c[k];
    __syncthreads();
    dd acc[kGroup];
    for (uint32_t t = 0; t < kGroup; ++t) acc[t] = dd{0.0, 0.0};
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = tid(); i < dim; i += stride) {
        double x = w[i];
        for (uint32_t k = 0; k < nv; ++k) x -= cs[k] * V[k * ld + i];
        w[i] = x;
        qf::neumaier(acc[0], x * x);
    }
    reduce_group(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc[0];
}

__global__ __launch_bounds__(qf::kThreads) void k_divide(double* __restrict__ w, uint64_t dim, double b) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = tid(); i < dim; i += stride) w[i] = w[i] / b;
}

// Lowest eigenvalue of the symmetric tridiagonal matrix (a on the diagonal,
// b[i] between rows i and i+1) by Sturm-sequence bisection (Golub and Van
// Loan, Matrix Computations, sec. 8.4): count(x) = number of negative pivots
// of T - x I; the smallest x with count(x) >= 1 is found to the last bits.
double tridiagonal_lowest(const std::vector<double>& a, const std::vector<double>& b) {
    const size_t n = a.size();
    double lo = a[0], hi = a[0];
    for (size_t i = 0; i < n; ++i) {
        const double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(b[i]) : 0.0);
        lo = std::min(lo, a[i] - r);
        hi = std::max(hi, a[i] + r);
    }
    auto below = [&](double x) {
        size_t count = 0;
        double d = a[0] - x;
        const double tiny = DBL_MIN * 4.0;
        for (size_t i = 0; i < n; ++i) {
            if (i > 0) d = a[i] - x - b[i - 1] * b[i - 1] / d;
            if (d == 0.0) d = -tiny;
            if (d < 0.0) ++count;
        }
        return count;
    };
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (below(mid) >= 1) hi = mid;
        else lo = mid;
    }
    return 0.5 * (lo + hi);
}

struct Work {
    double* V = nullptr;   // (m + 1) x dim basis
    dd* partial = nullptr;  // multi-dot and norm partials
    double* out = nullptr;  // reduced dots
    ~Work() {
        if (V) (void)hipFree(V);
        if (partial) (void)hipFree(partial);
        if (out) (void)hipFree(out);
    }
};

// One Lanczos run; `op(v, w, vprev, beta)` launches the operator.  tfim
// selects the TFIM routine's stopping rule and model size.
template <class Op>
qf_status lanczos(qf_backend* be, uint64_t dim, size_t m, bool tfim, double tolerance, Op op, double* lowest,
                  double* alpha_out, double* beta_out, uint64_t* size_out) {
    Work wk;
    const size_t vbytes = size_t(dim) * sizeof(double);
    qf_status s = device_alloc(reinterpret_cast<void**>(&wk.V), vbytes * (m + 1));
    if (s) return s;
    // Multi-dot partials of up to kMaxSteps + 1 vectors, then the update pass's norm partials.
    if ((s = device_alloc(reinterpret_cast<void**>(&wk.partial),
                          sizeof(dd) * ((kMaxSteps + 2) * size_t(kDotBlocks) + kNormBlocks))))
        return s;
    if ((s = device_alloc(reinterpret_cast<void**>(&wk.out), sizeof(double) * (kMaxSteps + 2) * 2))) return s;
    hipStream_t st = be->stream;
    const unsigned gstart = unsigned((dim + 64 * qf::kThreads - 1) / (64 * qf::kThreads));
    const unsigned gvec = std::min<unsigned>(grid_for(dim), 8192);
    const unsigned gnorm = std::min<unsigned>(gvec, kNormBlocks);
    const unsigned gdot = std::min<unsigned>(gvec, kDotBlocks);

    // |w|^2 of basis slot k over all of it; out[slot] receives the sum.
    auto norm2 = [&](double* w, double* host) -> qf_status {
        hipLaunchKernelGGL(k_multidot, dim3(gdot, 1), dim3(qf::kThreads), 0, st, wk.V, dim, 0u, w, dim, wk.partial);
        hipLaunchKernelGGL(k_finish, dim3(1), dim3(qf::kThreads), 0, st, wk.partial, gdot, wk.out);
        qf_status r = launch_check("k_multidot");
        if (r) return r;
        QF_HIP_TRY(hipMemcpyAsync(host, wk.out, sizeof(double), hipMemcpyDeviceToHost, st));
        QF_HIP_TRY(hipStreamSynchronize(st));
        return QF_OK;
    };
    // Start vector, normalized.
    hipLaunchKernelGGL(k_lanczos_start, dim3(gstart), dim3(qf::kThreads), 0, st, wk.V, dim);
    double n2 = 0.0;
    if ((s = norm2(wk.V, &n2))) return s;
    hipLaunchKernelGGL(k_divide, dim3(gvec), dim3(qf::kThreads), 0, st, wk.V, dim, std::sqrt(n2));

    std::vector<double> alpha(m, 0.0), beta(m, 0.0), host(kMaxSteps + 2);
    for (size_t j = 0; j < m; ++j) {
        double* v = wk.V + j * dim;
        double* w = wk.V + (j + 1) * dim;
        op(v, w, j > 0 ? wk.V + (j - 1) * dim : nullptr, j > 0 ? beta[j - 1] : 0.0);
        if ((s = launch_check("Lanczos operator"))) return s;
        const uint32_t nv = uint32_t(j + 1);
        double a = 0.0, b2 = 0.0;
        for (int pass = 0; pass < 2; ++pass) {
            hipLaunchKernelGGL(k_multidot, dim3(gdot, (nv + kGroup) / kGroup), dim3(qf::kThreads), 0, st, wk.V, dim,
                               nv, w, dim, wk.partial);
            hipLaunchKernelGGL(k_finish, dim3(nv + 1), dim3(qf::kThreads), 0, st, wk.partial, gdot, wk.out);
            hipLaunchKernelGGL(k_update, dim3(gnorm), dim3(qf::kThreads), 0, st, wk.V, dim, nv, wk.out, w, dim,
                               wk.partial + size_t(nv + 1) * gdot);
            hipLaunchKernelGGL(k_finish, dim3(1), dim3(qf::kThreads), 0, st, wk.partial + size_t(nv + 1) * gdot,
                               gnorm, wk.out + nv + 1);
            if ((s = launch_check("Lanczos reorthogonalization"))) return s;
            QF_HIP_TRY(hipMemcpyAsync(host.data(), wk.out, sizeof(double) * (nv + 2), hipMemcpyDeviceToHost, st));
            QF_HIP_TRY(hipStreamSynchronize(st));
            a += host[j];
            // The second pass is for cancellation against the basis: compare
            // with the norm of the three-term remainder w - c_j v_j
            // (|w|^2 - c_j^2 for a unit v_j), not with |w|, whose drop by
            // alpha_j v_j is the recurrence itself.
            const double before = host[nv] - host[j] * host[j], after = host[nv + 1];
            b2 = after;
            if (!(after < 0.5 * before)) break;
        }
        const double b = std::sqrt(b2);
        alpha[j] = a;
        if (tfim || j + 1 < m) beta[j] = b;
        if (tfim ? b < 1e-12 : (b < tolerance || j + 1 == m)) break;
        hipLaunchKernelGGL(k_divide, dim3(gvec), dim3(qf::kThreads), 0, st, w, dim, b);
    }
    QF_HIP_TRY(hipStreamSynchronize(st));
    size_t size = m;
    if (!tfim)
        for (size_t j = 0; j < m; ++j)
            if (alpha[j] == 0.0) {
                size = j;
                break;
            }
    if (alpha_out) std::copy(alpha.begin(), alpha.end(), alpha_out);
    if (beta_out) std::copy(beta.begin(), beta.end(), beta_out);
    if (size_out) *size_out = size;
    if (size == 0) return fail(QF_INVALID_ARGUMENT, "the Lanczos model is empty (first alpha is exactly 0)");
    std::vector<double> ta(alpha.begin(), alpha.begin() + size), tb(beta.begin(), beta.begin() + size);
    *lowest = tridiagonal_lowest(ta, tb);
    return QF_OK;
}

}  // namespace

extern "C" qf_status qf_sparse_lanczos(qf_backend* be, const uint64_t* row_offsets, const uint64_t* column_indices,
                                      const qf_c128* values, uint64_t dim, uint64_t max_iterations, double tolerance,
                                      double* lowest, double* alpha, double* beta, uint64_t* size) {
    if (!be || !row_offsets || !lowest) return fail(QF_INVALID_ARGUMENT, "null argument");
    if (dim == 0) return fail(QF_INVALID_ARGUMENT, "the operator has dimension zero");
    if (max_iterations == 0) return fail(QF_INVALID_ARGUMENT, "max_iterations must be positive");
    if (dim >= (uint64_t(1) << 31)) return fail(QF_STATE_TOO_LARGE, "dimension must be below 2^31");
    const uint64_t nnz = row_offsets[dim];
    if (nnz && (!values || !column_indices)) return fail(QF_INVALID_ARGUMENT, "null argument");
    if (nnz >= (uint64_t(1) << 32)) return fail(QF_STATE_TOO_LARGE, "at most 2^32 - 1 stored values");
    if (row_offsets[0] != 0) return fail(QF_INVALID_ARGUMENT, "row_offsets[0] must be 0");
    std::vector<uint32_t> rows(dim + 1), cols(nnz);
    std::vector<double> vals(nnz);
    for (uint64_t i = 0; i < dim; ++i) {
        if (row_offsets[i + 1] < row_offsets[i]) return fail(QF_INVALID_ARGUMENT, "row_offsets must not decrease");
        rows[i] = uint32_t(row_offsets[i]);
    }
    rows[dim] = uint32_t(nnz);
    for (uint64_t k = 0; k < nnz; ++k) {
        if (column_indices[k] >= dim) return fail(QF_INVALID_ARGUMENT, "column index out of range");
        cols[k] = uint32_t(column_indices[k]);
        vals[k] = values[k].re;
    }
    qf_status s = activate(be);
    if (s) return s;
    uint32_t* d_rows = nullptr;
    uint32_t* d_cols = nullptr;
    double* d_vals = nullptr;
    auto cleanup = [&](qf_status r) {
        (void)hipStreamSynchronize(be->stream);
        if (d_rows) (void)hipFree(d_rows);
        if (d_cols) (void)hipFree(d_cols);
        if (d_vals) (void)hipFree(d_vals);
        return r;
    };
    if ((s = device_alloc(reinterpret_cast<void**>(&d_rows), rows.size() * 4))) return cleanup(s);
    if ((s = device_alloc(reinterpret_cast<void**>(&d_cols), std::max<size_t>(1, cols.size()) * 4))) return cleanup(s);
    if ((s = device_alloc(reinterpret_cast<void**>(&d_vals), std::max<size_t>(1, vals.size()) * 8))) return cleanup(s);
    if (hipMemcpy(d_rows, rows.data(), rows.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        (nnz && hipMemcpy(d_cols, cols.data(), cols.size() * 4, hipMemcpyHostToDevice) != hipSuccess) ||
        (nnz && hipMemcpy(d_vals, vals.data(), vals.size() * 8, hipMemcpyHostToDevice) != hipSuccess))
        return cleanup(fail(QF_DEVICE_ERROR, "upload of the CSR operator failed"));
    const double avg = double(nnz) / double(dim);
    const int lanes = avg <= 3.0 ? 1 : (avg <= 12.0 ? 4 : (avg <= 48.0 ? 16 : 64));
    const uint32_t d32 = uint32_t(dim);
    auto op = [&](const double* v, double* w, const double* vp, double b) {
        const uint64_t threads = dim * uint64_t(lanes);
        const dim3 grid(unsigned((threads + qf::kThreads - 1) / qf::kThreads));
        switch (lanes) {
            case 1: hipLaunchKernelGGL(k_csr_spmv<1>, grid, dim3(qf::kThreads), 0, be->stream, d_rows, d_cols, d_vals, v, w, d32, vp, b); break;
            case 4: hipLaunchKernelGGL(k_csr_spmv<4>, grid, dim3(qf::kThreads), 0, be->stream, d_rows, d_cols, d_vals, v, w, d32, vp, b); break;
            case 16: hipLaunchKernelGGL(k_csr_spmv<16>, grid, dim3(qf::kThreads), 0, be->stream, d_rows, d_cols, d_vals, v, w, d32, vp, b); break;
            default: hipLaunchKernelGGL(k_csr_spmv<64>, grid, dim3(qf::kThreads), 0, be->stream, d_rows, d_cols, d_vals, v, w, d32, vp, b); break;
        }
    };
    const size_t m = size_t(std::min<uint64_t>(std::min<uint64_t>(max_iterations, dim), kMaxSteps));
    return cleanup(lanczos(be, dim, m, false, tolerance, op, lowest, alpha, beta, size));
}

extern "C" qf_status qf_tfim_lanczos(qf_backend* be, uint32_t sites, double* lowest, double* alpha, double* beta,
                                    uint64_t* size) {
    if (!be || !lowest) return fail(QF_INVALID_ARGUMENT, "null argument");
    if (sites < 2 || sites > 30) return fail(QF_INVALID_ARGUMENT, "the TFIM chain needs 2 to 30 sites");
    qf_status s = activate(be);
    if (s) return s;
    const uint64_t dim = uint64_t(1) << sites;
    // QF_TFIM_PASSES=1 keeps the one-pass kernel at every size (the two
    // forms give the same vectors to the bit).
    const char* passes = std::getenv("QF_TFIM_PASSES");
    const bool split = sites >= kTfimSplit && !(passes && std::strcmp(passes, "1") == 0);
    const uint32_t lo = std::min(kTfimLow, sites - kTfimHigh);
    auto op = [&](const double* v, double* w, const double* vp, double b) {
        if (split) {
            hipLaunchKernelGGL(k_tfim_high, dim3(unsigned(dim >> (kTfimHigh + kTfimRun))), dim3(kTfimThreads), 0,
                               be->stream, v, w, sites);
            hipLaunchKernelGGL(k_tfim_low, dim3(unsigned(dim >> lo)), dim3(kTfimThreads), 0, be->stream, v, w, sites,
                               lo, vp, b);
        } else {
            hipLaunchKernelGGL(k_tfim, dim3(grid_for(dim)), dim3(qf::kThreads), 0, be->stream, v, w, sites, dim, vp,
                               b);
        }
    };
    const size_t m = size_t(std::min<uint64_t>(sites >= 24 ? 30 : 100, dim));
    return lanczos(be, dim, m, true, 0.0, op, lowest, alpha, beta, size);
}
