// Uneven batches: adjoint gradients of many parameterized circuits of
// different widths (and structures) in one device schedule, plus the tile
// autotuner.
//
// Every instance gets its own forward and reverse plan.  Launch j of the
// forward sweep runs pass j of every instance at once: a grouped launch whose
// persistent grid walks a device-resident table of (instance, tile block)
// entries.  Because the tables live in device memory and the grids are fixed,
// the whole per-wave sequence (zero-fill, forward passes, costate + energy,
// reverse passes, gradient finish) is captured once into a hipGraph and
// replayed for every memory wave after its tables are uploaded.  Instances are
// packed into waves by their psi + lambda footprint against the free device
// memory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "qf_qaoa.h"
#include "qf_regtile.h"
#include "qf_runtime.h"

using qf::cd;
using qf::cx;
using namespace qfrt;

namespace {

struct GroupHeader {
    uint64_t total_blocks;
    uint32_t nentries;
    uint32_t pad;
};

struct TileEntry {
    uint64_t state_off;    // amplitude offset of the instance in the wave buffers
    uint64_t block_start;  // first block of this entry in the launch
    uint64_t ops_off;      // first TileOp of this pass
    uint64_t cs_off;       // angle row (doubles)
    uint64_t part_off;     // gradient partial rows (doubles)
    uint64_t nblocks;
    qf::TileArgs args;
};

struct PauliEntry {
    uint64_t state_off;
    uint64_t block_start;
    uint64_t terms_off;
    uint64_t epart_off;    // dd pairs
    uint32_t n;
    uint32_t nblocks;
    uint32_t nterms;
    uint32_t pad;
};

struct InstEntry {         // per-instance finishing work
    uint64_t state_off;
    uint64_t epart_off;    // dd pairs
    uint64_t part_off;     // doubles
    uint64_t grad_off;     // output gradient offset (global)
    uint64_t out_index;    // output energy index (global)
    uint64_t pcsr_off;     // offset of the instance's CSR arrays
    uint32_t eblocks;      // energy partial blocks
    uint32_t nblk;         // reverse-sweep tile blocks
    uint32_t nslots;
    uint32_t nparams;
};

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

template <class R, int RB, int NS>
__global__ __launch_bounds__(qf::kThreads) void k_tile_reg_grouped(cx<R>* __restrict__ psi, cx<R>* __restrict__ lam,
                                                                   const unsigned char* __restrict__ launch,
                                                                   const qf::TileOp* __restrict__ ops,
                                                                   const double* __restrict__ cs,
                                                                   double* __restrict__ partial) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double red[qf::kRedDoubles];
    const GroupHeader* h = reinterpret_cast<const GroupHeader*>(launch);
    const TileEntry* E = reinterpret_cast<const TileEntry*>(launch + sizeof(GroupHeader));
    const uint64_t total = h->total_blocks;
    const uint32_t ne = h->nentries;
    for (uint64_t b = blockIdx.x; b < total; b += gridDim.x) {
        const TileEntry& e = E[find_entry(E, ne, b)];
        qf::tile_reg_body<R, RB, NS>(psi + e.state_off, NS == 2 ? lam + e.state_off : nullptr, e.args, ops + e.ops_off,
                                     cs + e.cs_off, NS == 2 ? partial + e.part_off : nullptr, b - e.block_start,
                                     reinterpret_cast<cx<R>*>(smem_raw), red);
        __syncthreads();
    }
}

template <class R>
__global__ __launch_bounds__(qf::kThreads) void k_zero_grouped(cx<R>* __restrict__ psi,
                                                               const uint64_t* __restrict__ total_amps) {
    const uint64_t n = *total_amps;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) psi[i] = cx<R>{R(0), R(0)};
}

template <class R>
__global__ __launch_bounds__(qf::kThreads) void k_ones_grouped(cx<R>* __restrict__ psi,
                                                               const unsigned char* __restrict__ inst) {
    const GroupHeader* h = reinterpret_cast<const GroupHeader*>(inst);
    const InstEntry* I = reinterpret_cast<const InstEntry*>(inst + sizeof(GroupHeader));
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < h->nentries; k += gridDim.x * blockDim.x)
        psi[I[k].state_off] = cx<R>{R(1), R(0)};
}

// lambda = O psi and energy partials, per instance (terms sorted by x mask).
template <class R>
__global__ __launch_bounds__(qf::kThreads) void k_pauli_grouped(const cx<R>* __restrict__ psi,
                                                                cx<R>* __restrict__ lam,
                                                                const unsigned char* __restrict__ launch,
                                                                const qf::PauliDev* __restrict__ terms,
                                                                qf::dd* __restrict__ epart) {
    const GroupHeader* h = reinterpret_cast<const GroupHeader*>(launch);
    const PauliEntry* E = reinterpret_cast<const PauliEntry*>(launch + sizeof(GroupHeader));
    for (uint64_t b = blockIdx.x; b < h->total_blocks; b += gridDim.x) {
        const PauliEntry& e = E[find_entry(E, h->nentries, b)];
        const uint64_t lb = b - e.block_start;
        const uint64_t dim = uint64_t(1) << e.n;
        const cx<R>* __restrict__ p = psi + e.state_off;
        cx<R>* __restrict__ l = lam ? lam + e.state_off : nullptr;  // null: energies only
        const qf::PauliDev* T = terms + e.terms_off;
        qf::dd er{0, 0}, ei{0, 0};
        const uint64_t stride = uint64_t(e.nblocks) * blockDim.x;
        for (uint64_t i = lb * blockDim.x + threadIdx.x; i < dim; i += stride) {
            double ar = 0.0, ai = 0.0, pr = 0.0, pi = 0.0;
            uint64_t cur = ~uint64_t(0);
            for (uint32_t t = 0; t < e.nterms; ++t) {
                if (T[t].x != cur) {
                    cur = T[t].x;
                    const cx<R> v = p[i ^ cur];
                    pr = v.re;
                    pi = v.im;
                }
                const double sg = (__popcll((i ^ cur) & T[t].z) & 1) ? -1.0 : 1.0;
                const double cr = sg * T[t].cr, ci = sg * T[t].ci;
                ar += cr * pr - ci * pi;
                ai += cr * pi + ci * pr;
            }
            if (lam) l[i] = cx<R>{R(ar), R(ai)};
            const cx<R> v = p[i];
            const double vr = v.re, vi = v.im;
            qf::neumaier(er, vr * ar + vi * ai);
            qf::neumaier(ei, vr * ai - vi * ar);
        }
        qf::block_reduce(er, ei);
        if (threadIdx.x == 0) {
            epart[e.epart_off + 2 * lb] = er;
            epart[e.epart_off + 2 * lb + 1] = ei;
        }
        __syncthreads();
    }
}

// Per instance: energy from its partials, gradients from the reverse-sweep
// partials (occurrences of a parameter in reverse circuit order, blocks in order).
__global__ __launch_bounds__(qf::kThreads) void k_finish_grouped(const unsigned char* __restrict__ inst,
                                                                 const qf::dd* __restrict__ epart,
                                                                 const double* __restrict__ partial,
                                                                 const uint32_t* __restrict__ pcsr,
                                                                 double* __restrict__ energies,
                                                                 double* __restrict__ grads) {
    const GroupHeader* h = reinterpret_cast<const GroupHeader*>(inst);
    const InstEntry* I = reinterpret_cast<const InstEntry*>(inst + sizeof(GroupHeader));
    for (uint32_t k = blockIdx.x; k < h->nentries; k += gridDim.x) {
        const InstEntry& e = I[k];
        qf::dd re{0, 0}, im{0, 0};
        for (uint32_t j = threadIdx.x; j < e.eblocks; j += blockDim.x) {
            re = qf::dd_add(re, epart[e.epart_off + 2 * j]);
            im = qf::dd_add(im, epart[e.epart_off + 2 * j + 1]);
        }
        qf::block_reduce(re, im);
        if (threadIdx.x == 0) energies[e.out_index] = re.s + re.c;
        const uint32_t* pstart = pcsr + e.pcsr_off;
        const uint32_t* pslots = pstart + e.nparams + 1;
        // One parameter at a time, its occurrences' tile partials spread over
        // the workgroup (an instance has few parameters and up to 2^16 tiles).
        for (uint32_t p = 0; p < e.nparams; ++p) {
            qf::dd acc{0, 0}, unused{0, 0};
            for (uint32_t q = pstart[p]; q < pstart[p + 1]; ++q) {
                const double* col = partial + e.part_off + pslots[q];
                for (uint32_t b = threadIdx.x; b < e.nblk; b += blockDim.x) qf::neumaier(acc, col[uint64_t(b) * e.nslots]);
            }
            __syncthreads();  // thread 0 has read the previous reduction's table
            qf::block_reduce(acc, unused);
            if (threadIdx.x == 0) grads[e.grad_off + p] = acc.s + acc.c;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Host
// ---------------------------------------------------------------------------
struct Instance {
    const qf_uneven_instance* src;
    std::vector<qf::PauliDev> terms;  // sorted by x
    qf::Plan fwd, rev;
    std::vector<double> row;          // angle table row
    std::vector<int32_t> slot_param;
    uint64_t grad_off = 0;
    uint64_t ops_fwd = 0, ops_rev = 0;  // offsets into the uploaded op buffer
    uint64_t cs_off = 0, terms_off = 0, pcsr_off = 0;
    std::vector<uint64_t> fwd_pass_off, rev_pass_off;  // op offset of every pass
};

struct WaveBuffers {
    void *psi = nullptr, *lam = nullptr, *part = nullptr, *epart = nullptr;
    void *tables = nullptr, *total_amps = nullptr;
    ~WaveBuffers() {
        for (void* p : {psi, lam, part, epart, tables, total_amps})
            if (p) (void)hipFree(p);
    }
};

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace

namespace qfrt {

// Shared by the uneven batch and qf_autotune: the best register-kernel tile
// shape for a circuit, by timing each candidate's forward plan on the device.
struct TuneResult {
    uint32_t tile = 0, low = 0, reg = 0;
    double ms = 0.0;
};

}  // namespace qfrt

namespace {

std::map<std::string, qfrt::TuneResult>& tune_cache() {
    static std::map<std::string, qfrt::TuneResult> c;
    return c;
}

std::map<std::string, uint32_t>& reverse_tile_cache() {
    static std::map<std::string, uint32_t> c;
    return c;
}

// Times a plan (built from `gates`) on a scratch state; returns ms per run.
// two_state: the reverse sweep's form (psi, the costate and the overlap
// partial table).
qf_status time_plan(qf_backend* be, const std::vector<qf::GateIn>& gates, uint32_t qubits, int precision,
                    const qf::PlanOptions& po, const std::vector<double>& row, uint32_t nslots, double* ms,
                    bool two_state = false) {
    qf::Plan plan;
    std::string e = qf::build_plan(gates, qubits, precision, po, plan);
    if (!e.empty()) return fail(QF_INVALID_ARGUMENT, e);
    void* sp = nullptr;
    qf_backend_stream(be, &sp);
    hipStream_t st = static_cast<hipStream_t>(sp);
    void* d = nullptr;
    qf_status s = device_alloc(&d, (size_t(1) << qubits) * amp_bytes(precision));
    if (s) return s;
    void* cs = nullptr;
    if ((s = device_alloc(&cs, std::max<size_t>(row.size(), 2) * sizeof(double)))) {
        (void)hipFree(d);
        return s;
    }
    void* lam = nullptr;
    void* part = nullptr;
    const uint32_t M = plan.passes.empty() ? qubits : plan.passes[0].tile_bits;
    const size_t part_bytes = (size_t(1) << (qubits - std::min(M, qubits))) * std::max<uint32_t>(nslots, 1) * 8;
    if (two_state && ((s = device_alloc(&lam, (size_t(1) << qubits) * amp_bytes(precision))) ||
                      (s = device_alloc(&part, part_bytes)))) {
        if (lam) (void)hipFree(lam);
        (void)hipFree(d);
        (void)hipFree(cs);
        return s;
    }
    std::vector<const qf::TileOp*> bases;
    hipEvent_t a = nullptr, b = nullptr;
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(st);
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
        (void)hipFree(d);
        (void)hipFree(cs);
        if (lam) (void)hipFree(lam);
        if (part) (void)hipFree(part);
    };
    if (!row.empty() && hipMemcpy(cs, row.data(), row.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        cleanup();
        return fail(QF_DEVICE_ERROR, "autotune: angle upload failed");
    }
    if ((s = upload_plans(be, {&plan}, bases))) {
        cleanup();
        return s;
    }
    LaunchCtx ctx;
    ctx.cs = static_cast<const double*>(cs);
    ctx.nslots = nslots;
    ctx.lam = lam;
    ctx.partial = static_cast<double*>(part);
    // The state content does not matter for timing; one warm run, then two timed runs.
    if ((s = launch_plan(be, d, precision, qubits, 1, plan, bases[0], ctx))) {
        cleanup();
        return s;
    }
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
        cleanup();
        return fail(QF_DEVICE_ERROR, "autotune: event creation failed");
    }
    (void)hipEventRecord(a, st);
    for (int r = 0; r < 2 && !s; ++r) s = launch_plan(be, d, precision, qubits, 1, plan, bases[0], ctx);
    (void)hipEventRecord(b, st);
    (void)hipEventSynchronize(b);
    float t = 0.f;
    (void)hipEventElapsedTime(&t, a, b);
    *ms = double(t) / 2.0;
    cleanup();
    return s;
}

// Candidate register-kernel shapes (tile bits, low bits, register bits).
std::vector<qfrt::TuneResult> candidates(int precision, uint32_t qubits) {
    std::vector<qfrt::TuneResult> c;
    const uint32_t shapes128[][3] = {{11, 5, 3}, {12, 5, 4}, {11, 4, 3}, {12, 6, 4}, {10, 4, 3}};
    const uint32_t shapes64[][3] = {{12, 6, 4}, {11, 5, 3}, {12, 5, 4}, {11, 6, 3}, {10, 5, 3}};
    for (int k = 0; k < 5; ++k) {
        const uint32_t* sh = precision == QF_C128 ? shapes128[k] : shapes64[k];
        if (sh[0] <= qubits) c.push_back(qfrt::TuneResult{sh[0], sh[1], sh[2], 0.0});
    }
    return c;
}

qf_status autotune_gates(qf_backend* be, const std::vector<qf::GateIn>& gates, uint32_t qubits, int precision,
                         const std::vector<double>& row, uint32_t nslots, const std::string& key,
                         qfrt::TuneResult* best) {
    auto it = tune_cache().find(key);
    if (it != tune_cache().end()) {
        *best = it->second;
        return QF_OK;
    }
    std::vector<qfrt::TuneResult> cand = candidates(precision, qubits);
    if (cand.empty()) {  // too few qubits for the register kernel: LDS tile kernel
        *best = qfrt::TuneResult{0, 0, 0, 0.0};
        return QF_OK;
    }
    bool have = false;
    for (auto& c : cand) {
        qf::PlanOptions po;
        po.fuse = 1;
        po.tile_qubits = c.tile;
        po.low_qubits = c.low;
        po.reg_bits = c.reg;
        po.kernel = 2;
        double ms = 0.0;
        qf_status s = time_plan(be, gates, qubits, precision, po, row, nslots, &ms);
        if (s) return s;
        c.ms = ms;
        if (!have || ms < best->ms) {
            *best = c;
            have = true;
        }
    }
    tune_cache()[key] = *best;
    return QF_OK;
}

}  // namespace

extern "C" qf_status qf_autotune(qf_backend* be, const qf_op* ops, size_t count, uint32_t qubits, int32_t precision,
                                 qf_plan_options* best, double* best_ms) {
    if (!be || (!ops && count) || !best) return fail(QF_INVALID_ARGUMENT, "null argument");
    if (precision != QF_C128 && precision != QF_C64) return fail(QF_INVALID_ARGUMENT, "unknown precision");
    qf_status s = check_qubits(qubits);
    if (s) return s;
    std::vector<qf::GateIn> g;
    if ((s = lower_ops(ops, count, qubits, g))) return s;
    if ((s = activate(be))) return s;
    const std::string key = "c:" + std::to_string(qubits) + ":" + std::to_string(precision) + ":" +
                            std::to_string(count) + ":" + std::to_string(std::hash<std::string>()(
                                std::string(reinterpret_cast<const char*>(ops), count * sizeof(qf_op))));
    qfrt::TuneResult r;
    if ((s = autotune_gates(be, g, qubits, precision, {}, 0, key, &r))) return s;
    std::memset(best, 0, sizeof(*best));
    best->fuse = 1;
    best->tile_qubits = r.tile;
    best->low_qubits = r.low;
    best->reg_bits = r.reg;
    best->kernel = r.tile ? 2 : 1;
    best->precision = precision;
    if (best_ms) *best_ms = r.ms;
    return QF_OK;
}

namespace {

uint32_t energy_blocks(uint32_t n) {
    const uint64_t dim = uint64_t(1) << n;
    uint64_t b = dim / 1024;
    if (b > 1024) b = 1024;
    return uint32_t(b ? b : 1);
}

qf::TileArgs pass_args(const qf::Pass& p, uint32_t n, uint32_t nslots) {
    qf::TileArgs a{};
    a.n = n;
    a.M = p.tile_bits;
    a.low = p.low;
    a.log_blocks = n - p.tile_bits;
    a.nops = uint32_t(p.ops.size());
    a.nslots = nslots;
    for (size_t k = 0; k < p.high.size(); ++k) a.high[k] = p.high[k];
    return a;
}

template <class R>
qf_status capture_wave_graph(hipStream_t st, uint32_t F, uint32_t Rv, uint32_t RBf, uint32_t Mf, uint32_t RBr,
                             uint32_t Mr, void* psi, void* lam, unsigned char* tables, size_t table_stride,
                             const qf::TileOp* ops, const double* cs, double* part, qf::dd* epart,
                             const qf::PauliDev* terms, const uint32_t* pcsr, double* energies, double* grads,
                             const uint64_t* total_amps, hipGraphExec_t* exec, uint32_t* nodes) {
    // The grouped tile kernels loop over their tiles; 8192 workgroups keep
    // every CU at its LDS-limited occupancy (the reverse kernel runs one
    // wave per workgroup).
    const dim3 tile_grid(8192), elem_grid(4096), inst_grid(512);
    const dim3 fblock(1u << (Mf - RBf)), rblock(1u << (Mr - RBr));
    const size_t flds = (size_t(1) << Mf) * sizeof(cx<R>), rlds = 2 * (size_t(1) << Mr) * sizeof(cx<R>);
    auto fwd_fn = RBf == 3 ? reinterpret_cast<const void*>(&k_tile_reg_grouped<R, 3, 1>)
                           : reinterpret_cast<const void*>(&k_tile_reg_grouped<R, 4, 1>);
    if (flds > 64 * 1024) QF_HIP_TRY(hipFuncSetAttribute(fwd_fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(flds)));
    if (rlds > 64 * 1024)
        QF_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tile_reg_grouped<R, 3, 2>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(rlds)));
    (void)RBr;
    hipGraph_t graph = nullptr;
    QF_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    cx<R>* p = static_cast<cx<R>*>(psi);
    cx<R>* l = static_cast<cx<R>*>(lam);
    // table layout: [0] instances, [1 .. F] forward, [F+1] Pauli, [F+2 .. F+1+Rv] reverse
    const unsigned char* t_inst = tables;
    hipLaunchKernelGGL(k_zero_grouped<R>, elem_grid, dim3(qf::kThreads), 0, st, p, total_amps);
    hipLaunchKernelGGL(k_ones_grouped<R>, dim3(4), dim3(qf::kThreads), 0, st, p, t_inst);
    uint32_t count = 2;
    for (uint32_t j = 0; j < F; ++j, ++count) {
        const unsigned char* t = tables + (1 + j) * table_stride;
        if (RBf == 3)
            hipLaunchKernelGGL((k_tile_reg_grouped<R, 3, 1>), tile_grid, fblock, flds, st, p, l, t, ops, cs, part);
        else
            hipLaunchKernelGGL((k_tile_reg_grouped<R, 4, 1>), tile_grid, fblock, flds, st, p, l, t, ops, cs, part);
    }
    hipLaunchKernelGGL(k_pauli_grouped<R>, dim3(2048), dim3(qf::kThreads), 0, st, p, l,
                       tables + (1 + F) * table_stride, terms, epart);
    ++count;
    for (uint32_t j = 0; j < Rv; ++j, ++count) {
        const unsigned char* t = tables + (2 + F + j) * table_stride;
        hipLaunchKernelGGL((k_tile_reg_grouped<R, 3, 2>), tile_grid, rblock, rlds, st, p, l, t, ops, cs, part);
    }
    hipLaunchKernelGGL(k_finish_grouped, inst_grid, dim3(qf::kThreads), 0, st, t_inst, epart, part, pcsr, energies,
                       grads);
    ++count;
    hipError_t le = hipGetLastError();
    hipError_t ce = hipStreamEndCapture(st, &graph);
    if (le != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return hip_fail(le, "grouped launch during capture");
    }
    if (ce != hipSuccess) return hip_fail(ce, "hipStreamEndCapture");
    hipError_t ie = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) return hip_fail(ie, "hipGraphInstantiate");
    *nodes = count;
    return QF_OK;
}

}  // namespace

namespace {
// Energies, and with want_grad the adjoint gradients, of an uneven batch.
// Without gradients the instances have no reverse plans, no costate buffer
// and no partial table: the per-wave graph is the forward passes, the energy
// reduction and the finishing kernel.
qf_status uneven_run(qf_backend* be, const qf_uneven_instance* inst, uint32_t count, int32_t precision,
                     double* energies, double* grads, qf_uneven_stats* stats, bool want_grad) {
    if (!be || !inst || !energies) return fail(QF_INVALID_ARGUMENT, "null argument");
    if (count == 0) return fail(QF_EMPTY_BATCH, "an uneven batch cannot be empty");
    if (precision != QF_C128 && precision != QF_C64) return fail(QF_INVALID_ARGUMENT, "unknown precision");
    qf_uneven_stats st_local{};
    qf_uneven_stats& S = stats ? *stats : st_local;
    std::memset(&S, 0, sizeof(S));
    qf_status s;
    // 1. validate every instance, exactly as qf_gradient would
    std::vector<Instance> I(count);
    std::vector<PCirc> pcs(count);
    uint64_t goff = 0;
    uint32_t widest = 0;
    for (uint32_t k = 0; k < count; ++k) {
        const qf_uneven_instance& q = inst[k];
        if ((!q.ops && q.nops) || (!q.params && q.nparams) || (!q.obs && q.nterms))
            return fail(QF_INVALID_ARGUMENT, "null argument in instance " + std::to_string(k));
        if ((s = check_qubits(q.qubits))) return s;
        pcs[k].ops = q.ops;
        pcs[k].nops = q.nops;
        pcs[k].nparams = q.nparams;
        pcs[k].ninputs = 0;
        if ((s = validate_values(pcs[k], nullptr, 0, q.params, q.nparams))) return s;
        std::vector<qf::PauliDev> t;
        if ((s = pauli_terms(q.obs, q.nterms, q.qubits, t))) return s;
        for (size_t i = 0; i < q.nops; ++i)
            if (q.ops[i].kind == 2) return fail(QF_INVALID_INPUT_COUNT, "expected 0 encoded inputs");
        if ((s = validate_ops(pcs[k], q.qubits))) return s;
        I[k].src = &q;
        I[k].terms = sorted_terms(t);
        I[k].grad_off = goff;
        goff += q.nparams;
    }
    if (want_grad && goff && !grads) return fail(QF_INVALID_ARGUMENT, "null gradient output");
    if ((s = activate(be))) return s;
    void* sp = nullptr;
    qf_backend_stream(be, &sp);
    hipStream_t st = static_cast<hipStream_t>(sp);

    // QAOA-form instances on their own engine; the rest below.
    std::vector<char> general(count, 1);
    {
        std::vector<qfq::Form> forms(count);
        std::vector<uint32_t> qids;
        std::vector<uint64_t> goffs(count);
        const char* sw = std::getenv("QF_QAOA_ENGINE");
        const bool on = !(sw && std::strcmp(sw, "0") == 0);
        for (uint32_t k = 0; k < count; ++k) {
            goffs[k] = I[k].grad_off;
            if (on && qfq::match(inst[k], pcs[k], forms[k])) {
                qids.push_back(k);
                general[k] = 0;
            }
        }
        if ((s = qfq::run(be, inst, qids, pcs, forms, goffs, precision, want_grad, energies, grads, S))) return s;
        if (qids.size() == count) return QF_OK;
    }
    for (uint32_t k = 0; k < count; ++k)
        if (general[k] && (!general[widest] || inst[k].qubits > inst[widest].qubits)) widest = k;

    // 2. autotune the forward tile shape on the widest instance
    qfrt::TuneResult tune;
    {
        const PCirc& pc = pcs[widest];
        std::vector<double> row(std::max<size_t>(pc.slot_param.size() * 2, 2));
        fill_row(pc, inst[widest].params, -1, 0.0, row.data());
        const std::string key = "u:" + std::to_string(inst[widest].qubits) + ":" + std::to_string(precision) + ":" +
                                std::to_string(inst[widest].nops);
        if ((s = autotune_gates(be, param_gates(pc, nullptr, false), inst[widest].qubits, precision, row,
                                uint32_t(pc.slot_param.size()), key, &tune)))
            return s;
    }
    S.tile_qubits = tune.tile;
    S.low_qubits = tune.low;
    S.reg_bits = tune.reg;
    S.tune_ms = tune.ms;
    const uint32_t Mf = tune.tile, RBf = tune.reg, RBr = 3;
    // The reverse sweep's tile (two states in LDS): the forward tile or one
    // qubit less, timed on the widest instance.
    uint32_t Mr = Mf ? Mf - 1 : 0;
    if (want_grad && Mf) {
        const PCirc& pc = pcs[widest];
        std::vector<double> row(std::max<size_t>(pc.slot_param.size() * 2, 2));
        fill_row(pc, inst[widest].params, -1, 0.0, row.data());
        const std::string key = "r:" + std::to_string(inst[widest].qubits) + ":" + std::to_string(precision) + ":" +
                                std::to_string(inst[widest].nops) + ":" + std::to_string(Mf);
        auto it = reverse_tile_cache().find(key);
        if (it != reverse_tile_cache().end()) {
            Mr = it->second;
        } else {
            double best = -1.0;
            for (uint32_t cand : {Mf - 1, Mf}) {
                if (cand < RBr + 6 || cand - RBr > 8 || cand > inst[widest].qubits ||
                    2 * (size_t(1) << cand) * amp_bytes(precision) > 64 * 1024)
                    continue;
                qf::PlanOptions po;
                po.fuse = 1;
                po.kernel = 2;
                po.tile_qubits = cand;
                po.low_qubits = tune.low;
                po.reg_bits = RBr;
                double ms = 0.0;
                if ((s = time_plan(be, param_gates(pc, nullptr, true), inst[widest].qubits, precision, po, row,
                                   uint32_t(pc.slot_param.size()), &ms, true)))
                    return s;
                if (best < 0 || ms < best) {
                    best = ms;
                    Mr = cand;
                }
            }
            reverse_tile_cache()[key] = Mr;
        }
    }
    S.reverse_tile_qubits = Mr;

    // 3. plans (on host threads: one forward and one reverse plan per
    // instance); instances too narrow for the shared register tile run on
    // their own
    std::vector<uint32_t> grouped, single;
    for (uint32_t k = 0; k < count; ++k) {
        if (!general[k]) continue;
        if (Mf == 0 || inst[k].qubits < Mf) single.push_back(k);
        else grouped.push_back(k);
    }
    {
        std::vector<std::string> errs(grouped.size());
        auto plan_one = [&](size_t g) {
            const uint32_t k = grouped[g];
            qf::PlanOptions po;
            po.fuse = 1;
            po.kernel = 2;
            po.tile_qubits = Mf;
            po.low_qubits = tune.low;
            po.reg_bits = RBf;
            errs[g] = qf::build_plan(param_gates(pcs[k], nullptr, false), inst[k].qubits, precision, po, I[k].fwd);
            if (!errs[g].empty() || !want_grad) return;
            po.tile_qubits = Mr;
            po.reg_bits = RBr;
            errs[g] = qf::build_plan(param_gates(pcs[k], nullptr, true), inst[k].qubits, precision, po, I[k].rev);
        };
        const size_t nt = std::min<size_t>(std::max(1u, std::min(16u, std::thread::hardware_concurrency())),
                                           grouped.size() / 8 + 1);
        std::vector<std::thread> pool;
        for (size_t w = 0; w < nt; ++w)
            pool.emplace_back([&, w] {
                for (size_t g = w; g < grouped.size(); g += nt) plan_one(g);
            });
        for (auto& t : pool) t.join();
        for (const std::string& e : errs)
            if (!e.empty()) return fail(QF_INVALID_ARGUMENT, e);
    }
    for (uint32_t k : single) {  // narrow instances: the batched path with one set each
        qf_gradient_options o{};
        o.strategy = QF_GRAD_ADJOINT;
        o.precision = precision;
        o.fd_step = 1e-3;
        if (want_grad)
            s = qf_batch_gradient(be, inst[k].ops, inst[k].nops, inst[k].nparams, 0, nullptr, inst[k].qubits,
                                  inst[k].params, 1, inst[k].obs, inst[k].nterms, &o, energies + k,
                                  inst[k].nparams ? grads + I[k].grad_off : nullptr);
        else
            s = qf_batch_expectation(be, inst[k].ops, inst[k].nops, inst[k].nparams, 0, nullptr, inst[k].qubits,
                                     inst[k].params, 1, inst[k].obs, inst[k].nterms, &o, energies + k);
        if (s) return s;
    }
    S.grouped_instances = uint32_t(grouped.size());
    if (grouped.empty()) return QF_OK;

    // 4. device-resident inputs for all grouped instances
    std::vector<const qf::Plan*> plans;
    for (uint32_t k : grouped) {
        plans.push_back(&I[k].fwd);
        plans.push_back(&I[k].rev);
    }
    std::vector<const qf::TileOp*> bases;
    if ((s = upload_plans(be, plans, bases))) return s;
    const qf::TileOp* ops0 = bases[0];
    uint32_t F = 0, Rv = 0;
    std::vector<double> cs_host;
    std::vector<qf::PauliDev> terms_host;
    std::vector<uint32_t> pcsr_host;
    for (size_t g = 0; g < grouped.size(); ++g) {
        Instance& X = I[grouped[g]];
        const PCirc& pc = pcs[grouped[g]];
        uint64_t off = uint64_t(bases[2 * g] - ops0);
        for (const auto& p : X.fwd.passes) {
            X.fwd_pass_off.push_back(off);
            off += p.ops.size();
        }
        off = uint64_t(bases[2 * g + 1] - ops0);
        for (const auto& p : X.rev.passes) {
            X.rev_pass_off.push_back(off);
            off += p.ops.size();
        }
        F = std::max<uint32_t>(F, uint32_t(X.fwd.passes.size()));
        Rv = std::max<uint32_t>(Rv, uint32_t(X.rev.passes.size()));
        const uint32_t nslots = uint32_t(pc.slot_param.size());
        X.cs_off = cs_host.size();
        cs_host.resize(cs_host.size() + std::max<size_t>(nslots * 2, 2));
        fill_row(pc, X.src->params, -1, 0.0, cs_host.data() + X.cs_off);
        X.terms_off = terms_host.size();
        terms_host.insert(terms_host.end(), X.terms.begin(), X.terms.end());
        X.pcsr_off = pcsr_host.size();
        std::vector<uint32_t> ps(pc.nparams + 1, 0), pl(nslots);
        for (uint32_t q = 0; q < nslots; ++q) ps[pc.slot_param[q] + 1]++;
        for (uint32_t p = 0; p < pc.nparams; ++p) ps[p + 1] += ps[p];
        std::vector<uint32_t> fill(ps.begin(), ps.end() - 1);
        for (uint32_t q = nslots; q-- > 0;) pl[fill[pc.slot_param[q]]++] = q;
        pcsr_host.insert(pcsr_host.end(), ps.begin(), ps.end());
        pcsr_host.insert(pcsr_host.end(), pl.begin(), pl.end());
    }
    struct Dev {
        void* p = nullptr;
        ~Dev() {
            if (p) (void)hipFree(p);
        }
    } d_cs, d_terms, d_pcsr, d_energy, d_grads;
    if ((s = device_alloc(&d_cs.p, cs_host.size() * sizeof(double) + 16))) return s;
    if ((s = device_alloc(&d_terms.p, terms_host.size() * sizeof(qf::PauliDev) + 16))) return s;
    if ((s = device_alloc(&d_pcsr.p, pcsr_host.size() * 4 + 16))) return s;
    if ((s = device_alloc(&d_energy.p, size_t(count) * sizeof(double)))) return s;
    if ((s = device_alloc(&d_grads.p, std::max<uint64_t>(goff, 1) * sizeof(double)))) return s;
    QF_HIP_TRY(hipMemcpy(d_cs.p, cs_host.data(), cs_host.size() * sizeof(double), hipMemcpyHostToDevice));
    if (!terms_host.empty())
        QF_HIP_TRY(hipMemcpy(d_terms.p, terms_host.data(), terms_host.size() * sizeof(qf::PauliDev),
                             hipMemcpyHostToDevice));
    QF_HIP_TRY(hipMemcpy(d_pcsr.p, pcsr_host.data(), pcsr_host.size() * 4, hipMemcpyHostToDevice));

    // 5. pack instances into memory waves (first fit, widest first)
    const size_t ab = amp_bytes(precision);
    auto footprint = [&](uint32_t k) {
        const uint32_t n = inst[k].qubits;
        const uint64_t nblk = uint64_t(1) << (n - Mr);
        return (want_grad ? 2 : 1) * (uint64_t(1) << n) * ab +
               (want_grad ? nblk * pcs[k].slot_param.size() * 8 : 0) + energy_blocks(n) * 32 + 256;
    };
    std::vector<uint32_t> order(grouped);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return inst[a].qubits > inst[b].qubits; });
    size_t freeb = 0, totalb = 0;
    QF_HIP_TRY(hipMemGetInfo(&freeb, &totalb));
    uint64_t budget = uint64_t(double(freeb) * 0.8);
    // QF_UNEVEN_WAVE_BYTES caps the bytes of one wave (a process sharing the
    // device, or a test that wants several waves on small instances).
    if (const char* cap = std::getenv("QF_UNEVEN_WAVE_BYTES")) {
        const unsigned long long v = std::strtoull(cap, nullptr, 10);
        if (v > 0 && v < budget) budget = v;
    }
    std::vector<std::vector<uint32_t>> waves;
    std::vector<uint64_t> wave_bytes;
    for (uint32_t k : order) {
        const uint64_t f = footprint(k);
        if (f > budget) return fail(QF_OUT_OF_MEMORY, "instance " + std::to_string(k) + " does not fit in device memory");
        bool placed = false;
        for (size_t w = 0; w < waves.size() && !placed; ++w)
            if (wave_bytes[w] + f <= budget) {
                waves[w].push_back(k);
                wave_bytes[w] += f;
                placed = true;
            }
        if (!placed) {
            waves.push_back({k});
            wave_bytes.push_back(f);
        }
    }
    S.waves += uint32_t(waves.size());

    // 6. wave buffers sized for the largest wave
    uint64_t max_amps = 0, max_part = 0, max_epart = 0;
    size_t max_entries = 1;
    for (const auto& w : waves) {
        uint64_t a = 0, p = 0, e = 0;
        for (uint32_t k : w) {
            const uint32_t n = inst[k].qubits;
            a += uint64_t(1) << n;
            if (want_grad) p += (uint64_t(1) << (n - Mr)) * pcs[k].slot_param.size();
            e += 2 * energy_blocks(n);
        }
        max_amps = std::max(max_amps, a);
        max_part = std::max(max_part, p);
        max_epart = std::max(max_epart, e);
        max_entries = std::max(max_entries, w.size());
    }
    const size_t entry_bytes = std::max(sizeof(TileEntry), std::max(sizeof(PauliEntry), sizeof(InstEntry)));
    const size_t table_stride = align_up(sizeof(GroupHeader) + max_entries * entry_bytes, 256);
    const size_t ntables = size_t(F) + Rv + 2;
    WaveBuffers B;
    if ((s = device_alloc(&B.psi, max_amps * ab))) return s;
    if (want_grad && (s = device_alloc(&B.lam, max_amps * ab))) return s;
    if ((s = device_alloc(&B.part, std::max<uint64_t>(max_part, 1) * 8))) return s;
    if ((s = device_alloc(&B.epart, std::max<uint64_t>(max_epart, 2) * sizeof(qf::dd)))) return s;
    if ((s = device_alloc(&B.tables, ntables * table_stride))) return s;
    if ((s = device_alloc(&B.total_amps, sizeof(uint64_t)))) return s;

    // 7. capture the per-wave graph once
    hipGraphExec_t exec = nullptr;
    uint32_t nodes = 0;
    QF_HIP_TRY(hipStreamSynchronize(st));
    if (precision == QF_C128)
        s = capture_wave_graph<double>(st, F, Rv, RBf, Mf, RBr, Mr, B.psi, B.lam, static_cast<unsigned char*>(B.tables),
                                       table_stride, ops0, static_cast<const double*>(d_cs.p),
                                       static_cast<double*>(B.part), static_cast<qf::dd*>(B.epart),
                                       static_cast<const qf::PauliDev*>(d_terms.p),
                                       static_cast<const uint32_t*>(d_pcsr.p), static_cast<double*>(d_energy.p),
                                       static_cast<double*>(d_grads.p), static_cast<const uint64_t*>(B.total_amps),
                                       &exec, &nodes);
    else
        s = capture_wave_graph<float>(st, F, Rv, RBf, Mf, RBr, Mr, B.psi, B.lam, static_cast<unsigned char*>(B.tables),
                                      table_stride, ops0, static_cast<const double*>(d_cs.p),
                                      static_cast<double*>(B.part), static_cast<qf::dd*>(B.epart),
                                      static_cast<const qf::PauliDev*>(d_terms.p),
                                      static_cast<const uint32_t*>(d_pcsr.p), static_cast<double*>(d_energy.p),
                                      static_cast<double*>(d_grads.p), static_cast<const uint64_t*>(B.total_amps),
                                      &exec, &nodes);
    if (s) return s;
    S.graph_nodes = std::max(S.graph_nodes, nodes);

    for (uint32_t k : grouped) {
        const double sweep = 2.0 * double(uint64_t(1) << inst[k].qubits) * double(ab);
        S.alg_bytes += sweep * (double(I[k].fwd.passes.size()) + 1.0 + 2.0 * double(I[k].rev.passes.size()));
    }
    hipEvent_t wave_a = nullptr, wave_b = nullptr;
    QF_HIP_TRY(hipEventCreate(&wave_a));
    if (hipEventCreate(&wave_b) != hipSuccess) {
        (void)hipEventDestroy(wave_a);
        return fail(QF_DEVICE_ERROR, "hipEventCreate");
    }

    // 8. per wave: build and upload the tables, replay the graph
    std::vector<unsigned char> host(ntables * table_stride);
    qf_status result = QF_OK;
    for (const auto& w : waves) {
        std::fill(host.begin(), host.end(), 0);
        auto header = [&](size_t t) { return reinterpret_cast<GroupHeader*>(host.data() + t * table_stride); };
        auto entries = [&](size_t t) { return host.data() + t * table_stride + sizeof(GroupHeader); };
        uint64_t amp_off = 0, part_off = 0, ep_off = 0;
        std::vector<uint64_t> aoff(w.size()), poff(w.size()), eoff(w.size());
        for (size_t i = 0; i < w.size(); ++i) {
            const uint32_t n = inst[w[i]].qubits;
            aoff[i] = amp_off;
            poff[i] = part_off;
            eoff[i] = ep_off;
            amp_off += uint64_t(1) << n;
            if (want_grad) part_off += (uint64_t(1) << (n - Mr)) * pcs[w[i]].slot_param.size();
            ep_off += 2 * energy_blocks(n);
        }
        // instances (table 0)
        {
            InstEntry* E = reinterpret_cast<InstEntry*>(entries(0));
            for (size_t i = 0; i < w.size(); ++i) {
                const uint32_t k = w[i], n = inst[k].qubits;
                E[i] = InstEntry{aoff[i], eoff[i], poff[i], I[k].grad_off, k, I[k].pcsr_off, energy_blocks(n),
                                 uint32_t(1) << (n - Mr), uint32_t(pcs[k].slot_param.size()),
                                 want_grad ? pcs[k].nparams : 0u};
            }
            header(0)->nentries = uint32_t(w.size());
            header(0)->total_blocks = w.size();
        }
        // forward passes (tables 1..F) and reverse passes (tables F+2..)
        for (int dir = 0; dir < 2; ++dir) {
            const uint32_t npass = dir == 0 ? F : Rv;
            for (uint32_t j = 0; j < npass; ++j) {
                const size_t t = dir == 0 ? 1 + j : 2 + F + j;
                TileEntry* E = reinterpret_cast<TileEntry*>(entries(t));
                uint32_t ne = 0;
                uint64_t blocks = 0;
                for (size_t i = 0; i < w.size(); ++i) {
                    const uint32_t k = w[i];
                    const qf::Plan& pl = dir == 0 ? I[k].fwd : I[k].rev;
                    if (j >= pl.passes.size()) continue;
                    const qf::Pass& p = pl.passes[j];
                    TileEntry e{};
                    e.state_off = aoff[i];
                    e.block_start = blocks;
                    e.ops_off = dir == 0 ? I[k].fwd_pass_off[j] : I[k].rev_pass_off[j];
                    e.cs_off = I[k].cs_off;
                    e.part_off = poff[i];
                    e.args = pass_args(p, inst[k].qubits, uint32_t(pcs[k].slot_param.size()));
                    e.nblocks = uint64_t(1) << e.args.log_blocks;
                    blocks += e.nblocks;
                    E[ne++] = e;
                }
                header(t)->nentries = ne;
                header(t)->total_blocks = blocks;
            }
        }
        // Pauli table (F+1)
        {
            PauliEntry* E = reinterpret_cast<PauliEntry*>(entries(1 + F));
            uint64_t blocks = 0;
            for (size_t i = 0; i < w.size(); ++i) {
                const uint32_t k = w[i], n = inst[k].qubits;
                E[i] = PauliEntry{aoff[i], blocks, I[k].terms_off, eoff[i], n, energy_blocks(n),
                                  uint32_t(I[k].terms.size()), 0};
                blocks += energy_blocks(n);
            }
            header(1 + F)->nentries = uint32_t(w.size());
            header(1 + F)->total_blocks = blocks;
        }
        if (hipMemcpy(B.tables, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(B.total_amps, &amp_off, sizeof(uint64_t), hipMemcpyHostToDevice) != hipSuccess) {
            result = fail(QF_DEVICE_ERROR, "uploading wave tables failed");
            break;
        }
        hipError_t e = hipEventRecord(wave_a, st);
        if (e == hipSuccess) e = hipGraphLaunch(exec, st);
        if (e == hipSuccess) e = hipEventRecord(wave_b, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, wave_a, wave_b);
        if (e != hipSuccess) {
            result = hip_fail(e, "hipGraphLaunch (uneven wave)");
            break;
        }
        S.device_ms += ms;
        ++S.graph_replays;
    }
    (void)hipEventDestroy(wave_a);
    (void)hipEventDestroy(wave_b);
    (void)hipGraphExecDestroy(exec);
    if (result) return result;
    std::vector<double> e_all(count);
    QF_HIP_TRY(hipMemcpy(e_all.data(), d_energy.p, size_t(count) * sizeof(double), hipMemcpyDeviceToHost));
    for (uint32_t k : grouped) energies[k] = e_all[k];
    if (want_grad && goff) {
        std::vector<double> g_all(goff);
        QF_HIP_TRY(hipMemcpy(g_all.data(), d_grads.p, goff * sizeof(double), hipMemcpyDeviceToHost));
        for (uint32_t k : grouped)
            std::memcpy(grads + I[k].grad_off, g_all.data() + I[k].grad_off, inst[k].nparams * sizeof(double));
    }
    return QF_OK;
}
}  // namespace

extern "C" qf_status qf_uneven_adjoint_gradient(qf_backend* be, const qf_uneven_instance* inst, uint32_t count,
                                                int32_t precision, double* energies, double* grads,
                                                qf_uneven_stats* stats) {
    return uneven_run(be, inst, count, precision, energies, grads, stats, true);
}

extern "C" qf_status qf_uneven_expectation(qf_backend* be, const qf_uneven_instance* inst, uint32_t count,
                                           int32_t precision, double* energies, qf_uneven_stats* stats) {
    return uneven_run(be, inst, count, precision, energies, nullptr, stats, false);
}
