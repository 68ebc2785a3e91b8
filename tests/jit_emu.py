"""CPU emulation of compiled-pass kernels (test infrastructure).

The pass compiler emits plain HIP C++ that uses only threadIdx, blockIdx,
gridDim, a static __shared__ array and __syncthreads().  Here that source is
compiled with g++ under a small emulation of those built-ins: every workgroup
runs as real host threads (one per work item) meeting at a std::barrier, one
workgroup after another.  This checks what the generated code computes (index
scatter, register renaming, phase exchanges through the shared array) against
the oracle without a GPU.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import os
import re
import subprocess
import tempfile

import numpy as np

PRELUDE = r"""
#include <barrier>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>
struct qf_dim3 { unsigned x, y, z; };
static thread_local qf_dim3 qf_tid, qf_bid;
static qf_dim3 qf_grid;
static std::barrier<>* qf_bar;
#define threadIdx qf_tid
#define blockIdx qf_bid
#define gridDim qf_grid
#define __global__
#define __launch_bounds__(...)
#define __shared__ static
#define __syncthreads() qf_bar->arrive_and_wait()
// Lane exchange of the wave reductions: every work item of the workgroup
// calls it at the same point.
static double qf_shfl[1024];
static inline double __shfl_xor(double v, int m, int) {
    qf_shfl[qf_tid.x] = v;
    qf_bar->arrive_and_wait();
    const double r = qf_shfl[qf_tid.x ^ unsigned(m)];
    qf_bar->arrive_and_wait();
    return r;
}
// The wave sum the generated adjoint code leaves in lane 63 (here in every
// lane), added in lane order.
#define QF_WAVE_SUM_DEFINED
static inline double qf_wave_sum(double v) {
    qf_shfl[qf_tid.x] = v;
    qf_bar->arrive_and_wait();
    double s = 0.0;
    for (unsigned l = qf_tid.x & ~63u; l < (qf_tid.x & ~63u) + 64u; ++l) s += qf_shfl[l];
    qf_bar->arrive_and_wait();
    return s;
}
"""

# Lane exchanges of the batched record sums: DPP quad permutes and row
# rotations (0x121..0x12f: lane i of a row reads lane i + n mod 16; the code
# only uses rotations whose sums do not depend on the direction), swizzles
# across lanes (bit mode: lane ^ xor mask within 32) and a permute across 32.
PRELUDE += r"""
static inline double qf_lane_read(double v, unsigned src) {
    qf_shfl[qf_tid.x] = v;
    qf_bar->arrive_and_wait();
    const double r = qf_shfl[(qf_tid.x & ~63u) | src];
    qf_bar->arrive_and_wait();
    return r;
}
template <int CTRL, int ROWS> static inline double qf_dpp(double v) {
    const unsigned ln = qf_tid.x & 63u;
    unsigned src = ln;
    if (CTRL < 0x100) src = (ln & ~3u) | ((unsigned(CTRL) >> (2 * (ln & 3u))) & 3u);
    else if (CTRL > 0x120 && CTRL < 0x130) src = (ln & ~15u) | ((ln + unsigned(CTRL - 0x120)) & 15u);
    return qf_lane_read(v, src);
}
template <int PAT> static inline double qf_swz(double v) {
    const unsigned ln = qf_tid.x & 63u;
    return qf_lane_read(v, (ln & ~31u) | ((ln & 31u) ^ ((unsigned(PAT) >> 10) & 31u)));
}
static inline double qf_bperm32(double v, unsigned ln) { return qf_lane_read(v, ln ^ 32u); }
#define QF_PIN(v) ((void)0)
#define __builtin_amdgcn_readfirstlane(x) (x)
"""

LAUNCHER = r"""
typedef void (*qf_kfn)(C*, const double*, unsigned);
typedef void (*qf_kfn2)(C*, const double*, unsigned, C*, double*);
extern "C" void qf_emu_launch(void* fn, C* s, const double* cs, unsigned nslots, C* lam, double* partial,
                              unsigned gx, unsigned gy, unsigned nt) {
    qf_grid = {gx, gy, 1};
    for (unsigned by = 0; by < gy; ++by)
        for (unsigned bx = 0; bx < gx; ++bx) {
            std::barrier<> bar(nt);
            qf_bar = &bar;
            std::vector<std::thread> th;
            th.reserve(nt);
            for (unsigned t = 0; t < nt; ++t)
                th.emplace_back([=] {
                    qf_tid = {t, 0, 0};
                    qf_bid = {bx, by, 0};
                    if (lam) reinterpret_cast<qf_kfn2>(fn)(s, cs, nslots, lam, partial);
                    else reinterpret_cast<qf_kfn>(fn)(s, cs, nslots);
                });
            for (auto& x : th) x.join();
        }
}
"""

KERNEL_RE = re.compile(r'extern "C" __global__ __launch_bounds__\((\d+)(?:, \d+)?\) void (\w+)\(')


class EmulatedProgram:
    """g++ build of a plan's generated source; run() applies every pass."""

    def __init__(self, source: str, workdir: str | None = None) -> None:
        self.kernels = [(name, int(nt)) for nt, name in KERNEL_RE.findall(source)]
        self.complex64 = "float re, im" in source
        full = PRELUDE + source + LAUNCHER
        key = hashlib.sha1(full.encode()).hexdigest()[:16]
        workdir = workdir or tempfile.mkdtemp(prefix="qf_emu_")
        src = os.path.join(workdir, f"emu_{key}.cpp")
        so = os.path.join(workdir, f"emu_{key}.so")
        if not os.path.exists(so):
            with open(src, "w") as f:
                f.write(full)
            subprocess.run(["g++", "-O1", "-std=c++20", "-shared", "-fPIC", "-pthread", "-ffp-contract=off",
                            "-w", src, "-o", so], check=True, capture_output=True, text=True)
        self.lib = C.CDLL(so)
        self.launch = self.lib.qf_emu_launch
        self.launch.restype = None
        self.launch.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint, C.c_void_p, C.c_void_p, C.c_uint,
                                C.c_uint, C.c_uint]
        self.adjoint = "double* __restrict__ partial" in source

    def run(self, amps: np.ndarray, qubits: int, tile_bits: list[int], count: int = 1,
            cs: np.ndarray | None = None, nslots: int = 0, lam: np.ndarray | None = None):
        """Applies the passes in order to `count` contiguous states (a copy);
        tile_bits[i] is the tile size of compiled kernel i.  Adjoint-form
        source: lam is the costate (a copy is carried along) and the call
        returns (psi, lambda, partial table [count * 2^(n - tile)][nslots])."""
        dt = np.complex64 if self.complex64 else np.complex128
        s = np.ascontiguousarray(amps.astype(dt, copy=True))
        csp = C.c_void_p(cs.ctypes.data) if cs is not None and cs.size else C.c_void_p(0)
        lp = pp = C.c_void_p(0)
        if self.adjoint:
            lam = np.ascontiguousarray(lam.astype(dt, copy=True))
            rows = count << (qubits - (tile_bits[0] if tile_bits else qubits))
            partial = np.zeros((rows, max(nslots, 1)))
            lp, pp = C.c_void_p(lam.ctypes.data), C.c_void_p(partial.ctypes.data)
        for (name, nt), M in zip(self.kernels, tile_bits):
            fn = C.cast(getattr(self.lib, name), C.c_void_p)
            logb = qubits - M
            gx_log = min(logb, 16)
            gy = count << (logb - gx_log)
            self.launch(fn, C.c_void_p(s.ctypes.data), csp, nslots, lp, pp, 1 << gx_log, gy, nt)
        return (s, lam, partial) if self.adjoint else s
