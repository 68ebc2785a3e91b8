"""Builds libqfhip.so (the HIP engine behind include/qf_hip.h) in-tree for gfx950.

    python -m quantumforge_amd.build        # or quantumforge_amd.build.build()

hipcc cross-compiles the device code without a GPU present.  All HIP code is
compiled with packed FP32 instructions disabled (-packed-fp32-ops).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libqfhip.so")
ARCH = os.environ.get("QF_OFFLOAD_ARCH", "gfx950")

SOURCES = ["qf_runtime.hip", "qf_gradients.hip", "qf_uneven.hip", "qf_qaoa.hip", "qf_plan.cpp", "qf_jit.cpp", "qf_shots.hip",
           "qf_density.hip", "qf_lanczos.hip"]

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=fast-honor-pragmas",
    # Keep LLVM from sinking the arms of the register-bit switches into one
    # block with computed array offsets: that turns the per-thread amplitude
    # registers of the tile kernels into scratch memory.
    "-mllvm", "-simplifycfg-sink-common=false",
    "-I", os.path.join(ROOT, "include"),
    # The pass compiler (qf_jit.cpp) compiles generated kernels with hiprtc.
    "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib",
]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found; the HIP engine cannot be built")
    return exe


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, *srcs, "-o", tmp]
    t0 = time.time()
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"hipcc failed ({proc.returncode}) building {LIB}")
    if verbose and (proc.stdout or proc.stderr):
        # The host pass warns that it does not know the packed-fp32 feature; harmless.
        lines = [l for l in (proc.stdout + proc.stderr).splitlines() if "packed-fp32-ops" not in l]
        if lines:
            sys.stderr.write("\n".join(lines) + "\n")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} in {time.time() - t0:.1f}s")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
