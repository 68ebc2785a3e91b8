// Pass compiler: turns each register-kernel tile pass of a plan into its own
// straight-line gfx950 kernel, compiled at run time with hiprtc.
//
// The register kernel (qf_regtile.h) interprets the op list of a pass: every
// thread decodes every op, branches on its kind, form and register bit, and
// the branch arms meet in register copies.  A compiled pass has none of that:
// register bits, matrix structure, index scatter masks and the LDS exchange
// pattern are constants of the source, fixed gate matrices are exact hex
// literals (zero entries vanish, +-1 entries become adds), and permutations
// (X, SWAP, CNOT with its control on a register bit) are renamings that cost
// no instruction.  Rotation angles stay run-time values read from the
// per-instance angle table, so one compiled plan serves every parameter set.
//
// Compiled code is cached in process (by source hash) and, when a cache
// directory is writable, on disk.
#pragma once
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "qf_plan.h"

namespace qf {

// Source of one compiled pass kernel `name` (precision 0 = complex128,
// 1 = complex64).  Returns an empty string and sets *err when the pass is not
// a register-kernel tile pass the compiler handles.
// ns = 2: the adjoint form (kernel arguments add the costate lambda and the
// partial table; TF_RECORD rotations record Im <lambda| G |psi> per tile).
// scaled: run-time X/Y rotations read (scale, ratio) angle rows
// (Plan::scaled_rotations).
std::string jit_pass_source(const Pass& p, uint32_t qubits, int precision, const std::string& name,
                            std::string* err, int ns = 1, bool scaled = false);

// The compiled kernels of one plan at one precision.
struct JitProgram {
    std::string key;                       // hash of the source
    std::vector<std::vector<char>> codes;  // gfx950 code objects (one per compile chunk)
    std::vector<int> code_of_kernel;       // code object holding each kernel
    std::vector<int> kernel_of_pass;       // index into names, -1 for passes run by the generic kernels
    std::vector<std::string> names;
    std::vector<uint32_t> threads;         // workgroup size per kernel
    int states = 1;                        // 2: adjoint form (psi and lambda)
    std::string prelude;                   // shared declarations of the source
    std::vector<std::string> bodies;       // source of each kernel
    double compile_ms = 0;
    bool from_disk = false;
};

// Generates the program source of a plan (kernels named, key set) into *prog
// without compiling; returns the full source, or "" with *err set.
std::string jit_plan_source(const Plan& plan, int precision, JitProgram* prog, std::string* err, int ns = 1);

// Generates, compiles (or fetches from the caches) the program of a plan.
// Returns nullptr and sets *err on failure.
std::shared_ptr<JitProgram> jit_compile_plan(const Plan& plan, int precision, std::string* err, int ns = 1);

// Default compile options (exposed for tests and the describe output).
std::vector<std::string> jit_compile_options();

}  // namespace qf
