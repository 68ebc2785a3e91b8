/*
 * qf_hip.h - C ABI of the MI355X (gfx950) state-vector engine.
 *
 * Plain C: opaque handles, plain pointers and sizes, int32 status codes.  No
 * C++ or torch types cross this boundary.  Every entry point states which
 * interface of the reference crate (quantumforge 0.2.10) it replaces, as
 * file:line in that crate.
 *
 * Conventions (identical to the reference):
 *   - qubit 0 is the least-significant bit of the basis index
 *     (src/lib.rs:674-676);
 *   - amplitudes are interleaved {re, im} complex128, the layout of the
 *     reference's repr(C) Complex64 (src/lib.rs:437-445); complex64 states
 *     use {float re, float im};
 *   - gate matrices are computed on the host with the C library's cos/sin,
 *     exactly as Gate::matrix does (src/lib.rs:2267-2333).
 *
 * Threading: a backend owns one HIP stream.  Calls on one backend (and on
 * states created from it) are stream-ordered; a handle must not be used from
 * two threads at once without external locking.  Gate calls validate their
 * arguments on the host before anything is enqueued (the validate-then-launch
 * order of src/cuda.rs:1297-1311), so a returned error leaves the device state
 * untouched.  Device faults surface at the next synchronizing call
 * (download, reductions, qf_backend_synchronize), as the reference's batch
 * path defers them to download (src/cuda.rs:1817-1818).
 *
 * Ownership: the caller destroys every handle it receives.  A state or batch
 * holds a reference on its backend, so it stays valid after
 * qf_backend_destroy (the Arc of src/cuda.rs:519-527, tests/quantum_core.rs:436).
 */
#ifndef QF_HIP_H
#define QF_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes: one per QuantumError / CudaError variant on the path        */
/* (QuantumError src/lib.rs:509-582, CudaError src/cuda.rs:3051-3080).       */
/* ------------------------------------------------------------------------ */
typedef int32_t qf_status;
enum {
    QF_OK = 0,
    QF_ZERO_QUBITS = 1,                 /* QuantumError::ZeroQubits */
    QF_QUBIT_OUT_OF_RANGE = 2,          /* QuantumError::QubitOutOfRange */
    QF_INVALID_QUBIT_PAIR = 3,          /* QuantumError::InvalidQubitPair */
    QF_NON_FINITE_ANGLE = 4,            /* CudaError::NonFiniteAngle */
    QF_STATE_TOO_LARGE = 5,             /* CudaError::StateTooLarge */
    QF_EMPTY_BATCH = 6,                 /* CudaError::EmptyBatch */
    QF_INCOMPATIBLE_BATCH = 7,          /* CudaError::IncompatibleBatch */
    QF_INVALID_PARAMETER_COUNT = 8,     /* QuantumError::InvalidParameterCount */
    QF_INVALID_INPUT_COUNT = 9,         /* QuantumError::InvalidInputCount */
    QF_NON_FINITE_PARAMETER = 10,       /* QuantumError::NonFiniteParameter */
    QF_OUT_OF_MEMORY = 11,              /* QuantumError::MemoryAllocationFailed */
    QF_DEVICE_ERROR = 12,               /* CudaError::Driver */
    QF_REQUIRES_DENSITY_MATRIX = 13,    /* QuantumError::OperationRequiresDensityMatrix */
    QF_REQUIRES_SHOT_EXECUTION = 14,    /* QuantumError::OperationRequiresShotExecution */
    QF_INVALID_ARGUMENT = 15,           /* null handle / bad enum value */
    QF_INCOMPATIBLE_DEVICE_STATE = 16,  /* CudaError::IncompatibleDeviceState */
    QF_STATE_SIZE_NOT_POWER_OF_TWO = 17,/* QuantumError::StateSizeNotPowerOfTwo */
    QF_INVALID_BATCH_SAMPLES = 18,      /* CudaError::InvalidBatchSamples */
    QF_MEMORY_BUDGET_EXCEEDED = 19,     /* QuantumError::MemoryBudgetExceeded */
    QF_INVALID_FD_STEP = 20,            /* QuantumError::InvalidFiniteDifferenceStep */
    QF_SIMULATION_SIZE_OVERFLOW = 21,   /* QuantumError::SimulationSizeOverflow */
    QF_INVALID_PROBABILITY = 22,        /* QuantumError::InvalidProbability */
    QF_INVALID_SHOT_COUNT = 23,         /* QuantumError::InvalidShotCount */
    QF_CLASSICAL_BIT_OUT_OF_RANGE = 24, /* QuantumError::ClassicalBitOutOfRange */
    QF_NON_HERMITIAN_DENSITY_MATRIX = 25,/* QuantumError::NonHermitianDensityMatrix (host-side entropy) */
    QF_INVALID_SUBSYSTEM_SELECTION = 26 /* QuantumError::InvalidSubsystemSelection (host-side partial trace) */
};

/* Gate ids, in the order of `enum Gate` (src/lib.rs:2239-2265). */
enum {
    QF_GATE_I = 0, QF_GATE_X = 1, QF_GATE_Y = 2, QF_GATE_Z = 3, QF_GATE_H = 4,
    QF_GATE_S = 5, QF_GATE_SDG = 6, QF_GATE_T = 7, QF_GATE_TDG = 8,
    QF_GATE_RX = 9, QF_GATE_RY = 10, QF_GATE_RZ = 11
};

/* Noise channels, in the declaration order of NoiseChannel
 * (src/lib.rs:1561-1600); parameters as NoiseChannel::params(). */
enum {
    QF_CHANNEL_DEPOLARIZING = 0, QF_CHANNEL_AMPLITUDE_DAMPING = 1, QF_CHANNEL_PHASE_DAMPING = 2,
    QF_CHANNEL_THERMAL_DAMPING = 3
};

/* Two-qubit op ids, numbered as the reference kernel's `op` argument
 * (0 CNOT, 1 CZ, 2 SWAP; src/cuda.rs:99-113). */
enum { QF_2Q_CNOT = 0, QF_2Q_CZ = 1, QF_2Q_SWAP = 2 };

/* Single-qubit observables (SingleQubitObservable, src/lib.rs:2931-2938). */
enum { QF_OBS_X = 0, QF_OBS_Y = 1, QF_OBS_Z = 2 };

/* Amplitude precision of a device state. */
enum { QF_C128 = 0, QF_C64 = 1 };

/* Gradient strategies (GradientStrategy, src/gradients.rs:48-71). */
enum {
    QF_GRAD_PARAMETER_SHIFT = 0, QF_GRAD_ADJOINT = 1, QF_GRAD_BACKPROP = 2,
    QF_GRAD_FINITE_DIFFERENCE = 3, QF_GRAD_AUTO = 4
};

typedef struct { double re, im; } qf_c128;
typedef struct { float re, im; } qf_c64;

/* One operation of a Circuit (CircuitOp, src/lib.rs:2386-2440).
 * kind: 0 single-qubit {gate, a=target, theta for Rx/Ry/Rz},
 *       1 CNOT {a=control, b=target}, 2 CZ {a, b}, 3 SWAP {a, b},
 *       4 channel {gate = QF_CHANNEL_*, a=target, theta = first parameter,
 *       theta2 = ThermalDamping's beta_delta} (NoiseChannel,
 *       src/lib.rs:1561-1600), 5 measure {a=target, b=classical bit},
 *       6 classically conditioned gate {gate, a=target, theta, b=classical bit
 *       | value << 16} (CircuitOp::Measure / ConditionalSingleQubit,
 *       src/lib.rs:2386-2440; value is 0 or 1, if_classical_bit keeps value & 1);
 *       4-6 are rejected by the state-vector runs with the reference's errors
 *       (src/lib.rs:2793-2798) and run by qf_run_circuit_shots,
 *       7 exact controlled phase diag(1,1,1,e^{i theta}) {a=control, b=target}
 *       (no CircuitOp counterpart: the crate lowers it to Rz/CNOT,
 *       src/algorithms.rs:139-145, which equals it times e^{-i theta/4}). */
typedef struct {
    int32_t kind;
    int32_t gate;
    uint32_t a, b;
    double theta;
    double theta2;
} qf_op;

/* One operation of a ParameterizedCircuit (ParameterizedOp, src/lib.rs:2942-2996).
 * kind: 0 fixed gate {gate, a, theta}, 1 rotation {gate=axis 0 X / 1 Y / 2 Z,
 *       a=target, index=parameter}, 2 encoding rotation {axis, a, index=input},
 *       3 CNOT {a=control, b=target}, 4 CZ, 5 SWAP, 6 channel (rejected). */
typedef struct {
    int32_t kind;
    int32_t gate;
    uint32_t a, b;
    uint32_t index;
    uint32_t reserved;
    double theta;
} qf_pop;

/* One weighted Pauli word of a PauliSum (src/analytic.rs:2549-2778): qubit q
 * carries X if only bit q of x_mask is set, Z if only bit q of z_mask is set,
 * Y if both are set.  Y follows apply_basis_index (src/analytic.rs:2629-2652):
 * +i on a 0 bit, -i on a 1 bit. */
typedef struct {
    double re, im;
    uint64_t x_mask, z_mask;
} qf_pauli_term;

/* One rotation of apply_rotation_chain (src/cuda.rs:1325-1347). */
typedef struct {
    int32_t axis;     /* 0 X, 1 Y, 2 Z */
    uint32_t qubit;
    double theta;
} qf_rotation;

/* Fusion / tiling options of the circuit planner (no reference counterpart:
 * the reference dispatches one kernel per gate, src/cuda.rs:1050-1083). */
typedef struct {
    int32_t fuse;          /* 0: one sweep per gate (reference order, bit-exact arithmetic); 1: tile passes */
    uint32_t tile_qubits;  /* qubits held per workgroup tile (0 = default) */
    uint32_t low_qubits;   /* contiguous low qubits always in a tile (0 = default) */
    uint32_t max_block;    /* kernel 4: qubits of a dense block, 4 or 5 (0 = 5) */
    int32_t kernel;        /* 0 default, 1 LDS op-by-op tile kernel, 2 register-resident tile kernel,
                              3 compiled passes (one generated kernel per pass; default from 18 qubits),
                              4 dense max_block-qubit blocks on the matrix cores (one block matrix per
                              pass, composed in FP64 on the host; fixed gates, at least max_block + 4
                              qubits) */
    uint32_t reg_bits;     /* register kernel: 2^reg_bits amplitudes per thread (0 = default) */
    int32_t precision;     /* QF_C128 or QF_C64: the amplitude type the defaults are chosen for */
} qf_plan_options;

/* Kernel families a plan launch can use (qf_plan_stats.kernels_run bits). */
enum {
    QF_KERNEL_SWEEP = 1,     /* one full-state sweep per gate (exact lowering) */
    QF_KERNEL_LDS_TILE = 2,  /* LDS tile kernel, op by op */
    QF_KERNEL_REG_TILE = 4,  /* interpreting register-tile kernel */
    QF_KERNEL_COMPILED = 8,  /* generated per-pass kernels */
    QF_KERNEL_MFMA = 16      /* dense-block matrix-core passes */
};

/* Summary of a plan (counts the HBM sweeps the plan streams). */
typedef struct {
    uint64_t passes;       /* full-state sweeps (kernel launches that stream the state once) */
    uint64_t ops_in;       /* operations in the circuit */
    uint64_t ops_applied;  /* operations after 1q merging */
    uint64_t dense_blocks; /* fused dense k-qubit blocks (kernel 4: one per pass) */
    uint64_t max_ops_per_pass;
    uint64_t phases;       /* register-kernel phases (LDS re-distributions + 1 per pass) */
    uint32_t kernels_run;  /* QF_KERNEL_* mask of what the last run of this plan launched (0: not run) */
    uint32_t runtime_rotations; /* rotations whose angles come from a per-set table (parameterized plans) */
} qf_plan_stats;

typedef struct qf_backend qf_backend;
typedef struct qf_state qf_state;
typedef struct qf_batch qf_batch;
typedef struct qf_plan qf_plan;
typedef struct qf_density qf_density;

/* ------------------------------------------------------------------------ */
/* Errors                                                                    */
/* ------------------------------------------------------------------------ */
/* Message of the last failed call on this thread (Display of CudaError /
 * QuantumError, src/cuda.rs:3081-3104, src/lib.rs:584-658). */
const char* qf_last_error_message(void);
/* Stable name of a status code ("QubitOutOfRange", ...). */
const char* qf_status_name(qf_status status);

/* ------------------------------------------------------------------------ */
/* Backend (CudaBackend, src/cuda.rs:513-790)                                */
/* ------------------------------------------------------------------------ */
/* CudaBackend::new(ordinal), src/cuda.rs:556. */
qf_status qf_backend_create(int32_t ordinal, qf_backend** out);
/* Drop for CudaBackend; live states keep the device alive. */
qf_status qf_backend_destroy(qf_backend* backend);
/* cuCtxSynchronize, src/cuda.rs:1983; reports deferred device errors. */
qf_status qf_backend_synchronize(qf_backend* backend);
/* No reference counterpart: frees the device arena the uneven-batch engine
 * keeps between calls (its psi, lambda and partial buffers), as a caching
 * allocator's trim.  A device allocation that fails releases every backend's
 * arena and tries once more by itself. */
qf_status qf_backend_release_cached(qf_backend* backend);
/* CudaBackend::device_ordinal. */
qf_status qf_backend_device_ordinal(const qf_backend* backend, int32_t* out);
/* Planner options used by qf_state_run_circuit when no plan is passed. */
qf_status qf_backend_set_plan_options(qf_backend* backend, const qf_plan_options* options);
/* The backend's HIP stream (hipStream_t as void*), for interop with callers
 * that enqueue their own work (e.g. a torch stream wait). */
qf_status qf_backend_stream(const qf_backend* backend, void** out);

/* ------------------------------------------------------------------------ */
/* Single resident state (CudaDeviceState, src/cuda.rs:1277-1484)            */
/* ------------------------------------------------------------------------ */
/* StateVector::zero (src/lib.rs:930) allocated directly on the device. */
qf_status qf_state_zero(qf_backend* backend, uint32_t qubits, int32_t precision, qf_state** out);
/* CudaBackend::upload, src/cuda.rs:844-867. */
/* StateVector::zero_with_memory_budget, src/lib.rs:940-966: |0...0> only when
 * the amplitude payload (2^qubits x 16 bytes complex128, 8 bytes complex64)
 * fits budget_bytes; errors in the reference's order: ZeroQubits,
 * SimulationSizeOverflow (64 or more qubits), MemoryBudgetExceeded, then the
 * device limits of qf_state_zero. */
qf_status qf_state_zero_with_memory_budget(qf_backend* backend, uint32_t qubits, int32_t precision,
                                           uint64_t budget_bytes, qf_state** out);
qf_status qf_state_upload(qf_backend* backend, const qf_c128* amps, uint32_t qubits, qf_state** out);
/* complex64 upload (no reference counterpart: the reference is complex128 only). */
qf_status qf_state_upload_c64(qf_backend* backend, const qf_c64* amps, uint32_t qubits, qf_state** out);
/* CudaDeviceState::download, src/cuda.rs:1459-1470 (complex64 states are widened). */
qf_status qf_state_download(const qf_state* state, qf_c128* out);
/* Narrow download of a complex64 state. */
qf_status qf_state_download_c64(const qf_state* state, qf_c64* out);
/* Drop for CudaDeviceState, src/cuda.rs:1473-1484. */
qf_status qf_state_destroy(qf_state* state);
/* CudaDeviceState::qubits, src/cuda.rs:1279. */
qf_status qf_state_qubits(const qf_state* state, uint32_t* out);
qf_status qf_state_precision(const qf_state* state, int32_t* out);
/* CudaDeviceState::copy_from_device, src/cuda.rs:1444-1456; states of different
 * precisions convert on the device (complex64 widened exactly, complex128
 * rounded to nearest float). */
qf_status qf_state_copy_from(qf_state* dst, const qf_state* src);
/* Device address of the amplitudes (interop, e.g. wrapping as a torch tensor). */
qf_status qf_state_device_pointer(const qf_state* state, void** out);

/* CudaDeviceState::apply_gate, src/cuda.rs:1296-1323 (StateVector::apply, src/lib.rs:1017). */
qf_status qf_state_apply_gate(qf_state* state, int32_t gate, uint32_t target, double theta);
/* apply_matrix_kernel with a caller matrix m = {m00, m01, m10, m11} as re,im pairs
 * (src/cuda.rs:114-128). */
qf_status qf_state_apply_matrix(qf_state* state, uint32_t target, const double* m8);
/* CudaDeviceState::apply_rotation_chain, src/cuda.rs:1325-1347. */
qf_status qf_state_apply_rotation_chain(qf_state* state, const qf_rotation* rotations, size_t count);
/* apply_cnot / apply_cz / apply_swap, src/cuda.rs:1349-1382 (src/lib.rs:1062-1122). */
qf_status qf_state_apply_2q(qf_state* state, int32_t op, uint32_t a, uint32_t b);
/* Exact controlled phase diag(1,1,1,e^{i phi}) (the gate algorithms.rs:139-145
 * lowers to Rz/CNOT up to the global phase e^{-i phi/4}). */
qf_status qf_state_apply_cp(qf_state* state, uint32_t control, uint32_t target, double phi);
/* CudaBackend::run_circuit / Circuit::run on a resident state
 * (src/cuda.rs:1050-1083, src/lib.rs:2786-2802).  plan may be NULL; a plan
 * must come from qf_plan_build of this very op list and width, otherwise
 * QF_INVALID_ARGUMENT is returned before anything is launched (a
 * parameterized or density plan, or one built for other gates or angles). */
qf_status qf_state_run_circuit(qf_state* state, const qf_op* ops, size_t count, const qf_plan* plan);

/* StateVector::norm_sqr, src/lib.rs:1000-1002 (FP64, compensated). */
/* Amplitudes at `count` basis indices as complex128 (StateVector::amplitudes()[i],
 * src/lib.rs:992-994) without downloading the whole state; an index >= 2^qubits
 * is QF_INVALID_ARGUMENT before any launch. */
qf_status qf_state_gather(const qf_state* state, const uint64_t* indices, size_t count, qf_c128* out);
/* The device's in-place streaming rate over the state's bytes (no reference
 * counterpart; a measurement aid): every amplitude is read and written back
 * unchanged, `launches` timed launches after one warm launch, rate =
 * 2 x state bytes / average launch time.  variant 0: one 16-byte word per lane
 * per grid-stride step, 1: nontemporal loads and stores, 2: four words per lane
 * per step, 3: one contiguous 64 KiB chunk per workgroup, 4: eight words per
 * lane per step. */
qf_status qf_state_stream_rate(qf_state* state, int32_t variant, uint32_t launches, double* gbs, double* avg_ms);
qf_status qf_state_norm_sqr(const qf_state* state, double* out);
/* StateVector::probabilities, src/lib.rs:1144-1146. */
qf_status qf_state_probabilities(const qf_state* state, double* out);
/* StateVector::expectation{,_x,_y,_z}, src/lib.rs:1149-1318. */
qf_status qf_state_expectation(const qf_state* state, int32_t observable, uint32_t qubit, double* out);
/* PauliSum::expectation_ket, src/analytic.rs:2775-2778 (FP64, compensated). */
qf_status qf_state_expectation_pauli_sum(const qf_state* state, const qf_pauli_term* terms, size_t count,
                                         double* out_re, double* out_im);
/* qf_state_expectation_pauli_sum without allocating: no reference
 * counterpart.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * the reduction partials, the term table and the result live in a fixed
 * device array of 164096 bytes per GPU (held until the process ends), so the
 * call makes no hipMalloc / hipFree.  At most 4096 terms (QF_INVALID_ARGUMENT
 * beyond); same kernels and summation order as
 * qf_state_expectation_pauli_sum, so the results are identical.  Every call
 * on one GPU shares that array: calls that overlap on one GPU (states of
 * different backends used from several threads) race. */
qf_status qf_state_expectation_pauli_sum_noalloc(const qf_state* state, const qf_pauli_term* terms, size_t count,
                                                 double* out_re, double* out_im);
/* <a|b> (Bra::inner, src/quantum.rs:327-337), FP64 compensated. */
qf_status qf_state_inner(const qf_state* a, const qf_state* b, double* out_re, double* out_im);

/* ------------------------------------------------------------------------ */
/* Batches (CudaDeviceStateBatch, src/cuda.rs:1486-2005)                     */
/* ------------------------------------------------------------------------ */
/* A batch of `count` states with widths qubits[0..count), stored back to back
 * (amps: the states' 2^qubits[i] amplitudes one after the other).  Widths may
 * differ (no reference counterpart: the reference batches equal widths only);
 * every batch operation then runs over each run of consecutive equal-width
 * states, and validates a gate against every width before anything launches.
 * EmptyBatch for count 0. */
qf_status qf_batch_upload(qf_backend* backend, const qf_c128* amps, const uint32_t* qubits, uint32_t count,
                          qf_batch** out);
/* CudaBackend::upload_batch, src/cuda.rs:871-954: equal widths only, with the
 * reference's EmptyBatch / IncompatibleBatch (src/cuda.rs:875-877) errors. */
qf_status qf_batch_upload_equal(qf_backend* backend, const qf_c128* amps, const uint32_t* qubits, uint32_t count,
                                qf_batch** out);
/* |0..0> batch of `count` equal-width states, and of states of the given widths. */
qf_status qf_batch_zero(qf_backend* backend, uint32_t qubits, uint32_t count, int32_t precision, qf_batch** out);
qf_status qf_batch_zero_uneven(qf_backend* backend, const uint32_t* qubits, uint32_t count, int32_t precision,
                               qf_batch** out);
qf_status qf_batch_destroy(qf_batch* batch);
/* CudaDeviceStateBatch::qubits / batch_size, src/cuda.rs:1488-1496 (qubits: the
 * widest state of an uneven batch). */
qf_status qf_batch_shape(const qf_batch* batch, uint32_t* qubits, uint32_t* count);
/* Width of every state (out is [count]). */
qf_status qf_batch_widths(const qf_batch* batch, uint32_t* out);
/* apply_gate / apply_cnot / apply_cz / apply_swap on every state, src/cuda.rs:1819-1936. */
qf_status qf_batch_apply_gate(qf_batch* batch, int32_t gate, uint32_t target, double theta);
qf_status qf_batch_apply_2q(qf_batch* batch, int32_t op, uint32_t a, uint32_t b);
/* CudaDeviceStateBatch::run_circuit, src/cuda.rs:1949-1970 (a plan only for an
 * equal-width batch: QF_INCOMPATIBLE_BATCH otherwise). */
qf_status qf_batch_run_circuit(qf_batch* batch, const qf_op* ops, size_t count, const qf_plan* plan);
/* CudaDeviceStateBatch::download, src/cuda.rs:1973-1992; out holds every state's
 * amplitudes back to back. */
qf_status qf_batch_download(const qf_batch* batch, qf_c128* out);
/* Per-state Pauli-sum energies (FP64 compensated). out is [count]. */
qf_status qf_batch_expectation_pauli_sum(const qf_batch* batch, const qf_pauli_term* terms, size_t nterms,
                                         double* out);

/* ------------------------------------------------------------------------ */
/* Gradients (src/gradients.rs, src/lib.rs:3484-3704)                        */
/* ------------------------------------------------------------------------ */
/* Options of a gradient call.  Defaults (all zero except fd_step) follow
 * GradientOptions::default, src/gradients.rs:155-165 (Auto, step 1e-3, 0 shots). */
typedef struct {
    int32_t strategy;      /* QF_GRAD_* */
    int32_t precision;     /* QF_C128 or QF_C64 amplitudes (reductions stay FP64) */
    double fd_step;        /* central-difference step (QF_GRAD_FINITE_DIFFERENCE) */
    uint64_t shots;        /* only informs QF_GRAD_AUTO (GradientStrategy::select, src/gradients.rs:81-89) */
    uint32_t tile_qubits;  /* forward tile (0 = default); the reverse sweep uses one qubit less */
    uint32_t chunk;        /* parameter sets per device wave (0 = as many as fit) */
} qf_gradient_options;

/* ParameterizedCircuit::gradient_with_inputs, src/gradients.rs:232-312, with
 * the observable as a Pauli sum (SingleQubitObservable = one term, weight 1).
 * value_out gets <obs>; strategy_used the resolved rule; evaluations the
 * GradientEstimate::evaluations count (2 for adjoint/backprop, 2 x rotation
 * occurrences for shift / finite differences).  Errors as the reference:
 * InvalidInputCount / NonFiniteParameter (inputs), InvalidParameterCount /
 * NonFiniteParameter (parameters), QubitOutOfRange (observable, gates),
 * OperationRequiresDensityMatrix (channels), InvalidFiniteDifferenceStep. */
qf_status qf_gradient(qf_backend* backend, const qf_pop* ops, size_t nops, uint32_t nparams, uint32_t ninputs,
                      const double* inputs, size_t inputs_len, const double* params, size_t params_len,
                      uint32_t qubits, const qf_pauli_term* obs, size_t nterms, const qf_gradient_options* options,
                      double* grad_out, double* value_out, int32_t* strategy_used, uint64_t* evaluations);

/* Batched gradients of one circuit at `sets` parameter sets (params is
 * [sets][nparams], grads [sets][nparams], energies [sets]) with shared inputs.
 * QF_GRAD_ADJOINT runs the two-state reverse sweep (psi is uncomputed, lambda
 * carried back; replaces adjoint_gradient_with_inputs, src/lib.rs:3564-3704);
 * QF_GRAD_PARAMETER_SHIFT runs 2 circuits per rotation occurrence
 * (src/lib.rs:3516-3546); QF_GRAD_FINITE_DIFFERENCE runs the central
 * difference (src/gradients.rs:344-375); QF_GRAD_BACKPROP the store-every-state
 * reverse mode (src/gradients.rs:392-506). */
qf_status qf_batch_gradient(qf_backend* backend, const qf_pop* ops, size_t nops, uint32_t nparams, uint32_t ninputs,
                            const double* inputs, uint32_t qubits, const double* params, uint32_t sets,
                            const qf_pauli_term* obs, size_t nterms, const qf_gradient_options* options,
                            double* energies, double* grads);

/* What the last batched gradient or expectation call of a backend ran (no
 * reference counterpart; a measurement record).  Pass counts and kernel
 * families of the forward and reverse plans, the costate kernel, the
 * algorithmic bytes of each phase (forward pass: 2 x 2^n x s per set, costate
 * lambda = O psi: 2 x 2^n x s, reverse two-state pass: 4 x 2^n x s; s = 16
 * complex128, 8 complex64), and with profiling on (qf_backend_set_profiling)
 * the device time of each phase from HIP events on the backend's stream,
 * summed over the call's memory waves. */
typedef struct {
    uint32_t sets, waves;
    uint32_t forward_passes, reverse_passes;
    uint32_t forward_kernels, reverse_kernels;  /* QF_KERNEL_* masks */
    uint32_t costate_kernel;                    /* 1 windowed coefficient tables, 2 per-term signs */
    uint32_t profiled;                          /* 1 when the *_ms fields were measured */
    double forward_bytes, costate_bytes, reverse_bytes;
    double forward_ms, costate_ms, reverse_ms, finish_ms;
} qf_gradient_profile;

/* Profiling of batched gradient calls on this backend (0 off, the default;
 * 1 on: HIP events around each phase, read back after each wave). */
qf_status qf_backend_set_profiling(qf_backend* backend, int32_t on);
qf_status qf_backend_last_gradient_profile(const qf_backend* backend, qf_gradient_profile* out);

/* <obs> at `sets` parameter sets (ParameterizedCircuit::run + expectation). */
qf_status qf_batch_expectation(qf_backend* backend, const qf_pop* ops, size_t nops, uint32_t nparams,
                               uint32_t ninputs, const double* inputs, uint32_t qubits, const double* params,
                               uint32_t sets, const qf_pauli_term* obs, size_t nterms,
                               const qf_gradient_options* options, double* energies);

/* ParameterizedCircuit::run_with_inputs, src/lib.rs:3108-3121, into a resident state. */
qf_status qf_state_run_parameterized(qf_state* state, const qf_pop* ops, size_t nops, uint32_t nparams,
                                     uint32_t ninputs, const double* inputs, size_t inputs_len, const double* params,
                                     size_t params_len);

/* One circuit of an uneven batch: its own ops, width, parameters and observable
 * (no inputs).  No reference counterpart: the reference batches equal widths
 * only (IncompatibleBatch, src/cuda.rs:875-877). */
typedef struct {
    const qf_pop* ops;
    size_t nops;
    uint32_t nparams;
    uint32_t qubits;
    const double* params;          /* [nparams] */
    const qf_pauli_term* obs;
    size_t nterms;
} qf_uneven_instance;

typedef struct {
    uint32_t waves;                /* device-memory waves the instances were packed into */
    uint32_t graph_nodes;          /* kernel launches in the captured per-wave graph */
    uint32_t graph_replays;        /* replays of that graph (one per wave) */
    uint32_t grouped_instances;    /* instances run through the grouped graph */
    uint32_t tile_qubits, low_qubits, reg_bits;  /* autotuned forward tile shape (fusion width = tile_qubits) */
    uint32_t reverse_tile_qubits;                /* autotuned reverse-sweep tile (two states per tile) */
    double tune_ms;                /* tuned forward time of the widest instance */
    double alg_bytes;              /* algorithmic HBM bytes of the grouped instances' passes: per instance
                                      forward passes x 2 x 2^n x s, costate 2 x 2^n x s, reverse passes
                                      x 4 x 2^n x s (s = 16 complex128, 8 complex64) */
    double device_ms;              /* device time of the graph replays (HIP events), summed over waves */
    uint32_t qaoa_instances;       /* instances in QAOA form, run on the QAOA tile engine (12 <= n <= 30) */
    uint32_t qaoa_forward_passes;  /* most forward passes of a QAOA-form instance */
} qf_uneven_stats;

/* Energies <obs_i> and adjoint gradients of `count` parameterized circuits of
 * different widths.  grads is the concatenation of every instance's [nparams].
 * Each instance is validated like qf_gradient.  stats may be NULL.  Instances
 * are packed into device-memory waves of at most 80% of the free memory, or of
 * the environment variable QF_UNEVEN_WAVE_BYTES bytes when that is smaller.
 * Instances in QAOA form (12 to 30 qubits; a Hadamard on every qubit, then
 * layers of cost rotations CNOT(u, v) Rz(v) CNOT(u, v) followed by an X
 * rotation on every qubit, every rotation one parameter occurrence; an
 * observable of real Z and Z Z words and a constant) run on a tile engine of
 * their own with the graph as run-time data; QF_QAOA_ENGINE=0 sends them
 * through the general engine. */
qf_status qf_uneven_adjoint_gradient(qf_backend* backend, const qf_uneven_instance* instances, uint32_t count,
                                     int32_t precision, double* energies, double* grads, qf_uneven_stats* stats);

/* Energies <obs_i> of `count` parameterized circuits of different widths (no
 * gradients): the uneven engine's forward passes and energy reductions only,
 * one psi per instance (no costate), so each wave holds twice the instances of
 * qf_uneven_adjoint_gradient.  Validation and stats as there. */
qf_status qf_uneven_expectation(qf_backend* backend, const qf_uneven_instance* instances, uint32_t count,
                                int32_t precision, double* energies, qf_uneven_stats* stats);

/* Times the candidate register-kernel tile shapes of a circuit on the device
 * and returns the fastest as plan options (cached per backend process). */
qf_status qf_autotune(qf_backend* backend, const qf_op* ops, size_t count, uint32_t qubits, int32_t precision,
                      qf_plan_options* best, double* best_ms);

/* ------------------------------------------------------------------------ */
/* Measurement and shots                                                      */
/* ------------------------------------------------------------------------ */
/* CudaDeviceState::measure, src/cuda.rs:1386-1442: P(1) of `qubit` (FP64),
 * outcome = uniform < P(1), InvalidProbability when the drawn branch has
 * probability <= DBL_EPSILON, collapse dividing the kept branch by
 * sqrt(probability).  The reference draws `uniform` from its SmallRng inside
 * the call; here the caller draws it (rng.next_f64()) and passes it. */
qf_status qf_state_measure(qf_state* state, uint32_t qubit, double uniform, uint8_t* outcome);
/* CudaDeviceStateBatch::measure_with_uniforms, src/cuda.rs:1511-1609: one
 * uniform in [0, 1) per state (else InvalidBatchSamples); every outcome is
 * checked before any state collapses. */
qf_status qf_batch_measure_with_uniforms(qf_batch* batch, uint32_t qubit, const double* uniforms, size_t count,
                                         uint8_t* outcomes);
/* CudaDeviceStateBatch::apply_gate_masked (the classically conditioned gate
 * of a shot batch, src/cuda.rs:1239-1249): the gate on states with enabled[i]. */
qf_status qf_batch_apply_gate_masked(qf_batch* batch, int32_t gate, uint32_t target, double theta,
                                     const uint8_t* enabled, size_t count);
/* One Kraus-trajectory step of a noise channel (QF_CHANNEL_*, p0 / p1 as
 * qf_density_apply_channel) on qubit `target` of every state, one uniform in
 * [0, 1) per state (else InvalidBatchSamples), with the branch rules of
 * StateVector::apply_channel_trajectory (src/lib.rs:1356-1505): depolarizing
 * applies X, Y or Z; the damping channels take each state's norm and P(1)
 * and renormalize the drawn branch (InvalidProbability for a branch of
 * probability <= DBL_EPSILON).  The per-batch step of
 * run_circuit_shots_batched (CudaDeviceStateBatch::
 * apply_channel_trajectory_with_uniforms, src/cuda.rs:1670-1790, which is
 * private there); ThermalDamping is included. */
qf_status qf_batch_apply_channel_trajectory(qf_batch* batch, int32_t channel, uint32_t target, double p0, double p1,
                                            const double* uniforms, size_t count);
/* CudaBackend::run_circuit_shots (batched = 0, src/cuda.rs:1085-1159) and
 * run_circuit_shots_batched (batched = 1, src/cuda.rs:1161-1269) from a
 * complex128 initial state: SmallRng(seed) uniforms in the CPU's shot-major
 * order (Circuit::run_shots, src/lib.rs:2843-2911), so the registers equal the
 * CPU's.  Writes the classical register of every shot, [shots][width] bytes
 * (least-significant classical bit first), into `registers`; the caller
 * counts them (ShotResult::counts).  With registers == NULL only validates and
 * sets *classical_width.  Noise channels: the batched form runs Kraus
 * trajectories (one uniform per channel and shot in the same shot-major order,
 * the initial state normalized first, branch rules of
 * StateVector::apply_channel_trajectory, src/lib.rs:1356-1505, so the
 * registers equal Circuit::run_shots_with_noise); the sequential form rejects
 * them with OperationRequiresDensityMatrix (src/cuda.rs:1109-1113).  Errors:
 * InvalidShotCount, ClassicalBitOutOfRange, the gate errors,
 * InvalidProbability. */
qf_status qf_run_circuit_shots(qf_backend* backend, const qf_c128* initial, uint32_t qubits, const qf_op* ops,
                               size_t nops, uint64_t shots, uint64_t seed, int32_t batched, uint8_t* registers,
                               uint32_t* classical_width);

/* ------------------------------------------------------------------------ */
/* Density matrices (CudaDeviceDensityMatrix, src/cuda.rs:2007-2237)         */
/* ------------------------------------------------------------------------ */
/* A density matrix of n qubits is held as its row-major entries, which is a
 * state vector of 2n qubits: entry (row, col) at index row * 2^n + col, row
 * qubit q at vector bit n + q, column qubit q at bit q.  Gates, channels and
 * whole circuits run in place (no scratch matrix) through the fused tile
 * planner: rho -> U rho U^dagger is U on the row bits and conj(U) on the
 * column bits, and a channel is one 4x4 superoperator on the bit pair
 * (q, n + q).  Complex128 only; 1 <= n <= 15 (validate_density_size,
 * src/cuda.rs:2239-2253: ZeroQubits, StateTooLarge). */

/* CudaBackend::upload_density, src/cuda.rs:969-1022: `entries` row major,
 * len = 4^qubits (else StateTooLarge, as the reference). */
qf_status qf_density_upload(qf_backend* backend, const qf_c128* entries, size_t len, uint32_t qubits,
                            qf_density** out);
/* DensityMatrix::from_state (src/lib.rs:1642-1656) of a resident complex128
 * state, computed on the device: entries = psi[row] * conj(psi[col]). */
qf_status qf_density_from_state(const qf_state* state, qf_density** out);
/* DensityMatrix::maximally_mixed (src/lib.rs:1659-1704) on the device. */
qf_status qf_density_maximally_mixed(qf_backend* backend, uint32_t qubits, qf_density** out);
/* DensityMatrix::maximally_mixed_with_memory_budget, src/lib.rs:1668-1700: the
 * budget covers the 4^qubits complex128 entries. */
qf_status qf_density_maximally_mixed_with_memory_budget(qf_backend* backend, uint32_t qubits, uint64_t budget_bytes,
                                                        qf_density** out);
qf_status qf_density_destroy(qf_density* density);
/* CudaDeviceDensityMatrix::qubits / download, src/cuda.rs:2009-2011, 2215-2227. */
qf_status qf_density_qubits(const qf_density* density, uint32_t* out);
qf_status qf_density_download(const qf_density* density, qf_c128* out, size_t len);
/* CudaDeviceDensityMatrix::apply_gate, src/cuda.rs:2015-2029: QubitOutOfRange,
 * NonFiniteAngle for a non-finite rotation angle. */
qf_status qf_density_apply_gate(qf_density* density, int32_t gate, uint32_t target, double theta);
/* apply_cnot / apply_cz / apply_swap, src/cuda.rs:2069-2121 (pair errors of
 * the CPU path: QubitOutOfRange, InvalidQubitPair). */
qf_status qf_density_apply_cnot(qf_density* density, uint32_t control, uint32_t target);
qf_status qf_density_apply_cz(qf_density* density, uint32_t first, uint32_t second);
qf_status qf_density_apply_swap(qf_density* density, uint32_t first, uint32_t second);
/* CudaDeviceDensityMatrix::apply_channel, src/cuda.rs:2123-2159, channel
 * QF_CHANNEL_*, p0 = first parameter, p1 = ThermalDamping's beta_delta:
 * QubitOutOfRange, InvalidProbability (validate_probability and
 * NoiseChannel::thermal_rates, src/lib.rs:2229-2235, 1611-1627).
 * ThermalDamping runs here as on the CPU path (the reference's CUDA backend
 * reports UnsupportedOperation for it). */
qf_status qf_density_apply_channel(qf_density* density, int32_t channel, uint32_t target, double p0, double p1);
/* CudaDeviceDensityMatrix::measure, src/cuda.rs:2161-2213, with the uniform
 * passed in (the caller's rng.next_f64()): outcome = uniform < P(1) (clamped
 * to [0, 1]), InvalidProbability when the drawn branch has probability <=
 * DBL_EPSILON, kept entries divided by the probability as
 * DensityMatrix::measure does (src/lib.rs:1924-1950). */
qf_status qf_density_measure(qf_density* density, uint32_t qubit, double uniform, uint8_t* outcome);
/* DensityMatrix::trace, purity, expectation and probability_one
 * (src/lib.rs:1732-1921) as FP64 reductions over the resident entries. */
qf_status qf_density_trace(const qf_density* density, double* re, double* im);
qf_status qf_density_purity(const qf_density* density, double* out);
qf_status qf_density_expectation(const qf_density* density, int32_t observable, uint32_t qubit, double* out);
qf_status qf_density_probability_one(const qf_density* density, uint32_t qubit, double* out);
/* Runs a circuit (gates, CNOT/CZ/SWAP, controlled phase, channels) on a
 * resident density matrix in place, fused into tile passes over the 2n-qubit
 * view.  Ops are validated in order before anything runs; measurements and
 * conditional gates give OperationRequiresShotExecution (Circuit::run_density,
 * src/lib.rs:2804-2819). */
qf_status qf_density_run_circuit(qf_density* density, const qf_op* ops, size_t nops);
/* CudaBackend::run_density_circuit, src/cuda.rs:1025-1047: upload, run,
 * download into `out` (4^qubits entries). */
qf_status qf_run_density_circuit(qf_backend* backend, const qf_c128* initial, size_t len, uint32_t qubits,
                                 const qf_op* ops, size_t nops, qf_c128* out, size_t out_len);
/* The plan qf_density_run_circuit would run (host only; for inspection,
 * qf_plan_compile and repeated runs), and a run of such a plan. */
qf_status qf_density_plan_build(const qf_op* ops, size_t nops, uint32_t qubits, const qf_plan_options* options,
                                qf_plan** out);
qf_status qf_density_run_plan(qf_density* density, const qf_plan* plan);

/* ------------------------------------------------------------------------ */
/* Sparse Lanczos (sparse_lanczos_gpu, src/quantum.rs:1853-2023; the TFIM     */
/* Lanczos of DMRG, src/tensor_network.rs:1535-1689 and 1692-2025)           */
/* ------------------------------------------------------------------------ */
/* Lowest Ritz value of a real symmetric operator in CSR form (the reference
 * keeps the real part of every value): the start vector of both reference
 * routines (seed 42, seed = seed * 6364136223846793005 + 1, x = seed / 2^64 -
 * 0.5, normalized), m = min(max_iterations, dim, 100) steps with full
 * reorthogonalization, stop when beta < tolerance, the tridiagonal model cut at
 * its first alpha equal to 0.0, and its lowest eigenvalue.  The Krylov basis
 * stays in device memory; reorthogonalization is one multi-dot and one fused
 * update-and-norm pass over the basis per step (a second pass when the norm
 * drops below 1/sqrt(2) of the three-term remainder's).  alpha / beta (capacity 100 each) and
 * size may be NULL.  Errors: InvalidArgument (dim or max_iterations zero,
 * inconsistent CSR arrays, a first alpha of exactly 0.0), StateTooLarge
 * (dim >= 2^31 or nnz >= 2^32), MemoryAllocationFailed. */
qf_status qf_sparse_lanczos(qf_backend* backend, const uint64_t* row_offsets, const uint64_t* column_indices,
                            const qf_c128* values, uint64_t dim, uint64_t max_iterations, double tolerance,
                            double* lowest, double* alpha, double* beta, uint64_t* size);
/* The TFIM chain H = -sum_i Z_i Z_{i+1} - sum_i X_i of `sites` sites (site i is
 * index bit sites-1-i), matrix-free: m = min(sites >= 24 ? 30 : 100, 2^sites)
 * steps, reorthogonalization against every stored vector, stop when beta <
 * 1e-12, a tridiagonal model of m rows (zero beyond the last step), its lowest
 * eigenvalue.  2 <= sites <= 30 (InvalidArgument below, MemoryAllocationFailed
 * when the basis does not fit). */
qf_status qf_tfim_lanczos(qf_backend* backend, uint32_t sites, double* lowest, double* alpha, double* beta,
                          uint64_t* size);

/* ------------------------------------------------------------------------ */
/* Planning (no reference counterpart; see qf_plan_options)                  */
/* ------------------------------------------------------------------------ */
/* Builds the fused tile-pass schedule of a circuit (host only, no device use).
 * options may be NULL (defaults). */
qf_status qf_plan_build(const qf_op* ops, size_t count, uint32_t qubits, const qf_plan_options* options,
                        qf_plan** out);
qf_status qf_plan_stats_get(const qf_plan* plan, qf_plan_stats* out);
/* JSON text of the plan's passes (inspection and planner tests).  Writes at
 * most cap bytes including the terminating NUL; *needed gets the full length + 1. */
qf_status qf_plan_describe(const qf_plan* plan, char* buf, size_t cap, size_t* needed);
/* Plan of a parameterized circuit as the gradient engine runs it (host only):
 * rotation occurrences read (cos, sin) of their half angle from a per-instance
 * table in occurrence order; reverse != 0 gives the inverted circuit in reverse
 * order with every occurrence recorded (the adjoint reverse sweep).  Validates
 * like qf_gradient.  For inspection (describe, source, compile). */
qf_status qf_plan_build_parameterized(const qf_pop* ops, size_t nops, uint32_t nparams, uint32_t ninputs,
                                      const double* inputs, size_t inputs_len, uint32_t qubits, int32_t reverse,
                                      const qf_plan_options* options, qf_plan** out);
/* Source of the plan's compiled passes (kernel = 3, or the default for states
 * of 18+ qubits): one generated kernel per register-kernel tile pass, in pass
 * order.  Same buffer convention as qf_plan_describe. */
qf_status qf_plan_source(const qf_plan* plan, int32_t precision, char* buf, size_t cap, size_t* needed);
/* The same in the adjoint form the gradient engine launches for a reverse
 * plan: each kernel also carries the costate lambda and writes, per tile, Im
 * <lambda| G |psi> of every recorded rotation into the partial table (row =
 * instance * 2^(n - tile) + tile block, column = occurrence slot).  For
 * inspection; no counterpart in the reference, whose adjoint runs on the CPU
 * (src/gradients.rs). */
qf_status qf_plan_adjoint_source(const qf_plan* plan, int32_t precision, char* buf, size_t cap, size_t* needed);
/* Compiles the plan's passes for gfx950 now (host only; otherwise the first
 * run compiles them).  compile_ms and from_cache may be NULL; from_cache is 1
 * when the code objects came from the on-disk cache (QF_JIT_CACHE, default
 * $HOME/.cache/quantumforge_amd/jit; "0" disables it). */
qf_status qf_plan_compile(const qf_plan* plan, int32_t precision, double* compile_ms, int32_t* from_cache);
qf_status qf_plan_destroy(qf_plan* plan);

#ifdef __cplusplus
}
#endif

#endif /* QF_HIP_H */
