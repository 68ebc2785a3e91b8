"""ctypes binding of libqfhip.so (the C ABI declared in include/qf_hip.h).

The library is built in-tree (``python -m quantumforge_amd.build``).  There is
no fallback: if the library is missing, importing anything that needs it
raises, so a GPU run can never silently take another path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libqfhip.so")

# Status codes, include/qf_hip.h
STATUS = {
    0: "Ok", 1: "ZeroQubits", 2: "QubitOutOfRange", 3: "InvalidQubitPair", 4: "NonFiniteAngle",
    5: "StateTooLarge", 6: "EmptyBatch", 7: "IncompatibleBatch", 8: "InvalidParameterCount",
    9: "InvalidInputCount", 10: "NonFiniteParameter", 11: "MemoryAllocationFailed", 12: "Driver",
    13: "OperationRequiresDensityMatrix", 14: "OperationRequiresShotExecution", 15: "InvalidArgument",
    16: "IncompatibleDeviceState", 17: "StateSizeNotPowerOfTwo", 18: "InvalidBatchSamples",
    19: "MemoryBudgetExceeded", 20: "InvalidFiniteDifferenceStep", 21: "SimulationSizeOverflow",
    22: "InvalidProbability", 23: "InvalidShotCount", 24: "ClassicalBitOutOfRange",
    25: "NonHermitianDensityMatrix", 26: "InvalidSubsystemSelection",
}

# Record layouts identical to the C structs (checked by tests/test_abi.py).
OP_DTYPE = np.dtype([("kind", "<i4"), ("gate", "<i4"), ("a", "<u4"), ("b", "<u4"), ("theta", "<f8"),
                     ("theta2", "<f8")])
POP_DTYPE = np.dtype([("kind", "<i4"), ("gate", "<i4"), ("a", "<u4"), ("b", "<u4"), ("index", "<u4"),
                      ("reserved", "<u4"), ("theta", "<f8")])
PAULI_DTYPE = np.dtype([("re", "<f8"), ("im", "<f8"), ("x_mask", "<u8"), ("z_mask", "<u8")])
ROT_DTYPE = np.dtype([("axis", "<i4"), ("qubit", "<u4"), ("theta", "<f8")])


class PlanOptions(C.Structure):
    _fields_ = [("fuse", C.c_int32), ("tile_qubits", C.c_uint32), ("low_qubits", C.c_uint32),
                ("max_block", C.c_uint32), ("kernel", C.c_int32), ("reg_bits", C.c_uint32),
                ("precision", C.c_int32)]


class PlanStats(C.Structure):
    _fields_ = [("passes", C.c_uint64), ("ops_in", C.c_uint64), ("ops_applied", C.c_uint64),
                ("dense_blocks", C.c_uint64), ("max_ops_per_pass", C.c_uint64), ("phases", C.c_uint64),
                ("kernels_run", C.c_uint32), ("runtime_rotations", C.c_uint32)]


# qf_plan_stats.kernels_run bits (QF_KERNEL_*)
KERNEL_BITS = {1: "sweep", 2: "lds_tile", 4: "register_tile", 8: "compiled", 16: "mfma"}


class GradOptions(C.Structure):
    _fields_ = [("strategy", C.c_int32), ("precision", C.c_int32), ("fd_step", C.c_double),
                ("shots", C.c_uint64), ("tile_qubits", C.c_uint32), ("chunk", C.c_uint32)]


class GradientProfile(C.Structure):
    _fields_ = [("sets", C.c_uint32), ("waves", C.c_uint32), ("forward_passes", C.c_uint32),
                ("reverse_passes", C.c_uint32), ("forward_kernels", C.c_uint32), ("reverse_kernels", C.c_uint32),
                ("costate_kernel", C.c_uint32), ("profiled", C.c_uint32), ("forward_bytes", C.c_double),
                ("costate_bytes", C.c_double), ("reverse_bytes", C.c_double), ("forward_ms", C.c_double),
                ("costate_ms", C.c_double), ("reverse_ms", C.c_double), ("finish_ms", C.c_double)]


class UnevenInstance(C.Structure):
    _fields_ = [("ops", C.c_void_p), ("nops", C.c_size_t), ("nparams", C.c_uint32), ("qubits", C.c_uint32),
                ("params", C.c_void_p), ("obs", C.c_void_p), ("nterms", C.c_size_t)]


class UnevenStats(C.Structure):
    _fields_ = [("waves", C.c_uint32), ("graph_nodes", C.c_uint32), ("graph_replays", C.c_uint32),
                ("grouped_instances", C.c_uint32), ("tile_qubits", C.c_uint32), ("low_qubits", C.c_uint32),
                ("reg_bits", C.c_uint32), ("reverse_tile_qubits", C.c_uint32), ("tune_ms", C.c_double),
                ("alg_bytes", C.c_double), ("device_ms", C.c_double), ("qaoa_instances", C.c_uint32),
                ("qaoa_forward_passes", C.c_uint32)]


P = C.c_void_p
PP = C.POINTER(C.c_void_p)
u32, i32, u64, dbl, sz = C.c_uint32, C.c_int32, C.c_uint64, C.c_double, C.c_size_t

# name -> (restype, argtypes); restype None means qf_status (int32).
SIGNATURES = {
    "qf_last_error_message": (C.c_char_p, []),
    "qf_status_name": (C.c_char_p, [i32]),
    "qf_backend_create": (None, [i32, PP]),
    "qf_backend_destroy": (None, [P]),
    "qf_backend_synchronize": (None, [P]),
    "qf_backend_release_cached": (None, [P]),
    "qf_backend_device_ordinal": (None, [P, C.POINTER(i32)]),
    "qf_backend_set_plan_options": (None, [P, C.POINTER(PlanOptions)]),
    "qf_backend_stream": (None, [P, PP]),
    "qf_backend_set_profiling": (None, [P, i32]),
    "qf_backend_last_gradient_profile": (None, [P, C.POINTER(GradientProfile)]),
    "qf_state_zero": (None, [P, u32, i32, PP]),
    "qf_state_zero_with_memory_budget": (None, [P, u32, i32, u64, PP]),
    "qf_state_gather": (None, [P, P, sz, P]),
    "qf_state_stream_rate": (None, [P, i32, u32, C.POINTER(dbl), C.POINTER(dbl)]),
    "qf_density_maximally_mixed_with_memory_budget": (None, [P, u32, u64, PP]),
    "qf_state_upload": (None, [P, P, u32, PP]),
    "qf_state_upload_c64": (None, [P, P, u32, PP]),
    "qf_state_download": (None, [P, P]),
    "qf_state_download_c64": (None, [P, P]),
    "qf_state_destroy": (None, [P]),
    "qf_state_qubits": (None, [P, C.POINTER(u32)]),
    "qf_state_precision": (None, [P, C.POINTER(i32)]),
    "qf_state_copy_from": (None, [P, P]),
    "qf_state_device_pointer": (None, [P, PP]),
    "qf_state_apply_gate": (None, [P, i32, u32, dbl]),
    "qf_state_apply_matrix": (None, [P, u32, P]),
    "qf_state_apply_rotation_chain": (None, [P, P, sz]),
    "qf_state_apply_2q": (None, [P, i32, u32, u32]),
    "qf_state_apply_cp": (None, [P, u32, u32, dbl]),
    "qf_state_run_circuit": (None, [P, P, sz, P]),
    "qf_state_norm_sqr": (None, [P, C.POINTER(dbl)]),
    "qf_state_probabilities": (None, [P, P]),
    "qf_state_expectation": (None, [P, i32, u32, C.POINTER(dbl)]),
    "qf_state_expectation_pauli_sum": (None, [P, P, sz, C.POINTER(dbl), C.POINTER(dbl)]),
    "qf_state_inner": (None, [P, P, C.POINTER(dbl), C.POINTER(dbl)]),
    "qf_state_expectation_pauli_sum_noalloc": (None, [P, P, sz, C.POINTER(dbl), C.POINTER(dbl)]),
    "qf_batch_upload": (None, [P, P, P, u32, PP]),
    "qf_batch_upload_equal": (None, [P, P, P, u32, PP]),
    "qf_batch_zero": (None, [P, u32, u32, i32, PP]),
    "qf_batch_zero_uneven": (None, [P, P, u32, i32, PP]),
    "qf_batch_widths": (None, [P, P]),
    "qf_batch_destroy": (None, [P]),
    "qf_batch_shape": (None, [P, C.POINTER(u32), C.POINTER(u32)]),
    "qf_batch_apply_gate": (None, [P, i32, u32, dbl]),
    "qf_batch_apply_2q": (None, [P, i32, u32, u32]),
    "qf_batch_run_circuit": (None, [P, P, sz, P]),
    "qf_batch_download": (None, [P, P]),
    "qf_batch_expectation_pauli_sum": (None, [P, P, sz, P]),
    "qf_plan_build": (None, [P, sz, u32, C.POINTER(PlanOptions), PP]),
    "qf_plan_stats_get": (None, [P, C.POINTER(PlanStats)]),
    "qf_plan_describe": (None, [P, C.c_char_p, sz, C.POINTER(sz)]),
    "qf_state_measure": (None, [P, u32, dbl, C.POINTER(C.c_uint8)]),
    "qf_batch_measure_with_uniforms": (None, [P, u32, P, sz, P]),
    "qf_batch_apply_gate_masked": (None, [P, i32, u32, dbl, P, sz]),
    "qf_run_circuit_shots": (None, [P, P, u32, P, sz, u64, u64, i32, P, C.POINTER(u32)]),
    "qf_batch_apply_channel_trajectory": (None, [P, i32, u32, dbl, dbl, P, sz]),
    "qf_density_upload": (None, [P, P, sz, u32, PP]),
    "qf_density_from_state": (None, [P, PP]),
    "qf_density_maximally_mixed": (None, [P, u32, PP]),
    "qf_density_destroy": (None, [P]),
    "qf_density_qubits": (None, [P, C.POINTER(u32)]),
    "qf_density_download": (None, [P, P, sz]),
    "qf_density_apply_gate": (None, [P, i32, u32, dbl]),
    "qf_density_apply_cnot": (None, [P, u32, u32]),
    "qf_density_apply_cz": (None, [P, u32, u32]),
    "qf_density_apply_swap": (None, [P, u32, u32]),
    "qf_density_apply_channel": (None, [P, i32, u32, dbl, dbl]),
    "qf_density_measure": (None, [P, u32, dbl, C.POINTER(C.c_uint8)]),
    "qf_density_trace": (None, [P, C.POINTER(dbl), C.POINTER(dbl)]),
    "qf_density_purity": (None, [P, C.POINTER(dbl)]),
    "qf_density_expectation": (None, [P, i32, u32, C.POINTER(dbl)]),
    "qf_density_probability_one": (None, [P, u32, C.POINTER(dbl)]),
    "qf_density_run_circuit": (None, [P, P, sz]),
    "qf_run_density_circuit": (None, [P, P, sz, u32, P, sz, P, sz]),
    "qf_density_plan_build": (None, [P, sz, u32, C.POINTER(PlanOptions), PP]),
    "qf_density_run_plan": (None, [P, P]),
    "qf_sparse_lanczos": (None, [P, P, P, P, u64, u64, dbl, C.POINTER(dbl), P, P, C.POINTER(u64)]),
    "qf_tfim_lanczos": (None, [P, u32, C.POINTER(dbl), P, P, C.POINTER(u64)]),
    "qf_plan_build_parameterized": (None, [P, sz, u32, u32, P, sz, u32, i32, C.POINTER(PlanOptions), PP]),
    "qf_plan_source": (None, [P, i32, C.c_char_p, sz, C.POINTER(sz)]),
    "qf_plan_adjoint_source": (None, [P, i32, C.c_char_p, sz, C.POINTER(sz)]),
    "qf_plan_compile": (None, [P, i32, C.POINTER(dbl), C.POINTER(i32)]),
    "qf_plan_destroy": (None, [P]),
    "qf_gradient": (None, [P, P, sz, u32, u32, P, sz, P, sz, u32, P, sz, C.POINTER(GradOptions), P,
                           C.POINTER(dbl), C.POINTER(i32), C.POINTER(u64)]),
    "qf_batch_gradient": (None, [P, P, sz, u32, u32, P, u32, P, u32, P, sz, C.POINTER(GradOptions), P, P]),
    "qf_batch_expectation": (None, [P, P, sz, u32, u32, P, u32, P, u32, P, sz, C.POINTER(GradOptions), P]),
    "qf_state_run_parameterized": (None, [P, P, sz, u32, u32, P, sz, P, sz]),
    "qf_uneven_adjoint_gradient": (None, [P, C.POINTER(UnevenInstance), u32, i32, P, P, C.POINTER(UnevenStats)]),
    "qf_uneven_expectation": (None, [P, C.POINTER(UnevenInstance), u32, i32, P, C.POINTER(UnevenStats)]),
    "qf_autotune": (None, [P, P, sz, u32, i32, C.POINTER(PlanOptions), C.POINTER(dbl)]),
}

_lib = None


def lib() -> C.CDLL:
    """Loads libqfhip.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build the HIP engine first "
                "(python -m quantumforge_amd.build or __graft_entry__.build())")
        handle = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name, None)
            if fn is None:
                continue
            fn.restype = res if res is not None else C.c_int32
            fn.argtypes = args
        _lib = handle
    return _lib


class QuantumError(Exception):
    """Typed error of the engine: ``kind`` is the reference variant name
    (QuantumError / CudaError, src/lib.rs:509-582, src/cuda.rs:3051-3080)."""

    def __init__(self, status: int, message: str):
        self.status = status
        self.kind = STATUS.get(status, "Unknown")
        super().__init__(f"{self.kind}: {message}")


def check(status: int) -> None:
    if status != 0:
        msg = lib().qf_last_error_message()
        raise QuantumError(status, msg.decode() if msg else "")


def ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data) if a is not None and a.size else C.c_void_p(0)
