"""HIP backend: the drop-in for the reference's `cuda.rs` (CudaBackend,
CudaDeviceState, CudaDeviceStateBatch, CudaDeviceDensityMatrix,
src/cuda.rs:484-2237), over the C ABI of include/qf_hip.h."""
from __future__ import annotations

import ctypes as C
import json
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _native as N
from ._native import QuantumError, check, lib, ptr
from .core import (Circuit, DensityMatrix, Gate, NoiseChannel, PauliSum, RotationAxis, ShotResult,
                   SingleQubitObservable, SparseOperator, StateVector)

C128, C64 = 0, 1
_PRECISION = {"complex128": C128, "c128": C128, C128: C128, "complex64": C64, "c64": C64, C64: C64}


def _precision(p) -> int:
    try:
        return _PRECISION[p]
    except KeyError:
        raise ValueError(f"unknown precision {p!r}; use 'complex128' or 'complex64'") from None


class Plan:
    """Fused execution plan of a circuit (qf_plan_build)."""

    def __init__(self, circuit: Circuit, qubits: int, fuse: bool = True, tile_qubits: int = 0,
                 low_qubits: int = 0, kernel: int = 0, reg_bits: int = 0, precision="complex128",
                 max_block: int = 0) -> None:
        """precision picks the default tile shape (complex64 tiles hold twice
        the amplitudes in the same LDS); the plan runs on either precision.
        kernel=4 plans dense max_block-qubit blocks (4 or 5, default 5) for the
        matrix cores."""
        self._ops = circuit.to_ops()
        opts = N.PlanOptions(int(fuse), tile_qubits, low_qubits, max_block, kernel, reg_bits, _precision(precision))
        h = C.c_void_p()
        check(lib().qf_plan_build(ptr(self._ops), len(self._ops), qubits, C.byref(opts), C.byref(h)))
        self._h = h
        self._lib = lib()  # kept for release at interpreter exit
        self.qubits = qubits

    @classmethod
    def parameterized(cls, pc, qubits: int, inputs=None, reverse: bool = False, tile_qubits: int = 0,
                      low_qubits: int = 0, kernel: int = 0, reg_bits: int = 0, precision="complex128") -> "Plan":
        """The plan the gradient engine runs for a ParameterizedCircuit
        (qf_plan_build_parameterized): rotations read their angles per
        instance; reverse gives the inverted, recorded reverse sweep."""
        self = cls.__new__(cls)
        self._ops = pc.to_pops()
        inp = np.ascontiguousarray(np.asarray(inputs if inputs is not None else [], dtype=np.float64))
        opts = N.PlanOptions(1, tile_qubits, low_qubits, 0, kernel, reg_bits, _precision(precision))
        h = C.c_void_p()
        check(lib().qf_plan_build_parameterized(ptr(self._ops), len(self._ops), pc.parameter_count,
                                                pc.input_count, ptr(inp), inp.size, qubits, int(reverse),
                                                C.byref(opts), C.byref(h)))
        self._h = h
        self._lib = lib()  # kept for release at interpreter exit
        self.qubits = qubits
        return self

    @classmethod
    def density(cls, circuit: Circuit, qubits: int, tile_qubits: int = 0, low_qubits: int = 0, kernel: int = 0,
                reg_bits: int = 0) -> "Plan":
        """The plan a density-matrix circuit runs as (qf_density_plan_build):
        tile passes over the 2n-qubit view of the row-major entries."""
        self = cls.__new__(cls)
        self._ops = circuit.to_ops()
        opts = N.PlanOptions(1, tile_qubits, low_qubits, 0, kernel, reg_bits, C128)
        h = C.c_void_p()
        check(lib().qf_density_plan_build(ptr(self._ops), len(self._ops), qubits, C.byref(opts), C.byref(h)))
        self._h = h
        self._lib = lib()  # kept for release at interpreter exit
        self.qubits = 2 * qubits
        return self

    def stats(self) -> dict:
        s = N.PlanStats()
        check(lib().qf_plan_stats_get(self._h, C.byref(s)))
        return {f: int(getattr(s, f)) for f, _ in N.PlanStats._fields_}

    def kernels_run(self) -> list:
        """Kernel families the last run of this plan launched
        (qf_plan_stats.kernels_run): "sweep", "lds_tile", "register_tile",
        "compiled", "mfma"; empty before the first run."""
        bits = self.stats()["kernels_run"]
        return [name for b, name in sorted(N.KERNEL_BITS.items()) if bits & b]

    def describe(self) -> dict:
        need = C.c_size_t()
        check(lib().qf_plan_describe(self._h, None, 0, C.byref(need)))
        buf = C.create_string_buffer(need.value)
        check(lib().qf_plan_describe(self._h, buf, need.value, C.byref(need)))
        return json.loads(buf.value.decode())

    @property
    def compiled(self) -> bool:
        """Whether the register passes run as compiled kernels."""
        return bool(self.describe().get("compiled", 0))

    def source(self, precision: str = "complex128", adjoint: bool = False) -> str:
        """Generated source of the compiled passes (qf_plan_source); adjoint:
        the two-state form the gradient engine launches for a reverse plan
        (qf_plan_adjoint_source)."""
        pr = 0 if precision == "complex128" else 1
        fn = lib().qf_plan_adjoint_source if adjoint else lib().qf_plan_source
        need = C.c_size_t()
        check(fn(self._h, pr, None, 0, C.byref(need)))
        buf = C.create_string_buffer(need.value)
        check(fn(self._h, pr, buf, need.value, C.byref(need)))
        return buf.value.decode()

    def compile(self, precision: str = "complex128") -> dict:
        """Compiles the passes for gfx950 now (host only); returns the compile
        time and whether the code came from the on-disk cache."""
        ms = C.c_double()
        cached = C.c_int32()
        check(lib().qf_plan_compile(self._h, 0 if precision == "complex128" else 1, C.byref(ms), C.byref(cached)))
        return {"compile_ms": ms.value, "from_cache": bool(cached.value)}

    def __del__(self) -> None:
        h = getattr(self, "_h", None)
        if h:
            getattr(self, '_lib', None) and self._lib.qf_plan_destroy(h)
            self._h = None


class HipBackend:
    """CudaBackend::new(ordinal), src/cuda.rs:556."""

    def __init__(self, ordinal: int = 0) -> None:
        h = C.c_void_p()
        check(lib().qf_backend_create(int(ordinal), C.byref(h)))
        self._h = h
        self._lib = lib()  # kept for release at interpreter exit
        self.ordinal = int(ordinal)

    def __del__(self) -> None:
        h = getattr(self, "_h", None)
        if h:
            getattr(self, '_lib', None) and self._lib.qf_backend_destroy(h)
            self._h = None

    def device_ordinal(self) -> int:
        return self.ordinal

    def synchronize(self) -> None:
        check(lib().qf_backend_synchronize(self._h))

    def release_cached(self) -> None:
        """Frees the device arena the uneven-batch engine keeps between calls
        (qf_backend_release_cached)."""
        check(lib().qf_backend_release_cached(self._h))

    def set_profiling(self, on: bool) -> None:
        """HIP-event timing of each phase of batched gradient calls."""
        check(lib().qf_backend_set_profiling(self._h, int(bool(on))))

    def last_gradient_profile(self) -> dict:
        """What the last batched gradient / expectation call ran: pass
        counts, kernel families, algorithmic bytes and (profiling on) device
        milliseconds per phase (qf_backend_last_gradient_profile)."""
        p = N.GradientProfile()
        check(lib().qf_backend_last_gradient_profile(self._h, C.byref(p)))
        d = {f: getattr(p, f) for f, _ in N.GradientProfile._fields_}
        for k in ("forward_kernels", "reverse_kernels"):
            d[k] = [name for b, name in sorted(N.KERNEL_BITS.items()) if d[k] & b]
        d["costate_kernel"] = {1: "windowed", 2: "per_term"}.get(d["costate_kernel"], "none")
        return d

    def set_plan_options(self, fuse: bool = True, tile_qubits: int = 0, low_qubits: int = 0, kernel: int = 0,
                         reg_bits: int = 0) -> None:
        o = N.PlanOptions(int(fuse), tile_qubits, low_qubits, 0, kernel, reg_bits)
        check(lib().qf_backend_set_plan_options(self._h, C.byref(o)))

    def stream(self) -> int:
        out = C.c_void_p()
        check(lib().qf_backend_stream(self._h, C.byref(out)))
        return out.value or 0

    # --- states -----------------------------------------------------------
    def zero_state(self, qubits: int, precision="complex128") -> "HipDeviceState":
        h = C.c_void_p()
        check(lib().qf_state_zero(self._h, qubits, _precision(precision), C.byref(h)))
        return HipDeviceState(h, self)

    def zero_state_with_memory_budget(self, qubits: int, budget_bytes: int,
                                      precision="complex128") -> "HipDeviceState":
        """StateVector::zero_with_memory_budget, src/lib.rs:940-966, on the
        device: MemoryBudgetExceeded when 2^qubits amplitudes do not fit."""
        h = C.c_void_p()
        check(lib().qf_state_zero_with_memory_budget(self._h, int(qubits), _precision(precision),
                                                     int(budget_bytes), C.byref(h)))
        return HipDeviceState(h, self)

    def upload(self, state: StateVector, precision="complex128") -> "HipDeviceState":
        """CudaBackend::upload, src/cuda.rs:844-867."""
        h = C.c_void_p()
        if _precision(precision) == C128:
            a = np.ascontiguousarray(state.amplitudes, dtype=np.complex128)
            check(lib().qf_state_upload(self._h, ptr(a), state.qubits, C.byref(h)))
        else:
            a = np.ascontiguousarray(state.amplitudes, dtype=np.complex64)
            check(lib().qf_state_upload_c64(self._h, ptr(a), state.qubits, C.byref(h)))
        return HipDeviceState(h, self)

    def upload_batch(self, states: Sequence[StateVector]) -> "HipDeviceStateBatch":
        """CudaBackend::upload_batch, src/cuda.rs:871-954 (EmptyBatch, IncompatibleBatch)."""
        if len(states) == 0:
            raise QuantumError(6, "a HIP state batch cannot be empty")
        qubits = np.array([s.qubits for s in states], dtype=np.uint32)
        if np.any(qubits != qubits[0]):
            raise QuantumError(7, "all states in a HIP batch must have equal dimensions")
        amps = np.ascontiguousarray(np.stack([s.amplitudes for s in states]), dtype=np.complex128)
        h = C.c_void_p()
        check(lib().qf_batch_upload_equal(self._h, ptr(amps), ptr(qubits), len(states), C.byref(h)))
        return HipDeviceStateBatch(h, self)

    def upload_uneven_batch(self, states: Sequence[StateVector]) -> "HipDeviceStateBatch":
        """A resident batch of states of different widths (qf_batch_upload):
        gates, circuits and energies apply to every state."""
        qubits = np.array([s.qubits for s in states], dtype=np.uint32)
        amps = np.ascontiguousarray(np.concatenate([np.asarray(s.amplitudes, dtype=np.complex128) for s in states])
                                    if len(states) else np.zeros(0, dtype=np.complex128))
        h = C.c_void_p()
        check(lib().qf_batch_upload(self._h, ptr(amps), ptr(qubits), len(states), C.byref(h)))
        return HipDeviceStateBatch(h, self)

    def zero_batch(self, qubits: int, count: int, precision="complex128") -> "HipDeviceStateBatch":
        h = C.c_void_p()
        check(lib().qf_batch_zero(self._h, qubits, count, _precision(precision), C.byref(h)))
        return HipDeviceStateBatch(h, self)

    def zero_uneven_batch(self, widths: Sequence[int], precision="complex128") -> "HipDeviceStateBatch":
        w = np.ascontiguousarray(np.asarray(widths, dtype=np.uint32))
        h = C.c_void_p()
        check(lib().qf_batch_zero_uneven(self._h, ptr(w), w.size, _precision(precision), C.byref(h)))
        return HipDeviceStateBatch(h, self)

    # --- convenience (src/cuda.rs:958-1083) --------------------------------
    def run_circuit(self, state: StateVector, circuit: Circuit, plan: Optional[Plan] = None) -> StateVector:
        dev = self.upload(state)
        dev.run_circuit(circuit, plan)
        return dev.download()

    def run_circuit_batch(self, states: Sequence[StateVector], circuit: Circuit) -> List[StateVector]:
        batch = self.upload_batch(states)
        batch.run_circuit(circuit)
        return batch.download()

    def _shots(self, initial: StateVector, circuit: Circuit, shots: int, seed: int, batched: bool) -> ShotResult:
        amps = np.ascontiguousarray(initial.amplitudes, dtype=np.complex128)
        ops = circuit.to_ops()
        width = C.c_uint32()
        check(lib().qf_run_circuit_shots(self._h, ptr(amps), initial.qubits, ptr(ops), len(ops), int(shots),
                                         int(seed), int(batched), None, C.byref(width)))
        regs = np.zeros((int(shots), width.value), dtype=np.uint8)
        buf = regs if regs.size else np.zeros(1, dtype=np.uint8)
        check(lib().qf_run_circuit_shots(self._h, ptr(amps), initial.qubits, ptr(ops), len(ops), int(shots),
                                         int(seed), int(batched), ptr(buf), C.byref(width)))
        return ShotResult.from_registers(regs)

    def run_circuit_shots(self, initial: StateVector, circuit: Circuit, shots: int, seed: int) -> ShotResult:
        """CudaBackend::run_circuit_shots, src/cuda.rs:1085-1159: shots one
        after another on a resident state, measurements and classical
        feed-forward on the device, SmallRng(seed) draws in the CPU's order."""
        return self._shots(initial, circuit, shots, seed, False)

    def run_circuit_shots_batched(self, initial: StateVector, circuit: Circuit, shots: int,
                                  seed: int) -> ShotResult:
        """CudaBackend::run_circuit_shots_batched, src/cuda.rs:1161-1269: all
        shots as one batch of states (split to fit device memory)."""
        return self._shots(initial, circuit, shots, seed, True)

    # --- density matrices (src/cuda.rs:969-1047) ---------------------------
    def upload_density(self, density: DensityMatrix) -> "HipDeviceDensityMatrix":
        """CudaBackend::upload_density, src/cuda.rs:969-1022."""
        e = np.ascontiguousarray(density.entries, dtype=np.complex128)
        h = C.c_void_p()
        check(lib().qf_density_upload(self._h, ptr(e), e.size, density.qubits, C.byref(h)))
        return HipDeviceDensityMatrix(h)

    def density_from_state(self, state: "HipDeviceState") -> "HipDeviceDensityMatrix":
        """|psi><psi| of a resident complex128 state, built on the device."""
        h = C.c_void_p()
        check(lib().qf_density_from_state(state._h, C.byref(h)))
        return HipDeviceDensityMatrix(h)

    def maximally_mixed_density(self, qubits: int) -> "HipDeviceDensityMatrix":
        h = C.c_void_p()
        check(lib().qf_density_maximally_mixed(self._h, int(qubits), C.byref(h)))
        return HipDeviceDensityMatrix(h)

    def maximally_mixed_density_with_memory_budget(self, qubits: int, budget_bytes: int) -> "HipDeviceDensityMatrix":
        """DensityMatrix::maximally_mixed_with_memory_budget, src/lib.rs:1668-1700."""
        h = C.c_void_p()
        check(lib().qf_density_maximally_mixed_with_memory_budget(self._h, int(qubits), int(budget_bytes),
                                                                  C.byref(h)))
        return HipDeviceDensityMatrix(h)

    def run_density_circuit(self, initial: DensityMatrix, circuit: Circuit) -> DensityMatrix:
        """CudaBackend::run_density_circuit, src/cuda.rs:1025-1047."""
        e = np.ascontiguousarray(initial.entries, dtype=np.complex128)
        out = np.empty_like(e)
        ops = circuit.to_ops()
        check(lib().qf_run_density_circuit(self._h, ptr(e), e.size, initial.qubits, ptr(ops), len(ops), ptr(out),
                                           out.size))
        return DensityMatrix(out, initial.qubits)

    # --- sparse Lanczos (src/quantum.rs:1853-2023, src/tensor_network.rs:1692-2025)
    def sparse_lanczos(self, op: SparseOperator, max_iterations: int, tolerance: float, details: bool = False):
        """sparse_lanczos_gpu: the lowest Ritz value of a real symmetric CSR
        operator; with details, (lowest, alpha, beta, model size)."""
        if op.rows != op.cols:
            raise QuantumError(15, "the operator must be square")
        rows = np.ascontiguousarray(op.row_offsets, dtype=np.uint64)
        cols = np.ascontiguousarray(op.column_indices, dtype=np.uint64)
        vals = np.ascontiguousarray(op.values, dtype=np.complex128)
        lo = C.c_double()
        a, b = np.zeros(100), np.zeros(100)
        size = C.c_uint64()
        check(lib().qf_sparse_lanczos(self._h, ptr(rows), ptr(cols), ptr(vals), op.rows, int(max_iterations),
                                      float(tolerance), C.byref(lo), ptr(a), ptr(b), C.byref(size)))
        return (lo.value, a, b, size.value) if details else lo.value

    def tfim_lanczos(self, sites: int, details: bool = False):
        """The DMRG TFIM Lanczos (H = -sum Z_i Z_{i+1} - sum X_i, open chain),
        matrix-free on the device; with details, (lowest, alpha, beta, m)."""
        lo = C.c_double()
        a, b = np.zeros(100), np.zeros(100)
        size = C.c_uint64()
        check(lib().qf_tfim_lanczos(self._h, int(sites), C.byref(lo), ptr(a), ptr(b), C.byref(size)))
        return (lo.value, a, b, size.value) if details else lo.value

    def apply_rotation_chain(self, state: StateVector, gates: Sequence[Tuple[RotationAxis, int, float]]) -> None:
        """Updates `state` in place (CudaBackend::apply_rotation_chain)."""
        dev = self.upload(state)
        dev.apply_rotation_chain(gates)
        state.amplitudes[:] = dev.download().amplitudes

    def apply_ry(self, state: StateVector, qubit: int, theta: float) -> None:
        self.apply_rotation_chain(state, [(RotationAxis.Y, qubit, theta)])


class HipDeviceState:
    """CudaDeviceState, src/cuda.rs:1277-1484: a resident device state."""

    def __init__(self, handle: C.c_void_p, backend: HipBackend) -> None:
        # The C handle holds a reference on its backend (freed with the last
        # state), so the Python backend object may be dropped first.
        self._h = handle
        self._lib = lib()  # kept for release at interpreter exit
        q = C.c_uint32()
        check(lib().qf_state_qubits(handle, C.byref(q)))
        self.qubits_ = int(q.value)
        p = C.c_int32()
        check(lib().qf_state_precision(handle, C.byref(p)))
        self.precision = int(p.value)

    def __del__(self) -> None:
        h = getattr(self, "_h", None)
        if h:
            getattr(self, '_lib', None) and self._lib.qf_state_destroy(h)
            self._h = None

    def qubits(self) -> int:
        return self.qubits_

    def device_pointer(self) -> int:
        out = C.c_void_p()
        check(lib().qf_state_device_pointer(self._h, C.byref(out)))
        return out.value or 0

    def apply_gate(self, gate: Gate, qubit: int) -> None:
        check(lib().qf_state_apply_gate(self._h, gate.id, int(qubit), float(gate.theta)))

    def apply_matrix(self, m: np.ndarray, qubit: int) -> None:
        mm = np.ascontiguousarray(np.asarray(m, dtype=np.complex128).reshape(4))
        check(lib().qf_state_apply_matrix(self._h, int(qubit), ptr(mm)))

    def apply_rotation(self, axis: RotationAxis, qubit: int, theta: float) -> None:
        self.apply_rotation_chain([(axis, qubit, theta)])

    def apply_rotation_chain(self, gates: Sequence[Tuple[RotationAxis, int, float]]) -> None:
        arr = np.zeros(len(gates), dtype=N.ROT_DTYPE)
        for i, (ax, q, t) in enumerate(gates):
            arr[i] = (int(ax), int(q), float(t))
        check(lib().qf_state_apply_rotation_chain(self._h, ptr(arr), len(arr)))

    def apply_cnot(self, control: int, target: int) -> None:
        check(lib().qf_state_apply_2q(self._h, 0, int(control), int(target)))

    def apply_cz(self, first: int, second: int) -> None:
        check(lib().qf_state_apply_2q(self._h, 1, int(first), int(second)))

    def apply_swap(self, first: int, second: int) -> None:
        check(lib().qf_state_apply_2q(self._h, 2, int(first), int(second)))

    def apply_cp(self, control: int, target: int, phi: float) -> None:
        check(lib().qf_state_apply_cp(self._h, int(control), int(target), float(phi)))

    def run_circuit(self, circuit: Circuit, plan: Optional[Plan] = None) -> None:
        ops = circuit.to_ops()
        check(lib().qf_state_run_circuit(self._h, ptr(ops), len(ops), plan._h if plan is not None else None))

    def copy_from_device(self, source: "HipDeviceState") -> None:
        check(lib().qf_state_copy_from(self._h, source._h))

    def download(self) -> StateVector:
        out = np.empty(1 << self.qubits_, dtype=np.complex128)
        check(lib().qf_state_download(self._h, ptr(out)))
        return StateVector(out, self.qubits_)

    def download_c64(self) -> np.ndarray:
        out = np.empty(1 << self.qubits_, dtype=np.complex64)
        check(lib().qf_state_download_c64(self._h, ptr(out)))
        return out

    def stream_rate(self, variant: int = 0, launches: int = 5) -> Tuple[float, float]:
        """(GB/s, ms per launch) of an in-place read-and-write-back sweep over
        this state's bytes (qf_state_stream_rate; the state is unchanged)."""
        g, ms = C.c_double(), C.c_double()
        check(lib().qf_state_stream_rate(self._h, int(variant), int(launches), C.byref(g), C.byref(ms)))
        return g.value, ms.value

    def gather(self, indices) -> np.ndarray:
        """Amplitudes at the given basis indices (complex128), read on the
        device without downloading the state (qf_state_gather)."""
        idx = np.ascontiguousarray(np.asarray(indices, dtype=np.uint64).reshape(-1))
        out = np.empty(idx.size, dtype=np.complex128)
        check(lib().qf_state_gather(self._h, ptr(idx), idx.size, ptr(out)))
        return out

    def measure(self, qubit: int, rng) -> int:
        """CudaDeviceState::measure, src/cuda.rs:1386-1442: draws
        rng.next_f64() and collapses the resident state."""
        o = C.c_uint8()
        check(lib().qf_state_measure(self._h, int(qubit), float(rng.next_f64()), C.byref(o)))
        return int(o.value)

    def norm_sqr(self) -> float:
        v = C.c_double()
        check(lib().qf_state_norm_sqr(self._h, C.byref(v)))
        return v.value

    def probabilities(self) -> np.ndarray:
        out = np.empty(1 << self.qubits_, dtype=np.float64)
        check(lib().qf_state_probabilities(self._h, ptr(out)))
        return out

    def expectation(self, observable: SingleQubitObservable, qubit: int) -> float:
        v = C.c_double()
        check(lib().qf_state_expectation(self._h, int(observable), int(qubit), C.byref(v)))
        return v.value

    def expectation_pauli_sum(self, h: PauliSum) -> complex:
        t = h.to_terms()
        re, im = C.c_double(), C.c_double()
        check(lib().qf_state_expectation_pauli_sum(self._h, ptr(t), len(t), C.byref(re), C.byref(im)))
        return complex(re.value, im.value)

    def inner(self, other: "HipDeviceState") -> complex:
        re, im = C.c_double(), C.c_double()
        check(lib().qf_state_inner(self._h, other._h, C.byref(re), C.byref(im)))
        return complex(re.value, im.value)


class HipDeviceStateBatch:
    """CudaDeviceStateBatch, src/cuda.rs:1486-2005: a resident batch (equal
    widths, or different widths from upload_uneven_batch / zero_uneven_batch)."""

    def __init__(self, handle: C.c_void_p, backend: HipBackend) -> None:
        self._h = handle
        self._lib = lib()  # kept for release at interpreter exit
        self._backend = backend
        q, n = C.c_uint32(), C.c_uint32()
        check(lib().qf_batch_shape(handle, C.byref(q), C.byref(n)))
        self.qubits_, self.count = int(q.value), int(n.value)
        w = np.zeros(max(1, self.count), dtype=np.uint32)
        check(lib().qf_batch_widths(handle, ptr(w)))
        self.widths = [int(x) for x in w[: self.count]]

    def __del__(self) -> None:
        h = getattr(self, "_h", None)
        if h:
            getattr(self, '_lib', None) and self._lib.qf_batch_destroy(h)
            self._h = None

    def qubits(self) -> int:
        return self.qubits_

    def batch_size(self) -> int:
        return self.count

    def apply_gate(self, gate: Gate, qubit: int) -> None:
        check(lib().qf_batch_apply_gate(self._h, gate.id, int(qubit), float(gate.theta)))

    def apply_cnot(self, control: int, target: int) -> None:
        check(lib().qf_batch_apply_2q(self._h, 0, int(control), int(target)))

    def apply_cz(self, first: int, second: int) -> None:
        check(lib().qf_batch_apply_2q(self._h, 1, int(first), int(second)))

    def apply_swap(self, first: int, second: int) -> None:
        check(lib().qf_batch_apply_2q(self._h, 2, int(first), int(second)))

    def run_circuit(self, circuit: Circuit, plan: Optional[Plan] = None) -> None:
        ops = circuit.to_ops()
        check(lib().qf_batch_run_circuit(self._h, ptr(ops), len(ops), plan._h if plan is not None else None))

    def measure(self, qubit: int, rng) -> List[int]:
        """CudaDeviceStateBatch::measure, src/cuda.rs:1500-1509: one
        rng.next_f64() per state, in state order."""
        return self.measure_with_uniforms(qubit, [rng.next_f64() for _ in range(self.count)])

    def measure_with_uniforms(self, qubit: int, uniforms: Sequence[float]) -> List[int]:
        """CudaDeviceStateBatch::measure_with_uniforms, src/cuda.rs:1511-1609."""
        u = np.ascontiguousarray(np.asarray(uniforms, dtype=np.float64))
        out = np.zeros(max(1, self.count), dtype=np.uint8)
        check(lib().qf_batch_measure_with_uniforms(self._h, int(qubit), ptr(u), u.size, ptr(out)))
        return [int(x) for x in out[: self.count]]

    def apply_gate_masked(self, gate: Gate, target: int, enabled: Sequence[int]) -> None:
        """The gate on the states whose enabled flag is set (classically
        conditioned gates of a shot batch)."""
        e = np.ascontiguousarray(np.asarray(enabled, dtype=np.uint8))
        check(lib().qf_batch_apply_gate_masked(self._h, gate.id, int(target), float(gate.theta), ptr(e), e.size))

    def apply_channel_trajectory_with_uniforms(self, channel: NoiseChannel, target: int,
                                               uniforms: Sequence[float]) -> None:
        """One Kraus-trajectory step per state (apply_channel_trajectory_with_uniforms,
        src/cuda.rs:1670-1790, with the CPU path's branch rules)."""
        u = np.ascontiguousarray(np.asarray(uniforms, dtype=np.float64))
        p = tuple(channel.params) + (0.0,) * (2 - len(channel.params))
        check(lib().qf_batch_apply_channel_trajectory(self._h, channel.kind, int(target), float(p[0]), float(p[1]),
                                                      ptr(u), u.size))

    def synchronize(self) -> None:
        self._backend.synchronize()

    def download(self) -> List[StateVector]:
        out = np.empty(sum(1 << q for q in self.widths), dtype=np.complex128)
        check(lib().qf_batch_download(self._h, ptr(out)))
        res, off = [], 0
        for q in self.widths:
            res.append(StateVector(out[off: off + (1 << q)].copy(), q))
            off += 1 << q
        return res

    def expectation_pauli_sum(self, h: PauliSum) -> np.ndarray:
        t = h.to_terms()
        out = np.empty(self.count, dtype=np.float64)
        check(lib().qf_batch_expectation_pauli_sum(self._h, ptr(t), len(t), ptr(out)))
        return out


class HipDeviceDensityMatrix:
    """CudaDeviceDensityMatrix, src/cuda.rs:2007-2237: a resident density
    matrix (row-major entries, complex128)."""

    def __init__(self, handle: C.c_void_p) -> None:
        self._h = handle
        self._lib = lib()  # kept for release at interpreter exit
        q = C.c_uint32()
        check(lib().qf_density_qubits(handle, C.byref(q)))
        self.qubits_ = int(q.value)

    def __del__(self) -> None:
        h = getattr(self, "_h", None)
        if h:
            getattr(self, '_lib', None) and self._lib.qf_density_destroy(h)
            self._h = None

    def qubits(self) -> int:
        return self.qubits_

    def apply_gate(self, gate: Gate, qubit: int) -> None:
        check(lib().qf_density_apply_gate(self._h, gate.id, int(qubit), float(gate.theta)))

    def apply_cnot(self, control: int, target: int) -> None:
        check(lib().qf_density_apply_cnot(self._h, int(control), int(target)))

    def apply_cz(self, first: int, second: int) -> None:
        check(lib().qf_density_apply_cz(self._h, int(first), int(second)))

    def apply_swap(self, first: int, second: int) -> None:
        check(lib().qf_density_apply_swap(self._h, int(first), int(second)))

    def apply_channel(self, channel: NoiseChannel, qubit: int) -> None:
        p = tuple(channel.params) + (0.0,) * (2 - len(channel.params))
        check(lib().qf_density_apply_channel(self._h, channel.kind, int(qubit), float(p[0]), float(p[1])))

    def run_circuit(self, circuit: Circuit, plan: Optional[Plan] = None) -> None:
        if plan is not None:
            check(lib().qf_density_run_plan(self._h, plan._h))
            return
        ops = circuit.to_ops()
        check(lib().qf_density_run_circuit(self._h, ptr(ops), len(ops)))

    def measure(self, qubit: int, rng) -> int:
        """CudaDeviceDensityMatrix::measure, src/cuda.rs:2161-2213: one rng.next_f64()."""
        return self.measure_with_uniform(qubit, rng.next_f64())

    def measure_with_uniform(self, qubit: int, uniform: float) -> int:
        o = C.c_uint8()
        check(lib().qf_density_measure(self._h, int(qubit), float(uniform), C.byref(o)))
        return int(o.value)

    def download(self) -> DensityMatrix:
        d = 1 << self.qubits_
        out = np.empty(d * d, dtype=np.complex128)
        check(lib().qf_density_download(self._h, ptr(out), out.size))
        return DensityMatrix(out, self.qubits_)

    def trace(self) -> complex:
        re, im = C.c_double(), C.c_double()
        check(lib().qf_density_trace(self._h, C.byref(re), C.byref(im)))
        return complex(re.value, im.value)

    def purity(self) -> float:
        v = C.c_double()
        check(lib().qf_density_purity(self._h, C.byref(v)))
        return v.value

    def expectation(self, observable: SingleQubitObservable, qubit: int) -> float:
        v = C.c_double()
        check(lib().qf_density_expectation(self._h, int(observable), int(qubit), C.byref(v)))
        return v.value

    def probability_one(self, qubit: int) -> float:
        v = C.c_double()
        check(lib().qf_density_probability_one(self._h, int(qubit), C.byref(v)))
        return v.value
