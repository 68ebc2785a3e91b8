"""Gradients of parameterized circuits on the HIP engine (src/gradients.rs).

`GradientStrategy`, `GradientOptions`, `GradientEstimate`, `CircuitProfile`
and `gradient` keep the reference's names and semantics: `Auto` resolves to
Adjoint (or ParameterShift when shots are requested, or the
OperationRequiresDensityMatrix error with a channel), `evaluations` is 2 for
the reverse-mode rules and 2 x rotation occurrences for shift / finite
differences.  The observable may be a single-qubit observable (as in the
reference) or a Pauli sum (the VQE energy).

`batch_gradient` runs many parameter sets of one circuit in one device batch.
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple, Union

import numpy as np

from . import _native as N
from ._native import QuantumError, check, lib, ptr
from .core import ParameterizedCircuit, PauliSum, SingleQubitObservable


class GradientStrategy(enum.IntEnum):
    """src/gradients.rs:46-124 (values are the QF_GRAD_* ids of include/qf_hip.h)."""

    ParameterShift = 0
    Adjoint = 1
    Backprop = 2
    FiniteDifference = 3
    Auto = 4

    def name_str(self) -> str:
        """GradientStrategy::name, src/gradients.rs:92-100."""
        return {0: "parameter-shift", 1: "adjoint", 2: "backprop", 3: "finite-difference", 4: "auto"}[int(self)]

    def is_exact(self) -> bool:
        """src/gradients.rs:105-107."""
        return self != GradientStrategy.FiniteDifference

    @staticmethod
    def parse(name: str) -> Optional["GradientStrategy"]:
        """GradientStrategy::parse, src/gradients.rs:114-123."""
        table = {"parameter-shift": 0, "shift": 0, "adjoint": 1, "backprop": 2, "backpropagation": 2,
                 "finite-difference": 3, "finite-diff": 3, "fd": 3, "auto": 4}
        v = table.get(name)
        return GradientStrategy(v) if v is not None else None

    @staticmethod
    def select(profile: "CircuitProfile") -> Optional["GradientStrategy"]:
        """GradientStrategy::select, src/gradients.rs:81-89."""
        if profile.has_channel:
            return None
        if profile.shots > 0:
            return GradientStrategy.ParameterShift
        return GradientStrategy.Adjoint


def gradient_strategies() -> List[str]:
    """Names accepted by `parse` (the Python binding's gradient_strategies())."""
    return ["auto", "parameter-shift", "adjoint", "backprop", "finite-difference"]


@dataclass
class CircuitProfile:
    """src/gradients.rs:128-140."""

    operations: int = 0
    parameters: int = 0
    parameterized_operations: int = 0
    has_channel: bool = False
    shots: int = 0

    @staticmethod
    def of(circuit: ParameterizedCircuit) -> "CircuitProfile":
        ops = circuit.operations()
        return CircuitProfile(len(ops), circuit.parameter_count, sum(1 for o in ops if o[0] == 1),
                              any(o[0] == 6 for o in ops), 0)


@dataclass
class GradientOptions:
    """src/gradients.rs:144-165 (defaults: Auto, step 1e-3, no shots)."""

    strategy: GradientStrategy = GradientStrategy.Auto
    finite_difference_step: float = 1e-3
    shots: int = 0
    precision: str = "complex128"
    tile_qubits: int = 0
    chunk: int = 0


@dataclass
class GradientEstimate:
    """src/gradients.rs:168-178."""

    gradient: np.ndarray
    strategy: GradientStrategy
    evaluations: int
    value: float = 0.0


Observable = Union[PauliSum, Tuple[SingleQubitObservable, int]]


def _terms(observable, qubits: int) -> np.ndarray:
    if isinstance(observable, PauliSum):
        return observable.to_terms()
    obs, q = observable
    if q >= qubits:
        raise QuantumError(2, f"qubit index {q} is out of range for {qubits} qubits")
    return PauliSum.single(SingleQubitObservable(obs), q, qubits).to_terms()


def _options(o: GradientOptions, strategy: Optional[int] = None) -> N.GradOptions:
    prec = 0 if o.precision in ("complex128", "c128") else 1
    return N.GradOptions(int(o.strategy if strategy is None else strategy), prec, float(o.finite_difference_step),
                         int(o.shots), int(o.tile_qubits), int(o.chunk))


def gradient(backend, circuit: ParameterizedCircuit, options: GradientOptions, parameters: Sequence[float],
             qubits: int, observable: Observable, inputs: Sequence[float] = ()) -> GradientEstimate:
    """ParameterizedCircuit::gradient_with_inputs, src/gradients.rs:232-312.

    `observable` is (SingleQubitObservable, qubit) or a PauliSum."""
    ops = circuit.to_pops()
    p = np.ascontiguousarray(parameters, dtype=np.float64)
    x = np.ascontiguousarray(inputs, dtype=np.float64)
    t = _terms(observable, qubits)
    g = np.zeros(max(circuit.parameter_count, 1), dtype=np.float64)
    val, used, evals = C.c_double(), C.c_int32(), C.c_uint64()
    o = _options(options)
    check(lib().qf_gradient(backend._h, ptr(ops), len(ops), circuit.parameter_count, circuit.input_count, ptr(x),
                            len(x), ptr(p), len(p), qubits, ptr(t), len(t), C.byref(o), ptr(g), C.byref(val),
                            C.byref(used), C.byref(evals)))
    return GradientEstimate(g[: circuit.parameter_count].copy(), GradientStrategy(used.value), int(evals.value),
                            val.value)


def parameter_shift_gradient(backend, circuit, parameters, qubits, observable, inputs=()) -> np.ndarray:
    """parameter_shift_gradient_with_inputs, src/lib.rs:3516-3546."""
    return gradient(backend, circuit, GradientOptions(GradientStrategy.ParameterShift), parameters, qubits,
                    observable, inputs).gradient


def adjoint_gradient(backend, circuit, parameters, qubits, observable, inputs=()) -> np.ndarray:
    """adjoint_gradient_with_inputs, src/lib.rs:3564-3704."""
    return gradient(backend, circuit, GradientOptions(GradientStrategy.Adjoint), parameters, qubits, observable,
                    inputs).gradient


def finite_difference_gradient_with_inputs(backend, circuit, inputs, parameters, qubits, observable,
                                           step: float) -> np.ndarray:
    """src/gradients.rs:344-375."""
    return gradient(backend, circuit, GradientOptions(GradientStrategy.FiniteDifference, step), parameters, qubits,
                    observable, inputs).gradient


def backprop_gradient_with_inputs(backend, circuit, inputs, parameters, qubits, observable) -> np.ndarray:
    """src/gradients.rs:392-506."""
    return gradient(backend, circuit, GradientOptions(GradientStrategy.Backprop), parameters, qubits, observable,
                    inputs).gradient


def batch_gradient(backend, circuit: ParameterizedCircuit, qubits: int, parameters: np.ndarray, observable,
                   strategy: GradientStrategy = GradientStrategy.Adjoint, precision: str = "complex128",
                   inputs: Sequence[float] = (), tile_qubits: int = 0, chunk: int = 0,
                   finite_difference_step: float = 1e-3) -> Tuple[np.ndarray, np.ndarray]:
    """Energies [sets] and gradients [sets, parameters] for many parameter sets in one device batch."""
    ops = circuit.to_pops()
    params = np.ascontiguousarray(np.atleast_2d(parameters), dtype=np.float64)
    sets = params.shape[0]
    if params.shape[1] != circuit.parameter_count:
        raise QuantumError(8, f"expected {circuit.parameter_count} circuit parameters, got {params.shape[1]}")
    x = np.ascontiguousarray(inputs, dtype=np.float64)
    t = _terms(observable, qubits)
    # Every element is written by the library (all strategies), so the
    # outputs skip the zero fill (8 MiB for 4096 sets x 256 parameters).
    energies = np.empty(sets, dtype=np.float64)
    grads = np.empty((sets, max(circuit.parameter_count, 1)), dtype=np.float64)
    o = _options(GradientOptions(strategy, finite_difference_step, 0, precision, tile_qubits, chunk))
    check(lib().qf_batch_gradient(backend._h, ptr(ops), len(ops), circuit.parameter_count, circuit.input_count,
                                  ptr(x), qubits, ptr(params), sets, ptr(t), len(t), C.byref(o), ptr(energies),
                                  ptr(grads)))
    return energies, grads[:, : circuit.parameter_count]


def batch_adjoint_gradient(backend, circuit, qubits, parameters, observable, precision="complex128", **kw):
    return batch_gradient(backend, circuit, qubits, parameters, observable, GradientStrategy.Adjoint, precision, **kw)


def batch_shift_gradient(backend, circuit, qubits, parameters, observable, precision="complex128", **kw):
    """The plain per-parameter parameter-shift GPU form (2 circuits per rotation occurrence)."""
    return batch_gradient(backend, circuit, qubits, parameters, observable, GradientStrategy.ParameterShift,
                          precision, **kw)


def batch_expectation(backend, circuit: ParameterizedCircuit, qubits: int, parameters: np.ndarray, observable,
                      precision: str = "complex128", inputs: Sequence[float] = (), tile_qubits: int = 0,
                      chunk: int = 0) -> np.ndarray:
    ops = circuit.to_pops()
    params = np.ascontiguousarray(np.atleast_2d(parameters), dtype=np.float64)
    x = np.ascontiguousarray(inputs, dtype=np.float64)
    t = _terms(observable, qubits)
    energies = np.empty(params.shape[0], dtype=np.float64)
    o = _options(GradientOptions(GradientStrategy.Adjoint, 1e-3, 0, precision, tile_qubits, chunk))
    check(lib().qf_batch_expectation(backend._h, ptr(ops), len(ops), circuit.parameter_count, circuit.input_count,
                                     ptr(x), qubits, ptr(params), params.shape[0], ptr(t), len(t), C.byref(o),
                                     ptr(energies)))
    return energies


def run_parameterized(backend, circuit: ParameterizedCircuit, parameters, qubits: int, inputs=(),
                      precision: str = "complex128"):
    """ParameterizedCircuit::run_with_inputs (src/lib.rs:3108-3121) into a new device state."""
    st = backend.zero_state(qubits, precision)
    ops = circuit.to_pops()
    p = np.ascontiguousarray(parameters, dtype=np.float64)
    x = np.ascontiguousarray(inputs, dtype=np.float64)
    check(lib().qf_state_run_parameterized(st._h, ptr(ops), len(ops), circuit.parameter_count, circuit.input_count,
                                           ptr(x), len(x), ptr(p), len(p)))
    return st


def uneven_adjoint_gradient(backend, instances: Sequence[Tuple[ParameterizedCircuit, int, Sequence[float], object]],
                            precision: str = "complex128"):
    """Energies and adjoint gradients of circuits of different widths in one
    device schedule (grouped launches replayed from one captured hipGraph per
    memory wave; tile shape from the autotuner).

    instances: (circuit, qubits, parameters, observable) tuples.
    Returns (energies [count], [gradient arrays], stats dict)."""
    keep = []
    recs = (N.UnevenInstance * len(instances))()
    for k, (pc, q, params, obs) in enumerate(instances):
        ops = pc.to_pops()
        p = np.ascontiguousarray(params, dtype=np.float64)
        t = _terms(obs, q)
        keep.append((ops, p, t))
        recs[k] = N.UnevenInstance(ops.ctypes.data if len(ops) else None, len(ops), pc.parameter_count, q,
                                   p.ctypes.data if p.size else None, t.ctypes.data if len(t) else None, len(t))
    energies = np.zeros(len(instances), dtype=np.float64)
    total = sum(pc.parameter_count for pc, _q, _p, _o in instances)
    grads = np.zeros(max(total, 1), dtype=np.float64)
    stats = N.UnevenStats()
    prec = 0 if precision in ("complex128", "c128") else 1
    check(lib().qf_uneven_adjoint_gradient(backend._h, recs, len(instances), prec, ptr(energies), ptr(grads),
                                           C.byref(stats)))
    out, off = [], 0
    for pc, _q, _p, _o in instances:
        out.append(grads[off: off + pc.parameter_count].copy())
        off += pc.parameter_count
    return energies, out, {f: getattr(stats, f) for f, _ in N.UnevenStats._fields_}


def uneven_expectation(backend, instances: Sequence[Tuple[ParameterizedCircuit, int, Sequence[float], object]],
                       precision: str = "complex128"):
    """Energies of circuits of different widths in one device schedule (the
    uneven engine without the reverse sweep).  instances: (circuit, qubits,
    parameters, observable) tuples.  Returns (energies [count], stats dict)."""
    keep = []
    recs = (N.UnevenInstance * len(instances))()
    for k, (pc, q, params, obs) in enumerate(instances):
        ops = pc.to_pops()
        p = np.ascontiguousarray(params, dtype=np.float64)
        t = _terms(obs, q)
        keep.append((ops, p, t))
        recs[k] = N.UnevenInstance(ops.ctypes.data if len(ops) else None, len(ops), pc.parameter_count, q,
                                   p.ctypes.data if p.size else None, t.ctypes.data if len(t) else None, len(t))
    energies = np.zeros(len(instances), dtype=np.float64)
    stats = N.UnevenStats()
    prec = 0 if precision in ("complex128", "c128") else 1
    check(lib().qf_uneven_expectation(backend._h, recs, len(instances), prec, ptr(energies), C.byref(stats)))
    return energies, {f: getattr(stats, f) for f, _ in N.UnevenStats._fields_}


def autotune(backend, circuit, qubits: int, precision: str = "complex128") -> Tuple[dict, float]:
    """Fastest register-kernel tile shape for a circuit (timed on the device)."""
    ops = circuit.to_ops()
    best = N.PlanOptions()
    ms = C.c_double()
    prec = 0 if precision in ("complex128", "c128") else 1
    check(lib().qf_autotune(backend._h, ptr(ops), len(ops), qubits, prec, C.byref(best), C.byref(ms)))
    return {f: int(getattr(best, f)) for f, _ in N.PlanOptions._fields_}, ms.value
