"""Circuit IR mirroring the reference crate's API (src/lib.rs).

`Gate`, `Circuit`, `ParameterizedCircuit`, `RotationAxis`,
`SingleQubitObservable`, `PauliSum` and `StateVector` keep the reference's names,
argument meaning and error variants.  These are plain host-side descriptions;
all amplitude arithmetic runs in the HIP engine (quantumforge_amd.backend).
"""
from __future__ import annotations

import enum
import math
from dataclasses import dataclass
from typing import Iterable, List, Optional, Sequence, Tuple

import numpy as np

from ._native import OP_DTYPE, PAULI_DTYPE, POP_DTYPE, QuantumError

# Gate ids of include/qf_hip.h, in the order of `enum Gate` (src/lib.rs:2239-2265).
_GATE_IDS = {"I": 0, "X": 1, "Y": 2, "Z": 3, "H": 4, "S": 5, "Sdg": 6, "T": 7, "Tdg": 8,
             "Rx": 9, "Ry": 10, "Rz": 11}
_GATE_NAMES = {v: k for k, v in _GATE_IDS.items()}


@dataclass(frozen=True)
class Gate:
    """One-qubit gate (`enum Gate`, src/lib.rs:2239-2265).  Rotations carry an angle."""

    name: str
    theta: float = 0.0

    @property
    def id(self) -> int:
        return _GATE_IDS[self.name]

    def is_rotation(self) -> bool:
        return self.name in ("Rx", "Ry", "Rz")

    def params(self) -> List[float]:
        """Gate::params, src/lib.rs:2335-2341."""
        return [self.theta] if self.is_rotation() else []

    def adjoint(self) -> "Gate":
        """adjoint_gate, src/lib.rs:2370-2382."""
        if self.is_rotation():
            return Gate(self.name, -self.theta)
        swap = {"S": "Sdg", "Sdg": "S", "T": "Tdg", "Tdg": "T"}
        return Gate(swap.get(self.name, self.name))

    def __str__(self) -> str:  # Display, src/lib.rs:2344-2361
        return f"{self.name}({self.theta})" if self.is_rotation() else self.name

    @staticmethod
    def Rx(theta: float) -> "Gate":
        return Gate("Rx", float(theta))

    @staticmethod
    def Ry(theta: float) -> "Gate":
        return Gate("Ry", float(theta))

    @staticmethod
    def Rz(theta: float) -> "Gate":
        return Gate("Rz", float(theta))


for _n in ("I", "X", "Y", "Z", "H", "S", "Sdg", "T", "Tdg"):
    setattr(Gate, _n, Gate(_n))


class RotationAxis(enum.IntEnum):
    """src/lib.rs:2921-2928."""
    X = 0
    Y = 1
    Z = 2


class SingleQubitObservable(enum.IntEnum):
    """src/lib.rs:2931-2938."""
    X = 0
    Y = 1
    Z = 2


def rotation_gate(axis: RotationAxis, theta: float) -> Gate:
    """rotation_gate, src/lib.rs:2363-2368."""
    return Gate(("Rx", "Ry", "Rz")[int(axis)], float(theta))


# ---------------------------------------------------------------------------
# Circuit (src/lib.rs:2386-2802)
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class NoiseChannel:
    """NoiseChannel, src/lib.rs:1561-1600: name is one of Depolarizing
    (probability), AmplitudeDamping (gamma), PhaseDamping (gamma),
    ThermalDamping (gamma_down, beta_delta)."""

    name: str
    params: Tuple[float, ...] = ()

    KINDS = ("Depolarizing", "AmplitudeDamping", "PhaseDamping", "ThermalDamping")

    @property
    def kind(self) -> int:
        return NoiseChannel.KINDS.index(self.name)


class Circuit:
    """Ordered list of operations (CircuitOp, src/lib.rs:2386-2440)."""

    def __init__(self) -> None:
        self._ops: List[tuple] = []
        self._channels: dict = {}  # op index -> NoiseChannel

    def push(self, gate: Gate, target: int) -> None:
        self._ops.append((0, gate.id, int(target), 0, float(gate.theta)))

    def cnot(self, control: int, target: int) -> None:
        self._ops.append((1, 0, int(control), int(target), 0.0))

    def cz(self, first: int, second: int) -> None:
        self._ops.append((2, 0, int(first), int(second), 0.0))

    def swap(self, first: int, second: int) -> None:
        self._ops.append((3, 0, int(first), int(second), 0.0))

    def cp(self, control: int, target: int, phi: float) -> None:
        """Exact controlled phase diag(1, 1, 1, e^{i phi}) (one diagonal op; the
        crate's Rz/CNOT lowering, src/algorithms.rs:139-145, equals it up to
        the global phase e^{-i phi/4})."""
        self._ops.append((7, 0, int(control), int(target), float(phi)))

    def channel(self, target: int, channel: Optional[NoiseChannel] = None) -> None:
        """A noise channel: kept in the list; state-vector and shot runs report
        OperationRequiresDensityMatrix."""
        ch = channel if isinstance(channel, NoiseChannel) else NoiseChannel("Depolarizing", (0.0,))
        self._channels[len(self._ops)] = ch
        self._ops.append((4, ch.kind, int(target), 0, float(ch.params[0]) if ch.params else 0.0))

    def measure(self, target: int, classical_bit: int) -> None:
        """Circuit::measure, src/lib.rs:2548-2553.  State-vector runs report
        OperationRequiresShotExecution; shot runs execute it."""
        self._ops.append((5, 0, int(target), int(classical_bit), 0.0))

    def if_classical_bit(self, classical_bit: int, value: int, gate: Gate, target: int) -> None:
        """Circuit::if_classical_bit, src/lib.rs:2567-2574 (value & 1): the
        gate runs when the classical bit equals value (shot runs only)."""
        self._ops.append((6, gate.id, int(target), int(classical_bit) | ((int(value) & 1) << 16), float(gate.theta)))

    def extend(self, other: "Circuit") -> None:
        for i, ch in other._channels.items():
            self._channels[len(self._ops) + i] = ch
        self._ops.extend(other._ops)

    def __len__(self) -> int:
        return len(self._ops)

    @property
    def operations(self) -> List[tuple]:
        return list(self._ops)

    def num_qubits(self) -> int:
        """Highest qubit touched + 1 (Circuit::num_qubits, src/lib.rs:2760-2784)."""
        top = -1
        for kind, _g, a, b, _t in self._ops:
            top = max(top, a, b if kind in (1, 2, 3, 7) else -1)
        return top + 1

    def to_ops(self) -> np.ndarray:
        """The op records of include/qf_hip.h (qf_op); a channel's second
        parameter (ThermalDamping beta_delta) goes to theta2."""
        arr = np.zeros(len(self._ops), dtype=OP_DTYPE)
        for i, (k, g, a, b, t) in enumerate(self._ops):
            ch = self._channels.get(i) if k == 4 else None
            t2 = float(ch.params[1]) if ch is not None and len(ch.params) > 1 else 0.0
            arr[i] = (k, g, a, b, t, t2)
        return arr

    @staticmethod
    def from_ops(ops: np.ndarray) -> "Circuit":
        c = Circuit()
        names = ops.dtype.names or ()
        for r in ops:
            k = int(r["kind"])
            if k == 4:
                kind = int(r["gate"])
                params = (float(r["theta"]),)
                if kind == 3:
                    params += (float(r["theta2"]) if "theta2" in names else 0.0,)
                c._channels[len(c._ops)] = NoiseChannel(NoiseChannel.KINDS[kind], params)
            c._ops.append((k, int(r["gate"]), int(r["a"]), int(r["b"]), float(r["theta"])))
        return c


class ShotResult:
    """ShotResult, src/lib.rs:2443-2452: shot count and the classical
    registers (tuples of bits, least-significant classical bit first) mapped
    to their counts."""

    def __init__(self, shots: int, counts: dict) -> None:
        self.shots = int(shots)
        self.counts = dict(counts)

    @staticmethod
    def from_registers(registers: np.ndarray) -> "ShotResult":
        counts: dict = {}
        for row in registers:
            key = tuple(int(x) for x in row)
            counts[key] = counts.get(key, 0) + 1
        return ShotResult(len(registers), counts)

    def __eq__(self, other) -> bool:
        return isinstance(other, ShotResult) and self.shots == other.shots and self.counts == other.counts

    def __repr__(self) -> str:
        return f"ShotResult(shots={self.shots}, counts={self.counts})"


# ---------------------------------------------------------------------------
# ParameterizedCircuit (src/lib.rs:2942-3121)
# ---------------------------------------------------------------------------
class ParameterizedCircuit:
    """Circuit with runtime-bound rotation parameters (src/lib.rs:2998-3104)."""

    def __init__(self, parameter_count: int, input_count: int = 0) -> None:
        self.parameter_count = int(parameter_count)
        self.input_count = int(input_count)
        self._ops: List[tuple] = []

    @staticmethod
    def with_inputs(parameter_count: int, input_count: int) -> "ParameterizedCircuit":
        return ParameterizedCircuit(parameter_count, input_count)

    def push(self, gate: Gate, target: int) -> None:
        self._ops.append((0, gate.id, int(target), 0, 0, float(gate.theta)))

    def rotation(self, axis: RotationAxis, target: int, parameter: int) -> None:
        self._ops.append((1, int(axis), int(target), 0, int(parameter), 0.0))

    def encode_rotation(self, axis: RotationAxis, target: int, input: int) -> None:
        self._ops.append((2, int(axis), int(target), 0, int(input), 0.0))

    def cnot(self, control: int, target: int) -> None:
        self._ops.append((3, 0, int(control), int(target), 0, 0.0))

    def cz(self, first: int, second: int) -> None:
        self._ops.append((4, 0, int(first), int(second), 0, 0.0))

    def swap(self, first: int, second: int) -> None:
        self._ops.append((5, 0, int(first), int(second), 0, 0.0))

    def channel(self, target: int, channel: object = None) -> None:
        self._ops.append((6, 0, int(target), 0, 0, 0.0))

    def operations(self) -> List[tuple]:
        return list(self._ops)

    def __len__(self) -> int:
        return len(self._ops)

    def rotation_count(self) -> int:
        """Parameterized occurrences (CircuitProfile::parameterized_operations)."""
        return sum(1 for o in self._ops if o[0] == 1)

    def has_channel(self) -> bool:
        return any(o[0] == 6 for o in self._ops)

    def to_pops(self) -> np.ndarray:
        """The op records of include/qf_hip.h (qf_pop)."""
        # One array construction from a list of records (assigning records
        # one by one costs about a microsecond each).
        return np.array([(k, g, a, b, idx, 0, t) for (k, g, a, b, idx, t) in self._ops], dtype=POP_DTYPE)

    def bind(self, parameters: Sequence[float], inputs: Sequence[float] = (),
             shift: Optional[Tuple[int, float]] = None) -> Circuit:
        """Fixed circuit for one parameter binding (run_with_shift, src/lib.rs:4677-4738)."""
        c = Circuit()
        for i, (k, g, a, b, idx, t) in enumerate(self._ops):
            if k == 0:
                c._ops.append((0, g, a, 0, t))
            elif k == 1:
                d = shift[1] if shift is not None and shift[0] == i else 0.0
                c.push(rotation_gate(RotationAxis(g), float(parameters[idx]) + d), a)
            elif k == 2:
                c.push(rotation_gate(RotationAxis(g), float(inputs[idx])), a)
            elif k == 3:
                c.cnot(a, b)
            elif k == 4:
                c.cz(a, b)
            elif k == 5:
                c.swap(a, b)
            else:
                c.channel(a)
        return c


# ---------------------------------------------------------------------------
# Pauli sums (src/analytic.rs:2549-2778)
# ---------------------------------------------------------------------------
class PauliSum:
    """Weighted sum of Pauli words.  A word is a string over I/X/Y/Z whose
    character k acts on qubit k (PauliWord(Vec<Pauli>), src/analytic.rs:2555)."""

    def __init__(self, qubits: int, terms: Iterable[Tuple[complex, str]] = ()) -> None:
        self.qubits = int(qubits)
        self.terms: List[Tuple[complex, str]] = []
        for c, w in terms:
            self.add(c, w)

    def add(self, coefficient: complex, word: str) -> None:
        if len(word) != self.qubits or any(ch not in "IXYZ" for ch in word):
            raise ValueError(f"Pauli word {word!r} must have {self.qubits} characters from IXYZ")
        self.terms.append((complex(coefficient), word))

    @staticmethod
    def single(observable: SingleQubitObservable, qubit: int, qubits: int) -> "PauliSum":
        w = ["I"] * qubits
        w[qubit] = "XYZ"[int(observable)]
        return PauliSum(qubits, [(1.0, "".join(w))])

    def to_terms(self) -> np.ndarray:
        """qf_pauli_term records: Y on q <=> bit q set in both masks."""
        recs = []
        for c, w in self.terms:
            x = z = 0
            for q, ch in enumerate(w):
                if ch in "XY":
                    x |= 1 << q
                if ch in "ZY":
                    z |= 1 << q
            recs.append((c.real, c.imag, x, z))
        return np.array(recs, dtype=PAULI_DTYPE)


# ---------------------------------------------------------------------------
# Host-side state container (StateVector, src/lib.rs:671-978)
# ---------------------------------------------------------------------------
class StateVector:
    """Host amplitudes, qubit 0 = least-significant index bit.  A plain
    container: gates and reductions run on a HIP backend (upload / download)."""

    def __init__(self, amplitudes: np.ndarray, qubits: int) -> None:
        self.amplitudes = amplitudes
        self.qubits = qubits

    @staticmethod
    def zero(qubits: int) -> "StateVector":
        """StateVector::zero, src/lib.rs:930-966."""
        if qubits == 0:
            raise QuantumError(1, "a state vector must contain at least one qubit")
        if qubits > 40:
            raise QuantumError(21, "dense simulation size exceeds the supported addressable range")
        a = np.zeros(1 << qubits, dtype=np.complex128)
        a[0] = 1.0
        return StateVector(a, qubits)

    @staticmethod
    def from_amplitudes(amplitudes: Sequence[complex]) -> "StateVector":
        """StateVector::from_amplitudes, src/lib.rs:969-978."""
        a = np.ascontiguousarray(np.asarray(amplitudes, dtype=np.complex128))
        n = a.shape[0]
        if n < 2:
            raise QuantumError(1, "a state vector must contain at least one qubit")
        if n & (n - 1):
            raise QuantumError(17, f"state vector length {n} is not a power of two")
        return StateVector(a, n.bit_length() - 1)

    def copy(self) -> "StateVector":
        return StateVector(self.amplitudes.copy(), self.qubits)


class DensityMatrix:
    """DensityMatrix, src/lib.rs:1553-2227: the row-major entries of a
    2^n x 2^n matrix.  A host value: gates, channels and measurement run on a
    HIP backend (HipBackend.upload_density / run_density_circuit); the
    queries below are host conveniences on downloaded matrices."""

    def __init__(self, entries: np.ndarray, qubits: int) -> None:
        d = 1 << qubits
        e = np.ascontiguousarray(np.asarray(entries, dtype=np.complex128).reshape(-1))
        if e.size != d * d:
            raise ValueError(f"a {qubits}-qubit density matrix has {d * d} entries, got {e.size}")
        self.entries = e
        self.qubits = qubits

    @staticmethod
    def from_entries(entries: Sequence[complex], qubits: int) -> "DensityMatrix":
        """DensityMatrix::from_entries, src/lib.rs:1707-1715."""
        return DensityMatrix(np.array(entries, dtype=np.complex128), qubits)

    @staticmethod
    def from_state(state: StateVector) -> "DensityMatrix":
        """DensityMatrix::from_state, src/lib.rs:1642-1656: |psi><psi|."""
        a = np.asarray(state.amplitudes, dtype=np.complex128)
        return DensityMatrix(np.outer(a, a.conj()), state.qubits)

    @staticmethod
    def maximally_mixed(qubits: int) -> "DensityMatrix":
        """DensityMatrix::maximally_mixed, src/lib.rs:1659-1704."""
        if qubits == 0:
            raise QuantumError(1, "a density matrix must contain at least one qubit")
        d = 1 << qubits
        return DensityMatrix(np.eye(d, dtype=np.complex128) / d, qubits)

    def dimension(self) -> int:
        return 1 << self.qubits

    def matrix(self) -> np.ndarray:
        d = self.dimension()
        return self.entries.reshape(d, d)

    def trace(self) -> complex:
        """DensityMatrix::trace, src/lib.rs:1732-1736."""
        return complex(np.trace(self.matrix()))

    def purity(self) -> float:
        """DensityMatrix::purity, src/lib.rs:1738-1742: sum of |rho_ij|^2."""
        return float(np.sum(np.abs(self.entries) ** 2))

    def _check(self, qubit: int) -> None:
        if qubit >= self.qubits:
            raise QuantumError(2, f"qubit index {qubit} is out of range for {self.qubits} qubits")

    def probability_one(self, qubit: int) -> float:
        """DensityMatrix::probability_one, src/lib.rs:1908-1921."""
        self._check(qubit)
        diag = np.real(np.diag(self.matrix()))
        ones = ((np.arange(self.dimension()) >> qubit) & 1) == 1
        return min(max(float(np.sum(diag[ones])), 0.0), 1.0)

    def expectation_z(self, qubit: int) -> float:
        """DensityMatrix::expectation_z, src/lib.rs:1858-1873."""
        self._check(qubit)
        diag = np.real(np.diag(self.matrix()))
        sign = 1.0 - 2.0 * ((np.arange(self.dimension()) >> qubit) & 1)
        return float(np.sum(sign * diag))

    def expectation(self, observable: "SingleQubitObservable", qubit: int) -> float:
        """DensityMatrix::expectation, src/lib.rs:1875-1905."""
        self._check(qubit)
        if observable == SingleQubitObservable.Z:
            return self.expectation_z(qubit)
        m = self.matrix()
        rows = np.arange(self.dimension())
        rows = rows[((rows >> qubit) & 1) == 0]
        coh = m[rows, rows | (1 << qubit)]
        return float(2.0 * np.sum(coh.real)) if observable == SingleQubitObservable.X else float(-2.0 * np.sum(coh.imag))

    def partial_trace(self, keep_qubits: Sequence[int]) -> "DensityMatrix":
        """DensityMatrix::partial_trace, src/lib.rs:1953-1981: traces out every
        qubit not listed; the kept qubits are ordered as given."""
        keep = [int(q) for q in keep_qubits]
        if any(q >= self.qubits for q in keep) or len(set(keep)) != len(keep):
            raise QuantumError(26, "invalid subsystem selection for the partial trace")
        traced = [q for q in range(self.qubits) if q not in keep]

        def expand(k: np.ndarray, e: np.ndarray) -> np.ndarray:
            full = np.zeros(np.broadcast(k, e).shape, dtype=np.int64)
            for pos, q in enumerate(keep):
                full |= ((k >> pos) & 1) << q
            for pos, q in enumerate(traced):
                full |= ((e >> pos) & 1) << q
            return full

        kept = np.arange(1 << len(keep))[:, None]
        env = np.arange(1 << len(traced))[None, :]
        full = expand(kept, env)  # [kept, env]
        m = self.matrix()
        out = m[full[:, None, :], full[None, :, :]].sum(axis=-1)
        return DensityMatrix(out, len(keep))

    def von_neumann_entropy(self) -> float:
        """DensityMatrix::von_neumann_entropy, src/lib.rs:1749-1796: -sum
        lambda ln lambda over the eigenvalues (Hermitian eigensolver),
        NonHermitianDensityMatrix beyond 1e-9."""
        m = self.matrix()
        if np.max(np.abs(m - m.conj().T), initial=0.0) > 1e-9:
            raise QuantumError(25, "density matrix is not Hermitian")
        ev = np.maximum(np.linalg.eigvalsh(m), 0.0)
        ev = ev[ev > 1e-15]
        return float(-np.sum(ev * np.log(ev)))


class SparseOperator:
    """SparseOperator, src/quantum.rs:784-880: compressed sparse rows built
    from (row, col, value) triplets; coordinates sorted, duplicates summed,
    explicit zeros dropped."""

    def __init__(self, rows: int, cols: int, entries: Iterable[Tuple[int, int, complex]]) -> None:
        maps: List[dict] = [dict() for _ in range(rows)]
        for r, c, v in entries:
            r, c = int(r), int(c)
            if r >= rows or c >= cols or r < 0 or c < 0:
                raise ValueError(f"sparse entry ({r}, {c}) is outside a {rows} x {cols} operator")
            maps[r][c] = maps[r].get(c, 0j) + complex(v)
        offsets, columns, values = [0], [], []
        for m in maps:
            for c in sorted(m):
                if m[c] != 0:
                    columns.append(c)
                    values.append(m[c])
            offsets.append(len(values))
        self.rows, self.cols = rows, cols
        self.row_offsets = np.array(offsets, dtype=np.uint64)
        self.column_indices = np.array(columns, dtype=np.uint64)
        self.values = np.array(values, dtype=np.complex128)

    @staticmethod
    def from_dense(matrix: np.ndarray, tolerance: float = 0.0) -> "SparseOperator":
        """SparseOperator::from_dense, src/quantum.rs:840-855: entries with
        magnitude above `tolerance`."""
        m = np.asarray(matrix, dtype=np.complex128)
        r, c = np.nonzero(np.abs(m) > tolerance)
        return SparseOperator(m.shape[0], m.shape[1], zip(r, c, m[r, c]))

    @staticmethod
    def tfim(sites: int) -> "SparseOperator":
        """The open TFIM chain H = -sum Z_i Z_{i+1} - sum X_i with site i at
        index bit sites-1-i (the DMRG Lanczos operator,
        src/tensor_network.rs:1702-1718), as CSR."""
        dim = 1 << sites
        a = np.arange(dim, dtype=np.int64)
        x = (a ^ (a >> 1)) & ((1 << (sites - 1)) - 1)
        differ = np.zeros(dim, dtype=np.int64)
        while np.any(x):
            differ += x & 1
            x >>= 1
        cols = np.empty((dim, sites + 1), dtype=np.int64)
        vals = np.full((dim, sites + 1), -1.0)
        cols[:, 0] = a
        vals[:, 0] = 2 * differ - (sites - 1)
        for i in range(sites):
            cols[:, 1 + i] = a ^ (1 << (sites - 1 - i))
        order = np.argsort(cols, axis=1)
        cols = np.take_along_axis(cols, order, axis=1)
        vals = np.take_along_axis(vals, order, axis=1)
        keep = vals != 0
        op = SparseOperator.__new__(SparseOperator)
        op.rows = op.cols = dim
        op.row_offsets = np.concatenate([[0], np.cumsum(keep.sum(axis=1))]).astype(np.uint64)
        op.column_indices = cols[keep].astype(np.uint64)
        op.values = vals[keep].astype(np.complex128)
        return op

    def nonzero_count(self) -> int:
        return int(self.values.size)

    def to_dense(self) -> np.ndarray:
        out = np.zeros((self.rows, self.cols), dtype=np.complex128)
        for r in range(self.rows):
            lo, hi = int(self.row_offsets[r]), int(self.row_offsets[r + 1])
            out[r, self.column_indices[lo:hi].astype(np.int64)] = self.values[lo:hi]
        return out

    def apply(self, vector: np.ndarray) -> np.ndarray:
        v = np.asarray(vector)
        out = np.zeros(self.rows, dtype=np.result_type(v.dtype, self.values.dtype))
        for r in range(self.rows):
            lo, hi = int(self.row_offsets[r]), int(self.row_offsets[r + 1])
            out[r] = np.dot(self.values[lo:hi], v[self.column_indices[lo:hi].astype(np.int64)])
        return out
