"""Python loader of the FP64 CPU oracle (liboracle.so, built from qf_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.  The
product package never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

OP_DTYPE = np.dtype([("kind", "<i4"), ("gate", "<i4"), ("a", "<u4"), ("b", "<u4"), ("theta", "<f8"),
                     ("theta2", "<f8")])
POP_DTYPE = np.dtype([("kind", "<i4"), ("gate", "<i4"), ("a", "<u4"), ("b", "<u4"), ("index", "<u4"),
                      ("reserved", "<u4"), ("theta", "<f8")])
PAULI_DTYPE = np.dtype([("re", "<f8"), ("im", "<f8"), ("x_mask", "<u8"), ("z_mask", "<u8")])


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P, u32, i32, dbl, sz, u64 = C.c_void_p, C.c_uint32, C.c_int32, C.c_double, C.c_size_t, C.c_uint64
        L.or_rng_seeded.restype = u64
        L.or_rng_seeded.argtypes = [u64]
        L.or_rng_next_u64.restype = u64
        L.or_rng_next_u64.argtypes = [C.POINTER(u64)]
        L.or_rng_next_f64.restype = dbl
        L.or_rng_next_f64.argtypes = [C.POINTER(u64)]
        L.or_gate_matrix.argtypes = [i32, dbl, P]
        L.or_gate_matrix.restype = None
        L.or_state_zero.restype = P
        L.or_state_zero.argtypes = [u32]
        L.or_state_from.restype = P
        L.or_state_from.argtypes = [P, u32]
        L.or_state_free.argtypes = [P]
        L.or_state_free.restype = None
        L.or_state_get.argtypes = [P, P]
        L.or_state_get.restype = None
        for n in ("or_apply_cnot", "or_apply_cz", "or_apply_swap"):
            getattr(L, n).argtypes = [P, u32, u32]
            getattr(L, n).restype = i32
        L.or_apply_gate.argtypes = [P, i32, dbl, u32]
        L.or_apply_gate.restype = i32
        L.or_apply_matrix.argtypes = [P, P, u32]
        L.or_apply_matrix.restype = i32
        L.or_run.argtypes = [P, P, sz]
        L.or_run.restype = i32
        L.or_time_run.argtypes = [u32, P, sz, i32]
        L.or_time_run.restype = dbl
        L.or_max_threads.restype = i32
        L.or_norm_sqr.argtypes = [P]
        L.or_norm_sqr.restype = dbl
        L.or_probabilities.argtypes = [P, P]
        L.or_probabilities.restype = None
        L.or_expectation.argtypes = [P, i32, u32, C.POINTER(dbl)]
        L.or_expectation.restype = i32
        L.or_pauli_expectation.argtypes = [P, P, sz, C.POINTER(dbl), C.POINTER(dbl)]
        L.or_pauli_expectation.restype = i32
        L.or_pc_run.argtypes = [P, sz, u32, u32, P, sz, P, sz, u32, C.c_int64, dbl, P]
        L.or_pc_run.restype = i32
        L.or_pc_expectation.argtypes = [P, sz, u32, u32, P, P, u32, P, sz, C.POINTER(dbl)]
        L.or_pc_expectation.restype = i32
        for n in ("or_pc_shift_gradient", "or_pc_adjoint_gradient"):
            getattr(L, n).argtypes = [P, sz, u32, u32, P, P, u32, P, sz, P]
            getattr(L, n).restype = i32
        L.or_pc_fd_gradient.argtypes = [P, sz, u32, u32, P, P, u32, P, sz, dbl, P]
        L.or_pc_fd_gradient.restype = i32
        L.or_fast_run.argtypes = [P, u32, P, sz]
        L.or_fast_run.restype = i32
        L.or_measure.argtypes = [P, u32, dbl, C.POINTER(C.c_uint8)]
        L.or_measure.restype = i32
        L.or_run_shots.argtypes = [P, u32, P, sz, u64, u64, P, C.POINTER(u32), i32]
        L.or_run_shots.restype = i32
        L.or_channel_trajectory.argtypes = [P, i32, dbl, dbl, u32, C.POINTER(u64)]
        L.or_channel_trajectory.restype = i32
        L.or_normalize.argtypes = [P]
        L.or_normalize.restype = i32
        L.or_density_from.restype = P
        L.or_density_from.argtypes = [P, u32]
        L.or_density_from_state.restype = P
        L.or_density_from_state.argtypes = [P, u32]
        L.or_density_free.argtypes = [P]
        L.or_density_free.restype = None
        L.or_density_get.argtypes = [P, P]
        L.or_density_get.restype = None
        L.or_density_apply_gate.argtypes = [P, i32, dbl, u32]
        L.or_density_apply_gate.restype = i32
        L.or_density_apply_2q.argtypes = [P, i32, u32, u32, dbl]
        L.or_density_apply_2q.restype = i32
        L.or_density_apply_channel.argtypes = [P, i32, dbl, dbl, u32]
        L.or_density_apply_channel.restype = i32
        L.or_density_measure.argtypes = [P, u32, dbl, C.POINTER(C.c_uint8)]
        L.or_density_measure.restype = i32
        L.or_density_probability_one.argtypes = [P, u32, C.POINTER(dbl)]
        L.or_density_probability_one.restype = i32
        L.or_density_run.argtypes = [P, P, sz]
        L.or_density_run.restype = i32
        L.or_lanczos_start.argtypes = [P, u64]
        L.or_lanczos_start.restype = None
        L.or_sparse_lanczos.argtypes = [P, P, P, u64, u64, dbl, P, P, C.POINTER(u64)]
        L.or_sparse_lanczos.restype = i32
        L.or_tfim_lanczos.argtypes = [u32, P, P, C.POINTER(u64)]
        L.or_tfim_lanczos.restype = i32
        _lib = L
    return _lib


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None and a.size else C.c_void_p(0)


def as_ops(ops) -> np.ndarray:
    """Op records in the current layout; arrays written with fewer fields
    (fixtures from before `theta2`) are copied over field by field."""
    ops = np.asarray(ops)
    if ops.dtype == OP_DTYPE:
        return np.ascontiguousarray(ops)
    out = np.zeros(ops.shape, dtype=OP_DTYPE)
    for name in ops.dtype.names:
        out[name] = ops[name]
    return out


class OracleError(Exception):
    def __init__(self, status):
        self.status = status
        super().__init__(f"oracle status {status}")


def _chk(s):
    if s != 0:
        raise OracleError(s)


class State:
    """Owned oracle state (clone-and-collect semantics per 1q gate)."""

    def __init__(self, qubits: int, amps: np.ndarray | None = None):
        self.qubits = qubits
        if amps is None:
            self.h = lib().or_state_zero(qubits)
        else:
            a = np.ascontiguousarray(amps, dtype=np.complex128)
            self.h = lib().or_state_from(_p(a), qubits)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_state_free(self.h)
            self.h = None

    def amplitudes(self) -> np.ndarray:
        out = np.empty(1 << self.qubits, dtype=np.complex128)
        lib().or_state_get(self.h, _p(out))
        return out

    def apply(self, gate_id: int, q: int, theta: float = 0.0):
        _chk(lib().or_apply_gate(self.h, gate_id, theta, q))

    def apply_matrix(self, m: np.ndarray, q: int):
        mm = np.ascontiguousarray(np.asarray(m, dtype=np.complex128).reshape(4))
        _chk(lib().or_apply_matrix(self.h, _p(mm), q))

    def cnot(self, c, t):
        _chk(lib().or_apply_cnot(self.h, c, t))

    def cz(self, a, b):
        _chk(lib().or_apply_cz(self.h, a, b))

    def swap(self, a, b):
        _chk(lib().or_apply_swap(self.h, a, b))

    def run(self, ops: np.ndarray):
        ops = as_ops(ops)
        _chk(lib().or_run(self.h, _p(ops), len(ops)))

    def measure(self, q: int, u: float) -> int:
        """StateVector::measure with the uniform draw passed in."""
        o = C.c_uint8()
        _chk(lib().or_measure(self.h, q, float(u), C.byref(o)))
        return int(o.value)

    def norm_sqr(self) -> float:
        return lib().or_norm_sqr(self.h)

    def normalize(self) -> None:
        _chk(lib().or_normalize(self.h))

    def probabilities(self) -> np.ndarray:
        out = np.empty(1 << self.qubits, dtype=np.float64)
        lib().or_probabilities(self.h, _p(out))
        return out

    def expectation(self, obs: int, q: int) -> float:
        v = C.c_double()
        _chk(lib().or_expectation(self.h, obs, q, C.byref(v)))
        return v.value

    def pauli_expectation(self, terms: np.ndarray) -> complex:
        t = np.ascontiguousarray(terms, dtype=PAULI_DTYPE)
        re, im = C.c_double(), C.c_double()
        _chk(lib().or_pauli_expectation(self.h, _p(t), len(t), C.byref(re), C.byref(im)))
        return complex(re.value, im.value)


def run_circuit(qubits: int, ops: np.ndarray) -> np.ndarray:
    s = State(qubits)
    s.run(ops)
    return s.amplitudes()


def gate_matrix(gate_id: int, theta: float = 0.0) -> np.ndarray:
    out = np.empty(4, dtype=np.complex128)
    lib().or_gate_matrix(gate_id, theta, _p(out))
    return out


def pc_run(pops, nparams, ninputs, params, inputs, qubits, shift_op=-1, delta=0.0):
    pops = np.ascontiguousarray(pops, dtype=POP_DTYPE)
    params = np.ascontiguousarray(params, dtype=np.float64)
    inputs = np.ascontiguousarray(inputs, dtype=np.float64)
    out = np.empty(1 << qubits, dtype=np.complex128)
    _chk(lib().or_pc_run(_p(pops), len(pops), nparams, ninputs, _p(params), len(params), _p(inputs), len(inputs),
                         qubits, shift_op, delta, _p(out)))
    return out


def _grad(fn, pops, nparams, ninputs, params, inputs, qubits, terms, *extra):
    pops = np.ascontiguousarray(pops, dtype=POP_DTYPE)
    params = np.ascontiguousarray(params, dtype=np.float64)
    inputs = np.ascontiguousarray(inputs, dtype=np.float64)
    terms = np.ascontiguousarray(terms, dtype=PAULI_DTYPE)
    if len(params) != nparams:
        raise OracleError(8)
    out = np.zeros(nparams, dtype=np.float64)
    _chk(fn(_p(pops), len(pops), nparams, ninputs, _p(params), _p(inputs), qubits, _p(terms), len(terms), *extra,
            _p(out)))
    return out


def pc_shift_gradient(pops, nparams, ninputs, params, inputs, qubits, terms):
    return _grad(lib().or_pc_shift_gradient, pops, nparams, ninputs, params, inputs, qubits, terms)


def pc_adjoint_gradient(pops, nparams, ninputs, params, inputs, qubits, terms):
    return _grad(lib().or_pc_adjoint_gradient, pops, nparams, ninputs, params, inputs, qubits, terms)


def pc_fd_gradient(pops, nparams, ninputs, params, inputs, qubits, terms, step):
    return _grad(lib().or_pc_fd_gradient, pops, nparams, ninputs, params, inputs, qubits, terms, C.c_double(step))


def pc_expectation(pops, nparams, ninputs, params, inputs, qubits, terms) -> float:
    pops = np.ascontiguousarray(pops, dtype=POP_DTYPE)
    params = np.ascontiguousarray(params, dtype=np.float64)
    inputs = np.ascontiguousarray(inputs, dtype=np.float64)
    terms = np.ascontiguousarray(terms, dtype=PAULI_DTYPE)
    v = C.c_double()
    _chk(lib().or_pc_expectation(_p(pops), len(pops), nparams, ninputs, _p(params), _p(inputs), qubits, _p(terms),
                                 len(terms), C.byref(v)))
    return v.value


def time_run(qubits: int, ops: np.ndarray, threads: int = 0) -> float:
    """Seconds for the restated reference CPU path to run `ops` from |0..0>."""
    ops = as_ops(ops)
    t = lib().or_time_run(qubits, _p(ops), len(ops), threads)
    if t < 0:
        raise OracleError(int(-t))
    return t


def max_threads() -> int:
    return int(lib().or_max_threads())


def fast_run(qubits: int, ops: np.ndarray, amps: np.ndarray | None = None) -> np.ndarray:
    """In-place OpenMP variant (same per-pair arithmetic), for larger checks."""
    if amps is None:
        amps = np.zeros(1 << qubits, dtype=np.complex128)
        amps[0] = 1
    ops = as_ops(ops)
    _chk(lib().or_fast_run(_p(amps), qubits, _p(ops), len(ops)))
    return amps


class Rng:
    def __init__(self, seed: int):
        self.s = C.c_uint64(lib().or_rng_seeded(seed))

    @property
    def state(self) -> int:
        return self.s.value

    def next_u64(self) -> int:
        return int(lib().or_rng_next_u64(C.byref(self.s)))

    def next_f64(self) -> float:
        return float(lib().or_rng_next_f64(C.byref(self.s)))


def run_shots(initial: np.ndarray, qubits: int, ops: np.ndarray, shots: int, seed: int,
              noise: bool = False) -> np.ndarray:
    """Circuit::run_shots (noise=False) or run_shots_with_noise (noise=True):
    the classical register of every shot, [shots][width] uint8,
    least-significant classical bit first."""
    initial = np.ascontiguousarray(initial, dtype=np.complex128)
    ops = as_ops(ops)
    w = C.c_uint32()
    nz = int(bool(noise))
    _chk(lib().or_run_shots(_p(initial), qubits, _p(ops), len(ops), shots, seed, None, C.byref(w), nz))
    out = np.zeros((shots, w.value), dtype=np.uint8)
    buf = out if w.value else np.zeros(1, np.uint8)
    _chk(lib().or_run_shots(_p(initial), qubits, _p(ops), len(ops), shots, seed, _p(buf), C.byref(w), nz))
    return out


def channel_trajectory(state: "State", kind: int, params, qubit: int, rng: "Rng") -> None:
    """StateVector::apply_channel_trajectory on an oracle state, drawing from `rng`."""
    p = tuple(params) + (0.0,) * (2 - len(params))
    _chk(lib().or_channel_trajectory(state.h, kind, p[0], p[1], qubit, C.byref(rng.s)))


class Density:
    """Owned oracle density matrix (row-major entries[row * dim + col])."""

    def __init__(self, qubits: int, entries: np.ndarray | None = None, state: np.ndarray | None = None):
        self.qubits = qubits
        if state is not None:
            a = np.ascontiguousarray(state, dtype=np.complex128)
            self.h = lib().or_density_from_state(_p(a), qubits)
        else:
            e = np.ascontiguousarray(entries, dtype=np.complex128).reshape(-1)
            self.h = lib().or_density_from(_p(e), qubits)
        if not self.h:
            raise OracleError(15)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_density_free(self.h)
            self.h = None

    def entries(self) -> np.ndarray:
        d = 1 << self.qubits
        out = np.zeros(d * d, dtype=np.complex128)
        lib().or_density_get(self.h, _p(out))
        return out

    def matrix(self) -> np.ndarray:
        d = 1 << self.qubits
        return self.entries().reshape(d, d)

    def apply(self, gate: int, qubit: int, theta: float = 0.0) -> None:
        _chk(lib().or_density_apply_gate(self.h, gate, theta, qubit))

    def apply_2q(self, kind: int, a: int, b: int, phi: float = 0.0) -> None:
        _chk(lib().or_density_apply_2q(self.h, kind, a, b, phi))

    def apply_channel(self, kind: int, params, qubit: int) -> None:
        p = tuple(params) + (0.0,) * (2 - len(params))
        _chk(lib().or_density_apply_channel(self.h, kind, p[0], p[1], qubit))

    def measure(self, qubit: int, u: float) -> int:
        o = C.c_uint8()
        _chk(lib().or_density_measure(self.h, qubit, u, C.byref(o)))
        return int(o.value)

    def probability_one(self, qubit: int) -> float:
        v = C.c_double()
        _chk(lib().or_density_probability_one(self.h, qubit, C.byref(v)))
        return v.value

    def run(self, ops: np.ndarray) -> None:
        ops = as_ops(ops)
        _chk(lib().or_density_run(self.h, _p(ops), len(ops)))




def tridiagonal_lowest(alpha: np.ndarray, beta: np.ndarray, size: int) -> float:
    """Lowest eigenvalue of the symmetric tridiagonal model (alpha on the
    diagonal, beta[i-1] beside it), with LAPACK's symmetric eigensolver in
    place of the reference's faer self_adjoint_evd."""
    T = np.diag(np.asarray(alpha[:size], dtype=np.float64))
    if size > 1:
        off = np.asarray(beta[: size - 1], dtype=np.float64)
        T += np.diag(off, 1) + np.diag(off, -1)
    return float(np.linalg.eigvalsh(T)[0]) if size else float("nan")


def lanczos_start(dim: int) -> np.ndarray:
    out = np.empty(dim, dtype=np.float64)
    lib().or_lanczos_start(_p(out), dim)
    return out


def sparse_lanczos(rows, cols, vals, dim: int, max_iterations: int, tolerance: float):
    """sparse_lanczos_gpu restated: (lowest, alpha, beta, size)."""
    rows = np.ascontiguousarray(rows, dtype=np.uint64)
    cols = np.ascontiguousarray(cols, dtype=np.uint64)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    m = max(1, min(max_iterations, dim, 100))
    a, b = np.zeros(m), np.zeros(m)
    size = C.c_uint64()
    _chk(lib().or_sparse_lanczos(_p(rows), _p(cols), _p(vals), dim, max_iterations, tolerance, _p(a), _p(b),
                                 C.byref(size)))
    return tridiagonal_lowest(a, b, size.value), a, b, size.value


def tfim_lanczos(n: int):
    """DMRG::run_sparse_lanczos on the TFIM chain restated: (lowest, alpha, beta, m)."""
    m = min(30 if n >= 24 else 100, 1 << n)
    a, b = np.zeros(m), np.zeros(m)
    size = C.c_uint64()
    _chk(lib().or_tfim_lanczos(n, _p(a), _p(b), C.byref(size)))
    return tridiagonal_lowest(a, b, size.value), a, b, size.value
