"""Brute-force NumPy oracle: dense 2^n x 2^n unitaries built from textbook
definitions (n <= 10).

TEST INFRASTRUCTURE ONLY.  Independent of qf_oracle.cpp: gate matrices come
from their mathematical definitions (rotations as matrix exponentials
exp(-i theta P / 2) via scipy), two-qubit gates as explicit permutation /
diagonal matrices, the state as U |0..0>.  Qubit 0 is the least-significant
index bit.  Used to pin the C++ restatement.
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import expm

I2 = np.eye(2, dtype=complex)
PX = np.array([[0, 1], [1, 0]], dtype=complex)
PY = np.array([[0, -1j], [1j, 0]], dtype=complex)
PZ = np.array([[1, 0], [0, -1]], dtype=complex)
PAULI = {"I": I2, "X": PX, "Y": PY, "Z": PZ}


def gate(gid: int, theta: float = 0.0) -> np.ndarray:
    """Gate id of include/qf_hip.h -> 2x2 matrix from the definitions."""
    s = 1 / np.sqrt(2)
    fixed = {
        0: I2, 1: PX, 2: PY, 3: PZ,
        4: s * np.array([[1, 1], [1, -1]], dtype=complex),
        5: np.diag([1, 1j]), 6: np.diag([1, -1j]),
        7: np.diag([1, np.exp(1j * np.pi / 4)]), 8: np.diag([1, np.exp(-1j * np.pi / 4)]),
    }
    if gid in fixed:
        return np.asarray(fixed[gid], dtype=complex)
    P = {9: PX, 10: PY, 11: PZ}[gid]
    return expm(-0.5j * theta * P)


def embed_1q(m: np.ndarray, q: int, n: int) -> np.ndarray:
    return np.kron(np.kron(np.eye(1 << (n - 1 - q)), m), np.eye(1 << q))


def perm_matrix(n: int, f) -> np.ndarray:
    d = 1 << n
    P = np.zeros((d, d), dtype=complex)
    for i in range(d):
        P[f(i), i] = 1.0
    return P


def cnot(c: int, t: int, n: int) -> np.ndarray:
    return perm_matrix(n, lambda i: i ^ (((i >> c) & 1) << t))


def swap(a: int, b: int, n: int) -> np.ndarray:
    def f(i):
        ba, bb = (i >> a) & 1, (i >> b) & 1
        return i if ba == bb else i ^ (1 << a) ^ (1 << b)
    return perm_matrix(n, f)


def cz(a: int, b: int, n: int) -> np.ndarray:
    return np.diag([-1.0 if ((i >> a) & 1) and ((i >> b) & 1) else 1.0 for i in range(1 << n)]).astype(complex)


def cp(c: int, t: int, phi: float, n: int) -> np.ndarray:
    return np.diag([np.exp(1j * phi) if ((i >> c) & 1) and ((i >> t) & 1) else 1.0 for i in range(1 << n)])


def op_unitary(op, n: int) -> np.ndarray:
    kind, g, a, b, th = int(op["kind"]), int(op["gate"]), int(op["a"]), int(op["b"]), float(op["theta"])
    if kind == 0:
        return embed_1q(gate(g, th), a, n)
    if kind == 1:
        return cnot(a, b, n)
    if kind == 2:
        return cz(a, b, n)
    if kind == 3:
        return swap(a, b, n)
    if kind == 7:
        return cp(a, b, th, n)
    raise ValueError("non-unitary op")


def run(ops, n: int, psi0: np.ndarray | None = None) -> np.ndarray:
    psi = np.zeros(1 << n, dtype=complex)
    psi[0] = 1.0
    if psi0 is not None:
        psi = psi0.astype(complex).copy()
    for op in ops:
        psi = op_unitary(op, n) @ psi
    return psi


def measure(psi: np.ndarray, q: int, u: float):
    """Projective measurement of qubit q with uniform u (StateVector::measure,
    src/lib.rs:1321-1350, as a projector): returns (outcome, collapsed state)."""
    n_idx = np.arange(len(psi))
    bit = (n_idx >> q) & 1
    p_one = min(max(float(np.sum(np.abs(psi[bit == 1]) ** 2)), 0.0), 1.0)
    o = int(u < p_one)
    p = p_one if o else 1.0 - p_one
    if p <= np.finfo(float).eps:
        raise ValueError("invalid probability")
    out = np.where(bit == o, psi, 0) / np.sqrt(p)
    return o, out


def channel_params(op) -> tuple:
    p = (float(op["theta"]),)
    if int(op["gate"]) == 3:
        p += (float(op["theta2"]) if "theta2" in (op.dtype.names or ()) else 0.0,)
    return p


def kraus(kind: int, params) -> list:
    """Kraus operators of a noise channel (NoiseChannel, src/lib.rs:1561-1600):
    0 depolarizing (1-p) rho + p/3 (X rho X + Y rho Y + Z rho Z), 1 amplitude
    damping |1> -> |0> with probability gamma, 2 phase damping, 3 thermal
    damping with gamma_up = gamma_down * exp(-beta_delta)."""
    p = float(params[0])
    if kind == 0:
        return [np.sqrt(1 - p) * I2, np.sqrt(p / 3) * PX, np.sqrt(p / 3) * PY, np.sqrt(p / 3) * PZ]
    if kind == 1:
        return [np.diag([1.0, np.sqrt(1 - p)]).astype(complex), np.array([[0, np.sqrt(p)], [0, 0]], dtype=complex)]
    if kind == 2:
        return [np.diag([1.0, np.sqrt(1 - p)]).astype(complex), np.diag([0.0, np.sqrt(p)]).astype(complex)]
    gu = p * np.exp(-float(params[1]))
    return [np.diag([np.sqrt(1 - gu), np.sqrt(1 - p)]).astype(complex),
            np.array([[0, np.sqrt(p)], [0, 0]], dtype=complex), np.array([[0, 0], [np.sqrt(gu), 0]], dtype=complex)]


def apply_channel(rho: np.ndarray, kind: int, params, q: int, n: int) -> np.ndarray:
    out = np.zeros_like(rho)
    for K in kraus(kind, params):
        E = embed_1q(K, q, n)
        out = out + E @ rho @ E.conj().T
    return out


def density_run(ops, n: int, rho: np.ndarray) -> np.ndarray:
    """Circuit::run_density (src/lib.rs:2804-2819) with dense matrices:
    U rho U^dagger for gates, sum_k E_k rho E_k^dagger for channels."""
    rho = np.asarray(rho, dtype=complex).reshape(1 << n, 1 << n).copy()
    for op in ops:
        k = int(op["kind"])
        if k == 4:
            rho = apply_channel(rho, int(op["gate"]), channel_params(op), int(op["a"]), n)
        elif k in (5, 6):
            raise ValueError("measurement needs shot execution")
        else:
            U = op_unitary(op, n)
            rho = U @ rho @ U.conj().T
    return rho


def density_measure(rho: np.ndarray, q: int, u: float):
    """Projective measurement of a density matrix (DensityMatrix::measure,
    src/lib.rs:1924-1950, as P rho P / p): returns (outcome, rho')."""
    d = rho.shape[0]
    bit = (np.arange(d) >> q) & 1
    p_one = min(max(float(np.sum(np.real(np.diag(rho))[bit == 1])), 0.0), 1.0)
    o = int(u < p_one)
    p = p_one if o else 1.0 - p_one
    if p <= np.finfo(float).eps:
        raise ValueError("invalid probability")
    P = np.diag((bit == o).astype(float))
    return o, P @ rho @ P / p


def trajectory(psi: np.ndarray, kind: int, params, q: int, rng) -> np.ndarray:
    """One Kraus-trajectory step (StateVector::apply_channel_trajectory,
    src/lib.rs:1356-1505) with vector operations: pick branch k with
    probability ||K_k psi||^2 / ||psi||^2 by the rule of that function (one
    uniform), return K_k psi normalized."""
    bit = ((np.arange(len(psi)) >> q) & 1).astype(bool)
    p = float(params[0])
    if kind == 0:
        u = rng.next_f64()
        if u >= 1.0 - p and p > 0.0:
            b = min(max(int(np.floor((u - (1.0 - p)) * (3.0 / p))), 0), 2)
            return embed_1q([PX, PY, PZ][b], q, int(np.log2(len(psi)))) @ psi
        return psi
    norm = float(np.sum(np.abs(psi) ** 2))
    p1 = float(np.sum(np.abs(psi[bit]) ** 2))
    if kind in (1, 2):
        jp = min(max(p * p1 / norm, 0.0), 1.0)
        jump = rng.next_f64() < jp
        bn = p * p1 if jump else norm - p * p1
        if bn <= np.finfo(float).eps:
            raise ValueError("invalid probability")
        inv = 1.0 / np.sqrt(bn)
        out = np.zeros_like(psi)
        if jump and kind == 1:
            out[~bit] = psi[bit] * (np.sqrt(p) * inv)
        elif jump:
            out[bit] = psi[bit] * (np.sqrt(p) * inv)
        else:
            out[~bit] = psi[~bit] * inv
            out[bit] = psi[bit] * (np.sqrt(1.0 - p) * inv)
        return out
    gd, gu = p, p * np.exp(-float(params[1]))
    pr1 = min(max(p1 / norm, 0.0), 1.0)
    pr0 = 1.0 - pr1
    down, up, stay = pr1 * gd, pr0 * gu, pr0 * (1 - gu) + pr1 * (1 - gd)
    u = rng.next_f64()
    out = np.zeros_like(psi)
    if u < down:
        out[~bit] = psi[bit] * (np.sqrt(gd) / np.sqrt(down))
    elif u < down + up:
        out[bit] = psi[~bit] * (np.sqrt(gu) / np.sqrt(up))
    else:
        out[~bit] = psi[~bit] * (np.sqrt(1 - gu) / np.sqrt(stay))
        out[bit] = psi[bit] * (np.sqrt(1 - gd) / np.sqrt(stay))
    return out


def run_shots(ops, n: int, psi0: np.ndarray, shots: int, seed: int, noise: bool = False) -> np.ndarray:
    """Circuit::run_shots / run_shots_with_noise, src/lib.rs:2822-2911, with
    dense unitaries: one SmallRng uniform per measurement and per channel,
    shot-major; with noise the initial state is normalized per shot."""
    from quantumforge_amd.workloads import SmallRng

    width = 0
    for op in ops:
        if int(op["kind"]) == 5:
            width = max(width, int(op["b"]) + 1)
        elif int(op["kind"]) == 6:
            width = max(width, (int(op["b"]) & 0xFFFF) + 1)
    rng = SmallRng(seed)
    regs = np.zeros((shots, width), dtype=np.uint8)
    for shot in range(shots):
        psi = psi0.astype(complex).copy()
        if noise:
            psi = psi / np.sqrt(float(np.sum(np.abs(psi) ** 2)))
        for op in ops:
            k = int(op["kind"])
            if k == 5:
                regs[shot, int(op["b"])], psi = measure(psi, int(op["a"]), rng.next_f64())
            elif k == 6:
                b = int(op["b"])
                if regs[shot, b & 0xFFFF] == ((b >> 16) & 1):
                    psi = embed_1q(gate(int(op["gate"]), float(op["theta"])), int(op["a"]), n) @ psi
            elif k == 4:
                if not noise:
                    raise ValueError("channels need run_shots_with_noise")
                psi = trajectory(psi, int(op["gate"]), channel_params(op), int(op["a"]), rng)
            else:
                psi = op_unitary(op, n) @ psi
    return regs


def pauli_word_matrix(word: str) -> np.ndarray:
    """Character k acts on qubit k (qubit 0 = least significant)."""
    m = np.eye(1)
    for ch in reversed(word):
        m = np.kron(m, PAULI[ch])
    return m


def pauli_sum_matrix(terms) -> np.ndarray:
    """terms: iterable of (coefficient, word)."""
    H = None
    for c, w in terms:
        M = c * pauli_word_matrix(w)
        H = M if H is None else H + M
    return H


def expectation(psi: np.ndarray, H: np.ndarray) -> complex:
    return complex(np.vdot(psi, H @ psi))


def pc_ops(pops, params, inputs=(), shift=None):
    """Binds a parameterized op list (qf_pop records) to fixed op dicts."""
    axis_gate = {0: 9, 1: 10, 2: 11}
    out = []
    for i, p in enumerate(pops):
        k = int(p["kind"])
        if k == 0:
            out.append(dict(kind=0, gate=int(p["gate"]), a=int(p["a"]), b=0, theta=float(p["theta"])))
        elif k in (1, 2):
            th = float(params[int(p["index"])] if k == 1 else inputs[int(p["index"])])
            if shift is not None and shift[0] == i:
                th += shift[1]
            out.append(dict(kind=0, gate=axis_gate[int(p["gate"])], a=int(p["a"]), b=0, theta=th))
        else:
            out.append(dict(kind=k - 2, gate=0, a=int(p["a"]), b=int(p["b"]), theta=0.0))
    return out


def pc_energy(pops, params, n, H, inputs=(), shift=None) -> float:
    return expectation(run(pc_ops(pops, params, inputs, shift), n), H).real


def pc_shift_gradient(pops, nparams, params, n, H, inputs=()):
    """Exact derivative of <H> by the two-term shift rule per occurrence."""
    g = np.zeros(nparams)
    for i, p in enumerate(pops):
        if int(p["kind"]) == 1:
            plus = pc_energy(pops, params, n, H, inputs, (i, np.pi / 2))
            minus = pc_energy(pops, params, n, H, inputs, (i, -np.pi / 2))
            g[int(p["index"])] += 0.5 * (plus - minus)
    return g
