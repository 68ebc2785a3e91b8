"""Seeded workload builders: the reference's generators and the synthetic
benchmark circuits, all driven by the reference's SmallRng.

Reference builders mirrored here: SmallRng (src/lib.rs:4881-4912),
HEABuilder (src/algorithms_templates.rs:189-282), qaoa_maxcut and
controlled_phase (src/algorithms.rs:106-145), qft_circuit
(src/algorithms.rs:270-286), MaxCutProblem (src/mps_qaoa.rs:42-211).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Sequence, Tuple

from .core import Circuit, Gate, NoiseChannel, ParameterizedCircuit, PauliSum, RotationAxis

MASK64 = (1 << 64) - 1


class SmallRng:
    """splitmix64-seeded xorshift64 (13, 7, 17), src/lib.rs:4891-4912."""

    def __init__(self, seed: int) -> None:
        z = (seed + 0x9E3779B97F4A7C15) & MASK64
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        z ^= z >> 31
        self.state = z if z != 0 else 0x9E3779B97F4A7C15

    @staticmethod
    def seeded(seed: int) -> "SmallRng":
        return SmallRng(seed)

    def next_u64(self) -> int:
        x = self.state
        x ^= (x << 13) & MASK64
        x ^= x >> 7
        x ^= (x << 17) & MASK64
        self.state = x
        return x

    def next_f64(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / float(1 << 53))


def fisher_yates(rng: SmallRng, items: list) -> list:
    """In-place shuffle as MaxCutProblem::random_3_regular does it
    (src/mps_qaoa.rs:69-72): for i from len-1 down to 1, j = next_u64 % (i+1)."""
    for i in range(len(items) - 1, 0, -1):
        j = rng.next_u64() % (i + 1)
        items[i], items[j] = items[j], items[i]
    return items


# ---------------------------------------------------------------------------
# Random layered circuits (20q / 30q depth-40 workloads)
# ---------------------------------------------------------------------------
_LAYER_GATES = ("H", "X", "Rx", "Ry", "Rz")


def random_layer_circuit(qubits: int, depth: int, seed: int, pairs_per_layer: int = 10) -> Circuit:
    """Depth layers; each applies one 1q gate per qubit drawn uniformly from
    {H, X, RX, RY, RZ} (index floor(5u); rotation angle 2*pi*u from a second
    draw), then `pairs_per_layer` two-qubit gates on disjoint pairs: a
    Fisher-Yates shuffle of 0..n-1 taken as consecutive pairs, each CNOT
    (first element is the control) if u < 0.5, else CZ."""
    if 2 * pairs_per_layer > qubits:
        raise ValueError("not enough qubits for disjoint pairs")
    rng = SmallRng(seed)
    c = Circuit()
    two_pi = 2.0 * math.pi
    for _ in range(depth):
        for q in range(qubits):
            g = _LAYER_GATES[min(int(5.0 * rng.next_f64()), 4)]
            if g in ("Rx", "Ry", "Rz"):
                c.push(Gate(g, two_pi * rng.next_f64()), q)
            else:
                c.push(Gate(g), q)
        perm = fisher_yates(rng, list(range(qubits)))
        for k in range(pairs_per_layer):
            a, b = perm[2 * k], perm[2 * k + 1]
            if rng.next_f64() < 0.5:
                c.cnot(a, b)
            else:
                c.cz(a, b)
    return c


def noisy_layer_circuit(qubits: int, layers: int, seed: int, strength: float = 0.02) -> Circuit:
    """Noisy layers for density matrices and trajectories: per layer an
    RX/RY/RZ rotation on every qubit (axis floor(3u), angle 2*pi*u), CNOT on
    (q, q+1) for even q, CZ(n-1, 0), then one channel on every qubit, cycling
    depolarizing (p), amplitude damping (gamma), phase damping (gamma) and
    thermal damping (gamma_down, beta_delta = 1) with strength p = gamma =
    gamma_down = `strength`."""
    rng = SmallRng(seed)
    c = Circuit()
    two_pi = 2.0 * math.pi
    chans = [NoiseChannel("Depolarizing", (strength,)), NoiseChannel("AmplitudeDamping", (strength,)),
             NoiseChannel("PhaseDamping", (strength,)), NoiseChannel("ThermalDamping", (strength, 1.0))]
    for layer in range(layers):
        for q in range(qubits):
            axis = ("Rx", "Ry", "Rz")[min(int(3.0 * rng.next_f64()), 2)]
            c.push(Gate(axis, two_pi * rng.next_f64()), q)
        for q in range(0, qubits - 1, 2):
            c.cnot(q, q + 1)
        if qubits > 2:
            c.cz(qubits - 1, 0)
        for q in range(qubits):
            c.channel(q, chans[(q + layer) % 4])
    return c


# ---------------------------------------------------------------------------
# QFT-like layers (33/34q workload)
# ---------------------------------------------------------------------------
def qft_like_circuit(qubits: int, layers: int, seed: int, max_distance: int = 8) -> Tuple[Circuit, Circuit]:
    """Per layer: H on every qubit; controlled phase CP(pi/2^d) on every pair
    at distance 1 <= d <= max_distance (control = higher qubit, as
    qft_circuit, src/algorithms.rs:278-284); RZ(2*pi*u) on every qubit.
    Returns (lowered, native): CP lowered the reference's way
    (controlled_phase, src/algorithms.rs:139-145) and the same circuit with
    exact CP ops; they agree up to the global phase e^{-i sum(phi)/4}."""
    rng = SmallRng(seed)
    c = Circuit()
    native = Circuit()
    for _ in range(layers):
        for q in range(qubits):
            c.push(Gate.H, q)
            native.push(Gate.H, q)
        for d in range(1, max_distance + 1):
            for t in range(qubits - d):
                ctl = t + d
                phi = math.pi / (1 << d)
                controlled_phase(c, ctl, t, phi)
                native.cp(ctl, t, phi)
        for q in range(qubits):
            th = 2.0 * math.pi * rng.next_f64()
            c.push(Gate.Rz(th), q)
            native.push(Gate.Rz(th), q)
    return c, native


def controlled_phase(c: Circuit, control: int, target: int, theta: float) -> None:
    """src/algorithms.rs:139-145: exact CP(theta) up to the global phase e^{-i theta/4}."""
    c.push(Gate.Rz(theta / 2.0), control)
    c.push(Gate.Rz(theta / 2.0), target)
    c.cnot(control, target)
    c.push(Gate.Rz(-theta / 2.0), target)
    c.cnot(control, target)


def qft_circuit(qubits: int) -> Circuit:
    """qft_circuit, src/algorithms.rs:270-286."""
    c = Circuit()
    for i in range(qubits // 2):
        c.swap(i, qubits - 1 - i)
    for i in range(qubits):
        c.push(Gate.H, i)
        for j in range(i + 1, qubits):
            controlled_phase(c, j, i, math.pi / (1 << (j - i)))
    return c


# ---------------------------------------------------------------------------
# Hardware-efficient ansatz + Heisenberg chain (VQE workload)
# ---------------------------------------------------------------------------
def hea_circuit(qubits: int, layers: int, entanglement: str = "linear") -> ParameterizedCircuit:
    """HEABuilder::build (src/algorithms_templates.rs:230-276) as a
    ParameterizedCircuit: per layer, Ry(p[idx]) then Rz(p[idx+1]) on each
    qubit, then the CNOT pattern (Linear: CNOT(i, i+1))."""
    pc = ParameterizedCircuit(layers * qubits * 2)
    idx = 0
    for _ in range(layers):
        for i in range(qubits):
            pc.rotation(RotationAxis.Y, i, idx)
            idx += 1
            pc.rotation(RotationAxis.Z, i, idx)
            idx += 1
        if entanglement == "linear":
            for i in range(qubits - 1):
                pc.cnot(i, i + 1)
        elif entanglement == "circular":
            for i in range(qubits):
                pc.cnot(i, (i + 1) % qubits)
        elif entanglement == "full":
            for i in range(qubits):
                for j in range(i + 1, qubits):
                    pc.cnot(i, j)
        else:
            raise ValueError(f"unknown entanglement {entanglement!r}")
    return pc


def heisenberg_hamiltonian(qubits: int, seed: int, lo: float = 0.5, hi: float = 1.5) -> PauliSum:
    """H = sum_i J_i (XX + YY + ZZ)_{i,i+1} (open chain) + sum_i h_i Z_i with
    J_i, h_i ~ U[lo, hi] drawn from SmallRng(seed): J_0..J_{n-2} then h_0..h_{n-1}."""
    rng = SmallRng(seed)
    J = [lo + (hi - lo) * rng.next_f64() for _ in range(qubits - 1)]
    h = [lo + (hi - lo) * rng.next_f64() for _ in range(qubits)]
    H = PauliSum(qubits)
    for i in range(qubits - 1):
        for p in "XYZ":
            w = ["I"] * qubits
            w[i] = w[i + 1] = p
            H.add(J[i], "".join(w))
    for i in range(qubits):
        w = ["I"] * qubits
        w[i] = "Z"
        H.add(h[i], "".join(w))
    return H


def uniform_parameters(count: int, sets: int, seed: int, lo: float = -math.pi, hi: float = math.pi):
    """`sets` x `count` parameters ~ U[lo, hi] from SmallRng(seed), row-major."""
    import numpy as np

    rng = SmallRng(seed)
    out = np.empty((sets, count), dtype=np.float64)
    for s in range(sets):
        for k in range(count):
            out[s, k] = lo + (hi - lo) * rng.next_f64()
    return out


# ---------------------------------------------------------------------------
# MaxCut / QAOA (uneven QAOA workload)
# ---------------------------------------------------------------------------
@dataclass
class MaxCutProblem:
    """src/mps_qaoa.rs:42-211."""

    num_qubits: int
    edges: List[Tuple[int, int, float]] = field(default_factory=list)

    @staticmethod
    def random_3_regular(num_qubits: int, seed: int) -> "MaxCutProblem":
        """Configuration model with rejection, src/mps_qaoa.rs:58-95 (even n >= 4)."""
        if num_qubits < 4 or num_qubits % 2:
            raise ValueError("3-regular graphs need an even number of vertices >= 4")
        rng = SmallRng(seed)
        for _ in range(10_000):
            stubs = [v for v in range(num_qubits) for _ in range(3)]
            fisher_yates(rng, stubs)
            edges: List[Tuple[int, int, float]] = []
            ok = True
            for k in range(0, len(stubs), 2):
                a, b = stubs[k], stubs[k + 1]
                if a == b:
                    ok = False
                    break
                u, v = (a, b) if a < b else (b, a)
                if any(eu == u and ev == v for eu, ev, _ in edges):
                    ok = False
                    break
                edges.append((u, v, 1.0))
            if ok:
                edges.sort(key=lambda e: (e[0], e[1]))
                return MaxCutProblem(num_qubits, edges)
        raise RuntimeError("failed to sample a 3-regular graph")

    def with_random_weights(self, seed: int, min_weight: float, max_weight: float) -> "MaxCutProblem":
        """src/mps_qaoa.rs:180-199."""
        if not (min_weight > 0.0 and max_weight > min_weight):
            raise ValueError("need 0 < min_weight < max_weight")
        rng = SmallRng(seed)
        return MaxCutProblem(self.num_qubits, [(u, v, min_weight + rng.next_f64() * (max_weight - min_weight))
                                               for u, v, _ in self.edges])

    def cut_value(self, bits: int) -> float:
        """src/mps_qaoa.rs:201-211."""
        return sum(w for u, v, w in self.edges if ((bits >> u) & 1) != ((bits >> v) & 1))

    def total_weight(self) -> float:
        return sum(w for _, _, w in self.edges)


def qaoa_maxcut(num_qubits: int, edges: Sequence[Tuple[int, int, float]], p: int, betas: Sequence[float],
                gammas: Sequence[float]) -> Circuit:
    """qaoa_maxcut, src/algorithms.rs:110-136: H layer, then per layer
    CNOT(u,v) Rz(2*gamma*w)_v CNOT(u,v) per edge and Rx(2*beta) per qubit."""
    c = Circuit()
    for i in range(num_qubits):
        c.push(Gate.H, i)
    for layer in range(p):
        for u, v, w in edges:
            c.cnot(u, v)
            c.push(Gate.Rz(2.0 * gammas[layer] * w), v)
            c.cnot(u, v)
        for i in range(num_qubits):
            c.push(Gate.Rx(2.0 * betas[layer]), i)
    return c


def qaoa_parameterized(num_qubits: int, edges: Sequence[Tuple[int, int, float]], p: int) -> ParameterizedCircuit:
    """The QAOA circuit with every rotation occurrence as its own parameter
    (the reference's Rotation ops carry no scale factor, src/lib.rs:2951-2958):
    occurrence order is, per layer, one Rz per edge then one Rx per qubit.
    Chain rule to (gamma, beta): see qaoa_chain_rule."""
    occ = p * (len(edges) + num_qubits)
    pc = ParameterizedCircuit(occ)
    for i in range(num_qubits):
        pc.push(Gate.H, i)
    k = 0
    for _ in range(p):
        for u, v, _w in edges:
            pc.cnot(u, v)
            pc.rotation(RotationAxis.Z, v, k)
            k += 1
            pc.cnot(u, v)
        for i in range(num_qubits):
            pc.rotation(RotationAxis.X, i, k)
            k += 1
    return pc


def qaoa_occurrence_angles(edges, num_qubits: int, p: int, gammas, betas):
    """Occurrence angles of qaoa_parameterized: 2*gamma_l*w_e and 2*beta_l."""
    import numpy as np

    out = []
    for layer in range(p):
        out.extend(2.0 * gammas[layer] * w for _u, _v, w in edges)
        out.extend(2.0 * betas[layer] for _ in range(num_qubits))
    return np.array(out, dtype=np.float64)


def qaoa_chain_rule(occ_grad, edges, num_qubits: int, p: int):
    """dE/dgamma_l = sum_e 2 w_e dE/dtheta_(l,e); dE/dbeta_l = sum_q 2 dE/dtheta_(l,q)."""
    import numpy as np

    g = np.zeros(p)
    b = np.zeros(p)
    k = 0
    for layer in range(p):
        for _u, _v, w in edges:
            g[layer] += 2.0 * w * occ_grad[k]
            k += 1
        for _ in range(num_qubits):
            b[layer] += 2.0 * occ_grad[k]
            k += 1
    return g, b


def maxcut_cost_hamiltonian(problem: MaxCutProblem) -> PauliSum:
    """Weighted ZZ cost sum_e w_e Z_u Z_v (the workload's convention; the
    reference's cut weight is sum_e w_e (1 - Z_u Z_v)/2, src/mps_qaoa.rs:13-15,
    i.e. total_weight/2 - (this)/2)."""
    H = PauliSum(problem.num_qubits)
    for u, v, w in problem.edges:
        word = ["I"] * problem.num_qubits
        word[u] = word[v] = "Z"
        H.add(w, "".join(word))
    return H


def qaoa_uneven_instances(count: int, seed: int, p: int = 6, qubit_choices=(18, 20, 22, 24, 26),
                          weight_range=(0.5, 1.0)):
    """The uneven QAOA MaxCut batch: instance i draws its width uniformly from
    `qubit_choices` (even widths: 3-regular graphs need an even vertex count,
    src/mps_qaoa.rs:59-62), a random 3-regular graph and U[0.5, 1.0) weights
    (with_random_weights), and gamma, beta ~ U[0, pi] per layer.  All draws
    come from SmallRng(seed) (graph and weight seeds are drawn from it too).
    Returns a list of dicts with problem, gammas, betas."""
    rng = SmallRng(seed)
    out = []
    for _ in range(count):
        n = qubit_choices[rng.next_u64() % len(qubit_choices)]
        gseed = rng.next_u64()
        wseed = rng.next_u64()
        prob = MaxCutProblem.random_3_regular(n, gseed).with_random_weights(wseed, *weight_range)
        gammas = [math.pi * rng.next_f64() for _ in range(p)]
        betas = [math.pi * rng.next_f64() for _ in range(p)]
        out.append({"problem": prob, "gammas": gammas, "betas": betas, "p": p})
    return out
