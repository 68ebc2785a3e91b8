"""GPU parity of the resident-state path against the oracle: the reference's
CUDA tests (tests/quantum_core.rs:401-599, 1e-12 per amplitude component),
bit-exactness of the exact (unfused) lowering, the error variants, batches
and the FP64 reductions."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from quantumforge_amd import Circuit, Gate, PauliSum, QuantumError, RotationAxis, SingleQubitObservable, StateVector
from quantumforge_amd.workloads import qft_circuit, random_layer_circuit

from .conftest import close

pytestmark = pytest.mark.gpu


def oracle_state(sv: StateVector) -> O.State:
    return O.State(sv.qubits, sv.amplitudes)


def test_rotation_chain_matches_cpu_reference(backend):
    # tests/quantum_core.rs:401-420 (cuda_ry_matches_cpu_reference).
    o = O.State(5)
    o.apply(4, 0)
    o.apply(10, 3, -0.2)
    sv = StateVector(o.amplitudes(), 5)
    backend.apply_rotation_chain(sv, [(RotationAxis.X, 2, 0.73), (RotationAxis.Y, 4, -0.2),
                                      (RotationAxis.Z, 1, 1.1)])
    o.apply(9, 2, 0.73)
    o.apply(10, 4, -0.2)
    o.apply(11, 1, 1.1)
    close(sv.amplitudes, o.amplitudes(), 1e-12)
    assert np.array_equal(sv.amplitudes, o.amplitudes())  # exact lowering: bit-identical


def test_device_state_stays_resident_and_outlives_backend():
    # tests/quantum_core.rs:422-447: the state keeps working after the backend is dropped.
    from quantumforge_amd.backend import HipBackend

    o = O.State(4)
    o.apply(4, 0)
    be = HipBackend(0)
    dev = be.upload(StateVector(o.amplitudes(), 4))
    dev.apply_rotation(RotationAxis.X, 2, 0.31)
    dev.apply_rotation_chain([(RotationAxis.Y, 1, -0.83), (RotationAxis.Z, 3, 1.27)])
    del be
    o.apply(9, 2, 0.31)
    o.apply(10, 1, -0.83)
    o.apply(11, 3, 1.27)
    close(dev.download().amplitudes, o.amplitudes(), 1e-12)


def test_entangling_gates_match_cpu_reference(backend):
    # tests/quantum_core.rs:449-498.
    o = O.State(5)
    for q in range(5):
        o.apply(4, q)
    o.apply(11, 2, 0.37)
    dev = backend.upload(StateVector(o.amplitudes(), 5))
    dev.apply_cnot(0, 4)
    o.cnot(0, 4)
    assert np.array_equal(dev.download().amplitudes, o.amplitudes())
    dev.apply_cz(1, 3)
    o.cz(1, 3)
    assert np.array_equal(dev.download().amplitudes, o.amplitudes())
    dev.apply_swap(0, 2)
    o.swap(0, 2)
    assert np.array_equal(dev.download().amplitudes, o.amplitudes())


def test_unitary_circuit_api(backend):
    # tests/quantum_core.rs:500-517 (cuda_executes_the_unitary_circuit_api).
    c = Circuit()
    c.push(Gate.H, 0)
    c.push(Gate.X, 1)
    c.push(Gate.Y, 2)
    c.push(Gate.Z, 3)
    c.push(Gate.Rx(0.19), 0)
    c.push(Gate.Ry(-0.57), 1)
    c.push(Gate.Rz(1.03), 2)
    c.cnot(0, 3)
    c.cz(1, 2)
    c.swap(0, 1)
    want = O.run_circuit(4, c.to_ops())
    for fuse in (False, True):
        backend.set_plan_options(fuse=fuse)
        got = backend.run_circuit(StateVector.zero(4), c)
        close(got.amplitudes, want, 1e-12)
    backend.set_plan_options(fuse=True)


@pytest.mark.parametrize("n", [1, 2, 3, 6, 9])
def test_every_gate_every_qubit_exact(backend, n):
    rng = np.random.default_rng(n)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    names = ["I", "X", "Y", "Z", "H", "S", "Sdg", "T", "Tdg", "Rx", "Ry", "Rz"]
    dev = backend.upload(StateVector(psi.copy(), n))
    o = O.State(n, psi)
    for q in range(n):
        for gid, name in enumerate(names):
            th = float(rng.uniform(-7, 7))
            dev.apply_gate(Gate(name, th), q)
            o.apply(gid, q, th)
    assert np.array_equal(dev.download().amplitudes, o.amplitudes())


def test_config1_random_circuit_exact_and_fused(backend):
    # 20 qubits, depth 40, seeded layers (the CPU-runnable workload).
    n = 20
    c = random_layer_circuit(n, 40, seed=2024)
    ops = c.to_ops()
    want = O.run_circuit(n, ops)
    backend.set_plan_options(fuse=False)
    got = backend.run_circuit(StateVector.zero(n), c)
    assert np.array_equal(got.amplitudes, want), "exact lowering must be bit-identical"
    backend.set_plan_options(fuse=True)
    got = backend.run_circuit(StateVector.zero(n), c)
    close(got.amplitudes.real, want.real, 1e-12)
    close(got.amplitudes.imag, want.imag, 1e-12)
    dev = backend.zero_state(n)
    dev.run_circuit(c)
    p = dev.probabilities()
    close(p, np.abs(want) ** 2, 1e-12)
    assert abs(dev.norm_sqr() - 1.0) < 1e-12


@pytest.mark.parametrize("tile,low,kernel,reg", [(6, 2, 1, 0), (8, 3, 1, 0), (10, 4, 1, 0), (11, 5, 1, 0),
                                                 (12, 5, 1, 0), (13, 6, 1, 0), (11, 5, 2, 3), (12, 5, 2, 4),
                                                 (10, 4, 2, 3), (9, 3, 2, 3), (11, 6, 2, 4), (12, 6, 2, 4)])
def test_fused_tile_shapes(backend, tile, low, kernel, reg):
    n = 14
    c = random_layer_circuit(n, 12, seed=tile * 7 + low, pairs_per_layer=7)
    c.swap(0, 13)
    c.swap(3, 9)
    c.swap(1, 2)
    want = O.run_circuit(n, c.to_ops())
    backend.set_plan_options(fuse=True, tile_qubits=tile, low_qubits=low, kernel=kernel, reg_bits=reg)
    got = backend.run_circuit(StateVector.zero(n), c)
    backend.set_plan_options(fuse=True)
    close(got.amplitudes, want, 1e-12)


@pytest.mark.parametrize("kernel", [1, 2])
def test_fused_batch_and_complex64_both_kernels(backend, kernel):
    n, B = 13, 3
    c = random_layer_circuit(n, 10, seed=kernel, pairs_per_layer=6)
    rng = np.random.default_rng(kernel)
    states = []
    for _ in range(B):
        a = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        states.append(StateVector(a / np.linalg.norm(a), n))
    backend.set_plan_options(fuse=True, kernel=kernel)
    out = backend.run_circuit_batch(states, c)
    for s, r in zip(states, out):
        o = O.State(n, s.amplitudes)
        o.run(c.to_ops())
        close(r.amplitudes, o.amplitudes(), 1e-12)
    dev = backend.zero_state(n, precision="complex64")
    dev.run_circuit(c)
    want = O.run_circuit(n, c.to_ops())
    got = dev.download().amplitudes
    assert 1 - abs(np.vdot(want, got)) ** 2 / np.vdot(got, got).real < 1e-9
    backend.set_plan_options(fuse=True)


def test_qft_fused(backend):
    n = 10
    c = qft_circuit(n)
    want = O.run_circuit(n, c.to_ops())
    got = backend.run_circuit(StateVector.zero(n), c)
    close(got.amplitudes, want, 1e-12)


def test_reductions(backend):
    n = 12
    c = random_layer_circuit(n, 6, seed=5, pairs_per_layer=6)
    o = O.State(n)
    o.run(c.to_ops())
    dev = backend.zero_state(n)
    dev.run_circuit(c)
    for q in (0, 5, 11):
        for obs in (0, 1, 2):
            assert abs(dev.expectation(SingleQubitObservable(obs), q) - o.expectation(obs, q)) < 1e-13
    H = PauliSum(n, [(0.5, "XX" + "I" * (n - 2)), (-0.25, "I" * 3 + "YY" + "I" * (n - 5)),
                     (1.5, "Z" * n), (0.75, "I" * (n - 2) + "XY"), (0.1 + 0.3j, "Y" + "I" * (n - 2) + "Z")])
    got = dev.expectation_pauli_sum(H)
    want = o.pauli_expectation(H.to_terms())
    assert abs(got - want) < 1e-13
    other = backend.zero_state(n)
    assert abs(dev.inner(other) - np.conj(o.amplitudes()[0])) < 1e-15


def test_error_variants(backend):
    dev = backend.zero_state(3)
    with pytest.raises(QuantumError) as e:
        dev.apply_gate(Gate.X, 3)
    assert e.value.kind == "QubitOutOfRange"
    with pytest.raises(QuantumError) as e:
        dev.apply_gate(Gate.Rx(float("nan")), 0)
    assert e.value.kind == "NonFiniteAngle"
    with pytest.raises(QuantumError) as e:
        dev.apply_cnot(2, 2)
    assert e.value.kind == "InvalidQubitPair"
    with pytest.raises(QuantumError) as e:
        dev.apply_cz(0, 7)
    assert e.value.kind == "QubitOutOfRange"
    c = Circuit()
    c.push(Gate.H, 0)
    c.channel(0)
    with pytest.raises(QuantumError) as e:
        dev.run_circuit(c)
    assert e.value.kind == "OperationRequiresDensityMatrix"
    c = Circuit()
    c.measure(0, 0)
    with pytest.raises(QuantumError) as e:
        dev.run_circuit(c)
    assert e.value.kind == "OperationRequiresShotExecution"
    # A rejected call leaves the state untouched.
    e0 = np.zeros(8, complex)
    e0[0] = 1
    assert np.array_equal(dev.download().amplitudes, e0)
    with pytest.raises(QuantumError) as e:
        backend.upload_batch([])
    assert e.value.kind == "EmptyBatch"
    with pytest.raises(QuantumError) as e:
        backend.upload_batch([StateVector.zero(2), StateVector.zero(3)])
    assert e.value.kind == "IncompatibleBatch"


def test_batch_matches_cpu(backend):
    # tests/quantum_core.rs:519-599: equal-width batch, one circuit for every state.
    n, B = 8, 6
    rng = np.random.default_rng(9)
    states = []
    for _ in range(B):
        a = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        states.append(StateVector(a / np.linalg.norm(a), n))
    c = random_layer_circuit(n, 10, seed=77, pairs_per_layer=4)
    for fuse in (False, True):
        backend.set_plan_options(fuse=fuse)
        out = backend.run_circuit_batch(states, c)
        for s, r in zip(states, out):
            o = O.State(n, s.amplitudes)
            o.run(c.to_ops())
            if fuse:
                close(r.amplitudes, o.amplitudes(), 1e-12)
            else:
                assert np.array_equal(r.amplitudes, o.amplitudes())
    backend.set_plan_options(fuse=True)
    b = backend.zero_batch(n, B)
    b.apply_gate(Gate.H, 0)
    b.apply_cnot(0, 1)
    got = b.download()
    for s in got:
        assert abs(s.amplitudes[0] - 1 / math.sqrt(2)) < 1e-15 and abs(s.amplitudes[3] - 1 / math.sqrt(2)) < 1e-15


def test_complex64_state(backend):
    n = 16
    c = random_layer_circuit(n, 20, seed=3, pairs_per_layer=8)
    want = O.run_circuit(n, c.to_ops())
    dev = backend.zero_state(n, precision="complex64")
    dev.run_circuit(c)
    got = dev.download().amplitudes
    fid = abs(np.vdot(want, got)) ** 2 / (np.vdot(got, got).real)
    assert 1 - fid < 1e-9, f"complex64 fidelity defect {1 - fid:.2e}"
    assert abs(dev.norm_sqr() - 1) < 1e-5
