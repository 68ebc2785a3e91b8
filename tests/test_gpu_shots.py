"""GPU measurement and seeded shots against the oracle: resident
measurement (tests/quantum_core.rs:601-619), shot circuits with classical
feed-forward in both execution forms (tests/quantum_core.rs:651-681), batch
measurement with explicit uniforms, masked gates and the error variants."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from quantumforge_amd import Circuit, Gate, QuantumError, ShotResult, StateVector
from quantumforge_amd.workloads import SmallRng, random_layer_circuit

from .conftest import close

pytestmark = pytest.mark.gpu


def random_state(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return a / np.linalg.norm(a)


def test_resident_measurement_matches_cpu_reference(backend):
    o = O.State(2)
    o.apply(4, 1)
    dev = backend.upload(StateVector(o.amplitudes(), 2))
    expected = o.measure(1, SmallRng(342).next_f64())
    assert dev.measure(1, SmallRng(342)) == expected
    close(dev.download().amplitudes, o.amplitudes(), 1e-12)


def test_measurement_on_random_states(backend):
    for seed in range(4):
        n = 12
        a0 = random_state(n, seed)
        q = (3 * seed + 1) % n
        o = O.State(n, a0)
        rng_cpu, rng_gpu = SmallRng(seed), SmallRng(seed)
        want = o.measure(q, rng_cpu.next_f64())
        dev = backend.upload(StateVector(a0, n))
        assert dev.measure(q, rng_gpu) == want
        close(dev.download().amplitudes, o.amplitudes(), 1e-12)


def feedforward_circuit() -> Circuit:
    c = Circuit()
    c.push(Gate.H, 0)
    c.push(Gate.H, 1)
    c.measure(0, 0)
    c.if_classical_bit(0, 1, Gate.X, 2)
    c.measure(1, 1)
    c.measure(2, 2)
    return c


def test_seeded_feedforward_shots_match_cpu(backend):
    c = feedforward_circuit()
    init = StateVector.zero(3)
    cpu = ShotResult.from_registers(O.run_shots(init.amplitudes, 3, c.to_ops(), 64, 721))
    result = backend.run_circuit_shots(init, c, 64, 721)
    batched = backend.run_circuit_shots_batched(init, c, 64, 721)
    assert result == cpu
    assert batched == cpu
    assert result.shots == 64 and sum(result.counts.values()) == 64
    assert set(result.counts) == {(0, 0, 0), (1, 1, 1), (0, 1, 0), (1, 0, 1)}


@pytest.mark.parametrize("seed", [3, 17])
def test_mid_circuit_measurement_shots_match_cpu(backend, seed):
    n = 10
    c = Circuit.from_ops(random_layer_circuit(n, 4, seed=seed, pairs_per_layer=5).to_ops())
    c.measure(2, 0)
    c.if_classical_bit(0, 1, Gate.Rx(0.7), 4)
    c.extend(random_layer_circuit(n, 3, seed=seed + 1, pairs_per_layer=5))
    c.measure(4, 1)
    c.if_classical_bit(1, 0, Gate.H, 0)
    c.measure(0, 3)
    c.measure(7, 2)
    init = StateVector(random_state(n, seed), n)
    cpu = O.run_shots(init.amplitudes, n, c.to_ops(), 150, seed)
    want = ShotResult.from_registers(cpu)
    assert backend.run_circuit_shots_batched(init, c, 150, seed) == want
    assert backend.run_circuit_shots(init, c, 150, seed) == want


def test_batch_measure_with_uniforms_and_masked_gates(backend):
    n = 8
    states = [random_state(n, s) for s in range(3)]
    batch = backend.upload_batch([StateVector(a, n) for a in states])
    u = [0.1, 0.9, 0.5]
    outs = batch.measure_with_uniforms(5, u)
    cpu = [O.State(n, a) for a in states]
    assert outs == [cpu[i].measure(5, u[i]) for i in range(3)]
    batch.apply_gate_masked(Gate.Ry(0.3), 2, [1, 0, 1])
    for i in (0, 2):
        cpu[i].apply(10, 2, 0.3)
    for got, o in zip(batch.download(), cpu):
        close(got.amplitudes, o.amplitudes(), 1e-12)


def test_measurement_errors(backend):
    dev = backend.zero_state(3)
    with pytest.raises(QuantumError) as e:
        dev.measure(3, SmallRng(1))
    assert e.value.kind == "QubitOutOfRange"
    a = np.zeros(2, complex)
    a[0], a[1] = math.sqrt(1 - 1e-18), 1e-9
    batch = backend.upload_batch([StateVector(a, 1)])
    with pytest.raises(QuantumError) as e:
        batch.measure_with_uniforms(0, [0.0])
    assert e.value.kind == "InvalidProbability"
    for bad in ([1.0], [float("nan")], [0.1, 0.2]):
        with pytest.raises(QuantumError) as e:
            batch.measure_with_uniforms(0, bad)
        assert e.value.kind == "InvalidBatchSamples"


def test_shot_errors(backend):
    init = StateVector.zero(2)
    c = Circuit()
    c.measure(0, 4096)
    with pytest.raises(QuantumError) as e:
        backend.run_circuit_shots(init, c, 4, 1)
    assert e.value.kind == "ClassicalBitOutOfRange"
    with pytest.raises(QuantumError) as e:
        backend.run_circuit_shots_batched(init, Circuit(), 0, 1)
    assert e.value.kind == "InvalidShotCount"
    c = Circuit()
    c.channel(0)
    c.measure(0, 0)
    with pytest.raises(QuantumError) as e:
        backend.run_circuit_shots(init, c, 4, 1)
    assert e.value.kind == "OperationRequiresDensityMatrix"
