"""Batches of different widths: qf_batch_upload / qf_batch_zero_uneven hold
states of several widths back to back; gates, circuits and Pauli energies
apply to every state, validated against every width before anything runs;
the equal-width entry point keeps the reference's IncompatibleBatch
(src/cuda.rs:871-954).  qf_uneven_expectation gives the energies of
parameterized circuits of different widths in one device schedule.  All
against the oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from quantumforge_amd import Gate, PauliSum, QuantumError, StateVector
from quantumforge_amd import gradients as G
from quantumforge_amd.backend import Plan
from quantumforge_amd.workloads import (hea_circuit, heisenberg_hamiltonian, maxcut_cost_hamiltonian,
                                        qaoa_occurrence_angles, qaoa_parameterized, qaoa_uneven_instances,
                                        random_layer_circuit, uniform_parameters)

from .conftest import close

pytestmark = pytest.mark.gpu


def random_state(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return a / np.linalg.norm(a)


def test_uneven_batch_gates_circuits_energies(backend):
    widths = [6, 9, 9, 7, 6]
    a0 = [random_state(n, 10 + i) for i, n in enumerate(widths)]
    batch = backend.upload_uneven_batch([StateVector(a, n) for a, n in zip(a0, widths)])
    assert batch.widths == widths and batch.qubits() == 9 and batch.batch_size() == 5
    oracle = [O.State(n, a) for a, n in zip(a0, widths)]
    batch.apply_gate(Gate.H, 2)
    batch.apply_cnot(0, 5)
    batch.apply_gate(Gate.Ry(0.37), 4)
    c = random_layer_circuit(6, 4, seed=5, pairs_per_layer=3)
    batch.run_circuit(c)
    for o in oracle:
        o.apply(4, 2)
        o.cnot(0, 5)
        o.apply(10, 4, 0.37)
        o.run(c.to_ops())
    for got, o in zip(batch.download(), oracle):
        close(got.amplitudes, o.amplitudes(), 1e-12)
    H = PauliSum(6, [(0.5, "XXIIII"), (-0.25, "ZIZIIY"), (1.0, "IIIIIZ")])
    e = batch.expectation_pauli_sum(H)
    t = H.to_terms()
    for k, o in enumerate(oracle):
        assert abs(e[k] - o.pauli_expectation(t).real) < 1e-12


def test_uneven_batch_errors_leave_the_states_untouched(backend):
    widths = [6, 8]
    a0 = [random_state(n, 20 + i) for i, n in enumerate(widths)]
    batch = backend.upload_uneven_batch([StateVector(a, n) for a, n in zip(a0, widths)])
    with pytest.raises(QuantumError) as e:
        batch.apply_gate(Gate.X, 7)  # fits the 8-qubit state, not the 6-qubit one
    assert e.value.kind == "QubitOutOfRange"
    c = random_layer_circuit(6, 2, seed=1, pairs_per_layer=2)
    with pytest.raises(QuantumError) as e:
        batch.run_circuit(c, Plan(c, 6))
    assert e.value.kind == "IncompatibleBatch"
    with pytest.raises(QuantumError) as e:
        batch.measure_with_uniforms(0, [0.5, 0.5])
    assert e.value.kind == "IncompatibleBatch"
    for got, a in zip(batch.download(), a0):
        assert np.array_equal(got.amplitudes, a)
    # The reference's equal-width entry point.
    with pytest.raises(QuantumError) as e:
        backend.upload_batch([StateVector(a, n) for a, n in zip(a0, widths)])
    assert e.value.kind == "IncompatibleBatch"
    with pytest.raises(QuantumError) as e:
        backend.upload_uneven_batch([])
    assert e.value.kind == "EmptyBatch"


def test_uneven_zero_batch_runs_compiled_segments(backend):
    widths = [18, 19, 18, 12]
    batch = backend.zero_uneven_batch(widths)
    for got, n in zip(batch.download(), widths):
        want = np.zeros(1 << n, dtype=complex)
        want[0] = 1
        assert np.array_equal(got.amplitudes, want)
    c = random_layer_circuit(12, 6, seed=9, pairs_per_layer=5)
    batch.run_circuit(c)
    for got, n in zip(batch.download(), widths):
        close(got.amplitudes, O.run_circuit(n, c.to_ops()), 1e-12)


def test_uneven_expectation_matches_oracle_and_adjoint_energies(backend):
    specs = qaoa_uneven_instances(3, seed=31, p=2, qubit_choices=(12, 14, 16))
    insts = []
    for s in specs:
        prob = s["problem"]
        pc = qaoa_parameterized(prob.num_qubits, prob.edges, s["p"])
        ang = qaoa_occurrence_angles(prob.edges, prob.num_qubits, s["p"], s["gammas"], s["betas"])
        insts.append((pc, prob.num_qubits, ang, maxcut_cost_hamiltonian(prob)))
    pc = hea_circuit(8, 2)  # narrower than the tile: the single-instance path
    insts.append((pc, 8, uniform_parameters(pc.parameter_count, 1, seed=8)[0], heisenberg_hamiltonian(8, seed=8)))
    e, stats = G.uneven_expectation(backend, insts)
    ea, _, _ = G.uneven_adjoint_gradient(backend, insts)
    for k, (pc, q, ang, H) in enumerate(insts):
        want = O.pc_expectation(pc.to_pops(), pc.parameter_count, 0, ang, [], q, H.to_terms())
        assert abs(e[k] - want) < 1e-11, k
        assert abs(e[k] - ea[k]) < 1e-12, k
    assert stats["waves"] >= 1 and stats["graph_replays"] == stats["waves"]
