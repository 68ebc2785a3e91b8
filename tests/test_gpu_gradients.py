"""GPU gradients against the oracle: the reference's gradient tests
(tests/gradients.rs, tests/quantum_core.rs:79-123) run through the HIP engine,
the batched adjoint against the oracle's shift rule and finite differences,
and the error variants."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from quantumforge_amd import ParameterizedCircuit, PauliSum, QuantumError, RotationAxis, SingleQubitObservable
from quantumforge_amd import gradients as G
from quantumforge_amd.workloads import (MaxCutProblem, hea_circuit, heisenberg_hamiltonian, maxcut_cost_hamiltonian,
                                        qaoa_occurrence_angles, qaoa_parameterized, uniform_parameters)

from .test_oracle_gradients import seven_qubit_circuit

pytestmark = pytest.mark.gpu
EXACT = (G.GradientStrategy.ParameterShift, G.GradientStrategy.Adjoint, G.GradientStrategy.Backprop)


def test_ry_gradient_matches_analytic(backend):
    pc = ParameterizedCircuit(1)
    pc.rotation(RotationAxis.Y, 0, 0)
    for st in EXACT:
        est = G.gradient(backend, pc, G.GradientOptions(st), [0.7], 1, (SingleQubitObservable.Z, 0))
        assert abs(est.gradient[0] + math.sin(0.7)) < 1e-10, st
        assert est.strategy == st
        assert abs(est.value - math.cos(0.7)) < 1e-14


def test_all_exact_strategies_agree_7q(backend):
    pc, params = seven_qubit_circuit()
    t = PauliSum.single(SingleQubitObservable.Z, 0, 7).to_terms()
    want = O.pc_shift_gradient(pc.to_pops(), 42, 0, params, [], 7, t)
    for st in EXACT:
        got = G.gradient(backend, pc, G.GradientOptions(st), params, 7, (SingleQubitObservable.Z, 0)).gradient
        assert np.max(np.abs(got - want)) < 1e-9, st
    # evaluation counts, tests/gradients.rs:324-361
    est = G.gradient(backend, pc, G.GradientOptions(G.GradientStrategy.ParameterShift), params, 7,
                     (SingleQubitObservable.Z, 0))
    assert est.evaluations == 2 * 42
    est = G.gradient(backend, pc, G.GradientOptions(), params, 7, (SingleQubitObservable.Z, 0))
    assert est.strategy == G.GradientStrategy.Adjoint and est.evaluations == 2
    est = G.gradient(backend, pc, G.GradientOptions(shots=100), params, 7, (SingleQubitObservable.Z, 0))
    assert est.strategy == G.GradientStrategy.ParameterShift


def test_adjoint_equals_shift_with_inputs_and_shared_parameters(backend):
    pc = ParameterizedCircuit.with_inputs(2, 1)
    pc.encode_rotation(RotationAxis.Y, 0, 0)
    pc.rotation(RotationAxis.X, 1, 0)
    pc.cnot(0, 1)
    pc.rotation(RotationAxis.Z, 0, 1)
    pc.rotation(RotationAxis.Y, 1, 0)
    for obs in (0, 1, 2):
        t = PauliSum.single(SingleQubitObservable(obs), 1, 2).to_terms()
        want = O.pc_adjoint_gradient(pc.to_pops(), 2, 1, [0.31, -0.82], [0.43], 2, t)
        for st in EXACT:
            got = G.gradient(backend, pc, G.GradientOptions(st), [0.31, -0.82], 2,
                             (SingleQubitObservable(obs), 1), inputs=[0.43]).gradient
            assert np.max(np.abs(got - want)) < 1e-12, (obs, st)
    pc.channel(0)
    with pytest.raises(QuantumError) as e:
        G.gradient(backend, pc, G.GradientOptions(), [0.31, -0.82], 2, (SingleQubitObservable.Z, 0), inputs=[0.43])
    assert e.value.kind == "OperationRequiresDensityMatrix"


def test_finite_difference_second_order(backend):
    pc = ParameterizedCircuit(2)
    pc.rotation(RotationAxis.Y, 0, 0)
    pc.rotation(RotationAxis.X, 1, 1)
    pc.cnot(0, 1)
    pc.rotation(RotationAxis.Y, 2, 0)
    exact = G.parameter_shift_gradient(backend, pc, [0.5, -0.4], 3, (SingleQubitObservable.Z, 0))
    prev = math.inf
    for h in (1e-2, 1e-3, 1e-4):
        fd = G.finite_difference_gradient_with_inputs(backend, pc, [], [0.5, -0.4], 3, (SingleQubitObservable.Z, 0), h)
        err = np.max(np.abs(fd - exact))
        assert err * 50 < prev
        prev = err
    assert prev < 1e-8
    for h in (0.0, -1.0, math.nan, math.inf):
        with pytest.raises(QuantumError) as e:
            G.finite_difference_gradient_with_inputs(backend, pc, [], [0.5, -0.4], 3, (SingleQubitObservable.Z, 0), h)
        assert e.value.kind == "InvalidFiniteDifferenceStep"


def test_validation_errors(backend):
    pc = ParameterizedCircuit.with_inputs(2, 1)
    pc.rotation(RotationAxis.Y, 0, 1)
    obs = (SingleQubitObservable.Z, 0)
    cases = [(dict(parameters=[0.1, 0.2], inputs=[]), "InvalidInputCount"),
             (dict(parameters=[0.1], inputs=[0.3]), "InvalidParameterCount"),
             (dict(parameters=[0.1, math.inf], inputs=[0.3]), "NonFiniteParameter"),
             (dict(parameters=[0.1, 0.2], inputs=[math.nan]), "NonFiniteParameter")]
    for kw, kind in cases:
        with pytest.raises(QuantumError) as e:
            G.gradient(backend, pc, G.GradientOptions(G.GradientStrategy.Adjoint), kw["parameters"], 1, obs,
                       inputs=kw["inputs"])
        assert e.value.kind == kind
    with pytest.raises(QuantumError) as e:
        G.gradient(backend, pc, G.GradientOptions(), [0.1, 0.2], 1, (SingleQubitObservable.Z, 3), inputs=[0.3])
    assert e.value.kind == "QubitOutOfRange"
    bad = ParameterizedCircuit(1)
    bad.rotation(RotationAxis.Y, 0, 5)
    with pytest.raises(QuantumError) as e:
        G.gradient(backend, bad, G.GradientOptions(), [0.1], 1, obs)
    assert e.value.kind == "InvalidParameterCount"


@pytest.mark.parametrize("n,layers,sets", [(6, 2, 8), (9, 3, 5), (12, 2, 3)])
def test_batched_adjoint_heisenberg_vs_oracle(backend, n, layers, sets):
    pc = hea_circuit(n, layers)
    H = heisenberg_hamiltonian(n, seed=n)
    params = uniform_parameters(pc.parameter_count, sets, seed=100 + n)
    e, g = G.batch_adjoint_gradient(backend, pc, n, params, H)
    t = H.to_terms()
    for s in range(sets):
        want_e = O.pc_expectation(pc.to_pops(), pc.parameter_count, 0, params[s], [], n, t)
        want_g = O.pc_shift_gradient(pc.to_pops(), pc.parameter_count, 0, params[s], [], n, t)
        assert abs(e[s] - want_e) < 1e-12
        assert np.max(np.abs(g[s] - want_g)) < 1e-10
    es, gs = G.batch_shift_gradient(backend, pc, n, params, H)
    assert np.max(np.abs(es - e)) < 1e-13
    assert np.max(np.abs(gs - g)) < 1e-10
    assert np.max(np.abs(G.batch_expectation(backend, pc, n, params, H) - e)) < 1e-13


def test_batched_adjoint_tile_shapes(backend):
    n = 10
    pc = hea_circuit(n, 2)
    H = heisenberg_hamiltonian(n, seed=1)
    params = uniform_parameters(pc.parameter_count, 4, seed=2)
    e0, g0 = G.batch_adjoint_gradient(backend, pc, n, params, H)
    for tile in (4, 6, 8, 10):
        for chunk in (0, 1, 3):
            e, g = G.batch_adjoint_gradient(backend, pc, n, params, H, tile_qubits=tile, chunk=chunk)
            assert np.max(np.abs(e - e0)) < 1e-13 and np.max(np.abs(g - g0)) < 1e-12, (tile, chunk)


@pytest.mark.parametrize("precision,tol", [("complex128", 1e-11), ("complex64", 2e-4)])
def test_large_batch_runs_the_compiled_adjoint_sweep(backend, precision, tol):
    # 256 sets of 14 qubits (2^22 amplitudes per launch): the forward and
    # reverse plans run as compiled passes, the reverse sweep in the
    # two-state adjoint form that records the rotation overlaps per tile.
    # The first sets must equal a 4-set batch (interpreting kernels) and the
    # oracle's adjoint gradient.
    n, sets = 14, 256
    pc = hea_circuit(n, 2)
    H = heisenberg_hamiltonian(n, seed=6)
    params = uniform_parameters(pc.parameter_count, sets, seed=17)
    e, g = G.batch_adjoint_gradient(backend, pc, n, params, H, precision=precision)
    e4, g4 = G.batch_adjoint_gradient(backend, pc, n, params[:4], H, precision=precision)
    assert np.max(np.abs(e[:4] - e4)) < tol and np.max(np.abs(g[:4] - g4)) < tol
    t = H.to_terms()
    for s in (0, 1, 255):
        want = O.pc_adjoint_gradient(pc.to_pops(), pc.parameter_count, 0, params[s], [], n, t)
        assert np.max(np.abs(g[s] - want)) < tol, s


def test_first_forward_pass_in_chunks_of_sets(backend):
    # 1030 sets of 14 qubits: the first forward pass runs over four chunks of
    # sets, 1/16, 2/16, 4/16 and 9/16 of them (the host fills the next chunk's
    # angle rows meanwhile), the other passes over all sets.  Sets on both
    # sides of every chunk boundary (64, 193, 450) against the oracle.
    n, sets = 14, 1030
    pc = hea_circuit(n, 2)
    H = heisenberg_hamiltonian(n, seed=8)
    params = uniform_parameters(pc.parameter_count, sets, seed=23)
    e, g = G.batch_adjoint_gradient(backend, pc, n, params, H)
    t = H.to_terms()
    for s in (0, 63, 64, 192, 193, 449, 450, 1029):
        want_e = O.pc_expectation(pc.to_pops(), pc.parameter_count, 0, params[s], [], n, t)
        want = O.pc_adjoint_gradient(pc.to_pops(), pc.parameter_count, 0, params[s], [], n, t)
        assert abs(e[s] - want_e) < 1e-11, s
        assert np.max(np.abs(g[s] - want)) < 1e-10, s


def test_kept_arena_runs_a_larger_call_in_as_many_waves(backend):
    # 100 sets leave an arena sized for 100; then 150 sets in chunks of 128
    # need 2 memory waves either way, so they run in the kept arena as waves
    # of 100 and 50 sets.  Sets on both sides of the wave boundary against
    # the oracle.
    n = 10
    pc = hea_circuit(n, 2)
    H = heisenberg_hamiltonian(n, seed=4)
    backend.release_cached()
    G.batch_adjoint_gradient(backend, pc, n, uniform_parameters(pc.parameter_count, 100, seed=31), H)
    params = uniform_parameters(pc.parameter_count, 150, seed=32)
    e, g = G.batch_adjoint_gradient(backend, pc, n, params, H, chunk=128)
    assert backend.last_gradient_profile()["waves"] == 2
    t = H.to_terms()
    for s in (0, 99, 100, 149):
        want = O.pc_adjoint_gradient(pc.to_pops(), pc.parameter_count, 0, params[s], [], n, t)
        assert abs(e[s] - O.pc_expectation(pc.to_pops(), pc.parameter_count, 0, params[s], [], n, t)) < 1e-11, s
        assert np.max(np.abs(g[s] - want)) < 1e-10, s


def test_adjoint_at_18_qubits_runs_compiled(backend):
    # From 18 qubits every gradient plan is compiled, one set included.
    n = 18
    pc = hea_circuit(n, 1)
    H = heisenberg_hamiltonian(n, seed=2)
    params = uniform_parameters(pc.parameter_count, 1, seed=3)
    e, g = G.batch_adjoint_gradient(backend, pc, n, params, H)
    t = H.to_terms()
    assert abs(e[0] - O.pc_expectation(pc.to_pops(), pc.parameter_count, 0, params[0], [], n, t)) < 1e-11
    want = O.pc_adjoint_gradient(pc.to_pops(), pc.parameter_count, 0, params[0], [], n, t)
    assert np.max(np.abs(g[0] - want)) < 1e-10


def test_complex64_adjoint_error_bound(backend):
    n = 10
    pc = hea_circuit(n, 3)
    H = heisenberg_hamiltonian(n, seed=4)
    params = uniform_parameters(pc.parameter_count, 4, seed=5)
    e64, g64 = G.batch_adjoint_gradient(backend, pc, n, params, H)
    e32, g32 = G.batch_adjoint_gradient(backend, pc, n, params, H, precision="complex64")
    assert np.max(np.abs(e32 - e64)) < 1e-4
    assert np.max(np.abs(g32 - g64)) < 1e-4


def test_qaoa_occurrence_gradients(backend):
    prob = MaxCutProblem.random_3_regular(10, 3).with_random_weights(9, 0.5, 1.0)
    p = 3
    pc = qaoa_parameterized(10, prob.edges, p)
    rng = np.random.default_rng(0)
    ang = qaoa_occurrence_angles(prob.edges, 10, p, rng.uniform(0, math.pi, p), rng.uniform(0, math.pi, p))
    H = maxcut_cost_hamiltonian(prob)
    want = O.pc_adjoint_gradient(pc.to_pops(), pc.parameter_count, 0, ang, [], 10, H.to_terms())
    e, g = G.batch_adjoint_gradient(backend, pc, 10, ang[None, :], H)
    assert np.max(np.abs(g[0] - want)) < 1e-11


def test_run_parameterized_matches_oracle(backend):
    pc, params = seven_qubit_circuit()
    st = G.run_parameterized(backend, pc, params, 7)
    want = O.pc_run(pc.to_pops(), 42, 0, params, [], 7)
    assert np.max(np.abs(st.download().amplitudes - want)) < 1e-12


def test_torch_autograd_vqe_energy(backend):
    import torch

    from quantumforge_amd.torch_ops import vqe_energy

    n = 8
    pc = hea_circuit(n, 2)
    H = heisenberg_hamiltonian(n, seed=8)
    params = torch.tensor(uniform_parameters(pc.parameter_count, 3, seed=8), dtype=torch.float64,
                          requires_grad=True)
    e = vqe_energy(params, backend, pc, n, H)
    w = torch.tensor([1.0, -2.0, 0.5], dtype=torch.float64)
    (e * w).sum().backward()
    t = H.to_terms()
    for s in range(3):
        p = params.detach().numpy()[s]
        assert abs(e[s].item() - O.pc_expectation(pc.to_pops(), pc.parameter_count, 0, p, [], n, t)) < 1e-12
        want = O.pc_shift_gradient(pc.to_pops(), pc.parameter_count, 0, p, [], n, t) * w[s].item()
        assert np.max(np.abs(params.grad.numpy()[s] - want)) < 1e-10


def test_uneven_qaoa_batch_vs_oracle(backend, monkeypatch):
    from quantumforge_amd.workloads import qaoa_occurrence_angles, qaoa_parameterized, qaoa_uneven_instances
    monkeypatch.setenv("QF_QAOA_ENGINE", "0")  # the general engine (the QAOA engine has its own tests)
    specs = qaoa_uneven_instances(7, seed=11, p=2, qubit_choices=(8, 12, 14, 16))
    instances, meta = [], []
    for s in specs:
        prob = s["problem"]
        pc = qaoa_parameterized(prob.num_qubits, prob.edges, s["p"])
        ang = qaoa_occurrence_angles(prob.edges, prob.num_qubits, s["p"], s["gammas"], s["betas"])
        instances.append((pc, prob.num_qubits, ang, maxcut_cost_hamiltonian(prob)))
    e, g, stats = G.uneven_adjoint_gradient(backend, instances)
    assert stats["graph_replays"] == stats["waves"] >= 1
    assert stats["grouped_instances"] + sum(1 for _, q, _, _ in instances if q < stats["tile_qubits"]) == 7
    for k, (pc, q, ang, H) in enumerate(instances):
        t = H.to_terms()
        want_e = O.pc_expectation(pc.to_pops(), pc.parameter_count, 0, ang, [], q, t)
        want_g = O.pc_adjoint_gradient(pc.to_pops(), pc.parameter_count, 0, ang, [], q, t)
        assert abs(e[k] - want_e) < 1e-11, k
        assert np.max(np.abs(g[k] - want_g)) < 1e-10, k


def test_uneven_heterogeneous_circuits(backend):
    # different circuit families and widths in one uneven call
    insts = []
    for n, layers in ((13, 2), (15, 1), (12, 3)):
        pc = hea_circuit(n, layers)
        H = heisenberg_hamiltonian(n, seed=n)
        insts.append((pc, n, uniform_parameters(pc.parameter_count, 1, seed=n)[0], H))
    e, g, stats = G.uneven_adjoint_gradient(backend, insts)
    for k, (pc, n, p, H) in enumerate(insts):
        eb, gb = G.batch_adjoint_gradient(backend, pc, n, p[None, :], H)
        assert abs(e[k] - eb[0]) < 1e-12
        assert np.max(np.abs(g[k] - gb[0])) < 1e-11


def test_uneven_qaoa_multi_wave_at_config_shape(backend, monkeypatch):
    """Config-shape QAOA (p = 6, 3-regular weighted graphs) through several
    memory waves: QF_UNEVEN_WAVE_BYTES caps a wave below two 18-qubit
    instances, so the captured per-wave graph is replayed with new tables over
    reused psi, lambda and partial buffers; a 10-qubit instance narrower than
    the tile (unless the tuner picks a 10-qubit tile) takes the single-instance
    path in the same call.  Energies and
    occurrence gradients against the oracle's store-every-state adjoint
    (adjoint_gradient_with_inputs, src/lib.rs:3564-3704) at 1e-10."""
    from quantumforge_amd.workloads import qaoa_occurrence_angles, qaoa_parameterized, qaoa_uneven_instances

    specs = qaoa_uneven_instances(3, seed=21, p=6, qubit_choices=(18,))
    specs += qaoa_uneven_instances(1, seed=22, p=6, qubit_choices=(16,))
    specs += qaoa_uneven_instances(1, seed=23, p=6, qubit_choices=(10,))
    instances = []
    for s in specs:
        prob = s["problem"]
        pc = qaoa_parameterized(prob.num_qubits, prob.edges, s["p"])
        ang = qaoa_occurrence_angles(prob.edges, prob.num_qubits, s["p"], s["gammas"], s["betas"])
        instances.append((pc, prob.num_qubits, ang, maxcut_cost_hamiltonian(prob)))
    # One 18-qubit instance needs 2 x 2^18 x 16 bytes of psi and lambda plus its partials.
    monkeypatch.setenv("QF_UNEVEN_WAVE_BYTES", str(12 << 20))
    monkeypatch.setenv("QF_QAOA_ENGINE", "0")  # the general engine (the QAOA engine has its own tests)
    e, g, stats = G.uneven_adjoint_gradient(backend, instances)
    print(f"[uneven multi-wave] stats {stats}", flush=True)
    assert stats["waves"] >= 3 and stats["graph_replays"] == stats["waves"]
    assert stats["grouped_instances"] >= 4
    for k, (pc, q, ang, H) in enumerate(instances):
        t = H.to_terms()
        want_e = O.pc_expectation(pc.to_pops(), pc.parameter_count, 0, ang, [], q, t)
        want_g = O.pc_adjoint_gradient(pc.to_pops(), pc.parameter_count, 0, ang, [], q, t)
        assert abs(e[k] - want_e) < 1e-10, (k, e[k], want_e)
        assert np.max(np.abs(g[k] - want_g)) < 1e-10, k


@pytest.mark.parametrize("n,sets", [(9, 5), (12, 5), (14, 9)])
def test_windowed_pauli_tables_on_general_sums(backend, n, sets):
    """The costate kernel's windowed coefficient tables on a Pauli sum with
    X, Y and Z mixed in one word, complex coefficients, several words per x
    mask, z masks wider than one table window (per-term signs) and a constant
    term: energies and adjoint gradients against the oracle (PauliSum
    expectation, src/analytic.rs:2751-2778), at three widths."""
    rng = np.random.default_rng(5 + n)
    H = PauliSum(n)
    H.add(0.7, "I" * n)
    for _ in range(40):
        w = ["I"] * n
        for q in rng.choice(n, size=int(rng.integers(1, 5)), replace=False):
            w[q] = "XYZ"[int(rng.integers(0, 3))]
        H.add(complex(rng.normal(), 0.0), "".join(w))
    # the same x mask with many z patterns, and a z span of 12 bits
    H.add(0.3, "X" + "I" * (n - 2) + "Z")
    H.add(-0.4, "Y" + "Z" * (n - 2) + "Z")
    H.add(0.25, "Z" * n)
    pc = hea_circuit(n, 2)
    params = uniform_parameters(pc.parameter_count, sets, seed=3)
    e, g = G.batch_adjoint_gradient(backend, pc, n, params, H)
    prof = backend.last_gradient_profile()
    assert prof["costate_kernel"] == "windowed"
    t = H.to_terms()
    for s in range(params.shape[0]):
        want_e = O.pc_expectation(pc.to_pops(), pc.parameter_count, 0, params[s], [], n, t)
        want_g = O.pc_adjoint_gradient(pc.to_pops(), pc.parameter_count, 0, params[s], [], n, t)
        assert abs(e[s] - want_e) < 1e-12, s
        assert np.max(np.abs(g[s] - want_g)) < 1e-11, s
