"""The QAOA tile engine of the uneven batch (instances in QAOA form: Hadamard
layer, cost rotations CNOT(u, v) Rz(v) CNOT(u, v), an X rotation on every
qubit, a diagonal observable) against the oracle's store-every-state adjoint
(adjoint_gradient_with_inputs, src/lib.rs:3564-3704) and its expectation
(PauliSum::expectation_ket, src/analytic.rs:2775-2778): energies at 1e-11,
every occurrence gradient at 1e-10 (complex128).  Cases: both tile-set
schedules (two sets with a cost layer every pass, three sets with one every
second pass), graphs whose edges fall on register, lane, wave and tile-index
bits in every pass, observables with single Z words and a constant, several
memory waves mixed with an instance of another form, the energy-only call and
complex64."""
import numpy as np
import pytest

from oracle import oracle as O
from quantumforge_amd import gradients as G
from quantumforge_amd.workloads import (MaxCutProblem, SmallRng, hea_circuit, heisenberg_hamiltonian,
                                        maxcut_cost_hamiltonian, qaoa_occurrence_angles, qaoa_parameterized,
                                        uniform_parameters)

pytestmark = pytest.mark.gpu


def qaoa_instance(n, p, seed, extra_edges=0, with_fields=False):
    """A weighted graph on n qubits (a 3-regular graph for even n, a ring
    otherwise, plus extra_edges random chords), p layers, angles from
    SmallRng(seed); observable: its weighted Z Z words, with_fields adds a Z
    word per qubit and a constant."""
    rng = SmallRng(seed)
    if n % 2 == 0:
        prob = MaxCutProblem.random_3_regular(n, seed).with_random_weights(seed + 1, 0.5, 1.0)
        edges = list(prob.edges)
    else:
        edges = [(i, (i + 1) % n, 0.5 + 0.5 * rng.next_f64()) for i in range(n)]
        edges = [(min(u, v), max(u, v), w) for u, v, w in edges]
    have = {(u, v) for u, v, _ in edges}
    while extra_edges:
        u, v = rng.next_u64() % n, rng.next_u64() % n
        if u == v or (min(u, v), max(u, v)) in have:
            continue
        have.add((min(u, v), max(u, v)))
        edges.append((min(u, v), max(u, v), 0.5 + 0.5 * rng.next_f64()))
        extra_edges -= 1
    gam = [np.pi * rng.next_f64() for _ in range(p)]
    bet = [np.pi * rng.next_f64() for _ in range(p)]
    pc = qaoa_parameterized(n, edges, p)
    ang = qaoa_occurrence_angles(edges, n, p, gam, bet)
    H = maxcut_cost_hamiltonian(MaxCutProblem(n, edges))
    if with_fields:
        for q in range(n):
            word = ["I"] * n
            word[q] = "Z"
            H.add(0.25 + 0.5 * rng.next_f64(), "".join(word))
        H.add(0.75, "I" * n)
    return pc, n, ang, H


def check_vs_oracle(instances, energies, grads, tol_e=1e-11, tol_g=1e-10):
    for k, (pc, n, ang, H) in enumerate(instances):
        t = H.to_terms()
        want_e = O.pc_expectation(pc.to_pops(), pc.parameter_count, 0, ang, [], n, t)
        assert abs(energies[k] - want_e) < tol_e, (k, energies[k], want_e)
        if grads is not None:
            want_g = O.pc_adjoint_gradient(pc.to_pops(), pc.parameter_count, 0, ang, [], n, t)
            err = np.max(np.abs(grads[k] - want_g))
            print(f"[qaoa engine] instance {k}: {n} qubits, {pc.parameter_count} occurrences, "
                  f"max |grad - oracle| = {err:.2e}", flush=True)
            assert err < tol_g, (k, err)


def test_two_tile_sets_vs_oracle(backend):
    insts = [qaoa_instance(12, 2, 1), qaoa_instance(14, 3, 2, extra_edges=3, with_fields=True),
             qaoa_instance(13, 2, 3, extra_edges=6), qaoa_instance(16, 2, 4, with_fields=True)]
    e, g, stats = G.uneven_adjoint_gradient(backend, insts)
    print(f"[qaoa engine] stats {stats}", flush=True)
    assert stats["qaoa_instances"] == len(insts)
    assert stats["grouped_instances"] == 0
    check_vs_oracle(insts, e, g)


def test_three_tile_sets_vs_oracle(backend, monkeypatch):
    monkeypatch.setenv("QF_QAOA_LAYER_PASSES", "2")
    insts = [qaoa_instance(14, 3, 5, with_fields=True), qaoa_instance(15, 2, 6, extra_edges=4),
             qaoa_instance(16, 1, 7)]
    e, g, stats = G.uneven_adjoint_gradient(backend, insts)
    print(f"[qaoa engine] stats {stats}", flush=True)
    assert stats["qaoa_instances"] == len(insts)
    # a cost layer every second pass (p = 3): at most 2p + 1 passes, fewer
    # when a narrow instance's tile sets repeat qubits
    assert 2 * 3 - 1 <= stats["qaoa_forward_passes"] <= 2 * 3 + 1
    check_vs_oracle(insts, e, g)


def test_multi_wave_with_another_form(backend, monkeypatch):
    """A wave cap below two 14-qubit instances: several waves replay the
    captured graph over rewritten tables; an HEA instance in the same call
    runs on the general engine."""
    insts = [qaoa_instance(14, 2, 8), qaoa_instance(14, 2, 9, with_fields=True), qaoa_instance(12, 3, 10),
             qaoa_instance(13, 1, 11, extra_edges=2)]
    pc = hea_circuit(12, 2)
    insts.append((pc, 12, uniform_parameters(pc.parameter_count, 1, seed=12)[0], heisenberg_hamiltonian(12, seed=12)))
    monkeypatch.setenv("QF_UNEVEN_WAVE_BYTES", "600000")
    e, g, stats = G.uneven_adjoint_gradient(backend, insts)
    print(f"[qaoa engine] stats {stats}", flush=True)
    assert stats["qaoa_instances"] == 4 and stats["waves"] >= 4
    check_vs_oracle(insts, e, g)


def test_energy_only(backend):
    insts = [qaoa_instance(12, 2, 13, with_fields=True), qaoa_instance(15, 2, 14, extra_edges=3)]
    e, stats = G.uneven_expectation(backend, insts)
    assert stats["qaoa_instances"] == 2
    check_vs_oracle(insts, e, None)


def test_complex64(backend):
    insts = [qaoa_instance(14, 2, 15, with_fields=True), qaoa_instance(12, 3, 16)]
    e, g, stats = G.uneven_adjoint_gradient(backend, insts, precision="complex64")
    assert stats["qaoa_instances"] == 2
    check_vs_oracle(insts, e, g, tol_e=2e-5, tol_g=5e-5)


def test_against_general_engine_at_22_qubits(backend, monkeypatch):
    """22 qubits (three tile sets by width): the QAOA engine against the
    general engine on the same call (which the oracle pins at smaller widths),
    and the energy against the oracle."""
    insts = [qaoa_instance(22, 2, 17)]
    e, g, stats = G.uneven_adjoint_gradient(backend, insts)
    assert stats["qaoa_instances"] == 1 and stats["qaoa_forward_passes"] == 5
    monkeypatch.setenv("QF_QAOA_ENGINE", "0")
    e2, g2, stats2 = G.uneven_adjoint_gradient(backend, insts)
    assert stats2["qaoa_instances"] == 0
    assert abs(e[0] - e2[0]) < 1e-11
    assert np.max(np.abs(g[0] - g2[0])) < 1e-10
    check_vs_oracle(insts, e, None)
