"""Density-matrix circuits as tile plans over the 2n-qubit view of the
row-major entries (qf_density_plan_build), on the CPU: the plans replayed in
NumPy and the generated compiled kernels run under the g++ emulation, each
against the oracle's DensityMatrix restatement (src/lib.rs:1553-2227)."""
import numpy as np
import pytest

from oracle import oracle as O
from quantumforge_amd import Circuit, Gate, NoiseChannel, QuantumError
from quantumforge_amd.backend import Plan

from .conftest import close
from .test_abi import check_tile_invariants, replay
from .test_jit import emulate
from .test_oracle_density import random_noisy_circuit, random_state


def oracle_density(c: Circuit, n: int, psi: np.ndarray) -> np.ndarray:
    d = O.Density(n, state=psi)
    d.run(c.to_ops())
    return d.entries()


def noisy_layers(n: int, layers: int, seed: int) -> Circuit:
    """Layers of rotations and CNOTs, each followed by every kind of channel."""
    rng = np.random.default_rng(seed)
    chans = [NoiseChannel("Depolarizing", (0.05,)), NoiseChannel("AmplitudeDamping", (0.08,)),
             NoiseChannel("PhaseDamping", (0.11,)), NoiseChannel("ThermalDamping", (0.07, 0.9))]
    c = Circuit()
    for _ in range(layers):
        for q in range(n):
            c.push([Gate.Rx, Gate.Ry, Gate.Rz][rng.integers(3)](float(rng.uniform(-3, 3))), q)
        for q in range(0, n - 1, 2):
            c.cnot(q, q + 1)
        c.cz(n - 1, 0)
        for q in range(n):
            c.channel(q, chans[(q + _) % 4])
    return c


@pytest.mark.parametrize("n,kernel,tile,reg", [(3, 1, 0, 0), (4, 1, 6, 0), (5, 1, 8, 0), (6, 2, 10, 3),
                                               (6, 2, 12, 4), (5, 2, 9, 3)])
def test_density_plan_replay_equals_oracle(n, kernel, tile, reg):
    c = random_noisy_circuit(n, 5, seed=n * 7 + tile)
    c.swap(0, n - 1)
    c.cp(1, n - 1, 0.61)
    plan = Plan.density(c, n, tile_qubits=tile, kernel=kernel, reg_bits=reg)
    desc = plan.describe()
    check_tile_invariants(desc)
    assert desc["qubits"] == 2 * n
    psi = random_state(n, n)
    rho0 = np.outer(psi, psi.conj()).reshape(-1)
    close(replay(desc, rho0), oracle_density(c, n, psi), 1e-13)


@pytest.mark.parametrize("n,M,RB", [(5, 9, 3), (6, 10, 4), (6, 12, 5)])
def test_compiled_density_passes_match_oracle(n, M, RB):
    c = noisy_layers(n, 3, seed=n + M)
    plan = Plan.density(c, n, kernel=3, tile_qubits=M, reg_bits=RB)
    assert plan.compiled
    src = plan.source()
    assert "v3r" in src  # the superoperators are dense 4x4 ops in the generated code
    psi = random_state(n, M)
    rho0 = np.outer(psi, psi.conj()).reshape(-1)
    close(emulate(plan, rho0, 2 * n), oracle_density(c, n, psi), 1e-13)


def test_noisy_layer_workload_plan_matches_oracle():
    from quantumforge_amd.workloads import noisy_layer_circuit

    n = 5
    c = noisy_layer_circuit(n, 2, seed=7, strength=0.2)
    assert sum(1 for op in c.to_ops() if op["kind"] == 4) == 2 * n
    psi = random_state(n, 1)
    rho0 = np.outer(psi, psi.conj()).reshape(-1)
    close(replay(Plan.density(c, n, kernel=2, tile_qubits=10, reg_bits=3).describe(), rho0),
          oracle_density(c, n, psi), 1e-13)


def test_density_default_plan_is_compiled_from_nine_qubits():
    c = noisy_layers(9, 1, seed=1)
    assert Plan.density(c, 9).compiled
    assert not Plan.density(noisy_layers(8, 1, seed=1), 8).compiled


def test_density_plan_errors():
    c = Circuit()
    c.measure(0, 0)
    with pytest.raises(QuantumError) as e:
        Plan.density(c, 2)
    assert e.value.kind == "OperationRequiresShotExecution"
    for ch, kind in ((NoiseChannel("AmplitudeDamping", (1.5,)), "InvalidProbability"),
                     (NoiseChannel("ThermalDamping", (0.2, -1.0)), "InvalidProbability")):
        c = Circuit()
        c.channel(0, ch)
        with pytest.raises(QuantumError) as e:
            Plan.density(c, 2)
        assert e.value.kind == kind
    c = Circuit()
    c.push(Gate.H, 3)
    with pytest.raises(QuantumError) as e:
        Plan.density(c, 3)
    assert e.value.kind == "QubitOutOfRange"
    c = Circuit()
    c.cnot(1, 1)
    with pytest.raises(QuantumError) as e:
        Plan.density(c, 3)
    assert e.value.kind == "InvalidQubitPair"
    for n, kind in ((0, "ZeroQubits"), (16, "StateTooLarge")):
        with pytest.raises(QuantumError) as e:
            Plan.density(Circuit(), n)
        assert e.value.kind == kind


def test_channel_superoperators_are_the_kraus_sums():
    # Each channel alone on a one-qubit density: the plan's 4x4 op applied to
    # the four basis matrices |r><c| reproduces the oracle's Kraus sum.
    for ch in (NoiseChannel("Depolarizing", (0.3,)), NoiseChannel("AmplitudeDamping", (0.4,)),
               NoiseChannel("PhaseDamping", (0.2,)), NoiseChannel("ThermalDamping", (0.35, 0.5))):
        c = Circuit()
        c.channel(0, ch)
        desc = Plan.density(c, 1).describe()
        for k in range(4):
            rho = np.zeros(4, complex)
            rho[k] = 1
            d = O.Density(1, entries=rho)
            d.run(c.to_ops())
            close(replay(desc, rho), d.entries(), 1e-15)
