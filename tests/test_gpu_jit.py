"""GPU parity of compiled passes (one generated gfx950 kernel per tile pass)
against the oracle: tile shapes, the default from 18 qubits, complex64,
batches, controlled phases, and the gradient engine's forward passes."""
import numpy as np
import pytest

from oracle import oracle as O
from quantumforge_amd import StateVector
from quantumforge_amd import gradients as Gr
from quantumforge_amd.backend import Plan
from quantumforge_amd.workloads import (heisenberg_hamiltonian, hea_circuit, qft_like_circuit, random_layer_circuit,
                                        uniform_parameters)

from .conftest import close

pytestmark = pytest.mark.gpu


def oracle_run(n, c, a0=None):
    st = O.State(n, a0) if a0 is not None else O.State(n)
    st.run(c.to_ops())
    return st.amplitudes()


def random_state(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return a / np.linalg.norm(a)


@pytest.mark.parametrize("M,RB", [(11, 3), (11, 4), (12, 4), (12, 5), (10, 4), (13, 5)])
def test_compiled_passes_match_oracle(backend, M, RB):
    n = 18
    c = random_layer_circuit(n, 12, seed=M * 10 + RB, pairs_per_layer=9)
    plan = Plan(c, n, kernel=3, tile_qubits=M, reg_bits=RB)
    assert plan.compiled
    a0 = random_state(n, RB)
    dev = backend.upload(StateVector(a0, n))
    dev.run_circuit(c, plan)
    close(dev.download().amplitudes, oracle_run(n, c, a0), 1e-12)


def test_default_plan_runs_compiled_from_18_qubits(backend):
    n = 20
    c = random_layer_circuit(n, 20, seed=99)
    plan = Plan(c, n)
    assert plan.describe()["compiled"] == 1
    dev = backend.zero_state(n)
    dev.run_circuit(c)  # no plan: the backend's default options compile too
    close(dev.download().amplitudes, oracle_run(n, c), 1e-12)
    dev2 = backend.zero_state(n)
    dev2.run_circuit(c, plan)
    close(dev2.download().amplitudes, oracle_run(n, c), 1e-12)


def test_compiled_complex64(backend):
    n = 18
    c = random_layer_circuit(n, 10, seed=5, pairs_per_layer=9)
    dev = backend.zero_state(n, precision="complex64")
    dev.run_circuit(c, Plan(c, n, kernel=3, tile_qubits=12, reg_bits=4))
    close(dev.download_c64(), oracle_run(n, c), 2e-6)


def test_compiled_batch_of_states(backend):
    n = 18
    c = random_layer_circuit(n, 8, seed=17, pairs_per_layer=9)
    states = [StateVector(random_state(n, s), n) for s in range(3)]
    batch = backend.upload_batch(states)
    batch.run_circuit(c, Plan(c, n, kernel=3))
    for s, out in zip(states, batch.download()):
        close(out.amplitudes, oracle_run(n, c, s.amplitudes), 1e-12)


def test_compiled_controlled_phases(backend):
    n = 20
    _, c = qft_like_circuit(n, 2, seed=3, max_distance=8)
    dev = backend.zero_state(n)
    dev.run_circuit(c, Plan(c, n, kernel=3))
    close(dev.download().amplitudes, oracle_run(n, c), 1e-12)


def test_gradient_forward_passes_run_compiled(backend):
    # 18 qubits: the forward plan of the adjoint engine compiles (rotation
    # angles from the per-instance table); the reverse sweep keeps the
    # register kernel.
    n = 18
    pc = hea_circuit(n, 1)
    H = heisenberg_hamiltonian(n, seed=3)
    assert Plan.parameterized(pc, n).describe()["compiled"] == 1
    params = uniform_parameters(pc.parameter_count, 2, seed=12)
    e, g = Gr.batch_adjoint_gradient(backend, pc, n, params, H)
    for i in range(2):
        ge = O.pc_adjoint_gradient(pc.to_pops(), pc.parameter_count, 0, params[i], [], n, H.to_terms())
        ee = O.pc_expectation(pc.to_pops(), pc.parameter_count, 0, params[i], [], n, H.to_terms())
        assert abs(e[i] - ee) < 1e-10
        close(g[i], ge, 1e-10)
