"""Full-size parity where the CPU oracle cannot follow: the fused, compiled
plan against the exact lowering (one sweep per gate in the reference's
arithmetic order, bit-identical to the oracle at 20 qubits, see
test_gpu_state.py) on the 30-qubit depth-40 random circuit, as a fidelity
|<exact|fused>|^2; complex64 against complex128; and an unitarity round trip
U^dagger U |0> on the same workload."""
import numpy as np
import pytest

from quantumforge_amd import Circuit, Gate
from quantumforge_amd.backend import Plan
from quantumforge_amd.workloads import random_layer_circuit

pytestmark = pytest.mark.gpu


def inverse(c: Circuit) -> Circuit:
    """Adjoint circuit: reversed order, each one-qubit gate replaced by its
    adjoint (adjoint_gate, src/lib.rs:2370-2382); CNOT, CZ and SWAP are
    self-inverse."""
    from quantumforge_amd.core import _GATE_NAMES

    ops = c.to_ops()[::-1].copy()
    for r in ops:
        if r["kind"] == 0:
            g = Gate(_GATE_NAMES[int(r["gate"])], float(r["theta"])).adjoint()
            r["gate"], r["theta"] = g.id, g.theta
        else:
            assert r["kind"] in (1, 2, 3)
    return Circuit.from_ops(ops)


def test_fused_compiled_plan_matches_exact_lowering_at_30_qubits(backend):
    n = 30
    c = random_layer_circuit(n, 40, seed=2024)
    exact = backend.zero_state(n)
    exact.run_circuit(c, Plan(c, n, fuse=False))
    fused = backend.zero_state(n)
    plan = Plan(c, n)
    assert plan.describe()["compiled"] == 1
    fused.run_circuit(c, plan)
    ov = exact.inner(fused)
    fidelity = abs(ov) ** 2
    assert abs(exact.norm_sqr() - 1) < 1e-12 and abs(fused.norm_sqr() - 1) < 1e-12
    assert fidelity >= 1 - 1e-10, fidelity
    # Not just a global phase: the overlap itself is 1.
    assert abs(ov - 1) < 1e-10, ov


def test_round_trip_returns_to_zero_state_at_30_qubits(backend):
    n = 30
    c = random_layer_circuit(n, 40, seed=7)
    st = backend.zero_state(n)
    st.run_circuit(c, Plan(c, n))
    inv = inverse(c)
    st.run_circuit(inv, Plan(inv, n))
    zero = backend.zero_state(n)
    assert abs(abs(zero.inner(st)) ** 2 - 1) < 1e-10


def test_complex64_fidelity_against_complex128_at_26_qubits(backend):
    n = 26
    c = random_layer_circuit(n, 40, seed=2024)
    a = backend.zero_state(n)
    a.run_circuit(c, Plan(c, n))
    b = backend.zero_state(n, precision="complex64")
    b.run_circuit(c, Plan(c, n, precision="complex64"))
    x = a.download().amplitudes
    y = b.download_c64().astype(np.complex128)
    f = abs(np.vdot(x, y)) ** 2 / (np.vdot(y, y).real)
    assert f >= 1 - 1e-6, f
