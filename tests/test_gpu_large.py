"""Parity beyond 2^32 amplitudes and for the QFT-like family in complex64, where
the CPU oracle cannot hold the state.

After one QFT-like layer from |0...0> (H on every qubit, exact CP(pi/2^d) on
every pair at distance d <= 8, RZ(theta_q) on every qubit) every amplitude has
the closed form

    psi(i) = 2^(-n/2) exp(i phi(i)),
    phi(i) = sum_{d<=8, t} (pi / 2^d) b_t b_{t+d} + sum_q theta_q (b_q - 1/2),

because H^(x)n spreads |0...0> uniformly and every later gate is diagonal
(CP multiplies by e^{i phi} when both bits are set; RZ(theta) is
diag(e^{-i theta/2}, e^{i theta/2}), Gate::matrix src/lib.rs:2324-2331).  The
tests read >= 10^5 amplitudes at random indices (about half of them >= 2^32 at
33 and 34 qubits) and at the index-width edges on the device (qf_state_gather)
and compare them with the closed form; then the two-layer bench circuit
followed by its inverse must return to |0...0>.  The 30-qubit limit of the
reference GPU path (src/cuda.rs:2256-2259) does not apply here: indices are
64-bit throughout.

Bounds: complex128 amplitudes within 1e-12 relative to 2^(-n/2) (that is
about 1e-17 absolute); complex64 within 5e-5 relative (float32 storage of
each pass's result, about 30 passes' worth of rounding of ~6e-8 each); the
round trip overlap |<0|U^dagger U|0>|^2 >= 1 - 1e-10 (complex128) and
>= 1 - 1e-4 (complex64).

Also: the QFT-like family in complex64 against the oracle at 20 qubits, and
the headline random-circuit family at 26 qubits against the oracle's in-place
OpenMP run (oracle fast_run, the reference's per-pair arithmetic) with the
whole state compared.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from quantumforge_amd import Circuit, Gate
from quantumforge_amd.backend import Plan
from quantumforge_amd.workloads import qft_like_circuit, random_layer_circuit

from .conftest import close

pytestmark = pytest.mark.gpu


def inverse(c: Circuit) -> Circuit:
    """Adjoint circuit: reversed order, one-qubit gates replaced by their
    adjoints (adjoint_gate, src/lib.rs:2370-2382), controlled phases by the
    opposite angle; CNOT, CZ and SWAP are self-inverse."""
    from quantumforge_amd.core import _GATE_NAMES

    ops = c.to_ops()[::-1].copy()
    for r in ops:
        if r["kind"] == 0:
            g = Gate(_GATE_NAMES[int(r["gate"])], float(r["theta"])).adjoint()
            r["gate"], r["theta"] = g.id, g.theta
        elif r["kind"] == 7:
            r["theta"] = -r["theta"]
        else:
            assert r["kind"] in (1, 2, 3)
    return Circuit.from_ops(ops)


def layer_one_phases(n: int, seed: int, idx: np.ndarray) -> np.ndarray:
    """phi(i) of the module docstring for the indices idx (uint64)."""
    _, native = qft_like_circuit(n, 1, seed=seed)
    ops = native.to_ops()
    rz = ops[(ops["kind"] == 0) & (ops["gate"] == Gate.Rz(0.0).id)]
    assert len(rz) == n and list(rz["a"]) == list(range(n))
    theta = rz["theta"].astype(np.float64)
    bits = ((idx[:, None] >> np.arange(n, dtype=np.uint64)[None, :]) & np.uint64(1)).astype(np.float64)
    phi = (bits - 0.5) @ theta
    for d in range(1, 9):
        phi += (math.pi / (1 << d)) * np.sum(bits[:, :n - d] * bits[:, d:], axis=1)
    return phi


def sample_indices(n: int, count: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    edges = [0, 1, (1 << n) - 1, (1 << n) - 2]
    for b in range(n):
        edges += [1 << b, ((1 << n) - 1) ^ (1 << b)]
    if n > 32:
        edges += [(1 << 32) - 1, 1 << 32, (1 << 32) + 1, (1 << 33) - 1]
    r = rng.integers(0, 1 << n, size=count, dtype=np.uint64)
    return np.unique(np.concatenate([np.array(edges, dtype=np.uint64), r]))


@pytest.mark.parametrize("n,precision,rel_tol,trip_tol", [
    (33, "complex128", 1e-12, 1e-10),
    (34, "complex64", 5e-5, 1e-4),
])
def test_qft_like_beyond_2_pow_32_closed_form_and_round_trip(backend, n, precision, rel_tol, trip_tol):
    idx = sample_indices(n, 100_000, seed=n)
    assert np.count_nonzero(idx >= np.uint64(1 << 32)) > 40_000
    print(f"[{n}q {precision}] layer-1 closed form at {idx.size} indices", flush=True)
    _, layer1 = qft_like_circuit(n, 1, seed=33)
    st = backend.zero_state(n, precision=precision)
    plan = Plan(layer1, n, precision=precision)
    st.run_circuit(layer1, plan)
    assert plan.kernels_run() == ["compiled"]
    got = st.gather(idx)
    scale = 2.0 ** (-n / 2)
    want = scale * np.exp(1j * layer_one_phases(n, 33, idx))
    rel = float(np.max(np.abs(got - want))) / scale
    print(f"[{n}q {precision}] max |amp - closed form| / 2^(-n/2) = {rel:.3e}", flush=True)
    assert rel <= rel_tol
    # Layer 1 undone, then the two-layer bench circuit (seed 33) and its inverse.
    inv1 = inverse(layer1)
    st.run_circuit(inv1, Plan(inv1, n, precision=precision))
    assert abs(abs(st.gather([0])[0]) ** 2 - 1) < trip_tol
    _, two = qft_like_circuit(n, 2, seed=33)
    inv = inverse(two)
    st.run_circuit(two, Plan(two, n, precision=precision))
    st.run_circuit(inv, Plan(inv, n, precision=precision))
    fid = abs(st.gather([0])[0]) ** 2
    print(f"[{n}q {precision}] |<0|U^dagger U|0>|^2 = {fid:.15f}", flush=True)
    assert fid >= 1 - trip_tol
    del st


def test_qft_like_complex64_matches_oracle_at_20_qubits(backend):
    n = 20
    _, c = qft_like_circuit(n, 2, seed=3, max_distance=8)
    dev = backend.zero_state(n, precision="complex64")
    plan = Plan(c, n, precision="complex64")
    dev.run_circuit(c, plan)
    assert plan.kernels_run() == ["compiled"]
    st = O.State(n)
    st.run(c.to_ops())
    want = st.amplitudes()
    err = close(dev.download_c64(), want, 2e-6)
    print(f"[20q complex64 QFT-like] max |amp - oracle| = {err:.3e}", flush=True)
    # Same circuit in complex128 to 1e-12.
    d128 = backend.zero_state(n)
    d128.run_circuit(c, Plan(c, n))
    close(d128.download().amplitudes, want, 1e-12)


def test_headline_family_matches_oracle_at_26_qubits(backend):
    n = 26
    c = random_layer_circuit(n, 40, seed=2024)
    plan = Plan(c, n)
    dev = backend.zero_state(n)
    dev.run_circuit(c, plan)
    assert plan.kernels_run() == ["compiled"]
    got = dev.download().amplitudes
    del dev
    print(f"[26q depth 40] running the oracle ({O.max_threads()} threads)", flush=True)
    want = O.fast_run(n, c.to_ops())
    err = float(np.max(np.abs(got - want)))
    fid = abs(np.vdot(want, got)) ** 2
    print(f"[26q depth 40] max |amp - oracle| = {err:.3e}, fidelity 1 - {1 - fid:.3e}", flush=True)
    assert err <= 1e-12
    assert fid >= 1 - 1e-10
