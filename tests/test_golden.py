"""The committed golden fixtures (tests/golden/golden.npz, written by
tests/golden/make_golden.py from the oracle and the dense brute force in
agreement) against the oracle now (CPU) and against the HIP engine (GPU)."""
import numpy as np
import pytest

from oracle import oracle as O
from quantumforge_amd.workloads import SmallRng

from .conftest import ROOT, close

G = np.load(f"{ROOT}/tests/golden/golden.npz", allow_pickle=False)
RANDOM = [k[: -len("_ops")] for k in G.files if k.startswith("random_") and k.endswith("_ops")]


def nq(amps):
    return int(np.log2(len(amps)))


@pytest.mark.parametrize("key", RANDOM)
def test_oracle_reproduces_random_circuit_fixtures(key):
    amps = G[key + "_amps"]
    assert np.array_equal(O.run_circuit(nq(amps), G[key + "_ops"]), amps)


def test_oracle_reproduces_other_fixtures():
    e5 = np.zeros(64, complex)
    e5[5] = 1
    s = O.State(6, e5)
    s.run(G["qft6_basis5_ops"])
    assert np.array_equal(s.amplitudes(), G["qft6_basis5_amps"])
    g = O.pc_shift_gradient(G["grad7_pops"], 42, 0, G["grad7_params"], [], 7,
                            np.array([(1.0, 0.0, 0, 1)], dtype=O.PAULI_DTYPE))
    close(g, G["grad7_gradient"], 1e-15)
    npar = len(G["hea6_params"])
    e = O.pc_expectation(G["hea6_pops"], npar, 0, G["hea6_params"], [], 6, G["hea6_terms"])
    assert abs(e - G["hea6_energy"][0]) < 1e-15
    r = SmallRng(2024)
    assert [r.next_u64() for _ in range(16)] == [int(x) for x in G["rng2024_u64"]]


@pytest.mark.gpu
@pytest.mark.parametrize("key", RANDOM)
def test_engine_matches_random_circuit_fixtures(backend, key):
    from quantumforge_amd import Circuit, StateVector

    amps = G[key + "_amps"]
    n = nq(amps)
    c = Circuit.from_ops(G[key + "_ops"])
    backend.set_plan_options(fuse=False)
    assert np.array_equal(backend.run_circuit(StateVector.zero(n), c).amplitudes, amps)
    backend.set_plan_options(fuse=True)
    close(backend.run_circuit(StateVector.zero(n), c).amplitudes, amps, 1e-12)


@pytest.mark.gpu
def test_engine_matches_gradient_fixtures(backend):
    from quantumforge_amd import gradients as Gr
    from quantumforge_amd.core import ParameterizedCircuit, PauliSum

    pops = G["hea6_pops"]
    pc = ParameterizedCircuit(len(G["hea6_params"]))
    pc._ops = [(int(r["kind"]), int(r["gate"]), int(r["a"]), int(r["b"]), int(r["index"]), float(r["theta"]))
               for r in pops]
    H = PauliSum(6)
    for r in G["hea6_terms"]:
        w = "".join("Y" if (r["x_mask"] >> q) & (r["z_mask"] >> q) & 1 else "X" if (r["x_mask"] >> q) & 1
                    else "Z" if (r["z_mask"] >> q) & 1 else "I" for q in range(6))
        H.add(complex(r["re"], r["im"]), w)
    e, g = Gr.batch_adjoint_gradient(backend, pc, 6, G["hea6_params"][None, :], H)
    assert abs(e[0] - G["hea6_energy"][0]) < 1e-12
    close(g[0], G["hea6_gradient"], 1e-10)
