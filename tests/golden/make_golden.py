"""Writes the golden fixtures under tests/golden/ (small inputs and expected outputs).

Each fixture is computed twice, by the C++ oracle (oracle/qf_oracle.cpp) and by
the independent dense NumPy brute force (oracle/np_oracle.py), and written only
if the two agree to 1e-13.  Inputs are seeded (SmallRng), so the fixture files
pin the oracle against silent drift.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import np_oracle as NP  # noqa: E402
from oracle import oracle as O  # noqa: E402
from quantumforge_amd import ParameterizedCircuit, PauliSum, RotationAxis, SingleQubitObservable  # noqa: E402
from quantumforge_amd.workloads import (MaxCutProblem, hea_circuit, heisenberg_hamiltonian,  # noqa: E402
                                        maxcut_cost_hamiltonian, qaoa_maxcut, qft_circuit, random_layer_circuit,
                                        uniform_parameters)

TOL = 1e-13


def seven_qubit():
    pc = ParameterizedCircuit(42)
    params = []
    for layer in range(3):
        for q in range(7):
            pc.rotation(RotationAxis.Y, q, len(params))
            params.append(0.3 + 0.1 * layer)
        for q in range(7):
            pc.rotation(RotationAxis.X, q, len(params))
            params.append(-0.2 + 0.05 * q)
        for q in range(6):
            pc.cnot(q, q + 1)
    return pc, np.array(params)


def main() -> None:
    out = {}
    # 1. random layered circuits (the depth-40 family at small widths)
    for n, depth, seed in [(6, 8, 1), (8, 12, 2), (10, 6, 3)]:
        c = random_layer_circuit(n, depth, seed=seed, pairs_per_layer=n // 2)
        a = O.run_circuit(n, c.to_ops())
        b = NP.run(c.to_ops(), n)
        assert np.max(np.abs(a - b)) < TOL
        out[f"random_n{n}_d{depth}_s{seed}_ops"] = c.to_ops()
        out[f"random_n{n}_d{depth}_s{seed}_amps"] = a
    # 2. QFT on |5>
    c = qft_circuit(6)
    e5 = np.zeros(64, complex)
    e5[5] = 1
    s = O.State(6, e5)
    s.run(c.to_ops())
    a = s.amplitudes()
    assert np.max(np.abs(a - NP.run(c.to_ops(), 6, e5))) < TOL
    out["qft6_basis5_ops"] = c.to_ops()
    out["qft6_basis5_amps"] = a
    # 3. the 7-qubit, 42-parameter gradient of tests/gradients.rs:51-110
    pc, params = seven_qubit()
    t = PauliSum.single(SingleQubitObservable.Z, 0, 7).to_terms()
    g = O.pc_shift_gradient(pc.to_pops(), 42, 0, params, [], 7, t)
    gd = NP.pc_shift_gradient(pc.to_pops(), 42, params, 7, NP.pauli_sum_matrix([(1.0, "Z" + "I" * 6)]))
    assert np.max(np.abs(g - gd)) < 1e-12
    out["grad7_pops"] = pc.to_pops()
    out["grad7_params"] = params
    out["grad7_gradient"] = g
    # 4. Heisenberg HEA energy and gradient, 6 qubits
    pc = hea_circuit(6, 2)
    H = heisenberg_hamiltonian(6, seed=6)
    p = uniform_parameters(pc.parameter_count, 1, seed=60)[0]
    Hd = NP.pauli_sum_matrix(H.terms)
    e = O.pc_expectation(pc.to_pops(), pc.parameter_count, 0, p, [], 6, H.to_terms())
    g = O.pc_adjoint_gradient(pc.to_pops(), pc.parameter_count, 0, p, [], 6, H.to_terms())
    assert abs(e - NP.pc_energy(pc.to_pops(), p, 6, Hd)) < TOL
    assert np.max(np.abs(g - NP.pc_shift_gradient(pc.to_pops(), pc.parameter_count, p, 6, Hd))) < 1e-12
    out["hea6_pops"] = pc.to_pops()
    out["hea6_terms"] = H.to_terms()
    out["hea6_params"] = p
    out["hea6_energy"] = np.array([e])
    out["hea6_gradient"] = g
    # 5. QAOA 8q on random_3_regular(8, 7) with the angles of tests/mps_qaoa.rs:37-50
    prob = MaxCutProblem.random_3_regular(8, 7)
    c = qaoa_maxcut(8, prob.edges, 2, [0.31, -0.17], [0.42, 0.28])
    st = O.State(8)
    st.run(c.to_ops())
    probs = st.probabilities()
    cut = sum(pr * prob.cut_value(x) for x, pr in enumerate(probs))
    out["qaoa8_edges"] = np.array([(u, v) for u, v, _ in prob.edges], dtype=np.int64)
    out["qaoa8_cut_expectation"] = np.array([cut])
    out["qaoa8_amps"] = st.amplitudes()
    assert np.max(np.abs(st.amplitudes() - NP.run(c.to_ops(), 8))) < TOL
    # 6. SmallRng streams
    from quantumforge_amd.workloads import SmallRng
    r = SmallRng(2024)
    out["rng2024_u64"] = np.array([r.next_u64() for _ in range(16)], dtype=np.uint64)

    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)
    meta = {k: list(v.shape) for k, v in out.items()}
    with open(os.path.join(HERE, "golden_manifest.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print(f"wrote {len(out)} arrays")


if __name__ == "__main__":
    main()
