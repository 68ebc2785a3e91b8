"""Compiled passes on the CPU: the generated kernel source, run under a g++
emulation of the HIP built-ins (tests/jit_emu.py), against the oracle; the
plan flags and errors of the pass compiler; hiprtc compiling a plan for gfx950
without a device."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from quantumforge_amd import Gate, ParameterizedCircuit, QuantumError, RotationAxis
from quantumforge_amd.backend import Plan
from oracle import np_oracle as NP
from quantumforge_amd.workloads import (MaxCutProblem, hea_circuit, heisenberg_hamiltonian, qaoa_parameterized,
                                        qft_like_circuit, random_layer_circuit, uniform_parameters)

from .conftest import ROOT, close
from .jit_emu import EmulatedProgram
from .test_abi import check_tile_invariants, replay

EMU_DIR = os.path.join(ROOT, "tests", ".emu_cache")


def emulate(plan: Plan, amps: np.ndarray, qubits: int, count: int = 1, cs=None, nslots=0, precision="complex128"):
    os.makedirs(EMU_DIR, exist_ok=True)
    tb = [p["tile_bits"] for p in plan.describe()["passes"] if p["kind"] == "tile"]
    emu = EmulatedProgram(plan.source(precision), EMU_DIR)
    return emu.run(amps, qubits, tb, count=count, cs=cs, nslots=nslots)


def random_state(n, seed, count=1):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(count, 1 << n)) + 1j * rng.normal(size=(count, 1 << n))
    return a / np.linalg.norm(a, axis=1, keepdims=True)


# (qubits, depth, seed, tile_qubits, reg_bits): register sets of 2..5 bits,
# 64..512 threads, tiles from 8 qubits to the whole state.
SHAPES = [(12, 10, 1, 9, 3), (13, 12, 2, 10, 4), (12, 8, 5, 8, 2), (11, 12, 7, 11, 5), (13, 8, 11, 12, 3),
          (14, 10, 12, 11, 3)]


@pytest.mark.parametrize("n,depth,seed,M,RB", SHAPES)
def test_generated_kernels_match_oracle_on_random_circuits(n, depth, seed, M, RB):
    c = random_layer_circuit(n, depth, seed=seed, pairs_per_layer=n // 2)
    plan = Plan(c, n, kernel=3, tile_qubits=M, reg_bits=RB)
    assert plan.compiled
    a0 = random_state(n, seed)[0]
    st = O.State(n, a0)
    st.run(c.to_ops())
    close(emulate(plan, a0, n), st.amplitudes(), 1e-13)


def test_half_buffer_exchanges_are_generated_and_exact():
    # Tiles of 11+ qubits with more than 64 threads exchange through half the
    # tile in two rounds; the emulated kernels must still equal the oracle,
    # and switching the half buffer off must give the same amplitudes.
    n = 14
    c = random_layer_circuit(n, 12, seed=9, pairs_per_layer=7)
    plan = Plan(c, n, kernel=3, tile_qubits=12, reg_bits=3)
    src = plan.source()
    assert "const unsigned h = " in src and "xb[2048]" in src
    a0 = random_state(n, 9)[0]
    st = O.State(n, a0)
    st.run(c.to_ops())
    close(emulate(plan, a0, n), st.amplitudes(), 1e-13)


@pytest.mark.parametrize("n,depth,seed,M,RB", [(12, 10, 1, 9, 3), (13, 12, 2, 10, 4), (12, 8, 5, 8, 2),
                                               (14, 10, 12, 11, 3), (16, 14, 3, 12, 3)])
def test_relabeled_passes_match_oracle(n, depth, seed, M, RB):
    # Compiled passes hand the contiguous low index bits to qubits with
    # pending gates: a pass stores its tile with some qubits on other index
    # bits (the "store" permutation) and the plan's last pass puts every
    # qubit back on its own bit.  The NumPy replay of the described plan and
    # the emulated kernels (store exchange included) both equal the oracle.
    c = random_layer_circuit(n, depth, seed=seed, pairs_per_layer=n // 2)
    plan = Plan(c, n, kernel=3, tile_qubits=M, reg_bits=RB)
    desc = plan.describe()
    assert desc["relabeled"] > 0
    check_tile_invariants(desc)
    for p in desc["passes"]:
        if p["store"]:
            tile = list(range(p["low"])) + p["high"]
            assert sorted(p["store"]) == sorted(tile)
            assert set(range(p["low"])) <= set(p["store"])
    a0 = random_state(n, seed)[0]
    st = O.State(n, a0)
    st.run(c.to_ops())
    close(replay(desc, a0), st.amplitudes(), 1e-13)
    close(emulate(plan, a0, n), st.amplitudes(), 1e-13)


def test_relabeling_saves_passes_and_can_be_switched_off(monkeypatch):
    # The 30-qubit, depth-40 random layer circuit (seed 2024) at the default
    # compiled shape: relabeling the low bits saves passes; QF_PLAN_RELABEL=0
    # keeps the identity layout.
    c = random_layer_circuit(30, 40, seed=2024)
    on = Plan(c, 30).describe()
    monkeypatch.setenv("QF_PLAN_RELABEL", "0")
    off = Plan(c, 30).describe()
    assert on["relabeled"] > 0 and off["relabeled"] == 0
    assert all(not p["store"] for p in off["passes"])
    assert len(on["passes"]) < len(off["passes"])
    monkeypatch.delenv("QF_PLAN_RELABEL")
    src = Plan(c, 30).source()
    assert src.count("xb[2048]") == src.count("__global__") == len(on["passes"])


def test_reverse_plans_relabel_only_when_compiled():
    # A reverse plan that runs through the interpreting adjoint kernel keeps
    # every qubit on its own bit; one that runs compiled (its adjoint form)
    # may relabel like any compiled plan.
    pc = hea_circuit(20, 3)
    assert Plan.parameterized(pc, 20, reverse=True, kernel=2).describe()["relabeled"] == 0
    assert Plan.parameterized(pc, 20, reverse=True).describe()["relabeled"] > 0


def test_exchanges_without_a_common_thread_bit_hop_through_an_intermediate_layout():
    # With 11-qubit tiles (8 thread bits) some consecutive phases share no
    # thread bit; such an exchange takes two half-buffer rounds through a
    # layout that shares one with each side.
    n = 14
    c = random_layer_circuit(n, 12, seed=8, pairs_per_layer=7)
    plan = Plan(c, n, kernel=3, tile_qubits=11, reg_bits=3)
    d = plan.describe()
    exchanges = d["stats"]["phases"] - len(d["passes"]) + d["relabeled"]
    src = plan.source()
    assert src.count("const unsigned h = ") > exchanges
    assert "xb[1024]" in src and "xb[2048]" not in src
    a0 = random_state(n, 8)[0]
    st = O.State(n, a0)
    st.run(c.to_ops())
    close(emulate(plan, a0, n), st.amplitudes(), 1e-13)


def test_generated_kernels_on_a_batch_and_controlled_phases():
    # Two instances (the instance index lands above the state's qubits) of a
    # QFT-like circuit: controlled phases become diagonal ops on any qubit.
    n = 12
    _, c = qft_like_circuit(n, 2, seed=4, max_distance=6)
    plan = Plan(c, n, kernel=3, tile_qubits=9, reg_bits=3)
    a0 = random_state(n, 8, count=2)
    out = emulate(plan, a0.ravel(), n, count=2).reshape(2, -1)
    for i in range(2):
        st = O.State(n, a0[i])
        st.run(c.to_ops())
        close(out[i], st.amplitudes(), 1e-13)


def test_generated_kernels_general_two_qubit_diagonals():
    # Rz on both qubits next to a controlled phase merge into one diagonal
    # with four distinct entries; on qubits outside the register set it is
    # applied as a product of single-select factors.
    import math

    from quantumforge_amd import Circuit, Gate

    n = 12
    c = Circuit()
    rng = np.random.default_rng(5)
    for q in range(n):
        c.push(Gate.H, q)
    for layer in range(3):
        for a in range(n):
            b = (a + 1 + 2 * layer) % n
            c.push(Gate.Rz(float(rng.uniform(0, 2 * math.pi))), a)
            c.push(Gate.Rz(float(rng.uniform(0, 2 * math.pi))), b)
            c.cp(a, b, float(rng.uniform(0, 2 * math.pi)))
        for q in range(0, n, 3):
            c.push(Gate.Rx(0.3 + q), q)
    plan = Plan(c, n, kernel=3, tile_qubits=9, reg_bits=3, low_qubits=2)
    kinds = [o["kind"] for p in plan.describe()["passes"] for o in p["ops"]]
    assert kinds.count(1) > 0
    a0 = random_state(n, 12)[0]
    st = O.State(n, a0)
    st.run(c.to_ops())
    close(emulate(plan, a0, n), st.amplitudes(), 1e-13)


def test_generated_kernels_complex64():
    n = 12
    c = random_layer_circuit(n, 8, seed=11, pairs_per_layer=6)
    plan = Plan(c, n, kernel=3, tile_qubits=10, reg_bits=4)
    a0 = random_state(n, 3)[0]
    st = O.State(n, a0)
    st.run(c.to_ops())
    out = emulate(plan, a0.astype(np.complex64), n, precision="complex64")
    assert out.dtype == np.complex64
    close(out, st.amplitudes(), 2e-5)


def mixed_rotations(n: int) -> ParameterizedCircuit:
    pc = ParameterizedCircuit(3 * n)
    k = 0
    for layer in range(3):
        for q in range(n):
            pc.rotation([RotationAxis.X, RotationAxis.Y, RotationAxis.Z][(q + layer) % 3], q, k)
            k += 1
        for q in range(0, n - 1, 2 if layer % 2 else 1):
            t = (q + 1 + layer) % n
            if t != q:
                pc.cnot(q, t)
    return pc


def angle_rows(pops, params, scaled: bool) -> np.ndarray:
    """One instance's angle-table row: (cos, sin) of each rotation
    occurrence's half angle; in a plan with scaled rotations an X or Y
    occurrence holds (f, u) instead, f the larger of the two in magnitude,
    u the other over f, the lowest mantissa bit of u set when f = sin."""
    out = []
    for r in pops:
        if r["kind"] != 1:
            continue
        th = float(params[int(r["index"])])
        c, s = np.cos(th / 2), np.sin(th / 2)
        if scaled and int(r["gate"]) != 2:
            by_s = abs(s) > abs(c)
            f, u = (s, c / s) if by_s else (c, s / c)
            ub = (np.array([u]).view(np.uint64) & ~np.uint64(1)) | np.uint64(by_s)
            out += [f, float(ub.view(np.float64)[0])]
        else:
            out += [c, s]
    return np.array(out)


def deep_xy(n: int, layers: int = 5) -> ParameterizedCircuit:
    # X and Y rotations on every qubit per layer and a CNOT ladder: with the
    # whole state in one tile, 50 scaled rotations in one pass (two words of
    # form flags).
    pc = ParameterizedCircuit(layers * n)
    k = 0
    for layer in range(layers):
        for q in range(n):
            pc.rotation(RotationAxis.X if (q + layer) % 2 else RotationAxis.Y, q, k)
            k += 1
        for q in range(n - 1):
            pc.cnot(q, q + 1)
    return pc


@pytest.mark.parametrize("scaled", [True, False])
@pytest.mark.parametrize("case", ["hea", "qaoa", "mixed", "deep"])
def test_generated_kernels_read_rotation_angles_per_instance(case, scaled, monkeypatch):
    # Rotations (X/Y on register bits, Z and ZZ on any qubit) read
    # (cos, sin) of their half angle from the instance's table row, or
    # (scale, ratio) for X/Y in the scaled form (both forms of the
    # workgroup-uniform branch: angles of pi, -pi, 0 and pi/2 among them).
    if not scaled:
        monkeypatch.setenv("QF_SCALED_ROTATIONS", "0")
    if case == "hea":
        n, pc, M, RB = 12, hea_circuit(12, 3), 9, 3
    elif case == "qaoa":
        n, M, RB = 12, 10, 4
        pc = qaoa_parameterized(n, MaxCutProblem.random_3_regular(n, 5).edges, 2)
    elif case == "mixed":
        n, pc, M, RB = 11, mixed_rotations(11), 9, 3
    else:
        n, pc, M, RB = 10, deep_xy(10), 10, 3
    P = pc.parameter_count
    plan = Plan.parameterized(pc, n, kernel=3, tile_qubits=M, reg_bits=RB)
    assert plan.describe()["scaled_rotations"] == int(scaled)
    if case == "deep":
        assert len(plan.describe()["passes"]) == 1
        assert ("fw1" in plan.source()) == scaled
    pops = pc.to_pops()
    slots = [int(r["index"]) for r in pops if r["kind"] == 1]
    sets = [np.array(x) for x in uniform_parameters(P, 2, seed=7)]
    sets[0][:4] = [np.pi, -np.pi, 0.0, np.pi / 2]
    cs = np.concatenate([angle_rows(pops, x, scaled) for x in sets])
    a0 = np.zeros(2 << n, complex)
    a0[0] = a0[1 << n] = 1
    out = emulate(plan, a0, n, count=2, cs=cs, nslots=len(slots)).reshape(2, -1)
    for i in range(2):
        close(out[i], O.pc_run(pops, P, 0, sets[i], [], n), 1e-13)


def with_fixed_gates(pc: ParameterizedCircuit, n: int) -> ParameterizedCircuit:
    # Fixed gates between the rotations: literal dense ops (H, Y) and
    # diagonal ones (T, S) in the compiled passes, scales carried out of them.
    for q in range(n):
        pc.push([Gate.H, Gate.T, Gate.Y, Gate.S][q % 4], q)
    return pc


@pytest.mark.parametrize("n,M,RB,precision,circ,scaled", [(11, 9, 3, "complex128", "hea", True),
                                                            (12, 10, 3, "complex128", "hea", True),
                                                            (12, 10, 4, "complex64", "hea", True),
                                                            (11, 9, 3, "complex128", "mixed", True),
                                                            (11, 9, 3, "complex128", "mixed", False),
                                                            (10, 10, 3, "complex128", "deep", True)])
def test_adjoint_form_gives_the_oracle_gradient(n, M, RB, precision, circ, scaled, monkeypatch):
    # The compiled reverse sweep of the adjoint gradient: the forward plan's
    # kernels make psi, lambda = H psi, then the reverse plan's kernels in
    # their two-state form carry psi and lambda back and record
    # Im <lambda| G |psi> per tile for every rotation.  Summing the partial
    # table per parameter gives the oracle's adjoint gradient, and psi
    # returns to |0>.  Scaled: X/Y rotations in the (scale, ratio) form,
    # their scales carried in the thread accumulator and weighed into the
    # records.
    if not scaled:
        monkeypatch.setenv("QF_SCALED_ROTATIONS", "0")
    if circ == "deep":
        pc = deep_xy(n)
    else:
        pc = hea_circuit(n, 2) if circ == "hea" else with_fixed_gates(mixed_rotations(n), n)
    if circ == "hea":
        with_fixed_gates(pc, n)
    P = pc.parameter_count
    params = np.array(uniform_parameters(P, 1, seed=n + M)[0])
    params[:3] = [np.pi, -np.pi / 2, 3.0]
    H = heisenberg_hamiltonian(n, seed=3)
    pops = pc.to_pops()
    slots = [int(r["index"]) for r in pops if r["kind"] == 1]
    cs = angle_rows(pops, params, scaled)
    fwd = Plan.parameterized(pc, n, kernel=3, tile_qubits=M, reg_bits=RB, precision=precision)
    rev = Plan.parameterized(pc, n, reverse=True, kernel=3, tile_qubits=M, reg_bits=RB, precision=precision)
    assert fwd.describe()["scaled_rotations"] == rev.describe()["scaled_rotations"] == int(scaled)
    src = rev.source(precision, adjoint=True)
    assert "qf_wave_sum" in src and "partial[b * nslots" in src
    a0 = np.zeros(1 << n, complex)
    a0[0] = 1
    psi = emulate(fwd, a0, n, cs=cs, nslots=len(slots), precision=precision)
    lam = NP.pauli_sum_matrix(H.terms) @ psi.astype(complex)
    tb = [p["tile_bits"] for p in rev.describe()["passes"]]
    psi2, _, part = EmulatedProgram(src, EMU_DIR).run(psi, n, tb, cs=cs, nslots=len(slots), lam=lam)
    per_slot = part.sum(axis=0)
    grad = np.zeros(P)
    for k, p in enumerate(slots):
        grad[p] += per_slot[k]
    want = O.pc_adjoint_gradient(pops, P, 0, params, [], n, H.to_terms())
    tol = 1e-12 if precision == "complex128" else 2e-4
    close(grad, want, tol)
    close(psi2, a0, 1e-13 if precision == "complex128" else 1e-5)


def test_compiled_flag_defaults_and_errors():
    c = random_layer_circuit(18, 2, seed=1, pairs_per_layer=9)
    assert Plan(c, 18).describe()["compiled"] == 1          # default from 18 qubits
    assert Plan(c, 18, kernel=2).describe()["compiled"] == 0
    c17 = random_layer_circuit(17, 2, seed=1, pairs_per_layer=8)
    assert Plan(c17, 17).describe()["compiled"] == 0
    assert Plan(c17, 17, kernel=3).describe()["compiled"] == 2
    assert Plan(c17, 17, fuse=False).describe()["compiled"] == 0
    with pytest.raises(QuantumError, match="compiled passes need"):
        Plan(c17, 17, kernel=3, tile_qubits=8, reg_bits=3)
    with pytest.raises(QuantumError, match="does not run compiled passes"):
        Plan(c17, 17, kernel=2).compile()


def test_hiprtc_compiles_a_plan_for_gfx950_without_a_device(tmp_path, monkeypatch):
    monkeypatch.setenv("QF_JIT_CACHE", str(tmp_path))
    c = random_layer_circuit(14, 3, seed=21, pairs_per_layer=7)
    plan = Plan(c, 14, kernel=3, tile_qubits=10, reg_bits=3)
    r = plan.compile()
    assert r["compile_ms"] > 0
    files = os.listdir(tmp_path)
    assert files and all(f.endswith(".co") for f in files)
    # The code object is an AMDGPU ELF.
    blob = open(os.path.join(tmp_path, files[0]), "rb").read()
    assert blob[:4] == b"\x7fELF"
