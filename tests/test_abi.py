"""CPU checks of the C ABI and the host planner (no device calls):
every entry point declared in include/qf_hip.h is exported, the record
layouts match the C structs, and fused plans replayed in NumPy reproduce the
oracle's result (so the schedule, the peephole merges and the tile/bit
mapping are right before any kernel runs)."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from quantumforge_amd import Circuit, Gate, QuantumError
from quantumforge_amd import _native as N
from quantumforge_amd.backend import Plan
from quantumforge_amd.workloads import qft_circuit, qft_like_circuit, qaoa_maxcut, random_layer_circuit, MaxCutProblem

from .conftest import ROOT, close

HEADER = os.path.join(ROOT, "include", "qf_hip.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(qf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 40
    out = subprocess.run(["nm", "-D", "--defined-only", N.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (qf_[a-z0-9_]+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, f"declared but not exported: {missing}"
    lib = N.lib()
    for n in names:
        assert getattr(lib, n) is not None


def test_record_layouts():
    assert N.OP_DTYPE.itemsize == 32
    assert N.POP_DTYPE.itemsize == 32
    assert N.PAULI_DTYPE.itemsize == 32
    assert N.ROT_DTYPE.itemsize == 16
    src = open(HEADER).read()
    assert "int32_t kind;\n    int32_t gate;\n    uint32_t a, b;\n    double theta;\n    double theta2;\n} qf_op;" in src
    assert N.OP_DTYPE.names == ("kind", "gate", "a", "b", "theta", "theta2")


def test_plan_stats_layout_and_identity_fields():
    import ctypes as C

    assert C.sizeof(N.PlanStats) == 56
    src = open(HEADER).read()
    assert "uint32_t kernels_run;" in src and "uint32_t runtime_rotations;" in src
    from quantumforge_amd.workloads import hea_circuit

    pc = hea_circuit(6, 2)
    ps = Plan.parameterized(pc, 6).stats()
    assert ps["runtime_rotations"] == pc.parameter_count and ps["kernels_run"] == 0
    c = random_layer_circuit(6, 2, seed=1, pairs_per_layer=2)
    s = Plan(c, 6).stats()
    assert s["runtime_rotations"] == 0 and s["kernels_run"] == 0


def test_status_names_match_reference_variants():
    lib = N.lib()
    for code, name in N.STATUS.items():
        assert lib.qf_status_name(code).decode() == name


def test_plan_errors_are_typed():
    c = Circuit()
    c.push(Gate.X, 4)
    with pytest.raises(QuantumError) as e:
        Plan(c, 3)
    assert e.value.kind == "QubitOutOfRange"
    c = Circuit()
    c.cnot(1, 1)
    with pytest.raises(QuantumError) as e:
        Plan(c, 3)
    assert e.value.kind == "InvalidQubitPair"
    c = Circuit()
    c.push(Gate.Rx(float("inf")), 0)
    with pytest.raises(QuantumError) as e:
        Plan(c, 3)
    assert e.value.kind == "NonFiniteAngle"
    with pytest.raises(QuantumError) as e:
        Plan(Circuit(), 0)
    assert e.value.kind == "ZeroQubits"


# ---------------------------------------------------------------------------
# NumPy replay of a plan (test-side executor)
# ---------------------------------------------------------------------------
def _c(m, k):
    return complex(m[2 * k], m[2 * k + 1])


def _mat1(psi, idx, q, m):
    mask = 1 << q
    i0 = idx[(idx & mask) == 0]
    a0, a1 = psi[i0].copy(), psi[i0 | mask].copy()
    psi[i0] = _c(m, 0) * a0 + _c(m, 1) * a1
    psi[i0 | mask] = _c(m, 2) * a0 + _c(m, 3) * a1


def _cnot(psi, idx, c, t):
    sel = idx[(((idx >> c) & 1) == 1) & (((idx >> t) & 1) == 0)]
    tmp = psi[sel].copy()
    psi[sel] = psi[sel | (1 << t)]
    psi[sel | (1 << t)] = tmp


def _swap(psi, idx, a, b):
    sel = idx[(((idx >> a) & 1) == 1) & (((idx >> b) & 1) == 0)]
    other = sel ^ (1 << a) ^ (1 << b)
    tmp = psi[sel].copy()
    psi[sel] = psi[other]
    psi[other] = tmp


def _mat2(psi, idx, a, b, m):
    """Dense 4x4 over (bit(b) * 2 + bit(a)), row-major 16 complex."""
    base = idx[(((idx >> a) & 1) == 0) & (((idx >> b) & 1) == 0)]
    sel = [base, base | (1 << a), base | (1 << b), base | (1 << a) | (1 << b)]
    v = [psi[s].copy() for s in sel]
    for r in range(4):
        psi[sel[r]] = sum(_c(m, 4 * r + c) * v[c] for c in range(4))


def _diag(psi, idx, q0, q1, m):
    kk = (idx >> q0) & 1
    if q1 >= 0:
        kk = kk | (((idx >> q1) & 1) << 1)
    d = np.array([_c(m, j) for j in range(4)])
    return psi * d[kk]


def replay(desc, psi):
    """Applies a described fused plan (LDS tile kernel or register kernel) to psi."""
    n = desc["qubits"]
    idx = np.arange(1 << n, dtype=np.int64)
    psi = psi.astype(complex).copy()
    for p in desc["passes"]:
        assert p["kind"] == "tile", "fused plans contain tile passes only"
        bits = list(range(p["low"])) + list(p["high"])  # tile bit -> global qubit
        reg = p["reg_bits"]
        regs = None
        for op in p["ops"]:
            k, m = op["kind"], op["m"]
            if reg:
                if k == 7:
                    regs = [bits[(op["q0"] >> (4 * i)) & 15] for i in range(reg)]
                    assert len(set(regs)) == reg
                    continue
                assert regs is not None, "register plan must open with a phase header"
                if k == 0:
                    _mat1(psi, idx, regs[op["q0"]], m)
                elif k == 1:
                    psi = _diag(psi, idx, op["q0"], op["q1"], m)
                elif k == 2:
                    _cnot(psi, idx, op["q0"], regs[op["q1"]])
                elif k == 3:
                    _swap(psi, idx, regs[op["q0"]], regs[op["q1"]])
                elif k == 4:
                    assert op["q0"] < op["q1"], "register-form dense 4x4 ops have q0 < q1"
                    _mat2(psi, idx, regs[op["q0"]], regs[op["q1"]], m)
                else:
                    raise AssertionError(f"unexpected op kind {k}")
            else:
                if k == 0:
                    _mat1(psi, idx, bits[op["q0"]], m)
                elif k == 1:
                    psi = _diag(psi, idx, op["q0"], op["q1"], m)
                elif k == 2:
                    _cnot(psi, idx, bits[op["q0"]], bits[op["q1"]])
                elif k == 3:
                    _swap(psi, idx, bits[op["q0"]], bits[op["q1"]])
                elif k == 4:
                    _mat2(psi, idx, bits[op["q0"]], bits[op["q1"]], m)
                else:
                    raise AssertionError(f"unexpected op kind {k}")
        if p.get("store"):
            psi = _store_permute(psi, idx, bits, p["store"])
    return psi


def _store_permute(psi, idx, bits, store):
    """A relabeling pass's store: tile bit t (loaded from index bit bits[t])
    is written to index bit store[t]; bits outside the tile stay put."""
    assert sorted(store) == sorted(bits), "a store permutes the tile's own bits"
    dst = idx.copy()
    for b in bits:
        dst &= ~(np.int64(1) << b)
    for src, to in zip(bits, store):
        dst |= ((idx >> src) & 1) << to
    out = np.empty_like(psi)
    out[dst] = psi
    return out


def check_tile_invariants(desc):
    n = desc["qubits"]
    for p in desc["passes"]:
        M, low, high, reg = p["tile_bits"], p["low"], p["high"], p["reg_bits"]
        assert len(high) == M - low and high == sorted(high) and all(low <= h < n for h in high)
        if reg:
            assert p["ops"][0]["kind"] == 7
            for op in p["ops"]:
                if op["kind"] in (0, 3, 4):
                    assert 0 <= op["q0"] < reg and (op["kind"] == 0 or 0 <= op["q1"] < reg)
                if op["kind"] == 2:
                    assert 0 <= op["q0"] < n and 0 <= op["q1"] < reg
            continue
        for op in p["ops"]:
            if op["kind"] in (0, 2, 3, 4):
                assert 0 <= op["q0"] < M and (op["kind"] == 0 or 0 <= op["q1"] < M)


@pytest.mark.parametrize("n,depth,tile,low,kernel,reg", [
    (8, 10, 4, 2, 1, 0), (12, 12, 6, 3, 1, 0), (14, 20, 8, 4, 1, 0), (16, 10, 11, 5, 1, 0), (10, 6, 10, 5, 1, 0),
    (5, 4, 11, 5, 0, 0), (16, 10, 11, 5, 2, 3), (16, 10, 12, 5, 2, 4), (14, 20, 10, 4, 2, 3), (12, 12, 9, 3, 2, 3),
    (12, 8, 10, 5, 2, 4), (11, 6, 11, 6, 2, 3)])
def test_fused_plan_replay_equals_oracle(n, depth, tile, low, kernel, reg):
    c = random_layer_circuit(n, depth, seed=n * 31 + depth, pairs_per_layer=n // 2)
    c.swap(0, n - 1)
    c.swap(1, 2)
    c.push(Gate.T, 2)
    c.push(Gate.Sdg, 2)
    plan = Plan(c, n, True, tile, low, kernel, reg)
    if kernel == 2:
        assert plan.describe()["passes"][0]["reg_bits"] == reg
    desc = plan.describe()
    check_tile_invariants(desc)
    st = plan.stats()
    assert st["ops_in"] == len(c) and st["passes"] == len(desc["passes"])
    psi0 = np.zeros(1 << n, complex)
    psi0[0] = 1
    close(replay(desc, psi0), O.run_circuit(n, c.to_ops()), 1e-12)


def test_qaoa_and_qft_peepholes_replay():
    prob = MaxCutProblem.random_3_regular(10, 2).with_random_weights(3, 0.5, 1.0)
    c = qaoa_maxcut(10, prob.edges, 3, [0.2, 0.5, 0.9], [0.4, 0.1, 0.7])
    plan = Plan(c, 10, True, 6, 3)
    desc = plan.describe()
    # every CNOT Rz CNOT cost term became one two-qubit diagonal: no CNOT remains
    assert all(op["kind"] != 2 for p in desc["passes"] for op in p["ops"])
    psi0 = np.zeros(1 << 10, complex)
    psi0[0] = 1
    close(replay(desc, psi0), O.run_circuit(10, c.to_ops()), 1e-12)
    c2 = qft_circuit(9)
    d2 = Plan(c2, 9, True, 5, 2).describe()
    basis5 = np.eye(1 << 9)[5]
    close(replay(d2, basis5), _oracle_from(9, c2, basis5), 1e-12)
    c3, _ = qft_like_circuit(10, 1, seed=1, max_distance=4)
    d3 = Plan(c3, 10, True, 7, 3).describe()
    close(replay(d3, psi0), O.run_circuit(10, c3.to_ops()), 1e-12)


def _oracle_from(n, c, psi):
    s = O.State(n, psi)
    s.run(c.to_ops())
    return s.amplitudes()


def test_30q_plan_is_fused():
    c = random_layer_circuit(30, 40, seed=2024)
    st = Plan(c, 30, True).stats()
    assert st["ops_in"] == 1600
    assert st["passes"] <= 60, st
    exact = Plan(c, 30, False).stats()
    assert exact["passes"] == 1600


# ---------------------------------------------------------------------------
# Dense-block plans (kernel = 4): replayed from their description
# ---------------------------------------------------------------------------
def replay_dense(desc, psi):
    """Applies a described dense-block plan: per pass the column and block-bit
    diagonal factors, then the block matrix U on the block's qubits (bit i of a
    block index is qubit block[i])."""
    n = desc["qubits"]
    idx = np.arange(1 << n, dtype=np.int64)
    psi = psi.astype(complex).copy()
    for p in desc["passes"]:
        assert p["kind"] == "dense"
        blk = p["block"]
        k = len(blk)
        assert blk == sorted(blk) and len(set(blk)) == k
        for e in p["pre_col"]:
            assert e["a"] not in blk and e["b"] not in blk
            psi = _diag(psi, idx, e["a"], e["b"], e["d"])
        for e in p["pre_blk"]:
            assert e["b"] not in blk
            psi = _diag(psi, idx, blk[e["a"]], e["b"], e["d"])
        U = np.array([complex(p["U"][2 * i], p["U"][2 * i + 1]) for i in range(1 << (2 * k))]).reshape(1 << k, 1 << k)
        # block index of every amplitude, and the column base
        b = np.zeros_like(idx)
        for i, q in enumerate(blk):
            b |= ((idx >> q) & 1) << i
        mask = sum(1 << q for q in blk)
        base = idx & ~mask
        cols = np.unique(base)
        off = np.array([sum(((bb >> i) & 1) << q for i, q in enumerate(blk)) for bb in range(1 << k)])
        A = psi[cols[:, None] + off[None, :]]  # [column][block index]
        psi[cols[:, None] + off[None, :]] = A @ U.T
    return psi


@pytest.mark.parametrize("k", [4, 5])
def test_dense_block_plans_replay_to_the_oracle(k):
    from quantumforge_amd.workloads import qft_like_circuit

    n = 11
    cases = [random_layer_circuit(n, 6, seed=3, pairs_per_layer=4), qft_like_circuit(n, 2, seed=4)[1],
             qft_like_circuit(n, 1, seed=5)[0], qft_circuit(n)]
    for c in cases:
        plan = Plan(c, n, kernel=4, max_block=k)
        d = plan.describe()
        st = plan.stats()
        assert st["dense_blocks"] == st["passes"] == len(d["passes"]) > 0
        assert all(len(p["block"]) == k for p in d["passes"])
        rng = np.random.default_rng(k)
        a0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        a0 /= np.linalg.norm(a0)
        o = O.State(n, a0)
        o.run(c.to_ops())
        close(replay_dense(d, a0), o.amplitudes(), 1e-12)


def test_dense_block_plan_counts_and_errors():
    from quantumforge_amd.workloads import hea_circuit, qft_like_circuit

    # QFT-like layers: one block per 5 qubits and layer, cross-block phases as
    # diagonal factors of the next pass.
    n = 33
    _, c = qft_like_circuit(n, 2, seed=33)
    st = Plan(c, n, kernel=4, max_block=5).stats()
    assert st["dense_blocks"] == st["passes"] <= 2 * 7 + 2, st
    with pytest.raises(QuantumError):
        Plan(c, n, kernel=4, max_block=3)
    with pytest.raises(QuantumError):
        Plan(random_layer_circuit(8, 2, seed=1, pairs_per_layer=2), 8, kernel=4, max_block=5)
    with pytest.raises(QuantumError):
        Plan.parameterized(hea_circuit(10, 1), 10, kernel=4)
