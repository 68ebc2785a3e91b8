"""GPU checks of the C ABI's guards and accessors: a plan runs only for the
circuit it was built from (no launch otherwise), the executed kernel family is
recorded in the plan stats, the budgeted constructors follow
StateVector::zero_with_memory_budget / DensityMatrix::maximally_mixed_with_memory_budget
(tests/simulation_budget.rs), and the amplitude gather reads what a download
reads."""
import numpy as np
import pytest

from oracle import oracle as O
from quantumforge_amd import Circuit, Gate, QuantumError, StateVector
from quantumforge_amd.backend import Plan
from quantumforge_amd.workloads import hea_circuit, random_layer_circuit

from .conftest import close

pytestmark = pytest.mark.gpu


def random_state(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return a / np.linalg.norm(a)


def refused(fn):
    with pytest.raises(QuantumError) as e:
        fn()
    assert e.value.kind == "InvalidArgument", e.value
    return str(e.value)


@pytest.mark.parametrize("n", [6, 20])
def test_plan_of_another_circuit_is_refused_before_any_launch(backend, n):
    a0 = random_state(n, 3)
    dev = backend.upload(StateVector(a0, n))
    c = random_layer_circuit(n, 3, seed=11, pairs_per_layer=n // 3)
    plan = Plan(c, n)
    # Same structure and op count, other rotation angles: a different circuit.
    other = random_layer_circuit(n, 3, seed=11, pairs_per_layer=n // 3)
    ops = other.to_ops()
    k = int(np.flatnonzero((ops["kind"] == 0) & (ops["gate"] >= 9))[0])
    ops[k]["theta"] += 0.25
    other = Circuit.from_ops(ops)
    assert "different circuit" in refused(lambda: dev.run_circuit(other, plan))
    # A plan of the same gates on a wider state.
    wide = Plan(c, n + 1)
    assert "qubit count" in refused(lambda: dev.run_circuit(c, wide))
    # A parameterized plan (its rotations read an angle table) and a density plan.
    pc = hea_circuit(n, 1)
    bound = pc.bind(np.linspace(-1, 1, pc.parameter_count))
    assert "qf_plan_build" in refused(lambda: dev.run_circuit(bound, Plan.parameterized(pc, n)))
    # A density plan of a unitary circuit on n/2 qubits is a plan over an
    # n-qubit view, but not one a state run may take.
    nd = n // 2
    dc = random_layer_circuit(nd, 1, seed=2, pairs_per_layer=1)
    assert "qf_plan_build" in refused(lambda: dev.run_circuit(dc, Plan.density(dc, nd)))
    # Nothing ran: the state is exactly the uploaded one.
    assert np.array_equal(dev.download().amplitudes, a0)
    # The right plan still runs, and records what it launched.
    dev.run_circuit(c, plan)
    close(dev.download().amplitudes, _oracle(n, c, a0), 1e-12)
    kinds = plan.kernels_run()
    if n >= 18:
        assert kinds == ["compiled"], kinds
    else:
        assert kinds == (["register_tile"] if plan.describe()["passes"][0].get("reg_bits") else ["lds_tile"]), kinds


def _oracle(n, c, a0):
    st = O.State(n, a0)
    st.run(c.to_ops())
    return st.amplitudes()


def test_batch_and_density_plans_are_checked(backend):
    n = 8
    c = random_layer_circuit(n, 2, seed=5, pairs_per_layer=3)
    batch = backend.upload_batch([StateVector(random_state(n, s), n) for s in range(2)])
    before = [s.amplitudes for s in batch.download()]
    other = random_layer_circuit(n, 2, seed=6, pairs_per_layer=3)
    assert len(other) == len(c)
    refused(lambda: batch.run_circuit(other, Plan(c, n)))
    for s, b in zip(batch.download(), before):
        assert np.array_equal(s.amplitudes, b)
    d = backend.density_from_state(backend.zero_state(4))
    refused(lambda: d.run_circuit(c, Plan(c, 8)))


def test_kernels_run_reports_the_exact_lowering(backend):
    n = 6
    c = random_layer_circuit(n, 2, seed=1, pairs_per_layer=2)
    plan = Plan(c, n, fuse=False)
    assert plan.kernels_run() == []
    dev = backend.zero_state(n)
    dev.run_circuit(c, plan)
    assert plan.kernels_run() == ["sweep"]
    assert np.array_equal(dev.download().amplitudes, O.run_circuit(n, c.to_ops()))


def test_zero_state_with_memory_budget(backend):
    # tests/simulation_budget.rs:41-49: 2^3 x 16 bytes fit 128 bytes, not 127.
    st = backend.zero_state_with_memory_budget(3, 128)
    out = st.download().amplitudes
    assert out[0] == 1 and not np.any(out[1:])
    with pytest.raises(QuantumError) as e:
        backend.zero_state_with_memory_budget(3, 127)
    assert e.value.kind == "MemoryBudgetExceeded"
    assert "at least 128 bytes" in str(e.value) and "127-byte" in str(e.value)
    # complex64 amplitudes are 8 bytes.
    backend.zero_state_with_memory_budget(3, 64, precision="complex64")
    with pytest.raises(QuantumError) as e:
        backend.zero_state_with_memory_budget(3, 63, precision="complex64")
    assert e.value.kind == "MemoryBudgetExceeded"
    with pytest.raises(QuantumError) as e:
        backend.zero_state_with_memory_budget(0, 1 << 40)
    assert e.value.kind == "ZeroQubits"
    # tests/simulation_budget.rs:62-70: an unrepresentable size.
    with pytest.raises(QuantumError) as e:
        backend.zero_state_with_memory_budget(64, (1 << 64) - 1)
    assert e.value.kind == "SimulationSizeOverflow"
    # Within budget but beyond the device kernels' 40-qubit limit.
    with pytest.raises(QuantumError) as e:
        backend.zero_state_with_memory_budget(41, (1 << 64) - 1)
    assert e.value.kind == "StateTooLarge"


def test_maximally_mixed_density_with_memory_budget(backend):
    # tests/simulation_budget.rs:51-59: 4^3 x 16 bytes = 1024.
    d = backend.maximally_mixed_density_with_memory_budget(3, 1024)
    tr = d.trace()
    assert abs(tr.real - 1) < 1e-12 and abs(tr.imag) < 1e-15
    with pytest.raises(QuantumError) as e:
        backend.maximally_mixed_density_with_memory_budget(3, 1023)
    assert e.value.kind == "MemoryBudgetExceeded"
    assert "at least 1024 bytes" in str(e.value)


@pytest.mark.parametrize("precision", ["complex128", "complex64"])
def test_gather_reads_what_download_reads(backend, precision):
    n = 14
    a0 = random_state(n, 9)
    dev = backend.upload(StateVector(a0, n), precision=precision)
    full = dev.download().amplitudes
    idx = np.concatenate([[0, (1 << n) - 1, 1 << (n - 1)], np.random.default_rng(1).integers(0, 1 << n, 5000)])
    got = dev.gather(idx)
    assert np.array_equal(got, full[idx])
    assert dev.gather([]).size == 0
    with pytest.raises(QuantumError) as e:
        dev.gather([1 << n])
    assert e.value.kind == "InvalidArgument"


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4])
def test_stream_rate_leaves_the_state_unchanged(backend, variant):
    n = 20
    a0 = random_state(n, 4)
    dev = backend.upload(StateVector(a0, n))
    gbs, ms = dev.stream_rate(variant, 2)
    assert gbs > 0 and ms > 0
    assert np.array_equal(dev.download().amplitudes, a0)
    with pytest.raises(QuantumError):
        dev.stream_rate(5, 1)


def test_copy_from_device_same_and_converted_precision(backend):
    # CudaDeviceState::copy_from_device, src/cuda.rs:1444-1456, plus conversion.
    n = 12
    a0 = random_state(n, 12)
    src = backend.upload(StateVector(a0, n))
    dst = backend.zero_state(n)
    dst.copy_from_device(src)
    assert np.array_equal(dst.download().amplitudes, a0)
    narrow = backend.zero_state(n, precision="complex64")
    narrow.copy_from_device(src)
    assert np.array_equal(narrow.download_c64(), a0.astype(np.complex64))
    wide = backend.zero_state(n)
    wide.copy_from_device(narrow)
    assert np.array_equal(wide.download().amplitudes, a0.astype(np.complex64).astype(np.complex128))
    with pytest.raises(QuantumError) as e:
        backend.zero_state(n + 1).copy_from_device(src)
    assert e.value.kind == "IncompatibleDeviceState"
