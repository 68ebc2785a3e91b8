"""The MQT benchmark fixtures (tests/mqt_circuits/*.qasm, copied under
tests/golden/mqt/) parsed by the QASM reader and run through the HIP path,
against the oracle on the same circuits; exported QASM and JSON read back
and run again."""
import glob
import os

import pytest

from oracle import oracle as O
from quantumforge_amd import ShotResult, StateVector
from quantumforge_amd.formats import export_qasm, from_json, parse_qasm, to_json

from .conftest import ROOT, close

pytestmark = pytest.mark.gpu

FIXTURES = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "mqt", "*.qasm")))


def load(path):
    with open(path) as f:
        c = parse_qasm(f.read())
    n = int(os.path.basename(path).split("_")[1].split(".")[0])
    return c, n


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_fixture_matches_oracle(backend, path):
    c, n = load(path)
    o = O.State(n)
    o.run(c.to_ops())
    got = backend.run_circuit(StateVector.zero(n), c)
    close(got.amplitudes, o.amplitudes(), 1e-12)
    again = backend.run_circuit(StateVector.zero(n), parse_qasm(export_qasm(c)))
    close(again.amplitudes, o.amplitudes(), 1e-12)
    again = backend.run_circuit(StateVector.zero(n), from_json(to_json(c)))
    close(again.amplitudes, o.amplitudes(), 1e-12)


def test_fixture_shots_match_oracle(backend):
    c, n = load(os.path.join(ROOT, "tests", "golden", "mqt", "ghz_8.qasm"))
    for q in range(n):
        c.measure(q, q)
    init = StateVector.zero(n)
    want = ShotResult.from_registers(O.run_shots(init.amplitudes, n, c.to_ops(), 200, 11))
    assert backend.run_circuit_shots(init, c, 200, 11) == want
    assert backend.run_circuit_shots_batched(init, c, 200, 11) == want
    assert set(want.counts) <= {(0,) * n, (1,) * n}
