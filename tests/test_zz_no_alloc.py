"""The no-allocation Pauli expectation (qf_state_expectation_pauli_sum_noalloc)
against the allocating entry point it varies (qf_state_expectation_pauli_sum),
through the C ABI: two calls in a row, the second on a smaller state with fewer
terms (the shared scratch keeps the first call's contents), in complex128 and
complex64, plus its limits.  Results must be identical: the same kernels run
in the same order over the same partial layout."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from quantumforge_amd import PauliSum, StateVector
from quantumforge_amd import _native as N
from quantumforge_amd._native import lib, ptr

pytestmark = pytest.mark.gpu


def random_state(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return a / np.linalg.norm(a)


def random_sum(n, count, seed):
    rng = np.random.default_rng(seed)
    H = PauliSum(n)
    for _ in range(count):
        H.add(complex(rng.normal(), rng.normal() * 0.1), "".join("IXYZ"[int(k)] for k in rng.integers(0, 4, n)))
    return H.to_terms()


def both(dev, terms):
    re, im = C.c_double(), C.c_double()
    st = lib().qf_state_expectation_pauli_sum(dev._h, ptr(terms), len(terms), C.byref(re), C.byref(im))
    assert st == 0, N.STATUS.get(st)
    re2, im2 = C.c_double(), C.c_double()
    st = lib().qf_state_expectation_pauli_sum_noalloc(dev._h, ptr(terms), len(terms), C.byref(re2), C.byref(im2))
    assert st == 0, (N.STATUS.get(st), lib().qf_last_error_message())
    return (re.value, im.value), (re2.value, im2.value)


@pytest.mark.parametrize("precision", ["complex128", "complex64"])
def test_noalloc_expectation_equals_the_allocating_entry_point(backend, precision):
    for n, nterms, seed in ((16, 300, 1), (9, 7, 2)):  # second call: smaller state, fewer terms
        a0 = random_state(n, seed)
        dev = backend.upload(StateVector(a0, n), precision=precision)
        terms = random_sum(n, nterms, seed)
        want, got = both(dev, terms)
        assert got == want, (n, got, want)
        if precision == "complex128":
            o = O.State(n, a0).pauli_expectation(terms)
            assert abs(got[0] - o.real) < 1e-12 and abs(got[1] - o.imag) < 1e-12


def test_noalloc_expectation_limits(backend):
    n = 6
    dev = backend.upload(StateVector(random_state(n, 3), n))
    terms = random_sum(n, 4097, 4)
    re, im = C.c_double(), C.c_double()
    st = lib().qf_state_expectation_pauli_sum_noalloc(dev._h, ptr(terms), len(terms), C.byref(re), C.byref(im))
    assert N.STATUS[st] == "InvalidArgument"
    assert b"at most 4096" in lib().qf_last_error_message()
    # 4096 terms is the limit and still runs.
    want, got = both(dev, terms[:4096])
    assert got == want
    bad = random_sum(n, 1, 5)
    bad["z_mask"][0] = 1 << n
    st = lib().qf_state_expectation_pauli_sum_noalloc(dev._h, ptr(bad), 1, C.byref(re), C.byref(im))
    assert N.STATUS[st] == "QubitOutOfRange"
    st = lib().qf_state_expectation_pauli_sum_noalloc(dev._h, ptr(terms), 0, C.byref(re), C.byref(im))
    assert st == 0 and re.value == 0.0 and im.value == 0.0
