"""Pins SmallRng (src/lib.rs:4881-4912): the C++ oracle and the product's
Python SmallRng against each other and against the published test vectors of
the two algorithms it composes (splitmix64, Marsaglia's xorshift64)."""
from oracle import oracle as O
from quantumforge_amd.workloads import MASK64, SmallRng


def splitmix64_stream(seed, n):
    # Vigna's splitmix64: state += golden; z = mix(state).
    out, s = [], seed
    for _ in range(n):
        s = (s + 0x9E3779B97F4A7C15) & MASK64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        out.append(z ^ (z >> 31))
    return out


def test_splitmix64_published_vector():
    # Published splitmix64 outputs for seed 1234567 (the reference generator's test vector).
    assert splitmix64_stream(1234567, 5) == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                                             4593380528125082431, 16408922859458223821]


def test_seeding_is_one_splitmix64_step():
    for seed in (0, 1, 2, 1234567, 2**63, MASK64):
        expect = splitmix64_stream(seed, 1)[0] or 0x9E3779B97F4A7C15
        assert SmallRng(seed).state == expect
        assert O.Rng(seed).state == expect


def test_xorshift64_marsaglia_vector():
    # Marsaglia (2003), xor64 with seed 88172645463325252.
    r = SmallRng(0)
    r.state = 88172645463325252
    assert [r.next_u64() for _ in range(3)] == [8748534153485358512, 3040900993826735515, 3453997556048239312]


def test_python_and_oracle_streams_agree():
    for seed in (0, 7, 42, 2024, 99999):
        a, b = SmallRng(seed), O.Rng(seed)
        for _ in range(200):
            assert a.next_u64() == b.next_u64()
        for _ in range(200):
            assert a.next_f64() == b.next_f64()


def test_next_f64_range_and_resolution():
    r = SmallRng(3)
    xs = [r.next_f64() for _ in range(10000)]
    assert all(0.0 <= x < 1.0 for x in xs)
    # (x >> 11) * 2^-53: every draw is a multiple of 2^-53.
    assert all((x * 2.0**53).is_integer() for x in xs)
    assert abs(sum(xs) / len(xs) - 0.5) < 0.02


def test_zero_state_is_remapped():
    # seeded() maps a zero finalizer output to the golden constant; xorshift never returns 0 from nonzero.
    r = SmallRng(5)
    assert all(r.next_u64() != 0 for _ in range(1000))
