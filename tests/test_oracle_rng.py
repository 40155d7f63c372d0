"""The serial oracle's random stream against published known answers.

xoshiro256++ (rng.rs:32-45) is pinned by the rand_xoshiro crate's test
vector for the state (1, 2, 3, 4); SplitMix64 (rng.rs:19-25) by the first
output of Vigna's splitmix64.c from state 0; `below` (rng.rs:53-62) by the
closed form of rejection sampling checked by brute force.
"""

import ctypes

import numpy as np
import pytest

import oracle


def test_xoshiro256pp_matches_the_published_vector(oracle_lib):
    s = (ctypes.c_uint64 * 4)(1, 2, 3, 4)
    expected = [41943041, 58720359, 3588806011781223, 3591011842654386,
                9228616714210784205, 9973669472204895162, 14011001112246962877,
                12406186145184390807, 15849039046786891736, 10450023813501588000]
    got = [int(oracle_lib.oracle_xoshiro_next(s)) for _ in expected]
    assert got == expected


def test_splitmix64_first_output_from_zero(oracle_lib):
    st = ctypes.c_uint64(0)
    assert int(oracle_lib.oracle_splitmix64_next(ctypes.byref(st))) == 0xE220A8397B1DCDAF


def test_seeding_runs_splitmix64_four_times(oracle_lib):
    # Xoshiro::seed fills s[i] with the i-th SplitMix64 output (rng.rs:17-29)
    st = ctypes.c_uint64(12345)
    outs = [int(oracle_lib.oracle_splitmix64_next(ctypes.byref(st))) for _ in range(4)]
    s = (ctypes.c_uint64 * 4)()
    oracle_lib.oracle_xoshiro_seed(12345, s)
    assert list(s) == outs


def test_floats_use_the_top_53_bits(oracle_lib):
    a = oracle.Xoshiro(7)
    b = oracle.Xoshiro(7)
    for _ in range(100):
        u = b.next_u64()
        assert a.next_f64() == (u >> 11) * 2.0 ** -53


def test_below_is_rejection_sampling(oracle_lib):
    # below(n) = v % n for the first v below the largest multiple of n
    for n in (1, 3, 5, 8, 26, 1000003, 2**40 + 7):
        a, b = oracle.Xoshiro(11), oracle.Xoshiro(11)
        zone = 2**64 - 1 - ((2**64 - 1) % n)
        for _ in range(50):
            want = None
            while want is None:
                v = b.next_u64()
                if v < zone:
                    want = v % n
            assert a.below(n) == want


def test_below_covers_its_range(oracle_lib):
    # rng.rs:79-88
    r = oracle.Xoshiro(11)
    seen = np.zeros(5, dtype=int)
    for _ in range(10000):
        seen[r.below(5)] += 1
    assert (seen > 1500).all()
