"""The oracle's M31 arithmetic against Python-integer closed forms (the field is
GF(2^31 - 1)), brute force on edge ranges, and the properties tests/ff.cpp checks in
the reference (f * k == f + ... + f, f * f^-1 == 1, byte round trip)."""
import random
import struct

import pytest

import oracle_lib as ol

P = ol.P31
EDGE = [0, 1, 2, 3, (1 << 30) - 1, 1 << 30, (1 << 30) + 1, P - 3, P - 2, P - 1]


def o():
    return ol.oracle()


def test_edge_values_exhaustive_pairs():
    for a in EDGE:
        for b in EDGE:
            assert o().oracle_m31_add(a, b) == (a + b) % P
            assert o().oracle_m31_sub(a, b) == (a - b) % P
            assert o().oracle_m31_mul(a, b) == (a * b) % P
        assert o().oracle_m31_neg(a) == (-a) % P


def test_brute_force_low_and_high_ranges():
    vals = list(range(0, 48)) + list(range(P - 48, P))
    for a in vals:
        for b in vals:
            assert o().oracle_m31_mul(a, b) == (a * b) % P
            assert o().oracle_m31_add(a, b) == (a + b) % P


def test_random_against_python_ints():
    rng = random.Random(789)
    for _ in range(20000):
        a, b = rng.randrange(P), rng.randrange(P)
        assert o().oracle_m31_mul(a, b) == (a * b) % P
        assert o().oracle_m31_add(a, b) == (a + b) % P
        assert o().oracle_m31_sub(a, b) == (a - b) % P


def test_ff_properties_from_reference_tests():
    # tests/ff.cpp:9-36: f * k == sum of k copies of f; f * f^-1 == 1; serialization round trip
    rng = random.Random(1234)
    for _ in range(200):
        f = rng.randrange(1, P)
        k = rng.randrange(100)
        s = 0
        for _ in range(k):
            s = o().oracle_m31_add(s, f)
        assert o().oracle_m31_mul(f, k) == s
        inv = pow(f, P - 2, P)
        assert o().oracle_m31_mul(f, inv) == 1
        assert o().oracle_m31_from_bytes(struct.pack("<I", f)) == f


def test_inv2_constant():
    # M31::INV_2 = 2^30 (include/field/M31.hpp:451) is the inverse of 2
    assert o().oracle_m31_mul(1 << 30, 2) == 1


@pytest.mark.parametrize("v", [0, 1, P - 1, P, P + 1, 1 << 31, (1 << 32) - 2, (1 << 32) - 1, 0x9abcdef0])
def test_from_bytes_fully_reduces(v):
    assert o().oracle_m31_from_bytes(struct.pack("<I", v)) == v % P


@pytest.mark.parametrize("v", [0, 5, P - 1, P, P + 1, 1 << 31, (1 << 32) - 2, (1 << 32) - 1])
def test_from_u32_is_congruent_and_keeps_p(v):
    # M31(uint32) folds once (include/field/M31.hpp:58-62): congruent to v, below 2^31 + 1,
    # and M31(p) stays p, a non-canonical representative of zero.
    r = o().oracle_m31_from_u32(v)
    assert r % P == v % P
    assert r <= (1 << 31)
    if v == P:
        assert r == P
