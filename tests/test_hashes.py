"""Transcript hashes: the oracle's SHA-256 and Keccak-256 and the library's host
implementations against known answers, hashlib, and the SHA-256 compression vectors of
the reference's own self-test (lib/btc_sha256/crypto/sha256.cpp:489-575); the transcript
framing (include/fiat_shamir/transcript.hpp:91-147) against a hashlib recomputation of the
reference's cross-platform transcript test (tests/transcript.cpp:26-69)."""
import ctypes as C
import hashlib
import os
import random
import struct

import numpy as np
import pytest

import gkrx
import oracle_lib as ol

P = ol.P31

# Input and expected states of the reference's SHA-256 Transform self-test: the state after
# compressing the first i * 64 bytes (no padding) of the data below, for i = 0..8.
SELFTEST_DATA = (
    b"Lorem ipsum dolor sit amet, consectetur adipiscing elit, sed do "
    b"eiusmod tempor incididunt ut labore et dolore magna aliqua. Et m"
    b"olestie ac feugiat sed lectus vestibulum mattis ullamcorper. Mor"
    b"bi blandit cursus risus at ultrices mi tempus imperdiet nulla. N"
    b"unc congue nisi vita suscipit tellus mauris. Imperdiet proin fer"
    b"mentum leo vel orci. Massa tempor nec feugiat nisl pretium fusce"
    b" id velit. Telus in metus vulputate eu scelerisque felis. Mi tem"
    b"pus imperdiet nulla malesuada pellentesque. Tristique magna sit.")
SELFTEST_STATES = [
    [0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a, 0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19],
    [0x91f8ec6b, 0x4da10fe3, 0x1c9c292c, 0x45e18185, 0x435cc111, 0x3ca26f09, 0xeb954cae, 0x402a7069],
    [0xcabea5ac, 0x374fb97c, 0x182ad996, 0x7bd69cbf, 0x450ff900, 0xc1d2be8a, 0x6a41d505, 0xe6212dc3],
    [0xbcff09d6, 0x3e76f36e, 0x3ecb2501, 0x78866e97, 0xe1c1e2fd, 0x32f4eaff, 0x8aa6c4e5, 0xdfc024bc],
    [0xa08c5d94, 0x0a862f93, 0x6b7f2f40, 0x8f9fae76, 0x6d40439f, 0x79dcee0c, 0x3e39ff3a, 0xdc3bdbb1],
    [0x216a0895, 0x9f1a3662, 0xe99946f9, 0x87ba4364, 0x0fb5db2c, 0x12bed3d3, 0x6689c0c7, 0x292f1b04],
    [0xca3067f8, 0xbc8c2656, 0x37cb7e0d, 0x9b6b8b0f, 0x46dc380b, 0xf1287f57, 0xc42e4b23, 0x3fefe94d],
    [0x3e4c4039, 0xbb6fca8c, 0x6f27d2f7, 0x301e44a4, 0x8352ba14, 0x5769ce37, 0x48a1155f, 0xc0e1c4c6],
    [0xfe2fa9dd, 0x69d0862b, 0x1ae0db23, 0x471f9244, 0xf55c0145, 0xc30f9c3b, 0x40a84ea0, 0x5b8a266c],
]

KECCAK256_KAT = {
    b"": "c5d2460186f7233c927e7db2dcc703c0e500b653ca82273b7bfad8045d85a470",
    b"abc": "4e03657aea45a94fc7d47ba826c8d667c0d1e6e33a64a036ec44f58fa12d6c45",
}
SHA256_KAT = {
    b"": "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855",
    b"abc": "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad",
    b"abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq":
        "248d6a61d20638b8e5c026930c3e6039a33ce45964ff2167f6ecedd419db06c1",
}


def test_sha256_compression_matches_reference_selftest():
    o = ol.oracle()
    o.oracle_sha256_blocks.argtypes = [C.c_void_p, C.c_char_p, C.c_size_t]
    for i, expect in enumerate(SELFTEST_STATES):
        st = (C.c_uint32 * 8)(*SELFTEST_STATES[0])
        o.oracle_sha256_blocks(st, SELFTEST_DATA, i)
        assert list(st) == expect


@pytest.mark.parametrize("msg,hexd", list(SHA256_KAT.items()))
def test_sha256_known_answers(msg, hexd):
    assert ol.sha256(msg).hex() == hexd
    assert gkrx.host_hash(gkrx.FS_SHA256, msg).hex() == hexd


@pytest.mark.parametrize("msg,hexd", list(KECCAK256_KAT.items()))
def test_keccak256_known_answers(msg, hexd):
    assert ol.keccak256(msg).hex() == hexd
    assert gkrx.host_hash(gkrx.FS_KECCAK256, msg).hex() == hexd


def test_sha256_random_lengths_against_hashlib():
    rng = random.Random(5)
    for n in list(range(0, 200)) + [511, 512, 513, 4096, 1 << 20, (1 << 20) + 37]:
        m = bytes(rng.getrandbits(8) for _ in range(min(n, 600))) * (n // 600 + 1)
        m = m[:n]
        ref = hashlib.sha256(m).digest()
        assert ol.sha256(m) == ref, n
        assert gkrx.host_hash(gkrx.FS_SHA256, m) == ref, n


def test_keccak_permutation_against_sha3_and_host_impl():
    """Same sponge with pad byte 0x06 is SHA3-256: pins the permutation, rate and
    multi-block absorption (lengths around the 136-byte rate)."""
    rng = random.Random(6)
    for n in list(range(0, 420)) + [1 << 16, (1 << 16) + 135]:
        m = os.urandom(n) if n > 420 else bytes(rng.getrandbits(8) for _ in range(n))
        assert ol.keccak256(m, pad=0x06) == hashlib.sha3_256(m).digest(), n
        assert gkrx.host_hash(gkrx.FS_KECCAK256, m) == ol.keccak256(m), n


def _transcript(kind):
    o = ol.oracle()
    o.oracle_transcript_new.restype = C.c_void_p
    o.oracle_transcript_new.argtypes = [C.c_int]
    o.oracle_transcript_append.argtypes = [C.c_void_p, C.c_char_p, C.c_size_t]
    o.oracle_transcript_challenge.argtypes = [C.c_void_p]
    o.oracle_transcript_challenge.restype = C.c_uint32
    o.oracle_transcript_free.argtypes = [C.c_void_p]
    return o, o.oracle_transcript_new(kind)


@pytest.mark.parametrize("kind", [ol.HASH_SHA256, ol.HASH_KECCAK256])
def test_transcript_framing_cross_platform_scenario(kind):
    """tests/transcript.cpp:26-69: append packed M31(0..7), draw alpha; append M31(8..15),
    draw beta, then gamma with nothing appended.  A challenge hashes only the bytes
    appended since the previous one (or the previous digest), and is the first four
    digest bytes as a little-endian u32 reduced mod p."""
    H = (lambda b: hashlib.sha256(b).digest()) if kind == ol.HASH_SHA256 else ol.keccak256
    o, t = _transcript(kind)
    f0 = struct.pack("<8I", *range(8))
    f1 = struct.pack("<8I", *range(8, 16))
    o.oracle_transcript_append(t, f0, len(f0))
    alpha = o.oracle_transcript_challenge(t)
    o.oracle_transcript_append(t, f1, len(f1))
    beta = o.oracle_transcript_challenge(t)
    gamma = o.oracle_transcript_challenge(t)
    o.oracle_transcript_free(t)
    d1 = H(f0)
    d2 = H(f1)
    d3 = H(d2)
    le = lambda d: struct.unpack("<I", d[:4])[0] % P
    assert (alpha, beta, gamma) == (le(d1), le(d2), le(d3))


@pytest.mark.parametrize("length", [0, 1, 31, 32, 135, 136, 137, 272, 1000, 65536 + 5])
def test_keccak_many_equals_one_at_a_time(length):
    """The eight-lane (AVX-512) Keccak-256 of the batched prover's transcript prefixes gives
    the scalar digest of every message, for partial groups too."""
    rng = random.Random(length)
    for n in (1, 3, 8, 11):
        msgs = [bytes(rng.getrandbits(8) for _ in range(length)) for _ in range(n)]
        got, _ = gkrx.host_hash_many(gkrx.FS_KECCAK256, msgs)
        assert got == [gkrx.host_hash(gkrx.FS_KECCAK256, m) for m in msgs]
    got, _ = gkrx.host_hash_many(gkrx.FS_SHA256, msgs)
    assert got == [hashlib.sha256(m).digest() for m in msgs]
