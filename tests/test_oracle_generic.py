"""The field-generic Python oracle (oracle/gkr_generic.py).  Its M31 instantiation must
reproduce the reference-generated fixtures byte for byte, which pins the transcript
framing, the sumcheck and the verifier that its BN254 instantiation shares; the BN254
field arithmetic is Python big-int arithmetic and is checked against closed forms."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import gkrx
import oracle_lib as ol

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import gkr_generic as gg  # noqa: E402

FIXTURES = json.load(open(os.path.join(HERE, "golden", "fixtures.json")))["cases"]
TINY = [c for c in FIXTURES if c["depth"] * (1 << c["log_width"]) * c["lanes"] <= 131072]
KIND = {"sha256": 0, "keccak256": 1}


def _layers(seed, depth, k, add_permille=700, locality=True):
    return [([tuple(int(x) for x in g) for g in m], [tuple(int(x) for x in g) for g in a])
            for m, a in gkrx.synth_circuit(seed, depth, k, add_permille, locality)]


def _wit(seed, n, lanes):
    return gkrx.synth_witness(seed, n, lanes).reshape(n, lanes).tolist()


@pytest.mark.parametrize("case", TINY, ids=lambda c: c["name"])
def test_generic_m31_reproduces_reference_fixture(case):
    F = gg.m31(case["lanes"])
    layers = _layers(case["seed"], case["depth"], case["log_width"], case["add_permille"], case["locality"])
    claimed, proof = gg.prove(F, layers, _wit(case["seed"], 1 << case["log_width"], case["lanes"]),
                              KIND[case["hash"]])
    assert len(proof) == case["proof_len"]
    assert hashlib.sha256(proof).hexdigest() == case["proof_sha256"]
    cl = np.array(claimed, dtype=np.uint32)
    assert hashlib.sha256(cl.tobytes()).hexdigest() == case["claimed_sha256"]
    assert gg.verify(F, layers, claimed, proof, KIND[case["hash"]])


@pytest.mark.parametrize("kind", [0, 1])
def test_generic_m31_equals_c_oracle(kind, tmp_path):
    seed, depth, k, lanes = 77, 3, 5, 8
    F = gg.m31(lanes)
    layers = _layers(seed, depth, k)
    mp, ap = str(tmp_path / "m.txt"), str(tmp_path / "a.txt")
    gkrx.write_text(layers, mp, ap)
    assert gg.load_text(mp, ap) == [[m, a] for m, a in layers]
    wit = gkrx.synth_witness(seed, 1 << k, lanes)
    oc, op = ol.OracleCircuit(mp, ap).prove(lanes, kind, wit)
    gc, gp = gg.prove(F, layers, wit.reshape(-1, lanes).tolist(), kind)
    assert gp == op and np.array_equal(np.array(gc, dtype=np.uint32), oc)


def test_bn254_field_facts():
    r = gg.BN254_R
    assert r.bit_length() == 254
    # r is the BN254 group order: r = 36u^4 + 36u^3 + 18u^2 + 6u + 1, u = 4965661367192848881
    u = 4965661367192848881
    assert r == 36 * u**4 + 36 * u**3 + 18 * u**2 + 6 * u + 1
    assert pow(5, r - 1, r) == 1 and pow(2, (r - 1) // 2, r) != 0
    F = gg.bn254()
    assert F.repetitions() == 1 and F.S == 32
    assert F.challenge(bytes([0xff] * 32)) == (2**256 - 1) % r


@pytest.mark.parametrize("kind", [0, 1])
def test_bn254_prove_verify_and_tamper(kind):
    F = gg.bn254()
    layers = _layers(5, 3, 4)
    import random
    rnd = random.Random(9)
    wit = [[rnd.randrange(F.p)] for _ in range(16)]
    claimed, proof = gg.prove(F, layers, wit, kind)
    R = 1
    n = [s[0] for s in gg.shape(layers)]
    assert len(proof) == 16 * 32 + 32 + sum(6 * x * R + 2 * R for x in n) * 32
    # claimed output equals the multilinear extension of the evaluated output layer
    assert gg.verify(F, layers, claimed, proof, kind)
    bad = bytearray(proof)
    bad[16 * 32 + 32 + 3] ^= 1
    assert not gg.verify(F, layers, claimed, bytes(bad), kind)
    bad = [[(claimed[0][0] + 1) % F.p]]
    assert not gg.verify(F, layers, bad, proof, kind)
