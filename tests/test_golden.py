"""The oracle against the golden fixtures produced by the REFERENCE prover
(tests/golden/make_golden.py), and -- where the reference drivers are built
(oracle/_ref) -- against the reference directly on fresh shapes, including acceptance of
oracle proofs by the reference verifier and rejection of tampered ones."""
import hashlib
import json
import os

import numpy as np
import pytest

import gkrx
import oracle_lib as ol

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
FIXTURES = json.load(open(os.path.join(GOLD, "fixtures.json")))["cases"]
SMALL = [c for c in FIXTURES if c["depth"] * (1 << c["log_width"]) <= (1 << 18)]
HASH = {"sha256": ol.HASH_SHA256, "keccak256": ol.HASH_KECCAK256}


def _files(case, tmp_path):
    layers = gkrx.synth_circuit(case["seed"], case["depth"], case["log_width"], case["add_permille"],
                                case["locality"])
    mp, ap = str(tmp_path / "mul.txt"), str(tmp_path / "add.txt")
    gkrx.write_text(layers, mp, ap)
    wit = gkrx.synth_witness(case["seed"], 1 << case["log_width"], case["lanes"])
    return layers, mp, ap, wit


@pytest.mark.parametrize("case", SMALL, ids=lambda c: c["name"])
def test_oracle_reproduces_reference_fixture(case, tmp_path):
    layers, mp, ap, wit = _files(case, tmp_path)
    assert sum(len(m) for m, _ in layers) == case["mul_gates"]
    assert sum(len(a) for _, a in layers) == case["add_gates"]
    claimed, proof = ol.OracleCircuit(mp, ap).prove(case["lanes"], HASH[case["hash"]], wit)
    assert len(proof) == case["proof_len"]
    assert hashlib.sha256(proof).hexdigest() == case["proof_sha256"]
    assert hashlib.sha256(claimed.astype(np.uint32).tobytes()).hexdigest() == case["claimed_sha256"]


def test_c1_files_byte_for_byte(tmp_path):
    """The committed c1 files (8 layers x 2^10, M31x8, SHA-256 -- the reference's shipped
    configuration) regenerate identically and the oracle reproduces the reference proof."""
    d = os.path.join(GOLD, "c1")
    case = next(c for c in FIXTURES if c["name"] == "c1_sha256")
    _, mp, ap, wit = _files(case, tmp_path)
    assert open(mp).read() == open(os.path.join(d, "mul.txt")).read()
    assert open(ap).read() == open(os.path.join(d, "add.txt")).read()
    assert np.array_equal(wit, np.fromfile(os.path.join(d, "witness.bin"), dtype=np.uint32))
    ref_proof = open(os.path.join(d, "proof.bin"), "rb").read()
    ref_claimed = np.fromfile(os.path.join(d, "claimed.bin"), dtype=np.uint32).reshape(-1, 8)
    oc = ol.OracleCircuit(os.path.join(d, "mul.txt"), os.path.join(d, "add.txt"))
    claimed, proof = oc.prove(8, ol.HASH_SHA256, wit)
    assert proof == ref_proof
    assert np.array_equal(claimed, ref_claimed)
    assert oc.verify(8, ol.HASH_SHA256, ref_claimed, ref_proof) == 1
    commit = (1 << 10) * 32
    for pos in (0, commit - 1, commit + 32 + 5, len(ref_proof) - 1):
        bad = bytearray(ref_proof)
        bad[pos] ^= 0x10
        assert oc.verify(8, ol.HASH_SHA256, ref_claimed, bytes(bad)) == 0, pos
    # Verifier::verify recomputes the grinding and steps over the prover's 32 grind bytes
    # without reading them (include/LinearGKR/LinearGKR.hpp:119-121): they are unchecked.
    bad = bytearray(ref_proof)
    bad[commit + 5] ^= 0x10
    assert oc.verify(8, ol.HASH_SHA256, ref_claimed, bytes(bad)) == 1
    if ol.ref_available(8, ol.HASH_SHA256):
        assert ol.ref_verify(8, ol.HASH_SHA256, os.path.join(d, "mul.txt"), os.path.join(d, "add.txt"),
                             ref_claimed, bytes(bad))


needs_ref = pytest.mark.skipif(not (ol.ref_available(8, 0) and ol.ref_available(16, 1)),
                               reason="reference drivers (oracle/_ref) not built or no AVX-512 on this host")


@needs_ref
@pytest.mark.parametrize("lanes,h,seed,D,k,addp,loc", [
    (8, 0, 101, 3, 7, 500, False), (8, 1, 102, 4, 5, 900, True), (16, 0, 103, 2, 9, 100, True),
    (16, 1, 104, 5, 4, 700, True), (16, 1, 105, 3, 11, 700, True)])
def test_oracle_equals_reference_fresh_shapes(lanes, h, seed, D, k, addp, loc, tmp_path):
    layers = gkrx.synth_circuit(seed, D, k, addp, loc)
    mp, ap = str(tmp_path / "m.txt"), str(tmp_path / "a.txt")
    gkrx.write_text(layers, mp, ap)
    wit = gkrx.synth_witness(seed, 1 << k, lanes)
    rc, rp = ol.ref_prove(lanes, h, mp, ap, wit)
    oc = ol.OracleCircuit(mp, ap)
    c, p = oc.prove(lanes, h, wit)
    assert p == rp and np.array_equal(c, rc)
    assert ol.ref_verify(lanes, h, mp, ap, c, p)
    bad = bytearray(p)
    bad[len(p) // 2] ^= 1
    assert not ol.ref_verify(lanes, h, mp, ap, c, bytes(bad))
    assert oc.verify(lanes, h, c, bytes(bad)) == 0


@needs_ref
def test_reference_rejects_wrong_claim(tmp_path):
    layers = gkrx.synth_circuit(7, 3, 6, 500, False)
    mp, ap = str(tmp_path / "m.txt"), str(tmp_path / "a.txt")
    gkrx.write_text(layers, mp, ap)
    wit = gkrx.synth_witness(7, 1 << 6, 8)
    c, p = ol.OracleCircuit(mp, ap).prove(8, 0, wit)
    c2 = c.copy()
    c2[1, 3] = (int(c2[1, 3]) + 1) % ol.P31
    assert ol.ref_verify(8, 0, mp, ap, c, p)
    assert not ol.ref_verify(8, 0, mp, ap, c2, p)
    assert ol.OracleCircuit(mp, ap).verify(8, 0, c2, p) == 0
