"""The library's host verifier (gkrx_verify_arrays: no context, no GPU) against proofs it
did not make: the reference-generated c1 proof, C-oracle proofs of the golden shapes
(whose bytes the fixtures pin to the reference), and BN254 proofs of the field-generic
Python oracle.  Every honest proof is accepted; tampered proofs and claims are not."""
import hashlib
import json
import os
import random
import sys

import numpy as np
import pytest

import gkrx
import oracle_lib as ol

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import gkr_generic as gg  # noqa: E402

FIXTURES = json.load(open(os.path.join(HERE, "golden", "fixtures.json")))["cases"]
SMALL = [c for c in FIXTURES if c["depth"] * (1 << c["log_width"]) <= (1 << 14)]
HASH = {"sha256": gkrx.FS_SHA256, "keccak256": gkrx.FS_KECCAK256}


def test_reference_c1_proof_accepted():
    d = os.path.join(HERE, "golden", "c1")
    case = next(c for c in FIXTURES if c["name"] == "c1_sha256")
    layers = gkrx.synth_circuit(case["seed"], case["depth"], case["log_width"], case["add_permille"],
                                case["locality"])
    proof = open(os.path.join(d, "proof.bin"), "rb").read()
    claimed = np.fromfile(os.path.join(d, "claimed.bin"), dtype=np.uint32)
    cfg = gkrx.Config(fs_hash=gkrx.FS_SHA256, nb_parallel=8)
    assert gkrx.verify_layers(cfg, layers, claimed, proof)
    bad = bytearray(proof)
    bad[(1 << 10) * 32 + 32 + 7] ^= 1
    assert not gkrx.verify_layers(cfg, layers, claimed, bytes(bad))
    cl = claimed.copy()
    cl[3] = (int(cl[3]) + 1) % gkrx.P31
    assert not gkrx.verify_layers(cfg, layers, cl, proof)


@pytest.mark.parametrize("case", SMALL, ids=lambda c: c["name"])
def test_fixture_proofs_accepted(case, tmp_path):
    layers = gkrx.synth_circuit(case["seed"], case["depth"], case["log_width"], case["add_permille"],
                                case["locality"])
    mp, ap = str(tmp_path / "m.txt"), str(tmp_path / "a.txt")
    gkrx.write_text(layers, mp, ap)
    wit = gkrx.synth_witness(case["seed"], 1 << case["log_width"], case["lanes"])
    claimed, proof = ol.OracleCircuit(mp, ap).prove(case["lanes"], HASH[case["hash"]], wit)
    assert hashlib.sha256(proof).hexdigest() == case["proof_sha256"]
    cfg = gkrx.Config(fs_hash=HASH[case["hash"]], nb_parallel=case["lanes"])
    assert gkrx.verify_layers(cfg, layers, claimed, proof)
    bad = bytearray(proof)
    bad[len(proof) - 5] ^= 0x10
    assert not gkrx.verify_layers(cfg, layers, claimed, bytes(bad))


def _int_layers(layers):
    return [([tuple(int(x) for x in g) for g in m], [tuple(int(x) for x in g) for g in a]) for m, a in layers]


@pytest.mark.parametrize("kind", [gkrx.FS_SHA256, gkrx.FS_KECCAK256])
@pytest.mark.parametrize("shape", [(11, 2, 3), (12, 4, 5), (13, 3, 7)])
def test_bn254_oracle_proofs_accepted(kind, shape):
    seed, depth, k = shape
    layers = gkrx.synth_circuit(seed, depth, k, 700, True)
    wit = gkrx.bn254_to_ints(gkrx.synth_witness_bn254(seed, 1 << k))
    F = gg.bn254()
    claimed, proof = gg.prove(F, _int_layers(layers), [[v] for v in wit], kind)
    cw = np.frombuffer(claimed[0][0].to_bytes(32, "little"), dtype=np.uint32)
    cfg = gkrx.Config(fs_hash=kind, field_type=gkrx.FIELD_BN254)
    assert gkrx.verify_layers(cfg, layers, cw, proof)
    rnd = random.Random(seed)
    for _ in range(4):
        pos = rnd.randrange((1 << k) * 32 + 32, len(proof))
        bad = bytearray(proof)
        bad[pos] ^= 1 << rnd.randrange(8)
        assert not gkrx.verify_layers(cfg, layers, cw, bytes(bad))
    wrong = np.frombuffer(((claimed[0][0] + 1) % gg.BN254_R).to_bytes(32, "little"), dtype=np.uint32)
    assert not gkrx.verify_layers(cfg, layers, wrong, proof)


def test_bn254_synth_witness():
    w = gkrx.synth_witness_bn254(2, 4096)
    vals = gkrx.bn254_to_ints(w)
    assert all(0 <= v < gg.BN254_R for v in vals)
    assert np.array_equal(w, gkrx.synth_witness_bn254(2, 4096))
    assert not np.array_equal(w, gkrx.synth_witness_bn254(3, 4096))
    # top limb is masked to 254 bits and values spread over [0, r): the mean is near r / 2
    assert max(v.bit_length() for v in vals) == 254
    assert abs(sum(vals) / len(vals) / gg.BN254_R - 0.5) < 0.03
