"""GPU proofs against the reference prover's own outputs (tests/golden/fixtures.json,
written by tests/golden/make_golden.py from the reference built from its sources),
including the full-size 24 x 2^20 M31x16 proofs of both transcript hashes."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import gkrx

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = json.load(open(os.path.join(HERE, "golden", "fixtures.json")))["cases"]
HASH = {"sha256": gkrx.FS_SHA256, "keccak256": gkrx.FS_KECCAK256}


@pytest.fixture(scope="module")
def contexts():
    ctxs = {}
    yield ctxs
    for c in ctxs.values():
        c.close()


@pytest.mark.parametrize("case", FIXTURES, ids=lambda c: c["name"])
def test_gpu_proof_equals_reference_fixture(case, contexts):
    print(f"[gpu] fixture {case['name']}", flush=True)
    key = (case["lanes"], case["hash"])
    if key not in contexts:
        contexts[key] = gkrx.Context(gkrx.Config(fs_hash=HASH[case["hash"]], nb_parallel=case["lanes"]))
    ctx = contexts[key]
    layers = gkrx.synth_circuit(case["seed"], case["depth"], case["log_width"], case["add_permille"],
                                case["locality"])
    circ = gkrx.Circuit.from_layers(ctx, layers)
    circ.set_witness(gkrx.synth_witness(case["seed"], 1 << case["log_width"], case["lanes"]))
    claimed, proof = gkrx.Prover(ctx).prove(circ)
    assert len(proof) == case["proof_len"]
    assert hashlib.sha256(bytes(proof)).hexdigest() == case["proof_sha256"]
    assert hashlib.sha256(claimed.astype(np.uint32).tobytes()).hexdigest() == case["claimed_sha256"]
    circ.close()
