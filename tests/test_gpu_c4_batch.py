"""The uneven-batch workload at its own shapes: a prefix of the 512-proof list (proof i has
seed 3 + i, width 2^k with k uniform in [10, 18], depth uniform in [4, 24], 70% add / 30%
mul, locality wiring, M31x16, Keccak-256), 32 proofs that include a 20-layer 2^18-wide one.
Every batched proof is byte-identical to the plain proof of the same circuit and accepted by
the verifier; sampled proofs (the largest included) are byte-identical to the C oracle's,
and the largest also to the reference prover built from its own sources (as
tests/GKR_Test.cpp:34-79 checks a proof: honest accept, tamper reject)."""
import os
import sys

import numpy as np
import pytest

import gkrx
import oracle_lib as ol

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402  (batch_specs: the workload's proof list)

pytestmark = pytest.mark.gpu

COUNT = 32          # proofs 0 .. 31 of the list; proof 31 is 20 layers x 2^18
SAMPLES = [31, 0, 1, 12]  # the largest, a 23 x 2^17 one and two 2^10-wide ones (7 and 24 layers)


@pytest.fixture(scope="module")
def c4():
    specs = bench.batch_specs(512)[:COUNT]
    assert max(k for _, _, k in specs) == 18 and (specs[31][1], specs[31][2]) == (20, 18)
    assert all(s == 3 + i for i, (s, _, _) in enumerate(specs))
    ctx = gkrx.Context(gkrx.Config(fs_hash=gkrx.FS_KECCAK256, nb_parallel=16))
    circs, layers = [], []
    for s, d, k in specs:
        ly = gkrx.synth_circuit(s, d, k, 700, True)
        c = gkrx.Circuit.from_layers(ctx, ly)
        c.set_witness(gkrx.synth_witness(s, 1 << k, 16))
        circs.append(c)
        layers.append(ly)
    print(f"[gpu] c4 slice: {COUNT} proofs, {sum(d << k for _, d, k in specs)} gates", flush=True)
    batch = gkrx.Prover(ctx).prove_batch(circs)
    yield ctx, specs, circs, layers, batch
    for c in circs:
        c.close()
    ctx.close()


def test_c4_batch_equals_plain_and_verifies(c4):
    ctx, specs, circs, _, batch = c4
    prover, verifier = gkrx.Prover(ctx), gkrx.Verifier(ctx)
    for i, ((claimed, proof), circ) in enumerate(zip(batch, circs)):
        circ.evaluate()
        c1, p1 = prover.prove(circ)
        assert np.array_equal(claimed, c1), i
        assert bytes(proof) == bytes(p1), i
        assert verifier.verify(circ, claimed, proof), i
    # tamper reject on the largest: one byte of a round polynomial, and a claim
    claimed, proof = batch[31]
    bad = np.array(proof, copy=True)
    bad[bad.size - 100] ^= 1
    assert not verifier.verify(circs[31], claimed, bad)
    badc = np.array(claimed, copy=True)
    badc.flat[0] = (int(badc.flat[0]) + 1) % gkrx.P31
    assert not verifier.verify(circs[31], badc, proof)


@pytest.mark.parametrize("idx", SAMPLES)
def test_c4_batch_matches_oracle(c4, tmp_path, idx):
    _, specs, _, layers, batch = c4
    s, d, k = specs[idx]
    print(f"[gpu] c4 proof {idx} ({d} x 2^{k}) vs the C oracle", flush=True)
    mp, ap = str(tmp_path / "mul.txt"), str(tmp_path / "add.txt")
    gkrx.write_text(layers[idx], mp, ap)
    wit = gkrx.synth_witness(s, 1 << k, 16)
    oc, op = ol.OracleCircuit(mp, ap).prove(16, ol.HASH_KECCAK256, wit)
    claimed, proof = batch[idx]
    assert np.array_equal(claimed, oc)
    assert bytes(proof) == op
    if idx == SAMPLES[0] and ol.ref_available(16, ol.HASH_KECCAK256):
        rc, rp = ol.ref_prove(16, ol.HASH_KECCAK256, mp, ap, wit)
        assert np.array_equal(claimed, rc) and bytes(proof) == rp
        assert ol.ref_verify(16, ol.HASH_KECCAK256, mp, ap, claimed, proof)
