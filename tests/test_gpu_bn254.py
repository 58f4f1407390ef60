"""BN254 on the GPU: every proof byte-identical to the field-generic Python oracle
(oracle/gkr_generic.py, whose structure is pinned on the M31 fixtures), claims equal,
accepted by the host verifier; the forward pass equals the oracle's evaluation; tampered
proofs are rejected.  The shapes cover layers small enough for the single-workgroup
rounds only (up to 512 rows) and layers wide enough (2^10, 2^13, 2^14) to take the grid
rounds first.  A batch of BN254 proofs (several streams at once) gives each proof's bytes
made alone."""
import os
import sys

import numpy as np
import pytest

import gkrx

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import gkr_generic as gg  # noqa: E402

pytestmark = pytest.mark.gpu

CASES = [  # (seed, depth, log_width, add_permille, locality, hash)
    (21, 2, 2, 500, False, gkrx.FS_SHA256),
    (22, 3, 4, 700, True, gkrx.FS_KECCAK256),
    (23, 4, 6, 0, True, gkrx.FS_SHA256),
    (24, 3, 8, 1000, False, gkrx.FS_KECCAK256),
    (27, 3, 10, 700, True, gkrx.FS_SHA256),
    (25, 2, 13, 700, True, gkrx.FS_KECCAK256),
    (26, 2, 14, 300, True, gkrx.FS_SHA256),
]


def _int_layers(layers):
    return [([tuple(int(x) for x in g) for g in m], [tuple(int(x) for x in g) for g in a]) for m, a in layers]


@pytest.fixture(scope="module")
def ctxs():
    out = {h: gkrx.Context(gkrx.Config(fs_hash=h, field_type=gkrx.FIELD_BN254)) for h in (0, 1)}
    yield out
    for c in out.values():
        c.close()


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"D{c[1]}_k{c[2]}_add{c[3]}_loc{int(c[4])}_h{c[5]}")
def test_bn254_proof_matches_oracle(ctxs, case):
    seed, depth, k, addp, loc, h = case
    print(f"[gpu] bn254 seed={seed} D={depth} k={k} hash={h}", flush=True)
    ctx = ctxs[h]
    layers = gkrx.synth_circuit(seed, depth, k, addp, loc)
    wit = gkrx.synth_witness_bn254(seed, 1 << k)
    circ = gkrx.Circuit.from_layers(ctx, layers)
    circ.set_witness(wit)
    circ.evaluate()
    out = circ.output()
    claimed, proof = gkrx.Prover(ctx).prove(circ)
    F = gg.bn254()
    il = _int_layers(layers)
    wi = [[v] for v in gkrx.bn254_to_ints(wit)]
    ovals = gg.evaluate(F, il, wi)
    assert gkrx.bn254_to_ints(out) == [v[0] for v in ovals[-1]]
    oc, op = gg.prove(F, il, wi, h)
    assert gkrx.bn254_to_ints(claimed) == [oc[0][0]]
    assert len(proof) == len(op)
    assert bytes(proof) == op
    assert gkrx.Verifier(ctx).verify(circ, claimed, proof)
    bad = bytearray(bytes(proof))
    bad[len(bad) // 2] ^= 4
    assert not gkrx.Verifier(ctx).verify(circ, claimed, bytes(bad))
    circ.close()


def test_bn254_repeat_and_batch_equal(ctxs):
    """Proving twice, and through the batch entry point, gives the same bytes."""
    print("[gpu] bn254 repeat + batch", flush=True)
    ctx = ctxs[1]
    circs = []
    for s, d, k in [(31, 3, 9), (32, 2, 12)]:
        c = gkrx.Circuit.from_layers(ctx, gkrx.synth_circuit(s, d, k, 700, True))
        c.set_witness(gkrx.synth_witness_bn254(s, 1 << k))
        circs.append(c)
    prover = gkrx.Prover(ctx)
    single = [prover.prove(c) for c in circs]
    again = [prover.prove(c) for c in circs]
    batch = prover.prove_batch(circs)
    for (c1, p1), (c2, p2), (c3, p3) in zip(single, again, batch):
        assert bytes(p1) == bytes(p2) == bytes(p3)
        assert np.array_equal(c1, c2) and np.array_equal(c1, c3)
    for c in circs:
        c.close()


def test_bn254_rejects_noncanonical_witness(ctxs):
    ctx = ctxs[0]
    circ = gkrx.Circuit.from_layers(ctx, gkrx.synth_circuit(41, 2, 3, 700, True))
    w = gkrx.synth_witness_bn254(41, 8)
    r_words = np.frombuffer(gg.BN254_R.to_bytes(32, "little"), dtype=np.uint32)
    w[8:16] = r_words
    with pytest.raises(gkrx.GkrxError):
        circ.set_witness(w)
    circ.close()


def test_bn254_batch_equals_alone(ctxs):
    """More proofs than the batch has streams, of uneven sizes; every proof and claim equal to
    the proof made alone, twice in a row (lanes and their tables reused)."""
    print("[gpu] bn254 batch vs alone", flush=True)
    ctx = ctxs[1]
    specs = [(60 + i, 2 + i % 3, [4, 12, 9, 14, 6, 10, 11][i]) for i in range(7)]
    circs = []
    for s_, d, k in specs:
        c = gkrx.Circuit.from_layers(ctx, gkrx.synth_circuit(s_, d, k, 700, True))
        c.set_witness(gkrx.synth_witness_bn254(s_, 1 << k))
        circs.append(c)
    prover = gkrx.Prover(ctx)
    alone = [prover.prove(c) for c in circs]
    for _ in range(2):
        batch = prover.prove_batch(circs)
        for (cl, pr), (c1, p1) in zip(batch, alone):
            assert np.array_equal(cl, c1)
            assert bytes(pr) == bytes(p1)
    assert gkrx.Verifier(ctx).verify(circs[3], batch[3][0], batch[3][1])
    for c in circs:
        c.close()
