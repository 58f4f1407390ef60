"""CIRCUIT2 circuits on the GPU: loaded from the binary file (gkrx_circuit_load_circuit2),
proved, and compared byte for byte with the oracle's proof of the flattened circuit (whose
flattening the CPU tests pin to the reference's reader)."""
import numpy as np
import pytest

import circuit2_gen as cg
import gkrx
import oracle_lib as ol

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", [(11, 3, 6, 8, 0), (12, 4, 9, 16, 1), (13, 2, 12, 16, 0)],
                         ids=lambda c: f"D{c[1]}_k{c[2]}_L{c[3]}_h{c[4]}")
def test_circuit2_proof_matches_oracle(case, tmp_path):
    seed, depth, k, lanes, h = case
    print(f"[gpu] circuit2 seed={seed} D={depth} k={k} L={lanes} hash={h}", flush=True)
    segs, layers, sentinel = cg.make(seed, depth, k)
    path = str(tmp_path / "c.circuit2")
    gkrx.write_circuit2(path, segs, layers, sentinel)
    mp, ap = str(tmp_path / "m.txt"), str(tmp_path / "a.txt")
    gkrx.circuit2_to_text(path, mp, ap, coef_mode=0)
    ctx = gkrx.Context(gkrx.Config(fs_hash=h, nb_parallel=lanes))
    circ = gkrx.Circuit.load_circuit2(ctx, path)
    wit = gkrx.synth_witness(seed, 1 << k, lanes)
    circ.set_witness(wit)
    claimed, proof = gkrx.Prover(ctx).prove(circ)
    oc, op = ol.OracleCircuit(mp, ap).prove(lanes, h, wit)
    assert bytes(proof) == op and np.array_equal(claimed, oc)
    assert gkrx.Verifier(ctx).verify(circ, claimed, proof)
    ctx.close()


@pytest.mark.parametrize("engine,block,lanes", [(gkrx.BATCH_LOCKSTEP, 64, 1), (gkrx.BATCH_LOCKSTEP, 256, 4),
                                                (gkrx.BATCH_PER_PROOF, 256, 4)],
                         ids=["lockstep_b64_v1", "lockstep_b256_v4", "per_proof"])
def test_circuit2_batch_equals_plain(engine, block, lanes, tmp_path):
    """Outputs with two gates (the CIRCUIT2 circuits have up to about a thousand per layer)
    through the batch engines' forward, g_x and h_y ops: every batched proof equals the
    proof of the same circuit made alone."""
    print(f"[gpu] circuit2 batch engine={engine} block={block} lanes={lanes}", flush=True)
    ctx = gkrx.Context(gkrx.Config(fs_hash=gkrx.FS_KECCAK256, nb_parallel=16))
    circs = []
    for seed, depth, k in [(12, 4, 9), (13, 2, 12), (14, 3, 11)]:
        segs, layers, sentinel = cg.make(seed, depth, k)
        path = str(tmp_path / f"c{seed}.circuit2")
        gkrx.write_circuit2(path, segs, layers, sentinel)
        c = gkrx.Circuit.load_circuit2(ctx, path)
        c.set_witness(gkrx.synth_witness(seed, 1 << k, 16))
        circs.append(c)
    p = gkrx.Prover(ctx)
    alone = [p.prove(c) for c in circs]
    alone = [(np.array(cl), bytes(pr)) for cl, pr in alone]
    ctx.set_batch_engine(engine, block)
    ctx.set_batch_lanes(lanes)
    for _ in range(2):
        res = p.prove_batch(circs)
        for (cl, pr), (ca, pa) in zip(res, alone):
            assert np.array_equal(cl, ca) and bytes(pr) == pa
    ctx.close()
