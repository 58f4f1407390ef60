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
