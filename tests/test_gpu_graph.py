"""HIP-graph replay of gkrx_prove: with graphs on, the device part of a prove is captured once
per circuit and replayed for new witnesses.  Every replay must give the bytes of the eager
prove of the same witness (made on a second context between the replays) and, on the first
and last witness, the oracle's bytes."""
import numpy as np
import pytest

import gkrx
import oracle_lib as ol

pytestmark = pytest.mark.gpu

# (depth, log_width, add_permille, locality, lanes, hash, seed)
SHAPES = [
    (4, 10, 700, True, 16, gkrx.FS_KECCAK256, 31),
    (3, 9, 500, False, 8, gkrx.FS_SHA256, 32),
    (5, 14, 700, True, 16, gkrx.FS_KECCAK256, 33),  # round pairs before the tails
]


def _say(msg):
    print(msg, flush=True)


def _pair(fs, lanes):
    eager = gkrx.Context(gkrx.Config(fs_hash=fs, nb_parallel=lanes), device=0)
    graph = gkrx.Context(gkrx.Config(fs_hash=fs, nb_parallel=lanes), device=0)
    graph.set_graphs(True)
    assert graph.graphs() and not eager.graphs()
    return eager, graph


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"D{s[0]}_k{s[1]}_L{s[4]}_h{s[5]}")
def test_graph_replays_equal_eager_and_oracle(shape, tmp_path):
    D, k, addp, loc, L, fs, seed = shape
    layers = gkrx.synth_circuit(seed, D, k, addp, loc)
    mp, ap = str(tmp_path / "mul.txt"), str(tmp_path / "add.txt")
    gkrx.write_text(layers, mp, ap)
    oc = ol.OracleCircuit(mp, ap)
    eager, graph = _pair(fs, L)
    ce_circ = gkrx.Circuit.load_extracted_gates(eager, mp, ap)
    cg_circ = gkrx.Circuit.load_extracted_gates(graph, mp, ap)
    pe, pg = gkrx.Prover(eager), gkrx.Prover(graph)
    for i in range(4):
        _say(f"[gpu] graph replay {i}: D={D} k={k} L={L} hash={fs} seed={seed}")
        wit = gkrx.synth_witness(seed * 10 + i, 1 << k, L)
        ce_circ.set_witness(wit)
        cg_circ.set_witness(wit)
        claimed_g, proof_g = pg.prove(cg_circ)  # captured at i = 0, replayed after
        claimed_e, proof_e = pe.prove(ce_circ)
        assert np.array_equal(claimed_g, claimed_e), f"claims of replay {i} differ from the eager prove"
        assert bytes(proof_g) == bytes(proof_e), f"proof of replay {i} differs from the eager prove"
        if i in (0, 3):
            oclaimed, oproof = oc.prove(L, fs, wit)
            assert np.array_equal(claimed_g, oclaimed) and bytes(proof_g) == oproof
    ce_circ.close()
    cg_circ.close()
    eager.close()
    graph.close()


def test_graph_follows_circuit_changes():
    """Alternating two circuits on one context recaptures the graph on each change; a larger
    circuit regrows the scratch (dropping the graph); turning graphs off and on again works."""
    eager, graph = _pair(gkrx.FS_KECCAK256, 16)
    shapes = [(41, 4, 11), (42, 3, 12), (41, 4, 11), (43, 6, 13), (42, 3, 12)]
    circs = {}
    for n, (seed, D, k) in enumerate(shapes):
        _say(f"[gpu] graph circuit change {n}: seed={seed} D={D} k={k}")
        if seed not in circs:
            layers = gkrx.synth_circuit(seed, D, k, 700, True)
            circs[seed] = (gkrx.Circuit.from_layers(eager, layers), gkrx.Circuit.from_layers(graph, layers), k)
        ce_circ, cg_circ, kk = circs[seed]
        wit = gkrx.synth_witness(seed + 100 * n, 1 << kk, 16)
        ce_circ.set_witness(wit)
        cg_circ.set_witness(wit)
        if n == 3:
            graph.set_graphs(False)
            graph.set_graphs(True)
        claimed_g, proof_g = gkrx.Prover(graph).prove(cg_circ)
        claimed_e, proof_e = gkrx.Prover(eager).prove(ce_circ)
        assert np.array_equal(claimed_g, claimed_e) and bytes(proof_g) == bytes(proof_e), f"step {n} differs"
    for ce_circ, cg_circ, _ in circs.values():
        ce_circ.close()
        cg_circ.close()
    eager.close()
    graph.close()
