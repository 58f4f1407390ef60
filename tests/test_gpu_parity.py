"""GPU prover parity: proofs from libgkrx on the MI355X are byte-identical to the oracle's
(which is itself byte-identical to the reference prover, see test_oracle_vs_reference.py
and the golden fixtures), the forward pass matches the oracle's evaluate, and the
library verifier accepts honest proofs and rejects tampered ones."""
import os
import sys

import numpy as np
import pytest

import gkrx
import oracle_lib as ol

pytestmark = pytest.mark.gpu

# (depth, log_width, add_permille, locality, lanes, hash, seed)
CASES = [
    (8, 10, 500, False, 8, gkrx.FS_SHA256, 0),      # the reference's CPU-runnable shape
    (8, 10, 500, False, 8, gkrx.FS_KECCAK256, 0),
    (3, 2, 500, False, 8, gkrx.FS_SHA256, 3),       # smallest valid width (max index 3)
    (2, 2, 500, False, 16, gkrx.FS_SHA256, 4),
    (4, 6, 700, True, 16, gkrx.FS_KECCAK256, 5),
    (3, 12, 700, True, 16, gkrx.FS_SHA256, 6),
    (2, 13, 300, True, 8, gkrx.FS_KECCAK256, 7),
    (5, 9, 1000, False, 16, gkrx.FS_SHA256, 8),     # add gates only
    (5, 9, 0, True, 16, gkrx.FS_KECCAK256, 9),      # mul gates only
]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def contexts():
    ctxs = {}
    yield ctxs
    for c in ctxs.values():
        c.close()


def _ctx(contexts, lanes, fs):
    key = (lanes, fs)
    if key not in contexts:
        contexts[key] = gkrx.Context(gkrx.Config(fs_hash=fs, nb_parallel=lanes), device=0)
    return contexts[key]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"D{c[0]}_k{c[1]}_add{c[2]}_loc{int(c[3])}_L{c[4]}_h{c[5]}")
def test_proof_matches_oracle(case, tmp_path, contexts):
    D, k, addp, loc, L, fs, seed = case
    _say(f"[gpu] prove D={D} k={k} add={addp} loc={loc} L={L} hash={fs} seed={seed}")
    layers = gkrx.synth_circuit(seed, D, k, addp, loc)
    mp, ap = str(tmp_path / "mul.txt"), str(tmp_path / "add.txt")
    gkrx.write_text(layers, mp, ap)
    wit = gkrx.synth_witness(seed, 1 << k, L)
    ctx = _ctx(contexts, L, fs)
    circ = gkrx.Circuit.load_extracted_gates(ctx, mp, ap)
    circ.set_witness(wit)
    circ.evaluate()
    oc = ol.OracleCircuit(mp, ap)
    ovals = oc.evaluate(L, wit)
    off = 0
    for l in range(D + 1):
        got = circ.layer_values(l)
        assert np.array_equal(got, ovals[off:off + got.size]), f"layer {l} values differ"
        off += got.size
    claimed, proof = gkrx.Prover(ctx).prove(circ)
    oclaimed, oproof = oc.prove(L, fs, wit)
    assert len(proof) == len(oproof)
    assert np.array_equal(claimed, oclaimed)
    if bytes(proof) != oproof:
        first = next(i for i in range(len(oproof)) if proof[i] != oproof[i])
        pytest.fail(f"proof differs from oracle at byte {first} of {len(oproof)}")
    v = gkrx.Verifier(ctx)
    assert v.verify(circ, claimed, proof)
    bad = np.array(proof, copy=True)
    bad[-7] ^= 1
    assert not v.verify(circ, claimed, bad)
    bad_claim = claimed.copy()
    bad_claim[0, 0] = (int(bad_claim[0, 0]) + 1) % gkrx.P31
    assert not v.verify(circ, bad_claim, proof)
    circ.close()


def test_width_two_layer_is_rejected(contexts):
    """A layer whose largest index is 1 gets 0 variables from the reference's
    _compute_nb_vars (next_pow_of_2(1) = 1, include/circuit/circuit.hpp:145-146), which
    its prover cannot handle; the library refuses such circuits instead."""
    ctx = _ctx(contexts, 8, gkrx.FS_SHA256)
    layers = gkrx.synth_circuit(3, 2, 1, 500, False)
    with pytest.raises(gkrx.GkrxError):
        gkrx.Circuit.from_layers(ctx, layers)


def test_repeat_prove_is_deterministic(tmp_path, contexts):
    """Proving twice on one context (scratch reuse) gives the same bytes."""
    _say("[gpu] repeat prove determinism")
    layers = gkrx.synth_circuit(21, 4, 11, 700, True)
    ctx = _ctx(contexts, 16, gkrx.FS_KECCAK256)
    circ = gkrx.Circuit.from_layers(ctx, layers)
    circ.set_witness(gkrx.synth_witness(21, 1 << 11, 16))
    p = gkrx.Prover(ctx)
    c1, pr1 = p.prove(circ)
    c2, pr2 = p.prove(circ)
    assert np.array_equal(c1, c2) and bytes(pr1) == bytes(pr2)
    circ.close()


def test_profile_classes_times_only_the_selected_class(contexts):
    """Class-masked timing records the selected kernel class only, with the same launch
    count as timing every class, and the proof bytes do not change."""
    _say("[gpu] class-masked kernel timing")
    layers = gkrx.synth_circuit(22, 3, 14, 700, True)
    ctx = _ctx(contexts, 16, gkrx.FS_KECCAK256)
    circ = gkrx.Circuit.from_layers(ctx, layers)
    circ.set_witness(gkrx.synth_witness(22, 1 << 14, 16))
    p = gkrx.Prover(ctx)
    plain = p.prove(circ)
    lib = gkrx.lib()
    try:
        lib.gkrx_ctx_profile(ctx._h, 1)
        every = p.prove(circ)
        all_stats = gkrx.kernel_stats(ctx)
        idx = gkrx.kernel_class_index(ctx, "k_pair_ph1")
        lib.gkrx_ctx_profile_classes(ctx._h, 1 << idx)
        masked = p.prove(circ)
        one = gkrx.kernel_stats(ctx)
    finally:
        lib.gkrx_ctx_profile(ctx._h, 0)
    assert set(one) == {"k_pair_ph1"} and len(all_stats) > 5
    assert one["k_pair_ph1"]["count"] == all_stats["k_pair_ph1"]["count"] > 0
    for cl, pr in (every, masked):
        assert np.array_equal(cl, plain[0]) and bytes(pr) == bytes(plain[1])
    circ.close()


@pytest.mark.parametrize("lanes,fs", [(16, gkrx.FS_KECCAK256), (8, gkrx.FS_SHA256)])
def test_heavy_gx_rows_match_oracle(lanes, fs, tmp_path, contexts):
    """A g_x row with many entries (300 gates share first input 7 in every layer) puts many
    LDS atomic adds on one row of one tile; the proof must still equal the oracle's."""
    _say(f"[gpu] heavy g_x rows L={lanes} hash={fs}")
    layers = gkrx.synth_circuit(51, 4, 11, 600, True)
    heavy = []
    for mul, add in layers:
        mul = np.array(mul, copy=True).reshape(-1, 4)
        add = np.array(add, copy=True).reshape(-1, 3)
        mul[:150, 0] = 7
        add[:150, 0] = 7
        heavy.append((mul, add))
    mp, ap = str(tmp_path / "mul.txt"), str(tmp_path / "add.txt")
    gkrx.write_text(heavy, mp, ap)
    wit = gkrx.synth_witness(51, 1 << 11, lanes)
    ctx = _ctx(contexts, lanes, fs)
    circ = gkrx.Circuit.load_extracted_gates(ctx, mp, ap)
    circ.set_witness(wit)
    claimed, proof = gkrx.Prover(ctx).prove(circ)
    oclaimed, oproof = ol.OracleCircuit(mp, ap).prove(lanes, fs, wit)
    assert np.array_equal(claimed, oclaimed)
    assert bytes(proof) == oproof
    circ.close()
