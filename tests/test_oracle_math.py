"""Oracle building blocks against brute-force mathematics: eq tables are the products
prod_i (b_i r_i + (1 - b_i)(1 - r_i)) and sum to their factor; eval_multilinear is the
multilinear extension sum_b eq(r, b) f(b); the forward pass is the gate-by-gate sum; and
the first sumcheck message satisfies p(0) + p(1) = claimed value."""
import os
import random

import numpy as np
import pytest

import gkrx
import oracle_lib as ol

P = ol.P31


def eq_brute(r, b, nbits):
    v = 1
    for i in range(nbits):
        bit = (b >> i) & 1
        v = v * (r[i] if bit else (1 - r[i])) % P
    return v


@pytest.mark.parametrize("n", [0, 1, 2, 5, 8])
def test_eq_table_closed_form(n):
    rng = random.Random(n)
    r = [rng.randrange(P) for _ in range(n)]
    fac = rng.randrange(P)
    out = np.zeros(1 << n, dtype=np.uint32)
    ol.oracle().oracle_eq_evals(np.array(r + [0], dtype=np.uint32), n, fac, out)
    for b in range(1 << n):
        assert int(out[b]) == fac * eq_brute(r, b, n) % P
    assert int(out.astype(object).sum()) % P == fac  # sum_b eq(r, b) = 1


@pytest.mark.parametrize("n,L", [(1, 8), (3, 8), (6, 16)])
def test_eval_multilinear_is_the_multilinear_extension(n, L):
    rng = random.Random(100 + n)
    f = np.array([rng.randrange(P) for _ in range((1 << n) * L)], dtype=np.uint32)
    r = [rng.randrange(P) for _ in range(n)]
    out = np.zeros(L, dtype=np.uint32)
    ol.oracle().oracle_eval_multilinear(f, L, np.array(r + [0], dtype=np.uint32), n, out)
    for lane in range(L):
        s = sum(eq_brute(r, b, n) * int(f[b * L + lane]) for b in range(1 << n)) % P
        assert int(out[lane]) == s
    # on Boolean points the extension is the table itself
    b = rng.randrange(1 << n)
    pt = np.array([(b >> i) & 1 for i in range(n)] + [0], dtype=np.uint32)
    ol.oracle().oracle_eval_multilinear(f, L, pt, n, out)
    assert np.array_equal(out, f[b * L:(b + 1) * L])


def test_forward_pass_brute_force(tmp_path):
    L, k, D = 8, 4, 3
    layers = gkrx.synth_circuit(77, D, k, 500, False)
    mp, ap = str(tmp_path / "m.txt"), str(tmp_path / "a.txt")
    gkrx.write_text(layers, mp, ap)
    w = gkrx.synth_witness(77, 1 << k, L)
    vals = ol.OracleCircuit(mp, ap).evaluate(L, w)
    cur = [int(x) for x in w]
    off = 0
    for mul, add in layers:
        assert list(vals[off:off + len(cur)]) == cur
        off += len(cur)
        out = [0] * ((1 << k) * L)
        for i0, i1, o, c in mul.tolist():
            for l in range(L):
                out[o * L + l] = (out[o * L + l] + cur[i0 * L + l] * cur[i1 * L + l] * c) % P
        for i0, o, c in add.tolist():
            for l in range(L):
                out[o * L + l] = (out[o * L + l] + cur[i0 * L + l] * c) % P
        cur = out
    assert list(vals[off:]) == cur


def test_first_sumcheck_message_sums_to_claim(tmp_path):
    """Output layer, rep 0: p(0) + p(1) = claimed_v[0] (alpha = 1, beta = 0), the first
    check of sumcheck_verify_gkr_layer (include/LinearGKR/sumcheck.hpp:152)."""
    L, k, D = 8, 5, 2
    layers = gkrx.synth_circuit(5, D, k, 500, False)
    mp, ap = str(tmp_path / "m.txt"), str(tmp_path / "a.txt")
    gkrx.write_text(layers, mp, ap)
    w = gkrx.synth_witness(5, 1 << k, L)
    claimed, proof = ol.OracleCircuit(mp, ap).prove(L, ol.HASH_SHA256, w)
    S = 4 * L
    at = (1 << k) * S + 32
    p0 = np.frombuffer(proof[at:at + S], dtype=np.uint32)
    p1 = np.frombuffer(proof[at + S:at + 2 * S], dtype=np.uint32)
    assert np.array_equal((p0.astype(np.uint64) + p1) % P, claimed[0].astype(np.uint64))


def test_proof_length_formula(tmp_path):
    """N0*S + 32 + sum_l (6 n_l R + 2R) S (the layout of Prover::prove)."""
    L, k, D = 16, 6, 4
    layers = gkrx.synth_circuit(9, D, k, 700, True)
    mp, ap = str(tmp_path / "m.txt"), str(tmp_path / "a.txt")
    gkrx.write_text(layers, mp, ap)
    _, proof = ol.OracleCircuit(mp, ap).prove(L, ol.HASH_KECCAK256, gkrx.synth_witness(9, 1 << k, L))
    S, R = 4 * L, 3
    assert len(proof) == (1 << k) * S + 32 + D * (6 * k * R + 2 * R) * S
