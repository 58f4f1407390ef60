"""The serving workload at its own mix, bounded to 10 s: Poisson arrivals at 80% of the
saturated rate; 95% of the requests are M31x16 proofs of the uneven-batch list (widths 2^10
.. 2^18, 4 .. 24 layers, seed 3 + id), 5% BN254 proofs of 2^14 .. 2^16 gates per layer and
8 .. 12 layers; witnesses resident in HBM; a 200 GB admission budget.  Every served proof is
compared with the proof of the same circuit made alone, and must be byte-identical; the
latency percentiles (overall and per kind) and the server's time split are printed."""
import json
import os
import random
import sys
import time

import pytest

import gkrx
import loadgen

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402  (batch_specs: the uneven-batch proof list)

pytestmark = pytest.mark.gpu

M31_POOL = 64     # circuits 0 .. 63 of the uneven-batch list
BN_POOL = 6
SECONDS = 10.0
BUDGET = 200 << 30


def test_c5_poisson_mix_equals_alone():
    mctx = gkrx.Context(gkrx.Config(fs_hash=gkrx.FS_KECCAK256, nb_parallel=16))
    bctx = gkrx.Context(gkrx.Config(fs_hash=gkrx.FS_KECCAK256, field_type=gkrx.FIELD_BN254))
    mcircs, mgates = [], []
    for s, d, k in bench.batch_specs(512)[:M31_POOL]:
        layers = gkrx.synth_circuit(s, d, k, 700, True)
        c = gkrx.Circuit.from_layers(mctx, layers)
        c.set_witness(gkrx.synth_witness(s, 1 << k, 16))
        mcircs.append(c)
        mgates.append(sum(len(m) + len(a) for m, a in layers))
    rng = random.Random(5)
    bcircs, bgates = [], []
    for i in range(BN_POOL):
        k, d = rng.randint(14, 16), rng.randint(8, 12)
        layers = gkrx.synth_circuit(20000 + i, d, k, 700, True)
        c = gkrx.Circuit.from_layers(bctx, layers)
        c.set_witness(gkrx.synth_witness_bn254(20000 + i, 1 << k))
        bcircs.append(c)
        bgates.append(sum(len(m) + len(a) for m, a in layers))
    print(f"[gpu] c5: {M31_POOL} M31x16 + {BN_POOL} BN254 circuits resident", flush=True)
    mprover, bprover = gkrx.Prover(mctx), gkrx.Prover(bctx)
    malone = {i: mprover.prove(c) for i, c in enumerate(mcircs)}
    t0 = time.perf_counter()
    balone = {i: bprover.prove(c) for i, c in enumerate(bcircs)}
    bn_s = (time.perf_counter() - t0) / BN_POOL
    # saturated M31 rate: one batch of the whole pool
    mprover.prove_batch(mcircs)
    t0 = time.perf_counter()
    mprover.prove_batch(mcircs)
    m31_s = (time.perf_counter() - t0) / M31_POOL
    rate = 0.8 / (0.95 * m31_s + 0.05 * bn_s)
    print(f"[gpu] c5: {rate:.0f} requests/s for {SECONDS}s (M31 {m31_s * 1e3:.2f} ms/proof batched, "
          f"BN254 {bn_s * 1e3:.1f} ms/proof alone)", flush=True)
    msrv = gkrx.Server(mctx, BUDGET, 2)
    bsrv = gkrx.Server(bctx, BUDGET // 8, 2)
    try:
        res = loadgen.run([loadgen.Kind("m31", msrv, mcircs, mgates, 0.95, malone),
                           loadgen.Kind("bn254", bsrv, bcircs, bgates, 0.05, balone)], rate, SECONDS, seed=7)
        mst, bst = msrv.stats(), bsrv.stats()
    finally:
        msrv.close()
        bsrv.close()
    res["m31_peak_admitted_gb"] = round(mst["peak_bytes"] / (1 << 30), 3)
    res["rate_per_s"] = round(rate, 1)
    print("[gpu] c5 result " + json.dumps(res), flush=True)
    assert not res["errors"], res["errors"]
    assert res["completed"] == res["requests"] > 0
    assert res["checked_vs_alone"] == res["completed"]
    assert res["mismatches_vs_alone"] == 0
    assert res["per_kind"]["bn254"]["requests"] > 0 and res["per_kind"]["m31"]["requests"] > 0
    assert mst["peak_bytes"] <= BUDGET and bst["peak_bytes"] <= BUDGET // 8
    for c in mcircs + bcircs:
        c.close()
    mctx.close()
    bctx.close()
