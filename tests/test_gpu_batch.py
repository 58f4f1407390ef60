"""Batched proving: one launch proves many circuits of uneven widths and depths; every
proof is byte-identical to the same circuit proved alone (and, for small circuits, to the
oracle), and the results do not depend on what else is in the batch."""
import os

import numpy as np
import pytest

import gkrx
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _mixed(seed, count, kmin, kmax, dmin, dmax):
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        k = int(rng.integers(kmin, kmax + 1))
        d = int(rng.integers(dmin, dmax + 1))
        specs.append((seed * 1000 + i, d, k))
    return specs


@pytest.fixture(scope="module")
def ctx():
    c = gkrx.Context(gkrx.Config(fs_hash=gkrx.FS_KECCAK256, nb_parallel=16))
    yield c
    c.close()


def _circuits(ctx, specs):
    out = []
    for s, d, k in specs:
        circ = gkrx.Circuit.from_layers(ctx, gkrx.synth_circuit(s, d, k, 700, True))
        circ.set_witness(gkrx.synth_witness(s, 1 << k, 16))
        out.append(circ)
    return out


def test_batch_equals_single_proofs(ctx):
    print("[gpu] batch of 24 uneven proofs vs one at a time", flush=True)
    specs = _mixed(3, 24, 2, 13, 1, 6)
    circs = _circuits(ctx, specs)
    prover = gkrx.Prover(ctx)
    batch = prover.prove_batch(circs)
    for (claimed, proof), circ in zip(batch, circs):
        c1, p1 = prover.prove(circ)
        assert np.array_equal(claimed, c1)
        assert bytes(proof) == bytes(p1)
        assert gkrx.Verifier(ctx).verify(circ, claimed, proof)
    for c in circs:
        c.close()


def test_batch_matches_oracle_small(ctx, tmp_path):
    print("[gpu] batch vs oracle", flush=True)
    specs = _mixed(4, 6, 3, 8, 2, 4)
    circs = _circuits(ctx, specs)
    batch = gkrx.Prover(ctx).prove_batch(circs)
    for (claimed, proof), (s, d, k) in zip(batch, specs):
        mp, ap = str(tmp_path / f"m{s}.txt"), str(tmp_path / f"a{s}.txt")
        gkrx.write_text(gkrx.synth_circuit(s, d, k, 700, True), mp, ap)
        oc, op = ol.OracleCircuit(mp, ap).prove(16, ol.HASH_KECCAK256, gkrx.synth_witness(s, 1 << k, 16))
        assert np.array_equal(claimed, oc) and bytes(proof) == op
    for c in circs:
        c.close()


def test_batch_wide_layers_equal_single(ctx):
    """Widths up to 2^18 (the largest of the uneven-batch workload) in one batch."""
    print("[gpu] batch of wide proofs (2^16 .. 2^18) vs one at a time", flush=True)
    specs = [(901, 2, 16), (902, 3, 17), (903, 2, 18), (904, 1, 10)]
    circs = _circuits(ctx, specs)
    prover = gkrx.Prover(ctx)
    batch = prover.prove_batch(circs)
    for (claimed, proof), circ in zip(batch, circs):
        c1, p1 = prover.prove(circ)
        assert np.array_equal(claimed, c1)
        assert bytes(proof) == bytes(p1)
    for c in circs:
        c.close()


def test_result_independent_of_batch_mix(ctx):
    """The same circuit proved alone in a batch of one, and inside a larger mix, gives the
    same bytes."""
    print("[gpu] batch mix independence", flush=True)
    specs = _mixed(5, 10, 4, 12, 2, 8)
    circs = _circuits(ctx, specs)
    prover = gkrx.Prover(ctx)
    mix = prover.prove_batch(circs)
    alone = prover.prove_batch([circs[7]])
    assert bytes(mix[7][1]) == bytes(alone[0][1]) and np.array_equal(mix[7][0], alone[0][0])
    for c in circs:
        c.close()


ENGINES = [(gkrx.BATCH_PER_PROOF, 256)] + [(e, b) for e in (gkrx.BATCH_LOCKSTEP, gkrx.BATCH_LOCKSTEP_GROUPS)
                                           for b in (64, 128, 256, 512)]


@pytest.fixture(scope="module")
def mixed_plain(ctx):
    """A mix with grid rounds (k >= 12), tail-only layers and two-variable layers (the
    narrowest the loader accepts), and each circuit's plain proof."""
    specs = _mixed(6, 14, 2, 14, 1, 5) + [(960, 3, 15), (961, 1, 2), (962, 2, 12)]
    circs = _circuits(ctx, specs)
    prover = gkrx.Prover(ctx)
    plain = [prover.prove(c) for c in circs]
    yield circs, plain
    for c in circs:
        c.close()


@pytest.mark.parametrize("engine,block", ENGINES, ids=[f"e{e}_b{b}" for e, b in ENGINES])
def test_every_engine_equals_plain(ctx, mixed_plain, engine, block):
    print(f"[gpu] batch engine {engine} block {block} vs plain proofs", flush=True)
    circs, plain = mixed_plain
    saved = ctx.batch_engine()
    ctx.set_batch_engine(engine, block)
    try:
        batch = gkrx.Prover(ctx).prove_batch(circs)
        again = gkrx.Prover(ctx).prove_batch(circs[::-1])[::-1]  # buffers reused, other order
    finally:
        ctx.set_batch_engine(*saved)
    for (claimed, proof), (c1, p1), (c2, p2) in zip(batch, plain, again):
        assert np.array_equal(claimed, c1) and bytes(proof) == bytes(p1)
        assert np.array_equal(c2, c1) and bytes(p2) == bytes(p1)


def test_autotune_picks_an_engine_and_keeps_results(ctx, mixed_plain):
    print("[gpu] batch autotune", flush=True)
    circs, plain = mixed_plain
    saved = ctx.batch_engine()
    try:
        engine, block, lanes, ms = ctx.batch_autotune(circs)
        assert (engine, block) in ENGINES and ctx.batch_engine() == (engine, block) and ctx.batch_lanes() == lanes
        assert len(ms) == 25 and all(v > 0 for v in ms.values())
        # the first setting (in the order of the dict) within 1% of the fastest
        tied = [k for k, v in ms.items() if v <= min(ms.values()) * 1.01]
        assert (engine, block, lanes) == tied[0]
        batch = gkrx.Prover(ctx).prove_batch(circs)
    finally:
        ctx.set_batch_engine(*saved)
        ctx.set_batch_lanes(4)
    for (claimed, proof), (c1, p1) in zip(batch, plain):
        assert np.array_equal(claimed, c1) and bytes(proof) == bytes(p1)


@pytest.mark.parametrize("engine,block,lanes", [(gkrx.BATCH_LOCKSTEP, 64, 1), (gkrx.BATCH_LOCKSTEP, 128, 2),
                                               (gkrx.BATCH_LOCKSTEP, 256, 1), (gkrx.BATCH_LOCKSTEP, 512, 2),
                                               (gkrx.BATCH_LOCKSTEP_GROUPS, 64, 4), (gkrx.BATCH_LOCKSTEP_GROUPS, 256, 1),
                                               (gkrx.BATCH_LOCKSTEP_GROUPS, 512, 2)])
def test_lanes_per_thread_equal_plain(ctx, mixed_plain, engine, block, lanes):
    print(f"[gpu] lockstep engine {engine} block {block} lanes {lanes} vs plain proofs", flush=True)
    circs, plain = mixed_plain
    saved = ctx.batch_engine()
    ctx.set_batch_engine(engine, block)
    ctx.set_batch_lanes(lanes)
    try:
        batch = gkrx.Prover(ctx).prove_batch(circs)
    finally:
        ctx.set_batch_engine(*saved)
        ctx.set_batch_lanes(4)
    for (claimed, proof), (c1, p1) in zip(batch, plain):
        assert np.array_equal(claimed, c1) and bytes(proof) == bytes(p1)


GRAPH_ENGINE = 3  # the lockstep launches replayed from HIP graphs (engine taken out)


@pytest.mark.skip(reason="HIP-graph replay of the lockstep batch is taken out: its first batch (three "
                         "streams joined by events, captured per segment of steps) ended the process with a "
                         "host segmentation fault inside the library call, cause not found")
def test_graph_replays_follow_new_witnesses(ctx):
    """The graph engine captures a batch's launches once and replays them: every replay,
    with plain proofs (eager launches on the same stream) and new witnesses between the
    replays, must equal the plain proofs."""
    print("[gpu] batch graph replays", flush=True)
    specs = _mixed(11, 6, 3, 14, 1, 5)
    circs = _circuits(ctx, specs)
    prover = gkrx.Prover(ctx)
    saved = ctx.batch_engine()
    ctx.set_batch_engine(GRAPH_ENGINE, 128)
    try:
        for rep in range(3):
            batch = prover.prove_batch(circs)
            for (cl, pr), c in zip(batch, circs):
                c.evaluate()
                c1, p1 = prover.prove(c)
                assert np.array_equal(cl, c1) and bytes(pr) == bytes(p1), f"replay {rep}"
            for (s, d, k), c in zip(specs, circs):
                c.set_witness(gkrx.synth_witness(s + 100 * (rep + 1), 1 << k, 16))
    finally:
        ctx.set_batch_engine(*saved)
    for c in circs:
        c.close()


def test_batch_beyond_the_lockstep_limit(ctx):
    """More than 2048 proofs (the lockstep engine's limit) run on the per-proof engine; every
    proof must still equal its plain proof."""
    print("[gpu] batch of 2100 small proofs", flush=True)
    specs = [(5000 + i, 1 + i % 3, 2 + i % 4) for i in range(2100)]
    circs = _circuits(ctx, specs)
    prover = gkrx.Prover(ctx)
    saved = ctx.batch_engine()
    ctx.set_batch_engine(gkrx.BATCH_LOCKSTEP, 256)
    try:
        batch = prover.prove_batch(circs)
    finally:
        ctx.set_batch_engine(*saved)
    for i in range(0, len(circs), 97):
        circs[i].evaluate()
        c1, p1 = prover.prove(circs[i])
        assert np.array_equal(batch[i][0], c1) and bytes(batch[i][1]) == bytes(p1), i
    for c in circs:
        c.close()


def test_engine_arguments_checked(ctx):
    with pytest.raises(gkrx.GkrxError):
        ctx.set_batch_engine(gkrx.BATCH_LOCKSTEP, 96)
    with pytest.raises(gkrx.GkrxError):
        ctx.set_batch_lanes(3)
    with pytest.raises(gkrx.GkrxError):
        ctx.set_batch_engine(7, 256)


def test_metadata_reuse_follows_witness_and_circuit_changes(ctx):
    """A repeated batch reuses the job table and schedule on the device; the proofs must
    still follow new witnesses and a circuit replaced by another of the same shape."""
    print("[gpu] batch metadata reuse", flush=True)
    specs = _mixed(7, 6, 3, 13, 1, 4)
    circs = _circuits(ctx, specs)
    prover = gkrx.Prover(ctx)
    first = prover.prove_batch(circs)
    for (s, d, k), c in zip(specs, circs):
        c.set_witness(gkrx.synth_witness(s + 77, 1 << k, 16))
    second = prover.prove_batch(circs)
    for (cl, pr), c in zip(second, circs):
        c1, p1 = prover.prove(c)
        assert np.array_equal(cl, c1) and bytes(pr) == bytes(p1)
    assert all(bytes(a[1]) != bytes(b[1]) for a, b in zip(first, second))
    s, d, k = specs[2]
    circs[2].close()
    circs[2] = gkrx.Circuit.from_layers(ctx, gkrx.synth_circuit(s + 5, d, k, 300, False))
    circs[2].set_witness(gkrx.synth_witness(s, 1 << k, 16))
    third = prover.prove_batch(circs)
    for (cl, pr), c in zip(third, circs):
        c1, p1 = prover.prove(c)
        assert np.array_equal(cl, c1) and bytes(pr) == bytes(p1)
    for c in circs:
        c.close()
