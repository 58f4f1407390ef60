"""Benchmark: proved gates/s of the MI355X GKR prover, batched against one proof per launch.

Headline workload (the `value` line): the uneven batch, 512 independent M31x16 proofs;
proof i has its own seed 3 + i, width 2^k with k uniform in [10, 18] and depth uniform in
[4, 24] (drawn from a generator seeded with 3), 70% add / 30% mul, 80% of input indices
within +-1024 of the output index, RAW commitment, Keccak-256 transcript.  A step is one
batched prove of the whole batch: the host transcript prefixes (commitment hashes,
grinding, output challenges) on host threads, overlapped with every proof's forward pass,
layer sumchecks and on-device transcript, and every proof's bytes back on the host.  The
witnesses are resident in HBM before timing starts.  Beside it: the same 512 proofs one
gkrx_prove at a time (the plain path), the serving loop under Poisson load, the single
24 x 2^20 M31x16 proof, the 12 x 2^18 BN254 proof, and the reference CPU prover on the
host cores.

Prints one JSON line (the first), then diagnostics on stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "expander-cpp_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

from loadgen import percentile  # noqa: E402

HBM_PEAK_GBS = 8000.0
TAIL_ROWS = 2048  # tables with at most this many rows end a phase in the small-table kernels

# HBM traffic per dispatch measured with two rocprofv3 --pmc runs (FETCH_SIZE, then
# WRITE_SIZE), each scaled by what the counter reads for a 256 MiB device-to-device copy
# (--calibration-copy-mb 256: bytes read = 2 x FETCH_SIZE KiB, written = 1 x WRITE_SIZE KiB).
# Keys: the workload and the batch engine setting the profiled run used.
# k_pair<16, phase 1, double-fold> of the single 24 x 2^20 proof
# (profiles/pmc_traffic_m31x16_24x2e20_keccak256.md):
SINGLE_PAIR_TRAFFIC = {(24, 20, 16): 137754769}
# BN254 fused round k_bn_round<true> of the 12 x 2^18 proof: int32 VALU wave instructions per
# dispatch (rocprofv3 --pmc SQ_INSTS_VALU_INT32, averaged over the proof's dispatches of both
# phases; profiles/valu_bn254_12x2e18_keccak256.md).  The int32 VALU peak is derived from the
# published vector FP32 rate: 157.3 TFLOP/s counts a packed FMA (2 lanes x 2 flops) per lane
# per cycle, so unpacked 32-bit lane operations peak at 157.3 / 4 = 39.3 T/s (256 CUs x 64
# lanes x 2.4 GHz).
BN254_ROUND_INT32 = {(12, 18): 459232}
INT32_VALU_PEAK_TOPS = 39.32
# k_ls_pair of the 512-proof batch (profiles/pmc_ls_pair_m31x16_512_uneven_keccak256.md):
# {(proofs, engine, block, lanes per thread): bytes per dispatch}; engine 4 (proof groups on
# their own streams) launches about three times as many, smaller, round-pair kernels
BATCH_PAIR_TRAFFIC = {
    (512, 2, 64, 1): 1252212537, (512, 2, 64, 2): 1215790856, (512, 2, 64, 4): 1213910904,
    (512, 2, 128, 1): 1215148902, (512, 2, 128, 2): 1188869878, (512, 2, 128, 4): 1179831817,
    (512, 2, 256, 1): 1184559781, (512, 2, 256, 2): 1170093154, (512, 2, 256, 4): 1167763312,
    (512, 2, 512, 1): 1171026623, (512, 2, 512, 2): 1163177863, (512, 2, 512, 4): 1164018785,
    (512, 4, 64, 1): 422053602, (512, 4, 64, 2): 411957245, (512, 4, 64, 4): 411074671,
    (512, 4, 128, 1): 411592693, (512, 4, 128, 2): 404512912, (512, 4, 128, 4): 402235282,
    (512, 4, 256, 1): 402466210, (512, 4, 256, 2): 398174594, (512, 4, 256, 4): 397908388,
    (512, 4, 512, 1): 398239225, (512, 4, 512, 2): 395819784, (512, 4, 512, 4): 397058799,
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# ----------------------------------------------------------------- algorithmic bytes
def pair_count(n):
    """Round pairs of a phase of a layer with n input variables: rounds run in pairs while
    the table has more than TAIL_ROWS rows; K is the first round whose table has at most
    TAIL_ROWS rows, rounded up to even, and the pairs are (0, 1), (2, 3), .., (K - 2, K - 1)."""
    k0 = 0
    while k0 < n and (1 << (n - k0)) > TAIL_ROWS:
        k0 += 1
    return ((k0 + 1) & ~1) // 2


def pair_task_bytes(n, R, S, phase, j):
    """Algorithmic HBM bytes of round pair j of one phase of a layer with n input variables,
    for all R repetitions.  f is the layer input (shared by the repetitions, counted once)
    for pairs 0 and 1, then table 2j - 2 of each repetition; h is the lane table hg (S bytes
    a row) in phase 1 and the scalar table s (4 bytes) in phase 2.  Pair 0 reads table 0;
    pair j >= 1 reads table 2j - 2 (N_k * 4 rows, N_k = 2^(n - 2j)) and writes table 2j
    (N_k rows) of f and h."""
    N = 1 << n
    hs = S if phase == 1 else 4
    if j == 0:
        return N * S + R * N * hs
    if j == 1:
        return N * S + R * N * hs + R * (N >> 2) * (S + hs)
    nk = N >> (2 * j)
    return R * 5 * nk * (S + hs)


def circuit_pair_bytes(nvars, R, S, phases=(1, 2), first=True):
    """Sum of pair_task_bytes over the layers (nb_input_vars list) of a circuit;
    first=False leaves out pair 0 of each phase."""
    return sum(pair_task_bytes(n, R, S, ph, j) for n in nvars for ph in phases
               for j in range(0 if first else 1, pair_count(n)))


# ----------------------------------------------------------------- host facts
def host_cpu():
    """Model name, physical cores and threads of the host (lscpu), and the CPUs this process
    may run on."""
    info = {"cpus_usable": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, check=True).stdout
        kv = dict(re.findall(r"^([^:\n]+):\s*(.+)$", out, flags=re.M))
        info["model"] = kv.get("Model name", "").strip()
        sockets = int(kv.get("Socket(s)", "1"))
        cores = int(kv.get("Core(s) per socket", "0"))
        info["physical_cores"] = sockets * cores
        info["threads_per_core"] = int(kv.get("Thread(s) per core", "1"))
    except (OSError, subprocess.CalledProcessError, ValueError):
        pass
    return info


# ----------------------------------------------------------------- the uneven batch
def batch_specs(count, seed=3):
    """The uneven batch: proof i has its own seed (3 + i), width 2^k with k uniform in
    [10, 18] and depth uniform in [4, 24] (drawn from a generator seeded with 3)."""
    rng = np.random.default_rng(seed)
    ks = rng.integers(10, 19, size=count)
    ds = rng.integers(4, 25, size=count)
    return [(seed + i, int(ds[i]), int(ks[i])) for i in range(count)]


def build_m31(gkrx, ctx, specs):
    from concurrent.futures import ThreadPoolExecutor

    def make(spec):
        s, d, k = spec
        layers = gkrx.synth_circuit(s, d, k, 700, True)
        c = gkrx.Circuit.from_layers(ctx, layers)
        c.set_witness(gkrx.synth_witness(s, 1 << k, 16))
        return c, sum(len(m) + len(a) for m, a in layers)

    with ThreadPoolExecutor(max_workers=16) as ex:
        made = list(ex.map(make, specs))
    return [c for c, _ in made], [g for _, g in made]


def batch_leg(gkrx, fs, count, steps, warmup, plain=True, fixed=None):
    """The headline: the uneven batch proved by the autotuned batch engine, K timed steps of
    one batched prove each, against the plain path (gkrx_prove once per proof) on the same
    proofs.  The round-pair kernel k_ls_pair (the batch's critical lane) is timed inside
    the timed steps by HIP events that each of its launches records itself."""
    L = 16
    specs = batch_specs(count)
    ctx = gkrx.Context(gkrx.Config(fs_hash=fs, nb_parallel=L), device=0)
    t0 = time.perf_counter()
    circs, gate_list = build_m31(gkrx, ctx, specs)
    gates = sum(gate_list)
    log(f"batch of {count} proofs, {gates} gates, built in {time.perf_counter() - t0:.1f}s")
    prover = gkrx.Prover(ctx)
    R = ctx.config.get_num_repetitions()
    outs = [np.empty(prover.proof_size(c), dtype=np.uint8) for c in circs]
    names = {gkrx.BATCH_PER_PROOF: "per_proof", gkrx.BATCH_LOCKSTEP: "lockstep",
             gkrx.BATCH_LOCKSTEP_GROUPS: "lockstep_groups"}
    if fixed:  # (engine, block, lanes) of a lockstep engine, no autotuning
        engine, block, lanes = fixed
        ctx.set_batch_engine(engine, block)
        ctx.set_batch_lanes(lanes)
        tuned = None
    else:
        engine, block, lanes, tune_ms = ctx.batch_autotune(circs)
        tuned = {f"{names[e]}_b{b}_v{v}": round(ms, 3) for (e, b, v), ms in tune_ms.items()}
    log(f"autotune: {tuned} -> {names[engine]} block {block} lanes {lanes}")
    for _ in range(warmup):
        prover.prove_batch(circs, outs)
    dom = "k_ls_work:pair"  # the class of k_ls_pair launches
    gkrx.lib().gkrx_ctx_profile_classes(ctx._h, 1 << gkrx.kernel_class_index(ctx, dom))
    splits = []
    t0 = time.perf_counter()
    for _ in range(steps):
        res = prover.prove_batch(circs, outs)
        splits.append(prover.last_timing())
    batched_s = (time.perf_counter() - t0) / steps
    st = gkrx.kernel_stats(ctx).get(dom, {"count": 0, "ms": 0.0})
    gkrx.lib().gkrx_ctx_profile_classes(ctx._h, 0)
    # the plain path on the same proofs, one gkrx_prove per proof; its proofs must equal
    # the batched ones
    plain_lat = []
    same = None
    buf = np.empty(max(o.size for o in outs), dtype=np.uint8)
    t0 = time.perf_counter()
    for i, c in enumerate(circs if plain else []):
        t1 = time.perf_counter()
        c.evaluate()
        cl, pr = prover.prove(c, out=buf)
        plain_lat.append((time.perf_counter() - t1) * 1e3)
        same = (same is not False) and bool(np.array_equal(cl, res[i][0])) and pr.size == res[i][1].size and \
            bool(np.array_equal(pr, res[i][1]))
    plain_s = time.perf_counter() - t0 if plain else None
    ok = all(gkrx.Verifier(ctx).verify(circs[i], res[i][0], res[i][1]) for i in range(0, count, max(1, count // 8)))
    plain_txt = f"{plain_s * 1e3:.1f} ms" if plain else "not run"
    log(f"batched {batched_s * 1e3:.1f} ms, plain {plain_txt}, identical={same}, verified={ok}")
    roof = None
    if engine in (gkrx.BATCH_LOCKSTEP, gkrx.BATCH_LOCKSTEP_GROUPS) and st["count"]:
        nv = [ctx_n for c in circs for ctx_n in c.nb_input_vars]
        algo = circuit_pair_bytes(nv, R, 4 * L) * steps / st["count"]
        avg_ms = st["ms"] / st["count"]
        achieved = algo / (avg_ms * 1e-3) / 1e9
        traffic = BATCH_PAIR_TRAFFIC.get((count, engine, block, lanes))
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_frac": round(traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
                "kernel": f"k_ls_pair<16,{block},{lanes}> (round pairs of every proof in a step)",
                "algorithmic_bytes_per_launch": int(algo), "avg_launch_us": round(avg_ms * 1e3, 2),
                "launches_per_step": st["count"] // steps, "pair_ms_per_step": round(st["ms"] / steps, 3)}
    out = {"proofs": count, "gates": gates, "batched_gates_per_s": gates / batched_s,
           "batched_ms": batched_s * 1e3, "plain_gates_per_s": gates / plain_s if plain else None,
           "plain_ms": plain_s * 1e3 if plain else None, "ratio": plain_s / batched_s if plain else None,
           "plain_latency_ms": {"p50": round(percentile(plain_lat, 50), 3),
                                "p99": round(percentile(plain_lat, 99), 3)} if plain else None,
           "identical_to_plain": same, "verified_sample": ok, "engine": names[engine], "block": block,
           "lanes_per_thread": lanes, "autotune_ms": tuned,
           "batched_split_ms": {k: round(float(np.mean([s[k] for s in splits])), 3) for k in splits[0]},
           "roofline": roof}
    # the context and circuits stay resident for the serving leg (the caller closes ctx)
    return out, (ctx, circs, gate_list, batched_s / count)


def serve_leg(gkrx, fs, m31_state, seconds, budget_gb, slots, bn_slots):
    """Serving under Poisson arrivals: 95% of the requests are M31x16 proofs drawn from the
    uneven-batch circuits (witnesses resident in HBM), 5% BN254 proofs of 2^14 .. 2^16 gates
    per layer and 8 .. 12 layers.  Two servers (one per field) share the GPU; arrivals come
    at 80% of the saturated rate of that mix (batched M31 time per proof and the BN254
    proof time measured alone).  A request's latency runs from its arrival to its proof on
    the host.  Completed proofs are checked against the proof of the same circuit made
    alone (every BN254 circuit, every 37th M31 circuit)."""
    import random

    import loadgen

    mctx, mcircs0, mgates0, m31_s = m31_state
    # a second resident copy of every M31 circuit (same wiring and witness), so that a
    # circuit with a request in flight does not hold back arrivals
    clones, _ = build_m31(gkrx, mctx, batch_specs(len(mcircs0)))
    mcircs, mgates = mcircs0 + clones, mgates0 + mgates0
    n0 = len(mcircs0)
    bctx = gkrx.Context(gkrx.Config(fs_hash=fs, field_type=gkrx.FIELD_BN254), device=0)
    rng = random.Random(5)
    bcircs, bgates = [], []
    for i in range(16):
        k, d = rng.randint(14, 16), rng.randint(8, 12)
        layers = gkrx.synth_circuit(20000 + i, d, k, 700, True)
        c = gkrx.Circuit.from_layers(bctx, layers)
        c.set_witness(gkrx.synth_witness_bn254(20000 + i, 1 << k))
        bcircs.append(c)
        bgates.append(sum(len(m) + len(a) for m, a in layers))
    mprover, bprover = gkrx.Prover(mctx), gkrx.Prover(bctx)
    check_m = {}
    for i in range(0, n0, 37):
        check_m[i] = check_m[i + n0] = mprover.prove(mcircs[i])
    t0 = time.perf_counter()
    check_b = {i: bprover.prove(c) for i, c in enumerate(bcircs)}
    bn_s = (time.perf_counter() - t0) / len(bcircs)
    rate = 0.8 / (0.95 * m31_s + 0.05 * bn_s)
    log(f"serving: {rate:.0f} requests/s for {seconds}s (M31 batched {m31_s * 1e3:.2f} ms/proof, "
        f"BN254 {bn_s * 1e3:.1f} ms/proof alone)")
    budget = int(budget_gb * (1 << 30))
    msrv, bsrv = gkrx.Server(mctx, budget, slots), gkrx.Server(bctx, budget // 8, bn_slots)
    try:
        res = loadgen.run([loadgen.Kind("m31", msrv, mcircs, mgates, 0.95, check_m),
                           loadgen.Kind("bn254", bsrv, bcircs, bgates, 0.05, check_b)], rate, seconds, log=log)
        ms, bs = msrv.stats(), bsrv.stats()
    finally:
        msrv.close()
        bsrv.close()
    bctx.close()
    for c in clones:
        c.close()
    lat = res["latency_ms"]
    log(f"served {res['requests']} requests in {res['elapsed_s']:.1f}s, p50 {lat['p50']:.1f} ms, "
        f"p99 {lat['p99']:.1f} ms; per kind {json.dumps({k: v['latency_ms'] for k, v in res['per_kind'].items()})}")
    res.update({"workload": f"Poisson arrivals for {seconds}s at 80% of the saturated rate of the mix; 95% M31x16 "
                            f"proofs of the uneven batch, 5% BN254 proofs (2^14..2^16 gates x 8..12 layers); HBM "
                            f"budget {budget_gb} GB; {slots} M31 batch slots, {bn_slots} BN254 slots; two resident "
                            f"copies of each M31 circuit",
                "rate_per_s": round(rate, 1), "m31_batches": ms["batches"], "m31_largest_batch": ms["max_batch"],
                "m31_peak_admitted_gb": round(ms["peak_bytes"] / (1 << 30), 2), "bn254_proofs": bs["completed"]})
    return res


def single_leg(gkrx, fs, hash_name, D, k, steps, warmup):
    """One large M31x16 proof per launch sequence: D layers x 2^k gates, seed 1.  The phase-1
    double-fold pair kernel (the proof's dominant bandwidth kernel) is timed inside the
    timed steps by events its launches record."""
    L, seed = 16, 1
    log(f"generating circuit {D} x 2^{k}")
    layers = gkrx.synth_circuit(seed, D, k, 700, True)
    gates = sum(len(m) + len(a) for m, a in layers)
    ctx = gkrx.Context(gkrx.Config(fs_hash=fs, nb_parallel=L), device=0)
    circ = gkrx.Circuit.from_layers(ctx, layers)
    circ.set_witness(gkrx.synth_witness(seed, 1 << k, L))
    prover = gkrx.Prover(ctx)
    R = ctx.config.get_num_repetitions()
    buf = np.empty(prover.proof_size(circ), dtype=np.uint8)
    for _ in range(warmup):
        circ.evaluate()
        prover.prove(circ, out=buf)
    dom = "k_pair_ph1"
    gkrx.lib().gkrx_ctx_profile_classes(ctx._h, 1 << gkrx.kernel_class_index(ctx, dom))
    lat, parts = [], []
    t_start = time.perf_counter()
    for _ in range(steps):
        t0 = time.perf_counter()
        circ.evaluate()
        claimed, proof = prover.prove(circ, out=buf)
        lat.append((time.perf_counter() - t0) * 1e3)
        parts.append(prover.last_timing())
    elapsed = time.perf_counter() - t_start
    st = gkrx.kernel_stats(ctx).get(dom, {"count": 0, "ms": 0.0})
    # per-kernel breakdown: two more proofs with every launch timed (outside the timed steps)
    gkrx.lib().gkrx_ctx_profile(ctx._h, 1)
    for _ in range(2):
        circ.evaluate()
        prover.prove(circ, out=buf)
    stats = gkrx.kernel_stats(ctx)
    gkrx.lib().gkrx_ctx_profile(ctx._h, 0)
    ok = gkrx.Verifier(ctx).verify(circ, claimed, proof)
    log(f"single proof: {elapsed * 1e3 / steps:.1f} ms, verifier {'accepted' if ok else 'REJECTED'}")
    roof = None
    if st["count"]:
        per_launch = circuit_pair_bytes(circ.nb_input_vars, R, 4 * L, phases=(1,), first=False) * steps / st["count"]
        avg_ms = st["ms"] / st["count"]
        achieved = per_launch / (avg_ms * 1e-3) / 1e9
        traffic = SINGLE_PAIR_TRAFFIC.get((D, k, L))
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_frac": round(traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
                "algorithmic_bytes_per_launch": int(per_launch), "kernel": "k_pair<16,phase1,double-fold>",
                "avg_launch_us": round(avg_ms * 1e3, 2), "launches": st["count"]}
    ctx.close()
    return {"workload": f"single M31x16 proof, one proof per launch sequence: {D} layers x 2^{k} gates, 70% add / "
                        f"30% mul, locality wiring, RAW commitment, {hash_name} transcript, seed 1",
            "gates": gates, "gates_per_s": round(gates * steps / elapsed, 1), "ms_per_proof": round(elapsed * 1e3 / steps, 3),
            "latency_ms": {"p50": round(percentile(lat, 50), 3), "p99": round(percentile(lat, 99), 3)},
            "breakdown_ms": {k2: round(float(np.mean([p[k2] for p in parts])), 3) for k2 in parts[0]},
            "kernels_ms_per_proof": {k2: round(v["ms"] / 2, 3) for k2, v in sorted(stats.items())},
            "roofline": roof, "verified": bool(ok)}


def bn254_leg(gkrx, fs, depth, log_width, steps, warmup):
    """One BN254 scalar-field proof (one lane, one repetition): depth layers of 2^log_width
    gates, 70% add / 30% mul, locality wiring, witness uniform in [0, r), seed 2."""
    seed = 2
    layers = gkrx.synth_circuit(seed, depth, log_width, 700, True)
    gates = sum(len(m) + len(a) for m, a in layers)
    ctx = gkrx.Context(gkrx.Config(fs_hash=fs, field_type=gkrx.FIELD_BN254), device=0)
    circ = gkrx.Circuit.from_layers(ctx, layers)
    circ.set_witness(gkrx.synth_witness_bn254(seed, 1 << log_width))
    prover = gkrx.Prover(ctx)
    buf = np.empty(prover.proof_size(circ), dtype=np.uint8)
    for _ in range(warmup):
        circ.evaluate()
        prover.prove(circ, out=buf)
    lat, parts = [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        circ.evaluate()
        claimed, proof = prover.prove(circ, out=buf)
        lat.append((time.perf_counter() - t0) * 1e3)
        parts.append(prover.last_timing())
    # the fused rounds k_bn_round<true> of phase 1, timed by events around their launches in
    # proofs of their own (the events lengthen a proof, so the steps above run without them)
    dom = "k_round_ph1"
    gkrx.lib().gkrx_ctx_profile_classes(ctx._h, 1 << gkrx.kernel_class_index(ctx, dom))
    for _ in range(steps):
        circ.evaluate()
        prover.prove(circ, out=buf)
    st = gkrx.kernel_stats(ctx).get(dom, {"count": 0, "ms": 0.0})
    gkrx.lib().gkrx_ctx_profile_classes(ctx._h, 0)
    roof = None
    insts = BN254_ROUND_INT32.get((depth, log_width))
    if st["count"] and insts:
        avg_s = st["ms"] / st["count"] * 1e-3
        achieved = insts * 64 / avg_s / 1e12
        roof = {"bound": "valu (int32)", "achieved": round(achieved, 3), "peak": INT32_VALU_PEAK_TOPS,
                "unit": "T int32 lane-ops/s", "frac": round(achieved / INT32_VALU_PEAK_TOPS, 4),
                "kernel": "k_bn_round<true> (fused BN254 round: 4 folds and 3 products, 8x8-limb CIOS each)",
                "int32_wave_insts_per_launch": insts, "avg_launch_us": round(avg_s * 1e6, 2),
                "launches": st["count"]}
    # per-kernel breakdown from one more proof with every launch timed
    gkrx.lib().gkrx_ctx_profile(ctx._h, 1)
    circ.evaluate()
    prover.prove(circ, out=buf)
    stats = gkrx.kernel_stats(ctx)
    gkrx.lib().gkrx_ctx_profile(ctx._h, 0)
    ok = gkrx.Verifier(ctx).verify(circ, claimed, proof)
    ctx.close()
    ms = float(np.mean(lat))
    log(f"bn254 {depth} x 2^{log_width}: {ms:.1f} ms per proof, verified={ok}")
    return {"workload": f"one BN254 scalar-field proof, {depth} layers x 2^{log_width} gates, 70% add / 30% mul, "
                        f"locality wiring, witness uniform in [0, r), seed 2", "gates": gates,
            "gates_per_s": round(gates / (ms * 1e-3), 1), "ms_per_proof": round(ms, 3),
            "proof_bytes": int(proof.size),
            "breakdown_ms": {k2: round(float(np.mean([p[k2] for p in parts])), 3) for k2 in parts[0]},
            "kernels_ms_per_proof": {k2: round(v["ms"], 3) for k2, v in sorted(stats.items())},
            "roofline": roof, "verified": bool(ok)}


# ----------------------------------------------------------------- CPU baseline
def ref_throughput(fs, specs, threads):
    """The reference prover (oracle/_ref, built from the reference sources) in its harness's
    throughput mode (bench_keccak.cpp:32-42, 82-92): `threads` threads of one process take the
    listed circuits one at a time; each circuit is built and evaluated, then prepare_mem +
    prove is timed.  Returns (gates, slowest thread's busy seconds, wall seconds), or None
    without the reference build."""
    import oracle_lib as ol

    if not ol.ref_available(16, fs):
        return None
    cmd = [ol.ref_binary(16, fs), "throughput", str(threads), "700", "1"] + [f"{s}:{d}:{k}" for s, d, k in specs]
    out = subprocess.run(cmd, check=True, capture_output=True, text=True).stdout.split()
    return int(out[0]), float(out[1]), float(out[2])


def cpu_baselines(fs, count, single_shape):
    host = host_cpu()
    threads = max(1, min(16, host.get("cpus_usable", 1)))
    desc_cpu = f"{host.get('model', 'host CPU')}, {host.get('physical_cores', '?')} physical cores"
    base = all_cores = single = None
    t0 = time.time()
    r = ref_throughput(fs, batch_specs(count), threads)
    if r:
        gates, busy, wall = r
        log(f"cpu baseline (reference, {threads} threads, {count} proofs): {gates} gates, slowest thread "
            f"{busy:.1f}s busy, wall {wall:.1f}s")
        base = {"value": gates / busy, "unit": "gates/s", "cores": threads, "kind": "reference",
                "sample": f"the whole {count}-proof uneven batch (every proof of the workload), the reference "
                          f"harness's throughput mode: {threads} threads of one process take one proof at a time "
                          f"(built and evaluated untimed, prepare_mem + prove timed); value = gates / the slowest "
                          f"thread's prove time; {desc_cpu}",
                "wall_s": round(wall, 2), "host": host}
    if single_shape:
        D, k = single_shape
        r = ref_throughput(fs, [(1, D, k)], 1)
        if r:
            gates, busy, wall = r
            log(f"cpu baseline (reference, 1 thread, {D} x 2^{k}): prove {busy:.1f}s")
            single = {"value": gates / busy, "unit": "gates/s", "cores": 1, "kind": "reference",
                      "sample": f"the full {D} x 2^{k} proof on one core (prepare_mem + prove, evaluate excluded "
                                f"as in the reference harness); {desc_cpu}", "prove_s": round(busy, 3)}
    log(f"cpu baselines took {time.time() - t0:.1f}s")
    return base, single


def calibration_copy(mb):
    """Device-to-device copies of mb MiB (3 times) through the HIP runtime's copy kernel:
    the known-size reference that scales the FETCH_SIZE / WRITE_SIZE counters of a
    rocprofv3 --pmc run into bytes."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_size_t(mb << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), n) or hip.hipMalloc(ctypes.byref(b), n):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, n)
    for _ in range(3):
        hip.hipMemcpy(b, a, n, 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"calibration: 3 device copies of {mb} MiB")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed batched proves of the uneven batch")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=512, help="proofs in the uneven batch")
    ap.add_argument("--hash", choices=["keccak256", "sha256"], default="keccak256")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain (one proof per launch) pass")
    ap.add_argument("--fixed-engine", type=str, default="",
                    help="ENGINE,BLOCK,LANES: this lockstep engine (2 lockstep, 4 lockstep with proof groups) with "
                         "this block size and lanes per thread, no autotuning")
    ap.add_argument("--single-layers", type=int, default=24, help="layers of the single-proof leg (0 = skip)")
    ap.add_argument("--single-log-width", type=int, default=20)
    ap.add_argument("--single-steps", type=int, default=5)
    ap.add_argument("--serve-seconds", type=float, default=20.0, help="serving leg duration (0 = skip)")
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--serve-slots", type=int, default=2)
    ap.add_argument("--serve-bn254-slots", type=int, default=2)
    ap.add_argument("--bn254-layers", type=int, default=12, help="layers of the BN254 leg (0 = skip)")
    ap.add_argument("--bn254-log-width", type=int, default=18)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibration-copy-mb", type=int, default=0,
                    help="also run device copies of this size (for scaling rocprofv3 --pmc byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path here (independent proofs would "
              "shard by proof id, one process per GPU, with no communication)", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1:
        print("--steps must be at least 1", file=sys.stderr)
        sys.exit(2)
    fixed = tuple(int(x) for x in args.fixed_engine.split(",")) if args.fixed_engine else None

    import gkrx

    if args.calibration_copy_mb > 0:
        calibration_copy(args.calibration_copy_mb)
    fs = gkrx.FS_KECCAK256 if args.hash == "keccak256" else gkrx.FS_SHA256

    # --batch 0 leaves out the headline (profiling runs of the other legs)
    batch = serve = None
    if args.batch > 0:
        batch, m31_state = batch_leg(gkrx, fs, args.batch, args.steps, args.warmup, not args.no_plain, fixed)
        if args.serve_seconds > 0:
            serve = serve_leg(gkrx, fs, m31_state, args.serve_seconds, args.serve_budget_gb, args.serve_slots,
                              args.serve_bn254_slots)
        m31_state[0].close()
    single = None
    if args.single_layers > 0:
        single = single_leg(gkrx, fs, args.hash, args.single_layers, args.single_log_width, args.single_steps, 2)
    bn254 = None
    if args.bn254_layers > 0:
        bn254 = bn254_leg(gkrx, fs, args.bn254_layers, args.bn254_log_width, 2, 1)
    cpu = None
    if not args.no_cpu_baseline:
        try:
            shape = (args.single_layers, args.single_log_width) if single else None
            cpu, cpu_single = cpu_baselines(fs, args.batch, shape) if args.batch > 0 else (None, None)
            if single is not None:
                single["cpu_baseline"] = cpu_single
        except Exception as e:  # reported, never fatal
            log(f"cpu baseline failed: {e}")
    if batch is None:
        print(json.dumps({"metric": "proved gates/s (M31x16) batched vs one-proof-per-launch; p50/p99 proof latency",
                          "value": None, "single_proof": single, "bn254": bn254}), flush=True)
        return
    value = batch["batched_gates_per_s"]
    line = {
        "metric": "proved gates/s (M31x16) batched vs one-proof-per-launch; p50/p99 proof latency",
        "value": round(value, 1), "unit": "gates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(batch["batched_ms"], 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32 (M31 field)", "data": "synthetic",
        "config": {"workload": f"uneven batch: {args.batch} independent M31x16 proofs, width 2^k with k uniform in "
                               f"[10, 18], depth uniform in [4, 24], 70% add / 30% mul, locality wiring, RAW "
                               f"commitment, {args.hash} transcript, seed 3 + proof id; one step = one batched prove "
                               f"of the whole batch",
                   "proofs": args.batch, "gates_per_step": batch["gates"], "lanes": 16, "repetitions": 3,
                   "method": f"batched ({batch['engine']} engine, block {batch['block']}, "
                             f"{batch['lanes_per_thread']} lanes per thread, autotuned)"},
        "lane_gates_per_s": round(value * 16, 1),
        "plain_gates_per_s": batch["plain_gates_per_s"] and round(batch["plain_gates_per_s"], 1),
        "batched_vs_plain": batch["ratio"] and round(batch["ratio"], 3),
        "latency_ms": serve["latency_ms"] if serve else None,
        "roofline": batch["roofline"],
        "cpu_baseline": cpu,
        "batch": {k: (round(v, 3) if isinstance(v, float) else v) for k, v in batch.items() if k != "roofline"},
        "serving": serve,
        "single_proof": single,
        "bn254": bn254,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
