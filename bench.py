"""Benchmark: proved gates/s of the MI355X GKR prover on the single large M31x16 proof.

Workload: one synthetic layered circuit of 24 layers x 2^20 gates (70% add / 30% mul,
80% of input indices within +-1024 of the output index), M31 with 16 data-parallel
lanes, RAW commitment, Keccak-256 transcript, seed 1.  A step is one whole proof:
forward pass on the device, transcript prefix (commitment hash, grinding, output
challenges) on a host core, every layer sumcheck with its on-device transcript, and the
proof bytes back on the host.  The witness is resident in HBM before timing starts.

Prints one JSON line (the first), then diagnostics on stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "expander-cpp_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
# HBM bytes per dispatch of the roofline kernel on the 24 x 2^20 M31x16 workload, measured
# with two rocprofv3 --pmc runs of this script (FETCH_SIZE, then WRITE_SIZE, per dispatch)
# and scaled by what the same counters read for a 256 MiB device-to-device copy
# (--calibration-copy-mb 256): bytes read = 2 x FETCH_SIZE KiB, written = 1 x WRITE_SIZE KiB.
# k_pair<16, phase 1, double-fold>: read 2 x FETCH_SIZE, written 1 x WRITE_SIZE (KiB) per
# dispatch (profiles/pmc_traffic_m31x16_24x2e20_keccak256.md).
ROOFLINE_TRAFFIC_BYTES = 140514307


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * q / 100.0
    lo, hi = int(np.floor(k)), int(np.ceil(k))
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


TAIL_ROWS = 2048  # tables with at most this many rows end a phase in the small-table kernels


def pair_kernel_bytes(n, R, S):
    """Algorithmic HBM bytes of the phase-1 double-fold round-pair launches (k_pair, MODE 2)
    of one layer with n input variables.  Rounds run in pairs (k, k+1) while the tables are
    large: K is the first round whose table has at most TAIL_ROWS rows, rounded up to even.
    The pass of pair k >= 2 reads f and hg of table k - 2 (2^(n-k+2) entries each) and writes
    table k (2^(n-k) entries each): 10 * 2^(n-k) * S bytes per rep."""
    k0 = 0
    while k0 < n and (1 << (n - k0)) > TAIL_ROWS:
        k0 += 1
    K = (k0 + 1) & ~1
    return sum(R * 10 * (1 << (n - k)) * S for k in range(2, K, 2))


def cpu_baseline(layers_all, L, fs, seed, n_layers):
    """The reference prover (oracle/_ref, built from the reference sources) on a bounded
    sample: the first n_layers layers of the workload circuit, single thread.  Also returns
    the reference harness's throughput mode (bench_keccak.cpp:82-92: one proof per thread on
    every core) on the same sample, or None without the reference build."""
    import oracle_lib as ol
    import gkrx

    sample = layers_all[:n_layers]
    k = int(np.log2(len(sample[0][0]) + len(sample[0][1])))
    wit = gkrx.synth_witness(seed, 1 << k, L)
    gates = sum(len(m) + len(a) for m, a in sample)
    desc = (f"first {n_layers} of the 24 layers (2^{k} gates each, same wiring and witness shape), "
            f"evaluate excluded as in the reference harness")
    allcores = None
    with tempfile.TemporaryDirectory() as d:
        mp, ap, wp = (os.path.join(d, x) for x in ("mul.txt", "add.txt", "w.bin"))
        gkrx.write_text(sample, mp, ap)
        wit.tofile(wp)
        kind, binpath = "reference", ol.ref_binary(L, fs)
        if ol.ref_available(L, fs):
            cmd = [binpath, "time", mp, ap, wp, "1"]
            t0 = time.time()
            out = subprocess.run(cmd, check=True, capture_output=True, text=True)
            secs = float(out.stdout.strip().splitlines()[-1])
            log(f"cpu baseline (reference) wall {time.time() - t0:.1f}s, prove {secs:.3f}s")
            # one proof per core at once, each in its own process
            ncores = min(16, os.cpu_count() or 1)
            t0 = time.time()
            procs = [subprocess.Popen(cmd, stdout=subprocess.PIPE, text=True) for _ in range(ncores)]
            each = [float(p.communicate()[0].strip().splitlines()[-1]) for p in procs]
            if any(p.returncode for p in procs):
                raise RuntimeError("reference prover failed")
            log(f"cpu baseline (reference, {ncores} processes) wall {time.time() - t0:.1f}s, "
                f"prove {min(each):.3f} .. {max(each):.3f}s")
            allcores = {"value": ncores * gates / max(each), "unit": "gates/s", "cores": ncores, "kind": kind,
                        "sample": desc + f"; {ncores} proofs at once, one process per core, the slowest prove time"}
        else:
            kind = "port"
            oc = ol.OracleCircuit(mp, ap)
            t0 = time.time()
            oc.prove(L, fs, wit)
            secs = time.time() - t0
    return {"value": gates / secs, "unit": "gates/s", "cores": 1, "kind": kind,
            "sample": desc + "; one proof on one core"}, allcores


def batch_specs(count, seed=3):
    """The uneven batch: proof i has its own seed (3 + i), width 2^k with k uniform in
    [10, 18] and depth uniform in [4, 24] (drawn from a generator seeded with 3)."""
    rng = np.random.default_rng(seed)
    ks = rng.integers(10, 19, size=count)
    ds = rng.integers(4, 25, size=count)
    return [(seed + i, int(ds[i]), int(ks[i])) for i in range(count)]


def batch_leg(gkrx, fs, count, steps, warmup):
    """Batched proving (the autotuned batch engine) against the plain path (gkrx_prove once
    per proof) on the same uneven batch of M31x16 proofs."""
    from concurrent.futures import ThreadPoolExecutor

    L = 16
    specs = batch_specs(count)
    ctx = gkrx.Context(gkrx.Config(fs_hash=fs, nb_parallel=L), device=0)
    t0 = time.perf_counter()

    def make(spec):
        s, d, k = spec
        layers = gkrx.synth_circuit(s, d, k, 700, True)
        c = gkrx.Circuit.from_layers(ctx, layers)
        c.set_witness(gkrx.synth_witness(s, 1 << k, L))
        return c, sum(len(m) + len(a) for m, a in layers)

    with ThreadPoolExecutor(max_workers=16) as ex:
        made = list(ex.map(make, specs))
    circs = [c for c, _ in made]
    gates = sum(g for _, g in made)
    log(f"batch of {count} proofs, {gates} gates, built in {time.perf_counter() - t0:.1f}s")
    prover = gkrx.Prover(ctx)
    outs = [np.empty(prover.proof_size(c), dtype=np.uint8) for c in circs]
    # autotuner: every batch engine and block size on this batch's device work
    engine, block, lanes, tune_ms = ctx.batch_autotune(circs)
    names = {gkrx.BATCH_PER_PROOF: "per_proof", gkrx.BATCH_LOCKSTEP: "lockstep"}
    tuned = {f"{names[e]}_b{b}_v{v}": round(ms, 3) for (e, b, v), ms in tune_ms.items()}
    log(f"autotune: {tuned} -> {names[engine]} block {block} lanes {lanes}")
    for _ in range(warmup):
        prover.prove_batch(circs, outs)
    t0 = time.perf_counter()
    for _ in range(steps):
        res = prover.prove_batch(circs, outs)
    batched_s = (time.perf_counter() - t0) / steps
    split = prover.last_timing()
    # the plain path on the same proofs, one gkrx_prove per proof; its proofs must equal
    # the batched ones
    plain_lat = []
    same = True
    buf = np.empty(max(o.size for o in outs), dtype=np.uint8)
    t0 = time.perf_counter()
    for i, c in enumerate(circs):
        t1 = time.perf_counter()
        c.evaluate()
        cl, pr = prover.prove(c, out=buf)
        plain_lat.append((time.perf_counter() - t1) * 1e3)
        same &= bool(np.array_equal(cl, res[i][0])) and pr.size == res[i][1].size and bool(np.array_equal(pr, res[i][1]))
    plain_s = time.perf_counter() - t0
    ok = all(gkrx.Verifier(ctx).verify(circs[i], res[i][0], res[i][1]) for i in range(0, count, max(1, count // 8)))
    log(f"batched {batched_s * 1e3:.1f} ms, plain {plain_s * 1e3:.1f} ms, identical={same}, verified={ok}")
    out = {"workload": f"{count} independent M31x16 proofs, width 2^k with k uniform in [10, 18], depth uniform "
                       f"in [4, 24], 70% add / 30% mul, locality wiring, seed 3 + proof id",
           "proofs": count, "gates": gates,
           "batched_gates_per_s": round(gates / batched_s, 1), "batched_ms": round(batched_s * 1e3, 3),
           "plain_gates_per_s": round(gates / plain_s, 1), "plain_ms": round(plain_s * 1e3, 3),
           "ratio": round(plain_s / batched_s, 3),
           "plain_latency_ms": {"p50": round(percentile(plain_lat, 50), 3), "p99": round(percentile(plain_lat, 99), 3)},
           "identical_to_plain": same, "verified_sample": ok,
           "engine": names[engine], "block": block, "lanes_per_thread": lanes, "autotune_device_ms": tuned,
           "batched_split_ms": {k: round(v, 3) for k, v in split.items()}}
    # the context and circuits stay resident for the serving leg (the caller closes ctx)
    return out, (ctx, circs, [g for _, g in made], batched_s / count)


def serve_leg(gkrx, fs, m31_state, seconds, budget_gb, slots):
    """Serving under Poisson arrivals: 95% of the requests are M31x16 proofs drawn from the
    uneven-batch circuits (witnesses resident in HBM), 5% BN254 proofs of 2^14 .. 2^16 gates
    per layer and 8 .. 12 layers.  Two servers (one per field) share the GPU; arrivals come
    at 80% of the saturated rate of that mix (batched M31 time per proof and the BN254
    proof time measured alone).  A request's latency runs from its arrival to its proof on
    the host.  Each completed proof is checked against the proof of the same circuit made
    alone."""
    import random
    import threading

    mctx, mcircs0, mgates0, m31_s = m31_state
    # a second resident copy of every M31 circuit (same wiring and witness), so that a
    # circuit with a request in flight does not hold back arrivals
    from concurrent.futures import ThreadPoolExecutor

    def clone(spec):
        s, d, k = spec
        c = gkrx.Circuit.from_layers(mctx, gkrx.synth_circuit(s, d, k, 700, True))
        c.set_witness(gkrx.synth_witness(s, 1 << k, 16))
        return c

    with ThreadPoolExecutor(max_workers=16) as ex:
        clones = list(ex.map(clone, batch_specs(len(mcircs0))))
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
    # reference proofs made alone, for the identity check
    mprover, bprover = gkrx.Prover(mctx), gkrx.Prover(bctx)
    check_m = {i: mprover.prove(mcircs[i]) for i in range(0, n0, 37)}
    t0 = time.perf_counter()
    check_b = {i: bprover.prove(c) for i, c in enumerate(bcircs)}
    bn_s = (time.perf_counter() - t0) / len(bcircs)
    rate = 0.8 / (0.95 * m31_s + 0.05 * bn_s)
    log(f"serving: {rate:.0f} requests/s for {seconds}s (M31 batched {m31_s * 1e3:.2f} ms/proof, "
        f"BN254 {bn_s * 1e3:.1f} ms/proof alone)")
    budget = int(budget_gb * (1 << 30))
    # M31: `slots` batch streams (+ one for circuits too wide to batch); BN254: one stream
    msrv, bsrv = gkrx.Server(mctx, budget, slots), gkrx.Server(bctx, budget // 8, 1)
    free = {"m": list(range(len(mcircs))), "b": list(range(len(bcircs)))}
    pending = []  # (server, ticket, kind, index, arrival)
    lock = threading.Lock()
    lat, done_gates, mismatches = [], [0], [0]
    stop = threading.Event()

    def reaper():
        while not (stop.is_set() and not pending):
            with lock:
                items = list(pending)
            progressed = False
            for it in items:
                srv, t, kind, i, arr, sub = it
                r = srv.wait(t, timeout_s=0)
                if r is None:
                    continue
                cl, pr, _ = r
                now = time.perf_counter()
                ref = check_m.get(i % n0) if kind == "m" else check_b.get(i)
                if ref is not None and (bytes(pr) != bytes(ref[1]) or not np.array_equal(cl, ref[0])):
                    mismatches[0] += 1
                with lock:
                    pending.remove(it)
                    free[kind].append(i)
                    lat.append((now - arr) * 1e3)
                    done_gates[0] += (mgates if kind == "m" else bgates)[i]
                progressed = True
            if not progressed:
                time.sleep(0.0005)

    th = threading.Thread(target=reaper)
    th.start()
    start = time.perf_counter()
    t_next = start
    t_log = start + 30.0
    arrivals = waited = 0
    while t_next - start < seconds:
        now = time.perf_counter()
        if now >= t_log:  # progress of a long serving run
            log(f"serving: {now - start:.0f}s, {arrivals} arrivals, {len(lat)} completed")
            t_log += 30.0
        if now < t_next:
            time.sleep(min(t_next - now, 0.002))
            continue
        kind = "b" if rng.random() < 0.05 else "m"
        while True:
            with lock:
                if free[kind]:
                    i = free[kind].pop(rng.randrange(len(free[kind])))
                    break
            waited += 1
            time.sleep(0.0002)
        srv, circ = (msrv, mcircs[i]) if kind == "m" else (bsrv, bcircs[i])
        t = srv.submit(circ)
        with lock:
            pending.append((srv, t, kind, i, t_next, time.perf_counter()))
        arrivals += 1
        t_next += rng.expovariate(rate)
    stop.set()
    th.join()
    elapsed = time.perf_counter() - start
    ms, bs = msrv.stats(), bsrv.stats()
    msrv.close()
    bsrv.close()
    bctx.close()
    for c in clones:
        c.close()
    log(f"served {arrivals} requests in {elapsed:.1f}s, p50 {percentile(lat, 50):.1f} ms, p99 {percentile(lat, 99):.1f} ms")
    return {"workload": f"Poisson arrivals for {seconds}s at 80% of the saturated rate of the mix; 95% M31x16 proofs "
                        f"of the uneven batch, 5% BN254 proofs (2^14..2^16 gates x 8..12 layers); HBM budget "
                        f"{budget_gb} GB; {slots} M31 batch slots; two resident copies of each M31 circuit", "rate_per_s": round(rate, 1), "requests": arrivals,
            "gates_per_s": round(done_gates[0] / elapsed, 1), "proofs_per_s": round(arrivals / elapsed, 1),
            "latency_ms": {"p50": round(percentile(lat, 50), 3), "p99": round(percentile(lat, 99), 3)},
            "m31_batches": ms["batches"], "m31_largest_batch": ms["max_batch"],
            "m31_peak_admitted_gb": round(ms["peak_bytes"] / (1 << 30), 2), "bn254_proofs": bs["completed"],
            "arrivals_waiting_for_a_free_circuit": waited, "mismatches_vs_alone": mismatches[0]}


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
            "verified": bool(ok)}


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
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--layers", type=int, default=24)
    ap.add_argument("--log-width", type=int, default=20)
    ap.add_argument("--hash", choices=["keccak256", "sha256"], default="keccak256")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-layers", type=int, default=2)
    ap.add_argument("--batch", type=int, default=512, help="proofs in the uneven batch leg (0 = skip)")
    ap.add_argument("--batch-steps", type=int, default=2)
    ap.add_argument("--serve-seconds", type=float, default=20.0, help="serving leg duration (0 = skip)")
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--serve-slots", type=int, default=2)
    ap.add_argument("--bn254-layers", type=int, default=12, help="layers of the BN254 leg (0 = skip)")
    ap.add_argument("--bn254-log-width", type=int, default=18)
    ap.add_argument("--calibration-copy-mb", type=int, default=0,
                    help="also run device copies of this size (for scaling rocprofv3 --pmc byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path here (independent proofs would "
              "shard by proof id, one process per GPU, with no communication)", file=sys.stderr)
        sys.exit(2)

    import gkrx

    if args.calibration_copy_mb > 0:
        calibration_copy(args.calibration_copy_mb)

    L, seed = 16, 1
    fs = gkrx.FS_KECCAK256 if args.hash == "keccak256" else gkrx.FS_SHA256
    D, k = args.layers, args.log_width
    log(f"generating circuit {D} x 2^{k}")
    layers = gkrx.synth_circuit(seed, D, k, 700, True)
    gates = sum(len(m) + len(a) for m, a in layers)
    wit = gkrx.synth_witness(seed, 1 << k, L)
    ctx = gkrx.Context(gkrx.Config(fs_hash=fs, nb_parallel=L), device=0)
    circ = gkrx.Circuit.from_layers(ctx, layers)
    circ.set_witness(wit)  # witness resident in HBM from here on
    prover = gkrx.Prover(ctx)
    R = ctx.config.get_num_repetitions()
    buf = np.empty(prover.proof_size(circ), dtype=np.uint8)
    for _ in range(args.warmup):
        circ.evaluate()
        prover.prove(circ, out=buf)
    # The timed steps time the roofline kernel's launches only, with HIP events that the
    # launch itself records (hipExtLaunchKernelGGL on the kernel's stream: the kernel's own
    # start and end); an event pair around every launch of a proof would add ~5 ms to it.
    dom = "k_pair_ph1"
    gkrx.lib().gkrx_ctx_profile_classes(ctx._h, 1 << gkrx.kernel_class_index(ctx, dom))
    lat, parts = [], []
    t_start = time.perf_counter()
    for _ in range(args.steps):
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
    log(f"verifier on the last proof: {'accepted' if ok else 'REJECTED'}")
    ms_per_step = elapsed * 1e3 / args.steps
    value = gates * args.steps / elapsed

    S = 4 * L
    bytes_per_proof = sum(pair_kernel_bytes(n, R, S) for n in circ.nb_input_vars)
    roof = None
    if st["count"]:
        per_launch = bytes_per_proof * args.steps / st["count"]
        avg_ms = st["ms"] / st["count"]
        achieved = per_launch / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": ROOFLINE_TRAFFIC_BYTES if (D, k, L) == (24, 20, 16) else None,
                "algorithmic_bytes_per_launch": int(per_launch),
                "kernel": "k_pair<16,phase1,double-fold>", "avg_launch_us": round(avg_ms * 1e3, 2),
                "launches": st["count"]}
    ctx.close()
    batch = serve = None
    if args.batch > 0:
        batch, m31_state = batch_leg(gkrx, fs, args.batch, args.batch_steps, 1)
        if args.serve_seconds > 0:
            serve = serve_leg(gkrx, fs, m31_state, args.serve_seconds, args.serve_budget_gb, args.serve_slots)
        m31_state[0].close()
    bn254 = None
    if args.bn254_layers > 0:
        bn254 = bn254_leg(gkrx, fs, args.bn254_layers, args.bn254_log_width, 2, 1)
    cpu = cpu_all = None
    if not args.no_cpu_baseline:
        try:
            cpu, cpu_all = cpu_baseline(layers, L, fs, seed, args.cpu_layers)
        except Exception as e:  # reported, never fatal
            log(f"cpu baseline failed: {e}")
    line = {
        "metric": "proved gates/s (M31x16) batched vs one-proof-per-launch; p50/p99 proof latency",
        "value": round(value, 1), "unit": "gates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32 (M31 field)", "data": "synthetic",
        "config": {"workload": f"single M31x16 proof, one proof per launch sequence: {D} layers x 2^{k} gates, "
                               f"70% add / 30% mul, locality wiring, RAW commitment, {args.hash} transcript, seed 1",
                   "gates_per_proof": gates, "lanes": L, "repetitions": R, "method": "plain (one proof per launch)"},
        "lane_gates_per_s": round(value * L, 1),
        "latency_ms": {"p50": round(percentile(lat, 50), 3), "p99": round(percentile(lat, 99), 3)},
        "roofline": roof,
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
        "breakdown_ms": {k2: round(float(np.mean([p[k2] for p in parts])), 3) for k2 in parts[0]},
        "kernels_ms_per_proof": {k2: round(v["ms"] / 2, 3) for k2, v in sorted(stats.items())},
        "verified": bool(ok),
        "batch": batch,
        "serving": serve,
        "bn254": bn254,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
