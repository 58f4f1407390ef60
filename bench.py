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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * q / 100.0
    lo, hi = int(np.floor(k)), int(np.ceil(k))
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


TAIL_ROWS = 2048  # rounds whose tables have at most this many rows run in k_tail


def round_kernel_bytes(n, R, S):
    """Algorithmic HBM bytes of the phase-1 fused round kernel (k_round) launches of one
    layer with n input variables.  Round k >= 1 reads f_{k-1} and hg_{k-1} (2^(n-k+1)
    entries each) and writes f_k and hg_k (2^(n-k) entries each): 3 * 2^(n-k+1) * S bytes
    per rep; only rounds with more than TAIL_ROWS rows run in k_round."""
    k0 = 0
    while k0 < n and (1 << (n - k0)) > TAIL_ROWS:
        k0 += 1
    return sum(R * 3 * (1 << (n - k + 1)) * S for k in range(1, k0))


def cpu_baseline(layers_all, L, fs, seed, n_layers):
    """The reference prover (oracle/_ref, built from the reference sources) on a bounded
    sample: the first n_layers layers of the workload circuit, single thread."""
    import oracle_lib as ol
    import gkrx

    sample = layers_all[:n_layers]
    k = int(np.log2(len(sample[0][0]) + len(sample[0][1])))
    wit = gkrx.synth_witness(seed, 1 << k, L)
    gates = sum(len(m) + len(a) for m, a in sample)
    with tempfile.TemporaryDirectory() as d:
        mp, ap, wp = (os.path.join(d, x) for x in ("mul.txt", "add.txt", "w.bin"))
        gkrx.write_text(sample, mp, ap)
        wit.tofile(wp)
        kind, binpath = "reference", ol.ref_binary(L, fs)
        if ol.ref_available(L, fs):
            t0 = time.time()
            out = subprocess.run([binpath, "time", mp, ap, wp, "1"], check=True, capture_output=True, text=True)
            secs = float(out.stdout.strip().splitlines()[-1])
            log(f"cpu baseline (reference) wall {time.time() - t0:.1f}s, prove {secs:.3f}s")
        else:
            kind = "port"
            oc = ol.OracleCircuit(mp, ap)
            t0 = time.time()
            oc.prove(L, fs, wit)
            secs = time.time() - t0
    return {"value": gates / secs, "unit": "gates/s", "cores": 1, "kind": kind,
            "sample": f"first {n_layers} of the 24 layers (2^{k} gates each, same wiring and witness shape), "
                      f"one proof on one core, evaluate excluded as in the reference harness"}


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
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (proofs shard by proof id "
              "with no communication; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import gkrx

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
    gkrx.lib().gkrx_ctx_profile(ctx._h, 1)
    lat, parts = [], []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        circ.evaluate()
        claimed, proof = prover.prove(circ, out=buf)
        lat.append((time.perf_counter() - t0) * 1e3)
        parts.append(prover.last_timing())
    elapsed = time.perf_counter() - t_start
    stats = gkrx.kernel_stats(ctx)
    gkrx.lib().gkrx_ctx_profile(ctx._h, 0)
    ok = gkrx.Verifier(ctx).verify(circ, claimed, proof)
    log(f"verifier on the last proof: {'accepted' if ok else 'REJECTED'}")
    ms_per_step = elapsed * 1e3 / args.steps
    value = gates * args.steps / elapsed

    S = 4 * L
    dom = "k_round_ph1"
    st = stats.get(dom, {"count": 0, "ms": 0.0})
    bytes_per_proof = sum(round_kernel_bytes(n, R, S) for n in circ.nb_input_vars)
    roof = None
    if st["count"]:
        per_launch = bytes_per_proof * args.steps / st["count"]
        avg_ms = st["ms"] / st["count"]
        achieved = per_launch / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_round<16,phase1,fused>", "avg_launch_us": round(avg_ms * 1e3, 2),
                "launches": st["count"]}
    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(layers, L, fs, seed, args.cpu_layers)
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
        "breakdown_ms": {k2: round(float(np.mean([p[k2] for p in parts])), 3) for k2 in parts[0]},
        "kernels_ms_per_proof": {k2: round(v["ms"] / args.steps, 3) for k2, v in sorted(stats.items())},
        "verified": bool(ok),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
