"""Poisson request load for the prove servers (gkrx.Server): the serving benchmark and its
test.  No reference counterpart (the reference proves one circuit per Prover::prove call).

Requests of several kinds (e.g. M31x16 and BN254 proofs, each kind on its own server)
arrive with exponential gaps at a given total rate; each request takes a circuit of its
kind that has no request in flight (a circuit is busy until its proof is back), submits it
and is reaped by a polling thread.  A request's latency runs from its scheduled arrival to
its proof on the host, so time spent waiting for a free circuit counts.  Results are checked
against proofs made alone where the caller supplies them."""
from __future__ import annotations

import random
import threading
import time

import numpy as np


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * q / 100.0
    lo, hi = int(np.floor(k)), int(np.ceil(k))
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


class Kind:
    """One request kind: its server, its circuits (each with its gate count), its share of
    the arrivals, and optionally the proof of circuit i made alone (alone[i] = (claimed,
    proof)), against which every completed request on circuit i is compared."""

    def __init__(self, name, server, circuits, gates, weight, alone=None):
        self.name, self.server, self.circuits, self.gates = name, server, circuits, gates
        self.weight, self.alone = weight, alone or {}


def run(kinds, rate_per_s, seconds, seed=5, log=None, progress_s=30.0):
    """Drives the servers for `seconds` at `rate_per_s` total arrivals; returns a dict of
    counts, overall and per-kind latency percentiles, the mean and p99 of each part of the
    server's time split per kind, gates/s and mismatches against the alone proofs."""
    rng = random.Random(seed)
    free = {k.name: list(range(len(k.circuits))) for k in kinds}
    by_name = {k.name: k for k in kinds}
    pending = []  # (kind, ticket, circuit index, arrival)
    lock = threading.Lock()
    lat = {k.name: [] for k in kinds}
    splits = {k.name: [] for k in kinds}
    done_gates = [0]
    checked = [0]
    mismatches = [0]
    errors = []
    stop = threading.Event()

    def reaper():
        while not (stop.is_set() and not pending):
            with lock:
                items = list(pending)
            progressed = False
            for it in items:
                kind, t, i, arr = it
                k = by_name[kind]
                try:
                    r = k.server.wait(t, timeout_s=0, split=True)
                except Exception as e:  # a failed request is reported, the run goes on
                    errors.append(f"{kind}[{i}]: {e}")
                    r = (None, None, 0.0, {})
                if r is None:
                    continue
                cl, pr, _, sp = r
                now = time.perf_counter()
                ref = k.alone.get(i)
                if ref is not None and cl is not None:
                    checked[0] += 1
                    if bytes(pr) != bytes(ref[1]) or not np.array_equal(cl, ref[0]):
                        mismatches[0] += 1
                with lock:
                    pending.remove(it)
                    free[kind].append(i)
                    lat[kind].append((now - arr) * 1e3)
                    splits[kind].append(sp)
                    done_gates[0] += k.gates[i]
                progressed = True
            if not progressed:
                time.sleep(0.0005)

    th = threading.Thread(target=reaper)
    th.start()
    names = [k.name for k in kinds]
    weights = [k.weight for k in kinds]
    start = time.perf_counter()
    t_next = start
    t_log = start + progress_s
    arrivals = waited = 0
    try:
        while t_next - start < seconds:
            now = time.perf_counter()
            if log and now >= t_log:  # progress of a long serving run
                log(f"serving: {now - start:.0f}s, {arrivals} arrivals, {sum(map(len, lat.values()))} completed")
                t_log += progress_s
            if now < t_next:
                time.sleep(min(t_next - now, 0.002))
                continue
            kind = rng.choices(names, weights)[0]
            while True:
                with lock:
                    if free[kind]:
                        i = free[kind].pop(rng.randrange(len(free[kind])))
                        break
                waited += 1
                time.sleep(0.0002)
            t = by_name[kind].server.submit(by_name[kind].circuits[i])
            with lock:
                pending.append((kind, t, i, t_next))
            arrivals += 1
            t_next += rng.expovariate(rate_per_s)
    finally:
        stop.set()
        th.join()
    elapsed = time.perf_counter() - start
    all_lat = [x for v in lat.values() for x in v]

    def pq(xs):
        return {"p50": round(percentile(xs, 50), 3), "p99": round(percentile(xs, 99), 3)} if xs else None

    per_kind = {}
    for k in kinds:
        sp = splits[k.name]
        parts = {}
        for key in ("prefix_queue_ms", "prefix_ms", "admission_ms", "device_ms", "collect_ms"):
            xs = [s[key] for s in sp if s]
            if xs:
                parts[key] = {"mean": round(float(np.mean(xs)), 3), "p99": round(percentile(xs, 99), 3)}
        # the time before the server saw the request: waiting for a free circuit of the kind
        per_kind[k.name] = {"requests": len(lat[k.name]), "latency_ms": pq(lat[k.name]), "split_ms": parts}
    return {"requests": arrivals, "completed": len(all_lat), "elapsed_s": round(elapsed, 3),
            "gates_per_s": round(done_gates[0] / elapsed, 1), "proofs_per_s": round(len(all_lat) / elapsed, 1),
            "latency_ms": pq(all_lat), "per_kind": per_kind, "arrivals_waiting_for_a_free_circuit": waited,
            "checked_vs_alone": checked[0], "mismatches_vs_alone": mismatches[0], "errors": errors[:5]}
