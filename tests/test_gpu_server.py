"""The continuous-batching server (gkrx_server_*): every proof it returns is byte-identical
to the same circuit proved alone, whether its request arrived alone or in a mix with
others in flight; admission keeps the admitted HBM under the budget; misuse is refused."""
import time

import numpy as np
import pytest

import gkrx

pytestmark = pytest.mark.gpu


def _m31_circuits(ctx, specs):
    out = []
    for s, d, k in specs:
        c = gkrx.Circuit.from_layers(ctx, gkrx.synth_circuit(s, d, k, 700, True))
        c.set_witness(gkrx.synth_witness(s, 1 << k, 16))
        out.append(c)
    return out


@pytest.fixture(scope="module")
def ctx():
    c = gkrx.Context(gkrx.Config(fs_hash=gkrx.FS_KECCAK256, nb_parallel=16))
    yield c
    c.close()


# small proofs (batch slots) and two large ones (2^17 x 4 and 2^18 x 2 rows: the single slot)
SPECS = [(100 + i, 2 + (i * 5) % 7, 6 + (i * 3) % 9) for i in range(14)] + [(150, 4, 17), (151, 2, 18)]


def test_server_results_equal_alone(ctx):
    print("[gpu] server: mix in waves vs alone", flush=True)
    circs = _m31_circuits(ctx, SPECS)
    prover = gkrx.Prover(ctx)
    alone = [prover.prove(c) for c in circs]
    srv = gkrx.Server(ctx, hbm_budget_bytes=64 << 30, slots=3)
    tickets = []
    for i, c in enumerate(circs):
        tickets.append(srv.submit(c))
        if i % 5 == 4:
            time.sleep(0.002)  # later arrivals land while earlier batches run
    results = [srv.wait(t, timeout_s=60) for t in tickets]
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(circs) and st["batches"] >= 2
    for (cl, pr, lat), (c1, p1) in zip(results, alone):
        assert np.array_equal(cl, c1)
        assert bytes(pr) == bytes(p1)
        assert lat > 0
    # and alone through the server: one request at a time
    srv = gkrx.Server(ctx, hbm_budget_bytes=64 << 30, slots=1)
    for c, (c1, p1) in zip(circs[:4], alone[:4]):
        cl, pr, _ = srv.wait(srv.submit(c), timeout_s=60)
        assert bytes(pr) == bytes(p1) and np.array_equal(cl, c1)
    srv.close()
    for c in circs:
        c.close()


def test_server_budget_admission(ctx):
    print("[gpu] server: tight HBM budget", flush=True)
    specs = [(200 + i, 3, 12) for i in range(6)]
    circs = _m31_circuits(ctx, specs)
    prover = gkrx.Prover(ctx)
    alone = [prover.prove(c) for c in circs]
    # learn one request's footprint from a 1-request run, then allow about two at a time
    probe = gkrx.Server(ctx, hbm_budget_bytes=64 << 30, slots=1)
    probe.wait(probe.submit(circs[0]), timeout_s=60)
    one = probe.stats()["peak_bytes"]
    probe.close()
    budget = 2 * one + one // 2
    srv = gkrx.Server(ctx, hbm_budget_bytes=budget, slots=4)
    tickets = [srv.submit(c) for c in circs]
    results = [srv.wait(t, timeout_s=60) for t in tickets]
    st = srv.stats()
    srv.close()
    assert st["peak_bytes"] <= budget and st["max_batch"] <= 2 and st["completed"] == len(circs)
    for (cl, pr, _), (c1, p1) in zip(results, alone):
        assert bytes(pr) == bytes(p1) and np.array_equal(cl, c1)
    for c in circs:
        c.close()


def test_server_refuses_misuse(ctx):
    print("[gpu] server: misuse", flush=True)
    big, small = _m31_circuits(ctx, [(300, 8, 14), (301, 2, 6)])
    srv = gkrx.Server(ctx, hbm_budget_bytes=64 << 30, slots=2)
    t = srv.submit(big)
    with pytest.raises(gkrx.GkrxError):
        srv.submit(big)  # already in flight
    with pytest.raises(gkrx.GkrxError):
        gkrx.Prover(ctx).prove(big)
    assert srv.wait(t, timeout_s=60) is not None
    tiny = gkrx.Server(ctx, hbm_budget_bytes=1 << 16, slots=1)
    with pytest.raises(gkrx.GkrxError):
        tiny.submit(small)  # needs more than the budget
    tiny.close()
    srv.close()
    big.close()
    small.close()


def test_server_bn254():
    print("[gpu] server: BN254", flush=True)
    ctx = gkrx.Context(gkrx.Config(fs_hash=gkrx.FS_KECCAK256, field_type=gkrx.FIELD_BN254))
    circs = []
    for s, d, k in [(400, 3, 10), (401, 2, 13), (402, 4, 8)]:
        c = gkrx.Circuit.from_layers(ctx, gkrx.synth_circuit(s, d, k, 700, True))
        c.set_witness(gkrx.synth_witness_bn254(s, 1 << k))
        circs.append(c)
    prover = gkrx.Prover(ctx)
    alone = [prover.prove(c) for c in circs]
    srv = gkrx.Server(ctx, hbm_budget_bytes=16 << 30, slots=2)
    results = [srv.wait(t, timeout_s=60) for t in [srv.submit(c) for c in circs]]
    srv.close()
    for (cl, pr, _), (c1, p1) in zip(results, alone):
        assert bytes(pr) == bytes(p1) and np.array_equal(cl, c1)
    ctx.close()
