"""gkrx_prove_noalloc, the allocation-free variant of gkrx_prove (its prover tables live in a
file-scope device array): called through the C ABI, its proof bytes and claims equal
gkrx_prove's on two calls in a row, the second on a smaller circuit (the array keeps the
first call's contents), and a circuit whose tables exceed the array is refused with an
error instead of being proved."""
import ctypes as C

import numpy as np
import pytest

import gkrx

pytestmark = pytest.mark.gpu


def _circuit(ctx, seed, depth, k):
    c = gkrx.Circuit.from_layers(ctx, gkrx.synth_circuit(seed, depth, k, 700, True))
    c.set_witness(gkrx.synth_witness(seed, 1 << k, ctx.config.words))
    return c


def _noalloc(ctx, circ):
    size = gkrx.lib().gkrx_proof_size(ctx._h, circ._h)
    proof = np.zeros(size, dtype=np.uint8)
    R = ctx.config.get_num_repetitions()
    claimed = np.zeros(R * ctx.config.words, dtype=np.uint32)
    n = C.c_size_t()
    rc = gkrx.lib().gkrx_prove_noalloc(ctx._h, circ._h, proof, proof.size, C.byref(n), claimed)
    return rc, claimed.reshape(R, -1), proof[:n.value]


@pytest.mark.parametrize("lanes,hash_kind", [(16, gkrx.FS_KECCAK256), (8, gkrx.FS_SHA256)])
def test_noalloc_equals_prove_twice_then_smaller(lanes, hash_kind):
    print(f"[gpu] noalloc prove, {lanes} lanes", flush=True)
    ctx = gkrx.Context(gkrx.Config(fs_hash=hash_kind, nb_parallel=lanes))
    prover = gkrx.Prover(ctx)
    big, small = _circuit(ctx, 71, 5, 16), _circuit(ctx, 72, 3, 9)
    for circ in (big, small):
        rc, cl, pr = _noalloc(ctx, circ)
        assert rc == 0, gkrx.lib().gkrx_last_error().decode()
        circ.evaluate()
        c1, p1 = prover.prove(circ)
        assert np.array_equal(cl, c1)
        assert pr.size == p1.size and bytes(pr) == bytes(p1)
        assert gkrx.Verifier(ctx).verify(circ, cl, pr)
    # the same small circuit again with a new witness: the array's old contents do not leak
    small.set_witness(gkrx.synth_witness(73, 1 << 9, lanes))
    rc, cl, pr = _noalloc(ctx, small)
    assert rc == 0
    c1, p1 = prover.prove(small)
    assert np.array_equal(cl, c1) and bytes(pr) == bytes(p1)
    big.close()
    small.close()
    ctx.close()


def test_noalloc_refuses_tables_above_the_array():
    print("[gpu] noalloc size limit", flush=True)
    ctx = gkrx.Context(gkrx.Config(fs_hash=gkrx.FS_KECCAK256, nb_parallel=16))
    circ = _circuit(ctx, 74, 1, 17)  # 2^17 rows x 16 lanes: about 69 MB of tables
    rc, _, _ = _noalloc(ctx, circ)
    assert rc == -1
    assert "allocation" in gkrx.lib().gkrx_last_error().decode() or "gkrx_prove_noalloc" in \
        gkrx.lib().gkrx_last_error().decode()
    circ.close()
    ctx.close()
