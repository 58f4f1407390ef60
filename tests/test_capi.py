"""The C-ABI library on the host: it loads, exports every function include/gkrx.h
declares, reports configs like the reference's Config, fails cleanly without a GPU, and
its synthetic generator has the documented shape.  No device compute here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import gkrx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(gkrx_\w+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = gkrx.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_config_mirrors_reference_defaults():
    c = gkrx.Config()
    assert c._c.security_bits == 100 and c._c.grinding_bits == 10
    assert c.get_num_repetitions() == 3            # ceil((100 - 10) / 31), config.hpp:61-65
    assert gkrx.Config(field_type=gkrx.FIELD_BN254).get_num_repetitions() == 1


def test_context_without_gpu_fails_cleanly():
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(gkrx.GkrxError, match="device"):
        gkrx.Context(gkrx.Config(), device=0)


def test_unsupported_config_rejected():
    with pytest.raises(gkrx.GkrxError):
        gkrx.Context(gkrx.Config(nb_parallel=4))
    with pytest.raises(gkrx.GkrxError):
        gkrx.Context(gkrx.Config(pc_type=1))


@pytest.mark.parametrize("k,addp,loc", [(10, 500, False), (12, 700, True), (6, 0, True), (6, 1000, False)])
def test_synthetic_layer_shape(k, addp, loc):
    mul, add = gkrx.synth_layer(3, 0, k, addp, loc)
    W = 1 << k
    outs = np.concatenate([mul[:, 2], add[:, 1]])
    assert sorted(outs.tolist()) == list(range(W))            # one gate per output index
    ins = np.concatenate([mul[:, 0], mul[:, 1], add[:, 0]])
    assert ins.max() == W - 1 and ins.min() >= 0             # forced max index -> exactly k vars
    assert (np.concatenate([mul[:, 3], add[:, 2]]) == 1).all()
    frac_add = len(add) / W
    assert abs(frac_add - addp / 1000) < 0.06
    if loc and k >= 12:
        near = np.abs(mul[:, 0].astype(np.int64) - mul[:, 2].astype(np.int64)) <= 1024
        assert near.mean() > 0.75


def test_synthetic_is_deterministic_and_seeded():
    a = gkrx.synth_layer(1, 2, 9, 700, True)
    b = gkrx.synth_layer(1, 2, 9, 700, True)
    c = gkrx.synth_layer(2, 2, 9, 700, True)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not np.array_equal(a[0], c[0])
    w = gkrx.synth_witness(5, 1000, 16)
    assert w.max() < gkrx.P31 and np.array_equal(w, gkrx.synth_witness(5, 1000, 16))
