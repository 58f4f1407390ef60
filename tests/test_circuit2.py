"""The CIRCUIT2 binary loader (gkrx_circuit2_to_text, the flattening gkrx_circuit_load_circuit2
uses): against the reference's own CircuitRaw reader and Circuit::from_circuit_raw (the
reference driver's `flatten`, built from the reference sources), against a Python
restatement for the coefficient-reading mode the reference does not have, and on the
malformed-file paths.  A proof of a CIRCUIT2 circuit is then the same bytes as the
reference prover's on the same file."""
import os
import struct
import subprocess

import numpy as np
import pytest

import circuit2_gen as cg
import gkrx
import oracle_lib as ol


def _gate_sets(layers):
    return [(sorted(map(tuple, np.asarray(m).reshape(-1, 4).tolist())),
             sorted(map(tuple, np.asarray(a).reshape(-1, 3).tolist()))) for m, a in layers]


@pytest.mark.parametrize("shape", [(1, 2, 4), (2, 3, 5), (3, 4, 6)])
def test_flatten_matches_python_restatement(shape, tmp_path):
    seed, depth, k = shape
    segs, layers, sentinel = cg.make(seed, depth, k)
    path = str(tmp_path / "c.circuit2")
    gkrx.write_circuit2(path, segs, layers, sentinel)
    for mode in (0, 1):
        mp, ap = str(tmp_path / f"m{mode}.txt"), str(tmp_path / f"a{mode}.txt")
        gkrx.circuit2_to_text(path, mp, ap, coef_mode=mode)
        got = _gate_sets(gkrx.read_text(mp, ap))
        want = cg.flatten(segs, layers, mode, sentinel)
        assert got == [(list(m), list(a)) for m, a in want]
    # constant gates are dropped and every layer has its intended variable counts
    c = gkrx.read_text(mp, ap)
    assert all(int(np.max(m[:, :2], initial=0)) <= (1 << k) - 1 for m, _ in c)


@pytest.mark.skipif(not ol.ref_available(8, ol.HASH_SHA256), reason="reference driver not built or no AVX-512")
@pytest.mark.parametrize("shape", [(4, 2, 4), (5, 3, 6), (6, 5, 8)])
def test_flatten_matches_reference_reader(shape, tmp_path):
    seed, depth, k = shape
    segs, layers, sentinel = cg.make(seed, depth, k)
    path = str(tmp_path / "c.circuit2")
    gkrx.write_circuit2(path, segs, layers, sentinel)
    rm, ra = str(tmp_path / "rm.txt"), str(tmp_path / "ra.txt")
    subprocess.run([ol.ref_binary(8, ol.HASH_SHA256), "flatten", path, rm, ra], check=True, capture_output=True)
    mp, ap = str(tmp_path / "m.txt"), str(tmp_path / "a.txt")
    gkrx.circuit2_to_text(path, mp, ap, coef_mode=0)
    assert _gate_sets(gkrx.read_text(mp, ap)) == _gate_sets(gkrx.read_text(rm, ra))


@pytest.mark.skipif(not ol.ref_available(8, ol.HASH_SHA256), reason="reference driver not built or no AVX-512")
def test_proof_of_circuit2_equals_reference(tmp_path):
    """The oracle proving the flattened text gives the reference prover's bytes for the
    same CIRCUIT2 file (prove2 reads the binary file itself)."""
    segs, layers, sentinel = cg.make(7, 3, 6)
    path = str(tmp_path / "c.circuit2")
    gkrx.write_circuit2(path, segs, layers, sentinel)
    mp, ap = str(tmp_path / "m.txt"), str(tmp_path / "a.txt")
    gkrx.circuit2_to_text(path, mp, ap, coef_mode=0)
    wit = gkrx.synth_witness(7, 1 << 6, 8)
    wp, pp, cp = str(tmp_path / "w.bin"), str(tmp_path / "p.bin"), str(tmp_path / "c.bin")
    wit.tofile(wp)
    subprocess.run([ol.ref_binary(8, ol.HASH_SHA256), "prove2", path, wp, pp, cp], check=True, capture_output=True)
    claimed, proof = ol.OracleCircuit(mp, ap).prove(8, ol.HASH_SHA256, wit)
    assert proof == open(pp, "rb").read()
    assert np.array_equal(claimed.ravel(), np.fromfile(cp, dtype=np.uint32))


def test_malformed_files_are_refused(tmp_path):
    segs, layers, sentinel = cg.make(9, 2, 4)
    good = str(tmp_path / "good.circuit2")
    gkrx.write_circuit2(good, segs, layers, sentinel)
    data = open(good, "rb").read()
    out = (str(tmp_path / "m.txt"), str(tmp_path / "a.txt"))
    cases = {
        "bad_magic": b"X" + data[1:],
        "truncated": data[:-40],
        "empty": b"",
        # a layer that points past the segment table
        "bad_layer": data[:-40] + struct.pack("<Q", 999) + data[-32:],
    }
    for name, blob in cases.items():
        p = str(tmp_path / f"{name}.circuit2")
        open(p, "wb").write(blob)
        with pytest.raises(gkrx.GkrxError):
            gkrx.circuit2_to_text(p, *out)
    # an out-of-range random-coefficient index (the reference exits, circuit_raw.hpp:374-375)
    bad = [dict(s) for s in segs]
    bad[0] = dict(bad[0], rand=[10 ** 6])
    p = str(tmp_path / "badrand.circuit2")
    gkrx.write_circuit2(p, bad, layers, sentinel)
    with pytest.raises(gkrx.GkrxError):
        gkrx.circuit2_to_text(p, *out)
    # a cycle in the segment tree
    cyc = [dict(s) for s in segs]
    cyc[1] = dict(cyc[1], children=[(layers[0], [(0, 0)])])
    p = str(tmp_path / "cycle.circuit2")
    gkrx.write_circuit2(p, cyc, layers, sentinel)
    with pytest.raises(gkrx.GkrxError):
        gkrx.circuit2_to_text(p, *out)
    assert os.path.exists(good)
