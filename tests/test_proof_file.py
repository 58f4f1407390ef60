"""Proof files ([u32 byte count][bytes], Proof::write_and_compress / read_proof in
include/fiat_shamir/transcript.hpp:54-85): the library writes what the reference reads and
reads what the reference writes (the reference driver, built from the reference sources)."""
import os
import subprocess

import numpy as np
import pytest

import gkrx
import oracle_lib as ol


def test_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    for n in (0, 1, 4097, 80416):
        data = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        p = str(tmp_path / f"p{n}.bin")
        gkrx.proof_write(p, data)
        raw = open(p, "rb").read()
        assert raw[:4] == n.to_bytes(4, "little") and raw[4:] == data
        assert gkrx.proof_read(p) == data


def test_truncated_file_refused(tmp_path):
    p = str(tmp_path / "t.bin")
    open(p, "wb").write((100).to_bytes(4, "little") + b"x" * 10)
    with pytest.raises(gkrx.GkrxError):
        gkrx.proof_read(p)


@pytest.mark.skipif(not ol.ref_available(8, ol.HASH_SHA256), reason="reference driver not built or no AVX-512")
def test_reference_reads_and_writes_the_same_files(tmp_path):
    ref = ol.ref_binary(8, ol.HASH_SHA256)
    data = open(os.path.join(os.path.dirname(__file__), "golden", "c1", "proof.bin"), "rb").read()
    raw = str(tmp_path / "raw.bin")
    open(raw, "wb").write(data)
    ours, back = str(tmp_path / "ours.proof"), str(tmp_path / "back.bin")
    gkrx.proof_write(ours, data)
    subprocess.run([ref, "proof_read", ours, back], check=True, capture_output=True)
    assert open(back, "rb").read() == data
    theirs = str(tmp_path / "theirs.proof")
    subprocess.run([ref, "proof_write", raw, theirs, str(tmp_path / "theirs.tar.gz")], check=True,
                   capture_output=True)
    assert gkrx.proof_read(theirs) == data
    assert open(theirs, "rb").read() == open(ours, "rb").read()
