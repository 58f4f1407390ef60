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


LIBDIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "expander-cpp_amd")
INCDIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")

SHIM_WRITER = r"""
#include "gkrx/expander.hpp"
int main(int argc, char **argv)
{
    gkr::Proof<gkr::M31_field::VectorizedM31> p;
    const unsigned char b[5] = {1, 2, 3, 4, 5};
    p.append_bytes(b, 5);
    p.write_and_compress(argv[1], argv[2]);
    return 0;
}
"""


@pytest.mark.skipif(not os.path.exists(os.path.join(LIBDIR, "libgkrx.so")), reason="libgkrx.so not built")
def test_shim_compress_quotes_paths(tmp_path):
    """Proof::write_and_compress hands both paths to tar through the shell; paths holding
    spaces, quotes, ';' and '$(...)' must name files, never run anything."""
    src = tmp_path / "w.cpp"
    src.write_text(SHIM_WRITER)
    exe = str(tmp_path / "w")
    subprocess.run(["g++", "-std=c++17", "-I" + INCDIR, "-o", exe, str(src), "-L" + LIBDIR, "-lgkrx",
                    "-Wl,-rpath," + LIBDIR], check=True, capture_output=True)
    d = tmp_path / "d ir"
    d.mkdir()
    marker = tmp_path / "ran"
    proof = str(d / "p'; touch ran; '.bin")
    tgz = str(d / "c $(touch ran) '.tar.gz")
    subprocess.run([exe, proof, tgz], check=True, capture_output=True, timeout=60, cwd=str(tmp_path))
    assert not marker.exists()
    assert gkrx.proof_read(proof) == bytes([1, 2, 3, 4, 5])
    listing = subprocess.run(["tar", "-tzf", tgz], check=True, capture_output=True, text=True).stdout
    assert listing.strip().endswith(os.path.basename(proof))
