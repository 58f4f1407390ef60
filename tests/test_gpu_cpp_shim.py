"""The reference-style C++ call site (expander-cpp_amd/examples/prove_example.cpp, built
against include/gkrx/expander.hpp) proves the committed c1 circuit on the GPU and
reproduces the reference proof bytes."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "expander-cpp_amd", "examples", "bin", "prove_example")
C1 = os.path.join(ROOT, "tests", "golden", "c1")

pytestmark = pytest.mark.gpu


def test_reference_style_call_site_reproduces_reference_proof():
    print("[gpu] C++ drop-in example on c1", flush=True)
    assert os.path.exists(EXE), "build the package first"
    r = subprocess.run([EXE, os.path.join(C1, "mul.txt"), os.path.join(C1, "add.txt"), os.path.join(C1, "witness.bin"),
                        os.path.join(C1, "proof.bin")], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "matches expected" in r.stdout and "accepted" in r.stdout
