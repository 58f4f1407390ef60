"""ctypes access to the CPU oracle (oracle/liboracle.so) and the reference drivers
(oracle/_ref/ref_m31x{8,16}_{sha256,keccak256}).  Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")
HASH_SHA256, HASH_KECCAK256 = 0, 1
P31 = (1 << 31) - 1

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_o = None


def oracle():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is missing (build with __graft_entry__.build())")
        o = C.CDLL(ORACLE_SO)
        for name in ("oracle_m31_add", "oracle_m31_sub", "oracle_m31_mul"):
            getattr(o, name).argtypes = [C.c_uint32, C.c_uint32]
            getattr(o, name).restype = C.c_uint32
        o.oracle_m31_neg.argtypes = [C.c_uint32]
        o.oracle_m31_neg.restype = C.c_uint32
        o.oracle_m31_from_u32.argtypes = [C.c_uint32]
        o.oracle_m31_from_u32.restype = C.c_uint32
        o.oracle_m31_from_bytes.argtypes = [C.c_char_p]
        o.oracle_m31_from_bytes.restype = C.c_uint32
        o.oracle_sha256.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p]
        o.oracle_keccak256.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p]
        o.oracle_keccak256_pad.argtypes = [C.c_char_p, C.c_size_t, C.c_uint8, C.c_char_p]
        o.oracle_circuit_load_text.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]
        o.oracle_circuit_free.argtypes = [C.c_void_p]
        o.oracle_evaluate.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]
        o.oracle_evaluate.restype = C.c_size_t
        o.oracle_eq_evals.argtypes = [_u32p, C.c_uint32, C.c_uint32, _u32p]
        o.oracle_eval_multilinear.argtypes = [_u32p, C.c_uint32, _u32p, C.c_uint32, _u32p]
        o.oracle_prove.argtypes = [C.c_void_p, C.c_uint32, C.c_int, _u32p, C.POINTER(C.POINTER(C.c_uint8)),
                                   C.POINTER(C.c_size_t), _u32p]
        o.oracle_verify.argtypes = [C.c_void_p, C.c_uint32, C.c_int, _u32p, C.c_char_p, C.c_size_t]
        o.oracle_free.argtypes = [C.c_void_p]
        o.oracle_num_repetitions.restype = C.c_uint32
        _o = o
    return _o


def sha256(data: bytes) -> bytes:
    out = C.create_string_buffer(32)
    oracle().oracle_sha256(data, len(data), out)
    return out.raw


def keccak256(data: bytes, pad: int = 0x01) -> bytes:
    out = C.create_string_buffer(32)
    oracle().oracle_keccak256_pad(data, len(data), pad, out)
    return out.raw


class OracleCircuit:
    def __init__(self, mul_path: str, add_path: str):
        h = C.c_void_p()
        rc = oracle().oracle_circuit_load_text(mul_path.encode(), add_path.encode(), C.byref(h))
        if rc != 0:
            raise RuntimeError(f"oracle load failed: {rc}")
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            oracle().oracle_circuit_free(self._h)
            self._h = None

    def evaluate(self, lanes: int, witness: np.ndarray) -> np.ndarray:
        w = np.ascontiguousarray(witness, dtype=np.uint32)
        n = oracle().oracle_evaluate(self._h, lanes, w.ctypes.data, None)
        out = np.empty(n, dtype=np.uint32)
        oracle().oracle_evaluate(self._h, lanes, w.ctypes.data, out.ctypes.data)
        return out

    def prove(self, lanes: int, hash_kind: int, witness: np.ndarray):
        R = oracle().oracle_num_repetitions()
        claimed = np.empty(R * lanes, dtype=np.uint32)
        p = C.POINTER(C.c_uint8)()
        n = C.c_size_t()
        rc = oracle().oracle_prove(self._h, lanes, hash_kind, np.ascontiguousarray(witness, dtype=np.uint32),
                                   C.byref(p), C.byref(n), claimed)
        if rc != 0:
            raise RuntimeError(f"oracle prove failed: {rc}")
        proof = C.string_at(p, n.value)
        oracle().oracle_free(p)
        return claimed.reshape(R, lanes), proof

    def verify(self, lanes: int, hash_kind: int, claimed: np.ndarray, proof: bytes) -> int:
        return oracle().oracle_verify(self._h, lanes, hash_kind, np.ascontiguousarray(claimed, dtype=np.uint32).ravel(),
                                      proof, len(proof))


def ref_binary(lanes: int, hash_kind: int) -> str:
    name = f"ref_m31x{lanes}_{'sha256' if hash_kind == HASH_SHA256 else 'keccak256'}"
    return os.path.join(REF_DIR, name)


def cpu_has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            flags = f.read()
        return " avx512f" in flags and " avx512vl" in flags
    except OSError:
        return False


def ref_available(lanes: int, hash_kind: int) -> bool:
    return os.path.exists(ref_binary(lanes, hash_kind)) and cpu_has_avx512()


def ref_prove(lanes: int, hash_kind: int, mul_path: str, add_path: str, witness: np.ndarray):
    with tempfile.TemporaryDirectory() as d:
        wp, pp, cp = (os.path.join(d, x) for x in ("w.bin", "p.bin", "c.bin"))
        np.ascontiguousarray(witness, dtype=np.uint32).tofile(wp)
        subprocess.run([ref_binary(lanes, hash_kind), "prove", mul_path, add_path, wp, pp, cp], check=True,
                       stdout=subprocess.DEVNULL)
        proof = open(pp, "rb").read()
        claimed = np.fromfile(cp, dtype=np.uint32)
    return claimed.reshape(-1, lanes), proof


def ref_verify(lanes: int, hash_kind: int, mul_path: str, add_path: str, claimed: np.ndarray, proof: bytes) -> bool:
    with tempfile.TemporaryDirectory() as d:
        pp, cp = os.path.join(d, "p.bin"), os.path.join(d, "c.bin")
        open(pp, "wb").write(bytes(proof))
        np.ascontiguousarray(claimed, dtype=np.uint32).tofile(cp)
        r = subprocess.run([ref_binary(lanes, hash_kind), "verify", mul_path, add_path, cp, pp],
                           stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    return r.returncode == 0
