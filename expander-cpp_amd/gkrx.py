"""ctypes binding of libgkrx.so (the C ABI declared in include/gkrx.h).

Thin host-side mirror of the reference's C++ surface (PolyhedraZK/Expander-cpp):
``Config`` (include/configuration/config.hpp:26-86), ``Circuit``
(include/circuit/circuit.hpp:118-305), ``Prover.prove`` and ``Verifier.verify``
(include/LinearGKR/LinearGKR.hpp:34-150).  Arrays are numpy uint32; proofs are bytes.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgkrx.so")

FIELD_M31, FIELD_BABYBEAR, FIELD_BN254 = 0, 1, 2
FS_SHA256, FS_KECCAK256 = 0, 1
BATCH_PER_PROOF, BATCH_LOCKSTEP, BATCH_LOCKSTEP_GROUPS = 1, 2, 4
PC_RAW = 0
P31 = (1 << 31) - 1

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


class GkrxError(RuntimeError):
    pass


class _Config(C.Structure):
    _fields_ = [("field_type", C.c_int32), ("fs_hash", C.c_int32), ("pc_type", C.c_int32),
                ("security_bits", C.c_int32), ("grinding_bits", C.c_int32), ("nb_parallel", C.c_int32)]


_lib = None


def lib():
    """Load libgkrx.so; raises if the native library has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GkrxError(f"{LIB_PATH} is missing: build it first (__graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.gkrx_last_error.restype = C.c_char_p
    L.gkrx_config_default.argtypes = [C.POINTER(_Config)]
    L.gkrx_config_num_repetitions.argtypes = [C.POINTER(_Config)]
    L.gkrx_ctx_create.argtypes = [C.POINTER(_Config), C.c_int, C.POINTER(vp)]
    L.gkrx_ctx_destroy.argtypes = [vp]
    L.gkrx_circuit_load_text.argtypes = [vp, C.c_char_p, C.c_char_p, C.POINTER(vp)]
    L.gkrx_circuit_parse_text.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.POINTER(C.c_uint32), vp, vp]
    L.gkrx_circuit_from_arrays.argtypes = [vp, C.c_uint32, _u32p, C.POINTER(C.c_void_p), _u32p,
                                           C.POINTER(C.c_void_p), C.POINTER(vp)]
    L.gkrx_circuit_shape.argtypes = [vp, C.POINTER(C.c_uint32), vp, vp]
    L.gkrx_circuit_load_circuit2.argtypes = [vp, C.c_char_p, C.c_int, C.POINTER(vp)]
    L.gkrx_circuit2_to_text.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_char_p, C.c_char_p]
    L.gkrx_proof_write.argtypes = [C.c_char_p, _u8p, C.c_size_t]
    L.gkrx_proof_read.argtypes = [C.c_char_p, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]
    L.gkrx_circuit_write_text.argtypes = [vp, C.c_char_p, C.c_char_p]
    L.gkrx_circuit_set_witness.argtypes = [vp, _u32p, C.c_size_t]
    L.gkrx_circuit_evaluate.argtypes = [vp]
    L.gkrx_circuit_get_output.argtypes = [vp, _u32p, C.c_size_t]
    L.gkrx_circuit_get_layer_values.argtypes = [vp, C.c_uint32, _u32p, C.c_size_t]
    L.gkrx_circuit_destroy.argtypes = [vp]
    L.gkrx_proof_size.argtypes = [vp, vp]
    L.gkrx_proof_size.restype = C.c_size_t
    L.gkrx_prove.argtypes = [vp, vp, _u8p, C.c_size_t, C.POINTER(C.c_size_t), _u32p]
    L.gkrx_prove_noalloc.argtypes = [vp, vp, _u8p, C.c_size_t, C.POINTER(C.c_size_t), _u32p]
    L.gkrx_verify.argtypes = [vp, vp, _u32p, _u8p, C.c_size_t]
    L.gkrx_verify_arrays.argtypes = [C.POINTER(_Config), C.c_uint32, _u32p, C.POINTER(C.c_void_p), _u32p,
                                     C.POINTER(C.c_void_p), _u32p, _u8p, C.c_size_t]
    L.gkrx_last_timing.argtypes = [vp, C.POINTER(C.c_double)]
    L.gkrx_synth_layer.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _u32p,
                                   C.POINTER(C.c_uint32), _u32p, C.POINTER(C.c_uint32)]
    L.gkrx_synth_witness_m31.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, _u32p]
    L.gkrx_synth_witness_bn254.argtypes = [C.c_uint64, C.c_uint64, _u32p]
    L.gkrx_hash.argtypes = [C.c_int, C.c_char_p, C.c_size_t, C.c_char_p]
    L.gkrx_hash_many.argtypes = [C.c_int, C.c_uint32, C.POINTER(C.c_char_p), C.c_size_t, C.c_char_p]
    L.gkrx_prove_batch.argtypes = [vp, C.c_uint32, C.POINTER(vp), C.POINTER(C.c_void_p), C.POINTER(C.c_size_t),
                                   C.POINTER(C.c_size_t), C.POINTER(C.c_void_p)]
    L.gkrx_server_create.argtypes = [vp, C.c_uint64, C.c_uint32, C.POINTER(vp)]
    L.gkrx_server_submit.argtypes = [vp, vp, _u8p, C.c_size_t, _u32p, C.POINTER(C.c_uint64)]
    L.gkrx_server_wait.argtypes = [vp, C.c_uint64, C.c_int64, C.POINTER(C.c_size_t), C.POINTER(C.c_double)]
    L.gkrx_server_wait_split.argtypes = [vp, C.c_uint64, C.c_int64, C.POINTER(C.c_size_t), C.POINTER(C.c_double)]
    L.gkrx_server_stats.argtypes = [vp, C.POINTER(C.c_uint64)]
    L.gkrx_server_destroy.argtypes = [vp]
    L.gkrx_ctx_profile.argtypes = [vp, C.c_int]
    L.gkrx_ctx_set_batch_engine.argtypes = [vp, C.c_int, C.c_uint32]
    L.gkrx_ctx_get_batch_engine.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_uint32)]
    L.gkrx_ctx_set_batch_lanes.argtypes = [vp, C.c_uint32]
    L.gkrx_ctx_profile_classes.argtypes = [vp, C.c_uint64]
    L.gkrx_ctx_get_batch_lanes.argtypes = [vp, C.POINTER(C.c_uint32)]
    L.gkrx_ctx_set_graphs.argtypes = [vp, C.c_int]
    L.gkrx_ctx_get_graphs.argtypes = [vp, C.POINTER(C.c_int)]
    L.gkrx_batch_autotune.argtypes = [vp, C.c_uint32, C.POINTER(vp), C.POINTER(C.c_int), C.POINTER(C.c_uint32),
                                      C.POINTER(C.c_uint32), C.POINTER(C.c_double)]
    L.gkrx_ctx_kernel_stat.argtypes = [vp, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_double)]
    _lib = L
    return L


def _check(rc: int, what: str) -> int:
    if rc < 0:
        raise GkrxError(f"{what} failed ({rc}): {lib().gkrx_last_error().decode(errors='replace')}")
    return rc


# ----------------------------------------------------------------- synthetic inputs
def synth_layer(seed: int, layer: int, log_width: int, add_permille: int, locality: bool):
    """Gates of one synthetic layer: (mul[u,4], add[a,3]) uint32 arrays."""
    W = 1 << log_width
    mul = np.empty(4 * W, dtype=np.uint32)
    add = np.empty(3 * W, dtype=np.uint32)
    nm, na = C.c_uint32(), C.c_uint32()
    _check(lib().gkrx_synth_layer(seed, layer, log_width, add_permille, int(bool(locality)), mul, C.byref(nm),
                                  add, C.byref(na)), "gkrx_synth_layer")
    return mul[:4 * nm.value].reshape(-1, 4).copy(), add[:3 * na.value].reshape(-1, 3).copy()


def synth_circuit(seed: int, depth: int, log_width: int, add_permille: int, locality: bool):
    """List of (mul, add) per layer, layer 0 = input layer."""
    return [synth_layer(seed, l, log_width, add_permille, locality) for l in range(depth)]


def synth_witness(seed: int, n_elems: int, lanes: int) -> np.ndarray:
    out = np.empty(n_elems * lanes, dtype=np.uint32)
    _check(lib().gkrx_synth_witness_m31(seed, n_elems, lanes, out), "gkrx_synth_witness_m31")
    return out


def synth_witness_bn254(seed: int, n_elems: int) -> np.ndarray:
    """n_elems BN254 scalars as a flat uint32 array (8 LE words per element)."""
    out = np.empty(n_elems * 8, dtype=np.uint32)
    _check(lib().gkrx_synth_witness_bn254(seed, n_elems, out), "gkrx_synth_witness_bn254")
    return out


def bn254_to_ints(words: np.ndarray) -> list:
    """Flat uint32 words (8 per element) -> Python integers."""
    w = np.ascontiguousarray(words, dtype=np.uint32).reshape(-1, 8)
    return [int.from_bytes(row.tobytes(), "little") for row in w]


def write_text(layers, mul_path: str, add_path: str) -> None:
    """Write (mul, add) layer lists in the reference text format
    (Circuit::load_extracted_gates, include/circuit/circuit.hpp:152-215): one line per
    layer, output layer first."""
    with open(mul_path, "w") as fm, open(add_path, "w") as fa:
        for mul, add in reversed(layers):
            fm.write(" ".join([str(len(mul))] + [str(int(x)) for x in np.asarray(mul).ravel()]) + "\n")
            fa.write(" ".join([str(len(add))] + [str(int(x)) for x in np.asarray(add).ravel()]) + "\n")


CIRCUIT2_MAGIC = 3626604230490605891  # b"CIRCUIT2" little-endian


def write_circuit2(path: str, segments, layers, rand_sentinel: int = 0) -> None:
    """Write a CIRCUIT2 binary circuit (the layout CircuitRaw operator<< writes,
    include/circuit/circuit_raw.hpp:205-291, 514-537).  segments: list of dicts with keys
    i_len, o_len, children [(segment id, [(i_offset, o_offset), ...])], mul [(in0, in1, out,
    coef)], add [(in0, out, coef)], const [(out, coef)], rand [gate indices with a random
    coefficient, counted over mul, then add, then const]; layers: segment id per layer.
    Field elements are written as 32 little-endian bytes."""
    import struct

    def u64(x):
        return struct.pack("<Q", int(x))

    def fe(x):
        return int(x).to_bytes(32, "little")

    out = [u64(CIRCUIT2_MAGIC), u64(len(segments))]
    for s in segments:
        out += [u64(s["i_len"]), u64(s["o_len"]), u64(len(s.get("children", [])))]
        for sid, allocs in s.get("children", []):
            out += [u64(sid), u64(len(allocs))] + [u64(v) for a in allocs for v in a]
        muls, adds, consts = s.get("mul", []), s.get("add", []), s.get("const", [])
        out.append(u64(len(muls)))
        for g in muls:
            out += [u64(g[0]), u64(g[1]), u64(g[2]), fe(g[3])]
        out.append(u64(len(adds)))
        for g in adds:
            out += [u64(g[0]), u64(g[1]), fe(g[2])]
        out.append(u64(len(consts)))
        for g in consts:
            out += [u64(g[0]), fe(g[1])]
        rand = s.get("rand", [])
        out += [u64(len(rand))] + [u64(i) for i in rand]
    out += [u64(len(layers))] + [u64(l) for l in layers] + [fe(rand_sentinel)]
    with open(path, "wb") as f:
        f.write(b"".join(out))


def circuit2_to_text(path: str, mul_path: str, add_path: str, field_type: int = FIELD_M31,
                     coef_mode: int = 0) -> None:
    """Flatten a CIRCUIT2 file into the text format with no GPU (gkrx_circuit2_to_text)."""
    _check(lib().gkrx_circuit2_to_text(path.encode(), field_type, coef_mode, mul_path.encode(), add_path.encode()),
           "gkrx_circuit2_to_text")


def parse_text_shape(mul_path: str, add_path: str, field_type: int = FIELD_M31):
    """Host-only read and check of a text-format circuit (gkrx_circuit_parse_text): returns
    (nb_input_vars, nb_output_vars) per layer, layer 0 first."""
    n = C.c_uint32(0)
    _check(lib().gkrx_circuit_parse_text(mul_path.encode(), add_path.encode(), field_type, C.byref(n), None, None),
           "gkrx_circuit_parse_text")
    ni = np.zeros(n.value, dtype=np.uint32)
    no = np.zeros(n.value, dtype=np.uint32)
    _check(lib().gkrx_circuit_parse_text(mul_path.encode(), add_path.encode(), field_type, C.byref(n),
                                         ni.ctypes.data, no.ctypes.data), "gkrx_circuit_parse_text")
    return ni.tolist(), no.tolist()


def read_text(mul_path: str, add_path: str):
    """The text format back into (mul[u,4], add[a,3]) per layer, layer 0 first."""
    def parse(p, w):
        lines = [ln for ln in open(p).read().split("\n") if ln.strip()]
        out = []
        for ln in lines:
            t = np.array([int(x) for x in ln.split()], dtype=np.int64)
            out.append(t[1:].reshape(-1, w).astype(np.uint32))
        return out[::-1]
    return list(zip(parse(mul_path, 4), parse(add_path, 3)))


def kernel_stats(ctx: "Context") -> dict:
    """{kernel class: {"count": launches, "ms": total ms}} recorded while profiling was on."""
    out = {}
    i = 0
    while True:
        name, cnt, ms = C.c_char_p(), C.c_uint64(), C.c_double()
        if lib().gkrx_ctx_kernel_stat(ctx._h, i, C.byref(name), C.byref(cnt), C.byref(ms)) != 0:
            break
        if cnt.value:
            out[name.value.decode()] = {"count": int(cnt.value), "ms": float(ms.value)}
        i += 1
    return out


def kernel_class_index(ctx: "Context", name: str) -> int:
    """Index of a kernel class by name (for gkrx_ctx_profile_classes masks)."""
    i = 0
    while True:
        nm = C.c_char_p()
        if lib().gkrx_ctx_kernel_stat(ctx._h, i, C.byref(nm), None, None) != 0:
            raise KeyError(name)
        if nm.value.decode() == name:
            return i
        i += 1


def host_hash(fs_hash: int, data: bytes) -> bytes:
    out = C.create_string_buffer(32)
    _check(lib().gkrx_hash(fs_hash, data, len(data), out), "gkrx_hash")
    return out.raw


def host_hash_many(fs_hash: int, msgs) -> tuple:
    """Hashes of equal-length messages at once; returns (digests, eight_lane_path_ran)."""
    n = len(msgs)
    ln = len(msgs[0]) if n else 0
    assert all(len(m) == ln for m in msgs)
    bufs = [C.create_string_buffer(bytes(m), max(ln, 1)) for m in msgs]
    ptrs = (C.c_char_p * n)(*[C.cast(b, C.c_char_p) for b in bufs])
    out = C.create_string_buffer(32 * max(n, 1))
    rc = _check(lib().gkrx_hash_many(fs_hash, n, ptrs, ln, out), "gkrx_hash_many")
    return [out.raw[32 * i:32 * i + 32] for i in range(n)], rc == 1


# ----------------------------------------------------------------- API mirror
class Config:
    """gkr::Config (include/configuration/config.hpp:26-86); nb_parallel = lanes."""

    def __init__(self, fs_hash: int = FS_SHA256, nb_parallel: int = 8, field_type: int = FIELD_M31,
                 grinding_bits: int = 10, security_bits: int = 100, pc_type: int = PC_RAW):
        self._c = _Config()
        lib().gkrx_config_default(C.byref(self._c))
        self._c.fs_hash = fs_hash
        self._c.nb_parallel = nb_parallel
        self._c.field_type = field_type
        self._c.grinding_bits = grinding_bits
        self._c.security_bits = security_bits
        self._c.pc_type = pc_type

    @property
    def lanes(self) -> int:
        """Data-parallel lanes of a field element (BN254: one)."""
        return 1 if self._c.field_type == FIELD_BN254 else self._c.nb_parallel

    @property
    def words(self) -> int:
        """u32 words of one field element: the lanes for M31, 8 for a BN254 scalar."""
        return 8 if self._c.field_type == FIELD_BN254 else self._c.nb_parallel

    def get_num_repetitions(self) -> int:
        return _check(lib().gkrx_config_num_repetitions(C.byref(self._c)), "num_repetitions")


class Context:
    """A prover context bound to one GPU (Prover::Prover, LinearGKR.hpp:42-47)."""

    def __init__(self, config: Config, device: int = 0):
        self.config = config
        self._h = None
        self._circuits = weakref.WeakSet()
        self._servers = weakref.WeakSet()
        h = C.c_void_p()
        _check(lib().gkrx_ctx_create(C.byref(config._c), device, C.byref(h)), "gkrx_ctx_create")
        self._h = h

    def close(self):
        """Destroys the context after every circuit created on it."""
        if self._h:
            for s in list(self._servers):
                s.close()
            for c in list(self._circuits):
                c.close()
            _check(lib().gkrx_ctx_destroy(self._h), "gkrx_ctx_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_batch_engine(self, engine: int, block: int = 256) -> None:
        """BATCH_PER_PROOF, BATCH_LOCKSTEP or BATCH_LOCKSTEP_GROUPS (block 64, 128, 256 or 512) for prove_batch."""
        _check(lib().gkrx_ctx_set_batch_engine(self._h, engine, block), "gkrx_ctx_set_batch_engine")

    def batch_engine(self):
        e, b = C.c_int(), C.c_uint32()
        _check(lib().gkrx_ctx_get_batch_engine(self._h, C.byref(e), C.byref(b)), "gkrx_ctx_get_batch_engine")
        return e.value, b.value

    def batch_autotune(self, circuits):
        """Times every batch engine / block size / lanes per thread on these circuits (host
        prefixes included, as prove_batch runs) and keeps the fastest: among settings within 1%
        of the fastest, the first in the order of the returned dict.  Returns (engine, block,
        lanes, {(engine, block, lanes): ms})."""
        n = len(circuits)
        hs = (C.c_void_p * n)(*[c._h.value for c in circuits])
        e, b, v = C.c_int(), C.c_uint32(), C.c_uint32()
        ms = (C.c_double * 25)()
        _check(lib().gkrx_batch_autotune(self._h, n, hs, C.byref(e), C.byref(b), C.byref(v), ms),
               "gkrx_batch_autotune")
        keys = [(BATCH_PER_PROOF, 256, 4)] + [(e, x, y) for e in (BATCH_LOCKSTEP, BATCH_LOCKSTEP_GROUPS)
                                              for x in (64, 128, 256, 512) for y in (1, 2, 4)]
        return e.value, b.value, v.value, {k: ms[i] for i, k in enumerate(keys)}

    def set_batch_lanes(self, lanes: int) -> None:
        """Lanes per thread (1, 2 or 4) of the lockstep engine's round-pair kernel."""
        _check(lib().gkrx_ctx_set_batch_lanes(self._h, lanes), "gkrx_ctx_set_batch_lanes")

    def batch_lanes(self) -> int:
        v = C.c_uint32()
        _check(lib().gkrx_ctx_get_batch_lanes(self._h, C.byref(v)), "gkrx_ctx_get_batch_lanes")
        return v.value

    def set_graphs(self, on: bool) -> None:
        """Replay the device part of Prover.prove as a HIP graph captured per circuit."""
        _check(lib().gkrx_ctx_set_graphs(self._h, 1 if on else 0), "gkrx_ctx_set_graphs")

    def graphs(self) -> bool:
        v = C.c_int()
        _check(lib().gkrx_ctx_get_graphs(self._h, C.byref(v)), "gkrx_ctx_get_graphs")
        return bool(v.value)


class Circuit:
    """gkr::Circuit on the device."""

    def __init__(self, ctx: Context, handle):
        self.ctx = ctx
        self._h = handle
        ctx._circuits.add(self)
        n = C.c_uint32()
        _check(lib().gkrx_circuit_shape(handle, C.byref(n), None, None), "shape")
        self.nb_layers = n.value
        ni = np.zeros(n.value, dtype=np.uint32)
        no = np.zeros(n.value, dtype=np.uint32)
        _check(lib().gkrx_circuit_shape(handle, C.byref(n), ni.ctypes.data, no.ctypes.data), "shape")
        self.nb_input_vars = [int(x) for x in ni]
        self.nb_output_vars = [int(x) for x in no]

    @classmethod
    def load_extracted_gates(cls, ctx: Context, mul_path: str, add_path: str) -> "Circuit":
        h = C.c_void_p()
        _check(lib().gkrx_circuit_load_text(ctx._h, mul_path.encode(), add_path.encode(), C.byref(h)),
               "gkrx_circuit_load_text")
        return cls(ctx, h)

    @classmethod
    def load_circuit2(cls, ctx: Context, path: str, coef_mode: int = 0) -> "Circuit":
        """CircuitRaw >> + Circuit::from_circuit_raw (circuit_raw.hpp:539-592,
        circuit.hpp:217-251).  coef_mode 0: every coefficient 1, as the reference reads them."""
        h = C.c_void_p()
        _check(lib().gkrx_circuit_load_circuit2(ctx._h, path.encode(), coef_mode, C.byref(h)),
               "gkrx_circuit_load_circuit2")
        return cls(ctx, h)

    @classmethod
    def from_layers(cls, ctx: Context, layers) -> "Circuit":
        D = len(layers)
        muls = [np.ascontiguousarray(m, dtype=np.uint32).ravel() for m, _ in layers]
        adds = [np.ascontiguousarray(a, dtype=np.uint32).ravel() for _, a in layers]
        nm = np.array([len(m) // 4 for m in muls], dtype=np.uint32)
        na = np.array([len(a) // 3 for a in adds], dtype=np.uint32)
        mp = (C.c_void_p * D)(*[m.ctypes.data for m in muls])
        ap = (C.c_void_p * D)(*[a.ctypes.data for a in adds])
        h = C.c_void_p()
        _check(lib().gkrx_circuit_from_arrays(ctx._h, D, nm, mp, na, ap, C.byref(h)), "gkrx_circuit_from_arrays")
        return cls(ctx, h)

    def set_witness(self, lanes: np.ndarray) -> None:
        w = np.ascontiguousarray(lanes, dtype=np.uint32).ravel()
        _check(lib().gkrx_circuit_set_witness(self._h, w, w.size), "gkrx_circuit_set_witness")

    def evaluate(self) -> None:
        _check(lib().gkrx_circuit_evaluate(self._h), "gkrx_circuit_evaluate")

    def layer_values(self, layer: int) -> np.ndarray:
        L = self.ctx.config.words
        vars_ = self.nb_input_vars[layer] if layer < self.nb_layers else self.nb_output_vars[-1]
        out = np.empty((1 << vars_) * L, dtype=np.uint32)
        _check(lib().gkrx_circuit_get_layer_values(self._h, layer, out, out.size), "get_layer_values")
        return out

    def output(self) -> np.ndarray:
        return self.layer_values(self.nb_layers)

    def write_text(self, mul_path: str, add_path: str) -> None:
        _check(lib().gkrx_circuit_write_text(self._h, mul_path.encode(), add_path.encode()), "write_text")

    def close(self):
        if self._h:
            lib().gkrx_circuit_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Prover:
    """gkr::Prover::prove (LinearGKR.hpp:58-90): returns (claimed_v[R, L], proof bytes)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx

    def proof_size(self, circuit: Circuit) -> int:
        return lib().gkrx_proof_size(self.ctx._h, circuit._h)

    def prove(self, circuit: Circuit, out: np.ndarray | None = None):
        size = self.proof_size(circuit)
        buf = out if out is not None else np.empty(size, dtype=np.uint8)
        R = self.ctx.config.get_num_repetitions()
        claimed = np.empty(R * self.ctx.config.words, dtype=np.uint32)
        n = C.c_size_t()
        _check(lib().gkrx_prove(self.ctx._h, circuit._h, buf, buf.size, C.byref(n), claimed), "gkrx_prove")
        return claimed.reshape(R, -1), buf[:n.value]

    def prove_batch(self, circuits, outs=None):
        """All circuits in one batched launch; returns [(claimed[R, L], proof)] in order.
        Each proof is byte-identical to prove() of the same circuit."""
        n = len(circuits)
        R, L = self.ctx.config.get_num_repetitions(), self.ctx.config.words
        bufs = outs if outs is not None else [np.empty(self.proof_size(c), dtype=np.uint8) for c in circuits]
        for b in bufs:
            if b.dtype != np.uint8 or not b.flags.c_contiguous:
                raise ValueError("proof buffers must be contiguous uint8 arrays")
        # the pointer tables as numpy arrays (one allocation each, no per-proof ctypes
        # objects); claimed values of proof i are row i of one [n, R, L] array
        claims = np.empty((n, R, L), dtype=np.uint32)
        hs = np.fromiter((c._h.value for c in circuits), dtype=np.uint64, count=n)
        ps = np.fromiter((b.ctypes.data for b in bufs), dtype=np.uint64, count=n)
        caps = np.fromiter((b.size for b in bufs), dtype=np.uint64, count=n)
        lens = np.zeros(n, dtype=np.uint64)
        cs = np.uint64(claims.ctypes.data) + np.arange(n, dtype=np.uint64) * np.uint64(R * L * 4)
        _check(lib().gkrx_prove_batch(self.ctx._h, n, hs.ctypes.data_as(C.POINTER(C.c_void_p)),
                                      ps.ctypes.data_as(C.POINTER(C.c_void_p)),
                                      caps.ctypes.data_as(C.POINTER(C.c_size_t)),
                                      lens.ctypes.data_as(C.POINTER(C.c_size_t)),
                                      cs.ctypes.data_as(C.POINTER(C.c_void_p))), "gkrx_prove_batch")
        ln = lens.tolist()
        return [(claims[i], bufs[i][:ln[i]]) for i in range(n)]

    def last_timing(self):
        ms = (C.c_double * 3)()
        _check(lib().gkrx_last_timing(self.ctx._h, ms), "last_timing")
        return {"host_prefix_ms": ms[0], "device_ms": ms[1], "total_ms": ms[2]}


class Verifier:
    """gkr::Verifier::verify (LinearGKR.hpp:104-149), on the host."""

    def __init__(self, ctx: Context):
        self.ctx = ctx

    def verify(self, circuit: Circuit, claimed_v: np.ndarray, proof) -> bool:
        cv = np.ascontiguousarray(claimed_v, dtype=np.uint32).ravel()
        pb = np.frombuffer(bytes(proof), dtype=np.uint8) if not isinstance(proof, np.ndarray) else proof
        pb = np.ascontiguousarray(pb, dtype=np.uint8)
        return _check(lib().gkrx_verify(self.ctx._h, circuit._h, cv, pb, pb.size), "gkrx_verify") == 1


def verify_layers(config: Config, layers, claimed_v, proof) -> bool:
    """Verifier::verify on gate arrays, with no context and no GPU (gkrx_verify_arrays)."""
    D = len(layers)
    muls = [np.ascontiguousarray(m, dtype=np.uint32).ravel() for m, _ in layers]
    adds = [np.ascontiguousarray(a, dtype=np.uint32).ravel() for _, a in layers]
    nm = np.array([len(m) // 4 for m in muls], dtype=np.uint32)
    na = np.array([len(a) // 3 for a in adds], dtype=np.uint32)
    mp = (C.c_void_p * D)(*[m.ctypes.data for m in muls])
    ap = (C.c_void_p * D)(*[a.ctypes.data for a in adds])
    cv = np.ascontiguousarray(claimed_v, dtype=np.uint32).ravel()
    pb = np.ascontiguousarray(np.frombuffer(bytes(proof), dtype=np.uint8))
    return _check(lib().gkrx_verify_arrays(C.byref(config._c), D, nm, mp, na, ap, cv, pb, pb.size),
                  "gkrx_verify_arrays") == 1


class Server:
    """Continuous-batching prove server on one context (gkrx_server_*): requests are
    admitted against an HBM budget and launched in batches on free slots as they arrive.
    Proofs are byte-identical to Prover.prove of the same circuit."""

    def __init__(self, ctx: Context, hbm_budget_bytes: int, slots: int = 4):
        self.ctx = ctx
        self._h = None
        h = C.c_void_p()
        _check(lib().gkrx_server_create(ctx._h, int(hbm_budget_bytes), slots, C.byref(h)), "gkrx_server_create")
        self._h = h
        self._pending = {}  # ticket -> (circuit, proof buffer, claims buffer)
        ctx._servers.add(self)

    def submit(self, circuit: Circuit) -> int:
        size = lib().gkrx_proof_size(self.ctx._h, circuit._h)
        buf = np.empty(size, dtype=np.uint8)
        claims = np.empty(self.ctx.config.get_num_repetitions() * self.ctx.config.words, dtype=np.uint32)
        t = C.c_uint64()
        _check(lib().gkrx_server_submit(self._h, circuit._h, buf, buf.size, claims, C.byref(t)), "gkrx_server_submit")
        self._pending[t.value] = (circuit, buf, claims)
        return t.value

    SPLIT_KEYS = ("prefix_queue_ms", "prefix_ms", "admission_ms", "device_ms", "collect_ms", "total_ms")

    def wait(self, ticket: int, timeout_s: float | None = None, split: bool = False):
        """(claimed[R, words], proof, latency_ms), or None if still pending at the timeout.
        With split, a fourth item: the request's time split (SPLIT_KEYS, ms)."""
        n, sp = C.c_size_t(), (C.c_double * 6)()
        us = -1 if timeout_s is None else int(timeout_s * 1e6)
        rc = lib().gkrx_server_wait_split(self._h, ticket, us, C.byref(n), sp)
        if rc == 1:
            return None
        circuit, buf, claims = self._pending.pop(ticket)
        _check(rc, "gkrx_server_wait")
        out = (claims.reshape(self.ctx.config.get_num_repetitions(), -1), buf[:n.value], sp[5])
        if split:
            out += (dict(zip(self.SPLIT_KEYS, list(sp))),)
        return out

    def stats(self) -> dict:
        out = (C.c_uint64 * 6)()
        _check(lib().gkrx_server_stats(self._h, out), "gkrx_server_stats")
        keys = ("submitted", "completed", "batches", "max_batch", "peak_bytes", "admitted_bytes")
        return dict(zip(keys, [int(x) for x in out]))

    def close(self):
        if self._h:
            _check(lib().gkrx_server_destroy(self._h), "gkrx_server_destroy")
            self._h = None
            self._pending.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def proof_write(path: str, proof) -> None:
    """Proof file: [u32 byte count][bytes] (Proof::write_and_compress / read_proof,
    include/fiat_shamir/transcript.hpp:54-85)."""
    b = np.ascontiguousarray(np.frombuffer(bytes(proof), dtype=np.uint8))
    _check(lib().gkrx_proof_write(path.encode(), b, b.size), "gkrx_proof_write")


def proof_read(path: str) -> bytes:
    n = C.c_size_t()
    _check(lib().gkrx_proof_read(path.encode(), None, 0, C.byref(n)), "gkrx_proof_read")
    buf = np.empty(max(n.value, 1), dtype=np.uint8)
    _check(lib().gkrx_proof_read(path.encode(), buf.ctypes.data, buf.size, C.byref(n)), "gkrx_proof_read")
    return bytes(buf[:n.value])
