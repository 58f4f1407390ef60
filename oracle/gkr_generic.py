"""Field-generic pure-Python restatement of the reference prover -- TEST INFRASTRUCTURE
ONLY (small circuits; the product never imports it).

It restates the same algorithm as oracle/gkr.c over any field given as a `Field` object,
so that the BN254 scalar field -- which the reference names (include/configuration/
config.hpp:12-16) but does not implement -- gets the reference's transcript framing and
sumcheck unchanged.  Pinning: the M31 instantiation is checked byte for byte against the
C oracle and the reference-generated fixtures (tests/test_oracle_generic.py); the BN254
instantiation shares every line except the field, whose arithmetic is Python big-int
modular arithmetic (exact by construction).  BN254 parity against the reference is
therefore unpinned: the reference has no BN254 to compare with.

Build-defined BN254 conventions (no reference counterpart): an element serializes as its
canonical value in 32 little-endian bytes; a challenge is the 32 digest bytes read as a
little-endian integer mod r; grinding draws 256 // 254 = 1 challenge written into the
32-byte buffer (ceil(254 / 8) = 32).
"""
from __future__ import annotations

import hashlib
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "tests"))

P31 = (1 << 31) - 1
BN254_R = 0x30644e72e131a029b85045b68181585d2833e84879b9709143e1f593f0000001


class Field:
    """One data-parallel field: `lanes` independent lanes of a prime field."""

    def __init__(self, name, modulus, lanes, field_bits, elem_bytes):
        self.name, self.p, self.lanes, self.field_bits, self.elem_bytes = name, modulus, lanes, field_bits, elem_bytes
        self.S = lanes * elem_bytes

    def to_bytes(self, v):  # v: list of lanes
        return b"".join(int(x).to_bytes(self.elem_bytes, "little") for x in v)

    def from_bytes(self, b):
        # M31::from_bytes (include/field/M31.hpp:111-116) reduces fully; BN254: mod r
        return [int.from_bytes(b[i * self.elem_bytes:(i + 1) * self.elem_bytes], "little") % self.p
                for i in range(self.lanes)]

    def challenge(self, digest):
        # M31: LE32 of the first four digest bytes mod p (transcript.hpp:139-147)
        if self.name == "m31":
            return int.from_bytes(digest[:4], "little") % self.p
        return int.from_bytes(digest, "little") % self.p

    def coef(self, c):
        # M31(uint32) folds once (M31.hpp:58-62), congruent to c; BN254: c mod r
        return c % self.p

    def repetitions(self, security=100, grinding=10):
        # config.hpp:61-65
        n = security - grinding
        return n // self.field_bits + (1 if n % self.field_bits else 0)


def m31(lanes):
    return Field("m31", P31, lanes, 31, 4)


def bn254():
    return Field("bn254", BN254_R, 1, 254, 32)


def _hash(kind, data):
    if kind == 0:
        return hashlib.sha256(data).digest()
    import oracle_lib
    return oracle_lib.keccak256(data)


class Transcript:
    """include/fiat_shamir/transcript.hpp:87-175: non-chained challenges."""

    def __init__(self, F, kind):
        self.F, self.kind = F, kind
        self.bytes = bytearray()
        self.hash_start = 0
        self.digest = bytes(32)

    def append(self, b):
        self.bytes += b

    def challenge(self):
        if len(self.bytes) > self.hash_start:
            self.digest = _hash(self.kind, bytes(self.bytes[self.hash_start:]))
            self.hash_start = len(self.bytes)
        else:
            self.digest = _hash(self.kind, self.digest)
        return self.F.challenge(self.digest)


def load_text(mul_path, add_path):
    """Circuit::load_extracted_gates (include/circuit/circuit.hpp:152-215); returns layers
    (input layer first) of (mul [(in0, in1, out, coef)], add [(in0, out, coef)])."""
    mul_lines = open(mul_path).read().split("\n")
    if mul_lines and mul_lines[-1] == "":
        mul_lines.pop()
    D = len(mul_lines)
    toks = open(mul_path).read().split()
    at = 0
    layers = [None] * D
    for i in range(D):
        n = int(toks[at]); at += 1
        g = []
        for _ in range(n):
            g.append(tuple(int(x) for x in toks[at:at + 4])); at += 4
        layers[D - 1 - i] = [g, None]
    toks = open(add_path).read().split()
    at = 0
    for i in range(D):
        n = int(toks[at]); at += 1
        g = []
        for _ in range(n):
            g.append(tuple(int(x) for x in toks[at:at + 3])); at += 3
        layers[D - 1 - i][1] = g
    return layers


def nb_vars(max_id):
    # next_pow_of_2 (include/utils/myutil.hpp:38-48) then ctz, as _compute_nb_vars (:145-146)
    if max_id == 0:
        return 0
    p = max_id if (max_id & (max_id - 1)) == 0 else 1 << max_id.bit_length()
    return (p & -p).bit_length() - 1


def shape(layers):
    out = []
    for mul, add in layers:
        mi = max([max(g[0], g[1]) for g in mul] + [g[0] for g in add] + [0])
        mo = max([g[2] for g in mul] + [g[1] for g in add] + [0])
        out.append((nb_vars(mi), nb_vars(mo)))
    return out


def evaluate(F, layers, witness):
    """CircuitLayer::evaluate (circuit.hpp:92-105): per layer the input values; the last
    element is the output layer."""
    p, L = F.p, F.lanes
    vals = [witness]
    sh = shape(layers)
    for (mul, add), (n, m) in zip(layers, sh):
        cur = vals[-1]
        out = [[0] * L for _ in range(1 << m)]
        for i0, i1, o, c in mul:
            cc = F.coef(c)
            for k in range(L):
                out[o][k] = (out[o][k] + cur[i0][k] * cur[i1][k] * cc) % p
        for i0, o, c in add:
            cc = F.coef(c)
            for k in range(L):
                out[o][k] = (out[o][k] + cur[i0][k] * cc) % p
        vals.append(out)
    return vals


def eq_table(p, r, factor):
    """_eq_evals_at_primitive (include/LinearGKR/sumcheck_common.hpp:16-33)."""
    e = [factor % p]
    for ri in r:
        e = [x * (1 - ri) % p for x in e] + [x * ri % p for x in e]
    return e


def eval_ml(F, table, r):
    """eval_multilinear (include/poly_commit/poly.hpp:88-104): LSB-first folding."""
    p = F.p
    s = [list(v) for v in table]
    for ri in r:
        s = [[(s[2 * i][k] + (s[2 * i + 1][k] - s[2 * i][k]) * ri) % p for k in range(F.lanes)]
             for i in range(len(s) // 2)]
    return s[0]


def grind(F, t, grinding_bits=10):
    """grind (include/LinearGKR/LinearGKR.hpp:13-32)."""
    hb = bytearray(32)
    step = (F.field_bits + 7) // 8
    for i in range(256 // F.field_bits):
        h = t.challenge()
        hb[i * step:i * step + step] = int(h).to_bytes(step, "little")
    hb = bytes(hb)
    for _ in range(1 << grinding_bits):
        hb = _hash(t.kind, hb)
    t.append(hb)


def _sumcheck_layer(F, t, mul, add, n, m, V, rz1, rz2, alpha, beta, R):
    """sumcheck_prove_gkr_layer (include/LinearGKR/sumcheck.hpp:52-111) with the helper of
    include/LinearGKR/sumcheck_helper.hpp (gate_exists dropped: hg is zero there)."""
    p, L = F.p, F.lanes
    N = 1 << n
    st = []
    for j in range(R):
        e = [(a + b) % p for a, b in zip(eq_table(p, rz1[j], alpha), eq_table(p, rz2[j], beta))]
        hg = [[0] * L for _ in range(N)]
        for i0, i1, o, c in mul:
            ce = F.coef(c) * e[o] % p
            for k in range(L):
                hg[i0][k] = (hg[i0][k] + V[i1][k] * ce) % p
        for i0, o, c in add:
            ce = F.coef(c) * e[o] % p
            for k in range(L):
                hg[i0][k] = (hg[i0][k] + ce) % p
        st.append({"eqz": e, "f": [list(v) for v in V], "hg": hg, "rx": [], "ry": []})
    for i_var in range(2 * n):
        for j in range(R):
            s = st[j]
            if i_var == n:  # _prepare_phase_two (sumcheck_helper.hpp:263-272)
                vx = s["f"][0]
                ex = eq_table(p, s["rx"], 1)
                hg = [[0] * L for _ in range(N)]
                for i0, i1, o, c in mul:
                    sc = s["eqz"][o] * ex[i0] % p * F.coef(c) % p
                    for k in range(L):
                        hg[i1][k] = (hg[i1][k] + vx[k] * sc) % p
                s["hg"], s["f"] = hg, [list(v) for v in V]
            f, hg = s["f"], s["hg"]
            p0, p1, p2 = [0] * L, [0] * L, [0] * L
            for i in range(len(f) // 2):
                for k in range(L):
                    f0, f1, h0, h1 = f[2 * i][k], f[2 * i + 1][k], hg[2 * i][k], hg[2 * i + 1][k]
                    p0[k] = (p0[k] + f0 * h0) % p
                    p1[k] = (p1[k] + f1 * h1) % p
                    p2[k] = (p2[k] + (f0 + f1) * (h0 + h1)) % p
            p2 = [(6 * p1[k] + 3 * p0[k] - 2 * p2[k]) % p for k in range(L)]
            t.append(F.to_bytes(p0) + F.to_bytes(p1) + F.to_bytes(p2))
            r = t.challenge()
            s["f"] = [[(f[2 * i][k] + (f[2 * i + 1][k] - f[2 * i][k]) * r) % p for k in range(L)]
                      for i in range(len(f) // 2)]
            s["hg"] = [[(hg[2 * i][k] + (hg[2 * i + 1][k] - hg[2 * i][k]) * r) % p for k in range(L)]
                       for i in range(len(hg) // 2)]
            (s["rx"] if i_var < n else s["ry"]).append(r)
            if i_var == n - 1:
                t.append(F.to_bytes(s["f"][0]))
    for j in range(R):
        t.append(F.to_bytes(st[j]["f"][0]))
    return [s["rx"] for s in st], [s["ry"] for s in st]


def prove(F, layers, witness, kind):
    """Prover::prove (include/LinearGKR/LinearGKR.hpp:58-90) -> (claimed_v, proof bytes).
    witness: 2^n0 elements, each a list of F.lanes ints."""
    p = F.p
    R = F.repetitions()
    sh = shape(layers)
    vals = evaluate(F, layers, witness)
    t = Transcript(F, kind)
    t.append(b"".join(F.to_bytes(v) for v in witness))
    grind(F, t)
    mlast = sh[-1][1]
    rz1 = [[] for _ in range(R)]
    rz2 = [[] for _ in range(R)]
    for _ in range(mlast):
        for j in range(R):
            rz1[j].append(t.challenge())
            rz2[j].append(0)
    alpha, beta = 1, 0
    claimed = [eval_ml(F, vals[-1], rz1[j]) for j in range(R)]
    for l in range(len(layers) - 1, -1, -1):
        n, m = sh[l]
        rz1, rz2 = _sumcheck_layer(F, t, layers[l][0], layers[l][1], n, m, vals[l], rz1, rz2, alpha, beta, R)
        alpha = t.challenge()
        beta = t.challenge()
    return claimed, bytes(t.bytes)


def verify(F, layers, claimed, proof, kind):
    """Verifier::verify (include/LinearGKR/LinearGKR.hpp:104-149)."""
    p, L, S = F.p, F.lanes, F.S
    R = F.repetitions()
    inv2 = (p + 1) // 2
    sh = shape(layers)
    n0 = sh[0][0]
    N0 = 1 << n0
    commit = [F.from_bytes(proof[i * S:(i + 1) * S]) for i in range(N0)]
    t = Transcript(F, kind)
    t.append(proof[:N0 * S])
    grind(F, t)
    idx = N0 * S + 32
    mlast = sh[-1][1]
    rz1 = [[] for _ in range(R)]
    rz2 = [[] for _ in range(R)]
    for _ in range(mlast):
        for j in range(R):
            rz1[j].append(t.challenge())
            rz2[j].append(0)
    alpha, beta = 1, 0
    cv1 = [list(c) for c in claimed]
    cv2 = [[0] * L for _ in range(R)]
    ok = True

    def nxt():
        nonlocal idx
        v = F.from_bytes(proof[idx:idx + S])
        idx += S
        return v

    for l in range(len(layers) - 1, -1, -1):
        n, m = sh[l]
        mul, add = layers[l]
        sums = [[(cv1[j][k] * alpha + cv2[j][k] * beta) % p for k in range(L)] for j in range(R)]
        rx = [[] for _ in range(R)]
        ry = [[] for _ in range(R)]
        vx = [None] * R
        for i_var in range(2 * n):
            for j in range(R):
                ev = [nxt(), nxt(), nxt()]
                t.append(b"".join(F.to_bytes(e) for e in ev))
                r = t.challenge()
                (rx if i_var < n else ry)[j].append(r)
                for k in range(L):
                    ok &= (ev[0][k] + ev[1][k]) % p == sums[j][k]
                    c2 = inv2 * (ev[2][k] - 2 * ev[1][k] + ev[0][k]) % p
                    c1 = (ev[1][k] - ev[0][k] - c2) % p
                    sums[j][k] = (ev[0][k] + (c2 * r + c1) * r) % p
                if i_var == n - 1:
                    vx[j] = nxt()
                    e1, e2, ex = eq_table(p, rz1[j], alpha), eq_table(p, rz2[j], beta), eq_table(p, rx[j], 1)
                    a = sum((e1[o] + e2[o]) * ex[i0] * F.coef(c) for i0, o, c in add) % p
                    sums[j] = [(sums[j][k] - vx[j][k] * a) % p for k in range(L)]
                    t.append(F.to_bytes(vx[j]))
        vy = []
        for j in range(R):
            vy.append(nxt())
            e1, e2 = eq_table(p, rz1[j], alpha), eq_table(p, rz2[j], beta)
            ex, ey = eq_table(p, rx[j], 1), eq_table(p, ry[j], 1)
            mv = sum((e1[o] + e2[o]) * ex[i0] * ey[i1] * F.coef(c) for i0, i1, o, c in mul) % p
            for k in range(L):
                ok &= sums[j][k] == vx[j][k] * vy[j][k] * mv % p
            t.append(F.to_bytes(vy[j]))
        alpha = t.challenge()
        beta = t.challenge()
        rz1, rz2, cv1, cv2 = rx, ry, vx, vy
    for j in range(R):
        ok &= eval_ml(F, commit, rz1[j]) == cv1[j]
        ok &= eval_ml(F, commit, rz2[j]) == cv2[j]
    return bool(ok)
