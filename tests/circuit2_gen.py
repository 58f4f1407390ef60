"""Seeded hierarchical CIRCUIT2 circuits for the loader tests: leaf segments with mul, add
and constant gates (some with random coefficients), middle segments that place leaves at
several offsets, and one top segment per layer that places middle segments and carries
one gate at the largest input and output index, so the flattened layer has exactly the
intended variable counts."""
import random


def make(seed: int, depth: int, log_width: int):
    rnd = random.Random(seed)
    W = 1 << log_width
    segments, layers = [], []
    for _ in range(depth):
        # leaf: 2^a inputs -> 2^a outputs, a few gates
        a = max(1, log_width - 3)
        leaf_len = 1 << a
        leaf = {"i_len": leaf_len, "o_len": leaf_len, "mul": [], "add": [], "const": [], "rand": []}
        for o in range(leaf_len):
            coef = rnd.randrange(1, 1 << 40)
            if rnd.random() < 0.5:
                leaf["mul"].append((rnd.randrange(leaf_len), rnd.randrange(leaf_len), o, coef))
            else:
                leaf["add"].append((rnd.randrange(leaf_len), o, coef))
        leaf["const"].append((rnd.randrange(leaf_len), rnd.randrange(1, 100)))
        nm, na = len(leaf["mul"]), len(leaf["add"])
        leaf["rand"] = sorted(rnd.sample(range(nm + na + 1), 3))
        leaf_id = len(segments)
        segments.append(leaf)
        # middle: two leaves side by side
        mid_len = 2 * leaf_len
        mid = {"i_len": mid_len, "o_len": mid_len, "children": [(leaf_id, [(0, 0), (leaf_len, leaf_len)])]}
        mid_id = len(segments)
        segments.append(mid)
        # top: the middle segment at every offset, plus one gate at the corner indices
        allocs = [(off, off) for off in range(0, W, mid_len)]
        top = {"i_len": W, "o_len": W, "children": [(mid_id, allocs)],
               "add": [(W - 1, W - 1, rnd.randrange(1, 1 << 20))]}
        if rnd.random() < 0.5:
            top["children"].append((leaf_id, [(rnd.randrange(0, W - leaf_len + 1), 0)]))
        layers.append(len(segments))
        segments.append(top)
    return segments, layers, rnd.randrange(1, 1 << 60)


def flatten(segments, layers, mode: int, sentinel: int, p: int = (1 << 31) - 1):
    """The flattening restated in Python (constant gates dropped; mode 0: every
    coefficient 1; mode 1: coefficients reduced mod p, random ones take the sentinel)."""
    def coef(c):
        return 1 if mode == 0 else c % p

    def leaves(sid):
        s = segments[sid]
        out = {}
        if s.get("mul") or s.get("add") or s.get("const"):
            out.setdefault(sid, []).append((0, 0))
        for cid, allocs in s.get("children", []):
            for lid, las in leaves(cid).items():
                for la in las:
                    for ca in allocs:
                        out.setdefault(lid, []).append((ca[0] + la[0], ca[1] + la[1]))
        return out

    result = []
    for lsid in layers:
        mul, add = [], []
        for lid, allocs in leaves(lsid).items():
            s = segments[lid]
            rand = set(s.get("rand", []))
            nm = len(s.get("mul", []))
            for io, oo in allocs:
                for k, (i0, i1, o, c) in enumerate(s.get("mul", [])):
                    mul.append((i0 + io, i1 + io, o + oo, coef(sentinel if k in rand else c)))
                for k, (i0, o, c) in enumerate(s.get("add", [])):
                    add.append((i0 + io, o + oo, coef(sentinel if (k + nm) in rand else c)))
        result.append((sorted(mul), sorted(add)))
    return result

