// CIRCUIT2 binary circuits: CircuitRaw operator>> (include/circuit/circuit_raw.hpp:539-592,
// Segment operator>> :296-372) and the flattening of Circuit::from_circuit_raw
// (include/circuit/circuit.hpp:217-251, Segment::scan_leaf_segments :84-117 and
// leaf_gate_muls / leaf_gate_adds :144-183).
//
// File layout (all integers u64 little-endian, a field element 32 bytes):
//   magic "CIRCUIT2"; #segments; per segment: i_len, o_len, #children, per child
//   (segment id, #allocations, per allocation (i_offset, o_offset)); #mul gates, per gate
//   (in0, in1, out, coef); #add gates, per gate (in0, out, coef); #const gates, per gate
//   (out, coef); #random-coefficient indices, the indices; then #layers, the layer segment
//   ids, and the random-coefficient sentinel element.
//
// A layer is the gates of every leaf segment reachable from its segment (a segment with
// gates is a leaf of itself at offset 0), each placed at the sum of the allocation offsets
// on the path; constant gates are dropped, as from_circuit_raw does.
//
// Coefficients: the reference's element reader (include/field/M31.hpp:124-135) consumes the
// 32 bytes and then sets the value to 1, so every coefficient of a CIRCUIT2 circuit --
// including the random ones, which become the sentinel read the same way -- is 1.  Mode 0
// keeps that behaviour (same circuit, same proofs as the reference); mode 1 reads the 32
// bytes as a little-endian integer reduced mod the field (M31) or requires it below 2^32
// (BN254, whose gate coefficients are u32 here).
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "gkrx.h"
#include "internal.hpp"

namespace gkrx {

namespace {

constexpr uint64_t kMagic = 3626604230490605891ull;  // b"CIRCUIT2" little-endian

struct Reader {
    const uint8_t *p, *end;
    bool ok = true;
    uint64_t u64()
    {
        if (end - p < 8) {
            ok = false;
            return 0;
        }
        uint64_t v;
        memcpy(&v, p, 8);
        p += 8;
        return v;
    }
    const uint8_t *bytes(size_t n)
    {
        if ((size_t)(end - p) < n) {
            ok = false;
            return nullptr;
        }
        const uint8_t *q = p;
        p += n;
        return q;
    }
    // a count whose records take at least `rec` bytes each: reject counts the file cannot hold
    uint64_t count(size_t rec)
    {
        const uint64_t n = u64();
        if (ok && rec && n > (uint64_t)(end - p) / rec) ok = false;
        return ok ? n : 0;
    }
};

struct Seg {
    std::vector<std::pair<uint64_t, std::vector<std::pair<uint64_t, uint64_t>>>> children;
    std::vector<uint64_t> mul;   // in0, in1, out
    std::vector<uint64_t> add;   // in0, out
    std::vector<uint32_t> mulc, addc;
    std::vector<uint8_t> mul_rand, add_rand;
    bool has_gates = false;
};

// 32-byte element -> gate coefficient
bool read_coef(const uint8_t *b, int field, int mode, uint32_t &out)
{
    if (mode == 0) {
        out = 1;
        return true;
    }
    if (field == GKRX_FIELD_M31) {
        // little-endian 256-bit value mod 2^31 - 1, Horner from the top byte
        uint64_t acc = 0;
        for (int i = 31; i >= 0; i--) acc = (acc * 256 + b[i]) % 0x7fffffffu;
        out = (uint32_t)acc;
        return true;
    }
    for (int i = 4; i < 32; i++)
        if (b[i]) return false;
    memcpy(&out, b, 4);
    return true;
}

int parse(const std::vector<uint8_t> &buf, int field, int mode, std::vector<Seg> &segs,
          std::vector<uint64_t> &layers)
{
    Reader r{buf.data(), buf.data() + buf.size()};
    if (r.u64() != kMagic || !r.ok) {
        set_error("not a CIRCUIT2 file (bad magic)");
        return GKRX_ERR_CIRCUIT;
    }
    const uint64_t nseg = r.count(48);
    segs.resize(nseg);
    for (uint64_t s = 0; s < nseg && r.ok; s++) {
        Seg &g = segs[s];
        r.u64();  // i_len
        r.u64();  // o_len
        const uint64_t nch = r.count(16);
        g.children.resize(nch);
        for (uint64_t c = 0; c < nch && r.ok; c++) {
            g.children[c].first = r.u64();
            const uint64_t na = r.count(16);
            g.children[c].second.resize(na);
            for (uint64_t a = 0; a < na && r.ok; a++) {
                g.children[c].second[a].first = r.u64();
                g.children[c].second[a].second = r.u64();
            }
        }
        const uint64_t nm = r.count(56);
        g.mul.resize(3 * nm);
        g.mulc.resize(nm);
        g.mul_rand.assign(nm, 0);
        for (uint64_t i = 0; i < nm && r.ok; i++) {
            for (int k = 0; k < 3; k++) g.mul[3 * i + k] = r.u64();
            const uint8_t *cb = r.bytes(32);
            if (cb && !read_coef(cb, field, mode, g.mulc[i])) {
                set_error("coefficient above 2^32 for a BN254 circuit");
                return GKRX_ERR_CIRCUIT;
            }
        }
        const uint64_t na = r.count(48);
        g.add.resize(2 * na);
        g.addc.resize(na);
        g.add_rand.assign(na, 0);
        for (uint64_t i = 0; i < na && r.ok; i++) {
            for (int k = 0; k < 2; k++) g.add[2 * i + k] = r.u64();
            const uint8_t *cb = r.bytes(32);
            if (cb && !read_coef(cb, field, mode, g.addc[i])) {
                set_error("coefficient above 2^32 for a BN254 circuit");
                return GKRX_ERR_CIRCUIT;
            }
        }
        const uint64_t nc = r.count(40);
        for (uint64_t i = 0; i < nc && r.ok; i++) {
            r.u64();
            r.bytes(32);
        }
        const uint64_t nr = r.count(8);
        for (uint64_t i = 0; i < nr && r.ok; i++) {
            const uint64_t idx = r.u64();
            if (idx < nm) g.mul_rand[idx] = 1;
            else if (idx < nm + na) g.add_rand[idx - nm] = 1;
            else if (idx >= nm + na + nc) {
                set_error("invalid random coefficient index");  // the reference exits (:374-375)
                return GKRX_ERR_CIRCUIT;
            }
        }
        g.has_gates = nm || na || nc;
    }
    const uint64_t nl = r.count(8);
    layers.resize(nl);
    for (uint64_t i = 0; i < nl && r.ok; i++) layers[i] = r.u64();
    const uint8_t *sent = r.bytes(32);
    if (!r.ok || !sent) {
        set_error("truncated CIRCUIT2 file");
        return GKRX_ERR_CIRCUIT;
    }
    uint32_t rand_coef = 1;
    if (!read_coef(sent, field, mode, rand_coef)) {
        set_error("random-coefficient sentinel above 2^32 for a BN254 circuit");
        return GKRX_ERR_CIRCUIT;
    }
    // random coefficients take the file's sentinel value (circuit_raw.hpp:563-587)
    for (auto &g : segs) {
        for (size_t i = 0; i < g.mulc.size(); i++)
            if (g.mul_rand[i]) g.mulc[i] = rand_coef;
        for (size_t i = 0; i < g.addc.size(); i++)
            if (g.add_rand[i]) g.addc[i] = rand_coef;
    }
    for (uint64_t l : layers)
        if (l >= nseg) {
            set_error("layer refers to a missing segment");
            return GKRX_ERR_CIRCUIT;
        }
    for (auto &g : segs)
        for (auto &ch : g.children)
            if (ch.first >= nseg) {
                set_error("child refers to a missing segment");
                return GKRX_ERR_CIRCUIT;
            }
    return GKRX_OK;
}

using Leaves = std::map<uint64_t, std::vector<std::pair<uint64_t, uint64_t>>>;

// scan_leaf_segments: leaf segment -> every (input, output) offset it is placed at
int scan(const std::vector<Seg> &segs, uint64_t cur, uint32_t depth, std::vector<Leaves> &memo,
         std::vector<uint8_t> &done)
{
    if (done[cur]) return GKRX_OK;
    if (depth > segs.size()) {
        set_error("segment graph has a cycle");
        return GKRX_ERR_CIRCUIT;
    }
    Leaves out;
    if (segs[cur].has_gates) out[cur].push_back({0, 0});
    for (const auto &ch : segs[cur].children) {
        int rc = scan(segs, ch.first, depth + 1, memo, done);
        if (rc) return rc;
        for (const auto &leaf : memo[ch.first])
            for (const auto &la : leaf.second)
                for (const auto &ca : ch.second) out[leaf.first].push_back({ca.first + la.first, ca.second + la.second});
    }
    memo[cur] = std::move(out);
    done[cur] = 1;
    return GKRX_OK;
}

}  // namespace

int circuit2_flatten(const char *path, int field, int mode, std::vector<std::vector<uint32_t>> &mul,
                     std::vector<std::vector<uint32_t>> &add)
{
    if (mode != 0 && mode != 1) return GKRX_ERR_ARG;
    FILE *f = fopen(path, "rb");
    if (!f) {
        set_error(std::string("cannot open ") + path);
        return GKRX_ERR_IO;
    }
    std::vector<uint8_t> buf;
    uint8_t tmp[1 << 16];
    size_t n;
    while ((n = fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
    fclose(f);
    std::vector<Seg> segs;
    std::vector<uint64_t> layers;
    int rc = parse(buf, field, mode, segs, layers);
    if (rc) return rc;
    if (layers.empty()) {
        set_error("circuit has no layers");
        return GKRX_ERR_CIRCUIT;
    }
    std::vector<Leaves> memo(segs.size());
    std::vector<uint8_t> done(segs.size(), 0);
    const size_t D = layers.size();
    mul.assign(D, {});
    add.assign(D, {});
    for (size_t i = 0; i < D; i++) {
        rc = scan(segs, layers[i], 0, memo, done);
        if (rc) return rc;
        for (const auto &leaf : memo[layers[i]]) {
            const Seg &g = segs[leaf.first];
            for (const auto &al : leaf.second) {
                for (size_t k = 0; k < g.mulc.size(); k++) {
                    const uint64_t in0 = g.mul[3 * k] + al.first, in1 = g.mul[3 * k + 1] + al.first,
                                   out = g.mul[3 * k + 2] + al.second;
                    if ((in0 | in1 | out) >> 32) {
                        set_error("gate index above 2^32");
                        return GKRX_ERR_CIRCUIT;
                    }
                    mul[i].insert(mul[i].end(), {(uint32_t)in0, (uint32_t)in1, (uint32_t)out, g.mulc[k]});
                }
                for (size_t k = 0; k < g.addc.size(); k++) {
                    const uint64_t in0 = g.add[2 * k] + al.first, out = g.add[2 * k + 1] + al.second;
                    if ((in0 | out) >> 32) {
                        set_error("gate index above 2^32");
                        return GKRX_ERR_CIRCUIT;
                    }
                    add[i].insert(add[i].end(), {(uint32_t)in0, (uint32_t)out, g.addc[k]});
                }
            }
        }
    }
    return GKRX_OK;
}

}  // namespace gkrx

extern "C" int gkrx_circuit2_to_text(const char *path, int field_type, int coef_mode, const char *mul_path,
                                     const char *add_path)
try {
    if (!path || !mul_path || !add_path) return GKRX_ERR_ARG;
    if (field_type != GKRX_FIELD_M31 && field_type != GKRX_FIELD_BN254) return GKRX_ERR_CONFIG;
    std::vector<std::vector<uint32_t>> mul, add;
    int rc = gkrx::circuit2_flatten(path, field_type, coef_mode, mul, add);
    if (rc) return rc;
    for (int which = 0; which < 2; which++) {
        FILE *f = fopen(which == 0 ? mul_path : add_path, "w");
        if (!f) {
            gkrx::set_error("cannot write circuit file");
            return GKRX_ERR_IO;
        }
        // text format: one line per layer, output layer first (circuit.hpp:152-215)
        for (size_t i = 0; i < mul.size(); i++) {
            const auto &g = which == 0 ? mul[mul.size() - 1 - i] : add[add.size() - 1 - i];
            const size_t w = which == 0 ? 4 : 3;
            fprintf(f, "%zu", g.size() / w);
            for (size_t k = 0; k < g.size(); k++) fprintf(f, " %u", g[k]);
            fputc('\n', f);
        }
        fclose(f);
    }
    return GKRX_OK;
} catch (const std::bad_alloc &) {
    gkrx::set_error("out of host memory");
    return GKRX_ERR_NOMEM;
}
