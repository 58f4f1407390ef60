// Host verifier: Verifier::verify (include/LinearGKR/LinearGKR.hpp:104-149) with
// gkr_verify (include/LinearGKR/gkr.hpp:75-122), sumcheck_verify_gkr_layer
// (include/LinearGKR/sumcheck.hpp:113-176), degree_2_eval and
// eval_sparse_circuit_connect_poly (include/LinearGKR/sumcheck_verifier_utils.hpp:10-56)
// and RawPC::verify (include/poly_commit/raw.hpp:75-78).  The verifier is the
// acceptance check of the library; it is O(gates + 2^vars) per layer and stays on the
// host, as in the reference.
#include <cstring>
#include <vector>

#include "host_hash.hpp"
#include "m31.hpp"
#include "verifier.hpp"

namespace gkrx {

namespace {

struct HostTranscript {
    int kind;
    std::vector<uint8_t> bytes;
    size_t hash_start = 0;
    uint8_t digest[32] = {0};

    void append(const void *p, size_t n)
    {
        const uint8_t *b = static_cast<const uint8_t *>(p);
        bytes.insert(bytes.end(), b, b + n);
    }
    uint32_t challenge()
    {
        if (bytes.size() > hash_start) {
            host_hash(kind, bytes.data() + hash_start, bytes.size() - hash_start, digest);
            hash_start = bytes.size();
        } else {
            host_hash(kind, digest, 32, digest);
        }
        uint32_t x;
        memcpy(&x, digest, 4);
        return m31_from_u32_full(x);
    }
};

void eq_table(const uint32_t *r, uint32_t n, uint32_t factor, std::vector<uint32_t> &out)
{
    out.assign((size_t)1 << n, 0);
    out[0] = factor;
    size_t cur = 1;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t e1 = r[i], e0 = m31_sub(1u, r[i]);
        for (size_t j = 0; j < cur; j++) {
            out[j + cur] = m31_mul(out[j], e1);
            out[j] = m31_mul(out[j], e0);
        }
        cur <<= 1;
    }
}

// eval_multilinear (include/poly_commit/poly.hpp:88-104) = sum_z eq(r, z) vals[z]
void eval_ml(const uint32_t *vals, uint32_t L, const uint32_t *r, uint32_t n, uint32_t *out)
{
    std::vector<uint32_t> eq;
    eq_table(r, n, 1u, eq);
    std::vector<uint64_t> acc(L, 0);
    for (size_t z = 0; z < eq.size(); z++)
        for (uint32_t k = 0; k < L; k++) acc[k] += m31_mul(eq[z], vals[z * L + k]);
    for (uint32_t k = 0; k < L; k++) out[k] = m31_reduce64(acc[k]);
}

uint32_t degree2(uint32_t v0, uint32_t v1, uint32_t v2, uint32_t x)
{
    const uint32_t inv2 = 1u << 30;
    const uint32_t c2 = m31_mul(inv2, m31_add(m31_sub(v2, m31_mul(v1, 2u)), v0));
    const uint32_t c1 = m31_sub(m31_sub(v1, v0), c2);
    return m31_add(v0, m31_mul(m31_add(m31_mul(c2, x), c1), x));
}

}  // namespace

int host_verify(const HostCircuitView &c, uint32_t L, uint32_t R, int kind, uint32_t grinding_bits,
                const uint32_t *claimed, const uint8_t *proof, size_t proof_len)
{
    const uint32_t D = c.D;
    const size_t S = (size_t)L * 4;
    const size_t N0 = (size_t)1 << c.n[0];
    size_t need = N0 * S + 32;
    for (uint32_t l = 0; l < D; l++) need += ((size_t)6 * c.n[l] * R + 2 * R) * S;
    if (proof_len < need) return 0;
    auto read_f = [&](size_t at, uint32_t *out) {
        for (uint32_t k = 0; k < L; k++) {
            uint32_t x;
            memcpy(&x, proof + at + 4 * k, 4);
            out[k] = m31_from_u32_full(x);
        }
    };
    std::vector<uint32_t> commit(N0 * L);
    for (size_t i = 0; i < N0; i++) read_f(i * S, &commit[i * L]);
    HostTranscript t{kind};
    t.bytes.reserve(need);
    t.append(proof, N0 * S);
    // grind (LinearGKR.hpp:13-32); the verifier appends its own grind bytes and skips the
    // prover's (LinearGKR.hpp:121)
    {
        uint8_t hb[32] = {0};
        for (int i = 0; i < 8; i++) {
            uint32_t h = t.challenge();
            memcpy(hb + 4 * i, &h, 4);
        }
        for (uint32_t i = 0; i < (1u << grinding_bits); i++) host_hash(kind, hb, 32, hb);
        t.append(hb, 32);
    }
    size_t idx = N0 * S + 32;
    const uint32_t mlast = c.m[D - 1];
    std::vector<uint32_t> rz1((size_t)R * 32, 0), rz2((size_t)R * 32, 0), rx((size_t)R * 32, 0),
        ry((size_t)R * 32, 0);
    for (uint32_t i = 0; i < mlast; i++)
        for (uint32_t j = 0; j < R; j++) rz1[j * 32 + i] = t.challenge();
    uint32_t alpha = 1, beta = 0;
    std::vector<uint32_t> cv1(claimed, claimed + (size_t)R * L), cv2((size_t)R * L, 0);
    std::vector<uint32_t> sum((size_t)R * L), ev(3 * L), vx((size_t)R * L), vy((size_t)R * L);
    bool ok = true;
    std::vector<uint32_t> e1, e2, ex, ey;
    for (int l = (int)D - 1; l >= 0; l--) {
        const uint32_t n = c.n[l], m = c.m[l];
        for (uint32_t j = 0; j < R; j++)
            for (uint32_t k = 0; k < L; k++)
                sum[j * L + k] = m31_add(m31_mul(cv1[j * L + k], alpha), m31_mul(cv2[j * L + k], beta));
        for (uint32_t i_var = 0; i_var < 2 * n; i_var++) {
            for (uint32_t j = 0; j < R; j++) {
                read_f(idx, &ev[0]);
                read_f(idx + S, &ev[L]);
                read_f(idx + 2 * S, &ev[2 * L]);
                idx += 3 * S;
                t.append(ev.data(), 3 * S);
                const uint32_t r = t.challenge();
                if (i_var < n) rx[j * 32 + i_var] = r;
                else ry[j * 32 + i_var - n] = r;
                for (uint32_t k = 0; k < L; k++) {
                    ok &= m31_add(ev[k], ev[L + k]) == sum[j * L + k];
                    sum[j * L + k] = degree2(ev[k], ev[L + k], ev[2 * L + k], r);
                }
                if (i_var == n - 1) {
                    read_f(idx, &vx[j * L]);
                    idx += S;
                    // alpha add(rz1, rx) + beta add(rz2, rx)
                    eq_table(&rz1[j * 32], m, alpha, e1);
                    eq_table(&rz2[j * 32], m, beta, e2);
                    eq_table(&rx[j * 32], n, 1u, ex);
                    uint32_t a = 0;
                    for (size_t g = 0; g < c.add[l].size() / 3; g++) {
                        const uint32_t *q = &c.add[l][3 * g];
                        a = m31_add(a, m31_mul(m31_mul(m31_add(e1[q[1]], e2[q[1]]), ex[q[0]]), q[2]));
                    }
                    for (uint32_t k = 0; k < L; k++) sum[j * L + k] = m31_sub(sum[j * L + k], m31_mul(vx[j * L + k], a));
                    t.append(&vx[j * L], S);
                }
            }
        }
        for (uint32_t j = 0; j < R; j++) {
            read_f(idx, &vy[j * L]);
            idx += S;
            eq_table(&rz1[j * 32], m, alpha, e1);
            eq_table(&rz2[j * 32], m, beta, e2);
            eq_table(&rx[j * 32], n, 1u, ex);
            eq_table(&ry[j * 32], n, 1u, ey);
            uint32_t mv = 0;
            for (size_t g = 0; g < c.mul[l].size() / 4; g++) {
                const uint32_t *q = &c.mul[l][4 * g];
                uint32_t prod = m31_mul(m31_mul(m31_add(e1[q[2]], e2[q[2]]), ex[q[0]]), ey[q[1]]);
                mv = m31_add(mv, m31_mul(prod, q[3]));
            }
            for (uint32_t k = 0; k < L; k++)
                ok &= sum[j * L + k] == m31_mul(m31_mul(vx[j * L + k], vy[j * L + k]), mv);
            t.append(&vy[j * L], S);
        }
        alpha = t.challenge();
        beta = t.challenge();
        rz1 = rx;
        rz2 = ry;
        cv1 = vx;
        cv2 = vy;
    }
    std::vector<uint32_t> e(L);
    for (uint32_t j = 0; j < R; j++) {
        eval_ml(commit.data(), L, &rz1[j * 32], c.n[0], e.data());
        ok &= memcmp(e.data(), &cv1[j * L], S) == 0;
        eval_ml(commit.data(), L, &rz2[j * 32], c.n[0], e.data());
        ok &= memcmp(e.data(), &cv2[j * L], S) == 0;
    }
    return ok ? 1 : 0;
}

}  // namespace gkrx
