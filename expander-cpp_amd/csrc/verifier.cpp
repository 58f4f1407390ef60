// Host verifier: Verifier::verify (include/LinearGKR/LinearGKR.hpp:104-149) with
// gkr_verify (include/LinearGKR/gkr.hpp:75-122), sumcheck_verify_gkr_layer
// (include/LinearGKR/sumcheck.hpp:113-176), degree_2_eval and
// eval_sparse_circuit_connect_poly (include/LinearGKR/sumcheck_verifier_utils.hpp:10-56)
// and RawPC::verify (include/poly_commit/raw.hpp:75-78).  The verifier is the
// acceptance check of the library; it is O(gates + 2^vars) per layer and stays on the
// host, as in the reference.  One template serves both fields: M31 with L lanes and the
// BN254 scalar field with one lane.
#include <cstring>
#include <vector>

#include "bn254.hpp"
#include "host_hash.hpp"
#include "m31.hpp"
#include "verifier.hpp"

namespace gkrx {

namespace {

// M31 (include/field/M31.hpp): canonical u32; from_bytes reduces fully (:111-116).
struct M31Ops {
    using E = uint32_t;
    static constexpr uint32_t kWords = 1;      // u32 words per serialized lane element
    static constexpr uint32_t kFieldBits = 31;  // config.hpp:54
    static E zero() { return 0; }
    static E one() { return 1; }
    static E add(E a, E b) { return m31_add(a, b); }
    static E sub(E a, E b) { return m31_sub(a, b); }
    static E mul(E a, E b) { return m31_mul(a, b); }
    static E inv2() { return 1u << 30; }
    static E coef(uint32_t c) { return c; }  // the library keeps canonical coefficients
    static E from_words(const uint32_t *w) { return m31_from_u32_full(w[0]); }
    static void to_words(E e, uint32_t *w) { w[0] = e; }
    static E challenge(const uint8_t d[32])
    {
        uint32_t x;
        memcpy(&x, d, 4);
        return m31_from_u32_full(x);
    }
    static bool eq(E a, E b) { return a == b; }
};

// BN254 Fr in Montgomery form; serialized as the canonical value in 32 LE bytes; a
// challenge is the digest read as a 256-bit LE integer reduced mod r.
struct BnOps {
    using E = Fr;
    static constexpr uint32_t kWords = 8;
    static constexpr uint32_t kFieldBits = 254;
    static E zero() { return fr_zero(); }
    static E one() { return Fr{GKRX_BN_ONE}; }
    static E add(const E &a, const E &b) { return fr_add(a, b); }
    static E sub(const E &a, const E &b) { return fr_sub(a, b); }
    static E mul(const E &a, const E &b) { return fr_mont_mul(a, b); }
    static E inv2()
    {
        // (r + 1) / 2
        const Fr h = {{0xf8000001u, 0xa1f0fac9u, 0x3cdcb848u, 0x9419f424u, 0x40c0ac2eu, 0xdc2822dbu, 0x7098d014u,
                       0x18322739u}};
        return fr_to_mont(h);
    }
    static E coef(uint32_t c) { return fr_mont_of_u32(c); }
    static E from_words(const uint32_t *w) { return fr_to_mont(fr_reduce256(w)); }
    static void to_words(const E &e, uint32_t *w)
    {
        const Fr c = fr_from_mont(e);
        memcpy(w, c.w, 32);
    }
    static E challenge(const uint8_t d[32])
    {
        uint32_t w[8];
        memcpy(w, d, 32);
        return fr_to_mont(fr_reduce256(w));
    }
    static bool eq(const E &a, const E &b) { return memcmp(a.w, b.w, 32) == 0; }
};

template <class F>
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
    void append_f(const typename F::E &e)
    {
        uint32_t w[F::kWords];
        F::to_words(e, w);
        append(w, sizeof w);
    }
    typename F::E challenge()
    {
        if (bytes.size() > hash_start) {
            host_hash(kind, bytes.data() + hash_start, bytes.size() - hash_start, digest);
            hash_start = bytes.size();
        } else {
            host_hash(kind, digest, 32, digest);
        }
        return F::challenge(digest);
    }
};

template <class F>
void eq_table(const typename F::E *r, uint32_t n, const typename F::E &factor, std::vector<typename F::E> &out)
{
    out.assign((size_t)1 << n, F::zero());
    out[0] = factor;
    size_t cur = 1;
    for (uint32_t i = 0; i < n; i++) {
        const typename F::E e1 = r[i], e0 = F::sub(F::one(), r[i]);
        for (size_t j = 0; j < cur; j++) {
            out[j + cur] = F::mul(out[j], e1);
            out[j] = F::mul(out[j], e0);
        }
        cur <<= 1;
    }
}

// eval_multilinear (include/poly_commit/poly.hpp:88-104) = sum_z eq(r, z) vals[z]
template <class F>
void eval_ml(const typename F::E *vals, uint32_t L, const typename F::E *r, uint32_t n, typename F::E *out)
{
    std::vector<typename F::E> eq;
    eq_table<F>(r, n, F::one(), eq);
    for (uint32_t k = 0; k < L; k++) out[k] = F::zero();
    for (size_t z = 0; z < eq.size(); z++)
        for (uint32_t k = 0; k < L; k++) out[k] = F::add(out[k], F::mul(eq[z], vals[z * L + k]));
}

template <class F>
typename F::E degree2(const typename F::E &v0, const typename F::E &v1, const typename F::E &v2,
                      const typename F::E &x)
{
    const typename F::E c2 = F::mul(F::inv2(), F::add(F::sub(v2, F::add(v1, v1)), v0));
    const typename F::E c1 = F::sub(F::sub(v1, v0), c2);
    return F::add(v0, F::mul(F::add(F::mul(c2, x), c1), x));
}

template <class F>
int verify_impl(const HostCircuitView &c, uint32_t L, uint32_t R, int kind, uint32_t grinding_bits,
                const uint32_t *claimed, const uint8_t *proof, size_t proof_len)
{
    using E = typename F::E;
    const uint32_t D = c.D;
    const size_t S = (size_t)L * F::kWords * 4;
    const size_t N0 = (size_t)1 << c.n[0];
    size_t need = N0 * S + 32;
    for (uint32_t l = 0; l < D; l++) need += ((size_t)6 * c.n[l] * R + 2 * R) * S;
    if (proof_len < need) return 0;
    auto read_f = [&](size_t at, E *out) {
        for (uint32_t k = 0; k < L; k++) {
            uint32_t w[F::kWords];
            memcpy(w, proof + at + 4 * F::kWords * k, sizeof w);
            out[k] = F::from_words(w);
        }
    };
    std::vector<E> commit(N0 * L);
    for (size_t i = 0; i < N0; i++) read_f(i * S, &commit[i * L]);
    HostTranscript<F> t{kind};
    t.bytes.reserve(need);
    t.append(proof, N0 * S);
    // grind (LinearGKR.hpp:13-32): 256 / field_bits challenges, each written at a stride of
    // ceil(field_bits / 8) bytes; the verifier appends its own grind bytes and skips the
    // prover's (LinearGKR.hpp:121)
    {
        uint8_t hb[32] = {0};
        const uint32_t step = (F::kFieldBits + 7) / 8;
        for (uint32_t i = 0; i < 256 / F::kFieldBits; i++) {
            const E h = t.challenge();
            uint32_t w[F::kWords];
            F::to_words(h, w);
            memcpy(hb + step * i, w, step);
        }
        for (uint32_t i = 0; i < (1u << grinding_bits); i++) host_hash(kind, hb, 32, hb);
        t.append(hb, 32);
    }
    size_t idx = N0 * S + 32;
    const uint32_t mlast = c.m[D - 1];
    std::vector<E> rz1((size_t)R * 32, F::zero()), rz2((size_t)R * 32, F::zero()), rx((size_t)R * 32, F::zero()),
        ry((size_t)R * 32, F::zero());
    for (uint32_t i = 0; i < mlast; i++)
        for (uint32_t j = 0; j < R; j++) rz1[j * 32 + i] = t.challenge();
    E alpha = F::one(), beta = F::zero();
    std::vector<E> cv1((size_t)R * L), cv2((size_t)R * L, F::zero());
    for (size_t i = 0; i < (size_t)R * L; i++) cv1[i] = F::from_words(claimed + i * F::kWords);
    std::vector<E> sum((size_t)R * L), ev(3 * L), vx((size_t)R * L), vy((size_t)R * L);
    bool ok = true;
    std::vector<E> e1, e2, ex, ey;
    for (int l = (int)D - 1; l >= 0; l--) {
        const uint32_t n = c.n[l], m = c.m[l];
        for (uint32_t j = 0; j < R; j++)
            for (uint32_t k = 0; k < L; k++)
                sum[j * L + k] = F::add(F::mul(cv1[j * L + k], alpha), F::mul(cv2[j * L + k], beta));
        for (uint32_t i_var = 0; i_var < 2 * n; i_var++) {
            for (uint32_t j = 0; j < R; j++) {
                // the reference re-serializes what it read (append_f of the parsed value)
                read_f(idx, &ev[0]);
                read_f(idx + S, &ev[L]);
                read_f(idx + 2 * S, &ev[2 * L]);
                for (uint32_t k = 0; k < 3 * L; k++) t.append_f(ev[k]);
                idx += 3 * S;
                const E r = t.challenge();
                if (i_var < n) rx[j * 32 + i_var] = r;
                else ry[j * 32 + i_var - n] = r;
                for (uint32_t k = 0; k < L; k++) {
                    ok &= F::eq(F::add(ev[k], ev[L + k]), sum[j * L + k]);
                    sum[j * L + k] = degree2<F>(ev[k], ev[L + k], ev[2 * L + k], r);
                }
                if (i_var == n - 1) {
                    read_f(idx, &vx[j * L]);
                    for (uint32_t k = 0; k < L; k++) t.append_f(vx[j * L + k]);
                    idx += S;
                    // alpha add(rz1, rx) + beta add(rz2, rx)
                    eq_table<F>(&rz1[j * 32], m, alpha, e1);
                    eq_table<F>(&rz2[j * 32], m, beta, e2);
                    eq_table<F>(&rx[j * 32], n, F::one(), ex);
                    E a = F::zero();
                    for (size_t g = 0; g < c.add[l].size() / 3; g++) {
                        const uint32_t *q = &c.add[l][3 * g];
                        a = F::add(a, F::mul(F::mul(F::add(e1[q[1]], e2[q[1]]), ex[q[0]]), F::coef(q[2])));
                    }
                    for (uint32_t k = 0; k < L; k++)
                        sum[j * L + k] = F::sub(sum[j * L + k], F::mul(vx[j * L + k], a));
                }
            }
        }
        for (uint32_t j = 0; j < R; j++) {
            read_f(idx, &vy[j * L]);
            for (uint32_t k = 0; k < L; k++) t.append_f(vy[j * L + k]);
            idx += S;
            eq_table<F>(&rz1[j * 32], m, alpha, e1);
            eq_table<F>(&rz2[j * 32], m, beta, e2);
            eq_table<F>(&rx[j * 32], n, F::one(), ex);
            eq_table<F>(&ry[j * 32], n, F::one(), ey);
            E mv = F::zero();
            for (size_t g = 0; g < c.mul[l].size() / 4; g++) {
                const uint32_t *q = &c.mul[l][4 * g];
                const E prod = F::mul(F::mul(F::add(e1[q[2]], e2[q[2]]), ex[q[0]]), ey[q[1]]);
                mv = F::add(mv, F::mul(prod, F::coef(q[3])));
            }
            for (uint32_t k = 0; k < L; k++)
                ok &= F::eq(sum[j * L + k], F::mul(F::mul(vx[j * L + k], vy[j * L + k]), mv));
        }
        alpha = t.challenge();
        beta = t.challenge();
        rz1 = rx;
        rz2 = ry;
        cv1 = vx;
        cv2 = vy;
    }
    std::vector<E> e(L);
    for (uint32_t j = 0; j < R; j++) {
        eval_ml<F>(commit.data(), L, &rz1[j * 32], c.n[0], e.data());
        for (uint32_t k = 0; k < L; k++) ok &= F::eq(e[k], cv1[j * L + k]);
        eval_ml<F>(commit.data(), L, &rz2[j * 32], c.n[0], e.data());
        for (uint32_t k = 0; k < L; k++) ok &= F::eq(e[k], cv2[j * L + k]);
    }
    return ok ? 1 : 0;
}

}  // namespace

int host_verify(const HostCircuitView &c, uint32_t L, uint32_t R, int kind, uint32_t grinding_bits,
                const uint32_t *claimed, const uint8_t *proof, size_t proof_len)
{
    return verify_impl<M31Ops>(c, L, R, kind, grinding_bits, claimed, proof, proof_len);
}

int host_verify_bn254(const HostCircuitView &c, int kind, uint32_t grinding_bits, const uint32_t *claimed,
                      const uint8_t *proof, size_t proof_len)
{
    return verify_impl<BnOps>(c, 1, 1, kind, grinding_bits, claimed, proof, proof_len);
}

}  // namespace gkrx
