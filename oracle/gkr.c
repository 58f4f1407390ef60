/*
 * Oracle GKR prover / verifier -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * A scalar C restatement of the reference's LinearGKR path for M31 with L data-parallel
 * lanes.  Lane arithmetic is the scalar M31 arithmetic: for canonical inputs the
 * reference's packed AVX lane ops (include/field/M31_avx.tcc:97-176) return the same
 * canonical values as the scalar ops (include/field/M31.hpp:64-116).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define P ORACLE_P31

/* ------------------------------------------------------------------ field */
/* M31::operator+ (include/field/M31.hpp:64-71) */
uint32_t oracle_m31_add(uint32_t a, uint32_t b)
{
    uint32_t s = a + b;
    return s >= P ? s - P : s;
}
/* M31::operator- unary (M31.hpp:84-88) */
uint32_t oracle_m31_neg(uint32_t a) { return a == 0 ? 0 : P - a; }
/* M31::operator- binary = a + (-b) (M31.hpp:90-93) */
uint32_t oracle_m31_sub(uint32_t a, uint32_t b) { return oracle_m31_add(a, oracle_m31_neg(b)); }
/* M31::operator* (M31.hpp:73-82): 64-bit product, one Mersenne fold, one conditional subtract */
uint32_t oracle_m31_mul(uint32_t a, uint32_t b)
{
    uint64_t x = (uint64_t)a * b;
    x = (x & P) + (x >> 31);
    if (x >= P) x -= P;
    return (uint32_t)x;
}
/* M31(uint32) (M31.hpp:58-62): one fold, no final subtract, so M31(p) stays p */
uint32_t oracle_m31_from_u32(uint32_t v) { return (v & P) + (v >> 31); }
/* M31::from_bytes (M31.hpp:111-116) */
uint32_t oracle_m31_from_bytes(const uint8_t *in)
{
    uint32_t x = (uint32_t)in[0] | ((uint32_t)in[1] << 8) | ((uint32_t)in[2] << 16) | ((uint32_t)in[3] << 24);
    x = (x & P) + (x >> 31);
    if (x >= P) x -= P;
    return x;
}

#define fadd oracle_m31_add
#define fsub oracle_m31_sub
#define fmul oracle_m31_mul

void oracle_free(void *p) { free(p); }

/* config.hpp:61-65 with security 100, grinding 10, field size 31 */
uint32_t oracle_num_repetitions(void) { return (100 - 10 + 30) / 31; }

/* ---------------------------------------------------------------- circuits */
void oracle_circuit_free(oracle_circuit *c)
{
    if (!c) return;
    for (uint32_t i = 0; i < c->nb_layers; i++) {
        free(c->mul[i]);
        free(c->add[i]);
    }
    free(c->mul); free(c->add); free(c->n_mul); free(c->n_add);
    free(c->nb_in_vars); free(c->nb_out_vars);
    free(c);
}

/* next_pow_of_2 (include/utils/myutil.hpp:33-48) */
static uint32_t next_pow_of_2(uint32_t n)
{
    if (n == 0 || (n & (n - 1)) == 0) return n;
    return 1u << (32 - __builtin_clz(n));
}

/* Circuit::_compute_nb_vars (include/circuit/circuit.hpp:124-149) */
static void compute_nb_vars(oracle_circuit *c)
{
    for (uint32_t l = 0; l < c->nb_layers; l++) {
        uint32_t mo = 0, mi = 0;
        for (uint32_t g = 0; g < c->n_mul[l]; g++) {
            const uint32_t *q = c->mul[l] + 4 * g;
            if (q[2] > mo) mo = q[2];
            if (q[0] > mi) mi = q[0];
            if (q[1] > mi) mi = q[1];
        }
        for (uint32_t g = 0; g < c->n_add[l]; g++) {
            const uint32_t *q = c->add[l] + 3 * g;
            if (q[1] > mo) mo = q[1];
            if (q[0] > mi) mi = q[0];
        }
        c->nb_in_vars[l] = mi > 0 ? (uint32_t)__builtin_ctz(next_pow_of_2(mi)) : 0;
        c->nb_out_vars[l] = mo > 0 ? (uint32_t)__builtin_ctz(next_pow_of_2(mo)) : 0;
    }
}

/* Circuit::load_extracted_gates (include/circuit/circuit.hpp:152-215): the mul file has
 * one line per layer, output layer first; each line is num_gates followed by
 * (in0 in1 out coef) tuples; the add file holds (in0 out coef) tuples per layer. */
int oracle_circuit_load_text(const char *mul_path, const char *add_path, oracle_circuit **out)
{
    FILE *f = fopen(mul_path, "r");
    if (!f) return -1;
    uint32_t nb = 0;
    int ch, last = '\n';
    while ((ch = fgetc(f)) != EOF) {
        if (ch == '\n') nb++;
        last = ch;
    }
    if (last != '\n') nb++; /* std::getline counts an unterminated last line */
    rewind(f);
    oracle_circuit *c = (oracle_circuit *)calloc(1, sizeof *c);
    c->nb_layers = nb;
    c->n_mul = (uint32_t *)calloc(nb, 4); c->n_add = (uint32_t *)calloc(nb, 4);
    c->mul = (uint32_t **)calloc(nb, sizeof(uint32_t *)); c->add = (uint32_t **)calloc(nb, sizeof(uint32_t *));
    c->nb_in_vars = (uint32_t *)calloc(nb, 4); c->nb_out_vars = (uint32_t *)calloc(nb, 4);
    for (uint32_t i = 0; i < nb; i++) {
        uint32_t l = nb - 1 - i;
        unsigned long ng = 0;
        if (fscanf(f, "%lu", &ng) != 1) { fclose(f); oracle_circuit_free(c); return -2; }
        c->n_mul[l] = (uint32_t)ng;
        c->mul[l] = (uint32_t *)malloc(4 * 4 * (ng ? ng : 1));
        for (unsigned long g = 0; g < ng; g++) {
            unsigned long i0, i1, o;
            unsigned coef;
            if (fscanf(f, "%lu%lu%lu%u", &i0, &i1, &o, &coef) != 4) { fclose(f); oracle_circuit_free(c); return -2; }
            uint32_t *q = c->mul[l] + 4 * g;
            q[0] = (uint32_t)i0; q[1] = (uint32_t)i1; q[2] = (uint32_t)o; q[3] = oracle_m31_from_u32(coef);
        }
    }
    fclose(f);
    f = fopen(add_path, "r");
    if (!f) { oracle_circuit_free(c); return -1; }
    for (uint32_t i = 0; i < nb; i++) {
        uint32_t l = nb - 1 - i;
        unsigned long ng = 0;
        if (fscanf(f, "%lu", &ng) != 1) { fclose(f); oracle_circuit_free(c); return -2; }
        c->n_add[l] = (uint32_t)ng;
        c->add[l] = (uint32_t *)malloc(4 * 3 * (ng ? ng : 1));
        for (unsigned long g = 0; g < ng; g++) {
            unsigned long i0, o;
            unsigned coef;
            if (fscanf(f, "%lu%lu%u", &i0, &o, &coef) != 3) { fclose(f); oracle_circuit_free(c); return -2; }
            uint32_t *q = c->add[l] + 3 * g;
            q[0] = (uint32_t)i0; q[1] = (uint32_t)o; q[2] = oracle_m31_from_u32(coef);
        }
    }
    fclose(f);
    compute_nb_vars(c);
    *out = c;
    return 0;
}

/* Shapes this restatement supports: every layer has >= 1 input var, consecutive layers
 * chain (n_in(l+1) == n_out(l)), every gate index is in range and every coefficient is
 * below 2^31 (where the reference's signed AVX lane multiply is exact). */
static int check_shape(const oracle_circuit *c)
{
    if (c->nb_layers == 0) return -3;
    for (uint32_t l = 0; l < c->nb_layers; l++) {
        uint32_t n = c->nb_in_vars[l], m = c->nb_out_vars[l];
        if (n < 1 || n > 30 || m > 30) return -3;
        if (l + 1 < c->nb_layers && c->nb_in_vars[l + 1] != m) return -3;
        for (uint32_t g = 0; g < c->n_mul[l]; g++) {
            const uint32_t *q = c->mul[l] + 4 * g;
            if (q[0] >> n || q[1] >> n || q[2] >> m || q[3] >= 0x80000000u) return -3;
        }
        for (uint32_t g = 0; g < c->n_add[l]; g++) {
            const uint32_t *q = c->add[l] + 3 * g;
            if (q[0] >> n || q[1] >> m || q[2] >= 0x80000000u) return -3;
        }
    }
    return 0;
}

/* CircuitLayer::evaluate (circuit.hpp:92-105): mul gates first, then add gates */
static void layer_evaluate(const oracle_circuit *c, uint32_t l, uint32_t L, const uint32_t *in, uint32_t *out)
{
    size_t M = (size_t)1 << c->nb_out_vars[l];
    memset(out, 0, M * L * 4);
    for (uint32_t g = 0; g < c->n_mul[l]; g++) {
        const uint32_t *q = c->mul[l] + 4 * g;
        for (uint32_t k = 0; k < L; k++) {
            uint32_t t = fmul(fmul(in[(size_t)q[0] * L + k], in[(size_t)q[1] * L + k]), q[3]);
            out[(size_t)q[2] * L + k] = fadd(out[(size_t)q[2] * L + k], t);
        }
    }
    for (uint32_t g = 0; g < c->n_add[l]; g++) {
        const uint32_t *q = c->add[l] + 3 * g;
        for (uint32_t k = 0; k < L; k++) {
            uint32_t t = fmul(in[(size_t)q[0] * L + k], q[2]);
            out[(size_t)q[1] * L + k] = fadd(out[(size_t)q[1] * L + k], t);
        }
    }
}

static size_t values_len(const oracle_circuit *c, uint32_t L)
{
    size_t tot = 0;
    for (uint32_t l = 0; l < c->nb_layers; l++) tot += ((size_t)1 << c->nb_in_vars[l]) * L;
    tot += ((size_t)1 << c->nb_out_vars[c->nb_layers - 1]) * L;
    return tot;
}

/* Circuit::evaluate (circuit.hpp:273-280) */
size_t oracle_evaluate(const oracle_circuit *c, uint32_t L, const uint32_t *witness, uint32_t *values)
{
    size_t tot = values_len(c, L);
    if (!values) return tot;
    size_t off = 0;
    size_t N0 = ((size_t)1 << c->nb_in_vars[0]) * L;
    memcpy(values, witness, N0 * 4);
    for (uint32_t l = 0; l < c->nb_layers; l++) {
        size_t Nl = ((size_t)1 << c->nb_in_vars[l]) * L;
        layer_evaluate(c, l, L, values + off, values + off + Nl);
        off += Nl;
    }
    return tot;
}

/* _eq_evals_at_primitive (include/LinearGKR/sumcheck_common.hpp:16-33), with
 * _eq(x,y) = 2xy - x - y + 1 (:9-14) evaluated at y = 0 and y = 1. */
void oracle_eq_evals(const uint32_t *r, uint32_t n, uint32_t mul_factor, uint32_t *out)
{
    out[0] = mul_factor;
    size_t cur = 1;
    for (uint32_t i = 0; i < n; i++) {
        /* _eq(r, 0) = 0*2 - r - 0 + 1 ; _eq(r, 1) = 2r - r - 1 + 1 */
        uint32_t e0 = fadd(fsub(fsub(fmul(fmul(r[i], 0), 2), r[i]), 0), 1);
        uint32_t e1 = fadd(fsub(fsub(fmul(fmul(r[i], 1), 2), r[i]), 1), 1);
        for (size_t j = 0; j < cur; j++) {
            out[j + cur] = fmul(out[j], e1);
            out[j] = fmul(out[j], e0);
        }
        cur <<= 1;
    }
}

/* eval_multilinear (include/poly_commit/poly.hpp:88-104): LSB-first folding */
void oracle_eval_multilinear(const uint32_t *evals, uint32_t L, const uint32_t *r, uint32_t nvars, uint32_t *out)
{
    size_t N = (size_t)1 << nvars;
    uint32_t *s = (uint32_t *)malloc(N * L * 4);
    memcpy(s, evals, N * L * 4);
    size_t cur = N >> 1;
    for (uint32_t v = 0; v < nvars; v++) {
        for (size_t i = 0; i < cur; i++)
            for (uint32_t k = 0; k < L; k++) {
                uint32_t a = s[(2 * i) * L + k], b = s[(2 * i + 1) * L + k];
                s[i * L + k] = fadd(a, fmul(fsub(b, a), r[v]));
            }
        cur >>= 1;
    }
    memcpy(out, s, L * 4);
    free(s);
}

/* -------------------------------------------------------------- transcript */
/* Transcript (include/fiat_shamir/transcript.hpp:87-175): the proof byte buffer, the
 * start of the not-yet-hashed suffix, and a 32-byte digest that starts at zero. */
typedef struct {
    uint8_t *bytes;
    size_t len, cap, hash_start;
    uint8_t digest[32];
    int hash_kind;
} transcript;

static void hash32(int kind, const uint8_t *in, size_t len, uint8_t out[32])
{
    uint8_t tmp[32];
    if (kind == ORACLE_HASH_KECCAK256) oracle_keccak256(in, len, tmp);
    else oracle_sha256(in, len, tmp);
    memcpy(out, tmp, 32);
}

static void tr_init(transcript *t, int kind, size_t reserve)
{
    memset(t, 0, sizeof *t);
    t->hash_kind = kind;
    t->cap = reserve ? reserve : 1024;
    t->bytes = (uint8_t *)malloc(t->cap);
}

/* append_bytes (transcript.hpp:127-130) */
static void tr_append(transcript *t, const void *p, size_t n)
{
    if (t->len + n > t->cap) {
        while (t->len + n > t->cap) t->cap *= 2;
        t->bytes = (uint8_t *)realloc(t->bytes, t->cap);
    }
    memcpy(t->bytes + t->len, p, n);
    t->len += n;
}

/* _hash_to_digest (transcript.hpp:91-103): hash only the bytes appended since the last
 * challenge; with none, re-hash the digest.  The digest is not chained otherwise. */
static void tr_hash_to_digest(transcript *t)
{
    if (t->len > t->hash_start) {
        hash32(t->hash_kind, t->bytes + t->hash_start, t->len - t->hash_start, t->digest);
        t->hash_start = t->len;
    } else {
        hash32(t->hash_kind, t->digest, 32, t->digest);
    }
}

/* challenge_f (transcript.hpp:139-147): M31::from_bytes of the digest's first 4 bytes */
static uint32_t tr_challenge(transcript *t)
{
    tr_hash_to_digest(t);
    return oracle_m31_from_bytes(t->digest);
}

void *oracle_transcript_new(int hash_kind)
{
    transcript *t = (transcript *)malloc(sizeof(transcript));
    tr_init(t, hash_kind, 1024);
    return t;
}
void oracle_transcript_append(void *t, const uint8_t *bytes, size_t len) { tr_append((transcript *)t, bytes, len); }
uint32_t oracle_transcript_challenge(void *t) { return tr_challenge((transcript *)t); }
void oracle_transcript_free(void *t)
{
    if (!t) return;
    free(((transcript *)t)->bytes);
    free(t);
}

/* grind (include/LinearGKR/LinearGKR.hpp:13-32): 256/31 = 8 challenges serialized at
 * 4-byte strides (ceil(31, 8) = 4) into 32 bytes, hashed 2^10 times in place, appended. */
static void grind(transcript *t)
{
    uint8_t hb[32];
    memset(hb, 0, 32);
    for (int i = 0; i < 256 / 31; i++) {
        uint32_t h = tr_challenge(t);
        memcpy(hb + 4 * i, &h, 4);
    }
    for (int i = 0; i < (1 << 10); i++) hash32(t->hash_kind, hb, 32, hb);
    tr_append(t, hb, 32);
}

/* --------------------------------------------------------------- prover */
typedef struct {
    /* one soundness repetition: GKRScratchPad (include/LinearGKR/scratch_pad.hpp:8-65) */
    uint32_t *f;      /* bookkeeping_f == v_evals, N * L */
    uint32_t *hg;     /* N * L */
    uint8_t *gate_exists;
    uint32_t *eqz;    /* combined alpha*eq(rz1) + beta*eq(rz2), M */
    uint32_t *eq2;    /* scratch M */
    uint32_t *eqx;    /* eq(rx), N */
    uint32_t *rx, *ry;
    uint32_t nrx, nry;
} rep_state;

/* _prepare_g_x_vals (include/LinearGKR/sumcheck_helper.hpp:172-231) */
static void prepare_g_x(const oracle_circuit *c, uint32_t l, uint32_t L, const uint32_t *v, const uint32_t *rz1,
                        const uint32_t *rz2, uint32_t alpha, uint32_t beta, rep_state *s)
{
    uint32_t n = c->nb_in_vars[l], m = c->nb_out_vars[l];
    size_t N = (size_t)1 << n, M = (size_t)1 << m;
    memset(s->hg, 0, N * L * 4);
    memset(s->gate_exists, 0, N);
    oracle_eq_evals(rz1, m, alpha, s->eqz);
    oracle_eq_evals(rz2, m, beta, s->eq2);
    for (size_t i = 0; i < M; i++) s->eqz[i] = fadd(s->eqz[i], s->eq2[i]);
    for (uint32_t g = 0; g < c->n_mul[l]; g++) {
        const uint32_t *q = c->mul[l] + 4 * g;
        uint32_t x = q[0], y = q[1], z = q[2];
        uint32_t ce = fmul(q[3], s->eqz[z]);
        for (uint32_t k = 0; k < L; k++)
            s->hg[(size_t)x * L + k] = fadd(s->hg[(size_t)x * L + k], fmul(v[(size_t)y * L + k], ce));
        s->gate_exists[x] = 1;
    }
    for (uint32_t g = 0; g < c->n_add[l]; g++) {
        const uint32_t *q = c->add[l] + 3 * g;
        uint32_t x = q[0], z = q[1];
        uint32_t ce = fmul(q[2], s->eqz[z]);
        for (uint32_t k = 0; k < L; k++) s->hg[(size_t)x * L + k] = fadd(s->hg[(size_t)x * L + k], ce);
        s->gate_exists[x] = 1;
    }
}

/* _prepare_h_y_vals (sumcheck_helper.hpp:233-261) with vx = v_evals[0] (:266-269) */
static void prepare_h_y(const oracle_circuit *c, uint32_t l, uint32_t L, rep_state *s)
{
    uint32_t n = c->nb_in_vars[l];
    size_t N = (size_t)1 << n;
    uint32_t *vx = (uint32_t *)malloc(L * 4);
    memcpy(vx, s->f, L * 4);
    memset(s->hg, 0, N * L * 4);
    memset(s->gate_exists, 0, N);
    oracle_eq_evals(s->rx, n, 1, s->eqx);
    for (uint32_t g = 0; g < c->n_mul[l]; g++) {
        const uint32_t *q = c->mul[l] + 4 * g;
        uint32_t x = q[0], y = q[1], z = q[2];
        uint32_t sc = fmul(fmul(s->eqz[z], s->eqx[x]), q[3]);
        for (uint32_t k = 0; k < L; k++)
            s->hg[(size_t)y * L + k] = fadd(s->hg[(size_t)y * L + k], fmul(vx[k], sc));
        s->gate_exists[y] = 1;
    }
    free(vx);
}

/* SumcheckMultiLinearProdHelper::poly_eval_at (sumcheck_helper.hpp:68-102) */
static void poly_eval_at(uint32_t L, uint32_t nb_vars, uint32_t var_idx, const uint32_t *src, const uint32_t *hg,
                         const uint8_t *ge, uint32_t *p /* 3*L */)
{
    size_t evalSize = (size_t)1 << (nb_vars - var_idx - 1);
    uint32_t *p0 = p, *p1 = p + L, *p2 = p + 2 * L;
    memset(p, 0, 3 * L * 4);
    for (size_t i = 0; i < evalSize; i++) {
        if (!ge[2 * i] && !ge[2 * i + 1]) continue;
        for (uint32_t k = 0; k < L; k++) {
            uint32_t f0 = src[(2 * i) * L + k], f1 = src[(2 * i + 1) * L + k];
            uint32_t h0 = hg[(2 * i) * L + k], h1 = hg[(2 * i + 1) * L + k];
            p0[k] = fadd(p0[k], fmul(f0, h0));
            p1[k] = fadd(p1[k], fmul(f1, h1));
            p2[k] = fadd(p2[k], fmul(fadd(f0, f1), fadd(h0, h1)));
        }
    }
    /* p2 = p1 * 6 + p0 * 3 - p2 * 2 (:100) */
    for (uint32_t k = 0; k < L; k++) p2[k] = fsub(fadd(fmul(p1[k], 6), fmul(p0[k], 3)), fmul(p2[k], 2));
}

/* SumcheckMultiLinearProdHelper::receive_challenge (sumcheck_helper.hpp:104-133) */
static void receive_challenge(uint32_t L, size_t cur_size, const uint32_t *src, uint32_t *f, uint32_t *hg, uint8_t *ge,
                              uint32_t r)
{
    for (size_t i = 0; i < (cur_size >> 1); i++) {
        for (uint32_t k = 0; k < L; k++) {
            uint32_t a = src[(2 * i) * L + k], b = src[(2 * i + 1) * L + k];
            f[i * L + k] = fadd(a, fmul(fsub(b, a), r));
        }
        if (!ge[2 * i] && !ge[2 * i + 1]) {
            ge[i] = 0;
            for (uint32_t k = 0; k < L; k++) hg[i * L + k] = 0;
        } else {
            ge[i] = 1;
            for (uint32_t k = 0; k < L; k++) {
                uint32_t a = hg[(2 * i) * L + k], b = hg[(2 * i + 1) * L + k];
                hg[i * L + k] = fadd(a, fmul(fsub(b, a), r));
            }
        }
    }
}

static void tr_append_f(transcript *t, const uint32_t *f, uint32_t L) { tr_append(t, f, (size_t)L * 4); }

/* sumcheck_prove_gkr_layer (include/LinearGKR/sumcheck.hpp:52-111) */
static void sumcheck_prove_layer(const oracle_circuit *c, uint32_t l, uint32_t L, uint32_t R, const uint32_t *v,
                                 uint32_t **rz1, uint32_t **rz2, uint32_t alpha, uint32_t beta, transcript *t,
                                 rep_state *st)
{
    uint32_t n = c->nb_in_vars[l];
    size_t N = (size_t)1 << n;
    uint32_t *p = (uint32_t *)malloc(3 * L * 4);
    for (uint32_t j = 0; j < R; j++) {
        prepare_g_x(c, l, L, v, rz1[j], rz2[j], alpha, beta, &st[j]);
        st[j].nrx = st[j].nry = 0;
    }
    for (uint32_t i_var = 0; i_var < 2 * n; i_var++) {
        for (uint32_t j = 0; j < R; j++) {
            rep_state *s = &st[j];
            uint32_t k = i_var < n ? i_var : i_var - n;
            if (i_var == n) prepare_h_y(c, l, L, s); /* poly_evals_at, sumcheck_helper.hpp:302-317 */
            const uint32_t *src = (k == 0) ? v : s->f;
            poly_eval_at(L, n, k, src, s->hg, s->gate_exists, p);
            tr_append_f(t, p, L);
            tr_append_f(t, p + L, L);
            tr_append_f(t, p + 2 * L, L);
            uint32_t r = tr_challenge(t);
            receive_challenge(L, N >> k, src, s->f, s->hg, s->gate_exists, r);
            if (i_var < n) s->rx[s->nrx++] = r;
            else s->ry[s->nry++] = r;
            if (i_var == n - 1) tr_append_f(t, s->f, L); /* vx_claim = v_evals[0] */
        }
    }
    for (uint32_t j = 0; j < R; j++) tr_append_f(t, st[j].f, L); /* vy_claim */
    free(p);
}

int oracle_prove(const oracle_circuit *c, uint32_t L, int hash_kind, const uint32_t *witness, uint8_t **proof,
                 size_t *proof_len, uint32_t *claimed)
{
    int rc = check_shape(c);
    if (rc) return rc;
    const uint32_t R = oracle_num_repetitions();
    const uint32_t D = c->nb_layers;
    size_t nvals = values_len(c, L);
    uint32_t *vals = (uint32_t *)malloc(nvals * 4);
    oracle_evaluate(c, L, witness, vals);
    size_t *off = (size_t *)malloc((D + 1) * sizeof(size_t));
    off[0] = 0;
    for (uint32_t l = 0; l < D; l++) off[l + 1] = off[l] + ((size_t)1 << c->nb_in_vars[l]) * L;

    uint32_t maxn = 0, maxm = 0;
    for (uint32_t l = 0; l < D; l++) {
        if (c->nb_in_vars[l] > maxn) maxn = c->nb_in_vars[l];
        if (c->nb_out_vars[l] > maxm) maxm = c->nb_out_vars[l];
    }
    size_t maxN = (size_t)1 << maxn, maxM = (size_t)1 << maxm;
    rep_state *st = (rep_state *)calloc(R, sizeof(rep_state));
    for (uint32_t j = 0; j < R; j++) {
        st[j].f = (uint32_t *)malloc(maxN * L * 4);
        st[j].hg = (uint32_t *)malloc(maxN * L * 4);
        st[j].gate_exists = (uint8_t *)malloc(maxN);
        st[j].eqz = (uint32_t *)malloc(maxM * 4);
        st[j].eq2 = (uint32_t *)malloc(maxM * 4);
        st[j].eqx = (uint32_t *)malloc(maxN * 4);
        st[j].rx = (uint32_t *)malloc(32 * 4);
        st[j].ry = (uint32_t *)malloc(32 * 4);
    }

    /* Prover::prove (LinearGKR.hpp:58-90): RAW commitment = the input-layer values */
    size_t N0 = (size_t)1 << c->nb_in_vars[0];
    transcript t;
    tr_init(&t, hash_kind, N0 * L * 4 + 4096);
    tr_append(&t, witness, N0 * L * 4);
    grind(&t);

    /* gkr_prove (include/LinearGKR/gkr.hpp:24-73) */
    uint32_t mlast = c->nb_out_vars[D - 1];
    uint32_t **rz1 = (uint32_t **)malloc(R * sizeof(uint32_t *));
    uint32_t **rz2 = (uint32_t **)malloc(R * sizeof(uint32_t *));
    for (uint32_t j = 0; j < R; j++) {
        rz1[j] = (uint32_t *)calloc(32, 4);
        rz2[j] = (uint32_t *)calloc(32, 4);
    }
    for (uint32_t i = 0; i < mlast; i++)
        for (uint32_t j = 0; j < R; j++) {
            rz1[j][i] = tr_challenge(&t);
            rz2[j][i] = 0;
        }
    uint32_t alpha = 1, beta = 0;
    for (uint32_t j = 0; j < R; j++) oracle_eval_multilinear(vals + off[D], L, rz1[j], mlast, claimed + j * L);
    for (int l = (int)D - 1; l >= 0; l--) {
        sumcheck_prove_layer(c, (uint32_t)l, L, R, vals + off[l], rz1, rz2, alpha, beta, &t, st);
        alpha = tr_challenge(&t);
        beta = tr_challenge(&t);
        for (uint32_t j = 0; j < R; j++) {
            memcpy(rz1[j], st[j].rx, st[j].nrx * 4);
            memcpy(rz2[j], st[j].ry, st[j].nry * 4);
        }
    }
    /* RAW openings are empty (include/poly_commit/raw.hpp:41-58) */
    *proof = t.bytes;
    *proof_len = t.len;
    for (uint32_t j = 0; j < R; j++) {
        free(rz1[j]); free(rz2[j]);
        free(st[j].f); free(st[j].hg); free(st[j].gate_exists); free(st[j].eqz); free(st[j].eq2);
        free(st[j].eqx); free(st[j].rx); free(st[j].ry);
    }
    free(rz1); free(rz2); free(st); free(vals); free(off);
    return 0;
}

/* ------------------------------------------------------------- verifier */
/* eval_sparse_circuit_connect_poly (include/LinearGKR/sumcheck_verifier_utils.hpp:22-56) */
static uint32_t eval_sparse(const uint32_t *gates, uint32_t ngates, int nb_input, uint32_t m, const uint32_t *rz1,
                            const uint32_t *rz2, uint32_t alpha, uint32_t beta, uint32_t n, const uint32_t *rx,
                            const uint32_t *ry)
{
    size_t M = (size_t)1 << m, N = (size_t)1 << n;
    uint32_t *e1 = (uint32_t *)malloc(M * 4), *e2 = (uint32_t *)malloc(M * 4);
    uint32_t *ex = (uint32_t *)malloc(N * 4), *ey = (uint32_t *)malloc(N * 4);
    oracle_eq_evals(rz1, m, alpha, e1);
    oracle_eq_evals(rz2, m, beta, e2);
    oracle_eq_evals(rx, n, 1, ex);
    if (nb_input == 2) oracle_eq_evals(ry, n, 1, ey);
    uint32_t v = 0;
    int stride = nb_input == 2 ? 4 : 3;
    for (uint32_t g = 0; g < ngates; g++) {
        const uint32_t *q = gates + (size_t)stride * g;
        uint32_t o = q[nb_input], coef = q[nb_input + 1];
        uint32_t prod = fadd(e1[o], e2[o]);
        prod = fmul(prod, ex[q[0]]);
        if (nb_input == 2) prod = fmul(prod, ey[q[1]]);
        v = fadd(v, fmul(prod, coef));
    }
    free(e1); free(e2); free(ex); free(ey);
    return v;
}

/* degree_2_eval (sumcheck_verifier_utils.hpp:10-18), per lane */
static uint32_t degree_2_eval(uint32_t v0, uint32_t v1, uint32_t v2, uint32_t x)
{
    const uint32_t inv2 = 1u << 30;
    uint32_t c2 = fmul(inv2, fadd(fsub(v2, fmul(v1, 2)), v0));
    uint32_t c1 = fsub(fsub(v1, v0), c2);
    return fadd(v0, fmul(fadd(fmul(c2, x), c1), x));
}

static void read_f(const uint8_t *p, uint32_t L, uint32_t *out)
{
    for (uint32_t k = 0; k < L; k++) out[k] = oracle_m31_from_bytes(p + 4 * k);
}

int oracle_verify(const oracle_circuit *c, uint32_t L, int hash_kind, const uint32_t *claimed, const uint8_t *proof,
                  size_t proof_len)
{
    int rc = check_shape(c);
    if (rc) return rc;
    const uint32_t R = oracle_num_repetitions();
    const uint32_t D = c->nb_layers;
    const size_t S = (size_t)L * 4;
    size_t N0 = (size_t)1 << c->nb_in_vars[0];
    size_t need = N0 * S + 32;
    for (uint32_t l = 0; l < D; l++) need += ((size_t)6 * c->nb_in_vars[l] * R + 2 * R) * S;
    if (proof_len < need) return 0;

    /* Verifier::verify (LinearGKR.hpp:104-149) */
    uint32_t *commit = (uint32_t *)malloc(N0 * S);
    for (size_t i = 0; i < N0; i++) read_f(proof + i * S, L, commit + i * L);
    transcript t;
    tr_init(&t, hash_kind, need + 1024);
    tr_append(&t, proof, N0 * S);
    grind(&t);
    size_t idx = N0 * S + 32;

    /* gkr_verify (gkr.hpp:75-122) */
    uint32_t mlast = c->nb_out_vars[D - 1];
    uint32_t *rz1 = (uint32_t *)calloc((size_t)R * 32, 4), *rz2 = (uint32_t *)calloc((size_t)R * 32, 4);
    uint32_t *rx = (uint32_t *)calloc((size_t)R * 32, 4), *ry = (uint32_t *)calloc((size_t)R * 32, 4);
    for (uint32_t i = 0; i < mlast; i++)
        for (uint32_t j = 0; j < R; j++) rz1[j * 32 + i] = tr_challenge(&t);
    uint32_t alpha = 1, beta = 0;
    uint32_t *cv1 = (uint32_t *)malloc(R * S), *cv2 = (uint32_t *)calloc(R, S);
    memcpy(cv1, claimed, R * S);
    uint32_t *sum = (uint32_t *)malloc(R * S), *ev = (uint32_t *)malloc(3 * S);
    uint32_t *vx = (uint32_t *)malloc(R * S), *vy = (uint32_t *)malloc(R * S);
    int ok = 1;
    for (int l = (int)D - 1; l >= 0; l--) {
        /* sumcheck_verify_gkr_layer (include/LinearGKR/sumcheck.hpp:113-176) */
        uint32_t n = c->nb_in_vars[l], m = c->nb_out_vars[l];
        for (uint32_t j = 0; j < R; j++)
            for (uint32_t k = 0; k < L; k++)
                sum[j * L + k] = fadd(fmul(cv1[j * L + k], alpha), fmul(cv2[j * L + k], beta));
        for (uint32_t i_var = 0; i_var < 2 * n; i_var++) {
            for (uint32_t j = 0; j < R; j++) {
                read_f(proof + idx, L, ev);
                read_f(proof + idx + S, L, ev + L);
                read_f(proof + idx + 2 * S, L, ev + 2 * L);
                idx += 3 * S;
                tr_append(&t, ev, 3 * S);
                uint32_t r = tr_challenge(&t);
                if (i_var < n) rx[j * 32 + i_var] = r;
                else ry[j * 32 + i_var - n] = r;
                for (uint32_t k = 0; k < L; k++) {
                    ok &= fadd(ev[k], ev[L + k]) == sum[j * L + k];
                    sum[j * L + k] = degree_2_eval(ev[k], ev[L + k], ev[2 * L + k], r);
                }
                if (i_var == n - 1) {
                    read_f(proof + idx, L, vx + j * L);
                    idx += S;
                    uint32_t a = eval_sparse(c->add[l], c->n_add[l], 1, m, rz1 + j * 32, rz2 + j * 32, alpha, beta,
                                             n, rx + j * 32, NULL);
                    for (uint32_t k = 0; k < L; k++) sum[j * L + k] = fsub(sum[j * L + k], fmul(vx[j * L + k], a));
                    tr_append(&t, vx + j * L, S);
                }
            }
        }
        for (uint32_t j = 0; j < R; j++) {
            read_f(proof + idx, L, vy + j * L);
            idx += S;
            uint32_t mv = eval_sparse(c->mul[l], c->n_mul[l], 2, m, rz1 + j * 32, rz2 + j * 32, alpha, beta, n,
                                      rx + j * 32, ry + j * 32);
            for (uint32_t k = 0; k < L; k++)
                ok &= sum[j * L + k] == fmul(fmul(vx[j * L + k], vy[j * L + k]), mv);
            tr_append(&t, vy + j * L, S);
        }
        alpha = tr_challenge(&t);
        beta = tr_challenge(&t);
        memcpy(rz1, rx, (size_t)R * 32 * 4);
        memcpy(rz2, ry, (size_t)R * 32 * 4);
        memcpy(cv1, vx, R * S);
        memcpy(cv2, vy, R * S);
    }
    /* RawPC::verify (include/poly_commit/raw.hpp:75-78) for rz1 and rz2 of every rep */
    uint32_t *e = (uint32_t *)malloc(S);
    uint32_t n0 = c->nb_in_vars[0];
    for (uint32_t j = 0; j < R; j++) {
        oracle_eval_multilinear(commit, L, rz1 + j * 32, n0, e);
        ok &= memcmp(e, cv1 + j * L, S) == 0;
        oracle_eval_multilinear(commit, L, rz2 + j * 32, n0, e);
        ok &= memcmp(e, cv2 + j * L, S) == 0;
    }
    free(e); free(commit); free(rz1); free(rz2); free(rx); free(ry); free(cv1); free(cv2);
    free(sum); free(ev); free(vx); free(vy); free(t.bytes);
    return ok ? 1 : 0;
}
