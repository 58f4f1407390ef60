// MI355X (gfx950) kernels for the GKR prove path over the BN254 scalar field Fr with one
// lane and one repetition (num_repetitions = ceil(90 / 254) = 1, config.hpp:61-65).
//
// The reference names BN254 only as a Field_type (include/configuration/config.hpp:12-16)
// and has no BN254 field or prover code; this path runs the same algorithm as the M31
// prover (prover_m31.hip) with the build's own Fr (bn254.hpp):
//   k_bn_forward <- CircuitLayer::evaluate             include/circuit/circuit.hpp:92-105
//   k_bn_eq      <- _eq_evals_at + the alpha/beta combination
//                                                      include/LinearGKR/sumcheck_common.hpp:16-53
//   k_bn_gx / k_bn_hy <- _prepare_g_x_vals / _prepare_h_y_vals
//                                                      include/LinearGKR/sumcheck_helper.hpp:172-261
//   k_bn_round + k_bn_fin, k_bn_tail <- poly_eval_at / receive_challenge and the transcript
//                                                      sumcheck_helper.hpp:68-133, sumcheck.hpp:72-103
//   k_bn_dot     <- eval_multilinear for claimed_v     include/poly_commit/poly.hpp:88-104
//
// Layout: one table entry is one Fr in Montgomery form (8 little-endian u32 limbs, 32
// bytes); a thread owns whole entries (two 16-byte loads).  Proof and transcript bytes
// are canonical values (32 LE bytes).  Every product is a CIOS Montgomery multiply on
// 32-bit limbs (v_mad_u64_u32 chains): this path is bound by integer multiply issue, not
// by HBM (a fused round moves 192 bytes per pair against 7 field multiplies).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "bn254.hpp"
#include "hash_dev.hpp"
#include "internal.hpp"
#include "xcd.hpp"

namespace gkrx {

namespace {

constexpr int kBnThreads = 256;
constexpr uint32_t kBnMaxBlocks = 1024;
constexpr uint32_t kBnTailThreads = 512;
constexpr uint32_t kBnTailRows = 1024;  // rounds with at most this many rows run in k_bn_tail
constexpr uint32_t kBnChalStride = 64;  // Fr per layer slot: rx in [0, 32), ry in [32, 64)

__device__ __forceinline__ Fr ld_fr(const uint32_t *p)
{
    const uint4 a = *reinterpret_cast<const uint4 *>(p), b = *reinterpret_cast<const uint4 *>(p + 4);
    return Fr{{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w}};
}
__device__ __forceinline__ void st_fr(uint32_t *p, const Fr &v)
{
    *reinterpret_cast<uint4 *>(p) = make_uint4(v.w[0], v.w[1], v.w[2], v.w[3]);
    *reinterpret_cast<uint4 *>(p + 4) = make_uint4(v.w[4], v.w[5], v.w[6], v.w[7]);
}
__device__ __forceinline__ Fr fr_one() { return Fr{GKRX_BN_ONE}; }
__device__ __forceinline__ Fr fold_fr(const Fr &a, const Fr &b, const Fr &r) { return fr_add(a, fr_mont_mul(fr_sub(b, a), r)); }
// x * c for a gate coefficient c (a small integer; x in Montgomery form stays so)
__device__ __forceinline__ Fr coef_mul(const Fr &x, uint32_t c)
{
    if (c == 1u) return x;
    if (c == 0u) return fr_zero();
    return fr_mont_mul(x, fr_mont_of_u32(c));
}
__device__ __forceinline__ Fr shfl_xor_fr(const Fr &v, int off)
{
    Fr o;
#pragma unroll
    for (int i = 0; i < 8; i++) o.w[i] = __shfl_xor(v.w[i], off, 64);
    return o;
}
__device__ __forceinline__ Fr wave_sum(Fr v)
{
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) v = fr_add(v, shfl_xor_fr(v, off));
    return v;
}

// ------------------------------------------------------------------ elementwise
__global__ void __launch_bounds__(kBnThreads) k_bn_to_mont(uint32_t *__restrict__ v, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kBnThreads + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBnThreads)
        st_fr(v + 8 * i, fr_to_mont(ld_fr(v + 8 * i)));
}

// ------------------------------------------------------------------ forward pass
__global__ void __launch_bounds__(kBnThreads) k_bn_forward(const uint32_t *__restrict__ vin, uint32_t *__restrict__ vout,
                                                           const uint32_t *__restrict__ ptr,
                                                           const uint32_t *__restrict__ ent, uint32_t M)
{
    const uint64_t tile = xcd_tile(blockIdx.x, gridDim.x);
    for (uint64_t o = tile * kBnThreads + threadIdx.x; o < M; o += (uint64_t)gridDim.x * kBnThreads) {
        Fr acc = fr_zero();
        const uint32_t e1 = ptr[o + 1];
        for (uint32_t e = ptr[o]; e < e1; e++) {
            const uint32_t a = ent[3 * (uint64_t)e], b = ent[3 * (uint64_t)e + 1], c = ent[3 * (uint64_t)e + 2];
            Fr va = ld_fr(vin + 8 * (uint64_t)a);
            if (b != kAddMark) va = fr_mont_mul(va, ld_fr(vin + 8 * (uint64_t)b));
            acc = fr_add(acc, coef_mul(va, c));
        }
        st_fr(vout + 8 * o, acc);
    }
}

// ------------------------------------------------------------------ eq tables
// out[z] = f1 eq(r1, z) + f2 eq(r2, z), z < 2^m (f1 = ab[0], f2 = ab[1]; without ab, f1 = 1
// and no second term), by halves: eq(r, z) = lo[z mod 2^h] hi[z >> h] with h = ceil(m / 2),
// lo[i] = prod_{b < h} (bit_b(i) ? r_b : 1 - r_b) and hi[j] = f prod_{b < m - h} (bit_b(j) ?
// r_{h+b} : 1 - r_{h+b}).  k_bn_eq_parts forms the small tables (lo1, hi1, then lo2, hi2), one
// entry per thread with a chain of at most ceil(m / 2) products; k_bn_eq takes one product per
// term per z.
__host__ __device__ inline uint32_t bn_eq_lo_bits(uint32_t m) { return (m + 1) / 2; }
__host__ __device__ inline uint64_t bn_eq_parts_fr(uint32_t m)  // Fr of one term's two tables
{
    return (1ull << bn_eq_lo_bits(m)) + (1ull << (m - bn_eq_lo_bits(m)));
}

__global__ void __launch_bounds__(kBnThreads) k_bn_eq_parts(uint32_t *__restrict__ parts, const uint32_t *__restrict__ r1,
                                                            const uint32_t *__restrict__ r2,
                                                            const uint32_t *__restrict__ ab, uint32_t m)
{
    const uint32_t h = bn_eq_lo_bits(m);
    const uint64_t per = bn_eq_parts_fr(m), nlo = 1ull << h;
    const uint64_t i = (uint64_t)blockIdx.x * kBnThreads + threadIdx.x;
    if (i >= per * (ab ? 2u : 1u)) return;
    const uint32_t term = (uint32_t)(i / per);
    const uint64_t j = i % per;
    const uint32_t *r = term ? r2 : r1;
    const bool lo = j < nlo;
    const uint64_t idx = lo ? j : j - nlo;
    const uint32_t b0 = lo ? 0u : h, nb = lo ? h : m - h;
    Fr a = (!lo && ab) ? ld_fr(ab + 8 * term) : fr_one();
    for (uint32_t b = 0; b < nb; b++) {
        const Fr x = ld_fr(r + 8 * (b0 + b));
        a = fr_mont_mul(a, (idx >> b) & 1u ? x : fr_sub(fr_one(), x));
    }
    st_fr(parts + 8 * i, a);
}

__global__ void __launch_bounds__(kBnThreads) k_bn_eq(uint32_t *__restrict__ out, const uint32_t *__restrict__ parts,
                                                      uint32_t m, uint32_t two)
{
    const uint32_t h = bn_eq_lo_bits(m);
    const uint64_t per = bn_eq_parts_fr(m), nlo = 1ull << h, M = 1ull << m, mask = nlo - 1;
    for (uint64_t z = (uint64_t)blockIdx.x * kBnThreads + threadIdx.x; z < M; z += (uint64_t)gridDim.x * kBnThreads) {
        const uint64_t li = z & mask, hi = nlo + (z >> h);
        Fr v = fr_mont_mul(ld_fr(parts + 8 * li), ld_fr(parts + 8 * hi));
        if (two) v = fr_add(v, fr_mont_mul(ld_fr(parts + 8 * (per + li)), ld_fr(parts + 8 * (per + hi))));
        st_fr(out + 8 * z, v);
    }
}

// ------------------------------------------------------------------ g_x and h_y
// out[x] = the sum of term(e) over the entries e of CSR row x, for the kBnThreads rows from
// x0 on.  The rows' entries are contiguous in the CSR; they are spread over the block's
// threads kBnThreads at a time, one term per thread (its gathers and products do not wait
// on the other entries of its row), staged in LDS, and thread t adds the staged terms of row
// x0 + t.  Every thread of the block calls it.
template <typename Term>
__device__ __forceinline__ void bn_row_sums(uint32_t *__restrict__ out, const uint32_t *__restrict__ ptr,
                                            const uint32_t *__restrict__ ent, uint64_t x0, uint32_t N, Term term)
{
    __shared__ Fr staged[kBnThreads];
    __shared__ uint32_t rp[kBnThreads + 1];
    const uint32_t t = threadIdx.x;
    const uint32_t rows = (uint32_t)(N - x0 < kBnThreads ? N - x0 : kBnThreads);
    if (t <= rows) rp[t] = ptr[x0 + t];
    if (t == 0) rp[rows] = ptr[x0 + rows];
    __syncthreads();
    const uint32_t e0 = rp[0], e1 = rp[rows];
    const uint32_t r0 = t < rows ? rp[t] : 0u, r1 = t < rows ? rp[t + 1] : 0u;
    Fr acc = fr_zero();
    for (uint32_t c = e0; c < e1; c += kBnThreads) {
        const uint32_t e = c + t;
        if (e < e1) staged[t] = term(ent + 4ull * e);
        __syncthreads();
        const uint32_t lo = r0 > c ? r0 : c, hi = r1 < c + kBnThreads ? r1 : c + kBnThreads;
        for (uint32_t j = lo; j < hi; j++) acc = fr_add(acc, staged[j - c]);
        __syncthreads();
    }
    if (t < rows) st_fr(out + 8 * (x0 + t), acc);
    __syncthreads();
}

// hg[x] = sum over the gates with first input x of c eqZ[z] v[y] (mul) or c eqZ[z] (add);
// CSR entry {y or kAddMark, z, c, x}
__global__ void __launch_bounds__(kBnThreads) k_bn_gx(uint32_t *__restrict__ hg, const uint32_t *__restrict__ v,
                                                      const uint32_t *__restrict__ eqz, const uint32_t *__restrict__ ptr,
                                                      const uint32_t *__restrict__ ent, uint32_t N)
{
    static_assert(kGxWords == 4, "g_x entry");
    const uint32_t ntiles = (N + kBnThreads - 1) / kBnThreads;
    for (uint32_t tile = xcd_tile(blockIdx.x, gridDim.x); tile < ntiles; tile += gridDim.x)
        bn_row_sums(hg, ptr, ent, (uint64_t)tile * kBnThreads, N, [&](const uint32_t *en) {
            const uint32_t b = en[0], z = en[1], c = en[2];
            Fr term = ld_fr(eqz + 8 * (uint64_t)z);
            if (b != kAddMark) term = fr_mont_mul(term, ld_fr(v + 8 * (uint64_t)b));
            return coef_mul(term, c);
        });
}

// s[y] = sum over the mul gates with second input y of c eqZ[z] eqX[x]; CSR entry {x, z, c, y}
__global__ void __launch_bounds__(kBnThreads) k_bn_hy(uint32_t *__restrict__ s, const uint32_t *__restrict__ eqz,
                                                      const uint32_t *__restrict__ eqx, const uint32_t *__restrict__ ptr,
                                                      const uint32_t *__restrict__ ent, uint32_t N)
{
    static_assert(kHyWords == 4, "h_y entry");
    const uint32_t ntiles = (N + kBnThreads - 1) / kBnThreads;
    for (uint32_t tile = xcd_tile(blockIdx.x, gridDim.x); tile < ntiles; tile += gridDim.x)
        bn_row_sums(s, ptr, ent, (uint64_t)tile * kBnThreads, N, [&](const uint32_t *en) {
            const uint32_t x = en[0], z = en[1], c = en[2];
            return coef_mul(fr_mont_mul(ld_fr(eqz + 8 * (uint64_t)z), ld_fr(eqx + 8 * (uint64_t)x)), c);
        });
}

// ------------------------------------------------------------------ rounds
// One pair i of round k: FUSED folds rows 4i..4i+3 of the round-(k-1) tables with r into
// rows 2i, 2i+1 of the round-k tables (stored); then p0 += f0 h0, p1 += f1 h1,
// p2 += (f0 + f1)(h0 + h1).
template <bool FUSED>
__device__ __forceinline__ void bn_round_item(const uint32_t *__restrict__ fs, uint32_t *__restrict__ fd,
                                              const uint32_t *__restrict__ hs, uint32_t *__restrict__ hd, uint64_t i,
                                              const Fr &r, Fr (&acc)[3])
{
    Fr f0, f1, h0, h1;
    if (FUSED) {
        const uint32_t *fq = fs + 32 * i, *hq = hs + 32 * i;
        f0 = fold_fr(ld_fr(fq), ld_fr(fq + 8), r);
        f1 = fold_fr(ld_fr(fq + 16), ld_fr(fq + 24), r);
        h0 = fold_fr(ld_fr(hq), ld_fr(hq + 8), r);
        h1 = fold_fr(ld_fr(hq + 16), ld_fr(hq + 24), r);
        st_fr(fd + 16 * i, f0);
        st_fr(fd + 16 * i + 8, f1);
        st_fr(hd + 16 * i, h0);
        st_fr(hd + 16 * i + 8, h1);
    } else {
        f0 = ld_fr(fs + 16 * i);
        f1 = ld_fr(fs + 16 * i + 8);
        h0 = ld_fr(hs + 16 * i);
        h1 = ld_fr(hs + 16 * i + 8);
    }
    acc[0] = fr_add(acc[0], fr_mont_mul(f0, h0));
    acc[1] = fr_add(acc[1], fr_mont_mul(f1, h1));
    acc[2] = fr_add(acc[2], fr_mont_mul(fr_add(f0, f1), fr_add(h0, h1)));
}

// Sums acc[3] over the block into out[0..3) (LDS or global, 8 words each).  red: LDS of
// blockDim / 64 * 3 Fr.  Every thread of the block calls it.
__device__ __forceinline__ void bn_block_sum3(Fr (&acc)[3], Fr *red, uint32_t *out)
{
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
#pragma unroll
    for (int e = 0; e < 3; e++) {
        const Fr s = wave_sum(acc[e]);
        if (lane == 0) red[3 * wave + e] = s;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        Fr s = fr_zero();
        for (uint32_t w = 0; w < nw; w++) s = fr_add(s, red[3 * w + threadIdx.x]);
        for (int i = 0; i < 8; i++) out[8 * threadIdx.x + i] = s.w[i];
    }
    __syncthreads();
}

struct BnRoundParams {
    const uint32_t *fs, *hs;
    uint32_t *fd, *hd;
    const uint32_t *rprev;  // challenge of round k-1 (Montgomery)
    uint32_t *partials;     // [block][3] Fr
    uint32_t items;         // pairs in round k
};

template <bool FUSED>
__global__ void __launch_bounds__(kBnThreads) k_bn_round(BnRoundParams p)
{
    __shared__ Fr red[3 * kBnThreads / 64];
    const Fr r = FUSED ? ld_fr(p.rprev) : fr_zero();
    Fr acc[3] = {fr_zero(), fr_zero(), fr_zero()};
    for (uint64_t i = (uint64_t)blockIdx.x * kBnThreads + threadIdx.x; i < p.items; i += (uint64_t)gridDim.x * kBnThreads)
        bn_round_item<FUSED>(p.fs, p.fd, p.hs, p.hd, i, r, acc);
    bn_block_sum3(acc, red, p.partials + 24 * (uint64_t)blockIdx.x);
}

// ------------------------------------------------------------------ transcript
// The transcript (include/fiat_shamir/transcript.hpp:87-175) over the proof tail on the
// device; tstate = [length, hash start, digest[8]] in bytes of the tail.
struct BnTsArgs {
    int phase;            // 1 = x variables, 2 = y variables
    uint32_t k;
    int last;             // k == n - 1
    uint32_t *chal;       // this layer's slot: rx at [0, 32), ry at [32, 64) Fr
    uint32_t *vx, *vy;    // one Fr each (Montgomery)
    const uint32_t *fin;  // round-k f table (rows 0 and 1 are folded after the last round)
    uint32_t *tstate;
    uint32_t *tail;
    uint32_t *ab_out;     // alpha, beta drawn after the layer (Montgomery)
    int hash_kind;
};

__device__ __forceinline__ Fr bn_challenge(const uint32_t d[8]) { return fr_to_mont(fr_reduce256(d)); }

// Block-wide transcript step of round k (blockDim >= 64): the message p(0), p(1), p(2)
// (p(2) = 6 p1 + 3 p0 - 2 p2, sumcheck_helper.hpp:100; times vx in phase 2) as canonical
// bytes, one challenge, then the x claim (phase 1) or the y claim, alpha and beta (phase
// 2) after the last round (sumcheck.hpp:90-103, gkr.hpp:65-66).  Hash inputs are staged
// in LDS; at most one element (the pending x claim) precedes a message.
__device__ __forceinline__ void bn_transcript_step(const uint32_t *sums, const BnTsArgs &a, Fr *r_out)
{
    __shared__ uint32_t ts[10], hb[8 + 24], dig[8];
    const uint32_t t = threadIdx.x, wave = t >> 6, lane = t & 63;
    if (t < 10) ts[t] = a.tstate[t];
    __syncthreads();
    const uint32_t base = ts[0], hs0 = ts[1];
    const uint32_t P = (base - hs0) / 4;
    if (t < P && P <= 8) hb[8 - P + t] = a.tail[hs0 / 4 + t];
    if (t < 3) {
        // thread e forms message element e: p0, p1, or p2 = 6 p1 + 3 p0 - 2 p2, times vx in
        // phase 2, as canonical bytes (three threads: the Montgomery products run side by side)
        const uint32_t e = t;
        const Fr s0 = ld_fr(sums), s1 = ld_fr(sums + 8), s2 = ld_fr(sums + 16);
        Fr q = e == 0 ? s0 : s1;
        if (e == 2) {
            const Fr s1x2 = fr_add(s1, s1), s1x3 = fr_add(s1x2, s1), s1x6 = fr_add(s1x3, s1x3);
            const Fr s0x3 = fr_add(fr_add(s0, s0), s0);
            q = fr_sub(fr_add(s1x6, s0x3), fr_add(s2, s2));
        }
        if (a.phase == 2) q = fr_mont_mul(q, ld_fr(a.vx));
        const Fr c = fr_from_mont(q);
        for (int i = 0; i < 8; i++) {
            a.tail[base / 4 + 8 * e + i] = c.w[i];
            hb[8 + 8 * e + i] = c.w[i];
        }
    }
    __syncthreads();
    if (wave == 0) {
        uint32_t d[8];
        hash_words_wave(a.hash_kind, P <= 8 ? hb + 8 - P : a.tail + hs0 / 4, P + 24, d);
        if (lane == 0) {
            const Fr r = bn_challenge(d);
            *r_out = r;
            for (int i = 0; i < 8; i++) a.chal[8 * ((a.phase == 2 ? 32u : 0u) + a.k) + i] = r.w[i];
            for (int i = 0; i < 8; i++) ts[2 + i] = d[i];
            ts[0] = base + 96;
            ts[1] = base + 96;
        }
    }
    __syncthreads();
    if (a.last) {
        const uint32_t len = ts[0];
        if (t == 0) {
            // the claim: f after the last fold (sumcheck_helper.hpp:302-341)
            const Fr v = fold_fr(ld_fr(a.fin), ld_fr(a.fin + 8), *r_out);
            st_fr(a.phase == 1 ? a.vx : a.vy, v);
            const Fr c = fr_from_mont(v);
            for (int i = 0; i < 8; i++) {
                a.tail[len / 4 + i] = c.w[i];
                hb[i] = c.w[i];
            }
            ts[0] = len + 32;
            ts[1] = a.phase == 1 ? len : len + 32;  // the x claim stays pending
        }
        __syncthreads();
        if (a.phase == 2) {
            if (wave == 0) {
                uint32_t d[8];
                hash_words_wave(a.hash_kind, hb, 8, d);
                if (lane == 0) {
                    st_fr(a.ab_out, bn_challenge(d));
                    for (int i = 0; i < 8; i++) dig[i] = d[i];
                }
            }
            __syncthreads();
            if (wave == 0) {
                uint32_t d[8];
                hash_words_wave(a.hash_kind, dig, 8, d);
                if (lane == 0) {
                    st_fr(a.ab_out + 8, bn_challenge(d));
                    for (int i = 0; i < 8; i++) ts[2 + i] = d[i];
                }
            }
        }
    }
    __syncthreads();
    if (t < 10) a.tstate[t] = ts[t];
    __syncthreads();
}

// Adds the [nb][3] Fr block partials into sums (LDS, 3 Fr): wave e < 3 takes element e.
__device__ __forceinline__ void bn_reduce_partials(const uint32_t *__restrict__ partials, uint32_t nb, uint32_t *sums)
{
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (wave < 3) {
        Fr s = fr_zero();
        for (uint32_t b = lane; b < nb; b += 64) s = fr_add(s, ld_fr(partials + 24 * (uint64_t)b + 8 * wave));
        s = wave_sum(s);
        if (lane == 0)
            for (int i = 0; i < 8; i++) sums[8 * wave + i] = s.w[i];
    }
    __syncthreads();
}

__global__ void __launch_bounds__(kBnThreads) k_bn_fin(const uint32_t *__restrict__ partials, uint32_t nb, BnTsArgs ts)
{
    __builtin_amdgcn_s_setprio(3);  // a hash chain: its waves issue first on a shared CU
    __shared__ uint32_t sums[24];
    __shared__ Fr rr;
    bn_reduce_partials(partials, nb, sums);
    bn_transcript_step(sums, ts, &rr);
}

// Rounds k0 .. n-1 of one phase in one workgroup of kBnTailThreads.  Table k lives in A for
// odd k and B for even k >= 2; round k0 (and round 1) reads f0 / h0.
struct BnTailParams {
    const uint32_t *f0, *h0;
    uint32_t *fA, *fB, *hA, *hB;
    uint32_t k0, n;
    BnTsArgs ts;
};

__global__ void __launch_bounds__(kBnTailThreads) k_bn_tail(BnTailParams p)
{
    __builtin_amdgcn_s_setprio(3);  // a hash chain: its waves issue first on a shared CU
    __shared__ Fr red[3 * kBnTailThreads / 64];
    __shared__ uint32_t sums[24];
    __shared__ Fr rr;
    if (threadIdx.x == 0) rr = p.k0 > 0 ? ld_fr(p.ts.chal + 8 * ((p.ts.phase == 2 ? 32u : 0u) + p.k0 - 1)) : fr_zero();
    __syncthreads();
    for (uint32_t k = p.k0; k < p.n; k++) {
        const uint64_t items = 1ull << (p.n - k - 1);
        const bool first = (k == p.k0) || (k == 1);
        const uint32_t *fs = first ? p.f0 : ((k - 1) & 1 ? p.fA : p.fB);
        const uint32_t *hs = first ? p.h0 : ((k - 1) & 1 ? p.hA : p.hB);
        uint32_t *fd = (k & 1) ? p.fA : p.fB;
        uint32_t *hd = (k & 1) ? p.hA : p.hB;
        const Fr r = rr;
        Fr acc[3] = {fr_zero(), fr_zero(), fr_zero()};
        for (uint64_t i = threadIdx.x; i < items; i += kBnTailThreads) {
            if (k == 0) bn_round_item<false>(fs, nullptr, hs, nullptr, i, r, acc);
            else bn_round_item<true>(fs, fd, hs, hd, i, r, acc);
        }
        bn_block_sum3(acc, red, sums);
        BnTsArgs a = p.ts;
        a.k = k;
        a.last = (k == p.n - 1);
        a.fin = (k == 0) ? fs : fd;
        bn_transcript_step(sums, a, &rr);
    }
}

// claimed_v = sum_z eq[z] * vals[z] (gkr.hpp:51-55): block partials, then k_bn_dot_fin
// writes the canonical value.
__global__ void __launch_bounds__(kBnThreads) k_bn_dot(const uint32_t *__restrict__ vals, const uint32_t *__restrict__ eq,
                                                       uint32_t M, uint32_t *__restrict__ partials)
{
    __shared__ Fr red[3 * kBnThreads / 64];
    Fr acc[3] = {fr_zero(), fr_zero(), fr_zero()};
    for (uint64_t z = (uint64_t)blockIdx.x * kBnThreads + threadIdx.x; z < M; z += (uint64_t)gridDim.x * kBnThreads)
        acc[0] = fr_add(acc[0], fr_mont_mul(ld_fr(vals + 8 * z), ld_fr(eq + 8 * z)));
    bn_block_sum3(acc, red, partials + 24 * (uint64_t)blockIdx.x);
}

__global__ void __launch_bounds__(kBnThreads) k_bn_dot_fin(const uint32_t *__restrict__ partials, uint32_t nb,
                                                           uint32_t *__restrict__ claims)
{
    __shared__ uint32_t sums[24];
    bn_reduce_partials(partials, nb, sums);
    if (threadIdx.x == 0) {
        const Fr c = fr_from_mont(ld_fr(sums));
        for (int i = 0; i < 8; i++) claims[i] = c.w[i];
    }
}

template <typename F>
inline void timed(Profiler *pf, int cls, hipStream_t st, F &&launch)
{
    if (pf) pf->begin(cls, st);
    launch();
    if (pf) pf->end(st);
}

uint32_t bn_grid(uint64_t work, uint64_t cap)
{
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((work + kBnThreads - 1) / kBnThreads, cap));
}

// eq over m variables into out (two terms when ab is given; see k_bn_eq_parts)
void bn_launch_eq(uint32_t *out, uint32_t *parts, const uint32_t *r1, const uint32_t *r2, const uint32_t *ab, uint32_t m,
                  hipStream_t st)
{
    const uint64_t nparts = bn_eq_parts_fr(m) * (ab ? 2u : 1u);
    hipLaunchKernelGGL(k_bn_eq_parts, dim3((uint32_t)((nparts + kBnThreads - 1) / kBnThreads)), dim3(kBnThreads), 0, st,
                       parts, r1, r2, ab, m);
    hipLaunchKernelGGL(k_bn_eq, dim3(bn_grid(1ull << m, 65536)), dim3(kBnThreads), 0, st, out,
                       (const uint32_t *)parts, m, ab ? 1u : 0u);
}


}  // namespace

// ------------------------------------------------------------------ host side
int dev_bn_to_mont(uint32_t *values, uint64_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_bn_to_mont, dim3(bn_grid(n, 65536)), dim3(kBnThreads), 0, st, values, n);
    GKRX_HIP_CHECK(hipGetLastError());
    return GKRX_OK;
}

int dev_bn_forward(const DevCircuit &dc, hipStream_t st, Profiler *pf)
{
    for (uint32_t l = 0; l < dc.D; l++) {
        const LayerDev &ly = dc.layers[l];
        const uint64_t M = 1ull << ly.m;
        timed(pf, kKForward, st, [&] {
            hipLaunchKernelGGL(k_bn_forward, dim3(bn_grid(M, 65536)), dim3(kBnThreads), 0, st, dc.values + ly.v_off,
                               dc.values + ly.out_off, ly.fw_ptr, ly.fw_ent, (uint32_t)M);
        });
    }
    GKRX_HIP_CHECK(hipGetLastError());
    return GKRX_OK;
}

void dev_bn_free_scratch(BnScratch &s)
{
    if (s.pool) (void)hipFree(s.pool);
    if (s.stage) (void)hipHostFree(s.stage);
    s = BnScratch{};
}

uint64_t dev_bn_scratch_bytes(const DevCircuit &dc, uint64_t tail_bytes)
{
    const uint64_t maxN = 1ull << dc.max_n, maxM = 1ull << dc.max_m;
    const uint64_t N2 = std::max<uint64_t>(2, maxN / 2), N4 = std::max<uint64_t>(2, maxN / 4);
    return 4 * (8 * maxM + 16 * maxN + 16 * N2 + 16 * N4 + 24ull * kBnMaxBlocks +
                16 * bn_eq_parts_fr(std::max(dc.max_n, dc.max_m)) + 8ull * (dc.D + 1) * kBnChalStride +
                16ull * (dc.D + 1) + 40 + tail_bytes / 4 + 64 + 15 * 64);
}

int dev_bn_ensure_scratch(BnScratch &s, const DevCircuit &dc, uint64_t tail_bytes)
{
    if (s.pool && s.max_n >= dc.max_n && s.max_m >= dc.max_m && s.D >= dc.D && s.tail_cap >= tail_bytes)
        return GKRX_OK;
    dev_bn_free_scratch(s);
    const uint64_t maxN = 1ull << dc.max_n, maxM = 1ull << dc.max_m;
    const uint64_t N2 = std::max<uint64_t>(2, maxN / 2), N4 = std::max<uint64_t>(2, maxN / 4);
    struct Piece { uint32_t **p; uint64_t u32; };
    Piece pieces[] = {
        {&s.eqz, 8 * maxM}, {&s.eqx, 8 * maxN}, {&s.h0, 8 * maxN}, {&s.hA, 8 * N2}, {&s.hB, 8 * N4},
        {&s.fA, 8 * N2}, {&s.fB, 8 * N4}, {&s.partials, 24ull * kBnMaxBlocks},
        {&s.eq_parts, 16 * bn_eq_parts_fr(std::max(dc.max_n, dc.max_m))},
        {&s.chal, 8ull * (dc.D + 1) * kBnChalStride}, {&s.ab, 16ull * (dc.D + 1)}, {&s.claims, 8},
        {&s.vxy, 16}, {&s.tstate, 16}, {&s.tail, (tail_bytes + 3) / 4 + 64}};
    uint64_t total = 0;
    for (auto &pc : pieces) total += (pc.u32 + 63) & ~63ull;
    GKRX_HIP_CHECK(hipMalloc(&s.pool, total * 4));
    uint64_t at = 0;
    for (auto &pc : pieces) {
        *pc.p = reinterpret_cast<uint32_t *>(s.pool) + at;
        at += (pc.u32 + 63) & ~63ull;
    }
    s.stage_u32 = 8 * kBnChalStride + 16 + 16 + 64;
    GKRX_HIP_CHECK(hipHostMalloc(&s.stage, s.stage_u32 * 4, hipHostMallocDefault));
    s.max_n = dc.max_n;
    s.max_m = dc.max_m;
    s.D = dc.D;
    s.tail_cap = tail_bytes;
    return GKRX_OK;
}

int dev_bn_prove(const DevCircuit &dc, BnScratch &s, int hash_kind, const uint32_t *rz1_host,
                 const uint32_t *tstate_host, const uint8_t *tail_prefix, uint32_t tail_prefix_len, hipStream_t st,
                 Profiler *pf)
{
    const uint32_t D = dc.D;
    const uint32_t mlast = dc.layers[D - 1].m;
    if (mlast > 32 || tail_prefix_len > 256) return GKRX_ERR_ARG;
    GKRX_HIP_CHECK(hipStreamSynchronize(st));  // the previous prove may still read the stage
    // seed slot D: rx = rz1 (Montgomery), ry = 0, alpha = 1, beta = 0
    uint32_t *seed = s.stage, *ab_seed = seed + 8 * kBnChalStride, *ts = ab_seed + 16, *pre = ts + 16;
    memset(seed, 0, 8 * kBnChalStride * 4);
    for (uint32_t i = 0; i < mlast; i++) {
        Fr c;
        memcpy(c.w, rz1_host + 8 * i, 32);
        const Fr mc = fr_to_mont(c);
        memcpy(seed + 8 * i, mc.w, 32);
    }
    const Fr one = {GKRX_BN_ONE};
    memcpy(ab_seed, one.w, 32);
    memset(ab_seed + 8, 0, 32);
    memcpy(ts, tstate_host, 40);
    memcpy(pre, tail_prefix, tail_prefix_len);
    auto chal_slot = [&](uint32_t l) { return s.chal + 8ull * kBnChalStride * l; };
    auto ab_slot = [&](uint32_t l) { return s.ab + 16ull * l; };
    GKRX_HIP_CHECK(hipMemcpyAsync(chal_slot(D), seed, 8 * kBnChalStride * 4, hipMemcpyHostToDevice, st));
    GKRX_HIP_CHECK(hipMemcpyAsync(ab_slot(D), ab_seed, 64, hipMemcpyHostToDevice, st));
    GKRX_HIP_CHECK(hipMemcpyAsync(s.tstate, ts, 40, hipMemcpyHostToDevice, st));
    if (tail_prefix_len) GKRX_HIP_CHECK(hipMemcpyAsync(s.tail, pre, tail_prefix_len, hipMemcpyHostToDevice, st));

    for (int l = (int)D - 1; l >= 0; l--) {
        const LayerDev &ly = dc.layers[l];
        const uint32_t n = ly.n, m = ly.m;
        const uint64_t N = 1ull << n, M = 1ull << m;
        const uint32_t *V = dc.values + ly.v_off;
        uint32_t *chal_l = chal_slot(l);
        const uint32_t *chal_p = chal_slot(l + 1);
        timed(pf, kKEq, st, [&] {
            bn_launch_eq(s.eqz, s.eq_parts, chal_p, chal_p + 8 * 32, ab_slot(l + 1), m, st);
        });
        if (l == (int)D - 1) {
            const uint32_t nb = bn_grid(M, kBnMaxBlocks);
            timed(pf, kKDot, st, [&] {
                hipLaunchKernelGGL(k_bn_dot, dim3(nb), dim3(kBnThreads), 0, st, dc.values + ly.out_off,
                                   (const uint32_t *)s.eqz, (uint32_t)M, s.partials);
                hipLaunchKernelGGL(k_bn_dot_fin, dim3(1), dim3(kBnThreads), 0, st, (const uint32_t *)s.partials, nb,
                                   s.claims);
            });
        }
        timed(pf, kKGx, st, [&] {
            hipLaunchKernelGGL(k_bn_gx, dim3(bn_grid(N, 65536)), dim3(kBnThreads), 0, st, s.h0, V,
                               (const uint32_t *)s.eqz, ly.gx_ptr, ly.gx_ent, (uint32_t)N);
        });
        for (int phase = 1; phase <= 2; phase++) {
            if (phase == 2) {
                // h0 (the phase-1 table 0) is no longer read: the phase-2 table s replaces it
                timed(pf, kKEq, st, [&] {
                    bn_launch_eq(s.eqx, s.eq_parts, chal_l, nullptr, nullptr, n, st);
                });
                timed(pf, kKHy, st, [&] {
                    hipLaunchKernelGGL(k_bn_hy, dim3(bn_grid(N, 65536)), dim3(kBnThreads), 0, st, s.h0,
                                       (const uint32_t *)s.eqz, (const uint32_t *)s.eqx, ly.hy_ptr, ly.hy_ent,
                                       (uint32_t)N);
                });
            }
            BnTsArgs ta{};
            ta.phase = phase;
            ta.chal = chal_l;
            ta.vx = s.vxy;
            ta.vy = s.vxy + 8;
            ta.tstate = s.tstate;
            ta.tail = s.tail;
            ta.ab_out = ab_slot(l);
            ta.hash_kind = hash_kind;
            uint32_t k0 = 0;
            while (k0 < n && (N >> k0) > kBnTailRows) k0++;
            for (uint32_t k = 0; k < k0; k++) {
                BnRoundParams rp{};
                rp.items = (uint32_t)(N >> (k + 1));
                rp.partials = s.partials;
                if (k == 0) {
                    rp.fs = V;
                    rp.hs = s.h0;
                } else {
                    rp.fs = (k == 1) ? V : ((k - 1) & 1 ? s.fA : s.fB);
                    rp.hs = (k == 1) ? s.h0 : ((k - 1) & 1 ? s.hA : s.hB);
                    rp.fd = (k & 1) ? s.fA : s.fB;
                    rp.hd = (k & 1) ? s.hA : s.hB;
                    rp.rprev = chal_l + 8 * ((phase == 2 ? 32u : 0u) + k - 1);
                }
                const uint32_t nb = bn_grid(rp.items, kBnMaxBlocks);
                const int cls = phase == 1 ? (k == 0 ? kKRoundPh1First : kKRoundPh1)
                                           : (k == 0 ? kKRoundPh2First : kKRoundPh2);
                timed(pf, cls, st, [&] {
                    if (k == 0) hipLaunchKernelGGL(k_bn_round<false>, dim3(nb), dim3(kBnThreads), 0, st, rp);
                    else hipLaunchKernelGGL(k_bn_round<true>, dim3(nb), dim3(kBnThreads), 0, st, rp);
                });
                BnTsArgs tk = ta;
                tk.k = k;
                tk.last = 0;
                timed(pf, kKRoundFin, st, [&] {
                    hipLaunchKernelGGL(k_bn_fin, dim3(1), dim3(kBnThreads), 0, st, (const uint32_t *)s.partials, nb, tk);
                });
            }
            BnTailParams tp{};
            tp.f0 = k0 <= 1 ? V : ((k0 - 1) & 1 ? s.fA : s.fB);
            tp.h0 = k0 <= 1 ? s.h0 : ((k0 - 1) & 1 ? s.hA : s.hB);
            tp.fA = s.fA;
            tp.fB = s.fB;
            tp.hA = s.hA;
            tp.hB = s.hB;
            tp.k0 = k0;
            tp.n = n;
            tp.ts = ta;
            timed(pf, phase == 1 ? kKTailPh1 : kKTailPh2, st, [&] {
                hipLaunchKernelGGL(k_bn_tail, dim3(1), dim3(kBnTailThreads), 0, st, tp);
            });
        }
    }
    GKRX_HIP_CHECK(hipGetLastError());
    return GKRX_OK;
}

}  // namespace gkrx
