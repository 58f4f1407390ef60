// MI355X (gfx950) kernels for the GKR prove path over M31 with L data-parallel lanes.
//
// Reference loops replaced (PolyhedraZK/Expander-cpp):
//   k_forward  <- CircuitLayer::evaluate              include/circuit/circuit.hpp:92-105
//   k_eq       <- _eq_evals_at / _primitive + combine include/LinearGKR/sumcheck_common.hpp:16-53,
//                                                     include/LinearGKR/sumcheck_helper.hpp:188-196
//   k_gx       <- _prepare_g_x_vals scatters          sumcheck_helper.hpp:198-229
//   k_hy       <- _prepare_h_y_vals scatter           sumcheck_helper.hpp:233-261
//   k_round    <- receive_challenge (fold of round k-1) fused with poly_eval_at (round k)
//                                                     sumcheck_helper.hpp:68-133
//   k_finish   <- the transcript steps of sumcheck_prove_gkr_layer / gkr_prove
//                                                     include/LinearGKR/sumcheck.hpp:72-103,
//                                                     include/LinearGKR/gkr.hpp:58-67
//   k_dot      <- eval_multilinear for claimed_v      include/poly_commit/poly.hpp:88-104
//
// Layout: every table keeps the reference's natural index order, element-major and
// lane-minor (one table entry = L consecutive u32 lanes = the bytes F::to_bytes writes).
// A thread owns 4 lanes of one entry (16-byte loads); L/4 threads cover one entry.
// Scatters of the reference become gathers over CSR lists built at circuit load, so no
// modular atomics are needed and every sum is exact field arithmetic (order-free).
// Folding is ping-pong between buffers (the reference folds in place, which is only
// safe sequentially).  The gate_exists skip of the reference is exact (hg is zero where
// no gate exists) and is therefore not needed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "hash_dev.hpp"
#include "internal.hpp"
#include "m31.hpp"

namespace gkrx {

namespace {

constexpr int kThreads = 256;
constexpr int kFinishThreads = 1024;
constexpr uint32_t kMaxBlocksPerRep = 256;
constexpr uint32_t kChalStride = 64;  // per rep: rx in [0, 32), ry in [32, 64)

__device__ __forceinline__ uint4 ld4(const uint32_t *p) { return *reinterpret_cast<const uint4 *>(p); }
__device__ __forceinline__ void st4(uint32_t *p, uint4 v) { *reinterpret_cast<uint4 *>(p) = v; }
__device__ __forceinline__ uint4 add4(uint4 a, uint4 b)
{
    return make_uint4(m31_add(a.x, b.x), m31_add(a.y, b.y), m31_add(a.z, b.z), m31_add(a.w, b.w));
}
__device__ __forceinline__ uint4 mul4(uint4 a, uint4 b)
{
    return make_uint4(m31_mul(a.x, b.x), m31_mul(a.y, b.y), m31_mul(a.z, b.z), m31_mul(a.w, b.w));
}
__device__ __forceinline__ uint4 mul4s(uint4 a, uint32_t s)
{
    return make_uint4(m31_mul(a.x, s), m31_mul(a.y, s), m31_mul(a.z, s), m31_mul(a.w, s));
}
__device__ __forceinline__ uint32_t fold1(uint32_t a, uint32_t b, uint32_t r) { return m31_add(a, m31_mul(m31_sub(b, a), r)); }
__device__ __forceinline__ uint4 fold4(uint4 a, uint4 b, uint32_t r)
{
    return make_uint4(fold1(a.x, b.x, r), fold1(a.y, b.y, r), fold1(a.z, b.z, r), fold1(a.w, b.w, r));
}

// ------------------------------------------------------------------ forward pass
template <int L>
__global__ void __launch_bounds__(kThreads) k_forward(const uint32_t *__restrict__ vin, uint32_t *__restrict__ vout,
                                                      const uint32_t *__restrict__ ptr,
                                                      const uint32_t *__restrict__ ent, uint32_t M)
{
    constexpr uint32_t LG = L / 4;
    const uint64_t total = (uint64_t)M * LG;
    for (uint64_t w = (uint64_t)blockIdx.x * kThreads + threadIdx.x; w < total; w += (uint64_t)gridDim.x * kThreads) {
        const uint32_t o = (uint32_t)(w / LG), g = (uint32_t)(w % LG);
        uint4 acc = make_uint4(0, 0, 0, 0);
        const uint32_t e1 = ptr[o + 1];
        for (uint32_t e = ptr[o]; e < e1; e++) {
            const uint32_t a = ent[3 * (uint64_t)e], b = ent[3 * (uint64_t)e + 1], c = ent[3 * (uint64_t)e + 2];
            uint4 va = ld4(vin + (uint64_t)a * L + 4 * g);
            if (b != kAddMark) va = mul4(va, ld4(vin + (uint64_t)b * L + 4 * g));
            acc = add4(acc, mul4s(va, c));
        }
        st4(vout + (uint64_t)o * L + 4 * g, acc);
    }
}

// ------------------------------------------------------------------ eq tables
// out[rep][z] = f1 * eq(r1_rep, z) + f2 * eq(r2_rep, z), z < 2^m, with
// eq(r, z) = prod_i (z_i ? r_i : 1 - r_i) (the reference's _eq at y = 0 / 1).
// f1, f2 come from ab[0], ab[1] when ab is given, else f1 = 1 and the second term is absent.
constexpr int kEqChunk = 16;  // z values per thread
__global__ void __launch_bounds__(kThreads) k_eq(uint32_t *__restrict__ out, uint64_t out_rs,
                                                 const uint32_t *__restrict__ chal, uint32_t r1_off, uint32_t r2_off,
                                                 const uint32_t *__restrict__ ab, uint32_t m)
{
    __shared__ uint32_t lo1[256], lo2[256], hi1[kEqChunk], hi2[kEqChunk];
    const uint32_t rep = blockIdx.y;
    const uint32_t *r1 = chal + (uint64_t)rep * kChalStride + r1_off;
    const uint32_t *r2 = chal + (uint64_t)rep * kChalStride + r2_off;
    const bool two = ab != nullptr;
    const uint32_t f1 = two ? ab[0] : 1u, f2 = two ? ab[1] : 0u;
    const uint32_t h = m < 8 ? m : 8, hb = m - h;
    const uint32_t t = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x * (kEqChunk * 256);
    if (t < (1u << h)) {
        uint32_t a = f1, b = f2;
        for (uint32_t i = 0; i < h; i++) {
            const uint32_t bit = (t >> i) & 1u;
            a = m31_mul(a, bit ? r1[i] : m31_sub(1u, r1[i]));
            if (two) b = m31_mul(b, bit ? r2[i] : m31_sub(1u, r2[i]));
        }
        lo1[t] = a;
        lo2[t] = b;
    }
    if (t < kEqChunk) {
        const uint64_t hi = (base >> h) + t;
        uint32_t a = 1u, b = 1u;
        for (uint32_t i = 0; i < hb; i++) {
            const uint32_t bit = (uint32_t)(hi >> i) & 1u;
            a = m31_mul(a, bit ? r1[h + i] : m31_sub(1u, r1[h + i]));
            if (two) b = m31_mul(b, bit ? r2[h + i] : m31_sub(1u, r2[h + i]));
        }
        hi1[t] = a;
        hi2[t] = b;
    }
    __syncthreads();
    const uint64_t M = 1ull << m;
    const uint32_t mask = (1u << h) - 1u;
    uint32_t *o = out + rep * out_rs;
    for (int c = 0; c < kEqChunk; c++) {
        const uint64_t z = base + (uint64_t)c * 256 + t;
        if (z < M) {
            const uint32_t li = (uint32_t)z & mask;
            const uint32_t hidx = (uint32_t)((z >> h) - (base >> h));
            uint32_t v = m31_mul(lo1[li], hi1[hidx]);
            if (two) v = m31_add(v, m31_mul(lo2[li], hi2[hidx]));
            o[z] = v;
        }
    }
}

// ------------------------------------------------------------------ g_x and h_y
template <int L>
__global__ void __launch_bounds__(kThreads) k_gx(uint32_t *__restrict__ hg, uint64_t hg_rs,
                                                 const uint32_t *__restrict__ v, const uint32_t *__restrict__ eqz,
                                                 uint64_t eqz_rs, const uint32_t *__restrict__ ptr,
                                                 const uint32_t *__restrict__ ent, uint32_t N)
{
    constexpr uint32_t LG = L / 4;
    const uint32_t rep = blockIdx.y;
    const uint32_t *ez = eqz + rep * eqz_rs;
    uint32_t *h = hg + rep * hg_rs;
    const uint64_t total = (uint64_t)N * LG;
    for (uint64_t w = (uint64_t)blockIdx.x * kThreads + threadIdx.x; w < total; w += (uint64_t)gridDim.x * kThreads) {
        const uint32_t x = (uint32_t)(w / LG), g = (uint32_t)(w % LG);
        uint4 acc = make_uint4(0, 0, 0, 0);
        const uint32_t e1 = ptr[x + 1];
        for (uint32_t e = ptr[x]; e < e1; e++) {
            const uint32_t b = ent[3 * (uint64_t)e], z = ent[3 * (uint64_t)e + 1], c = ent[3 * (uint64_t)e + 2];
            const uint32_t ce = m31_mul(c, ez[z]);
            if (b != kAddMark) {
                acc = add4(acc, mul4s(ld4(v + (uint64_t)b * L + 4 * g), ce));
            } else {
                acc = add4(acc, make_uint4(ce, ce, ce, ce));
            }
        }
        st4(h + (uint64_t)x * L + 4 * g, acc);
    }
}

__global__ void __launch_bounds__(kThreads) k_hy(uint32_t *__restrict__ s, uint64_t s_rs,
                                                 const uint32_t *__restrict__ eqz, uint64_t eqz_rs,
                                                 const uint32_t *__restrict__ eqx, uint64_t eqx_rs,
                                                 const uint32_t *__restrict__ ptr, const uint32_t *__restrict__ ent,
                                                 uint32_t N)
{
    const uint32_t rep = blockIdx.y;
    const uint32_t *ez = eqz + rep * eqz_rs;
    const uint32_t *ex = eqx + rep * eqx_rs;
    uint32_t *out = s + rep * s_rs;
    for (uint64_t y = (uint64_t)blockIdx.x * kThreads + threadIdx.x; y < N; y += (uint64_t)gridDim.x * kThreads) {
        uint32_t acc = 0;
        const uint32_t e1 = ptr[y + 1];
        for (uint32_t e = ptr[y]; e < e1; e++) {
            const uint32_t x = ent[3 * (uint64_t)e], z = ent[3 * (uint64_t)e + 1], c = ent[3 * (uint64_t)e + 2];
            acc = m31_add(acc, m31_mul(m31_mul(ez[z], ex[x]), c));
        }
        out[y] = acc;
    }
}

// ------------------------------------------------------------------ reductions
// Reduce 3 x 4 lane sums (u64) across the block.  Threads sharing t % LG hold the same
// lanes.  Result: partial[q] for q in [0, 3L) with q = pk * L + 4 * g + c.
template <int L>
__device__ __forceinline__ void block_reduce_store(uint64_t (&acc)[3][4], uint32_t *__restrict__ dst)
{
    constexpr uint32_t LG = L / 4;
    constexpr int kWaves = kThreads / 64;
    __shared__ uint32_t red[kWaves][3 * L];
    uint32_t v[3][4];
#pragma unroll
    for (int pk = 0; pk < 3; pk++)
#pragma unroll
        for (int c = 0; c < 4; c++) v[pk][c] = m31_reduce64(acc[pk][c]);
#pragma unroll
    for (uint32_t off = LG; off < 64; off <<= 1) {
#pragma unroll
        for (int pk = 0; pk < 3; pk++)
#pragma unroll
            for (int c = 0; c < 4; c++) v[pk][c] = m31_add(v[pk][c], __shfl_xor(v[pk][c], off, 64));
    }
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane < LG) {
#pragma unroll
        for (int pk = 0; pk < 3; pk++)
#pragma unroll
            for (int c = 0; c < 4; c++) red[wave][pk * L + 4 * lane + c] = v[pk][c];
    }
    __syncthreads();
    if (threadIdx.x < 3 * L) {
        uint32_t s = 0;
#pragma unroll
        for (int w = 0; w < kWaves; w++) s = m31_add(s, red[w][threadIdx.x]);
        dst[threadIdx.x] = s;
    }
}

// ------------------------------------------------------------------ sumcheck round
struct RoundParams {
    const uint32_t *fsrc;  // f_{k-1} (or f_0 when k == 0)
    uint64_t fsrc_rs;      // rep stride in u32 (0 when the source is the shared layer input)
    uint32_t *fdst;        // f_k (unused when k == 0)
    uint64_t fdst_rs;
    const uint32_t *hsrc;  // phase 1: lane table hg; phase 2: scalar table s
    uint64_t hsrc_rs;
    uint32_t *hdst;
    uint64_t hdst_rs;
    const uint32_t *rprev; // challenge of round k-1 for rep 0; rep stride kChalStride
    uint32_t *partials;    // [blockIdx.x][rep][3L]
    uint32_t items;        // pairs in round k = N_k / 2
    uint32_t R;
};

// Round k of one sumcheck phase for every repetition (blockIdx.y = rep).
// FUSED: fold round k-1 with its challenge, store f_k / h_k, and accumulate
//   p0 = sum f_k[2i] h_k[2i], p1 = sum f_k[2i+1] h_k[2i+1], p2 = sum (f0 + f1)(h0 + h1).
// PH2: h is the scalar table s (phase-2 hg = vx * s is rank one; vx multiplies later).
template <int L, bool PH2, bool FUSED>
__global__ void __launch_bounds__(kThreads) k_round(RoundParams p)
{
    constexpr uint32_t LG = L / 4;
    const uint32_t rep = blockIdx.y;
    const uint32_t g = threadIdx.x % LG;
    const uint32_t *fs = p.fsrc + rep * p.fsrc_rs;
    uint32_t *fd = p.fdst + rep * p.fdst_rs;
    const uint32_t *hs = p.hsrc + rep * p.hsrc_rs;
    uint32_t *hd = p.hdst + rep * p.hdst_rs;
    const uint32_t r = FUSED ? p.rprev[(uint64_t)rep * kChalStride] : 0u;
    uint64_t acc[3][4];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int c = 0; c < 4; c++) acc[a][c] = 0;
    const uint64_t total = (uint64_t)p.items * LG;
    for (uint64_t w = (uint64_t)blockIdx.x * kThreads + threadIdx.x; w < total; w += (uint64_t)gridDim.x * kThreads) {
        const uint64_t i = w / LG;
        uint4 f0, f1, h0, h1;
        uint32_t s0 = 0, s1 = 0;
        if (FUSED) {
            const uint32_t *src = fs + 4 * i * L + 4 * g;
            const uint4 a0 = ld4(src), a1 = ld4(src + L), a2 = ld4(src + 2 * L), a3 = ld4(src + 3 * L);
            f0 = fold4(a0, a1, r);
            f1 = fold4(a2, a3, r);
            st4(fd + 2 * i * L + 4 * g, f0);
            st4(fd + (2 * i + 1) * L + 4 * g, f1);
            if (PH2) {
                const uint32_t b0 = hs[4 * i], b1 = hs[4 * i + 1], b2 = hs[4 * i + 2], b3 = hs[4 * i + 3];
                s0 = fold1(b0, b1, r);
                s1 = fold1(b2, b3, r);
                if (g == 0) {
                    hd[2 * i] = s0;
                    hd[2 * i + 1] = s1;
                }
            } else {
                const uint32_t *hsrc = hs + 4 * i * L + 4 * g;
                const uint4 c0 = ld4(hsrc), c1 = ld4(hsrc + L), c2 = ld4(hsrc + 2 * L), c3 = ld4(hsrc + 3 * L);
                h0 = fold4(c0, c1, r);
                h1 = fold4(c2, c3, r);
                st4(hd + 2 * i * L + 4 * g, h0);
                st4(hd + (2 * i + 1) * L + 4 * g, h1);
            }
        } else {
            f0 = ld4(fs + 2 * i * L + 4 * g);
            f1 = ld4(fs + (2 * i + 1) * L + 4 * g);
            if (PH2) {
                s0 = hs[2 * i];
                s1 = hs[2 * i + 1];
            } else {
                h0 = ld4(hs + 2 * i * L + 4 * g);
                h1 = ld4(hs + (2 * i + 1) * L + 4 * g);
            }
        }
        if (PH2) {
            h0 = make_uint4(s0, s0, s0, s0);
            h1 = make_uint4(s1, s1, s1, s1);
        }
        const uint4 e0 = mul4(f0, h0), e1 = mul4(f1, h1), e2 = mul4(add4(f0, f1), add4(h0, h1));
        acc[0][0] += e0.x; acc[0][1] += e0.y; acc[0][2] += e0.z; acc[0][3] += e0.w;
        acc[1][0] += e1.x; acc[1][1] += e1.y; acc[1][2] += e1.z; acc[1][3] += e1.w;
        acc[2][0] += e2.x; acc[2][1] += e2.y; acc[2][2] += e2.z; acc[2][3] += e2.w;
    }
    block_reduce_store<L>(acc, p.partials + ((uint64_t)blockIdx.x * p.R + rep) * 3 * L);
}

// Dot product sum_z eq[rep][z] * vals[z] (lanes), for claimed_v.
template <int L>
__global__ void __launch_bounds__(kThreads) k_dot(const uint32_t *__restrict__ vals, const uint32_t *__restrict__ eq,
                                                  uint64_t eq_rs, uint32_t M, uint32_t *__restrict__ partials,
                                                  uint32_t R)
{
    constexpr uint32_t LG = L / 4;
    const uint32_t rep = blockIdx.y;
    const uint32_t g = threadIdx.x % LG;
    const uint32_t *e = eq + rep * eq_rs;
    uint64_t acc[3][4];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int c = 0; c < 4; c++) acc[a][c] = 0;
    const uint64_t total = (uint64_t)M * LG;
    for (uint64_t w = (uint64_t)blockIdx.x * kThreads + threadIdx.x; w < total; w += (uint64_t)gridDim.x * kThreads) {
        const uint64_t z = w / LG;
        const uint4 t = mul4s(ld4(vals + z * L + 4 * g), e[z]);
        acc[0][0] += t.x; acc[0][1] += t.y; acc[0][2] += t.z; acc[0][3] += t.w;
    }
    block_reduce_store<L>(acc, partials + ((uint64_t)blockIdx.x * R + rep) * 3 * L);
}

// ------------------------------------------------------------------ transcript
// Device transcript (include/fiat_shamir/transcript.hpp:87-175) over the proof bytes
// that follow the commitment.  state: [0] = len (bytes), [1] = hash_start, [2..9] digest.
struct DevTranscript {
    uint32_t len, hash_start;
    uint32_t digest[8];
    uint32_t *tail;
    int kind;

    __device__ void load(const uint32_t *st, uint32_t *t, int k)
    {
        len = st[0];
        hash_start = st[1];
        for (int i = 0; i < 8; i++) digest[i] = st[2 + i];
        tail = t;
        kind = k;
    }
    __device__ void store(uint32_t *st) const
    {
        st[0] = len;
        st[1] = hash_start;
        for (int i = 0; i < 8; i++) st[2 + i] = digest[i];
    }
    __device__ void append_words(const uint32_t *w, uint32_t n)
    {
        for (uint32_t i = 0; i < n; i++) tail[len / 4 + i] = w[i];
        len += 4 * n;
    }
    // challenge_f: hash the bytes appended since the last challenge (or the digest),
    // then M31::from_bytes of the first four digest bytes.
    __device__ uint32_t challenge()
    {
        uint32_t d[8];
        if (len > hash_start) {
            hash_words(kind, tail + hash_start / 4, (len - hash_start) / 4, d);
            hash_start = len;
        } else {
            hash_words(kind, digest, 8, d);
        }
        for (int i = 0; i < 8; i++) digest[i] = d[i];
        return m31_from_u32_full(digest[0]);
    }
};

struct FinishParams {
    const uint32_t *partials;
    uint32_t nbx, R;
    int mode;              // 0 = sumcheck round, 1 = claims (sums only)
    int phase;             // 1 or 2
    int last;              // last round of the phase
    int layer_end;         // last round of phase 2: append vy claims, draw alpha and beta
    uint32_t k;            // round index in the phase
    uint32_t *chal;        // this layer's challenge slots, rep stride kChalStride
    uint32_t *vx;          // R * L
    uint32_t *vy;          // R * L
    const uint32_t *fin;   // f table of this round (2 rows per rep at the last round)
    uint64_t fin_rs;
    uint32_t *tstate;
    uint32_t *tail;
    uint32_t *ab_out;      // alpha, beta drawn after this layer
    uint32_t *claims;      // mode 1 output, R * L
    int hash_kind;
};

template <int L>
__global__ void __launch_bounds__(kFinishThreads) k_finish(FinishParams p)
{
    constexpr uint32_t Q = 3 * L;
    __shared__ uint32_t part[kFinishThreads];
    __shared__ uint32_t sums[3 * Q];  // R <= 3
    const uint32_t items = p.R * Q;
    const uint32_t nparts = kFinishThreads / items;
    const uint32_t t = threadIdx.x;
    if (t < nparts * items) {
        const uint32_t item = t % items, pt = t / items;
        uint64_t s = 0;
        for (uint32_t b = pt; b < p.nbx; b += nparts) s += p.partials[(uint64_t)b * items + item];
        part[t] = m31_reduce64(s);
    }
    __syncthreads();
    if (t < items) {
        uint32_t s = 0;
        for (uint32_t pt = 0; pt < nparts; pt++) s = m31_add(s, part[pt * items + t]);
        sums[t] = s;
    }
    __syncthreads();
    if (p.mode == 1) {
        // claimed_v[rep] = first L sums of the rep
        if (t < p.R * L) p.claims[t] = sums[(t / L) * Q + (t % L)];
        return;
    }
    if (t != 0) return;
    DevTranscript tr;
    tr.load(p.tstate, p.tail, p.hash_kind);
    uint32_t msg[3 * L];
    for (uint32_t j = 0; j < p.R; j++) {
        const uint32_t *sj = sums + j * Q;
        for (uint32_t c = 0; c < L; c++) {
            uint32_t q0 = sj[c], q1 = sj[L + c], q2 = sj[2 * L + c];
            if (p.phase == 2) {
                const uint32_t vx = p.vx[j * L + c];
                q0 = m31_mul(q0, vx);
                q1 = m31_mul(q1, vx);
                q2 = m31_mul(q2, vx);
            }
            // p2 = p1 * 6 + p0 * 3 - p2 * 2 (sumcheck_helper.hpp:100)
            msg[c] = q0;
            msg[L + c] = q1;
            msg[2 * L + c] = m31_sub(m31_add(m31_mul(q1, 6u), m31_mul(q0, 3u)), m31_mul(q2, 2u));
        }
        tr.append_words(msg, 3 * L);
        const uint32_t r = tr.challenge();
        p.chal[j * kChalStride + (p.phase == 2 ? 32u : 0u) + p.k] = r;
        if (p.last) {
            const uint32_t *f = p.fin + j * p.fin_rs;
            uint32_t *dst = (p.phase == 1 ? p.vx : p.vy) + j * L;
            for (uint32_t c = 0; c < L; c++) dst[c] = fold1(f[c], f[L + c], r);
            if (p.phase == 1) tr.append_words(dst, L);  // vx_claim right after rep j's last x round
        }
    }
    if (p.layer_end) {
        for (uint32_t j = 0; j < p.R; j++) tr.append_words(p.vy + j * L, L);
        p.ab_out[0] = tr.challenge();
        p.ab_out[1] = tr.challenge();
    }
    tr.store(p.tstate);
}

template <int L>
int forward_impl(const DevCircuit &dc, hipStream_t st)
{
    for (uint32_t l = 0; l < dc.D; l++) {
        const LayerDev &ly = dc.layers[l];
        const uint64_t M = 1ull << ly.m;
        const uint64_t work = M * (L / 4);
        const uint32_t blocks = (uint32_t)std::min<uint64_t>((work + kThreads - 1) / kThreads, 8192);
        hipLaunchKernelGGL(k_forward<L>, dim3(blocks), dim3(kThreads), 0, st, dc.values + ly.v_off,
                           dc.values + ly.out_off, ly.fw_ptr, ly.fw_ent, (uint32_t)M);
    }
    GKRX_HIP_CHECK(hipGetLastError());
    return GKRX_OK;
}

uint32_t grid_for(uint64_t work)
{
    uint64_t b = (work + kThreads - 1) / kThreads;
    if (b < 1) b = 1;
    return (uint32_t)std::min<uint64_t>(b, kMaxBlocksPerRep);
}

template <int L>
int prove_impl(const DevCircuit &dc, DevScratch &s, uint32_t R, int hash_kind, hipStream_t st)
{
    constexpr uint32_t LG = L / 4;
    const uint32_t D = dc.D;
    const uint64_t maxN = 1ull << dc.max_n, maxM = 1ull << dc.max_m;
    const uint64_t eqz_rs = maxM, eqx_rs = maxN;
    const uint64_t hg0_rs = maxN * L, hgA_rs = (maxN / 2) * L, hgB_rs = std::max<uint64_t>(maxN / 4, 1) * L;
    const uint64_t fA_rs = (maxN / 2) * L, fB_rs = std::max<uint64_t>(maxN / 4, 1) * L;
    const uint64_t s0_rs = maxN, sA_rs = maxN / 2, sB_rs = std::max<uint64_t>(maxN / 4, 1);
    auto chal_slot = [&](uint32_t l) { return s.chal + (uint64_t)l * R * kChalStride; };
    auto ab_slot = [&](uint32_t l) { return s.ab + 2ull * l; };

    for (int l = (int)D - 1; l >= 0; l--) {
        const LayerDev &ly = dc.layers[l];
        const uint32_t n = ly.n, m = ly.m;
        const uint64_t N = 1ull << n, M = 1ull << m;
        const uint32_t *V = dc.values + ly.v_off;
        uint32_t *chal_l = chal_slot(l);
        const uint32_t *chal_prev = chal_slot(l + 1);
        // eqZ = alpha * eq(rz1, .) + beta * eq(rz2, .), rz1 = rx and rz2 = ry of the layer above
        {
            dim3 grid((uint32_t)((M + kEqChunk * 256 - 1) / (kEqChunk * 256)), R);
            hipLaunchKernelGGL(k_eq, grid, dim3(kThreads), 0, st, s.eqz, eqz_rs, chal_prev, 0u, 32u,
                               (const uint32_t *)ab_slot(l + 1), m);
        }
        if (l == (int)D - 1) {
            // claimed_v[j] = eval_multilinear(output, rz1[j]) = sum_z eq(rz1[j], z) out[z]
            const uint32_t nbx = grid_for(M * LG);
            hipLaunchKernelGGL(k_dot<L>, dim3(nbx, R), dim3(kThreads), 0, st, dc.values + ly.out_off, s.eqz, eqz_rs,
                               (uint32_t)M, s.partials, R);
            FinishParams fp{};
            fp.partials = s.partials; fp.nbx = nbx; fp.R = R; fp.mode = 1; fp.claims = s.claims;
            hipLaunchKernelGGL(k_finish<L>, dim3(1), dim3(kFinishThreads), 0, st, fp);
        }
        {
            dim3 grid(grid_for(N * LG) * 4, R);
            hipLaunchKernelGGL(k_gx<L>, grid, dim3(kThreads), 0, st, s.hg0, hg0_rs, V, s.eqz, eqz_rs, ly.gx_ptr,
                               ly.gx_ent, (uint32_t)N);
        }
        for (int phase = 1; phase <= 2; phase++) {
            if (phase == 2) {
                dim3 grid((uint32_t)((N + kEqChunk * 256 - 1) / (kEqChunk * 256)), R);
                hipLaunchKernelGGL(k_eq, grid, dim3(kThreads), 0, st, s.eqx, eqx_rs, (const uint32_t *)chal_l, 0u,
                                   0u, (const uint32_t *)nullptr, n);
                dim3 g2(grid_for(N) * 4, R);
                hipLaunchKernelGGL(k_hy, g2, dim3(kThreads), 0, st, s.s0, s0_rs, s.eqz, eqz_rs, s.eqx, eqx_rs,
                                   ly.hy_ptr, ly.hy_ent, (uint32_t)N);
            }
            for (uint32_t k = 0; k < n; k++) {
                RoundParams rp{};
                rp.items = (uint32_t)(N >> (k + 1));
                rp.R = R;
                rp.partials = s.partials;
                // f source / destination (ping-pong A: N/2 rows, B: N/4 rows)
                const uint32_t *fcur;
                uint64_t fcur_rs;
                if (k == 0) {
                    rp.fsrc = V; rp.fsrc_rs = 0;
                    fcur = V; fcur_rs = 0;
                } else {
                    rp.fsrc = (k == 1) ? V : ((k - 1) & 1 ? s.fA : s.fB);
                    rp.fsrc_rs = (k == 1) ? 0 : ((k - 1) & 1 ? fA_rs : fB_rs);
                    rp.fdst = (k & 1) ? s.fA : s.fB;
                    rp.fdst_rs = (k & 1) ? fA_rs : fB_rs;
                    fcur = rp.fdst; fcur_rs = rp.fdst_rs;
                    rp.rprev = chal_l + (phase == 2 ? 32 : 0) + (k - 1);
                }
                if (phase == 1) {
                    if (k == 0) { rp.hsrc = s.hg0; rp.hsrc_rs = hg0_rs; }
                    else {
                        rp.hsrc = (k == 1) ? s.hg0 : ((k - 1) & 1 ? s.hgA : s.hgB);
                        rp.hsrc_rs = (k == 1) ? hg0_rs : ((k - 1) & 1 ? hgA_rs : hgB_rs);
                        rp.hdst = (k & 1) ? s.hgA : s.hgB;
                        rp.hdst_rs = (k & 1) ? hgA_rs : hgB_rs;
                    }
                } else {
                    if (k == 0) { rp.hsrc = s.s0; rp.hsrc_rs = s0_rs; }
                    else {
                        rp.hsrc = (k == 1) ? s.s0 : ((k - 1) & 1 ? s.sA : s.sB);
                        rp.hsrc_rs = (k == 1) ? s0_rs : ((k - 1) & 1 ? sA_rs : sB_rs);
                        rp.hdst = (k & 1) ? s.sA : s.sB;
                        rp.hdst_rs = (k & 1) ? sA_rs : sB_rs;
                    }
                }
                const uint32_t nbx = grid_for((uint64_t)rp.items * LG);
                dim3 grid(nbx, R);
                if (phase == 1) {
                    if (k == 0) hipLaunchKernelGGL((k_round<L, false, false>), grid, dim3(kThreads), 0, st, rp);
                    else hipLaunchKernelGGL((k_round<L, false, true>), grid, dim3(kThreads), 0, st, rp);
                } else {
                    if (k == 0) hipLaunchKernelGGL((k_round<L, true, false>), grid, dim3(kThreads), 0, st, rp);
                    else hipLaunchKernelGGL((k_round<L, true, true>), grid, dim3(kThreads), 0, st, rp);
                }
                FinishParams fp{};
                fp.partials = s.partials; fp.nbx = nbx; fp.R = R; fp.mode = 0;
                fp.phase = phase; fp.last = (k == n - 1); fp.layer_end = (phase == 2 && k == n - 1);
                fp.k = k; fp.chal = chal_l; fp.vx = s.vxy; fp.vy = s.vxy + R * L;
                fp.fin = fcur; fp.fin_rs = fcur_rs;
                fp.tstate = s.tstate; fp.tail = s.tail; fp.ab_out = ab_slot(l); fp.hash_kind = hash_kind;
                hipLaunchKernelGGL(k_finish<L>, dim3(1), dim3(kFinishThreads), 0, st, fp);
            }
        }
    }
    GKRX_HIP_CHECK(hipGetLastError());
    return GKRX_OK;
}

}  // namespace

// ------------------------------------------------------------------ host side
int dev_forward(const DevCircuit &dc, hipStream_t st)
{
    if (dc.L == 8) return forward_impl<8>(dc, st);
    if (dc.L == 16) return forward_impl<16>(dc, st);
    return GKRX_ERR_CONFIG;
}

int dev_prove(const DevCircuit &dc, DevScratch &s, uint32_t R, int hash_kind, const uint32_t *rz1_host,
              const uint32_t *tstate_host, const uint8_t *tail_prefix, uint32_t tail_prefix_len, hipStream_t st)
{
    if (R > 3) return GKRX_ERR_CONFIG;
    const uint32_t D = dc.D;
    const uint32_t mlast = dc.layers[D - 1].m;
    // seed slot D: rx = rz1, ry = 0 (rz2), alpha = 1, beta = 0
    std::vector<uint32_t> seed((size_t)R * kChalStride, 0u);
    for (uint32_t j = 0; j < R; j++)
        for (uint32_t i = 0; i < mlast; i++) seed[(size_t)j * kChalStride + i] = rz1_host[(size_t)j * mlast + i];
    GKRX_HIP_CHECK(hipMemcpyAsync(s.chal + (uint64_t)D * R * kChalStride, seed.data(), seed.size() * 4,
                                  hipMemcpyHostToDevice, st));
    const uint32_t ab_seed[2] = {1u, 0u};
    GKRX_HIP_CHECK(hipMemcpyAsync(s.ab + 2ull * D, ab_seed, 8, hipMemcpyHostToDevice, st));
    GKRX_HIP_CHECK(hipMemcpyAsync(s.tstate, tstate_host, 10 * 4, hipMemcpyHostToDevice, st));
    if (tail_prefix_len)
        GKRX_HIP_CHECK(hipMemcpyAsync(s.tail, tail_prefix, tail_prefix_len, hipMemcpyHostToDevice, st));
    if (dc.L == 8) return prove_impl<8>(dc, s, R, hash_kind, st);
    if (dc.L == 16) return prove_impl<16>(dc, s, R, hash_kind, st);
    return GKRX_ERR_CONFIG;
}

// CSR for one key: counting sort of entries by key.
static void build_csr(uint32_t nrows, const std::vector<uint32_t> &keys, const std::vector<uint32_t> &ent3,
                      std::vector<uint32_t> &ptr, std::vector<uint32_t> &out)
{
    ptr.assign((size_t)nrows + 1, 0);
    for (uint32_t k : keys) ptr[k + 1]++;
    for (uint32_t i = 0; i < nrows; i++) ptr[i + 1] += ptr[i];
    out.resize(ent3.size());
    std::vector<uint32_t> pos(ptr.begin(), ptr.end() - 1);
    for (size_t e = 0; e < keys.size(); e++) {
        uint32_t at = pos[keys[e]]++;
        out[3 * (size_t)at] = ent3[3 * e];
        out[3 * (size_t)at + 1] = ent3[3 * e + 1];
        out[3 * (size_t)at + 2] = ent3[3 * e + 2];
    }
}

int dev_upload_circuit(DevCircuit &dc, uint32_t L, const std::vector<uint32_t> &n, const std::vector<uint32_t> &m,
                       const std::vector<std::vector<uint32_t>> &mul, const std::vector<std::vector<uint32_t>> &add)
{
    dc.D = (uint32_t)n.size();
    dc.L = L;
    dc.layers.assign(dc.D, LayerDev{});
    dc.max_n = dc.max_m = 0;
    // host CSR images, then one device pool
    std::vector<std::vector<uint32_t>> host(6 * dc.D);
    uint64_t off = 0;
    for (uint32_t l = 0; l < dc.D; l++) {
        LayerDev &ly = dc.layers[l];
        ly.n = n[l];
        ly.m = m[l];
        dc.max_n = std::max(dc.max_n, ly.n);
        dc.max_m = std::max(dc.max_m, ly.m);
        ly.v_off = off;
        off += (1ull << ly.n) * L;
        const size_t U = mul[l].size() / 4, A = add[l].size() / 3;
        std::vector<uint32_t> kf, ef, kg, eg, kh, eh;
        kf.reserve(U + A); ef.reserve(3 * (U + A));
        kg.reserve(U + A); eg.reserve(3 * (U + A));
        kh.reserve(U); eh.reserve(3 * U);
        for (size_t g = 0; g < U; g++) {
            const uint32_t *q = &mul[l][4 * g];
            kf.push_back(q[2]); ef.insert(ef.end(), {q[0], q[1], q[3]});
            kg.push_back(q[0]); eg.insert(eg.end(), {q[1], q[2], q[3]});
            kh.push_back(q[1]); eh.insert(eh.end(), {q[0], q[2], q[3]});
        }
        for (size_t g = 0; g < A; g++) {
            const uint32_t *q = &add[l][3 * g];
            kf.push_back(q[1]); ef.insert(ef.end(), {q[0], kAddMark, q[2]});
            kg.push_back(q[0]); eg.insert(eg.end(), {kAddMark, q[1], q[2]});
        }
        build_csr(1u << ly.m, kf, ef, host[6 * l + 0], host[6 * l + 1]);
        build_csr(1u << ly.n, kg, eg, host[6 * l + 2], host[6 * l + 3]);
        build_csr(1u << ly.n, kh, eh, host[6 * l + 4], host[6 * l + 5]);
    }
    dc.layers[dc.D - 1].out_off = off;
    for (uint32_t l = 0; l + 1 < dc.D; l++) dc.layers[l].out_off = dc.layers[l + 1].v_off;
    off += (1ull << dc.layers[dc.D - 1].m) * L;
    dc.values_u32 = off;
    uint64_t pool = 0;
    for (auto &v : host) pool += (v.size() + 3) & ~3ull;  // 16-byte aligned pieces
    GKRX_HIP_CHECK(hipMalloc(&dc.csr_pool, std::max<uint64_t>(pool, 4) * 4));
    GKRX_HIP_CHECK(hipMalloc(&dc.values, dc.values_u32 * 4));
    uint64_t at = 0;
    for (uint32_t l = 0; l < dc.D; l++) {
        uint32_t **dst[6] = {&dc.layers[l].fw_ptr, &dc.layers[l].fw_ent, &dc.layers[l].gx_ptr,
                             &dc.layers[l].gx_ent, &dc.layers[l].hy_ptr, &dc.layers[l].hy_ent};
        for (int k = 0; k < 6; k++) {
            auto &v = host[6 * l + k];
            *dst[k] = dc.csr_pool + at;
            if (!v.empty()) GKRX_HIP_CHECK(hipMemcpy(dc.csr_pool + at, v.data(), v.size() * 4, hipMemcpyHostToDevice));
            at += (v.size() + 3) & ~3ull;
        }
    }
    return GKRX_OK;
}

void dev_free_circuit(DevCircuit &dc)
{
    if (dc.csr_pool) (void)hipFree(dc.csr_pool);
    if (dc.values) (void)hipFree(dc.values);
    dc.csr_pool = nullptr;
    dc.values = nullptr;
}

int dev_ensure_scratch(DevScratch &s, const DevCircuit &dc, uint32_t R, uint64_t tail_bytes)
{
    if (s.pool && s.max_n >= dc.max_n && s.max_m >= dc.max_m && s.L == dc.L && s.D >= dc.D &&
        s.tail_cap >= tail_bytes)
        return GKRX_OK;
    dev_free_scratch(s);
    const uint32_t L = dc.L;
    const uint64_t maxN = 1ull << dc.max_n, maxM = 1ull << dc.max_m;
    const uint64_t N2 = std::max<uint64_t>(maxN / 2, 1), N4 = std::max<uint64_t>(maxN / 4, 1);
    struct Piece { uint32_t **p; uint64_t u32; };
    Piece pieces[] = {
        {&s.eqz, R * maxM}, {&s.eqx, R * maxN}, {&s.hg0, R * maxN * L}, {&s.hgA, R * N2 * L},
        {&s.hgB, R * N4 * L}, {&s.fA, R * N2 * L}, {&s.fB, R * N4 * L}, {&s.s0, R * maxN},
        {&s.sA, R * N2}, {&s.sB, R * N4}, {&s.partials, (uint64_t)R * kMaxBlocksPerRep * 3 * L},
        {&s.chal, (uint64_t)(dc.D + 1) * R * kChalStride}, {&s.ab, 2ull * (dc.D + 1)}, {&s.claims, R * L},
        {&s.vxy, 2ull * R * L}, {&s.tstate, 16}, {&s.tail, (tail_bytes + 3) / 4 + 64}};
    uint64_t total = 0;
    for (auto &pc : pieces) total += (pc.u32 + 63) & ~63ull;
    GKRX_HIP_CHECK(hipMalloc(&s.pool, total * 4));
    uint64_t at = 0;
    for (auto &pc : pieces) {
        *pc.p = reinterpret_cast<uint32_t *>(s.pool) + at;
        at += (pc.u32 + 63) & ~63ull;
    }
    s.bytes = total * 4;
    s.max_n = dc.max_n;
    s.max_m = dc.max_m;
    s.L = L;
    s.D = dc.D;
    s.tail_cap = tail_bytes;
    return GKRX_OK;
}

void dev_free_scratch(DevScratch &s)
{
    if (s.pool) (void)hipFree(s.pool);
    s = DevScratch{};
}

}  // namespace gkrx
