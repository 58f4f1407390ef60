// MI355X (gfx950) kernels for the GKR prove path over M31 with L data-parallel lanes.
//
// Reference loops replaced (PolyhedraZK/Expander-cpp):
//   k_forward  <- CircuitLayer::evaluate              include/circuit/circuit.hpp:92-105
//   k_eq       <- _eq_evals_at / _primitive + combine include/LinearGKR/sumcheck_common.hpp:16-53,
//                                                     include/LinearGKR/sumcheck_helper.hpp:188-196
//   k_gx       <- _prepare_g_x_vals scatters          sumcheck_helper.hpp:198-229
//   k_hy       <- _prepare_h_y_vals scatter           sumcheck_helper.hpp:233-261
//   k_pair     <- receive_challenge (folds of rounds k-2, k-1) fused with poly_eval_at
//                 (rounds k and k+1, the latter as a quadratic in r_k)
//                                                     sumcheck_helper.hpp:68-133, on the whole GPU
//   k_tail_rep, k_tail <- the small last rounds of a phase one at a time, one workgroup
//                 per repetition (each round's pass behind the previous round's hash),
//                 then the last round of the phase in one workgroup
//   k_ls_*     <- all of the above for a batch of independent proofs, one step of every
//                 proof per launch (no reference counterpart: the reference proves one
//                 proof per thread, bench_keccak.cpp:32-42)
//   transcript_step <- the transcript steps of sumcheck_prove_gkr_layer / gkr_prove
//                                                     include/LinearGKR/sumcheck.hpp:72-103,
//                                                     include/LinearGKR/gkr.hpp:58-67
//   k_dot      <- eval_multilinear for claimed_v      include/poly_commit/poly.hpp:88-104
//
// Layout: every table keeps the reference's natural index order, element-major and
// lane-minor (one table entry = L consecutive u32 lanes = the bytes F::to_bytes writes).
// A thread owns V = 1, 2 or 4 lanes of one entry (4-, 8- or 16-byte loads); L/V threads
// cover one entry.
// Scatters of the reference become gathers over CSR lists built at circuit load, so no
// modular atomics are needed and every sum is exact field arithmetic (order-free).
// Folding is ping-pong between buffers (the reference folds in place, which is only
// safe sequentially).  The gate_exists skip of the reference is exact (hg is zero where
// no gate exists) and is therefore not needed.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <string>

#include "hash_dev.hpp"
#include "internal.hpp"
#include "m31.hpp"
#include "xcd.hpp"

namespace gkrx {

namespace {

constexpr int kThreads = 256;
constexpr uint32_t kMaxBlocksPerRep = 256;
constexpr uint32_t kChalStride = 64;  // per rep: rx in [0, 32), ry in [32, 64)

// The transcript kernels are chains of dependent single-wave hashes: on a CU they share
// with bandwidth kernels of other streams (batches, server slots), their waves take the
// SIMDs' issue slots first.
__device__ __forceinline__ void latency_critical() { __builtin_amdgcn_s_setprio(3); }

__device__ __forceinline__ uint4 ld4(const uint32_t *p) { return *reinterpret_cast<const uint4 *>(p); }
__device__ __forceinline__ void st4(uint32_t *p, uint4 v) { *reinterpret_cast<uint4 *>(p) = v; }
// One entry of an interleaved eq table (kEqWords words: rep 0, 1, 2), moved as one 12-byte
// access (global_load_dwordx3 / global_store_dwordx3: 4-byte alignment is enough).
constexpr uint32_t kEqWords = 3;
struct Eq3 {
    uint32_t r0, r1, r2;
};
__device__ __forceinline__ Eq3 ld3(const uint32_t *p)
{
    Eq3 v;
    __builtin_memcpy(&v, p, sizeof v);
    return v;
}
__device__ __forceinline__ void st3(uint32_t *p, const Eq3 &v) { __builtin_memcpy(p, &v, sizeof v); }
__device__ __forceinline__ uint4 add4(uint4 a, uint4 b)
{
    return make_uint4(m31_add(a.x, b.x), m31_add(a.y, b.y), m31_add(a.z, b.z), m31_add(a.w, b.w));
}
__device__ __forceinline__ uint4 mul4(uint4 a, uint4 b)
{
    return make_uint4(m31_mul(a.x, b.x), m31_mul(a.y, b.y), m31_mul(a.z, b.z), m31_mul(a.w, b.w));
}
__device__ __forceinline__ uint4 sub4(uint4 a, uint4 b)
{
    return make_uint4(m31_sub(a.x, b.x), m31_sub(a.y, b.y), m31_sub(a.z, b.z), m31_sub(a.w, b.w));
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

// V consecutive lanes of one table entry (V = 1, 2 or 4 lanes per thread: one 4-, 8- or
// 16-byte access).
template <int V>
struct Lanes {
    uint32_t v[V];
};
template <int V>
__device__ __forceinline__ Lanes<V> ldv(const uint32_t *p)
{
    Lanes<V> r;
    if constexpr (V == 4) {
        const uint4 x = ld4(p);
        r.v[0] = x.x; r.v[1] = x.y; r.v[2] = x.z; r.v[3] = x.w;
    } else if constexpr (V == 2) {
        const uint2 x = *reinterpret_cast<const uint2 *>(p);
        r.v[0] = x.x; r.v[1] = x.y;
    } else {
        r.v[0] = *p;
    }
    return r;
}
template <int V>
__device__ __forceinline__ void stv(uint32_t *p, const Lanes<V> &a)
{
    if constexpr (V == 4) st4(p, make_uint4(a.v[0], a.v[1], a.v[2], a.v[3]));
    else if constexpr (V == 2) *reinterpret_cast<uint2 *>(p) = make_uint2(a.v[0], a.v[1]);
    else *p = a.v[0];
}
template <int V>
__device__ __forceinline__ Lanes<V> addv(const Lanes<V> &a, const Lanes<V> &b)
{
    Lanes<V> r;
#pragma unroll
    for (int c = 0; c < V; c++) r.v[c] = m31_add(a.v[c], b.v[c]);
    return r;
}
template <int V>
__device__ __forceinline__ Lanes<V> subv(const Lanes<V> &a, const Lanes<V> &b)
{
    Lanes<V> r;
#pragma unroll
    for (int c = 0; c < V; c++) r.v[c] = m31_sub(a.v[c], b.v[c]);
    return r;
}
template <int V>
__device__ __forceinline__ Lanes<V> foldv(const Lanes<V> &a, const Lanes<V> &b, uint32_t rr)
{
    Lanes<V> r;
#pragma unroll
    for (int c = 0; c < V; c++) r.v[c] = fold1(a.v[c], b.v[c], rr);
    return r;
}

// ------------------------------------------------------------------ forward pass
// Work item w: output o = w / LG, lanes 4g .. 4g+3 (g = w % LG).
template <int L>
__device__ __forceinline__ void forward_item(const uint32_t *__restrict__ vin, uint32_t *__restrict__ vout,
                                             const uint32_t *__restrict__ ptr, const uint32_t *__restrict__ ent,
                                             uint64_t w)
{
    constexpr uint32_t LG = L / 4;
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

// The work items w0 + u BS (u < U) of one thread.  Each of an item's three dependent loads
// (its entry range, its first CSR entry, the input rows that entry names) is issued for all
// U items before any is used, so a thread has U gathers in flight instead of one.  The loads
// take no branches: an item past the end, or an output with no entry, reads output 0's
// range, an entry from vin and input row 0 (valid addresses, and row 0 stays cached) and is
// dropped; an add gate's second row is row 0 too.  Outputs with more than one entry add the
// others afterwards.
template <int L, int U, int BS>
__device__ __forceinline__ void forward_items(const uint32_t *__restrict__ vin, uint32_t *__restrict__ vout,
                                              const uint32_t *__restrict__ ptr, const uint32_t *__restrict__ ent,
                                              uint64_t w0, uint64_t total)
{
    constexpr uint32_t LG = L / 4;
    uint32_t o[U], g[U], e0[U], e1[U], a[U], b[U], c[U];
    uint4 va[U], vb[U], acc[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const uint64_t w = w0 + (uint64_t)u * BS;
        o[u] = w < total ? (uint32_t)(w / LG) : 0u;
        g[u] = (uint32_t)(w % LG);
        e0[u] = ptr[o[u]];
        e1[u] = ptr[o[u] + 1];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        if (w0 + (uint64_t)u * BS >= total) e1[u] = e0[u];
        const bool has = e0[u] < e1[u];
        const uint32_t *pe = has ? ent + 3 * (uint64_t)e0[u] : vin;
        const uint32_t x = pe[0], y = pe[1], z = pe[2];
        a[u] = has ? x : 0u;
        b[u] = has ? y : kAddMark;
        c[u] = has ? z : 0u;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        va[u] = ld4(vin + (uint64_t)a[u] * L + 4 * g[u]);
        vb[u] = ld4(vin + (uint64_t)(b[u] != kAddMark ? b[u] : 0u) * L + 4 * g[u]);
    }
#pragma unroll
    for (int u = 0; u < U; u++) acc[u] = mul4s(b[u] != kAddMark ? mul4(va[u], vb[u]) : va[u], c[u]);
#pragma unroll
    for (int u = 0; u < U; u++)
        for (uint32_t e = e0[u] + 1; e < e1[u]; e++) {
            const uint32_t x = ent[3 * (uint64_t)e], y = ent[3 * (uint64_t)e + 1], z = ent[3 * (uint64_t)e + 2];
            uint4 v = ld4(vin + (uint64_t)x * L + 4 * g[u]);
            if (y != kAddMark) v = mul4(v, ld4(vin + (uint64_t)y * L + 4 * g[u]));
            acc[u] = add4(acc[u], mul4s(v, z));
        }
#pragma unroll
    for (int u = 0; u < U; u++)
        if (w0 + (uint64_t)u * BS < total) st4(vout + (uint64_t)o[u] * L + 4 * g[u], acc[u]);
}

// The forward pass of one layer by nb blocks of BS threads: block `tile` takes the items
// [tile U BS, (tile + 1) U BS), then the same range nb U BS further on, and so on.
template <int L, int U, int BS>
__device__ __forceinline__ void forward_range(const uint32_t *__restrict__ vin, uint32_t *__restrict__ vout,
                                              const uint32_t *__restrict__ ptr, const uint32_t *__restrict__ ent,
                                              uint64_t total, uint32_t tile, uint32_t nb)
{
    for (uint64_t w0 = (uint64_t)tile * U * BS + threadIdx.x; w0 < total; w0 += (uint64_t)nb * U * BS)
        forward_items<L, U, BS>(vin, vout, ptr, ent, w0, total);
}

// work items per thread in flight in the forward pass (k_forward, and the batch engine's
// forward ops); measured with 2, 4 and 8: the 512-proof batch 177.3 / 178.4 / 185.0 ms,
// the 24 x 2^20 proof's forward pass 0.893 / 0.900 / 0.902 ms
constexpr int kFwdItems = 2;
constexpr int kLsGatherItems = kFwdItems;
template <int L>
__global__ void __launch_bounds__(kThreads) k_forward(const uint32_t *__restrict__ vin, uint32_t *__restrict__ vout,
                                                      const uint32_t *__restrict__ ptr,
                                                      const uint32_t *__restrict__ ent, uint32_t M)
{
    forward_range<L, kFwdItems, kThreads>(vin, vout, ptr, ent, (uint64_t)M * (L / 4), xcd_tile(blockIdx.x, gridDim.x),
                                          gridDim.x);
}

// ------------------------------------------------------------------ eq tables
// The eq tables of the R <= 3 repetitions are interleaved: entry z is kEqWords words,
// out[3 z + j] = f1 * eq(r1_j, z) + f2 * eq(r2_j, z) for rep j < R (words R .. 2 are
// zero), so that a gather of entry z (g_x, h_y: entries in wiring order) is one 12-byte
// load for every repetition.  eq(r, z) = prod_i (z_i ? r_i : 1 - r_i) (the
// reference's _eq at y = 0 / 1, include/LinearGKR/sumcheck_common.hpp:9-53).  f1, f2 come
// from ab[0], ab[1] when ab is given, else f1 = 1 and the second term is absent.  r1_j /
// r2_j are chal + j kChalStride + r1_off / r2_off.  A block of BS threads writes z in
// [chunk * CH * BS, (chunk + 1) * CH * BS): the low log2(BS) bits of z index a table of BS
// partial products per rep, the high bits one of CH, both in LDS.
constexpr int kEqChunk = 16;      // z values per thread in the batch engine's eq ops
constexpr int kEqChunkOne = 8;    // in k_eq (one table per launch: more blocks; 2, 4, 16: slower)
template <int BS, int CH = kEqChunk>
__device__ __forceinline__ void eq_block(uint32_t *__restrict__ o, const uint32_t *__restrict__ chal, uint32_t r1_off,
                                         uint32_t r2_off, const uint32_t *__restrict__ ab, uint32_t m, uint32_t chunk,
                                         uint32_t R)
{
    constexpr uint32_t LOG = BS == 64 ? 6 : BS == 128 ? 7 : BS == 256 ? 8 : 9;
    static_assert((1 << LOG) == BS, "block size must be 64, 128, 256 or 512");
    // the challenges r1_j, r2_j (at most kMaxVars each) are staged in LDS, so the product
    // chains below read them without a memory round trip per factor
    __shared__ uint32_t rv[2][3][kMaxVars];
    __shared__ uint32_t lo1[3][BS], lo2[3][BS], hi1[3][CH], hi2[3][CH];
    const bool two = ab != nullptr;
    const uint32_t f1 = two ? ab[0] : 1u, f2 = two ? ab[1] : 0u;
    const uint32_t h = m < LOG ? m : LOG, hb = m - h;
    const uint32_t t = threadIdx.x;
    const uint64_t base = (uint64_t)chunk * (CH * BS);
    for (uint32_t u = t; u < 2 * 3 * kMaxVars; u += BS) {
        const uint32_t w = u / (3 * kMaxVars), j = (u / kMaxVars) % 3, i = u % kMaxVars;
        rv[w][j][i] = (j < R && i < m && (w == 0 || two)) ? chal[j * kChalStride + (w ? r2_off : r1_off) + i] : 0u;
    }
    __syncthreads();
    // low tables: entry e of rep j is f prod_{i < h} (e_i ? r_i : 1 - r_i)
    for (uint32_t u = t; u < (R << h); u += BS) {
        const uint32_t j = u >> h, e = u & ((1u << h) - 1u);
        uint32_t a = f1, b = f2;
#pragma unroll
        for (uint32_t i = 0; i < LOG; i++)
            if (i < h) {
                const uint32_t bit = (e >> i) & 1u, x = rv[0][j][i], y = rv[1][j][i];
                a = m31_mul(a, bit ? x : m31_sub(1u, x));
                if (two) b = m31_mul(b, bit ? y : m31_sub(1u, y));
            }
        lo1[j][e] = a;
        lo2[j][e] = b;
    }
    // high tables: the CH high indices of this block's z range, bits h .. m - 1
    for (uint32_t u = t; u < R * CH; u += BS) {
        const uint32_t j = u / CH, c = u % CH;
        const uint64_t hi = (base >> h) + c;
        uint32_t a = 1u, b = 1u;
#pragma unroll
        for (uint32_t i = 0; i < kMaxVars; i++)
            if (i < hb) {
                const uint32_t bit = (uint32_t)(hi >> i) & 1u, x = rv[0][j][h + i], y = rv[1][j][h + i];
                a = m31_mul(a, bit ? x : m31_sub(1u, x));
                if (two) b = m31_mul(b, bit ? y : m31_sub(1u, y));
            }
        hi1[j][c] = a;
        hi2[j][c] = b;
    }
    __syncthreads();
    const uint64_t M = 1ull << m;
    const uint32_t mask = (1u << h) - 1u;
    for (int c = 0; c < CH; c++) {
        const uint64_t z = base + (uint64_t)c * BS + t;
        if (z < M) {
            const uint32_t li = (uint32_t)z & mask;
            const uint32_t hidx = (uint32_t)((z >> h) - (base >> h));
            uint32_t v[3] = {0u, 0u, 0u};
#pragma unroll
            for (uint32_t j = 0; j < 3; j++)
                if (j < R) {
                    v[j] = m31_mul(lo1[j][li], hi1[j][hidx]);
                    if (two) v[j] = m31_add(v[j], m31_mul(lo2[j][li], hi2[j][hidx]));
                }
            st3(o + kEqWords * z, Eq3{v[0], v[1], v[2]});
        }
    }
}

__global__ void __launch_bounds__(kThreads) k_eq(uint32_t *__restrict__ out, const uint32_t *__restrict__ chal,
                                                 uint32_t r1_off, uint32_t r2_off, const uint32_t *__restrict__ ab,
                                                 uint32_t m, uint32_t R)
{
    eq_block<kThreads, kEqChunkOne>(out, chal, r1_off, r2_off, ab, m, blockIdx.x, R);
}

// ------------------------------------------------------------------ g_x and h_y
// g_x and h_y as segmented reductions over tiles of rows.  The wiring entries of a tile
// are contiguous in the CSR and each carries its row, so they are spread evenly over the
// workgroup whatever the fan-in of each row; every entry's contributions are added into
// u64 LDS accumulators (sums of canonical values: exact, and order-free, so the result
// does not depend on the schedule), then the workgroup writes every row of the tile (zero
// where no gate has that input).
//
// hg_j[x] = sum over the gates with first input x of v[y] * (c * eqZ_j[z]) (mul) or
// c * eqZ_j[z] (add, broadcast to every lane), for every rep j at once: each wiring entry
// and each gathered v[y] row is read once for all R repetitions.  A tile for a workgroup
// of BS threads is tile_lanes(BS) lane-rows (4 per thread, at most 1024: 64 rows of 16
// lanes); its u64 accumulators are [rep][row][lane].
constexpr uint32_t tile_lanes(uint32_t bs) { return 4 * bs < 1024 ? 4 * bs : 1024; }
template <int L, int BS>
constexpr uint32_t gx_rows() { return tile_lanes(BS) / L; }
// A row of the g_x accumulators is L + 1 u64 (one of padding): a wave's 64 atomics of one
// instruction hit up to 16 rows at the same lane offset, and with 8L-byte rows (a multiple
// of the 128 bytes the 32 LDS banks span) they all fell on the same two banks; the padding
// shifts each row by two banks (the bank-conflict cycles of k_gx were 57 % of its LDS
// cycles before it).
template <int L, int BS>
constexpr uint32_t gx_acc_words() { return 3 * gx_rows<L, BS>() * (L + 1); }
template <int L, int BS>
constexpr uint32_t tile_acc_words() { return gx_acc_words<L, BS>() > 3 * tile_lanes(BS) ? gx_acc_words<L, BS>() : 3 * tile_lanes(BS); }

// The tiles tfirst, tfirst + tstep, ... (< tiles) of g_x, software-pipelined: while a tile's
// contributions are added and its rows written, the gathers of the next tile, the CSR
// entries of the one after and the entry range of the third are in flight, so a block
// waits for one memory latency per tile instead of three dependent ones (row range, then
// entries, then the v rows and eqZ triples they name).  The first kGxPipe items of a tile
// per thread go through the pipeline; a tile with more work adds the rest directly.
constexpr int kGxPipe = 2;
constexpr uint32_t kGxBlocks = 1536;
template <int L, int BS>
__device__ __forceinline__ void gx_tiles(unsigned long long *__restrict__ acc, uint32_t *__restrict__ hg, uint64_t hg_rs,
                                         const uint32_t *__restrict__ v, const uint32_t *__restrict__ eqz,
                                         const uint32_t *__restrict__ ptr, const uint32_t *__restrict__ ent, uint32_t N,
                                         uint32_t R, uint32_t tfirst, uint32_t tstep, uint32_t tiles)
{
    constexpr uint32_t LG = L / 4, ROWS = gx_rows<L, BS>(), RS = L + 1;
    constexpr int NI = kGxPipe;
    const uint32_t t = threadIdx.x;
    // The loads below are issued without branches (an item out of range loads from the
    // start of v or eqZ, always valid, and is not used), so that the compiler can count
    // them and let them stay in flight across the tile's work.
    auto range = [&](uint32_t tile) {  // entry range of a tile; tiles past the end read row 0
        const uint32_t x0 = tile < tiles ? tile * ROWS : 0u, x1 = N - x0 < ROWS ? N : x0 + ROWS;
        return make_uint2(ptr[x0], ptr[x1]);
    };
    auto count = [&](uint32_t tile, uint2 r) { return tile < tiles ? (uint64_t)(r.y - r.x) * LG : 0ull; };
    auto entries = [&](uint64_t cnt, uint2 r, uint4 (&en)[NI]) {
#pragma unroll
        for (int i = 0; i < NI; i++) {
            const uint64_t w = t + (uint64_t)i * BS;
            en[i] = ld4(w < cnt ? ent + kGxWords * (r.x + w / LG) : v);
        }
    };
    auto gathers = [&](uint64_t cnt, const uint4 (&en)[NI], uint4 (&vb)[NI], Eq3 (&ez)[NI]) {
#pragma unroll
        for (int i = 0; i < NI; i++) {
            const uint64_t w = t + (uint64_t)i * BS;
            const bool ok = w < cnt;
            vb[i] = ld4(ok && en[i].x != kAddMark ? v + (uint64_t)en[i].x * L + 4 * (uint32_t)(w % LG) : v);
            ez[i] = ld3(ok ? eqz + kEqWords * (uint64_t)en[i].y : eqz);
        }
    };
    auto add = [&](uint32_t x0, const uint4 &en, uint4 vb, const Eq3 &ez, uint64_t w) {
        const uint32_t g = (uint32_t)(w % LG);
        if (en.x == kAddMark) vb = make_uint4(1u, 1u, 1u, 1u);
        const uint32_t c[3] = {m31_mul(en.z, ez.r0), R > 1 ? m31_mul(en.z, ez.r1) : 0u, R > 2 ? m31_mul(en.z, ez.r2) : 0u};
        unsigned long long *a = acc + (uint64_t)(en.w - x0) * RS + 4 * g;
#pragma unroll
        for (uint32_t j = 0; j < 3; j++)
            if (j < R) {
                unsigned long long *aj = a + j * ROWS * RS;
                atomicAdd(aj, (unsigned long long)m31_mul_wide(vb.x, c[j]));
                atomicAdd(aj + 1, (unsigned long long)m31_mul_wide(vb.y, c[j]));
                atomicAdd(aj + 2, (unsigned long long)m31_mul_wide(vb.z, c[j]));
                atomicAdd(aj + 3, (unsigned long long)m31_mul_wide(vb.w, c[j]));
            }
    };
    // prologue: the ranges of the first three tiles, the entries of the first two, the
    // gathers of the first
    uint32_t tg = tfirst, te = tfirst + tstep, tp = te + tstep;
    uint2 rg = range(tg), re = range(te), rp = range(tp);
    uint4 eng[NI], ene[NI], vbg[NI];
    Eq3 ezg[NI];
    entries(count(tg, rg), rg, eng);
    entries(count(te, re), re, ene);
    gathers(count(tg, rg), eng, vbg, ezg);
    while (tg < tiles) {
        const uint32_t x0 = tg * ROWS;
        const uint32_t rows = N - x0 < ROWS ? N - x0 : ROWS;
#pragma unroll
        for (uint32_t k = 0; k < (3 * ROWS * RS + BS - 1) / BS; k++)
            if (t + k * BS < R * ROWS * RS) acc[t + k * BS] = 0;
        __syncthreads();
        const uint64_t work = count(tg, rg);
        // items past the pipelined ones (a tile with many entries), loaded here directly:
        // before the next tiles' loads are issued, so that waiting for these does not wait
        // for those
        for (uint64_t w = t + (uint64_t)NI * BS; w < work; w += BS) {
            const uint4 en = ld4(ent + kGxWords * (rg.x + w / LG));
            const uint4 vb = en.x != kAddMark ? ld4(v + (uint64_t)en.x * L + 4 * (uint32_t)(w % LG)) : make_uint4(1u, 1u, 1u, 1u);
            add(x0, en, vb, ld3(eqz + kEqWords * (uint64_t)en.y), w);
        }
        // the next tiles' loads, in flight through this tile's additions and writes
        const uint32_t tq = tp + tstep;
        uint4 vbn[NI], enp[NI];
        Eq3 ezn[NI];
        gathers(count(te, re), ene, vbn, ezn);
        entries(count(tp, rp), rp, enp);
        const uint2 rq = range(tq);
#pragma unroll
        for (int i = 0; i < NI; i++) {
            const uint64_t w = t + (uint64_t)i * BS;
            if (w < work) add(x0, eng[i], vbg[i], ezg[i], w);
        }
        __syncthreads();
        // straight-line code (no loop), so that the loads above stay in flight through it
#pragma unroll
        for (uint32_t k = 0; k < (3 * ROWS * LG + BS - 1) / BS; k++) {
            const uint32_t i = t + k * BS;
            if (i < R * rows * LG) {
                const uint32_t j = i / (rows * LG), rem = i % (rows * LG), r = rem / LG, g = rem % LG;
                const unsigned long long *aj = acc + (j * ROWS + r) * RS + 4 * g;
                st4(hg + j * hg_rs + (uint64_t)(x0 + r) * L + 4 * g,
                    make_uint4(m31_reduce64(aj[0]), m31_reduce64(aj[1]), m31_reduce64(aj[2]), m31_reduce64(aj[3])));
            }
        }
        __syncthreads();
        tg = te; te = tp; tp = tq;
        rg = re; re = rp; rp = rq;
#pragma unroll
        for (int i = 0; i < NI; i++) {
            eng[i] = ene[i];
            ene[i] = enp[i];
            vbg[i] = vbn[i];
            ezg[i] = ezn[i];
        }
    }
}

template <int L>
__global__ void __launch_bounds__(kThreads) k_gx(uint32_t *__restrict__ hg, uint64_t hg_rs,
                                                 const uint32_t *__restrict__ v, const uint32_t *__restrict__ eqz,
                                                 const uint32_t *__restrict__ ptr,
                                                 const uint32_t *__restrict__ ent, uint32_t N, uint32_t R)
{
    __shared__ unsigned long long acc[gx_acc_words<L, kThreads>()];
    constexpr uint32_t ROWS = gx_rows<L, kThreads>();
    const uint32_t tiles = (N + ROWS - 1) / ROWS;
    gx_tiles<L, kThreads>(acc, hg, hg_rs, v, eqz, ptr, ent, N, R, xcd_tile(blockIdx.x, gridDim.x), gridDim.x, tiles);
}

// s_j[y] = sum over the mul gates with second input y of eqZ_j[z] * eqX_j[x] * c, for every
// rep j at once (phase-2 hg_j = vx_j * s_j); a tile is tile_lanes(BS) rows.
constexpr uint32_t kHyAccWords = 3 * 1024;

// The tiles tfirst, tfirst + tstep, ... (< tiles) of h_y, pipelined like gx_tiles: while a
// tile's products are added and its rows written, the eq gathers of the next tile, the CSR
// entries of the one after and the entry range of the third are in flight.
constexpr int kHyPipe = 2;
template <int BS>
__device__ __forceinline__ void hy_tiles(unsigned long long *__restrict__ acc, uint32_t *__restrict__ s, uint64_t s_rs,
                                         const uint32_t *__restrict__ eqz, const uint32_t *__restrict__ eqx,
                                         const uint32_t *__restrict__ ptr, const uint32_t *__restrict__ ent, uint32_t N,
                                         uint32_t R, uint32_t tfirst, uint32_t tstep, uint32_t tiles)
{
    constexpr uint32_t ROWS = tile_lanes(BS);
    constexpr int NI = kHyPipe;
    const uint32_t t = threadIdx.x;
    // branch-free loads: an item out of range reads the start of eqX (always valid) instead
    auto range = [&](uint32_t tile) {
        const uint32_t y0 = tile < tiles ? tile * ROWS : 0u, y1 = N - y0 < ROWS ? N : y0 + ROWS;
        return make_uint2(ptr[y0], ptr[y1]);
    };
    auto count = [&](uint32_t tile, uint2 r) { return tile < tiles ? r.y - r.x : 0u; };
    auto entries = [&](uint32_t cnt, uint2 r, uint4 (&en)[NI]) {  // {x, z, coef, y}
#pragma unroll
        for (int i = 0; i < NI; i++) {
            const uint32_t w = t + i * BS;
            en[i] = ld4(w < cnt ? ent + kHyWords * (uint64_t)(r.x + w) : eqx);
        }
    };
    auto gathers = [&](uint32_t cnt, const uint4 (&en)[NI], Eq3 (&a)[NI], Eq3 (&b)[NI]) {
#pragma unroll
        for (int i = 0; i < NI; i++) {
            const bool ok = t + i * BS < cnt;
            a[i] = ld3(ok ? eqz + kEqWords * (uint64_t)en[i].y : eqz);
            b[i] = ld3(ok ? eqx + kEqWords * (uint64_t)en[i].x : eqx);
        }
    };
    auto add = [&](uint32_t y0, const uint4 &en, const Eq3 &a, const Eq3 &b) {
        const uint32_t q[3] = {m31_mul(a.r0, b.r0), R > 1 ? m31_mul(a.r1, b.r1) : 0u, R > 2 ? m31_mul(a.r2, b.r2) : 0u};
#pragma unroll
        for (uint32_t j = 0; j < 3; j++)
            if (j < R) atomicAdd(acc + j * ROWS + (en.w - y0), (unsigned long long)m31_mul_wide(q[j], en.z));
    };
    uint32_t tg = tfirst, te = tfirst + tstep, tp = te + tstep;
    uint2 rg = range(tg), re = range(te), rp = range(tp);
    uint4 eng[NI], ene[NI];
    Eq3 ag[NI], bg[NI];
    entries(count(tg, rg), rg, eng);
    entries(count(te, re), re, ene);
    gathers(count(tg, rg), eng, ag, bg);
    while (tg < tiles) {
        const uint32_t y0 = tg * ROWS;
        const uint32_t rows = N - y0 < ROWS ? N - y0 : ROWS;
#pragma unroll
        for (uint32_t k = 0; k < (3 * ROWS + BS - 1) / BS; k++)
            if (t + k * BS < R * ROWS) acc[t + k * BS] = 0;
        __syncthreads();
        const uint32_t work = count(tg, rg);
        // items past the pipelined ones, before the next tiles' loads are issued
        for (uint32_t w = t + NI * BS; w < work; w += BS) {
            const uint4 en = ld4(ent + kHyWords * (uint64_t)(rg.x + w));
            add(y0, en, ld3(eqz + kEqWords * (uint64_t)en.y), ld3(eqx + kEqWords * (uint64_t)en.x));
        }
        const uint32_t tq = tp + tstep;
        uint4 enp[NI];
        Eq3 an[NI], bn[NI];
        gathers(count(te, re), ene, an, bn);
        entries(count(tp, rp), rp, enp);
        const uint2 rq = range(tq);
#pragma unroll
        for (int i = 0; i < NI; i++)
            if (t + i * BS < work) add(y0, eng[i], ag[i], bg[i]);
        __syncthreads();
        // four rows per store where the tile is whole in fours (rows is a power of two or 1024)
        if ((rows & 3) == 0) {
#pragma unroll
            for (uint32_t k = 0; k < (3 * ROWS / 4 + BS - 1) / BS; k++) {
                const uint32_t i = t + k * BS;
                if (i < R * rows / 4) {
                    const uint32_t j = i / (rows / 4), r = 4 * (i % (rows / 4));
                    const unsigned long long *aj = acc + j * ROWS + r;
                    st4(s + j * s_rs + y0 + r, make_uint4(m31_reduce64(aj[0]), m31_reduce64(aj[1]),
                                                          m31_reduce64(aj[2]), m31_reduce64(aj[3])));
                }
            }
        } else {
            for (uint32_t i = t; i < R * rows; i += BS) {
                const uint32_t j = i / rows, r = i % rows;
                s[j * s_rs + y0 + r] = m31_reduce64(acc[j * ROWS + r]);
            }
        }
        __syncthreads();
        tg = te; te = tp; tp = tq;
        rg = re; re = rp; rp = rq;
#pragma unroll
        for (int i = 0; i < NI; i++) {
            eng[i] = ene[i];
            ene[i] = enp[i];
            ag[i] = an[i];
            bg[i] = bn[i];
        }
    }
}

__global__ void __launch_bounds__(kThreads) k_hy(uint32_t *__restrict__ s, uint64_t s_rs,
                                                 const uint32_t *__restrict__ eqz,
                                                 const uint32_t *__restrict__ eqx,
                                                 const uint32_t *__restrict__ ptr, const uint32_t *__restrict__ ent,
                                                 uint32_t N, uint32_t R)
{
    __shared__ unsigned long long acc[kHyAccWords];
    constexpr uint32_t ROWS = tile_lanes(kThreads);
    const uint32_t tiles = (N + ROWS - 1) / ROWS;
    hy_tiles<kThreads>(acc, s, s_rs, eqz, eqx, ptr, ent, N, R, xcd_tile(blockIdx.x, gridDim.x), gridDim.x, tiles);
}

// ------------------------------------------------------------------ reductions
// Reduces the 3 x 4 lane sums (u64) of the 256 threads [t0, t0 + 256) that share
// t % LG.  red is this group's LDS scratch [4][3L]; the result goes to out[q * ostride],
// q = pk * L + 4 * g + c, written by threads t0 .. t0 + 3L - 1.  Every thread of the block
// must call it (it synchronizes the block).
template <int L>
__device__ __forceinline__ void group_reduce(uint64_t (&acc)[3][4], uint32_t (*red)[3 * L], uint32_t *out,
                                             uint32_t ostride = 1)
{
    constexpr uint32_t LG = L / 4;
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
    const uint32_t lane = threadIdx.x & 63, wave = (threadIdx.x >> 6) & 3;
    if (lane < LG) {
#pragma unroll
        for (int pk = 0; pk < 3; pk++)
#pragma unroll
            for (int c = 0; c < 4; c++) red[wave][pk * L + 4 * lane + c] = v[pk][c];
    }
    __syncthreads();
    const uint32_t tg = threadIdx.x & 255;
    if (tg < 3 * L) {
        uint32_t s = 0;
#pragma unroll
        for (int w = 0; w < 4; w++) s = m31_add(s, red[w][tg]);
        out[tg * ostride] = s;
    }
}

// ------------------------------------------------------------------ sumcheck rounds
// One item of round k: the pair (2i, 2i+1) of the round-k tables for lanes 4g..4g+3.
// FUSED: the round-k tables are produced here by folding rows 4i..4i+3 of the round-(k-1)
// tables with challenge r (receive_challenge, sumcheck_helper.hpp:104-133) and stored.
// Accumulates p0 += f0 h0, p1 += f1 h1, p2 += (f0 + f1)(h0 + h1) (poly_eval_at, :68-102).
// PH2: h is the phase-2 scalar table s; hg = vx * s is rank one and vx multiplies later.
template <int L, bool PH2, bool FUSED>
__device__ __forceinline__ void round_item(const uint32_t *__restrict__ fs, uint32_t *__restrict__ fd,
                                           const uint32_t *__restrict__ hs, uint32_t *__restrict__ hd, uint64_t i,
                                           uint32_t g, uint32_t r, uint64_t (&acc)[3][4])
{
    uint4 f0, f1, h0, h1;
    uint32_t s0 = 0, s1 = 0;
    if (FUSED) {
        const uint32_t *src = fs + 4 * i * L + 4 * g;
        const uint4 a0 = ld4(src), a1 = ld4(src + L), a2 = ld4(src + 2 * L), a3 = ld4(src + 3 * L);
        if (PH2) {
            const uint4 b = *reinterpret_cast<const uint4 *>(hs + 4 * i);
            s0 = fold1(b.x, b.y, r);
            s1 = fold1(b.z, b.w, r);
            if (g == 0) *reinterpret_cast<uint2 *>(hd + 2 * i) = make_uint2(s0, s1);
        } else {
            const uint32_t *hsrc = hs + 4 * i * L + 4 * g;
            const uint4 c0 = ld4(hsrc), c1 = ld4(hsrc + L), c2 = ld4(hsrc + 2 * L), c3 = ld4(hsrc + 3 * L);
            h0 = fold4(c0, c1, r);
            h1 = fold4(c2, c3, r);
            st4(hd + 2 * i * L + 4 * g, h0);
            st4(hd + (2 * i + 1) * L + 4 * g, h1);
        }
        f0 = fold4(a0, a1, r);
        f1 = fold4(a2, a3, r);
        st4(fd + 2 * i * L + 4 * g, f0);
        st4(fd + (2 * i + 1) * L + 4 * g, f1);
    } else {
        f0 = ld4(fs + 2 * i * L + 4 * g);
        f1 = ld4(fs + (2 * i + 1) * L + 4 * g);
        if (PH2) {
            const uint2 b = *reinterpret_cast<const uint2 *>(hs + 2 * i);
            s0 = b.x;
            s1 = b.y;
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

// Round k of a phase whose source is table k - 2 (the first small-table round after the
// round pairs below): rows 2i, 2i+1 of table k are folds of rows 8i .. 8i+7 of table k - 2
// by ra = r_{k-2} and then rb = r_{k-1}; they are stored and their round-k sums added.
template <int L, bool PH2>
__device__ __forceinline__ void round_item_dbl(const uint32_t *__restrict__ fs, uint32_t *__restrict__ fd,
                                               const uint32_t *__restrict__ hs, uint32_t *__restrict__ hd, uint64_t i,
                                               uint32_t g, uint32_t ra, uint32_t rb, uint64_t (&acc)[3][4])
{
    uint4 f[2], h[2];
#pragma unroll
    for (int r = 0; r < 2; r++) {
        const uint32_t *src = fs + (8 * i + 4 * r) * L + 4 * g;
        f[r] = fold4(fold4(ld4(src), ld4(src + L), ra), fold4(ld4(src + 2 * L), ld4(src + 3 * L), ra), rb);
        st4(fd + (2 * i + r) * L + 4 * g, f[r]);
        if (PH2) {
            const uint4 b = *reinterpret_cast<const uint4 *>(hs + 8 * i + 4 * r);
            const uint32_t sv = fold1(fold1(b.x, b.y, ra), fold1(b.z, b.w, ra), rb);
            if (g == 0) hd[2 * i + r] = sv;
            h[r] = make_uint4(sv, sv, sv, sv);
        } else {
            const uint32_t *hsrc = hs + (8 * i + 4 * r) * L + 4 * g;
            h[r] = fold4(fold4(ld4(hsrc), ld4(hsrc + L), ra), fold4(ld4(hsrc + 2 * L), ld4(hsrc + 3 * L), ra), rb);
            st4(hd + (2 * i + r) * L + 4 * g, h[r]);
        }
    }
    const uint4 e0 = mul4(f[0], h[0]), e1 = mul4(f[1], h[1]), e2 = mul4(add4(f[0], f[1]), add4(h[0], h[1]));
    acc[0][0] += e0.x; acc[0][1] += e0.y; acc[0][2] += e0.z; acc[0][3] += e0.w;
    acc[1][0] += e1.x; acc[1][1] += e1.y; acc[1][2] += e1.z; acc[1][3] += e1.w;
    acc[2][0] += e2.x; acc[2][1] += e2.y; acc[2][2] += e2.z; acc[2][3] += e2.w;
}

// ------------------------------------------------------------------ round pairs
// Rounds k and k+1 of a phase in one pass over the round-k tables T (f, and h: hg in phase
// 1, the scalar s in phase 2).  Round k+1's pairs are folds of rows (4i, 4i+1) and
// (4i+2, 4i+3) by the challenge r_k that is not drawn yet, so each of its three sums is a
// polynomial of degree 2 in r_k.  With x0..x3 = T[4i .. 4i+3] (of f and of h) those are
//   s0(r) = sum (x0 + r (x1 - x0))_f (..)_h,   s1(r) = the same of x2, x3,
//   s2(r) = sum ((x0 + x2) + r ((x1 + x3) - (x0 + x2)))_f (..)_h,
// and with u = x0, x1, x1 - x0 (and v, w likewise for s1, s2) every coefficient, and every
// sum of round k, is a linear combination of nine products per lane:
//   P0 = f0 h0, P1 = f1 h1, P2 = (f1 - f0)(h1 - h0), P3..P5 the same of rows 2, 3,
//   P6 = (f0 + f2)(h0 + h2), P7 = (f1 + f3)(h1 + h3), P8 = (f1 + f3 - f0 - f2)(h1 + h3 - h0 - h2);
//   s0(r) = P0 + r (P1 - P0 - P2) + r^2 P2 (likewise s1 from P3..P5, s2 from P6..P8), and round
//   k: p0 = P0 + P3, p1 = P1 + P4, sum (f0 + f1)(h0 + h1) = 2 P0 + 2 P1 - P2 (+ rows 2, 3).
// Once r_k is drawn the transcript kernel evaluates them: the same field values as a pass
// over the folded round-(k+1) tables, which are never stored.  T is the layer input (MODE 0:
// rounds 0 and 1) or is made here by folding rows 16i .. 16i+15 of table k - 2 by
// ra = r_{k-2} and then rb = r_{k-1} (MODE 2), or rows 8i .. 8i+7 of table k - 1 by ra =
// r_{k-1} (MODE 1, the small-table rounds), and stored for the next pass.  Sums are
// canonical (u32 accumulators keep the register count, and with it the occupancy, up).
constexpr uint32_t kPairVals = 9;

__device__ __forceinline__ void pair_add(uint32_t &acc, uint32_t a, uint32_t b) { acc = m31_add(acc, m31_mul(a, b)); }
// lazy form: one 64-bit multiply-add per product; the caller folds (pair_fold) before the
// sum can reach 2^64 (a product of canonical values is below 2^62)
__device__ __forceinline__ void pair_add(uint64_t &acc, uint32_t a, uint32_t b) { acc += (uint64_t)a * b; }
__device__ __forceinline__ void pair_fold(uint64_t &acc) { acc = (acc & kP) + (acc >> 31); }  // < 2^34

// g indexes the thread's V lanes [V g, V g + V) of an entry (L / V threads per entry).  A is
// the accumulator: canonical u32, or u64 sums of products (see pair_add).
template <int L, bool PH2, int MODE, int V = 4, typename A = uint32_t>
__device__ __forceinline__ void pair_item(const uint32_t *__restrict__ fs, uint32_t *__restrict__ fd,
                                          const uint32_t *__restrict__ hs, uint32_t *__restrict__ hd, uint64_t i,
                                          uint32_t g, uint32_t ra, uint32_t rb, A (&acc)[kPairVals][V])
{
    Lanes<V> f[4];
    Lanes<V> h[4];    // phase 1
    uint32_t sv[4];   // phase 2: the scalar h of each row
    if (MODE == 0) {
#pragma unroll
        for (int r = 0; r < 4; r++) f[r] = ldv<V>(fs + (4 * i + r) * L + V * g);
        if (PH2) {
            const uint4 b = *reinterpret_cast<const uint4 *>(hs + 4 * i);
            sv[0] = b.x; sv[1] = b.y; sv[2] = b.z; sv[3] = b.w;
        } else {
#pragma unroll
            for (int r = 0; r < 4; r++) h[r] = ldv<V>(hs + (4 * i + r) * L + V * g);
        }
    } else if (MODE == 1) {
        // table k from rows 8i .. 8i+7 of table k - 1, folded once by ra
        uint4 b[2];
        if (PH2) {
            b[0] = *reinterpret_cast<const uint4 *>(hs + 8 * i);
            b[1] = *reinterpret_cast<const uint4 *>(hs + 8 * i + 4);
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const uint32_t *src = fs + (8 * i + 2 * r) * L + V * g;
            f[r] = foldv(ldv<V>(src), ldv<V>(src + L), ra);
            stv(fd + (4 * i + r) * L + V * g, f[r]);
            if (PH2) {
                sv[r] = (r & 1) ? fold1(b[r >> 1].z, b[r >> 1].w, ra) : fold1(b[r >> 1].x, b[r >> 1].y, ra);
            } else {
                const uint32_t *hsrc = hs + (8 * i + 2 * r) * L + V * g;
                h[r] = foldv(ldv<V>(hsrc), ldv<V>(hsrc + L), ra);
                stv(hd + (4 * i + r) * L + V * g, h[r]);
            }
        }
        if (PH2 && g == 0) *reinterpret_cast<uint4 *>(hd + 4 * i) = make_uint4(sv[0], sv[1], sv[2], sv[3]);
    } else {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const uint32_t *src = fs + (16 * i + 4 * r) * L + V * g;
            f[r] = foldv(foldv(ldv<V>(src), ldv<V>(src + L), ra), foldv(ldv<V>(src + 2 * L), ldv<V>(src + 3 * L), ra),
                         rb);
            stv(fd + (4 * i + r) * L + V * g, f[r]);
            if (PH2) {
                const uint4 b = *reinterpret_cast<const uint4 *>(hs + 16 * i + 4 * r);
                sv[r] = fold1(fold1(b.x, b.y, ra), fold1(b.z, b.w, ra), rb);
            } else {
                const uint32_t *hsrc = hs + (16 * i + 4 * r) * L + V * g;
                h[r] = foldv(foldv(ldv<V>(hsrc), ldv<V>(hsrc + L), ra),
                             foldv(ldv<V>(hsrc + 2 * L), ldv<V>(hsrc + 3 * L), ra), rb);
                stv(hd + (4 * i + r) * L + V * g, h[r]);
            }
        }
        if (PH2 && g == 0) *reinterpret_cast<uint4 *>(hd + 4 * i) = make_uint4(sv[0], sv[1], sv[2], sv[3]);
    }
    // the nine factors of f (lanes) and of h (lanes, or one scalar in phase 2)
    const Lanes<V> f02 = addv(f[0], f[2]), f13 = addv(f[1], f[3]);
    const Lanes<V> ff[9] = {f[0], f[1], subv(f[1], f[0]), f[2], f[3], subv(f[3], f[2]), f02, f13, subv(f13, f02)};
    if (PH2) {
        const uint32_t s02 = m31_add(sv[0], sv[2]), s13 = m31_add(sv[1], sv[3]);
        const uint32_t hh[9] = {sv[0], sv[1], m31_sub(sv[1], sv[0]), sv[2], sv[3], m31_sub(sv[3], sv[2]),
                                s02, s13, m31_sub(s13, s02)};
#pragma unroll
        for (int q = 0; q < 9; q++)
#pragma unroll
            for (int c = 0; c < V; c++) pair_add(acc[q][c], ff[q].v[c], hh[q]);
    } else {
        const Lanes<V> h02 = addv(h[0], h[2]), h13 = addv(h[1], h[3]);
        const Lanes<V> hh[9] = {h[0], h[1], subv(h[1], h[0]), h[2], h[3], subv(h[3], h[2]), h02, h13, subv(h13, h02)};
#pragma unroll
        for (int q = 0; q < 9; q++)
#pragma unroll
            for (int c = 0; c < V; c++) pair_add(acc[q][c], ff[q].v[c], hh[q].v[c]);
    }
}

// ------------------------------------------------------------------ transcript
// The transcript (include/fiat_shamir/transcript.hpp:87-175) continues on the device over
// the proof bytes after the commitment.  tstate: [0] length in bytes, [1] start of the
// unhashed suffix, [2..9] digest.  A challenge hashes only the bytes appended since the
// previous challenge (or the previous digest when there are none), so within a round the
// R repetitions' challenges are independent: each hashes its own 3 field elements.  The
// exceptions are the last x round, where rep j's message begins with vx_{j-1}
// (sumcheck.hpp:90-95), and the layer end (vy claims, alpha, beta: sumcheck.hpp:98-103,
// gkr.hpp:65-66); those run in order on one lane.
struct TsArgs {
    uint32_t R;
    int phase;            // 1 = x variables, 2 = y variables
    uint32_t k;           // round index within the phase
    int last;             // k == n - 1
    uint32_t *chal;       // this layer's challenge slots, rep stride kChalStride
    uint32_t *vx, *vy;    // R * L each
    const uint32_t *fin;  // round-k f table (rows 0 and 1 are used after the last round)
    uint64_t fin_rs;
    uint32_t *tstate;
    uint32_t *tail;
    uint32_t *ab_out;     // alpha, beta drawn after the layer
    int hash_kind;
};

__device__ __forceinline__ uint32_t challenge_from(const uint32_t d[8]) { return m31_from_u32_full(d[0]); }

// Message word q (of 3L) of rep j from the round sums: p0, p1 and p2 = 6 p1 + 3 p0 - 2 p2
// (sumcheck_helper.hpp:100); in phase 2 each sum is first multiplied by vx.
template <int L>
__device__ __forceinline__ uint32_t msg_word(const uint32_t *sj, const uint32_t *vxj, int phase, uint32_t q)
{
    const uint32_t pk = q / L, c = q % L;
    uint32_t q0 = sj[c], q1 = sj[L + c], q2 = sj[2 * L + c];
    if (phase == 2) {
        const uint32_t v = vxj[c];
        q0 = m31_mul(q0, v);
        q1 = m31_mul(q1, v);
        q2 = m31_mul(q2, v);
    }
    if (pk == 0) return q0;
    if (pk == 1) return q1;
    return m31_sub(m31_add(m31_mul(q1, 6u), m31_mul(q0, 3u)), m31_mul(q2, 2u));
}

// Block-wide transcript step for round k of one phase, all reps.  sums: LDS [R][3L].
// r_out: LDS [R] receives the challenges.  Every hash is computed by a whole wavefront:
// wave j hashes rep j when the reps are independent.  The words a hash reads are staged
// in LDS as they are produced (they also go to the proof tail), so a hash does not wait
// on a global-memory round trip.  At most L words can be pending before a round (the last
// x claim); hb keeps L lead words in front of every message.  Every thread of the block
// must call it; the block has at least 64 R threads.
template <int L>
__device__ __forceinline__ void transcript_step(const uint32_t *sums, const TsArgs &a, uint32_t *r_out)
{
    constexpr uint32_t Q = 3 * L, S = L + Q;
    __shared__ uint32_t ts[10];
    __shared__ uint32_t dig[8];
    __shared__ uint32_t hb[3 * S];
    const uint32_t t = threadIdx.x, wave = t >> 6, lane = t & 63;
    if (t < 10) ts[t] = a.tstate[t];
    __syncthreads();
    const uint32_t base = ts[0], hs0 = ts[1];
    const uint32_t P = (base - hs0) / 4;  // unhashed words in front of this round's messages
    const bool staged = P <= L;
    if (t < P && staged) hb[L - P + t] = a.tail[hs0 / 4 + t];
    if (!(a.phase == 1 && a.last)) {
        for (uint32_t w = t; w < a.R * Q; w += blockDim.x) {
            const uint32_t j = w / Q, q = w % Q;
            const uint32_t v = msg_word<L>(sums + j * Q, a.vx + j * L, a.phase, q);
            a.tail[base / 4 + w] = v;
            hb[j * S + L + q] = v;
        }
        __syncthreads();
        if (wave < a.R) {
            const uint32_t j = wave;
            const uint32_t lead = j == 0 ? P : 0;
            const uint32_t *m = staged ? hb + j * S + L - lead : a.tail + (j == 0 ? hs0 / 4 : base / 4 + j * Q);
            uint32_t d[8];
            hash_words_wave(a.hash_kind, m, lead + Q, d);
            if (lane == 0) {
                const uint32_t r = challenge_from(d);
                r_out[j] = r;
                a.chal[j * kChalStride + (a.phase == 2 ? 32u : 0u) + a.k] = r;
                if (j == a.R - 1) {
                    for (int i = 0; i < 8; i++) ts[2 + i] = d[i];
                    ts[0] = base + a.R * 4 * Q;
                    ts[1] = ts[0];
                }
            }
        }
        __syncthreads();
        if (a.last) {
            // phase 2 ends the layer: vy_claim = f after the last fold for every rep, then all
            // vy claims are appended and alpha, beta drawn (sumcheck.hpp:98-103, gkr.hpp:65-66)
            const uint32_t len = ts[0];
            if (t < a.R * L) {
                const uint32_t j = t / L, c = t % L;
                const uint32_t *f = a.fin + j * a.fin_rs;
                const uint32_t v = fold1(f[c], f[L + c], r_out[j]);
                a.vy[t] = v;
                a.tail[len / 4 + t] = v;
                hb[t] = v;
            }
            __syncthreads();
            if (wave == 0) {
                uint32_t d[8];
                hash_words_wave(a.hash_kind, hb, a.R * L, d);
                if (lane == 0) {
                    a.ab_out[0] = challenge_from(d);
                    for (int i = 0; i < 8; i++) dig[i] = d[i];
                }
            }
            __syncthreads();
            if (wave == 0) {
                uint32_t d[8];
                hash_words_wave(a.hash_kind, dig, 8, d);
                if (lane == 0) {
                    a.ab_out[1] = challenge_from(d);
                    for (int i = 0; i < 8; i++) ts[2 + i] = d[i];
                    ts[0] = len + 4 * a.R * L;
                    ts[1] = ts[0];
                }
            }
        }
    } else {
        // last x round, in order: rep j hashes (vx_{j-1} ||) its message, then appends vx_j
        // (sumcheck.hpp:90-95).  hb[0, L) holds the lead words, hb[L, L + Q) the message.
        uint32_t len = base, hs = hs0;
        for (uint32_t j = 0; j < a.R; j++) {
            const uint32_t lead = (len - hs) / 4;
            for (uint32_t q = t; q < Q; q += blockDim.x) {
                const uint32_t v = msg_word<L>(sums + j * Q, a.vx + j * L, 1, q);
                a.tail[len / 4 + q] = v;
                hb[L + q] = v;
            }
            __syncthreads();
            if (wave == 0) {
                const uint32_t *m = (lead <= L) ? hb + L - lead : a.tail + hs / 4;
                uint32_t d[8];
                hash_words_wave(a.hash_kind, m, lead + Q, d);
                if (lane == 0) {
                    const uint32_t r = challenge_from(d);
                    r_out[j] = r;
                    a.chal[j * kChalStride + a.k] = r;
                    for (int i = 0; i < 8; i++) ts[2 + i] = d[i];
                }
            }
            __syncthreads();
            len += 4 * Q;
            hs = len;
            if (t < L) {
                const uint32_t *f = a.fin + j * a.fin_rs;
                const uint32_t v = fold1(f[t], f[L + t], r_out[j]);
                a.vx[j * L + t] = v;
                a.tail[len / 4 + t] = v;
                hb[t] = v;
            }
            __syncthreads();
            len += 4 * L;
        }
        if (t == 0) {
            ts[0] = len;
            ts[1] = hs;
        }
    }
    __syncthreads();
    if (t < 10) a.tstate[t] = ts[t];
    __syncthreads();
}

struct RoundParams {
    const uint32_t *fsrc;  // table k - 2 of f, or the layer input (rep stride 0) for the first two pairs
    uint64_t fsrc_rs;      // rep stride in u32
    uint32_t *fdst;        // table k of f (unused for the pair of rounds 0 and 1)
    uint64_t fdst_rs;
    const uint32_t *hsrc;  // phase 1: lane table hg; phase 2: scalar table s
    uint64_t hsrc_rs;
    uint32_t *hdst;
    uint64_t hdst_rs;
    const uint32_t *rprev; // r_{k-2} of rep 0 (r_{k-1} follows it); rep stride kChalStride
    uint32_t *partials;    // block-major: [rep][block][kPairVals L]
    uint32_t items;        // items of the pair: N_k / 4 (four rows of table k each)
    uint32_t R;
};

// Workgroup b of a pair launch of R * nb workgroups (nb a multiple of 8) -> (rep, tile): the
// R repetitions of a tile are consecutive workgroups of one XCD (workgroup b runs on XCD
// b mod 8), so the rows of the layer input that they share are read from HBM once.
__device__ __forceinline__ uint2 pair_block(uint32_t b, uint32_t R)
{
    const uint32_t idx = b >> 3;
    return make_uint2(idx % R, (idx / R) * 8 + (b & 7));
}

// Adds the NV x V canonical lane sums of a block of BS threads (V lanes per thread; the
// threads sharing t % (L / V) hold the same lanes) into out[v * ostride], v = q L + lane <
// NV L.  Every thread must call it.
template <int L, int BS, int NV, int V = 4>
__device__ __forceinline__ void block_reduce_vals(uint32_t (&acc)[NV][V], uint32_t *out, uint32_t ostride)
{
    constexpr uint32_t LG = L / V, W = BS / 64;
    __shared__ uint32_t red[W][NV * L];
#pragma unroll
    for (uint32_t off = LG; off < 64; off <<= 1) {
#pragma unroll
        for (int q = 0; q < NV; q++)
#pragma unroll
            for (int c = 0; c < V; c++) acc[q][c] = m31_add(acc[q][c], __shfl_xor(acc[q][c], off, 64));
    }
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane < LG) {
#pragma unroll
        for (int q = 0; q < NV; q++)
#pragma unroll
            for (int c = 0; c < V; c++) red[wave][q * L + V * lane + c] = acc[q][c];
    }
    __syncthreads();
    for (uint32_t v = threadIdx.x; v < NV * L; v += BS) {
        uint32_t x = 0;
#pragma unroll
        for (uint32_t w = 0; w < W; w++) x = m31_add(x, red[w][v]);
        out[v * ostride] = x;
    }
}

// Adds the block-major partials [nb][nvals] of k_pair (one repetition: nvals = kPairVals L)
// or k_dot (nvals = R 3L) into sums[nvals] (LDS), with NT threads.  nvals is a multiple of 4
// and at most kFinMaxVals; nb is at most kMaxBlocksPerRep.  Each block's partials are one
// contiguous, 16-byte aligned row, written whole by the block (a value-major layout would
// have every block write one word of lines that neighbouring blocks on other XCDs also
// write, and each XCD's L2 write its own partial copy of them back).  The sums are u64 (at
// most 256 canonical values: exact) and the group sums meet in LDS.  Every thread must call
// it.
//   NT = kFinThreads: thread t adds value t % nvals of rows t / nvals, t / nvals + G, ...
//     (G = NT / nvals groups, at least 7); all of a thread's loads are issued before the
//     first add, so the partials cost one memory latency.
//   smaller NT: a row is W = nvals / 4 16-byte words; G = NT / W groups of W threads read
//     rows grp, grp + G, ... with 16-byte loads, kFinBatch rows per thread in flight.
constexpr uint32_t kFinThreads = 1024;   // single-proof ends: one workgroup per repetition, alone
constexpr uint32_t kLsFinThreads = 256;  // batch ends: beside the other lanes' kernels (see k_ls_fin)
constexpr uint32_t kFinMaxVals = 144;
constexpr uint32_t kFinBatch = 16;
template <uint32_t NT>
__device__ __forceinline__ void reduce_partials(const uint32_t *__restrict__ partials, uint32_t nb, uint32_t nvals,
                                                uint32_t *sums)
{
    static_assert(NT >= kFinMaxVals, "a workgroup must cover one row of partials");
    if constexpr (NT >= kFinThreads) {
        constexpr uint32_t kGroupsMin = NT / kFinMaxVals;
        constexpr uint32_t kPerThread = (kMaxBlocksPerRep + kGroupsMin - 1) / kGroupsMin;
        __shared__ unsigned long long part1[NT];
        const uint32_t t = threadIdx.x, G = NT / nvals, v = t % nvals, grp = t / nvals;
        uint32_t x[kPerThread];
#pragma unroll
        for (uint32_t i = 0; i < kPerThread; i++) {
            const uint32_t b = grp + i * G;
            x[i] = (grp < G && b < nb) ? partials[(uint64_t)b * nvals + v] : 0u;
        }
        unsigned long long a = 0;
#pragma unroll
        for (uint32_t i = 0; i < kPerThread; i++) a += x[i];
        part1[t] = a;
        __syncthreads();
        if (t < nvals) {
            unsigned long long sum = 0;
            for (uint32_t g = 0; g < G; g++) sum += part1[g * nvals + t];
            sums[t] = m31_reduce64(sum);
        }
        __syncthreads();
    } else {
        __shared__ unsigned long long part[4 * NT];  // [group][value]: G nvals <= 4 NT
        const uint32_t t = threadIdx.x, W = nvals / 4, G = NT / W, grp = t / W, w = t % W;
        if (grp < G) {
            const uint4 *p = reinterpret_cast<const uint4 *>(partials) + w;
            unsigned long long a0 = 0, a1 = 0, a2 = 0, a3 = 0;
            for (uint32_t b0 = grp; b0 < nb; b0 += kFinBatch * G) {
                uint4 x[kFinBatch];
#pragma unroll
                for (uint32_t i = 0; i < kFinBatch; i++) {
                    const uint32_t b = b0 + i * G;
                    x[i] = b < nb ? p[(uint64_t)b * W] : make_uint4(0u, 0u, 0u, 0u);
                }
#pragma unroll
                for (uint32_t i = 0; i < kFinBatch; i++) {
                    a0 += x[i].x;
                    a1 += x[i].y;
                    a2 += x[i].z;
                    a3 += x[i].w;
                }
            }
            unsigned long long *q = part + grp * nvals + 4 * w;
            q[0] = a0;
            q[1] = a1;
            q[2] = a2;
            q[3] = a3;
        }
        __syncthreads();
        for (uint32_t v = t; v < nvals; v += NT) {
            unsigned long long sum = 0;
            for (uint32_t g = 0; g < G; g++) sum += part[g * nvals + v];
            sums[v] = m31_reduce64(sum);
        }
        __syncthreads();
    }
}

// Round k's three sums per lane from the nine product sums vals[q L + lane] of a round
// pair: u = P0 + P3, v = P1 + P4, and the sum at 2 = 2u + 2v - (P2 + P5).
template <int L>
__device__ __forceinline__ void pair_sums_first(const uint32_t *vals, uint32_t *sums)
{
    for (uint32_t c = threadIdx.x; c < L; c += blockDim.x) {
        const uint32_t *P = vals + c;  // P[q L] = product q of lane c
        const uint32_t u = m31_add(P[0], P[3 * L]), v = m31_add(P[L], P[4 * L]), w = m31_add(P[2 * L], P[5 * L]);
        sums[c] = u;                                                       // sum f0 h0
        sums[L + c] = v;                                                   // sum f1 h1
        sums[2 * L + c] = m31_sub(m31_add(m31_add(u, u), m31_add(v, v)), w);  // sum (f0 + f1)(h0 + h1)
    }
}

// Round k+1's sums of a round pair: each a quadratic in r = r_k over three products.
template <int L>
__device__ __forceinline__ void pair_sums_second(const uint32_t *vals, uint32_t r, uint32_t *sums)
{
    constexpr uint32_t Q = 3 * L;
    for (uint32_t q = threadIdx.x; q < Q; q += blockDim.x) {
        const uint32_t s = q / L, c = q % L;
        const uint32_t *P = vals + 3 * s * L + c;  // the three products of s_s
        const uint32_t c1 = m31_sub(P[L], m31_add(P[0], P[2 * L]));
        sums[q] = m31_add(P[0], m31_mul(r, m31_add(c1, m31_mul(r, P[2 * L]))));
    }
}

// One transcript step of repetition rep, run by one wave: the round's message (from sums,
// times vx in phase 2) goes to the proof tail at byte `base` and to hb (behind at most L
// pending words), then the bytes from `start` (below `base` only for the first message of
// rep 0 after other appends) to the end of the message are hashed.  Returns the challenge;
// d receives the digest.
template <int L>
__device__ __forceinline__ uint32_t rep_step(const uint32_t *sums, const TsArgs &ts, int phase, uint32_t rep,
                                             uint32_t base, uint32_t start, uint32_t *hb, uint32_t (&d)[8])
{
    constexpr uint32_t Q = 3 * L;
    const uint32_t lane = threadIdx.x & 63, lead = (base - start) / 4;
    for (uint32_t w = lane; w < Q; w += 64) {
        const uint32_t v = msg_word<L>(sums, ts.vx + rep * L, phase, w);
        ts.tail[base / 4 + w] = v;
        hb[L + w] = v;
    }
    if (lane < lead && lead <= L) hb[L - lead + lane] = ts.tail[start / 4 + lane];
    __threadfence_block();
    hash_words_wave(ts.hash_kind, lead <= L ? hb + L - lead : ts.tail + start / 4, lead + Q, d);
    return challenge_from(d);
}

// The end of a pair of grid rounds for repetition `rep` (one workgroup of NT threads per
// repetition): adds rep's partials (the nine products P0..P8 of pair_item per lane), forms
// round k's sums and draws r_k, evaluates round k+1's sums at r_k and draws r_{k+1}.  Before
// the last x round each repetition's challenge hashes only its own message (rep j's message
// of round k sits at len0 + 4 Q j, of round k+1 at len0 + 4 Q (R + j); rep 0 of round k also
// hashes the unhashed bytes from hs0 on), so the R workgroups run at once and read no
// transcript state: the host tracks len0 and hs0, and the workgroup of the last rep writes
// the state after round k+1.  ts.k is round k.
template <int L, uint32_t NT>
__device__ __forceinline__ void pair_fin_rep_body(const uint32_t *__restrict__ partials, uint32_t nb, const TsArgs &ts,
                                  uint32_t len0, uint32_t hs0, uint32_t rep)
{
    constexpr uint32_t Q = 3 * L, V = kPairVals * L;
    __shared__ uint32_t vals[V], sums[Q], hb[L + Q], rr;
    const uint32_t R = ts.R, t = threadIdx.x, wave = t >> 6, lane = t & 63;
    reduce_partials<NT>(partials + (uint64_t)rep * V * nb, nb, V, vals);
    pair_sums_first<L>(vals, sums);
    __syncthreads();
    uint32_t *chal = ts.chal + rep * kChalStride + (ts.phase == 2 ? 32u : 0u);
    uint32_t d[8] = {};
    for (uint32_t step = 0; step < 2; step++) {
        if (wave == 0) {
            const uint32_t base = len0 + step * 4 * Q * R + 4 * Q * rep;
            const uint32_t c = rep_step<L>(sums, ts, ts.phase, rep, base, (rep == 0 && step == 0) ? hs0 : base, hb, d);
            if (lane == 0) {
                chal[ts.k + step] = c;
                rr = c;
            }
        }
        __syncthreads();
        if (step == 0) {
            pair_sums_second<L>(vals, rr, sums);
            __syncthreads();
        }
    }
    if (rep == R - 1 && t == 0) {
        const uint32_t len = len0 + 2 * 4 * Q * R;
        ts.tstate[0] = len;
        ts.tstate[1] = len;
        for (int i = 0; i < 8; i++) ts.tstate[2 + i] = d[i];
    }
}

template <int L>
__global__ void __launch_bounds__(kFinThreads) k_pair_fin(const uint32_t *__restrict__ partials, uint32_t nb,
                                                          TsArgs ts, uint32_t len0, uint32_t hs0)
{
    latency_critical();
    pair_fin_rep_body<L, kFinThreads>(partials, nb, ts, len0, hs0, blockIdx.x);
}

// Rounds k0 .. n-1 of one phase in a single workgroup of 768 threads: the tables are
// small (at most kTailRows rows), so the launch and cross-block reduction cost of a grid
// would dominate.  Threads [256 j, 256 j + 256) work on rep j.
constexpr uint32_t kTailThreads = 768;  // 3 groups of 256: one per repetition
constexpr uint32_t kTailRows = 2048;
struct TailParams {
    const uint32_t *f0;       // round-k0 source of f (the layer input when k0 <= 1)
    uint64_t f0_rs;
    const uint32_t *h0;       // round-k0 source of h
    uint64_t h0_rs;
    uint32_t *fA, *fB, *hA, *hB;
    uint64_t fA_rs, fB_rs, hA_rs, hB_rs;
    uint32_t k0, n;
    uint32_t dbl;             // f0 / h0 are table k0 - 2 (after round pairs): round k0 folds twice
    TsArgs ts;
};

// Rounds p.k0 .. p.n - 1 of one phase, run by one workgroup (blockDim a multiple of 256,
// at least 256 R): group j = threads [256 j, 256 j + 256) works on rep j.
template <int L, bool PH2>
__device__ __forceinline__ void phase_rounds(const TailParams &p)
{
    constexpr uint32_t LG = L / 4, Q = 3 * L;
    __shared__ uint32_t red[4][4][3 * L];
    __shared__ uint32_t sums[3 * Q], rr[3], idle[3 * L];
    const uint32_t R = p.ts.R;
    const uint32_t grp = threadIdx.x >> 8, tg = threadIdx.x & 255, g = tg % LG;
    if (threadIdx.x < R) rr[threadIdx.x] = p.k0 > 0 ? p.ts.chal[threadIdx.x * kChalStride + (PH2 ? 32u : 0u) + p.k0 - 1] : 0u;
    __syncthreads();
    for (uint32_t k = p.k0; k < p.n; k++) {
        const uint64_t items = 1ull << (p.n - k - 1);
        // table k lives in A for odd k, B for even k >= 2; table k-1 is the source.  Round
        // k0 reads f0 / h0, and so does round 1 (table 0 is the layer input, hg0 or s0).
        const bool first = (k == p.k0) || (k == 1);
        const uint32_t *fs = first ? p.f0 : ((k - 1) & 1 ? p.fA : p.fB);
        const uint64_t fs_rs = first ? p.f0_rs : ((k - 1) & 1 ? p.fA_rs : p.fB_rs);
        const uint32_t *hs = first ? p.h0 : ((k - 1) & 1 ? p.hA : p.hB);
        const uint64_t hs_rs = first ? p.h0_rs : ((k - 1) & 1 ? p.hA_rs : p.hB_rs);
        uint32_t *fd = (k & 1) ? p.fA : p.fB;
        const uint64_t fd_rs = (k & 1) ? p.fA_rs : p.fB_rs;
        uint32_t *hd = (k & 1) ? p.hA : p.hB;
        const uint64_t hd_rs = (k & 1) ? p.hA_rs : p.hB_rs;
        uint64_t acc[3][4] = {};
        if (grp < R) {
            const uint32_t rep = grp;
            const uint32_t *f = fs + rep * fs_rs;
            const uint32_t *h = hs + rep * hs_rs;
            for (uint64_t w = tg; w < items * LG; w += 256) {
                if (k == 0) round_item<L, PH2, false>(f, nullptr, h, nullptr, w / LG, g, 0u, acc);
                else round_item<L, PH2, true>(f, fd + rep * fd_rs, h, hd + rep * hd_rs, w / LG, g, rr[rep], acc);
            }
        }
        group_reduce<L>(acc, red[grp], grp < R ? sums + grp * Q : idle);
        __syncthreads();
        TsArgs a = p.ts;
        a.k = k;
        a.last = (k == p.n - 1);
        a.fin = (k == 0) ? fs : fd;
        a.fin_rs = (k == 0) ? fs_rs : fd_rs;
        transcript_step<L>(sums, a, rr);
    }
}

template <int L, bool PH2>
__global__ void __launch_bounds__(kTailThreads) k_tail(TailParams p)
{
    latency_critical();
    phase_rounds<L, PH2>(p);
}

// Rounds k0 .. kend - 1 (all before the last round of the phase) with one workgroup of
// kRepThreads per repetition (blockIdx.x = rep).  Before the last x round the R challenge
// chains are independent: rep j's message of round k sits at base_k + 4 Q j, where
// base_k = len0 + (k - k0) 4 Q R, and its challenge hashes from there to the end of the
// message (rep 0 of round k0 also takes the unhashed bytes from hs0 on).  The host tracks
// len0 and hs0 (the proof layout is fixed by the circuit shape), so this kernel reads no
// transcript state; the workgroup of the last rep writes the final state.
constexpr uint32_t kRepThreads = 512;
constexpr int kTailV = 2;  // lanes per thread in the small-table passes
constexpr int kTailPreV = 4;  // in the first pass of k_tail_rep (table k0 from a table up to 4x larger)
struct TailRepParams {
    TailParams t;
    uint32_t kend;
    uint32_t len0, hs0;  // transcript length and hash start (bytes of the tail) at round k0
    // when not null: table k0 and its nine products were made by a grid pass (k_pair,
    // block-major partials [rep][pre_nb][kPairVals L]); the kernel only adds the partials
    const uint32_t *pre_part;
    uint32_t pre_nb;
};

// Adds the nine product sums of a wave's threads (V lanes per thread; the threads sharing
// t % (L / V) hold the same lanes) into out[q L + lane] (canonical).
template <int L, int V>
__device__ __forceinline__ void wave_reduce_pairs(uint32_t (&acc)[kPairVals][V], uint32_t *out)
{
    constexpr uint32_t LG = L / V;
#pragma unroll
    for (uint32_t off = LG; off < 64; off <<= 1)
#pragma unroll
        for (int q = 0; q < (int)kPairVals; q++)
#pragma unroll
            for (int c = 0; c < V; c++) acc[q][c] = m31_add(acc[q][c], __shfl_xor(acc[q][c], off, 64));
    const uint32_t lane = threadIdx.x & 63;
    if (lane < LG) {
#pragma unroll
        for (int q = 0; q < (int)kPairVals; q++)
#pragma unroll
            for (int c = 0; c < V; c++) out[q * L + V * lane + c] = acc[q][c];
    }
}

// The body of k_tail_rep for repetition `rep`; blockDim is NT (a multiple of 64, >= 128);
// k0 < kend <= n - 1.
// Each round's sums come from the nine products of a pass over the table's groups of four
// rows (the round-pair products above): round k's sums are their first combination, round
// k+1's a quadratic in r_k.  So the pass that makes table k+1 and its products (for round
// k+2) needs only r_k and runs on waves 1 .. NT/64 - 1 while wave 0 hashes round k+1's
// message, whose sums it evaluates from the previous pass at r_k; a round costs one hash
// and one barrier instead of a pass, a reduction and a hash in turn.  Before round k0 the
// whole workgroup makes table k0 (from the layer input, or folding table k0 - 1 once or
// table k0 - 2 twice) and its products, unless a grid pass made them already (pre_part:
// then it only adds that pass's partial rows).  The product sums of a pass live per wave in
// red[buffer][wave], two buffers used in turn (wave 0 reads one while the pass fills the
// other).  Table k (k >= 1) is stored in the A buffers for odd k, B for even k, where the
// last round (k_tail) finds table kend - 1.
template <int L, bool PH2, uint32_t NT = kRepThreads, int V = kTailV>
__device__ __forceinline__ void tail_rep_body(const TailRepParams &q, uint32_t rep)
{
    constexpr uint32_t LG = L / V, Q = 3 * L, NV = kPairVals * L, W = NT / 64;
    static_assert(NT % 64 == 0 && NT >= 128, "tail workgroup shape");
    const TailParams &p = q.t;
    __shared__ uint32_t red[2][W][NV];
    __shared__ uint32_t sums[Q], rr[2], hb[L + Q];
    const uint32_t R = p.ts.R, k0 = p.k0;
    const uint32_t t = threadIdx.x, wave = t >> 6, lane = t & 63;
    uint32_t *chal = p.ts.chal + rep * kChalStride + (PH2 ? 32u : 0u);
    const uint32_t *f0 = p.f0 + rep * p.f0_rs, *h0 = p.h0 + rep * p.h0_rs;
    auto tab_f = [&](uint32_t kk) { return ((kk & 1) ? p.fA : p.fB) + rep * ((kk & 1) ? p.fA_rs : p.fB_rs); };
    auto tab_h = [&](uint32_t kk) { return ((kk & 1) ? p.hA : p.hB) + rep * ((kk & 1) ? p.hA_rs : p.hB_rs); };
    if (q.pre_part) {
        // the products of table k0 from the grid pass: its few partial rows added (all
        // loads in flight at once) into slot 0; the other slots zero
        const uint32_t *pp = q.pre_part + (uint64_t)rep * q.pre_nb * NV;
        for (uint32_t v = t; v < W * NV; v += NT) {
            uint32_t x = 0;
            if (v < NV)
                for (uint32_t b = 0; b < q.pre_nb; b++) x = m31_add(x, pp[(uint64_t)b * NV + v]);
            red[0][v / NV][v % NV] = x;
        }
    } else {
        // table k0 and its products, with PV lanes per thread: one workgroup reads up to 16
        // rows per item of a table four times larger, so it keeps wider loads in flight
        constexpr int PV = NT >= kRepThreads ? kTailPreV : V;
        constexpr uint32_t PLG = L / PV;
        uint32_t acc[kPairVals][PV] = {};
        const uint64_t items = (1ull << (p.n - k0 - 2)) * PLG;
        if (k0 == 0) {
            for (uint64_t w = t; w < items; w += NT)
                pair_item<L, PH2, 0, PV>(f0, nullptr, h0, nullptr, w / PLG, (uint32_t)(w % PLG), 0u, 0u, acc);
        } else if (p.dbl) {
            const uint32_t ra = chal[k0 - 2], rb = chal[k0 - 1];
            for (uint64_t w = t; w < items; w += NT)
                pair_item<L, PH2, 2, PV>(f0, tab_f(k0), h0, tab_h(k0), w / PLG, (uint32_t)(w % PLG), ra, rb, acc);
        } else {
            const uint32_t ra = chal[k0 - 1];
            for (uint64_t w = t; w < items; w += NT)
                pair_item<L, PH2, 1, PV>(f0, tab_f(k0), h0, tab_h(k0), w / PLG, (uint32_t)(w % PLG), ra, 0u, acc);
        }
        wave_reduce_pairs<L, PV>(acc, red[0][wave]);
    }
    __syncthreads();
    uint32_t d[8] = {};
    for (uint32_t k = k0; k < q.kend; k++) {
        const uint32_t rprev = rr[(k - k0 + 1) & 1];  // r_{k-1}, for k > k0
        if (wave == 0) {
            // round k's sums from the products of table k0 (k = k0) or table k - 1 at r_{k-1}
            const uint32_t (*P)[NV] = red[(k - k0 + 1) & 1];
            if (k == k0) P = red[0];
            for (uint32_t x = lane; x < Q; x += 64) {
                const uint32_t sidx = x / L, c = x % L;
                uint32_t v[6];
#pragma unroll
                for (int j = 0; j < 6; j++) v[j] = 0u;
                const uint32_t q0 = k == k0 ? 0u : 3 * sidx;  // first: P0..P5; later: P(3s)..P(3s+2)
                const int nq = k == k0 ? 6 : 3;
                for (uint32_t w = 0; w < W; w++)
#pragma unroll
                    for (int j = 0; j < 6; j++)
                        if (j < nq) v[j] = m31_add(v[j], P[w][(q0 + j) * L + c]);
                uint32_t sv;
                if (k == k0) {
                    const uint32_t u = m31_add(v[0], v[3]), vv = m31_add(v[1], v[4]), ww = m31_add(v[2], v[5]);
                    sv = sidx == 0 ? u : sidx == 1 ? vv : m31_sub(m31_add(m31_add(u, u), m31_add(vv, vv)), ww);
                } else {
                    const uint32_t c1 = m31_sub(v[1], m31_add(v[0], v[2]));
                    sv = m31_add(v[0], m31_mul(rprev, m31_add(c1, m31_mul(rprev, v[2]))));
                }
                sums[x] = sv;
            }
            // wave 0 takes no part in the pass: its slot of the buffer the pass fills is zero
            if (k > k0)
                for (uint32_t x = lane; x < NV; x += 64) red[(k - k0) & 1][0][x] = 0u;
            __threadfence_block();
            const uint32_t base = q.len0 + (k - k0) * 4 * Q * R + 4 * Q * rep;
            const uint32_t start = (rep == 0 && k == k0) ? q.hs0 : base;
            const uint32_t c = rep_step<L>(sums, p.ts, PH2 ? 2 : 1, rep, base, start, hb, d);
            if (lane == 0) {
                chal[k] = c;
                rr[(k - k0) & 1] = c;  // the other slot may still be read by a pass wave
            }
        } else if (k > k0) {
            // table k from table k - 1 by r_{k-1}, and its products for round k + 1
            const uint32_t *fs = k == 1 ? f0 : tab_f(k - 1);
            const uint32_t *hs = k == 1 ? h0 : tab_h(k - 1);
            uint32_t acc[kPairVals][V] = {};
            const uint64_t items = (1ull << (p.n - k - 2)) * LG;
            for (uint64_t w = t - 64; w < items; w += NT - 64)
                pair_item<L, PH2, 1, V>(fs, tab_f(k), hs, tab_h(k), w / LG, (uint32_t)(w % LG), rprev, 0u, acc);
            if (k + 1 < q.kend) wave_reduce_pairs<L, V>(acc, red[(k - k0) & 1][wave]);
        }
        __syncthreads();
    }
    if (rep == R - 1 && t == 0) {
        const uint32_t len = q.len0 + (q.kend - k0) * 4 * Q * R;
        p.ts.tstate[0] = len;
        p.ts.tstate[1] = len;
        for (int i = 0; i < 8; i++) p.ts.tstate[2 + i] = d[i];
    }
}

template <int L, bool PH2>
__global__ void __launch_bounds__(kRepThreads) k_tail_rep(TailRepParams q)
{
    latency_critical();
    tail_rep_body<L, PH2>(q, blockIdx.x);
}

// Dot product sum_z eq[rep][z] * vals[z] (lanes): claimed_v = eval_multilinear(output, rz1)
// (include/LinearGKR/gkr.hpp:51-55).  Each block leaves one row of partials [block][rep][3L]
// (the p0 slot of the round layout; the other slots are zero) that k_dot_fin adds.
template <int L>
__global__ void __launch_bounds__(kThreads) k_dot(const uint32_t *__restrict__ vals, const uint32_t *__restrict__ eq,
                                                  uint32_t M, uint32_t *__restrict__ partials)
{
    constexpr uint32_t LG = L / 4, Q = 3 * L;
    __shared__ uint32_t red[4][3 * L];
    const uint32_t rep = blockIdx.y;
    const uint32_t g = threadIdx.x % LG;
    const uint32_t *e = eq + rep;  // interleaved eq table: rep j of entry z at 3 z + j
    uint64_t acc[3][4] = {};
    const uint64_t total = (uint64_t)M * LG;
    for (uint64_t w = (uint64_t)blockIdx.x * kThreads + threadIdx.x; w < total; w += (uint64_t)gridDim.x * kThreads) {
        const uint64_t z = w / LG;
        const uint4 t = mul4s(ld4(vals + z * L + 4 * g), e[kEqWords * z]);
        acc[0][0] += t.x; acc[0][1] += t.y; acc[0][2] += t.z; acc[0][3] += t.w;
    }
    group_reduce<L>(acc, red, partials + ((uint64_t)blockIdx.x * gridDim.y + rep) * Q, 1);
}

template <int L>
__global__ void __launch_bounds__(kFinThreads) k_dot_fin(const uint32_t *__restrict__ partials, uint32_t nb, uint32_t R,
                                                         uint32_t *__restrict__ claims)
{
    latency_critical();
    constexpr uint32_t Q = 3 * L;
    __shared__ uint32_t sums[3 * Q];
    reduce_partials<kFinThreads>(partials, nb, R * Q, sums);
    if (threadIdx.x < R * L) claims[threadIdx.x] = sums[(threadIdx.x / L) * Q + threadIdx.x % L];
}

// ------------------------------------------------------------------ batched proving
// A batch of independent proofs runs as ONE launch: workgroup b proves job order[b]
// from the forward pass to the last transcript byte, with no synchronization between
// workgroups (each proof's transcript is its own serial chain, so proofs only share the
// GPU).  Jobs are ordered largest first so the long proofs start early.  Every step reuses
// the block-level building blocks of the single-proof path.
constexpr uint32_t kBatchThreads = 768;  // 3 groups of 256: one per repetition
constexpr uint32_t kBatchMaxVars = 20;   // eq tables split 10 + 10 bits in LDS

struct BatchJob {
    uint32_t D;
    const LayerDev *layers;  // device array of D layers
    uint32_t *values;
    uint32_t *eqz, *eqx;     // interleaved eq tables (kEqWords words per entry)
    uint32_t *hg0, *hgA, *hgB, *fA, *fB, *s0, *sA, *sB;
    uint64_t hg0_rs, hgA_rs, hgB_rs, fA_rs, fB_rs, s0_rs, sA_rs, sB_rs;
    uint32_t *chal;          // (D + 1) * R * kChalStride
    uint32_t *ab;            // 2 (D + 1)
    uint32_t *vxy;           // 2 R L
    uint32_t *tstate;        // 16
    uint32_t *tail;          // proof bytes after the commitment
    uint32_t *claims;        // R L
    const uint32_t *seed;    // rz1 in slot layout [R * kChalStride], tstate [10], grind bytes [8]
    uint32_t *partials;      // R * kPairVals L * kMaxBlocksPerRep block partials (lockstep engine)
};

template <int L>
__device__ void block_forward(const uint32_t *__restrict__ vin, uint32_t *__restrict__ vout,
                              const uint32_t *__restrict__ ptr, const uint32_t *__restrict__ ent, uint32_t M)
{
    constexpr uint32_t LG = L / 4;
    for (uint64_t w = threadIdx.x; w < (uint64_t)M * LG; w += blockDim.x) forward_item<L>(vin, vout, ptr, ent, w);
}

// out[3 z + j] = f1 eq(r1_j, z) + f2 eq(r2_j, z) for z < 2^m (m <= 20), the interleaved
// layout of eq_block, rep by rep, from 2^10-entry low / high tables in LDS.  ab == nullptr:
// f1 = 1 and no second term.
__device__ void block_eq(uint32_t *__restrict__ out, const uint32_t *__restrict__ chal, uint32_t r1_off,
                         uint32_t r2_off, const uint32_t *__restrict__ ab, uint32_t m, uint32_t R)
{
    __shared__ uint32_t lo1[1024], lo2[1024], hi1[1024], hi2[1024];
    const bool two = ab != nullptr;
    const uint32_t h = m < 10 ? m : 10, hb = m - h;
    for (uint32_t j = 0; j < R; j++) {
        const uint32_t *r1 = chal + j * kChalStride + r1_off, *r2 = chal + j * kChalStride + r2_off;
        const uint32_t f1 = two ? ab[0] : 1u, f2 = two ? ab[1] : 0u;
        for (uint32_t t = threadIdx.x; t < (1u << h); t += blockDim.x) {
            uint32_t a = f1, b = f2;
            for (uint32_t i = 0; i < h; i++) {
                const uint32_t bit = (t >> i) & 1u;
                a = m31_mul(a, bit ? r1[i] : m31_sub(1u, r1[i]));
                if (two) b = m31_mul(b, bit ? r2[i] : m31_sub(1u, r2[i]));
            }
            lo1[t] = a;
            lo2[t] = b;
        }
        for (uint32_t t = threadIdx.x; t < (1u << hb); t += blockDim.x) {
            uint32_t a = 1u, b = 1u;
            for (uint32_t i = 0; i < hb; i++) {
                const uint32_t bit = (t >> i) & 1u;
                a = m31_mul(a, bit ? r1[h + i] : m31_sub(1u, r1[h + i]));
                if (two) b = m31_mul(b, bit ? r2[h + i] : m31_sub(1u, r2[h + i]));
            }
            hi1[t] = a;
            hi2[t] = b;
        }
        __syncthreads();
        const uint32_t mask = (1u << h) - 1u;
        for (uint32_t z = threadIdx.x; z < (1u << m); z += blockDim.x) {
            uint32_t v = m31_mul(lo1[z & mask], hi1[z >> h]);
            if (two) v = m31_add(v, m31_mul(lo2[z & mask], hi2[z >> h]));
            out[kEqWords * (uint64_t)z + j] = v;
        }
        __syncthreads();
    }
}

// claimed_v[j] = sum_z eq_j[z] * out[z], group j on rep j
template <int L>
__device__ void block_claims(const uint32_t *__restrict__ vals, const uint32_t *__restrict__ eq, uint32_t M, uint32_t R,
                             uint32_t *__restrict__ claims)
{
    constexpr uint32_t LG = L / 4, Q = 3 * L;
    __shared__ uint32_t red[4][4][3 * L];
    __shared__ uint32_t sums[3 * Q], idle[3 * L];
    const uint32_t grp = threadIdx.x >> 8, tg = threadIdx.x & 255, g = tg % LG;
    uint64_t acc[3][4] = {};
    if (grp < R) {
        const uint32_t *e = eq + grp;  // interleaved eq table
        for (uint64_t w = tg; w < (uint64_t)M * LG; w += 256) {
            const uint64_t z = w / LG;
            const uint4 t = mul4s(ld4(vals + z * L + 4 * g), e[kEqWords * z]);
            acc[0][0] += t.x; acc[0][1] += t.y; acc[0][2] += t.z; acc[0][3] += t.w;
        }
    }
    group_reduce<L>(acc, red[grp], grp < R ? sums + grp * Q : idle);
    __syncthreads();
    if (threadIdx.x < R * L) claims[threadIdx.x] = sums[(threadIdx.x / L) * Q + threadIdx.x % L];
    __syncthreads();
}


// A persistent grid over a device work queue: each workgroup takes the next slot of `order`
// (largest proofs first) from the queue head until the slots run out, and proves that job
// whole.  *queue must be 0 at launch.
template <int L>
__device__ void prove_batch_job(const BatchJob &J, uint32_t R, int hash_kind, unsigned long long *tacc);

template <int L>
__global__ void __launch_bounds__(kBatchThreads) k_prove_batch(const BatchJob *__restrict__ jobs,
                                                               const uint32_t *__restrict__ order, uint32_t nj,
                                                               uint32_t *queue, uint32_t R, int hash_kind)
{
    __shared__ unsigned long long tacc[tile_acc_words<L, kBatchThreads>()];  // g_x and h_y tiles
    __shared__ uint32_t slot;
    for (;;) {
        if (threadIdx.x == 0) slot = atomicAdd(queue, 1u);
        __syncthreads();
        const uint32_t q = slot;
        __syncthreads();  // every thread has read the slot before thread 0 takes the next
        if (q >= nj) return;
        const BatchJob J = jobs[order[q]];
        prove_batch_job<L>(J, R, hash_kind, tacc);
        __syncthreads();
    }
}

__global__ void k_zero_u32(uint32_t *p) { *p = 0u; }

// The transcript length each job's device transcript reached (tstate[0]), gathered so that
// one copy brings every job's back to the host.
__global__ void k_batch_tlen(const BatchJob *__restrict__ jobs, uint32_t nj, uint32_t *__restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nj) out[i] = jobs[i].tstate[0];
}

template <int L>
__device__ void prove_batch_job(const BatchJob &J, uint32_t R, int hash_kind, unsigned long long *tacc)
{
    const uint32_t D = J.D;
    auto chal_slot = [&](uint32_t l) { return J.chal + (uint64_t)l * R * kChalStride; };
    auto ab_slot = [&](uint32_t l) { return J.ab + 2ull * l; };
    // transcript seed from the host prefix: rz1 -> slot D, alpha = 1, beta = 0, state, grind bytes
    for (uint32_t t = threadIdx.x; t < R * kChalStride; t += blockDim.x) chal_slot(D)[t] = J.seed[t];
    if (threadIdx.x < 10) J.tstate[threadIdx.x] = J.seed[R * kChalStride + threadIdx.x];
    if (threadIdx.x < 8) J.tail[threadIdx.x] = J.seed[R * kChalStride + 10 + threadIdx.x];
    if (threadIdx.x == 0) {
        ab_slot(D)[0] = 1u;
        ab_slot(D)[1] = 0u;
    }
    // forward pass (Circuit::evaluate)
    for (uint32_t l = 0; l < D; l++) {
        const LayerDev &ly = J.layers[l];
        block_forward<L>(J.values + ly.v_off, J.values + ly.out_off, ly.fw_ptr, ly.fw_ent, 1u << ly.m);
        __syncthreads();
    }
    for (int l = (int)D - 1; l >= 0; l--) {
        const LayerDev &ly = J.layers[l];
        const uint32_t n = ly.n, m = ly.m;
        const uint32_t *V = J.values + ly.v_off;
        uint32_t *chal_l = chal_slot(l);
        block_eq(J.eqz, chal_slot(l + 1), 0u, 32u, ab_slot(l + 1), m, R);
        if (l == (int)D - 1) block_claims<L>(J.values + ly.out_off, J.eqz, 1u << m, R, J.claims);
        gx_tiles<L, kBatchThreads>(tacc, J.hg0, J.hg0_rs, V, J.eqz, ly.gx_ptr, ly.gx_ent, 1u << n, R, 0u, 1u,
                                   ((1u << n) + gx_rows<L, kBatchThreads>() - 1) / gx_rows<L, kBatchThreads>());
        __syncthreads();
        TsArgs ts{};
        ts.R = R; ts.chal = chal_l; ts.vx = J.vxy; ts.vy = J.vxy + R * L;
        ts.tstate = J.tstate; ts.tail = J.tail; ts.ab_out = ab_slot(l); ts.hash_kind = hash_kind;
        TailParams tp{};
        tp.f0 = V; tp.f0_rs = 0; tp.fA = J.fA; tp.fB = J.fB; tp.fA_rs = J.fA_rs; tp.fB_rs = J.fB_rs;
        tp.k0 = 0; tp.n = n;
        tp.h0 = J.hg0; tp.h0_rs = J.hg0_rs; tp.hA = J.hgA; tp.hB = J.hgB; tp.hA_rs = J.hgA_rs; tp.hB_rs = J.hgB_rs;
        tp.ts = ts;
        tp.ts.phase = 1;
        phase_rounds<L, false>(tp);
        block_eq(J.eqx, chal_l, 0u, 0u, nullptr, n, R);
        hy_tiles<kBatchThreads>(tacc, J.s0, J.s0_rs, J.eqz, J.eqx, ly.hy_ptr, ly.hy_ent, 1u << n, R, 0u, 1u,
                                ((1u << n) + tile_lanes(kBatchThreads) - 1) / tile_lanes(kBatchThreads));
        __syncthreads();
        tp.h0 = J.s0; tp.h0_rs = J.s0_rs; tp.hA = J.sA; tp.hB = J.sB; tp.hA_rs = J.sA_rs; tp.hB_rs = J.sB_rs;
        tp.ts.phase = 2;
        phase_rounds<L, true>(tp);
    }
}

// ------------------------------------------------------------------ lockstep batched proving
// The batch advances in steps.  Every proof is a fixed sequence of ops (its forward pass,
// then per layer: eqZ, g_x, the phase-1 rounds, eqX, h_y, the phase-2 rounds); step t runs
// op t of every proof that has one.  One launch of k_ls_work runs the wide ops of a step,
// each task on its own range of workgroups, so a wide proof gets as many workgroups as its
// tables need; then one launch of k_ls_fin runs the short ends (the partial sums and
// transcript step of a grid round, the claims).  The serial chains (the rounds of a phase
// whose tables are small, and the last round of a phase) run in k_ls_chain on a second
// stream, beside the next step's wide ops.  Launch boundaries and events order the steps,
// so no workgroup waits on another.  The schedule depends only on the circuit shapes: the
// host builds it and uploads it once.
enum LsOp : uint32_t {
    kOpSeed = 0,  // work, 1 block: transcript seed from the host prefix
    kOpFwd,       // work: forward pass of layer l
    kOpEqz,       // work, nb blocks: eqZ of layer l (every rep, interleaved)
    kOpGx,        // work: g_x of layer l
    kOpDot,       // work, R x nb blocks: claimed_v partials
    kOpPair,      // work, R x nb blocks: grid rounds k, k+1 of a phase
    kOpEqx,       // work, nb blocks: eqX of layer l (every rep, interleaved)
    kOpHy,        // work: h_y of layer l
    kOpPairFin,   // fin, 1 block: partials and transcript steps of grid rounds k, k+1
    kOpDotFin,    // fin, 1 block: claimed_v
    kOpTailRep,   // chain, R blocks: rounds k .. kend - 1 of a phase, one block per rep
    kOpLast,      // chain, 1 block: the last round of a phase
};

struct LsFwd {
    const uint32_t *vin;
    uint32_t *vout;
    const uint32_t *ptr, *ent;
    uint32_t M;
};
struct LsGx {
    uint32_t *hg;
    uint64_t hg_rs;
    const uint32_t *v, *eqz;
    const uint32_t *ptr, *ent;
    uint32_t N;
};
struct LsHy {
    uint32_t *s;
    uint64_t s_rs;
    const uint32_t *eqz, *eqx;
    const uint32_t *ptr, *ent;
    uint32_t N;
};
struct LsEq {
    uint32_t *out;            // interleaved (kEqWords words per entry)
    const uint32_t *chal;     // r1 / r2 of rep j at chal + j kChalStride + r1_off / r2_off
    uint32_t r1_off, r2_off;
    const uint32_t *ab;       // alpha, beta, or nullptr for a single eq table
    uint32_t m;
};
struct LsDot {
    const uint32_t *vals, *eq;
    uint32_t M;
    uint32_t *partials;
};

// One op of one proof in one step, with the pointers of a work op resolved on the host.
struct LsTask {
    uint32_t op, nb;     // nb: blocks per repetition (R x nb ops) or in total
    uint32_t job, l, phase;
    uint32_t k, kend;    // round (kOpPair: rounds k, k+1; kOpTailRep: rounds k .. kend - 1)
    uint32_t len0, hs0;  // kOpTailRep: transcript length and hash start at round k
    uint32_t dbl;        // kOpTailRep: the source is table k - 2 (round pairs ran before)
    union {
        LsFwd fwd;
        LsGx gx;
        LsHy hy;
        LsEq eq;
        LsDot dot;
        RoundParams rp;
    };
};

// Task of workgroup b in a step, and b's index within it.  start[i] is the first
// workgroup of task i, strictly increasing from start[0] = 0; ntasks <= 64 * 64.  Wave 0
// narrows the range with two 64-way ballots (two dependent loads).
__device__ __forceinline__ uint2 ls_find(const uint32_t *__restrict__ start, uint32_t ntasks, uint32_t b)
{
    __shared__ uint32_t found[2];
    if (threadIdx.x < 64) {
        const uint32_t lane = threadIdx.x;
        const uint32_t stride = (ntasks + 63) / 64;
        const uint32_t i = lane * stride;
        const uint32_t si = i < ntasks ? start[i] : 0xffffffffu;
        const uint64_t m1 = __ballot(si <= b);
        const uint32_t lo = (uint32_t)(__popcll(m1) - 1) * stride;
        const uint32_t j = lo + lane;
        const uint32_t sj = (lane < stride && j < ntasks) ? start[j] : 0xffffffffu;
        const uint64_t m2 = __ballot(sj <= b);
        const uint32_t pos = (uint32_t)__popcll(m2) - 1;
        const uint32_t s_at = __shfl(sj, (int)pos, 64);
        if (lane == 0) {
            found[0] = lo + pos;
            found[1] = b - s_at;
        }
    }
    __syncthreads();
    return make_uint2(__builtin_amdgcn_readfirstlane(found[0]), __builtin_amdgcn_readfirstlane(found[1]));
}

// Adds the 3 x 4 lane sums of a block of BS threads (lanes sharing t % LG) into
// out[q * ostride], q < 3L.  Every thread must call it.
template <int L, int BS>
__device__ __forceinline__ void block_reduce_to(uint64_t (&acc)[3][4], uint32_t *out, uint32_t ostride)
{
    constexpr uint32_t LG = L / 4, W = BS / 64;
    static_assert(BS >= 3 * L, "block too small for the round sums");
    __shared__ uint32_t red[W][3 * L];
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
        for (uint32_t w = 0; w < W; w++) s = m31_add(s, red[w][threadIdx.x]);
        out[threadIdx.x * ostride] = s;
    }
}

// Workgroup lb of a round-pair task (rep = lb / nb): k_pair on a block range, V lanes per
// thread.
template <int L, int BS, bool PH2, int MODE, int V>
__device__ __forceinline__ void ls_pair(const RoundParams &p, uint32_t nb, uint32_t lb)
{
    constexpr uint32_t LG = L / V;
    const uint2 rb_ = pair_block(lb, p.R);  // pair tasks start on multiples of 8 workgroups
    const uint32_t rep = rb_.x, bx = rb_.y, g = threadIdx.x % LG;
    const uint32_t *fs = p.fsrc + rep * p.fsrc_rs;
    const uint32_t *hs = p.hsrc + rep * p.hsrc_rs;
    uint32_t *fd = MODE ? p.fdst + rep * p.fdst_rs : nullptr;
    uint32_t *hd = MODE ? p.hdst + rep * p.hdst_rs : nullptr;
    const uint32_t ra = MODE ? p.rprev[(uint64_t)rep * kChalStride] : 0u;
    const uint32_t rb = MODE ? p.rprev[(uint64_t)rep * kChalStride + 1] : 0u;
    uint32_t acc[kPairVals][V] = {};
    const uint64_t total = (uint64_t)p.items * LG;
    if constexpr (MODE == 0) {
        // 64-bit product sums, folded every third item (as in k_pair)
        uint64_t a64[kPairVals][V] = {};
        uint32_t since = 0;
        for (uint64_t w = (uint64_t)bx * BS + threadIdx.x; w < total; w += (uint64_t)nb * BS) {
            pair_item<L, PH2, MODE, V, uint64_t>(fs, fd, hs, hd, w / LG, g, ra, rb, a64);
            if (++since == 3) {
                since = 0;
#pragma unroll
                for (int q = 0; q < kPairVals; q++)
#pragma unroll
                    for (int c = 0; c < V; c++) pair_fold(a64[q][c]);
            }
        }
#pragma unroll
        for (int q = 0; q < kPairVals; q++)
#pragma unroll
            for (int c = 0; c < V; c++) acc[q][c] = m31_reduce64(a64[q][c]);
    } else {
        for (uint64_t w = (uint64_t)bx * BS + threadIdx.x; w < total; w += (uint64_t)nb * BS)
            pair_item<L, PH2, MODE, V>(fs, fd, hs, hd, w / LG, g, ra, rb, acc);
    }
    block_reduce_vals<L, BS, kPairVals, V>(acc, p.partials + ((uint64_t)rep * nb + bx) * kPairVals * L, 1);
}

// Rounds k and k+1 of one phase for every repetition, on many blocks (pair_item): each
// block leaves its partial sums; k_pair_fin adds them and runs the two transcript steps.
// The launch boundary between the two orders the partials, so no device-scope fence is
// needed (on a multi-XCD part a release fence writes back the whole L2, once per block).
// Block size and lanes per thread by phase, measured on the 24 x 2^20 M31x16 proof: phase
// 1 (f and h both 16-lane rows) with 512 threads of one lane each (59 VGPRs, so 8 waves per
// SIMD hide the row gathers: the double-fold pass 31.1 -> 27.3 us, 0.54 -> 0.61 of HBM
// peak); phase 2 (h a scalar per row) with 256 threads of four lanes (one lane per thread
// was slower there: 1.82 -> 1.85 ms of pairs, 0.68 -> 0.75 ms of first pairs per proof).
template <int L, bool PH2, int MODE, int BS, int V>
__global__ void __launch_bounds__(BS) k_pair(RoundParams p)
{
    ls_pair<L, BS, PH2, MODE, V>(p, gridDim.x / p.R, blockIdx.x);
}
constexpr uint32_t kPairBS1 = 512, kPairV1 = 1, kPairBS2 = 256, kPairV2 = 4;

template <int L, int BS>
__global__ void __launch_bounds__(BS) k_ls_work(const BatchJob *__restrict__ jobs, const LsTask *__restrict__ tasks,
                                                const uint32_t *__restrict__ start, uint32_t ntasks, uint32_t R,
                                                uint32_t block_base)
{
    constexpr uint32_t LG = L / 4, Q = 3 * L;
    __shared__ unsigned long long tacc[tile_acc_words<L, BS>()];
    // block_base: the launch covers the step's workgroups from block_base on (a multiple of
    // 8, so that blockIdx.x and the step's workgroup index sit on the same XCD)
    const uint2 f = ls_find(start, ntasks, blockIdx.x + block_base);
    const LsTask &T = tasks[f.x];
    const uint32_t lb = f.y, t = threadIdx.x;
    switch (T.op) {
    case kOpSeed: {
        // rz1 -> chal slot D, alpha = 1, beta = 0, transcript state, grind bytes
        const BatchJob &J = jobs[T.job];
        uint32_t *slot = J.chal + (uint64_t)J.D * R * kChalStride;
        for (uint32_t i = t; i < R * kChalStride; i += BS) slot[i] = J.seed[i];
        if (t < 10) J.tstate[t] = J.seed[R * kChalStride + t];
        if (t < 8) J.tail[t] = J.seed[R * kChalStride + 10 + t];
        if (t == 0) {
            J.ab[2ull * J.D] = 1u;
            J.ab[2ull * J.D + 1] = 0u;
        }
        break;
    }
    case kOpFwd: {
        const LsFwd &a = T.fwd;
        forward_range<L, kLsGatherItems, BS>(a.vin, a.vout, a.ptr, a.ent, (uint64_t)a.M * LG, xcd_tile(lb, T.nb), T.nb);
        break;
    }
    case kOpEqz:
    case kOpEqx: {
        const LsEq &a = T.eq;
        eq_block<BS>(a.out, a.chal, a.r1_off, a.r2_off, a.ab, a.m, lb, R);
        break;
    }
    case kOpGx: {
        const LsGx &a = T.gx;
        constexpr uint32_t ROWS = gx_rows<L, BS>();
        const uint32_t tiles = (a.N + ROWS - 1) / ROWS;
        gx_tiles<L, BS>(tacc, a.hg, a.hg_rs, a.v, a.eqz, a.ptr, a.ent, a.N, R, xcd_tile(lb, T.nb), T.nb, tiles);
        break;
    }
    case kOpHy: {
        const LsHy &a = T.hy;
        constexpr uint32_t ROWS = tile_lanes(BS);
        const uint32_t tiles = (a.N + ROWS - 1) / ROWS;
        hy_tiles<BS>(tacc, a.s, a.s_rs, a.eqz, a.eqx, a.ptr, a.ent, a.N, R, xcd_tile(lb, T.nb), T.nb, tiles);
        break;
    }
    case kOpDot: {
        const LsDot &a = T.dot;
        const uint32_t rep = lb / T.nb, bx = lb % T.nb, g = t % LG;
        const uint32_t *e = a.eq + rep;  // interleaved eq table
        uint64_t acc[3][4] = {};
        for (uint64_t w = (uint64_t)bx * BS + t; w < (uint64_t)a.M * LG; w += (uint64_t)T.nb * BS) {
            const uint64_t z = w / LG;
            const uint4 v = mul4s(ld4(a.vals + z * L + 4 * g), e[kEqWords * z]);
            acc[0][0] += v.x; acc[0][1] += v.y; acc[0][2] += v.z; acc[0][3] += v.w;
        }
        block_reduce_to<L, BS>(acc, a.partials + ((uint64_t)bx * R + rep) * Q, 1);
        break;
    }
    default:
        break;
    }
}

// The round pairs of a step (their own kernel and stream: a pair item holds many more
// registers than the other ops, which would lower their occupancy), V lanes per thread.
template <int L, int BS, int V>
__global__ void __launch_bounds__(BS) k_ls_pair(const LsTask *__restrict__ tasks, const uint32_t *__restrict__ start,
                                                uint32_t ntasks)
{
    const uint2 f = ls_find(start, ntasks, blockIdx.x);
    const LsTask &T = tasks[f.x];
    if (T.phase == 1) {
        if (T.k == 0) ls_pair<L, BS, false, 0, V>(T.rp, T.nb, f.y);
        else ls_pair<L, BS, false, 2, V>(T.rp, T.nb, f.y);
    } else {
        if (T.k == 0) ls_pair<L, BS, true, 0, V>(T.rp, T.nb, f.y);
        else ls_pair<L, BS, true, 2, V>(T.rp, T.nb, f.y);
    }
}

// e0 / e1: when not null, the launch records the kernel's own start and end in them
template <int L, int BS>
void ls_launch_pair_v(uint32_t V, uint32_t blocks, hipStream_t st, const LsTask *tasks, const uint32_t *start,
                      uint32_t n, hipEvent_t e0, hipEvent_t e1)
{
    const dim3 g(blocks);
    auto go = [&](auto kernel) {
        if (e0) hipExtLaunchKernelGGL(kernel, g, dim3(BS), 0, st, e0, e1, 0, tasks, start, n);
        else hipLaunchKernelGGL(kernel, g, dim3(BS), 0, st, tasks, start, n);
    };
    switch (V) {
    case 1: go(k_ls_pair<L, BS, 1>); break;
    case 2: go(k_ls_pair<L, BS, 2>); break;
    default: go(k_ls_pair<L, BS, 4>); break;
    }
}

// k_ls_pair for block size BS in {64, 128, 256, 512} and V in {1, 2, 4} lanes per thread
template <int L>
void ls_launch_pair(uint32_t BS, uint32_t V, uint32_t blocks, hipStream_t st, const LsTask *tasks,
                    const uint32_t *start, uint32_t n, hipEvent_t e0, hipEvent_t e1)
{
    switch (BS) {
    case 64: ls_launch_pair_v<L, 64>(V, blocks, st, tasks, start, n, e0, e1); break;
    case 128: ls_launch_pair_v<L, 128>(V, blocks, st, tasks, start, n, e0, e1); break;
    case 512: ls_launch_pair_v<L, 512>(V, blocks, st, tasks, start, n, e0, e1); break;
    default: ls_launch_pair_v<L, 256>(V, blocks, st, tasks, start, n, e0, e1); break;
    }
}

template <int L>
__device__ __forceinline__ TsArgs ls_ts(const BatchJob &J, uint32_t l, uint32_t phase, uint32_t R, int hash_kind)
{
    TsArgs ts{};
    ts.R = R;
    ts.phase = (int)phase;
    ts.chal = J.chal + (uint64_t)l * R * kChalStride;
    ts.vx = J.vxy;
    ts.vy = J.vxy + R * L;
    ts.tstate = J.tstate;
    ts.tail = J.tail;
    ts.ab_out = J.ab + 2ull * l;
    ts.hash_kind = hash_kind;
    return ts;
}

// The small-table rounds of layer l, phase `phase`, from round k0 on (sources as in the
// single-proof path: table k - 1 feeds round k, table 0 is the layer input and hg0 / s0).
template <int L>
__device__ __forceinline__ TailParams ls_tail_params(const BatchJob &J, uint32_t l, uint32_t phase, uint32_t R, int hash_kind,
                                     uint32_t k0, uint32_t dbl)
{
    const LayerDev &ly = J.layers[l];
    const bool p2 = phase == 2;
    uint32_t *hA = p2 ? J.sA : J.hgA, *hB = p2 ? J.sB : J.hgB;
    const uint64_t hA_rs = p2 ? J.sA_rs : J.hgA_rs, hB_rs = p2 ? J.sB_rs : J.hgB_rs;
    TailParams tp{};
    tp.fA = J.fA; tp.fB = J.fB; tp.hA = hA; tp.hB = hB;
    tp.fA_rs = J.fA_rs; tp.fB_rs = J.fB_rs; tp.hA_rs = hA_rs; tp.hB_rs = hB_rs;
    tp.n = ly.n;
    tp.k0 = k0;
    tp.dbl = dbl;
    if (dbl) {
        // table k0 - 2: the layer input, or where the last round pair stored it (the A tables)
        tp.f0 = k0 == 2 ? J.values + ly.v_off : J.fA;
        tp.f0_rs = k0 == 2 ? 0 : J.fA_rs;
        tp.h0 = k0 == 2 ? (p2 ? J.s0 : J.hg0) : hA;
        tp.h0_rs = k0 == 2 ? (p2 ? J.s0_rs : J.hg0_rs) : hA_rs;
    } else if (k0 <= 1) {
        tp.f0 = J.values + ly.v_off;
        tp.f0_rs = 0;
        tp.h0 = p2 ? J.s0 : J.hg0;
        tp.h0_rs = p2 ? J.s0_rs : J.hg0_rs;
    } else {
        tp.f0 = (k0 - 1) & 1 ? J.fA : J.fB;
        tp.f0_rs = (k0 - 1) & 1 ? J.fA_rs : J.fB_rs;
        tp.h0 = (k0 - 1) & 1 ? hA : hB;
        tp.h0_rs = (k0 - 1) & 1 ? hA_rs : hB_rs;
    }
    tp.ts = ls_ts<L>(J, l, phase, R, hash_kind);
    return tp;
}

// The short ends of a step (one workgroup each): the partials and the two transcript
// steps after a pair of grid rounds, and the claims.  Workgroups of kLsFinThreads (one wave
// per SIMD): they start as soon as one slot per SIMD is free beside the other lanes'
// kernels, where a workgroup of 1 024 threads at this register count needs a whole CU.
template <int L>
__global__ void __launch_bounds__(kLsFinThreads) k_ls_fin(const BatchJob *__restrict__ jobs,
                                                        const LsTask *__restrict__ tasks,
                                                        const uint32_t *__restrict__ start, uint32_t ntasks,
                                                        uint32_t R, int hash_kind)
{
    latency_critical();
    constexpr uint32_t Q = 3 * L;
    __shared__ uint32_t sums[3 * Q];
    const uint2 f = ls_find(start, ntasks, blockIdx.x);
    const LsTask &T = tasks[f.x];
    const BatchJob &J = jobs[T.job];
    if (T.op == kOpPairFin) {
        TsArgs ts = ls_ts<L>(J, T.l, T.phase, R, hash_kind);
        ts.k = T.k;
        pair_fin_rep_body<L, kLsFinThreads>(J.partials, T.nb, ts, T.len0, T.hs0, f.y);  // one workgroup per rep
    } else {
        reduce_partials<kLsFinThreads>(J.partials, T.nb, R * Q, sums);
        if (threadIdx.x < R * L) J.claims[threadIdx.x] = sums[(threadIdx.x / L) * Q + threadIdx.x % L];
    }
}

// The last round of a phase (one pair per repetition, so no reduction across threads)
// and its transcript step: the ordered x-claim hashing of phase 1, or the layer end of
// phase 2.  The same work as phase_rounds from k0 = n - 1, for any blockDim of at least
// 64 R threads.
template <int L, bool PH2>
__device__ __forceinline__ void last_round_body(const TailParams &p)
{
    constexpr uint32_t LG = L / 4, Q = 3 * L;
    __shared__ uint32_t sums[3 * Q], rr[3];
    const uint32_t R = p.ts.R, k = p.n - 1, t = threadIdx.x;
    if (t < R) rr[t] = k > 0 ? p.ts.chal[t * kChalStride + (PH2 ? 32u : 0u) + k - 1] : 0u;
    __syncthreads();
    // round k's sources (table k - 1; table 0 is the layer input and h0) and destinations
    const uint32_t *fs = p.f0, *hs = p.h0;
    const uint64_t fs_rs = p.f0_rs, hs_rs = p.h0_rs;
    uint32_t *fd = (k & 1) ? p.fA : p.fB, *hd = (k & 1) ? p.hA : p.hB;
    const uint64_t fd_rs = (k & 1) ? p.fA_rs : p.fB_rs, hd_rs = (k & 1) ? p.hA_rs : p.hB_rs;
    if (t < R * LG) {
        const uint32_t rep = t / LG, g = t % LG;
        uint64_t acc[3][4] = {};
        if (k == 0) round_item<L, PH2, false>(fs + rep * fs_rs, nullptr, hs + rep * hs_rs, nullptr, 0, g, 0u, acc);
        else
            round_item<L, PH2, true>(fs + rep * fs_rs, fd + rep * fd_rs, hs + rep * hs_rs, hd + rep * hd_rs, 0, g,
                                     rr[rep], acc);
#pragma unroll
        for (int pk = 0; pk < 3; pk++)
#pragma unroll
            for (int c = 0; c < 4; c++) sums[rep * Q + pk * L + 4 * g + c] = m31_reduce64(acc[pk][c]);
    }
    __syncthreads();
    TsArgs a = p.ts;
    a.k = k;
    a.last = 1;
    a.fin = k == 0 ? fs : fd;
    a.fin_rs = k == 0 ? fs_rs : fd_rs;
    transcript_step<L>(sums, a, rr);
}

// The serial chains of a step: the small-table rounds of a phase (one workgroup per
// repetition) and the last round of a phase.  They run on a second stream, beside the
// wide ops of the next step of the other proofs; small workgroups let several chains share
// a CU.
constexpr uint32_t kChainThreads = 256;
template <int L>
__global__ void __launch_bounds__(kChainThreads) k_ls_chain(const BatchJob *__restrict__ jobs,
                                                            const LsTask *__restrict__ tasks,
                                                            const uint32_t *__restrict__ start, uint32_t ntasks,
                                                            uint32_t R, int hash_kind)
{
    latency_critical();
    const uint2 f = ls_find(start, ntasks, blockIdx.x);
    const LsTask &T = tasks[f.x];
    const BatchJob &J = jobs[T.job];
    if (T.op == kOpTailRep) {
        TailRepParams q{};
        q.t = ls_tail_params<L>(J, T.l, T.phase, R, hash_kind, T.k, T.dbl);
        q.kend = T.kend;
        q.len0 = T.len0;
        q.hs0 = T.hs0;
        if (T.phase == 1) tail_rep_body<L, false, kChainThreads>(q, f.y);
        else tail_rep_body<L, true, kChainThreads>(q, f.y);
    } else {
        const TailParams tp = ls_tail_params<L>(J, T.l, T.phase, R, hash_kind, J.layers[T.l].n - 1, 0u);
        if (T.phase == 1) last_round_body<L, false>(tp);
        else last_round_body<L, true>(tp);
    }
}

// Records a HIP event pair around one launch when profiling is on.
template <typename F>
inline void timed(Profiler *pf, int cls, hipStream_t st, F &&launch)
{
    if (pf) pf->begin(cls, st);
    launch();
    if (pf) pf->end(st);
}

// Launches kernel<<<grid, block, 0, st>>>(args); when class cls is timed, the launch records
// the kernel's own start and end in a profiler event pair.
template <typename K, typename... A>
void launch_timed(Profiler *pf, int cls, K kernel, dim3 grid, dim3 block, hipStream_t st, A... args)
{
    hipEvent_t e0, e1;
    if (pf && pf->kernel_events(cls, e0, e1)) hipExtLaunchKernelGGL(kernel, grid, block, 0, st, e0, e1, 0, args...);
    else hipLaunchKernelGGL(kernel, grid, block, 0, st, args...);
}

template <int L>
int forward_impl(const DevCircuit &dc, hipStream_t st, Profiler *pf)
{
    for (uint32_t l = 0; l < dc.D; l++) {
        const LayerDev &ly = dc.layers[l];
        const uint64_t M = 1ull << ly.m;
        const uint64_t work = M * (L / 4);
        const uint32_t blocks = (uint32_t)std::min<uint64_t>((work + kThreads * kFwdItems - 1) / (kThreads * kFwdItems), 65536);
        timed(pf, kKForward, st, [&] {
            hipLaunchKernelGGL(k_forward<L>, dim3(blocks), dim3(kThreads), 0, st, dc.values + ly.v_off,
                               dc.values + ly.out_off, ly.fw_ptr, ly.fw_ent, (uint32_t)M);
        });
    }
    GKRX_HIP_CHECK(hipGetLastError());
    return GKRX_OK;
}

// rows reserved per rep in a ping-pong table: at least 4 so that every rep base of a
// scalar table stays 16-byte aligned
inline uint64_t scratch_rows(uint64_t rows) { return std::max<uint64_t>(4, (rows + 3) & ~3ull); }

// blocks of kThreads covering `work` threads, at most cap (grid-stride loops take the rest)
uint32_t grid_cover(uint64_t work, uint64_t cap)
{
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((work + kThreads - 1) / kThreads, cap));
}

uint32_t grid_for(uint64_t work)
{
    uint64_t b = (work + kThreads - 1) / kThreads;
    if (b < 1) b = 1;
    return (uint32_t)std::min<uint64_t>(b, kMaxBlocksPerRep);
}

// ts_len, ts_hs: transcript length and hash start (bytes of the tail) when the device
// transcript starts; the host follows them through the fixed proof layout.
template <int L>
int prove_impl(const DevCircuit &dc, DevScratch &s, uint32_t R, int hash_kind, uint64_t ts_len, uint64_t ts_hs,
               hipStream_t st, Profiler *pf)
{
    constexpr uint32_t LG = L / 4, Q = 3 * L;
    const uint32_t D = dc.D;
    const uint64_t maxN = 1ull << dc.max_n;
    const uint64_t hg0_rs = maxN * L, hgA_rs = scratch_rows(maxN / 2) * L, hgB_rs = scratch_rows(maxN / 4) * L;
    const uint64_t fA_rs = scratch_rows(maxN / 2) * L, fB_rs = scratch_rows(maxN / 4) * L;
    const uint64_t s0_rs = scratch_rows(maxN), sA_rs = scratch_rows(maxN / 2), sB_rs = scratch_rows(maxN / 4);
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
            dim3 grid((uint32_t)((M + kEqChunkOne * kThreads - 1) / (kEqChunkOne * kThreads)));
            timed(pf, kKEq, st, [&] {
                hipLaunchKernelGGL(k_eq, grid, dim3(kThreads), 0, st, s.eqz, (const uint32_t *)chal_prev, 0u, 32u,
                                   (const uint32_t *)ab_slot(l + 1), m, R);
            });
        }
        if (l == (int)D - 1) {
            const uint32_t nbx = grid_for(M * LG);
            timed(pf, kKDot, st, [&] {
                hipLaunchKernelGGL(k_dot<L>, dim3(nbx, R), dim3(kThreads), 0, st, dc.values + ly.out_off, s.eqz,
                                   (uint32_t)M, s.partials);
                hipLaunchKernelGGL(k_dot_fin<L>, dim3(1), dim3(kFinThreads), 0, st, (const uint32_t *)s.partials, nbx,
                                   R, s.claims);
            });
        }
        {
            // a resident grid (six blocks per CU, limited by the LDS sums) walks the tiles with
            // the next tiles' loads in flight
            dim3 grid((uint32_t)std::min<uint64_t>((N + gx_rows<L, kThreads>() - 1) / gx_rows<L, kThreads>(), kGxBlocks));
            timed(pf, kKGx, st, [&] {
                hipLaunchKernelGGL(k_gx<L>, grid, dim3(kThreads), 0, st, s.hg0, hg0_rs, V, s.eqz, ly.gx_ptr,
                                   ly.gx_ent, (uint32_t)N, R);
            });
        }
        for (int phase = 1; phase <= 2; phase++) {
            if (phase == 2) {
                dim3 grid((uint32_t)((N + kEqChunkOne * kThreads - 1) / (kEqChunkOne * kThreads)));
                timed(pf, kKEq, st, [&] {
                    hipLaunchKernelGGL(k_eq, grid, dim3(kThreads), 0, st, s.eqx, (const uint32_t *)chal_l, 0u, 0u,
                                       (const uint32_t *)nullptr, n, R);
                });
                dim3 g2((uint32_t)std::min<uint64_t>((N + tile_lanes(kThreads) - 1) / tile_lanes(kThreads), 65536));
                timed(pf, kKHy, st, [&] {
                    hipLaunchKernelGGL(k_hy, g2, dim3(kThreads), 0, st, s.s0, s0_rs, s.eqz, s.eqx,
                                       ly.hy_ptr, ly.hy_ent, (uint32_t)N, R);
                });
            }
            uint32_t *hA = phase == 1 ? s.hgA : s.sA, *hB = phase == 1 ? s.hgB : s.sB;
            const uint32_t *h0 = phase == 1 ? s.hg0 : s.s0;
            const uint64_t hA_rs = phase == 1 ? hgA_rs : sA_rs, hB_rs = phase == 1 ? hgB_rs : sB_rs;
            const uint64_t h0_rs = phase == 1 ? hg0_rs : s0_rs;
            TsArgs ts{};
            ts.R = R; ts.phase = phase; ts.chal = chal_l; ts.vx = s.vxy; ts.vy = s.vxy + R * L;
            ts.tstate = s.tstate; ts.tail = s.tail; ts.ab_out = ab_slot(l); ts.hash_kind = hash_kind;
            // rounds [0, K) run on the whole GPU in pairs (K: the first round with at most
            // kTailRows rows, rounded up to even), the rest in k_tail_rep / k_tail.  Pair j stores table 2j (j >= 1); the tables
            // alternate so that the last one, table K - 2, is in the A buffers, where the first
            // small-table round (which folds it twice) reads it.
            uint32_t k0 = 0;
            while (k0 < n && (N >> k0) > kTailRows) k0++;
            const uint32_t K = (k0 + 1) & ~1u, J = K / 2;
            auto pair_in_A = [&](uint32_t j) { return ((J - 1 - j) & 1) == 0; };
            for (uint32_t j = 0; j < J; j++) {
                const uint32_t k = 2 * j;
                RoundParams rp{};
                rp.items = (uint32_t)(N >> (k + 2));
                rp.partials = s.partials;
                TsArgs tk = ts;
                tk.k = k;
                tk.last = 0;
                if (j <= 1) {
                    rp.fsrc = V; rp.fsrc_rs = 0;
                    rp.hsrc = h0; rp.hsrc_rs = h0_rs;
                } else {
                    const bool a = pair_in_A(j - 1);
                    rp.fsrc = a ? s.fA : s.fB; rp.fsrc_rs = a ? fA_rs : fB_rs;
                    rp.hsrc = a ? hA : hB; rp.hsrc_rs = a ? hA_rs : hB_rs;
                }
                if (j >= 1) {
                    const bool a = pair_in_A(j);
                    rp.fdst = a ? s.fA : s.fB; rp.fdst_rs = a ? fA_rs : fB_rs;
                    rp.hdst = a ? hA : hB; rp.hdst_rs = a ? hA_rs : hB_rs;
                    rp.rprev = chal_l + (phase == 2 ? 32 : 0) + (k - 2);
                }
                const uint32_t pbs = phase == 1 ? kPairBS1 : kPairBS2, pv = phase == 1 ? kPairV1 : kPairV2;
                const uint64_t pw = (uint64_t)rp.items * (L / pv);  // work items: (item, lane group)
                const uint32_t nbx = (uint32_t)((std::min<uint64_t>(std::max<uint64_t>(1, (pw + pbs - 1) / pbs),
                                                                    kMaxBlocksPerRep) + 7) & ~7ull);  // a multiple of 8
                rp.R = R;
                dim3 grid(nbx * R);
                const int cls = phase == 1 ? (j == 0 ? kKPairPh1First : kKPairPh1) : (j == 0 ? kKPairPh2First : kKPairPh2);
                const dim3 blk(pbs);
                if (phase == 1) {
                    if (j == 0) launch_timed(pf, cls, k_pair<L, false, 0, kPairBS1, kPairV1>, grid, blk, st, rp);
                    else launch_timed(pf, cls, k_pair<L, false, 2, kPairBS1, kPairV1>, grid, blk, st, rp);
                } else {
                    if (j == 0) launch_timed(pf, cls, k_pair<L, true, 0, kPairBS2, kPairV2>, grid, blk, st, rp);
                    else launch_timed(pf, cls, k_pair<L, true, 2, kPairBS2, kPairV2>, grid, blk, st, rp);
                }
                timed(pf, kKPairFin, st, [&] {
                    hipLaunchKernelGGL(k_pair_fin<L>, dim3(R), dim3(kFinThreads), 0, st, (const uint32_t *)s.partials,
                                       nbx, tk, (uint32_t)ts_len, (uint32_t)ts_hs);
                });
                ts_len += 2ull * R * 4 * Q;
                ts_hs = ts_len;
            }
            k0 = K;
            // round kk's sources: table kk - 1 (table 0 = the layer input V and h0)
            TailParams tp{};
            tp.fA = s.fA; tp.fB = s.fB; tp.hA = hA; tp.hB = hB;
            tp.fA_rs = fA_rs; tp.fB_rs = fB_rs; tp.hA_rs = hA_rs; tp.hB_rs = hB_rs;
            tp.n = n; tp.ts = ts;
            auto set_src = [&](TailParams &t, uint32_t kk) {
                t.k0 = kk;
                if (kk <= 1) {
                    t.f0 = V; t.f0_rs = 0; t.h0 = h0; t.h0_rs = h0_rs;
                } else {
                    t.f0 = (kk - 1) & 1 ? s.fA : s.fB;
                    t.f0_rs = (kk - 1) & 1 ? fA_rs : fB_rs;
                    t.h0 = (kk - 1) & 1 ? hA : hB;
                    t.h0_rs = (kk - 1) & 1 ? hA_rs : hB_rs;
                }
            };
            const uint32_t klast = n - 1;
            if (k0 < klast) {
                TailRepParams q{};
                set_src(q.t, k0);
                if (K >= 2) {
                    // table K - 2: the layer input, or the A buffers of the last pair
                    q.t.dbl = 1;
                    q.t.f0 = K == 2 ? V : s.fA;
                    q.t.f0_rs = K == 2 ? 0 : fA_rs;
                    q.t.h0 = K == 2 ? h0 : hA;
                    q.t.h0_rs = K == 2 ? h0_rs : hA_rs;
                }
                q.t.fA = s.fA; q.t.fB = s.fB; q.t.hA = hA; q.t.hB = hB;
                q.t.fA_rs = fA_rs; q.t.fB_rs = fB_rs; q.t.hA_rs = hA_rs; q.t.hB_rs = hB_rs;
                q.t.n = n; q.t.ts = ts;
                q.kend = klast;
                q.len0 = (uint32_t)ts_len;
                q.hs0 = (uint32_t)ts_hs;
                const int cls = phase == 1 ? kKTailPh1 : kKTailPh2;
                if (K >= 2) {
                    // table K (double-folded from table K - 2) and its nine products on the
                    // whole GPU: the round-pair pass of rounds K, K + 1 without its transcript
                    // end; k_tail_rep adds its partials instead of making the table in one
                    // workgroup per repetition (19 us there, a memory latency per row group)
                    RoundParams rp{};
                    rp.items = (uint32_t)(N >> (K + 2));
                    rp.partials = s.partials;
                    rp.fsrc = q.t.f0; rp.fsrc_rs = q.t.f0_rs;
                    rp.hsrc = q.t.h0; rp.hsrc_rs = q.t.h0_rs;
                    rp.fdst = s.fB; rp.fdst_rs = fB_rs;  // table K (even): the B buffers
                    rp.hdst = hB; rp.hdst_rs = hB_rs;
                    rp.rprev = chal_l + (phase == 2 ? 32 : 0) + (K - 2);
                    rp.R = R;
                    const uint32_t pbs = phase == 1 ? kPairBS1 : kPairBS2, pv = phase == 1 ? kPairV1 : kPairV2;
                    const uint64_t pw = (uint64_t)rp.items * (L / pv);
                    const uint32_t nbx = (uint32_t)((std::min<uint64_t>(std::max<uint64_t>(1, (pw + pbs - 1) / pbs),
                                                                        kMaxBlocksPerRep) + 7) & ~7ull);
                    if (phase == 1) launch_timed(pf, cls, k_pair<L, false, 2, kPairBS1, kPairV1>, dim3(nbx * R), dim3(pbs), st, rp);
                    else launch_timed(pf, cls, k_pair<L, true, 2, kPairBS2, kPairV2>, dim3(nbx * R), dim3(pbs), st, rp);
                    q.pre_part = s.partials;
                    q.pre_nb = nbx;
                }
                timed(pf, cls, st, [&] {
                    if (phase == 1) hipLaunchKernelGGL((k_tail_rep<L, false>), dim3(R), dim3(kRepThreads), 0, st, q);
                    else hipLaunchKernelGGL((k_tail_rep<L, true>), dim3(R), dim3(kRepThreads), 0, st, q);
                });
                ts_len += (uint64_t)(klast - k0) * R * 4 * Q;
                ts_hs = ts_len;
            }
            // the last round of the phase: ordered x-claim hashing (phase 1) or the layer end
            // (phase 2), one workgroup for all reps
            set_src(tp, klast);
            timed(pf, phase == 1 ? kKTailPh1 : kKTailPh2, st, [&] {
                if (phase == 1) hipLaunchKernelGGL((k_tail<L, false>), dim3(1), dim3(kTailThreads), 0, st, tp);
                else hipLaunchKernelGGL((k_tail<L, true>), dim3(1), dim3(kTailThreads), 0, st, tp);
            });
            if (phase == 1) {
                ts_len += (uint64_t)R * (4 * Q + 4 * L);
                ts_hs = ts_len - 4 * L;
            } else {
                ts_len += (uint64_t)R * (4 * Q + 4 * L);
                ts_hs = ts_len;
            }
        }
    }
    GKRX_HIP_CHECK(hipGetLastError());
    return GKRX_OK;
}

}  // namespace

// ------------------------------------------------------------------ host side
int dev_forward(const DevCircuit &dc, hipStream_t st, Profiler *pf)
{
    if (dc.L == 8) return forward_impl<8>(dc, st, pf);
    if (dc.L == 16) return forward_impl<16>(dc, st, pf);
    return GKRX_ERR_CONFIG;
}

int dev_prove(const DevCircuit &dc, DevScratch &s, uint32_t R, int hash_kind, const uint32_t *rz1_host,
              const uint32_t *tstate_host, const uint8_t *tail_prefix, uint32_t tail_prefix_len, hipStream_t st,
              Profiler *pf, bool graph)
{
    if (R > 3) return GKRX_ERR_CONFIG;
    const uint32_t D = dc.D;
    const uint32_t mlast = dc.layers[D - 1].m;
    // seed slot D: rx = rz1, ry = 0 (rz2), alpha = 1, beta = 0; staged in pinned memory
    // because the copies complete after this function returns
    const uint64_t seed_u32 = (uint64_t)R * kChalStride;
    const uint64_t need = seed_u32 + 2 + 10 + (tail_prefix_len + 3) / 4;
    if (need > s.stage_u32) return GKRX_ERR_ARG;
    GKRX_HIP_CHECK(hipStreamSynchronize(st));  // the previous prove may still read the stage
    uint32_t *seed = s.stage, *ab_seed = seed + seed_u32, *ts = ab_seed + 2, *pre = ts + 10;
    memset(seed, 0, seed_u32 * 4);
    for (uint32_t j = 0; j < R; j++)
        for (uint32_t i = 0; i < mlast; i++) seed[(size_t)j * kChalStride + i] = rz1_host[(size_t)j * mlast + i];
    ab_seed[0] = 1u;
    ab_seed[1] = 0u;
    memcpy(ts, tstate_host, 40);
    if (tail_prefix_len) memcpy(pre, tail_prefix, tail_prefix_len);
    // the copies from the pinned stage and every launch; in a graph the copies read the stage
    // when the graph runs, so a replay picks up the seed written above
    auto enqueue = [&]() -> int {
        GKRX_HIP_CHECK(hipMemcpyAsync(s.chal + (uint64_t)D * R * kChalStride, seed, seed_u32 * 4,
                                      hipMemcpyHostToDevice, st));
        GKRX_HIP_CHECK(hipMemcpyAsync(s.ab + 2ull * D, ab_seed, 8, hipMemcpyHostToDevice, st));
        GKRX_HIP_CHECK(hipMemcpyAsync(s.tstate, ts, 40, hipMemcpyHostToDevice, st));
        if (tail_prefix_len)
            GKRX_HIP_CHECK(hipMemcpyAsync(s.tail, pre, tail_prefix_len, hipMemcpyHostToDevice, st));
        if (dc.L == 8) return prove_impl<8>(dc, s, R, hash_kind, tstate_host[0], tstate_host[1], st, pf);
        if (dc.L == 16) return prove_impl<16>(dc, s, R, hash_kind, tstate_host[0], tstate_host[1], st, pf);
        return GKRX_ERR_CONFIG;
    };
    if (!graph || pf) return enqueue();
    const bool same = s.graph && s.graph_serial == dc.serial && s.graph_R == R && s.graph_kind == hash_kind &&
                      s.graph_len == tstate_host[0] && s.graph_hs == tstate_host[1] && s.graph_pre == tail_prefix_len;
    if (!same) {
        // capture the sequence once (nothing runs while capturing), then keep the executable
        // graph; everything it touches (scratch, circuit tables, stage) stays at its address
        // until the scratch or the circuit changes, which changes the key
        dev_drop_graph(s);
        GKRX_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue();
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(st, &g);
        if (rc != GKRX_OK || ec != hipSuccess || !g) {
            if (g) (void)hipGraphDestroy(g);
            if (rc != GKRX_OK) return rc;
            set_error(std::string("graph capture: ") + hipGetErrorString(ec));
            return GKRX_ERR_DEVICE;
        }
        hipGraphExec_t x = nullptr;
        const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess || !x) {
            set_error(std::string("graph instantiate: ") + hipGetErrorString(ei));
            return GKRX_ERR_DEVICE;
        }
        s.graph = x;
        s.graph_serial = dc.serial;
        s.graph_R = R;
        s.graph_kind = hash_kind;
        s.graph_len = tstate_host[0];
        s.graph_hs = tstate_host[1];
        s.graph_pre = tail_prefix_len;
    }
    GKRX_HIP_CHECK(hipGraphLaunch(s.graph, st));
    return GKRX_OK;
}

void dev_drop_graph(DevScratch &s)
{
    if (s.graph) (void)hipGraphExecDestroy(s.graph);
    s.graph = nullptr;
    s.graph_serial = 0;
}

// CSR for one key: counting sort of entries by key.
static void build_csr(uint32_t nrows, const std::vector<uint32_t> &keys, const std::vector<uint32_t> &ent, uint32_t W,
                      std::vector<uint32_t> &ptr, std::vector<uint32_t> &out)
{
    ptr.assign((size_t)nrows + 1, 0);
    for (uint32_t k : keys) ptr[k + 1]++;
    for (uint32_t i = 0; i < nrows; i++) ptr[i + 1] += ptr[i];
    out.resize(ent.size());
    std::vector<uint32_t> pos(ptr.begin(), ptr.end() - 1);
    for (size_t e = 0; e < keys.size(); e++) {
        const uint32_t at = pos[keys[e]]++;
        for (uint32_t w = 0; w < W; w++) out[W * (size_t)at + w] = ent[W * e + w];
    }
}

int dev_upload_circuit(DevCircuit &dc, uint32_t L, const std::vector<uint32_t> &n, const std::vector<uint32_t> &m,
                       const std::vector<std::vector<uint32_t>> &mul, const std::vector<std::vector<uint32_t>> &add)
{
    static std::atomic<uint64_t> serials{0};
    dc.serial = ++serials;
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
        kf.reserve(U + A); ef.reserve(kFwWords * (U + A));
        kg.reserve(U + A); eg.reserve(kGxWords * (U + A));
        kh.reserve(U); eh.reserve(kHyWords * U);
        for (size_t g = 0; g < U; g++) {
            const uint32_t *q = &mul[l][4 * g];  // in0, in1, out, coef
            kf.push_back(q[2]); ef.insert(ef.end(), {q[0], q[1], q[3]});
            kg.push_back(q[0]); eg.insert(eg.end(), {q[1], q[2], q[3], q[0]});
            kh.push_back(q[1]); eh.insert(eh.end(), {q[0], q[2], q[3], q[1]});
        }
        for (size_t g = 0; g < A; g++) {
            const uint32_t *q = &add[l][3 * g];  // in0, out, coef
            kf.push_back(q[1]); ef.insert(ef.end(), {q[0], kAddMark, q[2]});
            kg.push_back(q[0]); eg.insert(eg.end(), {kAddMark, q[1], q[2], q[0]});
        }
        build_csr(1u << ly.m, kf, ef, kFwWords, host[6 * l + 0], host[6 * l + 1]);
        build_csr(1u << ly.n, kg, eg, kGxWords, host[6 * l + 2], host[6 * l + 3]);
        build_csr(1u << ly.n, kh, eh, kHyWords, host[6 * l + 4], host[6 * l + 5]);
    }
    dc.layers[dc.D - 1].out_off = off;
    for (uint32_t l = 0; l + 1 < dc.D; l++) dc.layers[l].out_off = dc.layers[l + 1].v_off;
    off += (1ull << dc.layers[dc.D - 1].m) * L;
    dc.values_u32 = off;
    uint64_t pool = 0;
    for (auto &v : host) pool += (v.size() + 3) & ~3ull;  // 16-byte aligned pieces
    dc.csr_u32 = std::max<uint64_t>(pool, 4);
    GKRX_HIP_CHECK(hipMalloc(&dc.csr_pool, dc.csr_u32 * 4));
    GKRX_HIP_CHECK(hipMalloc(&dc.d_layers, sizeof(LayerDev) * dc.D));
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
    GKRX_HIP_CHECK(hipMemcpy(dc.d_layers, dc.layers.data(), sizeof(LayerDev) * dc.D, hipMemcpyHostToDevice));
    return GKRX_OK;
}

void dev_free_circuit(DevCircuit &dc)
{
    if (dc.csr_pool) (void)hipFree(dc.csr_pool);
    if (dc.values) (void)hipFree(dc.values);
    if (dc.d_layers) (void)hipFree(dc.d_layers);
    dc.csr_pool = nullptr;
    dc.values = nullptr;
    dc.d_layers = nullptr;
}

namespace {

// Carves the prover tables of circuits up to dc's shape out of `base` (nullptr: only
// measures) and returns their size in u32; s's table pointers are set when base is given.
uint64_t scratch_layout(DevScratch &s, const DevCircuit &dc, uint32_t R, uint64_t tail_bytes, uint32_t *base)
{
    const uint32_t L = dc.L;
    const uint64_t maxN = 1ull << dc.max_n, maxM = 1ull << dc.max_m;
    const uint64_t N2 = scratch_rows(maxN / 2), N4 = scratch_rows(maxN / 4), N1 = scratch_rows(maxN);
    struct Piece { uint32_t **p; uint64_t u32; };
    Piece pieces[] = {
        {&s.eqz, kEqWords * maxM}, {&s.eqx, kEqWords * maxN}, {&s.hg0, R * maxN * L}, {&s.hgA, R * N2 * L},
        {&s.hgB, R * N4 * L}, {&s.fA, R * N2 * L}, {&s.fB, R * N4 * L}, {&s.s0, R * N1},
        {&s.sA, R * N2}, {&s.sB, R * N4}, {&s.partials, (uint64_t)R * kMaxBlocksPerRep * kPairVals * L},
        {&s.chal, (uint64_t)(dc.D + 1) * R * kChalStride}, {&s.ab, 2ull * (dc.D + 1)}, {&s.claims, R * L},
        {&s.vxy, 2ull * R * L}, {&s.tstate, 16}, {&s.tail, (tail_bytes + 3) / 4 + 64}};
    uint64_t at = 0;
    for (auto &pc : pieces) {
        if (base) *pc.p = base + at;
        at += (pc.u32 + 63) & ~63ull;  // 256-byte aligned pieces
    }
    if (base) {
        s.bytes = at * 4;
        s.max_n = dc.max_n;
        s.max_m = dc.max_m;
        s.L = L;
        s.D = dc.D;
        s.tail_cap = tail_bytes;
    }
    return at;
}

}  // namespace

int dev_ensure_scratch(DevScratch &s, const DevCircuit &dc, uint32_t R, uint64_t tail_bytes)
{
    if (s.pool && s.max_n >= dc.max_n && s.max_m >= dc.max_m && s.L == dc.L && s.D >= dc.D &&
        s.tail_cap >= tail_bytes)
        return GKRX_OK;
    dev_free_scratch(s);
    const uint64_t total = scratch_layout(s, dc, R, tail_bytes, nullptr);
    GKRX_HIP_CHECK(hipMalloc(&s.pool, total * 4));
    scratch_layout(s, dc, R, tail_bytes, reinterpret_cast<uint32_t *>(s.pool));
    s.stage_u32 = 1024;
    GKRX_HIP_CHECK(hipHostMalloc(&s.stage, s.stage_u32 * 4, hipHostMallocDefault));
    return GKRX_OK;
}

// ------------------------------------------------------------------ allocation-free prove
// The prover tables of gkrx_prove_noalloc live in this file-scope device array (each GPU
// the library runs on holds its own copy until the process ends) instead of memory the
// library allocates.  Host code never uses the array's name as an address: dev_noalloc_init
// looks up the array's device address on each device once (hipGetSymbolAddress, with the
// device current) and keeps it; kernels get that address.  Every call on a GPU shares the
// GPU's copy, so calls that overlap on one GPU race.
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace {
constexpr int kNoAllocDevices = 64;
std::mutex g_noalloc_mu;
uint32_t *g_noalloc_ptr[kNoAllocDevices];

// The transcript seed from the host prefix as a kernel argument: rz1 of every repetition
// (R * mlast words, repetition-major), the transcript state and the 32 grind bytes.
struct NoAllocSeed {
    uint32_t rz1[3 * kMaxVars];
    uint32_t tstate[10];
    uint32_t grind[8];
    uint32_t R, mlast;
};

// Writes the seed slot D (rz1 in the slot layout, zero elsewhere), alpha = 1, beta = 0, the
// transcript state and the grind bytes at the start of the tail: every scratch word a
// prove reads before writing it.
__global__ void k_install_seed(NoAllocSeed a, uint32_t *__restrict__ chal_slot, uint32_t *__restrict__ ab,
                               uint32_t *__restrict__ tstate, uint32_t *__restrict__ tail)
{
    const uint32_t t = threadIdx.x;
    for (uint32_t i = t; i < a.R * kChalStride; i += blockDim.x) {
        const uint32_t j = i / kChalStride, v = i % kChalStride;
        chal_slot[i] = v < a.mlast ? a.rz1[j * a.mlast + v] : 0u;
    }
    if (t < 10) tstate[t] = a.tstate[t];
    if (t < 8) tail[t] = a.grind[t];
    if (t == 0) {
        ab[0] = 1u;
        ab[1] = 0u;
    }
}
}  // namespace

int dev_noalloc_init(int device)
{
    if (device < 0 || device >= kNoAllocDevices) {
        set_error("device ordinal outside the allocation-free prover's table");
        return GKRX_ERR_DEVICE;
    }
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return GKRX_OK;
    void *p = nullptr;
    GKRX_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
        set_error("hipGetSymbolAddress gave no address for the scratch array");
        return GKRX_ERR_DEVICE;
    }
    g_noalloc_ptr[device] = reinterpret_cast<uint32_t *>(p);
    return GKRX_OK;
}

uint64_t dev_noalloc_bytes(const DevCircuit &dc, uint32_t R, uint64_t tail_bytes)
{
    DevScratch s;
    return 4 * scratch_layout(s, dc, R, tail_bytes, nullptr);
}

int dev_prove_noalloc(const DevCircuit &dc, uint32_t R, int hash_kind, const uint32_t *rz1_host,
                      const uint32_t *tstate_host, const uint8_t *grind, uint64_t tail_bytes, int device,
                      hipStream_t st, DevScratch &view)
{
    if (R > 3 || (dc.L != 8 && dc.L != 16)) return GKRX_ERR_CONFIG;
    uint32_t *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_noalloc_mu);
        base = (device >= 0 && device < kNoAllocDevices) ? g_noalloc_ptr[device] : nullptr;
    }
    if (!base) {
        set_error("the allocation-free prover was not set up on this device");
        return GKRX_ERR_STATE;
    }
    const uint64_t need = dev_noalloc_bytes(dc, R, tail_bytes);
    if (need > kNoAllocScratchBytes) {
        set_error("circuit needs " + std::to_string(need) + " bytes of prover tables, above the " +
                  std::to_string(kNoAllocScratchBytes) + " bytes of the allocation-free prover");
        return GKRX_ERR_ARG;
    }
    view = DevScratch{};
    scratch_layout(view, dc, R, tail_bytes, base);
    const uint32_t D = dc.D, mlast = dc.layers[D - 1].m;
    NoAllocSeed a{};
    for (uint32_t i = 0; i < R * mlast; i++) a.rz1[i] = rz1_host[i];
    memcpy(a.tstate, tstate_host, sizeof a.tstate);
    memcpy(a.grind, grind, sizeof a.grind);
    a.R = R;
    a.mlast = mlast;
    hipLaunchKernelGGL(k_install_seed, dim3(1), dim3(256), 0, st, a, view.chal + (uint64_t)D * R * kChalStride,
                       view.ab + 2ull * D, view.tstate, view.tail);
    GKRX_HIP_CHECK(hipGetLastError());
    if (dc.L == 8) return prove_impl<8>(dc, view, R, hash_kind, tstate_host[0], tstate_host[1], st, nullptr);
    return prove_impl<16>(dc, view, R, hash_kind, tstate_host[0], tstate_host[1], st, nullptr);
}

void dev_free_scratch(DevScratch &s)
{
    dev_drop_graph(s);
    if (s.pool) (void)hipFree(s.pool);
    if (s.stage) (void)hipHostFree(s.stage);
    s = DevScratch{};
}

// ------------------------------------------------------------------ batched prove (host)
namespace {

// u32 of one proof's tables in the batch pool, in the order dev_prove_batch hands them out
uint64_t batch_tables_u32(const DevCircuit &dc, uint32_t R)
{
    const uint32_t L = dc.L;
    const uint64_t maxN = 1ull << dc.max_n, maxM = 1ull << dc.max_m;
    const uint64_t N1 = scratch_rows(maxN), N2 = scratch_rows(maxN / 2), N4 = scratch_rows(maxN / 4);
    uint64_t u = 0;
    auto add = [&](uint64_t n) { u += (n + 63) & ~63ull; };
    add(kEqWords * maxM); add(kEqWords * maxN); add(R * maxN * L); add(R * N2 * L); add(R * N4 * L); add(R * N2 * L);
    add(R * N4 * L); add(R * N1); add(R * N2); add(R * N4); add((uint64_t)(dc.D + 1) * R * kChalStride);
    add(2ull * (dc.D + 1)); add(2ull * R * L); add(16); add((uint64_t)R * kPairVals * L * kMaxBlocksPerRep);
    return u;
}

constexpr uint32_t kLsMaxJobs = 2048;   // at most two work tasks per proof in a step, 4096 per step
constexpr uint32_t kLsGroups = 3;           // grouped lockstep engine: groups of proofs (one stream each)
constexpr uint32_t kLsGxTilesPerBlock = 4;  // g_x tiles per workgroup (pipelined: the next tiles' loads in flight)
constexpr uint32_t kLsHyTilesPerBlock = 4;  // h_y tiles per workgroup, likewise

uint32_t ls_grid(uint64_t work, uint32_t bs)
{
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((work + bs - 1) / bs, kMaxBlocksPerRep));
}

// Blocks of a gather task: about kLsGatherItems items per thread, rounded up to a multiple
// of 8 so that every gather task (they lead their step) starts on the first XCD and
// xcd_tile applies to its block index.
uint32_t ls_gather_blocks(uint64_t work, uint32_t bs)
{
    uint64_t b = (work + (uint64_t)bs * kLsGatherItems - 1) / ((uint64_t)bs * kLsGatherItems);
    b = std::max<uint64_t>(1, std::min<uint64_t>(b, 65536));
    return (uint32_t)((b + 7) & ~7ull);
}

// Blocks of a tiled gather task (g_x, h_y): one per tile up to 65536, a multiple of 8.
uint32_t ls_tile_blocks(uint64_t tiles)
{
    const uint64_t b = std::max<uint64_t>(1, std::min<uint64_t>(tiles, 65536));
    return (uint32_t)((b + 7) & ~7ull);
}

uint32_t ls_work_blocks(const LsTask &t, uint32_t R)
{
    return (t.op == kOpDot || t.op == kOpPair) ? R * t.nb : t.nb;
}

uint32_t ls_fin_blocks(const LsTask &t, uint32_t R) { return t.op == kOpTailRep || t.op == kOpPairFin ? R : 1u; }

LsTask ls_task(uint32_t op, uint32_t job, uint32_t l)
{
    LsTask t;
    memset(&t, 0, sizeof t);
    t.op = op;
    t.job = job;
    t.l = l;
    return t;
}

// One step of one proof: up to two independent work tasks and one fin task.
struct LsEntry {
    LsTask w[2];
    uint32_t nw = 0;
    LsTask f;
    uint32_t nf = 0;
};

// The steps of one proof, in the order of the single-proof path (prove_impl): the same
// kernels, round split and transcript positions.  The first D entries are the forward
// pass; the rest (the sumcheck) needs the transcript seed (kOpSeed, placed by the caller)
// one step before it.  J holds the device addresses of the proof's tables; ts_len / ts_hs
// is the transcript position after the host prefix.
void ls_job_ops(const DevCircuit &dc, const BatchJob &J, uint32_t job, uint32_t R, uint32_t BS, uint32_t LPT,
                uint64_t ts_len, uint64_t ts_hs, std::vector<LsEntry> &ops)
{
    const uint32_t L = dc.L, LG = L / 4, Q = 3 * L, D = dc.D;
    auto push_w = [&](const LsTask &t) {
        LsEntry e;
        e.w[0] = t;
        e.nw = 1;
        ops.push_back(e);
    };
    auto push_f = [&](const LsTask &t) {
        LsEntry e;
        e.f = t;
        e.nf = 1;
        ops.push_back(e);
    };
    auto chal_slot = [&](uint32_t l) { return J.chal + (uint64_t)l * R * kChalStride; };
    for (uint32_t l = 0; l < D; l++) {
        const LayerDev &ly = dc.layers[l];
        LsEntry e;
        LsTask t = ls_task(kOpFwd, job, l);
        t.fwd = LsFwd{dc.values + ly.v_off, dc.values + ly.out_off, ly.fw_ptr, ly.fw_ent, 1u << ly.m};
        t.nb = ls_gather_blocks((1ull << ly.m) * LG, BS);
        e.w[0] = t;
        e.nw = 1;
        ops.push_back(e);
    }
    for (int li = (int)D - 1; li >= 0; li--) {
        const uint32_t l = (uint32_t)li;
        const LayerDev &ly = dc.layers[l];
        const uint32_t n = ly.n, m = ly.m;
        const uint64_t N = 1ull << n, M = 1ull << m;
        const uint32_t *V = dc.values + ly.v_off;
        {
            LsTask t = ls_task(kOpEqz, job, l);
            t.eq = LsEq{J.eqz, chal_slot(l + 1), 0u, 32u, J.ab + 2ull * (l + 1), m};
            t.nb = (uint32_t)((M + kEqChunk * BS - 1) / (kEqChunk * BS));
            push_w(t);
        }
        {
            LsEntry e;
            LsTask t = ls_task(kOpGx, job, l);
            t.gx = LsGx{J.hg0, J.hg0_rs, V, J.eqz, ly.gx_ptr, ly.gx_ent, (uint32_t)N};
            t.nb = ls_tile_blocks(((N * L + tile_lanes(BS) - 1) / tile_lanes(BS) + kLsGxTilesPerBlock - 1) / kLsGxTilesPerBlock);
            e.w[0] = t;
            e.nw = 1;
            if (l == D - 1) {
                LsTask d = ls_task(kOpDot, job, l);
                d.dot = LsDot{dc.values + ly.out_off, J.eqz, (uint32_t)M, J.partials};
                d.nb = ls_grid(M * LG, BS);
                e.w[1] = d;
                e.nw = 2;
                LsTask fz = ls_task(kOpDotFin, job, l);
                fz.nb = d.nb;
                e.f = fz;
                e.nf = 1;
            }
            ops.push_back(e);
        }
        for (uint32_t phase = 1; phase <= 2; phase++) {
            const bool p2 = phase == 2;
            if (p2) {
                LsTask t = ls_task(kOpEqx, job, l);
                t.eq = LsEq{J.eqx, chal_slot(l), 0u, 0u, nullptr, n};
                t.nb = (uint32_t)((N + kEqChunk * BS - 1) / (kEqChunk * BS));
                push_w(t);
                LsTask h = ls_task(kOpHy, job, l);
                h.hy = LsHy{J.s0, J.s0_rs, J.eqz, J.eqx, ly.hy_ptr, ly.hy_ent, (uint32_t)N};
                h.nb = ls_tile_blocks(((N + tile_lanes(BS) - 1) / tile_lanes(BS) + kLsHyTilesPerBlock - 1) /
                                      kLsHyTilesPerBlock);
                push_w(h);
            }
            uint32_t *h0 = p2 ? J.s0 : J.hg0, *hA = p2 ? J.sA : J.hgA, *hB = p2 ? J.sB : J.hgB;
            const uint64_t h0_rs = p2 ? J.s0_rs : J.hg0_rs, hA_rs = p2 ? J.sA_rs : J.hgA_rs,
                           hB_rs = p2 ? J.sB_rs : J.hgB_rs;
            // rounds [0, K) in pairs on the whole GPU, the rest in the chains (as prove_impl)
            uint32_t k0 = 0;
            while (k0 < n && (N >> k0) > kTailRows) k0++;
            const uint32_t K = (k0 + 1) & ~1u, Jp = K / 2;
            auto pair_in_A = [&](uint32_t j) { return ((Jp - 1 - j) & 1) == 0; };
            for (uint32_t j = 0; j < Jp; j++) {
                const uint32_t k = 2 * j;
                RoundParams rp{};
                rp.items = (uint32_t)(N >> (k + 2));
                rp.partials = J.partials;
                if (j <= 1) {
                    rp.fsrc = V; rp.fsrc_rs = 0;
                    rp.hsrc = h0; rp.hsrc_rs = h0_rs;
                } else {
                    const bool a = pair_in_A(j - 1);
                    rp.fsrc = a ? J.fA : J.fB; rp.fsrc_rs = a ? J.fA_rs : J.fB_rs;
                    rp.hsrc = a ? hA : hB; rp.hsrc_rs = a ? hA_rs : hB_rs;
                }
                if (j >= 1) {
                    const bool a = pair_in_A(j);
                    rp.fdst = a ? J.fA : J.fB; rp.fdst_rs = a ? J.fA_rs : J.fB_rs;
                    rp.hdst = a ? hA : hB; rp.hdst_rs = a ? hA_rs : hB_rs;
                    rp.rprev = chal_slot(l) + (p2 ? 32 : 0) + (k - 2);
                }
                rp.R = R;
                LsEntry e;
                LsTask t = ls_task(kOpPair, job, l);
                t.phase = phase;
                t.k = k;
                t.rp = rp;
                t.nb = (ls_grid((uint64_t)rp.items * (L / LPT), BS) + 7) & ~7u;  // a multiple of 8 (pair_block)
                e.w[0] = t;
                e.nw = 1;
                LsTask fz = ls_task(kOpPairFin, job, l);
                fz.phase = phase;
                fz.k = k;
                fz.nb = t.nb;
                fz.len0 = (uint32_t)ts_len;
                fz.hs0 = (uint32_t)ts_hs;
                e.f = fz;
                e.nf = 1;
                ops.push_back(e);
                ts_len += 2ull * R * 4 * Q;
                ts_hs = ts_len;
            }
            k0 = K;
            const uint32_t klast = n - 1;
            if (k0 < klast) {
                LsTask fz = ls_task(kOpTailRep, job, l);
                fz.phase = phase;
                fz.k = k0;
                fz.kend = klast;
                fz.len0 = (uint32_t)ts_len;
                fz.hs0 = (uint32_t)ts_hs;
                fz.dbl = K >= 2;
                fz.nb = 1;
                push_f(fz);
                ts_len += (uint64_t)(klast - k0) * R * 4 * Q;
                ts_hs = ts_len;
            }
            LsTask fz = ls_task(kOpLast, job, l);
            fz.phase = phase;
            fz.nb = 1;
            push_f(fz);
            ts_len += (uint64_t)R * (4 * Q + 4 * L);
            ts_hs = p2 ? ts_len : ts_len - 4 * L;
        }
    }
}

template <int L>
void ls_launch_work(uint32_t BS, uint32_t blocks, hipStream_t st, const BatchJob *jobs, const LsTask *tasks,
                    const uint32_t *start, uint32_t n, uint32_t R, uint32_t base)
{
    const dim3 g(blocks);
    switch (BS) {
    case 64: hipLaunchKernelGGL((k_ls_work<L, 64>), g, dim3(64), 0, st, jobs, tasks, start, n, R, base); break;
    case 128: hipLaunchKernelGGL((k_ls_work<L, 128>), g, dim3(128), 0, st, jobs, tasks, start, n, R, base); break;
    case 512: hipLaunchKernelGGL((k_ls_work<L, 512>), g, dim3(512), 0, st, jobs, tasks, start, n, R, base); break;
    default: hipLaunchKernelGGL((k_ls_work<L, 256>), g, dim3(256), 0, st, jobs, tasks, start, n, R, base); break;
    }
}

// Work ops in the order they take within a step: the gathers first (their block counts
// are multiples of 8), then the rest.  With profiling on, each op's range of a step is
// launched (and timed) on its own.
constexpr uint32_t kLsWorkOrder[] = {kOpFwd, kOpGx, kOpHy, kOpSeed, kOpEqz, kOpEqx, kOpDot};

int ls_op_class(uint32_t op)
{
    switch (op) {
    case kOpFwd: return kKLsFwd;
    case kOpGx: return kKLsGx;
    case kOpHy: return kKLsHy;
    case kOpEqz:
    case kOpEqx: return kKLsEq;
    default: return kKLsWork;
    }
}

// The task lists of a step: first task, first start, count, workgroups.
struct LsList {
    uint32_t task, start, n, blocks;
};
// One step: the wide work other than round pairs and the claims (main stream), the round
// pairs and their transcript ends (pair stream), the serial chains (chain stream).
enum { kLsWork = 0, kLsDotFin, kLsPairs, kLsPairFin, kLsChain, kLsLists };
struct LsStep {
    LsList l[kLsLists];
    // wait[x]: the lanes (bit 0 main, 1 pairs, 2 chains) whose previous work lane x's ops of
    // this step follow: a proof's op at step t depends only on its op at step t - 1
    uint32_t wait[3];
};

bool ls_is_chain(uint32_t op) { return op == kOpTailRep || op == kOpLast; }

// The lane an entry of a proof's schedule runs on: 0 main, 1 pairs, 2 chains.
int ls_lane(const LsEntry &e)
{
    if (e.nw && e.w[0].op == kOpPair) return 1;
    if (!e.nw && e.nf && ls_is_chain(e.f.op)) return 2;
    return 0;
}

// The device addresses of every job's tables in the batch pool (job i's tables start at
// pool + job_off[i]; its tail and claims sit in the tail and claim areas).
void batch_job_table(const std::vector<BatchItem> &items, uint32_t R, uint32_t *pool,
                     const std::vector<uint64_t> &job_off, uint32_t *tails, const std::vector<uint64_t> &tail_off,
                     uint32_t *claims, std::vector<BatchJob> &hjobs)
{
    const uint32_t L = items[0].dc->L;
    for (size_t i = 0; i < items.size(); i++) {
        const DevCircuit &dc = *items[i].dc;
        const uint64_t maxN = 1ull << dc.max_n, maxM = 1ull << dc.max_m;
        const uint64_t N1 = scratch_rows(maxN), N2 = scratch_rows(maxN / 2), N4 = scratch_rows(maxN / 4);
        uint32_t *at = pool + job_off[i];
        auto take = [&](uint64_t n) { uint32_t *p = at; at += (n + 63) & ~63ull; return p; };
        BatchJob &J = hjobs[i];
        J = BatchJob{};
        J.D = dc.D;
        J.layers = dc.d_layers;
        J.values = dc.values;
        J.eqz = take(kEqWords * maxM);
        J.eqx = take(kEqWords * maxN);
        J.hg0_rs = maxN * L; J.hg0 = take(R * maxN * L);
        J.hgA_rs = N2 * L; J.hgA = take(R * N2 * L);
        J.hgB_rs = N4 * L; J.hgB = take(R * N4 * L);
        J.fA_rs = N2 * L; J.fA = take(R * N2 * L);
        J.fB_rs = N4 * L; J.fB = take(R * N4 * L);
        J.s0_rs = N1; J.s0 = take(R * N1);
        J.sA_rs = N2; J.sA = take(R * N2);
        J.sB_rs = N4; J.sB = take(R * N4);
        J.chal = take((uint64_t)(dc.D + 1) * R * kChalStride);
        J.ab = take(2ull * (dc.D + 1));
        J.vxy = take(2ull * R * L);
        J.tstate = take(16);
        J.partials = take((uint64_t)R * kPairVals * L * kMaxBlocksPerRep);
        J.tail = tails + tail_off[i];
        J.claims = claims + i * (uint64_t)R * L;
    }
}

// The lockstep schedule of a batch: for every step its task lists (tasks and the prefix
// arrays of their first workgroups) and, per job, the step that installs its transcript seed.
void ls_build_schedule(const std::vector<BatchItem> &items, const std::vector<BatchJob> &hjobs,
                       const std::vector<uint32_t> &members, uint32_t R, uint32_t BS, uint32_t LPT,
                       std::vector<LsTask> &tasks, std::vector<uint32_t> &starts, std::vector<LsStep> &steps,
                       std::vector<uint32_t> &seed_step)
{
    // the schedule of the jobs in `members` (indices into items / hjobs); index m below is a
    // member, members[m] its job
    const size_t nj = members.size();
    // The transcript position after the host prefix follows from the shapes alone: the 32
    // grind bytes are in the buffer, hashed once if output challenges were drawn.
    auto ts_hashed = [&](size_t i) -> uint64_t {
        const DevCircuit &dc = *items[members[i]].dc;
        return (uint64_t)R * dc.layers[dc.D - 1].m ? 32 : 0;
    };
    // Every proof's forward pass first (steps [0, D)): it needs no transcript.  Then
    // the sumchecks right-aligned: every proof ends at the last step, so a proof with
    // fewer sumcheck steps starts later and needs its host prefix later.  A proof's
    // seed runs one step before its first sumcheck op.
    std::vector<std::vector<LsEntry>> ops(nj);
    std::vector<LsEntry> seed_entry(nj);
    std::vector<size_t> base(nj);
    size_t dmax = 0, smax = 0, ntask = 0;
    for (size_t i = 0; i < nj; i++) {
        const uint32_t D = items[members[i]].dc->D;
        ops[i].reserve((size_t)D * (8 + 2 * items[members[i]].dc->max_n));
        ls_job_ops(*items[members[i]].dc, hjobs[members[i]], members[i], R, BS, LPT, 32, ts_hashed(i), ops[i]);
        dmax = std::max<size_t>(dmax, D);
        smax = std::max(smax, ops[i].size() - D);
        ntask += 2 * ops[i].size() + 1;
    }
    const size_t nsteps = dmax + smax;
    for (size_t i = 0; i < nj; i++) {
        const uint32_t D = items[members[i]].dc->D;
        base[i] = nsteps - (ops[i].size() - D);  // >= dmax >= D
        seed_step[members[i]] = (uint32_t)(base[i] - 1);
        LsTask sd = ls_task(kOpSeed, members[i], 0);
        sd.nb = 1;
        if (base[i] == D) {  // the seed shares the last forward step
            LsEntry &e = ops[i][D - 1];
            e.w[e.nw++] = sd;
        } else {
            seed_entry[i].w[0] = sd;
            seed_entry[i].nw = 1;
        }
    }
    // job i's entry at step s
    auto at = [&](size_t i, size_t s) -> const LsEntry * {
        const uint32_t D = items[members[i]].dc->D;
        if (s < D) return &ops[i][s];
        if (s + 1 == base[i]) return &seed_entry[i];
        if (s >= base[i]) return &ops[i][D + (s - base[i])];
        return nullptr;
    };
    tasks.reserve(ntask);
    starts.reserve(ntask + 3 * nsteps);
    steps.reserve(nsteps);
    for (size_t s = 0; s < nsteps; s++) {
        LsStep S{};
        // appends the tasks that `take` selects (in job order) as list `li`
        auto list = [&](int li, auto &&take, auto &&nblocks) {
            LsList &ls = S.l[li];
            ls.task = (uint32_t)tasks.size();
            ls.start = (uint32_t)starts.size();
            uint32_t blocks = 0;
            for (size_t i = 0; i < nj; i++) {
                const LsEntry *e = at(i, s);
                if (!e) continue;
                take(*e, [&](const LsTask &t) {
                    starts.push_back(blocks);
                    tasks.push_back(t);
                    blocks += nblocks(t);
                });
            }
            ls.n = (uint32_t)tasks.size() - ls.task;
            ls.blocks = blocks;
            starts.push_back(blocks);
        };
        auto work_blocks = [&](const LsTask &t) { return ls_work_blocks(t, R); };
        auto fin_blocks = [&](const LsTask &t) { return ls_fin_blocks(t, R); };
        // wide work: by op (the gathers first: their block counts are multiples of 8)
        {
            LsList &ls = S.l[kLsWork];
            ls.task = (uint32_t)tasks.size();
            ls.start = (uint32_t)starts.size();
            uint32_t blocks = 0;
            for (uint32_t op : kLsWorkOrder)
                for (size_t i = 0; i < nj; i++) {
                    const LsEntry *e = at(i, s);
                    if (!e) continue;
                    for (uint32_t q = 0; q < e->nw; q++)
                        if (e->w[q].op == op) {
                            starts.push_back(blocks);
                            tasks.push_back(e->w[q]);
                            blocks += ls_work_blocks(e->w[q], R);
                        }
                }
            ls.n = (uint32_t)tasks.size() - ls.task;
            ls.blocks = blocks;
            starts.push_back(blocks);
        }
        list(kLsDotFin, [](const LsEntry &e, auto &&put) { if (e.nf && e.f.op == kOpDotFin) put(e.f); },
             fin_blocks);
        // round pairs: the tasks whose workgroups run longest (most items per thread) first, so
        // that the launch's last workgroups are short ones (workgroups start in index order)
        {
            LsList &ls = S.l[kLsPairs];
            ls.task = (uint32_t)tasks.size();
            ls.start = (uint32_t)starts.size();
            std::vector<const LsTask *> pt;
            for (size_t i = 0; i < nj; i++) {
                const LsEntry *e = at(i, s);
                if (!e) continue;
                for (uint32_t q = 0; q < e->nw; q++)
                    if (e->w[q].op == kOpPair) pt.push_back(&e->w[q]);
            }
            auto per_block = [](const LsTask *t) { return (uint64_t)t->rp.items * 1024 / t->nb; };
            std::stable_sort(pt.begin(), pt.end(),
                             [&](const LsTask *a, const LsTask *b) { return per_block(a) > per_block(b); });
            uint32_t blocks = 0;
            for (const LsTask *t : pt) {
                starts.push_back(blocks);
                tasks.push_back(*t);
                blocks += work_blocks(*t);
            }
            ls.n = (uint32_t)tasks.size() - ls.task;
            ls.blocks = blocks;
            starts.push_back(blocks);
        }
        list(kLsPairFin, [](const LsEntry &e, auto &&put) { if (e.nf && e.f.op == kOpPairFin) put(e.f); },
             fin_blocks);
        list(kLsChain, [](const LsEntry &e, auto &&put) { if (e.nf && ls_is_chain(e.f.op)) put(e.f); },
             fin_blocks);
        for (size_t i = 0; s > 0 && i < nj; i++) {
            const LsEntry *e = at(i, s), *p = at(i, s - 1);
            if (e && p) S.wait[ls_lane(*e)] |= 1u << ls_lane(*p);
        }
        steps.push_back(S);
    }
}

inline uint64_t align256(uint64_t x) { return (x + 255) & ~255ull; }

}  // namespace

int dev_prove_batch(const std::vector<BatchItem> &items, uint32_t R, int hash_kind, BatchScratch &bs, hipStream_t st,
                    Profiler *pf, BatchPrefixes *pre)
{
    const size_t nj = items.size();
    if (nj == 0) return GKRX_OK;
    if (R > 3) return GKRX_ERR_CONFIG;
    const uint32_t L = items[0].dc->L;
    const bool lockstep = bs.engine != kBatchPerProof && nj <= kLsMaxJobs;
    // the grouped lockstep engine: the proofs in kLsGroups groups, each group's steps on a
    // stream of its own (no events between them); otherwise one group on three op lanes
    const bool grouped = lockstep && bs.engine == kBatchLockstepGrouped;
    const uint32_t BS = (bs.block == 64 || bs.block == 128 || bs.block == 512) ? bs.block : 256u;
    const uint32_t LPT = (bs.lanes == 1 || bs.lanes == 2) ? bs.lanes : 4u;  // round-pair lanes per thread
    // sizes of every job's scratch, one pool
    std::vector<uint64_t> job_off(nj + 1, 0), tail_off(nj + 1, 0);
    const uint64_t seed_u32 = (uint64_t)R * kChalStride + 10 + 8;
    for (size_t i = 0; i < nj; i++) {
        const DevCircuit &dc = *items[i].dc;
        if (dc.L != L) return GKRX_ERR_ARG;
        job_off[i + 1] = job_off[i] + batch_tables_u32(dc, R);
        tail_off[i + 1] = tail_off[i] + ((items[i].tail_bytes + 3) / 4 + 63) / 64 * 64;
    }
    const uint64_t need = job_off[nj] + tail_off[nj] + nj * (uint64_t)R * L + nj * seed_u32 + nj + 64;
    GKRX_HIP_CHECK(hipStreamSynchronize(st));  // the previous batch may still use the buffers
    if (need > bs.pool_u32) {
        if (bs.pool) (void)hipFree(bs.pool);
        bs.pool = nullptr;
        bs.pool_u32 = 0;
        bs.meta_key.clear();  // the job tables point into the pool
        GKRX_HIP_CHECK(hipMalloc(&bs.pool, need * 4));
        bs.pool_u32 = need;
    }
    if (tail_off[nj] * 4 > bs.tails_cap) {
        if (bs.tails_host) (void)hipHostFree(bs.tails_host);
        bs.tails_host = nullptr;
        bs.tails_cap = 0;
        GKRX_HIP_CHECK(hipHostMalloc(&bs.tails_host, tail_off[nj] * 4, hipHostMallocDefault));
        bs.tails_cap = tail_off[nj] * 4;
    }
    bs.tail_stage_off.assign(tail_off.begin(), tail_off.end());
    uint32_t *tails = bs.pool + job_off[nj];
    uint32_t *claims = tails + tail_off[nj];
    // seed records (rz1 in chal-slot layout, transcript state, grind bytes), one per job in
    // the order in which the schedule needs them
    uint32_t *seeds = claims + nj * (uint64_t)R * L;
    uint32_t *tlens = seeds + nj * seed_u32;  // each job's final transcript length
    // The job table and the schedule depend only on the circuits' shapes, their device
    // addresses and the pool: a batch with the same key reuses the metadata already on the
    // device.
    std::vector<uint64_t> key = {(uint64_t)lockstep + 2 * (uint64_t)grouped, BS, LPT, L, R, nj,
                                 (uint64_t)(uintptr_t)bs.pool};
    for (size_t i = 0; i < nj; i++) {
        const DevCircuit &dc = *items[i].dc;
        key.insert(key.end(), {(uint64_t)(uintptr_t)dc.values, (uint64_t)(uintptr_t)dc.csr_pool,
                               (uint64_t)(uintptr_t)dc.d_layers, dc.D, items[i].tail_bytes});
        for (uint32_t l = 0; l < dc.D; l++) {
            const LayerDev &ly = dc.layers[l];
            key.insert(key.end(), {((uint64_t)ly.n << 32) | ly.m, ly.v_off, ly.out_off, (uint64_t)(uintptr_t)ly.fw_ptr,
                                   (uint64_t)(uintptr_t)ly.fw_ent, (uint64_t)(uintptr_t)ly.gx_ptr,
                                   (uint64_t)(uintptr_t)ly.gx_ent, (uint64_t)(uintptr_t)ly.hy_ptr,
                                   (uint64_t)(uintptr_t)ly.hy_ent});
        }
    }
    const bool reuse = bs.meta && bs.meta_used && key == bs.meta_key;
    if (!reuse) {
        bs.meta_key.clear();
        bs.meta_used = 0;
        std::vector<BatchJob> hjobs(nj);
        batch_job_table(items, R, bs.pool, job_off, tails, tail_off, claims, hjobs);
        // per-proof engine: workgroup order, largest proof first; lockstep: the step schedule
        std::vector<uint32_t> horder;
        std::vector<LsTask> tasks;
        std::vector<uint32_t> starts;
        std::vector<LsStep> steps;
        bs.meta_seed_step.assign(nj, 0);
        if (!lockstep) {
            std::vector<std::pair<uint64_t, uint32_t>> work(nj);  // (input rows of all layers, job)
            for (size_t i = 0; i < nj; i++) {
                uint64_t w = 0;
                for (uint32_t l = 0; l < items[i].dc->D; l++) w += (1ull << items[i].dc->layers[l].n);
                work[i] = {w, (uint32_t)i};
            }
            std::stable_sort(work.begin(), work.end(), [](auto &x, auto &y) { return x.first > y.first; });
            horder.resize(nj);
            for (size_t i = 0; i < nj; i++) horder[i] = work[i].second;
        }
        // groups of proofs: the heaviest first, each to the group with the fewest table rows
        const uint32_t G = grouped ? (uint32_t)std::min<size_t>(kLsGroups, nj) : 1u;
        std::vector<uint32_t> group(nj, 0);
        if (G > 1) {
            std::vector<std::pair<uint64_t, uint32_t>> work(nj);
            for (size_t i = 0; i < nj; i++) {
                uint64_t w = 0;
                for (uint32_t l = 0; l < items[i].dc->D; l++) w += 1ull << items[i].dc->layers[l].n;
                work[i] = {w, (uint32_t)i};
            }
            std::stable_sort(work.begin(), work.end(), [](auto &x, auto &y) { return x.first > y.first; });
            std::vector<uint64_t> load(G, 0);
            for (auto &w : work) {
                const uint32_t g = (uint32_t)(std::min_element(load.begin(), load.end()) - load.begin());
                group[w.second] = g;
                load[g] += w.first;
            }
        }
        bs.meta_group_off.assign(1, 0);
        if (lockstep)
            for (uint32_t g = 0; g < G; g++) {
                std::vector<uint32_t> members;
                for (size_t i = 0; i < nj; i++)
                    if (group[i] == g) members.push_back((uint32_t)i);
                ls_build_schedule(items, hjobs, members, R, BS, LPT, tasks, starts, steps, bs.meta_seed_step);
                bs.meta_group_off.push_back((uint32_t)steps.size());
            }
        // seed slots: by group, each group's in the order of need (a group uploads its own
        // slots, a contiguous range); the prefixes are computed in the order of need overall
        std::vector<uint32_t> &sorder = bs.meta_seed_order;
        sorder.resize(nj);
        for (size_t i = 0; i < nj; i++) sorder[i] = (uint32_t)i;
        std::vector<uint32_t> &norder = bs.meta_need_order;
        norder = sorder;
        std::stable_sort(norder.begin(), norder.end(),
                         [&](uint32_t a, uint32_t b) { return bs.meta_seed_step[a] < bs.meta_seed_step[b]; });
        std::stable_sort(sorder.begin(), sorder.end(), [&](uint32_t a, uint32_t b) {
            return group[a] != group[b] ? group[a] < group[b] : bs.meta_seed_step[a] < bs.meta_seed_step[b];
        });
        bs.meta_group_slot.assign(G + 1, 0);
        for (size_t i = 0; i < nj; i++) bs.meta_group_slot[group[i] + 1]++;
        for (uint32_t g = 0; g < G; g++) bs.meta_group_slot[g + 1] += bs.meta_group_slot[g];
        for (size_t q = 0; q < nj; q++) hjobs[sorder[q]].seed = seeds + q * seed_u32;
        // device metadata: [jobs][order][tasks][starts]; the pinned staging adds the seeds
        bs.o_order = align256(nj * sizeof(BatchJob));
        bs.o_tasks = align256(bs.o_order + 4 * horder.size());
        bs.o_starts = align256(bs.o_tasks + sizeof(LsTask) * tasks.size());
        const uint64_t meta_need = align256(bs.o_starts + 4 * starts.size());
        if (meta_need > bs.meta_bytes) {
            if (bs.meta) (void)hipFree(bs.meta);
            bs.meta = nullptr;
            bs.meta_bytes = 0;
            GKRX_HIP_CHECK(hipMalloc(&bs.meta, meta_need));
            bs.meta_bytes = meta_need;
        }
        const uint64_t stage_need = meta_need + nj * seed_u32 * 4;
        if (stage_need > bs.host_meta_bytes) {
            if (bs.host_meta) (void)hipHostFree(bs.host_meta);
            bs.host_meta = nullptr;
            bs.host_meta_bytes = 0;
            GKRX_HIP_CHECK(hipHostMalloc(&bs.host_meta, stage_need, hipHostMallocDefault));
            bs.host_meta_bytes = stage_need;
        }
        memcpy(bs.host_meta, hjobs.data(), nj * sizeof(BatchJob));
        if (!horder.empty()) memcpy(bs.host_meta + bs.o_order, horder.data(), 4 * horder.size());
        if (!tasks.empty()) memcpy(bs.host_meta + bs.o_tasks, tasks.data(), sizeof(LsTask) * tasks.size());
        if (!starts.empty()) memcpy(bs.host_meta + bs.o_starts, starts.data(), 4 * starts.size());
        GKRX_HIP_CHECK(hipMemcpyAsync(bs.meta, bs.host_meta, meta_need, hipMemcpyHostToDevice, st));
        bs.meta_steps.resize(steps.size() * (sizeof(LsStep) / 4));
        if (!steps.empty()) memcpy(bs.meta_steps.data(), steps.data(), steps.size() * sizeof(LsStep));
        bs.meta_used = meta_need;
        bs.meta_key.swap(key);
    }
    static_assert(sizeof(LsStep) % 4 == 0, "step records are whole words");
    std::vector<LsStep> steps(bs.meta_steps.size() / (sizeof(LsStep) / 4));
    if (!steps.empty()) memcpy(steps.data(), bs.meta_steps.data(), steps.size() * sizeof(LsStep));
    // The seeds change with every batch.  upload(seeded, q, sx) copies the records of slots
    // [seeded, q) (waiting for their prefixes first) through the pinned staging after the
    // metadata, on stream sx.
    const std::vector<uint32_t> &sorder = bs.meta_seed_order;
    const std::vector<uint32_t> &gslot = bs.meta_group_slot;
    uint32_t *stage = reinterpret_cast<uint32_t *>(bs.host_meta + bs.meta_used);
    if (pre) pre->start(bs.meta_need_order);
    auto upload = [&](size_t &seeded, size_t q, hipStream_t sx) -> int {
        if (q <= seeded) return GKRX_OK;
        for (size_t k = seeded; k < q; k++) {
            const size_t i = sorder[k];
            if (pre) pre->wait((uint32_t)i);
            const DevCircuit &dc = *items[i].dc;
            uint32_t *sd = stage + k * seed_u32;
            const uint32_t mlast = dc.layers[dc.D - 1].m;
            memset(sd, 0, seed_u32 * 4);
            for (uint32_t j = 0; j < R; j++)
                for (uint32_t v = 0; v < mlast; v++) sd[j * kChalStride + v] = items[i].rz1[(size_t)j * mlast + v];
            memcpy(sd + R * kChalStride, items[i].tstate, 40);
            memcpy(sd + R * kChalStride + 10, items[i].grind, 32);
        }
        GKRX_HIP_CHECK(hipMemcpyAsync(seeds + seeded * seed_u32, stage + seeded * seed_u32,
                                      (q - seeded) * seed_u32 * 4, hipMemcpyHostToDevice, sx));
        seeded = q;
        return GKRX_OK;
    };
    // with the prefixes still in progress, the lockstep engine uploads them step by step
    size_t seeded_all = 0;
    const bool seeds_now = !(pre && lockstep);
    if (seeds_now)
        if (int rc = upload(seeded_all, nj, st)) return rc;
    const BatchJob *djobs = reinterpret_cast<const BatchJob *>(bs.meta);
    if (!lockstep) {
        const uint32_t *dorder = reinterpret_cast<const uint32_t *>(bs.meta + bs.o_order);
        if (!bs.queue) GKRX_HIP_CHECK(hipMalloc(&bs.queue, 256));
        // two workgroups per CU stay resident (768 threads each) and drain the queue
        int dev = 0, cus = 0;
        GKRX_HIP_CHECK(hipGetDevice(&dev));
        GKRX_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        const uint32_t grid = (uint32_t)std::min<uint64_t>(nj, 2ull * (uint64_t)std::max(cus, 1));
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, st, bs.queue);
        timed(pf, kKBatch, st, [&] {
            if (L == 8)
                hipLaunchKernelGGL(k_prove_batch<8>, dim3(grid), dim3(kBatchThreads), 0, st, djobs, dorder,
                                   (uint32_t)nj, bs.queue, R, hash_kind);
            else
                hipLaunchKernelGGL(k_prove_batch<16>, dim3(grid), dim3(kBatchThreads), 0, st, djobs, dorder,
                                   (uint32_t)nj, bs.queue, R, hash_kind);
        });
    } else {
        const LsTask *dtasks = reinterpret_cast<const LsTask *>(bs.meta + bs.o_tasks);
        const uint32_t *dstarts = reinterpret_cast<const uint32_t *>(bs.meta + bs.o_starts);
        // timing any class but the round pairs (whose launches record their own start and
        // end) puts every lane on one stream and times each op class of the work alone
        const bool serial = pf && (pf->mask & ~(1ull << kKLsPair)) != 0;
        const bool multi = bs.chain_stream && !serial;
        if (multi && !bs.aux) {
            GKRX_HIP_CHECK(hipStreamCreateWithFlags(&bs.aux, hipStreamNonBlocking));
            GKRX_HIP_CHECK(hipStreamCreateWithFlags(&bs.aux2, hipStreamNonBlocking));
            GKRX_HIP_CHECK(hipEventCreateWithFlags(&bs.ev_main, hipEventDisableTiming));
            GKRX_HIP_CHECK(hipEventCreateWithFlags(&bs.ev_aux, hipEventDisableTiming));
            GKRX_HIP_CHECK(hipEventCreateWithFlags(&bs.ev_aux2, hipEventDisableTiming));
        }
        hipStream_t streams[3] = {st, multi ? bs.aux2 : st, multi ? bs.aux : st};
        hipEvent_t events[3] = {bs.ev_main, bs.ev_aux2, bs.ev_aux};
        const std::vector<uint32_t> &goff = bs.meta_group_off;
        const uint32_t G = (uint32_t)goff.size() - 1;
        // next slot to upload, per group (all uploaded already when the prefixes were ready)
        std::vector<size_t> gseeded(gslot.begin() + (seeds_now ? 1 : 0), gslot.end() - (seeds_now ? 0 : 1));
        // the seeds of the SEED tasks of group g's step t, on stream sx
        auto seed_step_t = [&](uint32_t g, size_t t, hipStream_t sx) -> int {
            size_t q = gseeded[g];
            while (q < gslot[g + 1] && bs.meta_seed_step[sorder[q]] <= t) q++;
            return upload(gseeded[g], q, sx);
        };
        auto fin = [&](const LsList &ls, hipStream_t sx, int cls) {
            timed(pf, cls, sx, [&] {
                if (L == 8)
                    hipLaunchKernelGGL(k_ls_fin<8>, dim3(ls.blocks), dim3(kLsFinThreads), 0, sx, djobs, dtasks + ls.task,
                                       dstarts + ls.start, ls.n, R, hash_kind);
                else
                    hipLaunchKernelGGL(k_ls_fin<16>, dim3(ls.blocks), dim3(kLsFinThreads), 0, sx, djobs, dtasks + ls.task,
                                       dstarts + ls.start, ls.n, R, hash_kind);
            });
        };
        auto chain = [&](const LsList &ls, hipStream_t cs) {
            timed(pf, kKLsChain, cs, [&] {
                if (L == 8)
                    hipLaunchKernelGGL(k_ls_chain<8>, dim3(ls.blocks), dim3(kChainThreads), 0, cs, djobs,
                                       dtasks + ls.task, dstarts + ls.start, ls.n, R, hash_kind);
                else
                    hipLaunchKernelGGL(k_ls_chain<16>, dim3(ls.blocks), dim3(kChainThreads), 0, cs, djobs,
                                       dtasks + ls.task, dstarts + ls.start, ls.n, R, hash_kind);
            });
        };
        // the round pairs of a step, then their transcript ends
        auto pairs = [&](const LsStep &S, hipStream_t ps) {
            const LsList &ls = S.l[kLsPairs];
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (pf && !pf->kernel_events(kKLsPair, e0, e1)) e0 = e1 = nullptr;
            if (L == 8) ls_launch_pair<8>(BS, LPT, ls.blocks, ps, dtasks + ls.task, dstarts + ls.start, ls.n, e0, e1);
            else ls_launch_pair<16>(BS, LPT, ls.blocks, ps, dtasks + ls.task, dstarts + ls.start, ls.n, e0, e1);
            if (S.l[kLsPairFin].n) fin(S.l[kLsPairFin], ps, kKPairFin);
        };
        // the wide work of a step (one launch; profiling: one per run of tasks of an op
        // class), then the claims
        auto work = [&](const LsStep &S, hipStream_t ws) {
            const LsList &W = S.l[kLsWork];
            auto go = [&](uint32_t blocks, uint32_t base) {
                if (L == 8) ls_launch_work<8>(BS, blocks, ws, djobs, dtasks + W.task, dstarts + W.start, W.n, R, base);
                else ls_launch_work<16>(BS, blocks, ws, djobs, dtasks + W.task, dstarts + W.start, W.n, R, base);
            };
            if (W.n && !serial) {
                go(W.blocks, 0);
            } else if (W.n) {
                const LsTask *ht = reinterpret_cast<const LsTask *>(bs.host_meta + bs.o_tasks) + W.task;
                const uint32_t *hs = reinterpret_cast<const uint32_t *>(bs.host_meta + bs.o_starts) + W.start;
                for (uint32_t a = 0; a < W.n;) {
                    uint32_t b = a + 1;
                    while (b < W.n && ls_op_class(ht[b].op) == ls_op_class(ht[a].op)) b++;
                    timed(pf, ls_op_class(ht[a].op), ws, [&] { go(hs[b] - hs[a], hs[a]); });
                    a = b;
                }
            }
            if (S.l[kLsDotFin].n) fin(S.l[kLsDotFin], ws, kKLsFin);
        };
        if (grouped) {
            // group g's steps in order on streams[g], every op of a step after the step
            // before: the groups share no tables, so nothing orders one group after another
            // except the fork from the main stream (the metadata copies) and the final join
            if (multi)
                for (uint32_t g = 1; g < G; g++) {
                    if (g == 1) GKRX_HIP_CHECK(hipEventRecord(events[0], st));
                    GKRX_HIP_CHECK(hipStreamWaitEvent(streams[g], events[0], 0));
                }
            size_t tmax = 0;
            for (uint32_t g = 0; g < G; g++) tmax = std::max<size_t>(tmax, goff[g + 1] - goff[g]);
            for (size_t t = 0; t < tmax; t++)
                for (uint32_t g = 0; g < G; g++) {
                    if (t >= goff[g + 1] - goff[g]) continue;
                    const LsStep &S = steps[goff[g] + t];
                    hipStream_t sx = streams[g];
                    if (int rc = seed_step_t(g, t, sx)) return rc;
                    work(S, sx);
                    if (S.l[kLsPairs].n) pairs(S, sx);
                    if (S.l[kLsChain].n) chain(S.l[kLsChain], sx);
                }
            if (multi)
                for (uint32_t g = 1; g < G; g++) {
                    GKRX_HIP_CHECK(hipEventRecord(events[g], streams[g]));
                    GKRX_HIP_CHECK(hipStreamWaitEvent(st, events[g], 0));
                }
        } else {
            // Three lanes: 0 main (wide work, claims), 1 pairs (+ their transcript ends), 2
            // chains.  With chain_stream they run on three streams: at step t each lane first
            // waits for what the lanes its ops follow (LsStep::wait) recorded at the end of
            // steps before t, and every lane that launched records its event after all lanes
            // of step t are enqueued.  Otherwise (and with profiling on, so that each op class
            // is timed alone) every lane is the main stream.
            uint64_t rec[3] = {1, 0, 0}, seen[3][3] = {};
            if (multi) GKRX_HIP_CHECK(hipEventRecord(events[0], st));  // the fork: st's work so far
            auto join = [&](int x, uint32_t mask) -> int {
                if (!multi) return GKRX_OK;
                for (int o = 0; o < 3; o++)
                    if (o != x && ((mask >> o) & 1) && rec[o] > seen[x][o]) {
                        GKRX_HIP_CHECK(hipStreamWaitEvent(streams[x], events[o], 0));
                        seen[x][o] = rec[o];
                    }
                return GKRX_OK;
            };
            const size_t nsteps = goff[1];
            for (size_t s2 = 0; s2 < nsteps; s2++) {
                const LsStep &S = steps[s2];
                bool launched[3] = {false, false, false};
                if (S.l[kLsChain].n) {
                    if (int rc = join(2, S.wait[2])) return rc;
                    chain(S.l[kLsChain], streams[2]);
                    launched[2] = true;
                }
                if (S.l[kLsPairs].n) {
                    if (int rc = join(1, S.wait[1])) return rc;
                    pairs(S, streams[1]);
                    launched[1] = true;
                }
                if (S.l[kLsWork].n || S.l[kLsDotFin].n) {
                    // the seeds of this step's SEED tasks (the other lanes are enqueued already)
                    if (int rc = seed_step_t(0, s2, st)) return rc;
                    if (int rc = join(0, S.wait[0])) return rc;
                    work(S, st);
                    launched[0] = true;
                }
                if (multi)
                    for (int x = 0; x < 3; x++)
                        if (launched[x]) {
                            GKRX_HIP_CHECK(hipEventRecord(events[x], streams[x]));
                            rec[x]++;
                        }
            }
            // the copies below (and the next batch, which synchronizes st) follow every lane
            if (int rc = join(0, 7u)) return rc;
        }
        for (uint32_t g = 0; g < G; g++)
            if (gseeded[g] != gslot[g + 1]) return GKRX_ERR_STATE;  // every proof has a SEED task
    }
    hipLaunchKernelGGL(k_batch_tlen, dim3((uint32_t)((nj + 255) / 256)), dim3(256), 0, st, djobs, (uint32_t)nj, tlens);
    GKRX_HIP_CHECK(hipGetLastError());
    GKRX_HIP_CHECK(hipMemcpyAsync(bs.tails_host, tails, tail_off[nj] * 4, hipMemcpyDeviceToHost, st));
    uint32_t *hc = bs.host_claims(nj * R * L + nj);
    if (!hc) return GKRX_ERR_NOMEM;
    GKRX_HIP_CHECK(hipMemcpyAsync(hc, claims, nj * (uint64_t)R * L * 4, hipMemcpyDeviceToHost, st));
    GKRX_HIP_CHECK(hipMemcpyAsync(hc + nj * (uint64_t)R * L, tlens, nj * 4, hipMemcpyDeviceToHost, st));
    bs.claims_words = nj * (uint64_t)R * L;
    return GKRX_OK;
}

int dev_batch_collect(const std::vector<BatchItem> &items, const BatchScratch &bs)
{
    // every job's device transcript must have ended where the proof layout says
    const uint32_t *tlen = bs.claims_host + bs.claims_words;
    for (size_t i = 0; i < items.size(); i++)
        if (tlen[i] != items[i].tail_bytes) {
            set_error("batch item " + std::to_string(i) + ": device transcript length " + std::to_string(tlen[i]) +
                      " != " + std::to_string(items[i].tail_bytes));
            return GKRX_ERR_DEVICE;
        }
    parallel_for(items.size(), 16, [&](size_t i) {
        memcpy(items[i].tail_out, bs.tails_host + bs.tail_stage_off[i] * 4, items[i].tail_bytes);
    });
    return GKRX_OK;
}

uint64_t dev_batch_job_bytes(const DevCircuit &dc, uint32_t R, uint64_t tail_bytes)
{
    uint64_t u = batch_tables_u32(dc, R);
    auto add = [&](uint64_t n) { u += (n + 63) & ~63ull; };
    add((tail_bytes + 3) / 4);
    add((uint64_t)dc.L * R + (uint64_t)R * kChalStride + 18);
    return 4 * u;
}

int dev_batch_supported(const DevCircuit &dc)
{
    return dc.max_n <= kBatchMaxVars && dc.max_m <= kBatchMaxVars;
}

void dev_free_batch(BatchScratch &bs)
{
    for (hipStream_t x : {bs.aux, bs.aux2})
        if (x) {
            (void)hipStreamSynchronize(x);
            (void)hipStreamDestroy(x);
        }
    for (hipEvent_t e : {bs.ev_main, bs.ev_aux, bs.ev_aux2})
        if (e) (void)hipEventDestroy(e);
    if (bs.pool) (void)hipFree(bs.pool);
    if (bs.meta) (void)hipFree(bs.meta);
    if (bs.queue) (void)hipFree(bs.queue);
    if (bs.host_meta) (void)hipHostFree(bs.host_meta);
    if (bs.claims_host) (void)hipHostFree(bs.claims_host);
    if (bs.tails_host) (void)hipHostFree(bs.tails_host);
    bs = BatchScratch{};
}

}  // namespace gkrx
