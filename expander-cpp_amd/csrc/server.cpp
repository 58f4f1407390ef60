// Continuous-batching prove server (gkrx_server_*): the serving loop of the library.
//
// The reference proves one circuit per Prover::prove call and has no serving layer; this
// is the build's own.  A request is a circuit whose witness is resident in HBM.
//
//   1. Host prefix.  A pool of host threads computes each request's transcript prefix
//      (the commitment hash, grinding and output challenges: a serial hash chain per
//      proof) as soon as it arrives, so that work overlaps everything on the GPU.
//   2. Admission.  Requests whose prefix is ready are admitted in arrival order while the
//      HBM they need -- the circuit's resident values and wiring plus the prover tables,
//      both computed from the circuit shape -- fits the budget.
//   3. Launch.  Slots are streams with their own scratch.  A batch slot launches every
//      admitted M31 request at once with the context's batch engine (the lockstep engine
//      by default: each proof's step on as many workgroups as its tables need); a single
//      slot runs one BN254 proof, or one M31 proof too wide for the batch engines, with
//      the multi-block path.  Whichever slot is free takes the next work, so requests that
//      arrive while others run start without waiting for them, and their kernels overlap
//      on the device.
//   4. Completion is observed by polling each slot's event (a stream host-function
//      callback is only delivered while some host thread is inside the runtime, and the
//      server's threads sleep between events).
//
// Each proof is the same bytes as gkrx_prove of the same circuit: a proof's transcript
// depends only on its own circuit and witness.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "handles.hpp"
#include "host_hash.hpp"

using Clock = std::chrono::steady_clock;

namespace {

constexpr uint32_t kMaxBatch = 1024;
// With the one-workgroup-per-proof batch engine, M31 requests with at least this many
// table rows over all layers (sum of 2^n) take a single slot: in one workgroup they would
// run for hundreds of milliseconds.
constexpr uint64_t kBigRows = 1ull << 19;
constexpr auto kPollPeriod = std::chrono::microseconds(100);

struct SrvReq {
    uint64_t id = 0;
    gkrx_circuit *c = nullptr;
    uint8_t *proof = nullptr;
    uint32_t *claims = nullptr;
    uint64_t bytes = 0;   // admission footprint
    bool single = false;  // runs on a single slot (BN254, or M31 outside the batch engine)
    // submitted; host prefix started and done; launched; completion seen on the device
    Clock::time_point t_submit, t_prefix0, t_prefix1, t_launch, t_seen;
    int state = 0;        // 0 queued, 1 on the GPU, 2 done
    int rc = 0;
    size_t len = 0;
    double lat_ms = 0;
    double split_ms[6] = {};  // see gkrx_server_wait_split
    std::vector<uint32_t> rz1;
    uint32_t tstate[10] = {};
    uint8_t grind[32] = {};
};

struct Slot {
    bool single = false;
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;  // recorded after the slot's last copy
    gkrx::BatchScratch bs;    // batch slot
    gkrx::DevScratch ds;      // single slot, M31
    gkrx::BnScratch bn;       // single slot, BN254
    uint8_t *tail_host = nullptr;  // pinned staging of a single proof's tail
    uint64_t tail_cap = 0;
    uint32_t *small_host = nullptr;  // pinned: claims (up to 64 words) then tstate (16 words)
    std::vector<SrvReq *> reqs;
    std::vector<gkrx::BatchItem> items;
    bool busy = false;
};

}  // namespace

struct gkrx_server {
    gkrx_ctx *ctx = nullptr;
    uint64_t budget = 0, in_use = 0, peak = 0;
    std::vector<std::unique_ptr<Slot>> slots;
    std::mutex m;
    std::condition_variable cv, done_cv, pcv;
    std::deque<SrvReq *> pq;     // waiting for the host prefix
    std::deque<SrvReq *> ready;  // prefix done, waiting for admission
    uint32_t prefixing = 0;      // prefixes being computed
    std::unordered_map<uint64_t, SrvReq *> reqs;
    uint64_t next_id = 1;
    bool stop = false;
    uint64_t submitted = 0, completed = 0, batches = 0, max_batch = 0;
    std::thread sched;
    std::vector<std::thread> workers;
};

using namespace gkrx;

namespace {

// HBM a request holds while admitted: its circuit plus the prover tables of one proof (the
// single-slot M31 path keeps the same tables as one batch job)
uint64_t footprint(const gkrx_ctx *ctx, const gkrx_circuit *c)
{
    const uint64_t tb = tail_bytes(ctx, c);
    return dev_circuit_bytes(c->dev) +
           (ctx->bn254 ? dev_bn_scratch_bytes(c->dev, tb) : dev_batch_job_bytes(c->dev, ctx->R, tb));
}

void compute_prefix(gkrx_ctx *ctx, SrvReq *r)
{
    const int kind = ctx->cfg.fs_hash == GKRX_FS_KECCAK256 ? 1 : 0;
    gkrx_circuit *c = r->c;
    const uint32_t mlast = c->m[c->D - 1];
    const size_t cl = commit_bytes(ctx, c);
    if (ctx->bn254) {
        r->rz1.resize((size_t)8 * mlast);
        host_transcript_prefix_bn254(kind, reinterpret_cast<const uint8_t *>(c->witness.data()), cl, mlast,
                                     (uint32_t)ctx->cfg.grinding_bits, r->grind, r->rz1.data(), r->tstate);
    } else {
        r->rz1.resize((size_t)ctx->R * mlast);
        host_transcript_prefix(kind, reinterpret_cast<const uint8_t *>(c->witness.data()), cl, ctx->R, mlast,
                               (uint32_t)ctx->cfg.grinding_bits, r->grind, r->rz1.data(), r->tstate);
    }
    memcpy(r->proof, c->witness.data(), cl);  // RAW commitment bytes (raw.hpp:19-27)
}

void prefix_worker(gkrx_server *S)
{
    std::unique_lock<std::mutex> lk(S->m);
    for (;;) {
        S->pcv.wait(lk, [&] { return S->stop || !S->pq.empty(); });
        if (S->pq.empty()) return;  // stopping and nothing left
        SrvReq *r = S->pq.front();
        S->pq.pop_front();
        S->prefixing++;
        lk.unlock();
        r->t_prefix0 = Clock::now();
        compute_prefix(S->ctx, r);
        r->t_prefix1 = Clock::now();
        lk.lock();
        S->prefixing--;
        S->ready.push_back(r);
        S->cv.notify_all();
    }
}

int ensure_pinned(Slot &sl, uint64_t tail_bytes)
{
    if (tail_bytes > sl.tail_cap) {
        if (sl.tail_host) (void)hipHostFree(sl.tail_host);
        sl.tail_host = nullptr;
        sl.tail_cap = 0;
        GKRX_HIP_CHECK(hipHostMalloc(&sl.tail_host, tail_bytes, hipHostMallocDefault));
        sl.tail_cap = tail_bytes;
    }
    if (!sl.small_host) GKRX_HIP_CHECK(hipHostMalloc(&sl.small_host, 80 * 4, hipHostMallocDefault));
    return GKRX_OK;
}

// Enqueues the slot's work (prefixes are ready); nothing is left enqueued on failure
// except work that the caller drains.
int launch(gkrx_server *S, Slot &sl)
{
    gkrx_ctx *ctx = S->ctx;
    const int kind = ctx->cfg.fs_hash == GKRX_FS_KECCAK256 ? 1 : 0;
    if (!sl.st) {
        GKRX_HIP_CHECK(hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking));
        GKRX_HIP_CHECK(hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming));
    }
    if (sl.single) {
        SrvReq *r = sl.reqs[0];
        gkrx_circuit *c = r->c;
        const uint64_t tb = tail_bytes(ctx, c);
        int rc = ensure_pinned(sl, tb);
        if (rc) return rc;
        if (ctx->bn254) {
            rc = dev_bn_ensure_scratch(sl.bn, c->dev, tb);
            if (!rc) rc = dev_bn_forward(c->dev, sl.st, nullptr);
            if (!rc) rc = dev_bn_prove(c->dev, sl.bn, kind, r->rz1.data(), r->tstate, r->grind, 32, sl.st, nullptr);
            if (rc) return rc;
            GKRX_HIP_CHECK(hipMemcpyAsync(sl.tail_host, sl.bn.tail, tb, hipMemcpyDeviceToHost, sl.st));
            GKRX_HIP_CHECK(hipMemcpyAsync(sl.small_host, sl.bn.claims, 32, hipMemcpyDeviceToHost, sl.st));
            GKRX_HIP_CHECK(hipMemcpyAsync(sl.small_host + 64, sl.bn.tstate, 40, hipMemcpyDeviceToHost, sl.st));
        } else {
            rc = dev_ensure_scratch(sl.ds, c->dev, ctx->R, tb);
            if (!rc) rc = dev_forward(c->dev, sl.st, nullptr);
            if (!rc)
                rc = dev_prove(c->dev, sl.ds, ctx->R, kind, r->rz1.data(), r->tstate, r->grind, 32, sl.st, nullptr);
            if (rc) return rc;
            GKRX_HIP_CHECK(hipMemcpyAsync(sl.tail_host, sl.ds.tail, tb, hipMemcpyDeviceToHost, sl.st));
            GKRX_HIP_CHECK(hipMemcpyAsync(sl.small_host, sl.ds.claims, (size_t)ctx->R * ctx->W * 4,
                                          hipMemcpyDeviceToHost, sl.st));
            GKRX_HIP_CHECK(hipMemcpyAsync(sl.small_host + 64, sl.ds.tstate, 40, hipMemcpyDeviceToHost, sl.st));
        }
    } else {
        sl.items.resize(sl.reqs.size());
        for (size_t i = 0; i < sl.reqs.size(); i++) {
            SrvReq *r = sl.reqs[i];
            sl.items[i] = BatchItem{&r->c->dev, r->rz1.data(), r->tstate, r->grind, tail_bytes(ctx, r->c),
                                    r->proof + commit_bytes(ctx, r->c)};
        }
        int rc = dev_prove_batch(sl.items, ctx->R, kind, sl.bs, sl.st, nullptr);
        if (rc) return rc;
    }
    GKRX_HIP_CHECK(hipEventRecord(sl.ev, sl.st));
    return GKRX_OK;
}

// Hands the results of a finished slot to its requests (S->m not held); returns a status.
int collect(gkrx_server *S, Slot &sl)
{
    gkrx_ctx *ctx = S->ctx;
    const size_t cw = (size_t)ctx->R * ctx->W;
    if (sl.single) {
        SrvReq *r = sl.reqs[0];
        const uint64_t tb = tail_bytes(ctx, r->c);
        if (sl.small_host[64] != tb) {
            set_error("device transcript length mismatch");
            return GKRX_ERR_DEVICE;
        }
        memcpy(r->proof + commit_bytes(ctx, r->c), sl.tail_host, tb);
        memcpy(r->claims, sl.small_host, cw * 4);
    } else {
        if (int rc = dev_batch_collect(sl.items, sl.bs)) return rc;
        for (size_t i = 0; i < sl.reqs.size(); i++) memcpy(sl.reqs[i]->claims, sl.bs.claims_host + i * cw, cw * 4);
    }
    return GKRX_OK;
}

double ms_between(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// seen: when the slot's completion was observed (before its results were collected)
void finish_reqs(gkrx_server *S, std::vector<SrvReq *> &rs, int rc, Clock::time_point seen)
{
    const auto now = Clock::now();
    for (SrvReq *r : rs) {
        r->rc = rc;
        r->len = rc ? 0 : commit_bytes(S->ctx, r->c) + tail_bytes(S->ctx, r->c);
        r->lat_ms = ms_between(r->t_submit, now);
        r->t_seen = seen;
        const Clock::time_point marks[6] = {r->t_submit, r->t_prefix0, r->t_prefix1, r->t_launch, seen, now};
        for (int i = 0; i < 5; i++) r->split_ms[i] = std::max(0.0, ms_between(marks[i], marks[i + 1]));
        r->split_ms[5] = r->lat_ms;
        r->state = 2;
        if (!rc) r->c->evaluated = true;
        r->c->in_server = false;
        S->in_use -= r->bytes;
        S->completed++;
    }
}

// Takes the work for a free slot from the ready queue, oldest first: the first request of
// the slot's kind that fits (single slot), or every request of its kind in order while the
// budget lasts (batch slot).  A request that does not fit stops its kind's admission, so
// large requests are not starved by small ones behind them.
std::vector<SrvReq *> pick(gkrx_server *S, const Slot &sl)
{
    std::vector<SrvReq *> out;
    const uint32_t cap = sl.single ? 1u : kMaxBatch;
    for (auto it = S->ready.begin(); it != S->ready.end() && out.size() < cap;) {
        SrvReq *r = *it;
        if (r->single != sl.single) {
            ++it;
            continue;
        }
        if (S->in_use + r->bytes > S->budget) break;
        S->in_use += r->bytes;
        r->state = 1;
        out.push_back(r);
        it = S->ready.erase(it);
    }
    return out;
}

void scheduler(gkrx_server *S)
{
    (void)hipSetDevice(S->ctx->device);
    std::unique_lock<std::mutex> lk(S->m);
    for (;;) {
        // completions: slots whose event has been reached (queried without the lock)
        std::vector<Slot *> busy;
        for (auto &sp : S->slots)
            if (sp->busy) busy.push_back(sp.get());
        bool progressed = false;
        if (!busy.empty()) {
            lk.unlock();
            struct Done {
                Slot *sl;
                int rc;
                Clock::time_point seen;
            };
            std::vector<Done> finished;
            for (Slot *sl : busy) {
                const hipError_t q = hipEventQuery(sl->ev);
                const auto seen = Clock::now();
                if (q == hipSuccess) {
                    finished.push_back({sl, collect(S, *sl), seen});
                } else if (q != hipErrorNotReady) {
                    (void)hipStreamSynchronize(sl->st);
                    finished.push_back({sl, GKRX_ERR_DEVICE, seen});
                }
            }
            lk.lock();
            for (auto &f : finished) {
                finish_reqs(S, f.sl->reqs, f.rc, f.seen);
                f.sl->reqs.clear();
                f.sl->busy = false;
                progressed = true;
            }
        }
        if (progressed) S->done_cv.notify_all();
        // launches on free slots
        bool launched = false;
        for (auto &sp : S->slots) {
            Slot &sl = *sp;
            if (sl.busy || S->ready.empty()) continue;
            std::vector<SrvReq *> batch = pick(S, sl);
            if (batch.empty()) continue;
            S->peak = std::max(S->peak, S->in_use);
            S->batches++;
            S->max_batch = std::max<uint64_t>(S->max_batch, batch.size());
            sl.busy = true;
            sl.reqs = batch;
            const auto t_launch = Clock::now();
            for (SrvReq *r : batch) r->t_launch = t_launch;
            lk.unlock();
            int rc = launch(S, sl);
            if (rc) (void)hipStreamSynchronize(sl.st);  // no completion event: drain the slot
            lk.lock();
            if (rc) {
                finish_reqs(S, sl.reqs, rc, Clock::now());
                sl.reqs.clear();
                sl.busy = false;
                S->done_cv.notify_all();
            }
            launched = true;
        }
        if (launched) continue;
        bool any_busy = false;
        for (auto &sp : S->slots) any_busy |= sp->busy;
        if (S->stop && S->pq.empty() && S->ready.empty() && S->prefixing == 0 && !any_busy) break;
        if (any_busy) S->cv.wait_for(lk, kPollPeriod);
        else S->cv.wait(lk);
    }
}

void free_slot(Slot &sl)
{
    if (sl.st) (void)hipStreamSynchronize(sl.st);
    dev_free_batch(sl.bs);
    dev_free_scratch(sl.ds);
    dev_bn_free_scratch(sl.bn);
    if (sl.tail_host) (void)hipHostFree(sl.tail_host);
    if (sl.small_host) (void)hipHostFree(sl.small_host);
    if (sl.ev) (void)hipEventDestroy(sl.ev);
    if (sl.st) (void)hipStreamDestroy(sl.st);
}

}  // namespace

extern "C" int gkrx_server_create(gkrx_ctx *ctx, uint64_t hbm_budget_bytes, uint32_t slots, gkrx_server **out)
try {
    if (!ctx || !out || hbm_budget_bytes == 0 || slots == 0 || slots > 16) return GKRX_ERR_ARG;
    GKRX_HIP_CHECK(hipSetDevice(ctx->device));
    auto *S = new (std::nothrow) gkrx_server;
    if (!S) return GKRX_ERR_NOMEM;
    S->ctx = ctx;
    S->budget = hbm_budget_bytes;
    // M31: `slots` batch slots and one single slot for large proofs; BN254: `slots` single slots
    const uint32_t total = ctx->bn254 ? slots : slots + 1;
    for (uint32_t i = 0; i < total; i++) {
        auto sl = std::make_unique<Slot>();
        sl->single = ctx->bn254 || i == slots;
        // the context's batch engine (gkrx_ctx_set_batch_engine); a slot runs its batch on one
        // stream, where the grouped lockstep engine would only serialize its groups' separate
        // launches, so that engine runs as the plain lockstep engine here
        sl->bs.engine = ctx->batch.engine == kBatchLockstepGrouped ? kBatchLockstep : ctx->batch.engine;
        sl->bs.block = ctx->batch.block;
        sl->bs.lanes = ctx->batch.lanes;
        sl->bs.chain_stream = false;  // slots already overlap; stay within the hardware queues
        // an M31 single slot (circuits wider than the batch engines take) gets its stream on
        // first use, so that an idle one holds no hardware queue
        if (sl->single && !ctx->bn254) {
            S->slots.push_back(std::move(sl));
            continue;
        }
        if (hipStreamCreateWithFlags(&sl->st, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl->ev, hipEventDisableTiming) != hipSuccess) {
            free_slot(*sl);
            for (auto &p : S->slots) free_slot(*p);
            delete S;
            set_error("hipStreamCreate / hipEventCreate failed");
            return GKRX_ERR_DEVICE;
        }
        S->slots.push_back(std::move(sl));
    }
    ctx->live_circuits++;  // the context outlives the server
    const unsigned nw = std::max(1u, std::min(12u, usable_cpus()));
    for (unsigned i = 0; i < nw; i++) S->workers.emplace_back(prefix_worker, S);
    S->sched = std::thread(scheduler, S);
    *out = S;
    return GKRX_OK;
} catch (const std::bad_alloc &) {
    set_error("out of host memory");
    return GKRX_ERR_NOMEM;
}

extern "C" int gkrx_server_submit(gkrx_server *S, gkrx_circuit *c, uint8_t *proof, size_t proof_cap,
                                  uint32_t *claimed_v, uint64_t *ticket)
try {
    if (!S || !c || !proof || !claimed_v || !ticket || c->ctx != S->ctx) return GKRX_ERR_ARG;
    gkrx_ctx *ctx = S->ctx;
    if (!c->has_witness) {
        set_error("no witness set");
        return GKRX_ERR_STATE;
    }
    if (proof_cap < commit_bytes(ctx, c) + tail_bytes(ctx, c)) {
        set_error("proof buffer too small");
        return GKRX_ERR_ARG;
    }
    uint64_t rows = 0;
    for (uint32_t l = 0; l < c->D; l++) rows += 1ull << c->n[l];
    // the lockstep engine gives a large proof as many workgroups as its tables need, so it
    // joins a batch; the one-workgroup-per-proof engine sends it to a single slot
    const bool single = ctx->bn254 || !dev_batch_supported(c->dev) ||
                        (ctx->batch.engine == kBatchPerProof && rows >= kBigRows);
    const uint64_t bytes = footprint(ctx, c);
    if (bytes > S->budget) {
        set_error("request needs " + std::to_string(bytes) + " bytes of HBM, above the server budget");
        return GKRX_ERR_NOMEM;
    }
    bool expect = false;
    if (!c->in_server.compare_exchange_strong(expect, true)) {
        set_error("circuit already has a request in flight");
        return GKRX_ERR_STATE;
    }
    auto *r = new (std::nothrow) SrvReq;
    if (!r) {
        c->in_server = false;
        return GKRX_ERR_NOMEM;
    }
    r->c = c;
    r->proof = proof;
    r->claims = claimed_v;
    r->bytes = bytes;
    r->single = single;
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(S->m);
        if (S->stop) {
            c->in_server = false;
            delete r;
            set_error("server is shutting down");
            return GKRX_ERR_STATE;
        }
        r->id = S->next_id++;
        S->reqs[r->id] = r;
        S->pq.push_back(r);
        S->submitted++;
        *ticket = r->id;
    }
    S->pcv.notify_one();
    return GKRX_OK;
} catch (const std::bad_alloc &) {
    set_error("out of host memory");
    return GKRX_ERR_NOMEM;
}

extern "C" int gkrx_server_wait_split(gkrx_server *S, uint64_t ticket, int64_t timeout_us, size_t *proof_len,
                                      double split_ms[6])
{
    if (!S || !split_ms) return GKRX_ERR_ARG;
    std::unique_lock<std::mutex> lk(S->m);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) {
        set_error("unknown ticket");
        return GKRX_ERR_ARG;
    }
    SrvReq *r = it->second;
    auto ready = [&] { return r->state == 2; };
    if (timeout_us < 0) S->done_cv.wait(lk, ready);
    else if (!S->done_cv.wait_for(lk, std::chrono::microseconds(timeout_us), ready)) return 1;
    const int rc = r->rc;
    if (proof_len) *proof_len = r->len;
    for (int i = 0; i < 6; i++) split_ms[i] = r->split_ms[i];
    S->reqs.erase(it);
    delete r;
    return rc;
}

extern "C" int gkrx_server_wait(gkrx_server *S, uint64_t ticket, int64_t timeout_us, size_t *proof_len,
                                double *latency_ms)
{
    double split[6] = {};
    const int rc = gkrx_server_wait_split(S, ticket, timeout_us, proof_len, split);
    if (rc != 1 && latency_ms) *latency_ms = split[5];
    return rc;
}

extern "C" int gkrx_server_stats(gkrx_server *S, uint64_t out[6])
{
    if (!S || !out) return GKRX_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->m);
    out[0] = S->submitted;
    out[1] = S->completed;
    out[2] = S->batches;
    out[3] = S->max_batch;
    out[4] = S->peak;
    out[5] = S->in_use;
    return GKRX_OK;
}

extern "C" int gkrx_server_destroy(gkrx_server *S)
{
    if (!S) return GKRX_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(S->m);
        S->stop = true;
    }
    S->pcv.notify_all();
    S->cv.notify_all();
    for (auto &w : S->workers) w.join();
    if (S->sched.joinable()) S->sched.join();
    (void)hipSetDevice(S->ctx->device);
    for (auto &sp : S->slots) free_slot(*sp);
    for (auto &kv : S->reqs) delete kv.second;
    S->ctx->live_circuits--;
    delete S;
    return GKRX_OK;
}
