// Continuous-batching prove server (gkrx_server_*): the serving loop of the library.
//
// The reference proves one circuit per Prover::prove call and has no serving layer; this
// is the build's own.  Requests (a circuit with its witness resident in HBM) queue in
// arrival order.  A scheduler thread admits them while the HBM they need -- the circuit's
// resident values and wiring plus the prover tables, both computed from the circuit shape
// -- fits the budget, and launches everything admitted as one batch on the first free
// slot (a stream with its own scratch).  Several slots run at once, so requests that
// arrive while a batch is on the GPU start on the next free slot instead of waiting for
// that batch to finish, and their work overlaps on the device.  A batch is one
// k_prove_batch launch (one workgroup per proof, M31); BN254 requests run one per slot
// through the multi-block BN254 path.  Each proof is the same bytes as gkrx_prove of the
// same circuit: a proof's transcript depends only on its own circuit and witness.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "handles.hpp"
#include "host_hash.hpp"

using Clock = std::chrono::steady_clock;

namespace {

constexpr uint32_t kMaxBatch = 1024;

struct SrvReq {
    uint64_t id = 0;
    gkrx_circuit *c = nullptr;
    uint8_t *proof = nullptr;
    uint32_t *claims = nullptr;
    uint64_t bytes = 0;  // admission footprint
    Clock::time_point t_submit;
    int state = 0;       // 0 queued, 1 on the GPU, 2 done
    int rc = 0;
    size_t len = 0;
    double lat_ms = 0;
    std::vector<uint32_t> rz1;
    uint32_t tstate[10] = {};
    uint8_t grind[32] = {};
};

struct Slot {
    gkrx_server *srv = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;  // recorded after a batch's last copy; the scheduler polls it
    gkrx::BatchScratch bs;
    gkrx::BnScratch bn;
    uint8_t *bn_tail_host = nullptr;
    uint64_t bn_tail_cap = 0;
    uint32_t *bn_claims_host = nullptr;
    std::vector<SrvReq *> reqs;
    std::vector<gkrx::BatchItem> items;
    bool busy = false;
};

// How often the scheduler polls the slot events while batches are in flight.  Completion
// is observed by polling, not by a stream host-function callback: the runtime delivers such
// callbacks only while some host thread is inside it, and a server whose threads all sleep
// waiting for the callback never sees its batches finish.
constexpr auto kPollPeriod = std::chrono::microseconds(100);

}  // namespace

struct gkrx_server {
    gkrx_ctx *ctx = nullptr;
    uint64_t budget = 0, in_use = 0, peak = 0;
    std::vector<std::unique_ptr<Slot>> slots;
    std::mutex m;
    std::condition_variable cv, done_cv;
    std::deque<SrvReq *> queue;
    std::unordered_map<uint64_t, SrvReq *> reqs;
    uint64_t next_id = 1;
    bool stop = false;
    uint64_t submitted = 0, completed = 0, batches = 0, max_batch = 0;
    std::thread sched;
};

using namespace gkrx;

namespace {

uint64_t footprint(const gkrx_ctx *ctx, const gkrx_circuit *c)
{
    const uint64_t tb = tail_bytes(ctx, c);
    return dev_circuit_bytes(c->dev) +
           (ctx->bn254 ? dev_bn_scratch_bytes(c->dev, tb) : dev_batch_job_bytes(c->dev, ctx->R, tb));
}

void host_prefixes(gkrx_ctx *ctx, std::vector<SrvReq *> &rs)
{
    const int kind = ctx->cfg.fs_hash == GKRX_FS_KECCAK256 ? 1 : 0;
    std::atomic<size_t> next{0};
    auto work = [&] {
        for (size_t k; (k = next++) < rs.size();) {
            SrvReq *r = rs[k];
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
    };
    const size_t nt = std::min<size_t>(rs.size(), 16);
    std::vector<std::thread> pool;
    for (size_t t = 1; t < nt; t++) pool.emplace_back(work);
    work();
    for (auto &th : pool) th.join();
}

// Enqueues the slot's batch; returns a status (nothing is enqueued on failure).
int launch(gkrx_server *S, Slot &sl)
{
    gkrx_ctx *ctx = S->ctx;
    const int kind = ctx->cfg.fs_hash == GKRX_FS_KECCAK256 ? 1 : 0;
    if (ctx->bn254) {
        SrvReq *r = sl.reqs[0];
        gkrx_circuit *c = r->c;
        const uint64_t tb = tail_bytes(ctx, c);
        int rc = dev_bn_ensure_scratch(sl.bn, c->dev, tb);
        if (rc) return rc;
        if (tb > sl.bn_tail_cap) {
            if (sl.bn_tail_host) (void)hipHostFree(sl.bn_tail_host);
            sl.bn_tail_host = nullptr;
            sl.bn_tail_cap = 0;
            GKRX_HIP_CHECK(hipHostMalloc(&sl.bn_tail_host, tb, hipHostMallocDefault));
            sl.bn_tail_cap = tb;
        }
        if (!sl.bn_claims_host) GKRX_HIP_CHECK(hipHostMalloc(&sl.bn_claims_host, 32, hipHostMallocDefault));
        rc = dev_bn_forward(c->dev, sl.st, nullptr);  // overlaps the host prefix
        if (rc) return rc;
        host_prefixes(ctx, sl.reqs);
        rc = dev_bn_prove(c->dev, sl.bn, kind, r->rz1.data(), r->tstate, r->grind, 32, sl.st, nullptr);
        if (rc) return rc;
        GKRX_HIP_CHECK(hipMemcpyAsync(sl.bn_tail_host, sl.bn.tail, tb, hipMemcpyDeviceToHost, sl.st));
        GKRX_HIP_CHECK(hipMemcpyAsync(sl.bn_claims_host, sl.bn.claims, 32, hipMemcpyDeviceToHost, sl.st));
    } else {
        host_prefixes(ctx, sl.reqs);
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

// Hands the results of a finished slot to its requests (S->m not held).
void collect(gkrx_server *S, Slot &sl)
{
    gkrx_ctx *ctx = S->ctx;
    const size_t cw = (size_t)ctx->R * ctx->W;
    if (ctx->bn254) {
        SrvReq *r = sl.reqs[0];
        memcpy(r->proof + commit_bytes(ctx, r->c), sl.bn_tail_host, tail_bytes(ctx, r->c));
        memcpy(r->claims, sl.bn_claims_host, 32);
    } else {
        dev_batch_collect(sl.items, sl.bs);
        for (size_t i = 0; i < sl.reqs.size(); i++) memcpy(sl.reqs[i]->claims, sl.bs.claims_host + i * cw, cw * 4);
    }
}

void finish_reqs(gkrx_server *S, std::vector<SrvReq *> &rs, int rc)
{
    const auto now = Clock::now();
    for (SrvReq *r : rs) {
        r->rc = rc;
        r->len = rc ? 0 : commit_bytes(S->ctx, r->c) + tail_bytes(S->ctx, r->c);
        r->lat_ms = std::chrono::duration<double, std::milli>(now - r->t_submit).count();
        r->state = 2;
        if (!rc) r->c->evaluated = true;
        r->c->in_server = false;
        S->in_use -= r->bytes;
        S->completed++;
    }
}

void scheduler(gkrx_server *S)
{
    (void)hipSetDevice(S->ctx->device);
    std::unique_lock<std::mutex> lk(S->m);
    for (;;) {
        // completions: the slots whose event has been reached (queried without the lock)
        std::vector<Slot *> busy;
        for (auto &sp : S->slots)
            if (sp->busy) busy.push_back(sp.get());
        bool progressed = false;
        if (!busy.empty()) {
            lk.unlock();
            std::vector<Slot *> finished;
            for (Slot *sl : busy) {
                const hipError_t q = hipEventQuery(sl->ev);
                if (q == hipSuccess) {
                    collect(S, *sl);
                    finished.push_back(sl);
                } else if (q != hipErrorNotReady) {
                    (void)hipStreamSynchronize(sl->st);
                    finished.push_back(sl);
                    sl->reqs.front()->rc = GKRX_ERR_DEVICE;  // marks the whole slot failed
                }
            }
            lk.lock();
            for (Slot *sl : finished) {
                const int rc = sl->reqs.front()->rc;
                finish_reqs(S, sl->reqs, rc);
                sl->reqs.clear();
                sl->busy = false;
                progressed = true;
            }
        }
        if (progressed) S->done_cv.notify_all();
        Slot *free_slot = nullptr;
        for (auto &sp : S->slots)
            if (!sp->busy) {
                free_slot = sp.get();
                break;
            }
        if (free_slot && !S->queue.empty()) {
            std::vector<SrvReq *> batch;
            const uint32_t cap = S->ctx->bn254 ? 1u : kMaxBatch;
            while (!S->queue.empty() && batch.size() < cap && S->in_use + S->queue.front()->bytes <= S->budget) {
                SrvReq *r = S->queue.front();
                S->queue.pop_front();
                r->state = 1;
                S->in_use += r->bytes;
                batch.push_back(r);
            }
            if (!batch.empty()) {
                S->peak = std::max(S->peak, S->in_use);
                S->batches++;
                S->max_batch = std::max<uint64_t>(S->max_batch, batch.size());
                for (SrvReq *r : batch) r->rc = GKRX_OK;
                free_slot->busy = true;
                free_slot->reqs = batch;
                lk.unlock();
                const int rc = launch(S, *free_slot);
                lk.lock();
                if (rc) {
                    // the completion event was not recorded: drain the slot, fail its requests
                    lk.unlock();
                    (void)hipStreamSynchronize(free_slot->st);
                    lk.lock();
                    finish_reqs(S, free_slot->reqs, rc);
                    free_slot->reqs.clear();
                    free_slot->busy = false;
                    S->done_cv.notify_all();
                }
                continue;
            }
        }
        bool any_busy = false;
        for (auto &sp : S->slots) any_busy |= sp->busy;
        if (S->stop && S->queue.empty() && !any_busy) break;
        if (any_busy) S->cv.wait_for(lk, kPollPeriod);
        else S->cv.wait(lk);
    }
}

}  // namespace

extern "C" int gkrx_server_create(gkrx_ctx *ctx, uint64_t hbm_budget_bytes, uint32_t slots, gkrx_server **out)
{
    if (!ctx || !out || hbm_budget_bytes == 0 || slots == 0 || slots > 16) return GKRX_ERR_ARG;
    GKRX_HIP_CHECK(hipSetDevice(ctx->device));
    auto *S = new (std::nothrow) gkrx_server;
    if (!S) return GKRX_ERR_NOMEM;
    S->ctx = ctx;
    S->budget = hbm_budget_bytes;
    for (uint32_t i = 0; i < slots; i++) {
        auto sl = std::make_unique<Slot>();
        sl->srv = S;
        if (hipStreamCreateWithFlags(&sl->st, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl->ev, hipEventDisableTiming) != hipSuccess) {
            if (sl->st) (void)hipStreamDestroy(sl->st);
            for (auto &p : S->slots) {
                (void)hipEventDestroy(p->ev);
                (void)hipStreamDestroy(p->st);
            }
            delete S;
            set_error("hipStreamCreate / hipEventCreate failed");
            return GKRX_ERR_DEVICE;
        }
        S->slots.push_back(std::move(sl));
    }
    ctx->live_circuits++;  // the context outlives the server
    S->sched = std::thread(scheduler, S);
    *out = S;
    return GKRX_OK;
}

extern "C" int gkrx_server_submit(gkrx_server *S, gkrx_circuit *c, uint8_t *proof, size_t proof_cap,
                                  uint32_t *claimed_v, uint64_t *ticket)
{
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
    if (!ctx->bn254 && !dev_batch_supported(c->dev)) {
        set_error("layers wider than 2^20 are not served (prove them with gkrx_prove)");
        return GKRX_ERR_CIRCUIT;
    }
    auto *r = new (std::nothrow) SrvReq;
    if (!r) return GKRX_ERR_NOMEM;
    r->c = c;
    r->proof = proof;
    r->claims = claimed_v;
    r->bytes = footprint(ctx, c);
    if (r->bytes > S->budget) {
        const uint64_t need = r->bytes;
        delete r;
        set_error("request needs " + std::to_string(need) + " bytes of HBM, above the server budget");
        return GKRX_ERR_NOMEM;
    }
    bool expect = false;
    if (!c->in_server.compare_exchange_strong(expect, true)) {
        delete r;
        set_error("circuit already has a request in flight");
        return GKRX_ERR_STATE;
    }
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(S->m);
        if (S->stop) {
            c->in_server = false;
            delete r;
            return GKRX_ERR_STATE;
        }
        r->id = S->next_id++;
        S->reqs[r->id] = r;
        S->queue.push_back(r);
        S->submitted++;
        *ticket = r->id;
    }
    S->cv.notify_all();
    return GKRX_OK;
}

extern "C" int gkrx_server_wait(gkrx_server *S, uint64_t ticket, int64_t timeout_us, size_t *proof_len,
                                double *latency_ms)
{
    if (!S) return GKRX_ERR_ARG;
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
    if (latency_ms) *latency_ms = r->lat_ms;
    S->reqs.erase(it);
    delete r;
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
    S->cv.notify_all();
    if (S->sched.joinable()) S->sched.join();
    (void)hipSetDevice(S->ctx->device);
    for (auto &sp : S->slots) {
        Slot &sl = *sp;
        (void)hipStreamSynchronize(sl.st);
        dev_free_batch(sl.bs);
        dev_bn_free_scratch(sl.bn);
        if (sl.bn_tail_host) (void)hipHostFree(sl.bn_tail_host);
        if (sl.bn_claims_host) (void)hipHostFree(sl.bn_claims_host);
        (void)hipEventDestroy(sl.ev);
        (void)hipStreamDestroy(sl.st);
    }
    for (auto &kv : S->reqs) delete kv.second;
    S->ctx->live_circuits--;
    delete S;
    return GKRX_OK;
}
