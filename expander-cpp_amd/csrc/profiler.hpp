// Optional per-kernel-class timing with HIP events recorded on the launch stream.
// Enabled per context (gkrx_ctx_profile, gkrx_ctx_profile_classes); when off, or for a
// class outside the mask, begin/end are no-ops (each recorded event adds a marker packet
// to the stream, so timing every launch slows a proof down measurably).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace gkrx {

enum KernelClass : int {
    kKForward = 0,
    kKEq,
    kKGx,
    kKHy,
    kKDot,
    kKPairPh1First,
    kKPairPh1,
    kKPairPh2First,
    kKPairPh2,
    kKTailPh1,
    kKTailPh2,
    kKBatch,
    kKPairFin,
    kKLsWork,
    kKLsFin,
    kKLsChain,
    kKLsFwd,
    kKLsGx,
    kKLsHy,
    kKLsEq,
    kKLsPair,
    kKRoundPh1First,  // single rounds (BN254)
    kKRoundPh1,
    kKRoundPh2First,
    kKRoundPh2,
    kKRoundFin,
    kKClassCount
};

inline const char *kernel_class_name(int c)
{
    static const char *names[kKClassCount] = {"k_forward",         "k_eq",          "k_gx",
                                              "k_hy",              "k_dot",         "k_pair_ph1_first",
                                              "k_pair_ph1",        "k_pair_ph2_first", "k_pair_ph2",
                                              "k_tail_ph1",        "k_tail_ph2",    "k_prove_batch",
                                              "k_pair_fin",        "k_ls_work",     "k_ls_fin",
                                              "k_ls_chain",        "k_ls_work:forward", "k_ls_work:g_x",
                                              "k_ls_work:h_y",     "k_ls_work:eq",  "k_ls_work:pair",
                                              "k_round_ph1_first", "k_round_ph1",   "k_round_ph2_first",
                                              "k_round_ph2",       "k_round_fin"};
    return (c >= 0 && c < kKClassCount) ? names[c] : "";
}

struct Profiler {
    bool on = false;
    uint64_t mask = ~0ull;  // classes that are timed (bit c = class c)
    bool open = false;      // begin recorded an event that end must close
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    std::vector<int> cls;  // class of event pair i (events 2i, 2i+1)
    uint64_t count[kKClassCount] = {0};
    double ms[kKClassCount] = {0};

    hipEvent_t next()
    {
        if (used == pool.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            pool.push_back(e);
        }
        return pool[used++];
    }
    void begin(int c, hipStream_t st)
    {
        open = false;
        if (!on || !((mask >> c) & 1)) return;
        hipEvent_t e = next();
        if (e) {
            (void)hipEventRecord(e, st);
            cls.push_back(c);
            open = true;
        }
    }
    void end(hipStream_t st)
    {
        if (!open) return;
        open = false;
        hipEvent_t e = next();
        if (e) (void)hipEventRecord(e, st);
    }
    // A start / stop event pair for a launch that records them itself
    // (hipExtLaunchKernelGGL: the kernel's own start and end, no marker packets in the
    // stream); false when class c is not timed.
    bool kernel_events(int c, hipEvent_t &e0, hipEvent_t &e1)
    {
        if (!on || !((mask >> c) & 1)) return false;
        e0 = next();
        e1 = e0 ? next() : nullptr;
        if (!e1) {
            if (e0) used--;
            return false;
        }
        cls.push_back(c);
        return true;
    }
    // After a prove (its streams synchronized): the timings are read only once many event
    // pairs are pending, or when the statistics are asked for, so that a timed loop does
    // not read hundreds of events between its proves.
    static constexpr size_t kPendingEvents = 8192;
    void settle()
    {
        if (used >= kPendingEvents) collect();
    }
    // after the streams that recorded the pending events have been synchronized
    void collect()
    {
        for (size_t i = 0; i < cls.size() && 2 * i + 1 < used; i++) {
            float t = 0;
            if (hipEventElapsedTime(&t, pool[2 * i], pool[2 * i + 1]) == hipSuccess) {
                count[cls[i]]++;
                ms[cls[i]] += t;
            }
        }
        cls.clear();
        used = 0;
    }
    void reset()
    {
        cls.clear();
        used = 0;
        for (int i = 0; i < kKClassCount; i++) {
            count[i] = 0;
            ms[i] = 0;
        }
    }
    ~Profiler()
    {
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

}  // namespace gkrx
