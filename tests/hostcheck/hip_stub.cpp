// A host-only stand-in for the HIP runtime, so that the library's host code (the C ABI, the
// batch engines' scheduling, staging and stream/event handling, the server) can run under
// AddressSanitizer / UndefinedBehaviorSanitizer on a machine with no GPU.  Device memory is
// host memory, copies are memcpy, kernel launches only have their shape checked, and every
// stream, event and allocation handle is tracked: using one that was never created or was
// already destroyed aborts with a message (the lifetime bugs this build is meant to find).
// Nothing here runs a kernel, so the proofs the library returns are wrong by construction;
// the driver checks only that every call returns in an orderly way.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <set>

namespace {
std::mutex mu;
std::set<const void *> streams, events, dev_allocs, host_allocs;
int device = 0;

[[noreturn]] void fail(const char *what, const void *p)
{
    std::fprintf(stderr, "hip_stub: %s (%p)\n", what, p);
    std::abort();
}
void check_stream(hipStream_t s)
{
    if (!s) return;  // the null stream
    std::lock_guard<std::mutex> lk(mu);
    if (!streams.count(s)) fail("use of a stream that does not exist", s);
}
void check_event(hipEvent_t e)
{
    std::lock_guard<std::mutex> lk(mu);
    if (!events.count(e)) fail("use of an event that does not exist", e);
}
void check_dim(dim3 g, dim3 b)
{
    if (!g.x || !g.y || !g.z || !b.x || !b.y || !b.z) fail("launch with an empty grid or block", nullptr);
    if ((uint64_t)b.x * b.y * b.z > 1024) fail("launch with more than 1024 threads per block", nullptr);
    if (g.y > 65535 || g.z > 65535) fail("launch with grid y or z above 65535", nullptr);
}
}  // namespace

extern "C" {
// module registration: nothing to register
void **__hipRegisterFatBinary(const void *) { static void *h = nullptr; return &h; }
void __hipUnregisterFatBinary(void **) {}
void __hipRegisterFunction(void **, const void *, char *, const char *, unsigned int, void *, void *, void *, void *,
                           int *) {}
void __hipRegisterVar(void **, void *, char *, const char *, int, size_t, int, int) {}
hipError_t __hipPushCallConfiguration(dim3 g, dim3 b, size_t, hipStream_t st)
{
    check_dim(g, b);
    check_stream(st);
    return hipSuccess;
}
hipError_t __hipPopCallConfiguration(dim3 *g, dim3 *b, size_t *shm, hipStream_t *st)
{
    *g = dim3(1);
    *b = dim3(1);
    *shm = 0;
    *st = nullptr;
    return hipSuccess;
}
// a per-file fat binary symbol (see the makefile) points here
extern const char gkrx_stub_fatbin[16] = {};
}

hipError_t hipLaunchKernel(const void *, dim3 g, dim3 b, void **, size_t, hipStream_t st)
{
    check_dim(g, b);
    check_stream(st);
    return hipSuccess;
}
hipError_t hipExtLaunchKernel(const void *, dim3 g, dim3 b, void **, size_t, hipStream_t st, hipEvent_t e0,
                              hipEvent_t e1, int)
{
    check_dim(g, b);
    check_stream(st);
    if (e0) check_event(e0);
    if (e1) check_event(e1);
    return hipSuccess;
}

hipError_t hipMalloc(void **p, size_t n)
{
    *p = std::calloc(1, n ? n : 1);
    if (!*p) return hipErrorOutOfMemory;
    std::lock_guard<std::mutex> lk(mu);
    dev_allocs.insert(*p);
    return hipSuccess;
}
hipError_t hipFree(void *p)
{
    if (!p) return hipSuccess;
    {
        std::lock_guard<std::mutex> lk(mu);
        if (!dev_allocs.erase(p)) fail("hipFree of memory hipMalloc did not return", p);
    }
    std::free(p);
    return hipSuccess;
}
hipError_t hipHostMalloc(void **p, size_t n, unsigned int)
{
    *p = std::calloc(1, n ? n : 1);
    if (!*p) return hipErrorOutOfMemory;
    std::lock_guard<std::mutex> lk(mu);
    host_allocs.insert(*p);
    return hipSuccess;
}
hipError_t hipHostFree(void *p)
{
    if (!p) return hipSuccess;
    {
        std::lock_guard<std::mutex> lk(mu);
        if (!host_allocs.erase(p)) fail("hipHostFree of memory hipHostMalloc did not return", p);
    }
    std::free(p);
    return hipSuccess;
}
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind)
{
    if (n) std::memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t st)
{
    check_stream(st);
    if (n) std::memmove(d, s, n);
    return hipSuccess;
}

hipError_t hipStreamCreate(hipStream_t *s) { return hipStreamCreateWithFlags(s, 0); }
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned int)
{
    *s = reinterpret_cast<hipStream_t>(std::malloc(8));
    std::lock_guard<std::mutex> lk(mu);
    streams.insert(*s);
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s)
{
    {
        std::lock_guard<std::mutex> lk(mu);
        if (!streams.erase(s)) fail("hipStreamDestroy of a stream that does not exist", s);
    }
    std::free(s);
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s)
{
    check_stream(s);
    return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned int)
{
    check_stream(s);
    check_event(e);
    return hipSuccess;
}
hipError_t hipDeviceSynchronize() { return hipSuccess; }

hipError_t hipEventCreate(hipEvent_t *e) { return hipEventCreateWithFlags(e, 0); }
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned int)
{
    *e = reinterpret_cast<hipEvent_t>(std::malloc(8));
    std::lock_guard<std::mutex> lk(mu);
    events.insert(*e);
    return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e)
{
    {
        std::lock_guard<std::mutex> lk(mu);
        if (!events.erase(e)) fail("hipEventDestroy of an event that does not exist", e);
    }
    std::free(e);
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s)
{
    check_event(e);
    check_stream(s);
    return hipSuccess;
}
hipError_t hipEventQuery(hipEvent_t e)
{
    check_event(e);
    return hipSuccess;
}
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b)
{
    check_event(a);
    check_event(b);
    *ms = 0.f;
    return hipSuccess;
}

hipError_t hipSetDevice(int d)
{
    if (d != 0) return hipErrorInvalidDevice;
    device = d;
    return hipSuccess;
}
hipError_t hipGetDevice(int *d)
{
    *d = device;
    return hipSuccess;
}
hipError_t hipGetDeviceCount(int *n)
{
    *n = 1;
    return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t, int)
{
    *v = 256;
    return hipSuccess;
}
hipError_t hipGetLastError() { return hipSuccess; }
const char *hipGetErrorString(hipError_t) { return "hip_stub error"; }
hipError_t hipGetSymbolAddress(void **p, const void *sym)
{
    *p = const_cast<void *>(sym);  // the host shadow of a __device__ array is host memory here
    return hipSuccess;
}

// HIP graphs: not provided (the driver leaves them off)
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorNotSupported; }
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *) { return hipErrorNotSupported; }
hipError_t hipGraphInstantiate(hipGraphExec_t *, hipGraph_t, hipGraphNode_t *, char *, size_t)
{
    return hipErrorNotSupported;
}
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorNotSupported; }
hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
