// Library-owned device resources and their teardown: the work vectors of the
// plain form, the node counters of the persistent kernels, and gb_shutdown,
// which releases these together with the batch buffer sets of the
// branch-and-bound driver and the fork streams of gb_relax_instances.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <map>
#include <mutex>

#include "gb_host.h"
#include "gb_slots.h"

namespace gb {

namespace {

constexpr int kMaxDevices = 64;

// Grow-only workspace for the per-node work vectors, one per (device,
// stream): launches on different streams never share work vectors, and
// launches on one stream are ordered.
struct Workspace {
    std::mutex mtx;
    double* ptr = nullptr;
    size_t bytes = 0;
};
std::mutex g_ws_mtx;
std::map<std::pair<int, void*>, Workspace>* g_ws = new std::map<std::pair<int, void*>, Workspace>();

// Node counters: 64 bytes each, allocated in chunks, lent per call by the
// rules of SlotTable.
constexpr int kCounterBytes = 64;
constexpr int kCounterChunk = 256;     // counters per allocation
constexpr size_t kCaptureReserve = 32; // never-used counters kept for captured calls

struct CounterPool {
    std::mutex mtx;
    std::vector<char*> chunks;
    SlotTable<hipEvent_t> table;
};
CounterPool g_counters[kMaxDevices];

std::once_flag g_hook_once;

}  // namespace

void ensure_shutdown_hook() {
    // Registered after the HIP runtime has initialised (every caller has made
    // a HIP call before), so at exit it runs before the runtime's own
    // teardown and releases streams, events and memory while it is alive.
    std::call_once(g_hook_once, [] { std::atexit([] { (void)gb_shutdown(); }); });
}

gb_status workspace_get(size_t bytes, void* stream, double** out) {
    int dev = 0;
    GB_HIP(hipGetDevice(&dev));
    ensure_shutdown_hook();
    Workspace* wp = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_ws_mtx);
        wp = &(*g_ws)[{dev, stream}];
    }
    Workspace& w = *wp;
    std::lock_guard<std::mutex> lk(w.mtx);
    if (w.bytes < bytes) {
        if (w.ptr) {
            // Work already queued on this stream may still use the old buffer.
            GB_HIP(hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream)));
            GB_HIP(hipFree(w.ptr));
            w.ptr = nullptr;
            w.bytes = 0;
        }
        void* p = nullptr;
        GB_HIP(hipMalloc(&p, bytes));
        w.ptr = static_cast<double*>(p);
        w.bytes = bytes;
    }
    *out = w.ptr;
    return GB_OK;
}

static bool stream_capturing(void* stream) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(reinterpret_cast<hipStream_t>(stream), &cs) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return cs != hipStreamCaptureStatusNone;
}

gb_status counter_acquire(void* stream, CounterLease* out, void* capture_stream) {
    int dev = 0;
    GB_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(GB_ERR_RUNTIME, "device index out of range");
    ensure_shutdown_hook();
    const bool capturing = stream_capturing(capture_stream ? capture_stream : stream);
    CounterPool& P = g_counters[dev];
    std::lock_guard<std::mutex> lk(P.mtx);
    auto signalled = [](hipEvent_t e) {
        const hipError_t q = hipEventQuery(e);
        if (q == hipSuccess) return true;
        if (q != hipErrorNotReady) (void)hipGetLastError();
        return false;
    };
    auto grow = [&]() -> gb_status {
        void* p = nullptr;
        const hipError_t e = hipMalloc(&p, (size_t)kCounterChunk * kCounterBytes);
        if (e != hipSuccess) {
            if (capturing)
                return fail(GB_ERR_RUNTIME,
                            "no node counter is free for a captured call: make one eager call first");
            return hip_fail(e, __FILE__, __LINE__);
        }
        P.chunks.push_back(static_cast<char*>(p));
        P.table.grow(kCounterChunk);
        return GB_OK;
    };
    if (!capturing && P.table.fresh() < kCaptureReserve) {
        gb_status st = grow();
        if (st != GB_OK) return st;
    }
    int i = P.table.take(capturing, signalled);
    if (i < 0) {
        gb_status st = grow();
        if (st != GB_OK) return st;
        i = P.table.take(capturing, signalled);
        if (i < 0) return fail(GB_ERR_RUNTIME, "node counter pool exhausted");
    }
    out->device = dev;
    out->slot = i;
    out->capturing = capturing;
    out->ptr = reinterpret_cast<int*>(P.chunks[(size_t)i / kCounterChunk]
                                      + (size_t)(i % kCounterChunk) * kCounterBytes);
    return GB_OK;
}

void counter_release(CounterLease& lease, void* stream) {
    if (!lease.ptr) return;
    CounterPool& P = g_counters[lease.device];
    std::lock_guard<std::mutex> lk(P.mtx);
    P.table.give_back(lease.slot, lease.capturing, [&](hipEvent_t& ev) {
        if (!ev && hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            ev = nullptr;
            return false;
        }
        if (hipEventRecord(ev, reinterpret_cast<hipStream_t>(stream)) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        return true;
    });
    lease.ptr = nullptr;
}

static void release_counters_and_workspaces() {
    for (int d = 0; d < kMaxDevices; ++d) {
        CounterPool& P = g_counters[d];
        std::lock_guard<std::mutex> lk(P.mtx);
        if (P.chunks.empty()) continue;
        (void)hipSetDevice(d);
        (void)hipDeviceSynchronize();
        for (size_t i = 0; i < P.table.size(); ++i)
            if (P.table.marker(i)) (void)hipEventDestroy(P.table.marker(i));
        for (char* c : P.chunks) (void)hipFree(c);
        P.chunks.clear();
        P.table.clear();
    }
    std::lock_guard<std::mutex> lk(g_ws_mtx);
    for (auto& kv : *g_ws) {
        std::lock_guard<std::mutex> wl(kv.second.mtx);
        if (!kv.second.ptr) continue;
        (void)hipSetDevice(kv.first.first);
        (void)hipDeviceSynchronize();
        (void)hipFree(kv.second.ptr);
        kv.second.ptr = nullptr;
        kv.second.bytes = 0;
    }
    g_ws->clear();
}

}  // namespace gb

extern "C" gb_status gb_shutdown(void) {
    int prev = 0;
    const bool have_prev = hipGetDevice(&prev) == hipSuccess;
    if (!have_prev) {
        (void)hipGetLastError();
        return GB_OK;  // no runtime, nothing was created
    }
    gb::release_buffer_sets();
    gb::release_fork_streams();
    gb::release_counters_and_workspaces();
    (void)hipSetDevice(prev);
    (void)hipGetLastError();
    return GB_OK;
}
