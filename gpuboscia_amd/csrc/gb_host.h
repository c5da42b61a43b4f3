// Host-side internal helpers of the library (error state, limits, shared
// launch helpers).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "gb_internal.h"

namespace gb {

constexpr int kMaxDenseN = 8192;         // vector staged in LDS (64 KiB FP64)
constexpr int kMaxSimplexN = 2048;       // sort/scan arrays in LDS
constexpr int kMaxBirkhoffOrder = 256;   // Hungarian state in LDS
constexpr size_t kMaxLdsBytes = 160 * 1024;
constexpr int kMaxFastN = 512;           // batched form: 8 coordinates per lane
constexpr int kFastRefresh = 64;         // direct Q x recompute interval of the batched form

extern thread_local std::string g_last_error;

gb_status fail(gb_status code, const std::string& msg);
gb_status hip_fail(hipError_t e, const char* file, int line);
gb_status workspace_get(size_t bytes, void* stream, double** out);
// staged: the relaxation stages a dense objective's vector in LDS (n <=
// kMaxDenseN); MultiBlock reads it from global memory (any n).
gb_status validate_descriptors(const gb_objective_desc* o, const gb_algorithm_desc* a,
                               const gb_lmo_desc* l, bool staged = true,
                               const gb_plugin_desc* plugin = nullptr);
// gb_relax_batch_plugin: the caller's launcher runs its instantiation of the
// FP64 kernel on a batch validated and laid out here.
gb_status relax_batch_plugin(const gb_objective_desc* o, const gb_algorithm_desc* a, const gb_lmo_desc* l,
                             const gb_plugin_desc* plugin, int N, int max_iters, double* x, const double* lb,
                             const double* ub, const double* eps, int* iterations, uint8_t* status, double* gap,
                             double* value, double* best_vertex, double* best_vertex_val, const uint8_t* flags,
                             double ctol, int block_threads, void* stream);
int auto_block_threads(int n);
gb_status relax_batch_fp64(const gb_objective_desc* o, const gb_algorithm_desc* a,
                           const gb_lmo_desc* l, int N, int max_iters, double* x, const double* lb,
                           const double* ub, const double* eps, int* iterations, uint8_t* status,
                           double* gap, double* value, double* best_vertex, double* best_vertex_val,
                           const uint8_t* flags, double ctol, int block_threads, void* stream);

// The batched form takes a simplex LMO with a finite integral right side
// (its nodes must be 0/1; others come back with status 0xFF).
bool fast_simplex_ok(const gb_lmo_desc* l);
gb_status relax_batch_fast(const gb_objective_desc* o, const gb_algorithm_desc* a,
                           const gb_lmo_desc* l, int N, int max_iters, double* x, const double* lb,
                           const double* ub, const double* eps, int* iterations, uint8_t* status,
                           double* gap, double* value, double* best_vertex, double* best_vertex_val,
                           const uint8_t* flags, double ctol, const double* cutoff,
                           const gb_fast_options* opt, void* stream, int* counter = nullptr,
                           bool rescore = false, const int* count = nullptr);
// The mixed-precision form (mixed_api.cpp); count as relax_batch_fast's.
gb_status relax_mixed_impl(const gb_mixed_q* q, const gb_objective_desc* objective,
                           const gb_algorithm_desc* algorithm, const gb_lmo_desc* lmo, int N, int max_iters,
                           double* x, const double* lb, const double* ub, const double* eps, int* iterations,
                           uint8_t* status, double* gap, double* value, double* best_vertex,
                           double* best_vertex_val, const uint8_t* integer_flags, double candidate_tol,
                           const double* cutoff, const gb_mixed_options* options, void* stream,
                           int* counter, const int* count = nullptr);
// A node counter of the persistent kernels lent to one call (resources.cpp):
// no launch that may run at the same time uses it, and a counter lent while
// the stream is being captured stays with the captured graph.
struct CounterLease {
    int device = -1;
    int slot = -1;
    bool capturing = false;
    int* ptr = nullptr;
};
// capture_stream: the stream whose capture status decides (default: stream
// itself); a fork stream that will join the caller's capture passes the
// caller's stream, so its counter stays with the graph.
gb_status counter_acquire(void* stream, CounterLease* out, void* capture_stream = nullptr);
// Gives the counter back once the call has queued every launch that uses it
// on `stream` (a completion event is recorded there).
void counter_release(CounterLease& lease, void* stream);
// Lends a counter for the scope of one call on one stream.
class ScopedCounter {
  public:
    explicit ScopedCounter(void* stream) : stream_(stream) {}
    ScopedCounter(const ScopedCounter&) = delete;
    ScopedCounter& operator=(const ScopedCounter&) = delete;
    ~ScopedCounter() { counter_release(lease_, stream_); }
    gb_status acquire() { return counter_acquire(stream_, &lease_); }
    int* get() const { return lease_.ptr; }

  private:
    void* stream_;
    CounterLease lease_;
};
// MultiBlock relaxation (capi.cpp): the plain form with blocks_per_node
// co-resident workgroups per node; dense quadratic over a box.
gb_status relax_batch_multiblock(const gb_objective_desc* o, const gb_algorithm_desc* a, const gb_lmo_desc* l,
                                 int N, int max_iters, double* x, const double* lb, const double* ub,
                                 const double* eps, int* iterations, uint8_t* status, double* gap, double* value,
                                 double* best_vertex, double* best_vertex_val, const uint8_t* flags, double ctol,
                                 int block_threads, int blocks_per_node, void* stream);
// Node FW tolerance at a depth (the host driver's, from DeviceResidentConfig).
double depth_tolerance(double root, double decay, double minimum, int depth);
// The device-resident search of gb_solve (dsearch.cpp), settings validated.
gb_status device_search(const gb_problem* pr, const gb_objective_desc* o, const gb_lmo_desc* l,
                        const gb_settings* s, gb_result* res);
// Teardown hooks of gb_shutdown: the idle batch buffer sets of the
// branch-and-bound driver (bnb.cpp) and the fork streams of
// gb_relax_instances (mixed_api.cpp).
void release_buffer_sets();
void release_fork_streams();
// Registers gb_shutdown to run at process exit (once; after HIP is up).
void ensure_shutdown_hook();
// A stream on a hardware queue of its own (all CUs enabled), for work that
// must run concurrently with the library's other streams: plain streams
// share a small pool of hardware queues with every stream in the process
// (e.g. a framework's stream pool), and two streams on one queue run in
// order.  Falls back to a plain stream if the runtime refuses.
hipError_t create_dedicated_stream(hipStream_t* out);

#define GB_HIP(call)                                                   \
    do {                                                               \
        hipError_t _e = (call);                                        \
        if (_e != hipSuccess) return ::gb::hip_fail(_e, __FILE__, __LINE__); \
    } while (0)

}  // namespace gb
