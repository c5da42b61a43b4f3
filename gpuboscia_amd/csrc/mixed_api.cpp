// C-ABI of the mixed-precision batched relaxation: Q plane preparation,
// the two-pass relaxation (BF16-MFMA pass, then the FP64 batched kernel that
// polishes, certifies and re-scores), and the product diagnostic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "gb_fast.h"
#include "gb_host.h"
#include "gb_mixed.h"

struct gb_mixed_q : gb::MixedQ {};

namespace gb {
namespace {

gb_status check_mixed_scope(const gb_mixed_q* q, const gb_objective_desc* o, const gb_lmo_desc* l) {
    if (!q || !q->frag) return fail(GB_ERR_INVALID_ARGUMENT, "null mixed-precision Q");
    if (o && (o->kind != GB_OBJ_DENSE_QUADRATIC || o->n != q->n))
        return fail(GB_ERR_INVALID_ARGUMENT, "the mixed-precision Q does not match the objective");
    if (l && l->kind != GB_LMO_BOX && !fast_simplex_ok(l))
        return fail(GB_ERR_INVALID_ARGUMENT,
                    "the mixed-precision form supports the box LMO, or a simplex with an integral right side");
    return GB_OK;
}

}  // namespace

MixedShape mixed_default_shape(int rt, int N) {
    MixedShape sh;
    // Row split (measured on MI355X at n = 60 and 200): one row tile per
    // wave up to 8 waves, two up to 16 row tiles, four beyond (n > 256).
    sh.rw = rt <= 8 ? 1 : (rt <= kMixedSmallRowTiles ? 2 : 4);
    // Columns: enough tiles to cover the 256 CUs of the chip.
    const int want = (N + 255) / 256;
    sh.cols = want < 1 ? 1 : (want > 16 ? 16 : want);
    sh.blocks = (N + sh.cols - 1) / sh.cols;
    return sh;
}

}  // namespace gb

using namespace gb;

extern "C" gb_status gb_mixed_destroy(gb_mixed_q* q) {
    if (!q) return GB_OK;
    int prev = 0;
    const bool have_prev = hipGetDevice(&prev) == hipSuccess;
    (void)hipSetDevice(q->device);
    (void)hipFree(q->frag);
    (void)hipFree(q->lsc);
    if (have_prev) (void)hipSetDevice(prev);
    delete q;
    return GB_OK;
}

extern "C" gb_status gb_mixed_prepare(const gb_objective_desc* objective, int planes, void* stream,
                                      gb_mixed_q** out) {
    if (!out) return fail(GB_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    if (!objective || objective->kind != GB_OBJ_DENSE_QUADRATIC || !objective->Qt)
        return fail(GB_ERR_INVALID_ARGUMENT, "the mixed-precision form needs a dense quadratic objective");
    if (objective->n <= 0 || objective->n > kMixedMaxN)
        return fail(GB_ERR_INVALID_ARGUMENT, "the mixed-precision form supports 1 <= n <= 512");
    if (planes != 1 && planes != kMixedMaxPlanes)
        return fail(GB_ERR_INVALID_ARGUMENT, "planes must be 1 or 3");
    auto* q = new gb_mixed_q();
    q->n = objective->n;
    q->rt = 2 * ((objective->n + 31) / 32);
    q->planes = planes;
    hipError_t e = hipGetDevice(&q->device);
    const size_t frags = static_cast<size_t>(planes) * q->rt * (q->rt / 2) * 64;
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&q->frag), frags * sizeof(uint4));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&q->lsc), sizeof(float) * 16 * q->rt);
    if (e == hipSuccess)
        e = static_cast<hipError_t>(launch_mixed_prepare(objective->Qt, objective->linear,
                                                         objective->grad_scale, objective->linear_scale,
                                                         q->n, q->rt, planes, q->frag, q->lsc, stream));
    if (e != hipSuccess) {
        gb_mixed_destroy(q);
        return hip_fail(e, __FILE__, __LINE__);
    }
    *out = q;
    return GB_OK;
}

extern "C" gb_status gb_mixed_products(const gb_mixed_q* q, const double* V, int N, double* Y, void* stream) {
    gb_status st = check_mixed_scope(q, nullptr, nullptr);
    if (st != GB_OK) return st;
    if (N < 0 || (N > 0 && (!V || !Y))) return fail(GB_ERR_INVALID_ARGUMENT, "invalid product buffers");
    if (N == 0) return GB_OK;
    const int e = launch_mixed_products(*q, V, N, Y, stream);
    if (e != 0) return hip_fail(static_cast<hipError_t>(e), __FILE__, __LINE__);
    return GB_OK;
}

extern "C" gb_status gb_mixed_warm_products(const gb_objective_desc* objective, int N, const double* x,
                                            const uint8_t* status, const int* iterations, double* out,
                                            void* stream) {
    if (!objective || objective->kind != GB_OBJ_DENSE_QUADRATIC || !objective->Qt)
        return fail(GB_ERR_INVALID_ARGUMENT, "warm products need a dense quadratic objective");
    if (objective->n <= 0 || objective->n > kMixedMaxN)
        return fail(GB_ERR_INVALID_ARGUMENT, "the mixed-precision form supports 1 <= n <= 512");
    if (N < 0 || (N > 0 && (!x || !status || !iterations || !out)))
        return fail(GB_ERR_INVALID_ARGUMENT, "invalid warm-product buffers");
    if (N == 0) return GB_OK;
    const int e = launch_mixed_warm_products(objective->Qt, objective->n, N, x, status, iterations, out, stream);
    if (e != 0) return hip_fail(static_cast<hipError_t>(e), __FILE__, __LINE__);
    return GB_OK;
}

namespace gb {

gb_status relax_mixed_impl(const gb_mixed_q* q, const gb_objective_desc* objective,
                           const gb_algorithm_desc* algorithm, const gb_lmo_desc* lmo, int N, int max_iters,
                           double* x, const double* lb, const double* ub, const double* eps, int* iterations,
                           uint8_t* status, double* gap, double* value, double* best_vertex,
                           double* best_vertex_val, const uint8_t* integer_flags, double candidate_tol,
                           const double* cutoff, const gb_mixed_options* options, void* stream,
                           int* counter, const int* count) {
    {
        gb_status st = validate_descriptors(objective, algorithm, lmo);
        if (st != GB_OK) return st;
        st = check_mixed_scope(q, objective, lmo);
        if (st != GB_OK) return st;
        const bool simplex = lmo->kind == GB_LMO_SIMPLEX;
        if (simplex && algorithm->algorithm != GB_ALG_DICG)
            return fail(GB_ERR_INVALID_ARGUMENT, "the mixed-precision form runs DICG over a simplex");
        if (N < 0 || max_iters < 0) return fail(GB_ERR_INVALID_ARGUMENT, "N and max_iters must be non-negative");
        if (N == 0) return GB_OK;
        if (!x || !lb || !ub || !eps || !iterations || !status || !gap || !value || !best_vertex
                || !best_vertex_val)
            return fail(GB_ERR_INVALID_ARGUMENT, "null node buffer");
        gb_mixed_options op{0, 0, 0, 1e-6f, 0, nullptr, nullptr, nullptr};
        if (options) op = *options;
        if (op.row_tiles_per_wave < 0 || op.row_tiles_per_wave > kMixedMaxRowTilesPerWave || op.columns < 0
                || op.columns > 16 || op.blocks < 0 || !(op.rel_floor >= 0.0f))
            return fail(GB_ERR_INVALID_ARGUMENT, "invalid mixed-precision options");
        // One node counter serves both passes: they run in order on `stream`
        // and each resets it before use.
        ScopedCounter lent(stream);
        if (!counter) {
            st = lent.acquire();
            if (st != GB_OK) return st;
            counter = lent.get();
        }
        if (max_iters > 0) {
            MixedParams P{};
            P.n = q->n;
            P.rt = q->rt;
            P.planes = q->planes;
            P.frag = q->frag;
            P.lsc = q->lsc;
            P.gs = static_cast<float>(objective->grad_scale);
            P.kappa = objective->constant;
            P.N = N;
            P.max_iters = max_iters;
            P.x = x;
            P.lb = lb;
            P.ub = ub;
            P.eps = eps;
            P.iters = iterations;
            P.status = status;
            P.best_v = best_vertex;
            P.best_val = best_vertex_val;
            P.cutoff = cutoff;
            P.counter = counter;
            P.rel_floor = op.rel_floor;
            P.stats = op.mixed_iterations;
            P.count = count;
            P.simplex = simplex ? 1 : 0;
            P.equality = simplex && lmo->equality ? 1 : 0;
            P.rhs = simplex ? lmo->right_side : 0.0;
            P.is_int = simplex ? lmo->is_int : nullptr;
            MixedShape sh = mixed_default_shape(q->rt, N);
            if (op.row_tiles_per_wave) sh.rw = op.row_tiles_per_wave;
            if (op.columns) sh.cols = op.columns;
            // one row tile per wave supports up to 8 waves (n <= 128); more
            // than 16 row tiles (n > 256) run four per wave
            if (sh.rw == 1 && q->rt > 8) sh.rw = 2;
            if (q->rt > kMixedSmallRowTiles) sh.rw = 4;
            if (op.blocks) sh.blocks = op.blocks;
            else sh.blocks = (N + sh.cols - 1) / sh.cols;
            P.cols = sh.cols;
            if (!simplex && mixed_warm_start(sh.rw, q->rt > kMixedSmallRowTiles)) {
                // Nodes with a supplied point start there: Q x in FP64 goes
                // into their best-vertex rows, which pass 1 reads when it takes
                // the node and overwrites when it retires it.
                const int we = launch_mixed_warm_products(objective->Qt, q->n, N, x, status, iterations,
                                                          best_vertex, stream, count);
                if (we != 0) return hip_fail(static_cast<hipError_t>(we), __FILE__, __LINE__);
                P.qx0 = best_vertex;
            }
            const int e = launch_relax_mixed(P, algorithm->algorithm == GB_ALG_DICG, sh.rw, sh.blocks, stream);
            if (e != 0) return hip_fail(static_cast<hipError_t>(e), __FILE__, __LINE__);
        }
        if (op.fp64_max_iters < 0) return GB_OK;
        // The FP64 pass certifies and polishes: by default a sixteenth of the
        // budget (a node still open after it is returned unconverged, with its
        // FP64-certified bound, and resumes as a continuation).
        const int fp64_iters = op.fp64_max_iters > 0 ? op.fp64_max_iters
                                                     : (max_iters / 16 > 1 ? max_iters / 16 : 1);
        const gb_fast_options fo{64, 0, 0, op.fp64_stats, op.gradient};
        return relax_batch_fast(objective, algorithm, lmo, N, fp64_iters, x, lb, ub, eps, iterations,
                                status, gap, value, best_vertex, best_vertex_val, integer_flags,
                                candidate_tol, cutoff, &fo, stream, counter, true, count);
    }
}

}  // namespace gb

extern "C" gb_status gb_relax_batch_mixed(const gb_mixed_q* q, const gb_objective_desc* objective,
                                          const gb_algorithm_desc* algorithm, const gb_lmo_desc* lmo,
                                          int N, int max_iters, double* x, const double* lb,
                                          const double* ub, const double* eps, int* iterations,
                                          uint8_t* status, double* gap, double* value,
                                          double* best_vertex, double* best_vertex_val,
                                          const uint8_t* integer_flags, double candidate_tol,
                                          const double* cutoff, const gb_mixed_options* options,
                                          void* stream) {
    try {
        return relax_mixed_impl(q, objective, algorithm, lmo, N, max_iters, x, lb, ub, eps, iterations, status,
                                gap, value, best_vertex, best_vertex_val, integer_flags, candidate_tol, cutoff,
                                options, stream, nullptr);
    } catch (const std::exception& ex) {
        return fail(GB_ERR_RUNTIME, ex.what());
    }
}

namespace {

// Scratch copy of a batch for the autotuner.
struct TuneScratch {
    double *x = nullptr, *gap = nullptr, *val = nullptr, *bv = nullptr, *bvv = nullptr;
    int* it = nullptr;
    uint8_t* st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~TuneScratch() {
        for (void* p : {(void*)x, (void*)gap, (void*)val, (void*)bv, (void*)bvv, (void*)it, (void*)st})
            if (p) (void)hipFree(p);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

}  // namespace

extern "C" gb_status gb_mixed_autotune(const gb_mixed_q* q, const gb_objective_desc* objective,
                                       const gb_algorithm_desc* algorithm, const gb_lmo_desc* lmo, int N,
                                       int max_iters, const double* x, const double* lb, const double* ub,
                                       const double* eps, const int* iterations, const uint8_t* status,
                                       const uint8_t* integer_flags, double candidate_tol, const double* cutoff,
                                       gb_mixed_options* out, double* seconds, void* stream) {
    try {
        gb_status st = validate_descriptors(objective, algorithm, lmo);
        if (st != GB_OK) return st;
        st = check_mixed_scope(q, objective, lmo);
        if (st != GB_OK) return st;
        if (!out || N <= 0 || max_iters <= 0 || !x || !lb || !ub || !eps || !iterations || !status)
            return fail(GB_ERR_INVALID_ARGUMENT, "invalid autotune arguments");
        const int n = q->n;
        hipStream_t hs = reinterpret_cast<hipStream_t>(stream);
        TuneScratch S;
        const size_t xb = (size_t)N * n * sizeof(double);
        GB_HIP(hipMalloc(reinterpret_cast<void**>(&S.x), xb));
        GB_HIP(hipMalloc(reinterpret_cast<void**>(&S.bv), xb));
        GB_HIP(hipMalloc(reinterpret_cast<void**>(&S.gap), (size_t)N * 8));
        GB_HIP(hipMalloc(reinterpret_cast<void**>(&S.val), (size_t)N * 8));
        GB_HIP(hipMalloc(reinterpret_cast<void**>(&S.bvv), (size_t)N * 8));
        GB_HIP(hipMalloc(reinterpret_cast<void**>(&S.it), (size_t)N * 4));
        GB_HIP(hipMalloc(reinterpret_cast<void**>(&S.st), (size_t)N));
        GB_HIP(hipEventCreate(&S.e0));
        GB_HIP(hipEventCreate(&S.e1));
        const std::vector<double> none(N, 1e300);
        std::vector<uint8_t> hst(N);
        auto restore = [&]() -> gb_status {
            GB_HIP(hipMemcpyAsync(S.x, x, xb, hipMemcpyDeviceToDevice, hs));
            GB_HIP(hipMemcpyAsync(S.it, iterations, (size_t)N * 4, hipMemcpyDeviceToDevice, hs));
            GB_HIP(hipMemcpyAsync(S.st, status, (size_t)N, hipMemcpyDeviceToDevice, hs));
            GB_HIP(hipMemcpyAsync(S.bvv, none.data(), (size_t)N * 8, hipMemcpyHostToDevice, hs));
            GB_HIP(hipStreamSynchronize(hs));
            return GB_OK;
        };
        // Runs one relaxation of the scratch batch; returns its time.
        auto run = [&](const gb_mixed_options& op, double* sec) -> gb_status {
            gb_status r = restore();
            if (r != GB_OK) return r;
            GB_HIP(hipEventRecord(S.e0, hs));
            r = relax_mixed_impl(q, objective, algorithm, lmo, N, max_iters, S.x, lb, ub, eps, S.it, S.st, S.gap,
                                 S.val, S.bv, S.bvv, integer_flags, candidate_tol, cutoff, &op, stream, nullptr);
            if (r != GB_OK) return r;
            GB_HIP(hipEventRecord(S.e1, hs));
            GB_HIP(hipEventSynchronize(S.e1));
            float ms = 0.f;
            GB_HIP(hipEventElapsedTime(&ms, S.e0, S.e1));
            *sec = ms * 1e-3;
            return GB_OK;
        };
        gb_mixed_options op = *out;
        op.blocks = 0;
        op.mixed_iterations = nullptr;
        op.fp64_stats = nullptr;
        op.gradient = nullptr;
        // Stage 1: tile shape of the mixed pass.
        int best_rw = 0, best_cols = 0;
        double best_t = 1e300;
        for (int rw = 1; rw <= kMixedMaxRowTilesPerWave; ++rw) {
            if (rw == 1 && q->rt > 8) continue;
            if (rw != 4 && q->rt > kMixedSmallRowTiles) continue;
            for (int cols : {2, 4, 8, 16}) {
                op.row_tiles_per_wave = rw;
                op.columns = cols;
                op.fp64_max_iters = -1;
                double t0 = 0.0, t1 = 0.0;
                st = run(op, &t0);  // first launch of a variant also loads its code
                if (st != GB_OK) return st;
                st = run(op, &t1);
                if (st != GB_OK) return st;
                if (t1 < best_t) {
                    best_t = t1;
                    best_rw = rw;
                    best_cols = cols;
                }
            }
        }
        // Stage 2: FP64-pass budget, by time per certified-converged node.
        op.row_tiles_per_wave = best_rw;
        op.columns = best_cols;
        int best_budget = 0;
        double best_cost = 1e300, best_sec = 0.0;
        for (int div : {32, 16, 8}) {
            op.fp64_max_iters = max_iters / div > 1 ? max_iters / div : 1;
            double t = 0.0;
            st = run(op, &t);
            if (st != GB_OK) return st;
            GB_HIP(hipMemcpy(hst.data(), S.st, (size_t)N, hipMemcpyDeviceToHost));
            int conv = 0;
            for (uint8_t v : hst)
                if (v == GB_NODE_FRACTIONAL_CONVERGED || v == GB_NODE_INTEGRAL_CONVERGED
                        || v == GB_NODE_INFEASIBLE)
                    ++conv;
            const double cost = t / (conv > 0 ? conv : 1);
            if (cost < best_cost) {
                best_cost = cost;
                best_budget = op.fp64_max_iters;
                best_sec = t;
            }
        }
        out->row_tiles_per_wave = best_rw;
        out->columns = best_cols;
        out->fp64_max_iters = best_budget;
        if (seconds) *seconds = best_sec;
        return GB_OK;
    } catch (const std::exception& ex) {
        return fail(GB_ERR_RUNTIME, ex.what());
    }
}

namespace gb {
namespace {

// Internal fork streams of gb_relax_instances, per device, default priority.
struct ForkStreams {
    std::mutex mtx;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> done;
    hipEvent_t start = nullptr;
};
ForkStreams g_fork[64];

// Form of a gb_relax_instances entry: 2 mixed, 1 batched FP64 (dense over a
// box or a simplex with an integral right side, n <= 512), 0 plain FP64
// (everything else).
int instance_form(const gb_instance_batch& b) {
    if (b.mixed) return 2;
    if (b.objective->kind == GB_OBJ_DENSE_QUADRATIC && b.objective->n <= kMaxFastN
            && (b.lmo->kind == GB_LMO_BOX || fast_simplex_ok(b.lmo)))
        return 1;
    return 0;
}

}  // namespace

void release_fork_streams() {
    for (int d = 0; d < 64; ++d) {
        ForkStreams& F = g_fork[d];
        std::lock_guard<std::mutex> lk(F.mtx);
        if (F.streams.empty() && !F.start) continue;
        (void)hipSetDevice(d);
        for (hipStream_t s : F.streams) (void)hipStreamSynchronize(s);
        for (hipStream_t s : F.streams) (void)hipStreamDestroy(s);
        for (hipEvent_t e : F.done) (void)hipEventDestroy(e);
        if (F.start) (void)hipEventDestroy(F.start);
        F.streams.clear();
        F.done.clear();
        F.start = nullptr;
    }
}

}  // namespace gb

extern "C" gb_status gb_relax_instances(int count, const gb_instance_batch* batches,
                                        const gb_algorithm_desc* algorithm, int max_iters, double candidate_tol,
                                        int concurrency, void* stream) {
    try {
        if (count < 0 || (count > 0 && !batches) || !algorithm || concurrency < 1 || concurrency > 8 || max_iters < 0)
            return fail(GB_ERR_INVALID_ARGUMENT, "invalid instance batch arguments");
        for (int i = 0; i < count; ++i) {
            const gb_instance_batch& b = batches[i];
            if (!b.objective || !b.lmo) return fail(GB_ERR_INVALID_ARGUMENT, "instance batch without descriptors");
            gb_status st = validate_descriptors(b.objective, algorithm, b.lmo);
            if (st != GB_OK) return st;
            if (b.mixed) {
                st = check_mixed_scope(b.mixed, b.objective, b.lmo);
                if (st != GB_OK) return st;
            }
        }
        if (count == 0) return GB_OK;
        int dev = 0;
        GB_HIP(hipGetDevice(&dev));
        if (dev < 0 || dev >= 64) return fail(GB_ERR_RUNTIME, "device index out of range");
        ensure_shutdown_hook();
        ForkStreams& F = g_fork[dev];
        std::lock_guard<std::mutex> lk(F.mtx);
        while ((int)F.streams.size() < concurrency) {
            hipStream_t s = nullptr;
            hipEvent_t e = nullptr;
            GB_HIP(create_dedicated_stream(&s));  // batches run concurrently
            GB_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            F.streams.push_back(s);
            F.done.push_back(e);
        }
        if (!F.start) GB_HIP(hipEventCreateWithFlags(&F.start, hipEventDisableTiming));
        hipStream_t caller = reinterpret_cast<hipStream_t>(stream);
        GB_HIP(hipEventRecord(F.start, caller));
        const int k = concurrency < count ? concurrency : count;
        // One node counter per fork stream: the entries of one fork stream
        // run in order and each launch resets the counter before use.
        std::vector<CounterLease> lease(k);
        struct GiveBack {
            std::vector<CounterLease>& l;
            ForkStreams& f;
            ~GiveBack() {
                for (size_t j = 0; j < l.size(); ++j) counter_release(l[j], f.streams[j]);
            }
        } give_back{lease, F};
        for (int j = 0; j < k; ++j) {
            gb_status st = counter_acquire(F.streams[j], &lease[j], caller);
            if (st != GB_OK) return st;
        }
        int cus = 0;
        GB_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        // Entries sharing the chip widen their mixed tiles (below).  A graph
        // replay overlaps the fork streams' kernels less than eager streams
        // do (measured: 666 ms per replay of the 256-instance batch with
        // widened tiles, 511 ms at the one-stream width, 382 ms eager and
        // widened), so a call under capture keeps the one-stream width.
        hipStreamCaptureStatus capturing = hipStreamCaptureStatusNone;
        GB_HIP(hipStreamIsCapturing(caller, &capturing));
        const int sharing = capturing == hipStreamCaptureStatusNone ? k : 1;
        // Entries of the plain FP64 form use their fork stream's work
        // vectors: size them for the largest such entry first, so that no
        // entry grows them (and waits on its stream) mid-call.
        {
            std::vector<size_t> need(k, 0);
            for (int i = 0; i < count; ++i) {
                const gb_instance_batch& b = batches[i];
                if (instance_form(b) == 0)
                    need[i % k] = std::max(need[i % k], (size_t)b.N * kWorkVectors * b.objective->n * sizeof(double));
            }
            for (int j = 0; j < k; ++j)
                if (need[j]) {
                    double* w = nullptr;
                    gb_status st = workspace_get(need[j], F.streams[j], &w);
                    if (st != GB_OK) return st;
                }
        }
        for (int j = 0; j < k; ++j) GB_HIP(hipStreamWaitEvent(F.streams[j], F.start, 0));
        for (int i = 0; i < count; ++i) {
            const gb_instance_batch& b = batches[i];
            void* s = F.streams[i % k];
            int* counter = lease[i % k].ptr;
            gb_status st;
            if (b.mixed) {
                // `sharing` entries share the CUs: widen the tiles so that
                // together they cover the chip once, up to the MFMA's 16 node
                // columns.  A tile reads its Q fragments once per iteration
                // whatever its width, so wider tiles cut the fragment traffic
                // per node (at 1024 nodes and 8 streams: 16 columns instead
                // of the solo default's 4).  Each entry equals its solo call
                // with gb_mixed_options.columns =
                // clamp(ceil(N sharing / CUs), default, 16).
                gb_mixed_options mo{0, 0, 0, 1e-6f, 0, nullptr, nullptr, nullptr};
                const int def = mixed_default_shape(b.mixed->rt, b.N).cols;
                const int want = cus > 0 ? (int)(((long long)b.N * sharing + cus - 1) / cus) : def;
                mo.columns = want < def ? def : (want > 16 ? 16 : want);
                st = relax_mixed_impl(b.mixed, b.objective, algorithm, b.lmo, b.N, max_iters, b.x, b.lb, b.ub, b.eps,
                                      b.iterations, b.status, b.gap, b.value, b.best_vertex, b.best_vertex_val,
                                      b.integer_flags, candidate_tol, b.cutoff, &mo, s, counter);
            } else if (instance_form(b) == 1) {
                st = relax_batch_fast(b.objective, algorithm, b.lmo, b.N, max_iters, b.x, b.lb, b.ub, b.eps,
                                      b.iterations, b.status, b.gap, b.value, b.best_vertex, b.best_vertex_val,
                                      b.integer_flags, candidate_tol, b.cutoff, nullptr, s, counter);
            } else {
                st = relax_batch_fp64(b.objective, algorithm, b.lmo, b.N, max_iters, b.x, b.lb, b.ub, b.eps,
                                      b.iterations, b.status, b.gap, b.value, b.best_vertex, b.best_vertex_val,
                                      b.integer_flags, candidate_tol, 0, s);
            }
            if (st != GB_OK) return st;
        }
        for (int j = 0; j < k; ++j) {
            GB_HIP(hipEventRecord(F.done[j], F.streams[j]));
            GB_HIP(hipStreamWaitEvent(caller, F.done[j], 0));
        }
        return GB_OK;
    } catch (const std::exception& ex) {
        return fail(GB_ERR_RUNTIME, ex.what());
    }
}

extern "C" gb_status gb_candidate_features(int n, int N, const double* X, const double* G, const double* LB,
                                           const double* UB, const int* depth, const double* gap,
                                           const double* value, const double* pc_up, const double* pc_down,
                                           const double* Qt, double gs, const uint8_t* flags, double tol, float* F,
                                           void* stream) {
    if (n <= 0 || N < 0) return fail(GB_ERR_INVALID_ARGUMENT, "n must be positive and N non-negative");
    if (N == 0) return GB_OK;
    if (!X || !G || !LB || !UB || !depth || !gap || !value || !pc_up || !pc_down || !Qt || !flags || !F)
        return fail(GB_ERR_INVALID_ARGUMENT, "null feature buffer");
    const int e = launch_candidate_features(n, N, X, G, LB, UB, depth, gap, value, pc_up, pc_down, Qt,
                                            (size_t)n + 1, gs, flags, tol, F, stream);
    if (e != 0) return hip_fail(static_cast<hipError_t>(e), __FILE__, __LINE__);
    return GB_OK;
}
