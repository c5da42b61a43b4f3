// C-ABI of the mixed-precision batched relaxation: Q plane preparation,
// the two-pass relaxation (BF16-MFMA pass, then the FP64 batched kernel that
// polishes, certifies and re-scores), and the product diagnostic.
#include <hip/hip_runtime.h>

#include <string>

#include "gb_host.h"
#include "gb_mixed.h"

struct gb_mixed_q : gb::MixedQ {};

namespace gb {
namespace {

gb_status check_mixed_scope(const gb_mixed_q* q, const gb_objective_desc* o, const gb_lmo_desc* l) {
    if (!q || !q->frag) return fail(GB_ERR_INVALID_ARGUMENT, "null mixed-precision Q");
    if (o && (o->kind != GB_OBJ_DENSE_QUADRATIC || o->n != q->n))
        return fail(GB_ERR_INVALID_ARGUMENT, "the mixed-precision Q does not match the objective");
    if (l && l->kind != GB_LMO_BOX)
        return fail(GB_ERR_INVALID_ARGUMENT, "the mixed-precision form supports the box LMO");
    return GB_OK;
}

}  // namespace

MixedShape mixed_default_shape(int rt, int N) {
    MixedShape sh;
    // Row split: about 4 waves per tile.
    sh.rw = rt <= 4 ? 1 : (rt <= 8 ? 2 : 4);
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
        return fail(GB_ERR_INVALID_ARGUMENT, "the mixed-precision form supports 1 <= n <= 256");
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
        gb_status st = validate_descriptors(objective, algorithm, lmo);
        if (st != GB_OK) return st;
        st = check_mixed_scope(q, objective, lmo);
        if (st != GB_OK) return st;
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
        if (max_iters > 0) {
            int* counter = nullptr;
            st = counter_get(&counter);
            if (st != GB_OK) return st;
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
            MixedShape sh = mixed_default_shape(q->rt, N);
            if (op.row_tiles_per_wave) sh.rw = op.row_tiles_per_wave;
            if (op.columns) sh.cols = op.columns;
            // one row tile per wave supports up to 8 waves (n <= 128)
            if (sh.rw == 1 && q->rt > 8) sh.rw = 2;
            if (op.blocks) sh.blocks = op.blocks;
            else sh.blocks = (N + sh.cols - 1) / sh.cols;
            P.cols = sh.cols;
            const int e = launch_relax_mixed(P, algorithm->algorithm == GB_ALG_DICG, sh.rw, sh.blocks, stream);
            if (e != 0) return hip_fail(static_cast<hipError_t>(e), __FILE__, __LINE__);
        }
        if (op.fp64_max_iters < 0) return GB_OK;
        // The FP64 pass certifies and polishes: by default a sixteenth of the
        // budget (a node still open after it is returned unconverged, with its
        // FP64-certified bound, and resumes as a continuation).
        const int fp64_iters = op.fp64_max_iters > 0 ? op.fp64_max_iters
                                                     : (max_iters / 16 > 1 ? max_iters / 16 : 1);
        const gb_fast_options fo{64, 4, 0, op.fp64_stats, op.gradient};
        return relax_batch_fast(objective, algorithm, lmo, N, fp64_iters, x, lb, ub, eps, iterations,
                                status, gap, value, best_vertex, best_vertex_val, integer_flags,
                                candidate_tol, cutoff, &fo, stream, nullptr, true);
    } catch (const std::exception& ex) {
        return fail(GB_ERR_RUNTIME, ex.what());
    }
}
