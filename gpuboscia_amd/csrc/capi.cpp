// C-ABI entry points: descriptor validation and the FP64 batch relaxation
// launches (gb_relax_batch, its allocation-free variant, the batched form).
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "gb_barrier.h"
#include "gb_fast.h"
#include "gb_internal.h"
#include "gb_host.h"

namespace gb {

thread_local std::string g_last_error;

gb_status fail(gb_status code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

gb_status hip_fail(hipError_t e, const char* file, int line) {
    return fail(GB_ERR_DEVICE, std::string("HIP ") + file + ":" + std::to_string(line) + " - "
                                   + hipGetErrorString(e));
}

// LMO descriptor checks for n variables (obj_kind: the objective it is
// paired with; a: the algorithm, may be NULL; staged: the one-workgroup
// kernels hold the LMO's vectors in LDS).
static gb_status validate_lmo(const gb_lmo_desc* l, int n, int obj_kind, const gb_algorithm_desc* a,
                              bool staged = true) {
    if (l->kind == GB_LMO_SIMPLEX) {
        if (!l->is_int) return fail(GB_ERR_INVALID_ARGUMENT, "simplex LMO requires is_int");
        if (staged && n > kMaxSimplexN)
            return fail(GB_ERR_INVALID_ARGUMENT, "simplex LMO n exceeds the LDS limit (MultiBlock takes any n)");
    } else if (l->kind == GB_LMO_BIRKHOFF) {
        if (l->order <= 0 || l->order > kMaxBirkhoffOrder)
            return fail(GB_ERR_INVALID_ARGUMENT, "Birkhoff order must be in [1, 256]");
    } else if (l->kind == GB_LMO_PORTFOLIO) {
        if (obj_kind != GB_OBJ_DENSE_QUADRATIC || n % 2 != 0 || n > kMaxSimplexN)
            return fail(GB_ERR_INVALID_ARGUMENT,
                        "portfolio LMO needs a dense objective over (x, z) with n = 2 * assets <= 2048");
        if (!(l->cap > 0.0) || !(l->cardinality >= 0.0) || !(l->right_side > 0.0))
            return fail(GB_ERR_INVALID_ARGUMENT, "portfolio LMO needs cap > 0, cardinality >= 0, right_side > 0");
        if (a && a->algorithm == GB_ALG_DICG)
            return fail(GB_ERR_INVALID_ARGUMENT,
                        "the portfolio LMO has no in-face oracle: use the Frank-Wolfe algorithm");
    } else if (l->kind == GB_LMO_PLUGIN) {
        // checked against the plug-in in validate_descriptors
    } else if (l->kind != GB_LMO_BOX) {
        return fail(GB_ERR_INVALID_ARGUMENT, "unknown LMO kind");
    }
    return GB_OK;
}

gb_status validate_descriptors(const gb_objective_desc* o, const gb_algorithm_desc* a,
                               const gb_lmo_desc* l, bool staged, const gb_plugin_desc* plugin) {
    if (!o || !a || !l) return fail(GB_ERR_INVALID_ARGUMENT, "null descriptor");
    if (o->n <= 0) return fail(GB_ERR_INVALID_ARGUMENT, "objective n must be positive");
    // Device plug-ins (the reference's compile-time checks, gpuboscia.cuh:135-149)
    const bool pobj = o->kind == GB_OBJ_PLUGIN, plmo = l->kind == GB_LMO_PLUGIN,
               pheur = a->heuristic == GB_HEUR_PLUGIN;
    if ((pobj || plmo || pheur) && !plugin)
        return fail(GB_ERR_INVALID_ARGUMENT, "a plug-in objective, LMO or heuristic needs a gb_plugin_desc");
    if (plugin) {
        if (!plugin->launch) return fail(GB_ERR_INVALID_ARGUMENT, "plug-in without a launcher");
        if (pobj != (plugin->objective != 0) || plmo != (plugin->lmo != 0) || pheur != (plugin->heuristic != 0))
            return fail(GB_ERR_INVALID_ARGUMENT,
                        "plug-in flags must match the GB_OBJ_PLUGIN / GB_LMO_PLUGIN / GB_HEUR_PLUGIN kinds");
        if (a->algorithm == GB_ALG_DICG && plmo && !plugin->supports_dicg)
            return fail(GB_ERR_INVALID_ARGUMENT,
                        "DICG requires compute_inface_v, dicg_maximum_step_local and dicg_sum_maximum_step");
        if (a->line_search == GB_LS_QUADRATIC_EXACT && pobj && !plugin->supports_quadratic_exact)
            return fail(GB_ERR_INVALID_ARGUMENT, "QuadraticExactLS requires a quadratic objective");
        if (plugin->block_threads < 0) return fail(GB_ERR_INVALID_ARGUMENT, "plug-in block_threads must be >= 0");
    }
    if (pobj) {
        // n only; the device computations are the plug-in's
    } else if (o->kind == GB_OBJ_DENSE_QUADRATIC) {
        if (!o->Qt) return fail(GB_ERR_INVALID_ARGUMENT, "dense objective requires Qt");
        if (staged && o->n > kMaxDenseN)
            return fail(GB_ERR_INVALID_ARGUMENT,
                        "dense objective n exceeds the LDS-staged limit (MultiBlock takes any n)");
    } else if (o->kind == GB_OBJ_QAP) {
        if (o->order <= 0 || o->n != o->order * o->order || !o->A || !o->B)
            return fail(GB_ERR_INVALID_ARGUMENT, "QAP objective requires order, A, B and n == order^2");
    } else {
        return fail(GB_ERR_INVALID_ARGUMENT, "unknown objective kind");
    }
    gb_status st = validate_lmo(l, o->n, o->kind, a, staged);
    if (st != GB_OK) return st;
    if (a->algorithm != GB_ALG_FRANK_WOLFE && a->algorithm != GB_ALG_DICG)
        return fail(GB_ERR_INVALID_ARGUMENT, "unknown FW algorithm");
    if (a->line_search < GB_LS_OPEN_LOOP || a->line_search > GB_LS_QUADRATIC_EXACT)
        return fail(GB_ERR_INVALID_ARGUMENT, "unknown line search");
    // gpuboscia.cuh:146-149
    if (a->algorithm == GB_ALG_FRANK_WOLFE
            && (a->line_search == GB_LS_SCALED_OPEN_LOOP
                || a->line_search == GB_LS_SAFEGUARDED_SCALED_OPEN_LOOP))
        return fail(GB_ERR_INVALID_ARGUMENT, "scaled open-loop is available only for DICG");
    if (a->line_search == GB_LS_SECANT && a->secant_iters < 0)
        return fail(GB_ERR_INVALID_ARGUMENT, "secant iterations must be non-negative");
    if (a->heuristic != GB_HEUR_ROUND && a->heuristic != GB_HEUR_FLOOR_OFFSET && a->heuristic != GB_HEUR_PLUGIN)
        return fail(GB_ERR_INVALID_ARGUMENT, "unknown heuristic");
    return GB_OK;
}

int auto_block_threads(int n) {
    const int wanted = ((n + 63) / 64) * 64;
    return wanted < 64 ? 64 : (wanted > 256 ? 256 : wanted);
}

// Validation and launch parameters of the plain FP64 form (everything but
// the work vectors, which the caller provides).
struct Fp64Launch {
    RelaxParams P{};
    LdsLayout L{};
    int threads = 0;
    size_t work_bytes = 0;
};

static gb_status relax_fp64_setup(const gb_objective_desc* o, const gb_algorithm_desc* a,
                                  const gb_lmo_desc* l, int N, int max_iters, double* x, const double* lb,
                                  const double* ub, const double* eps, int* iterations, uint8_t* status,
                                  double* gap, double* value, double* best_vertex, double* best_vertex_val,
                                  const uint8_t* flags, double ctol, int block_threads, Fp64Launch* out,
                                  bool lds = true, const gb_plugin_desc* plugin = nullptr) {
    gb_status st = validate_descriptors(o, a, l, lds, plugin);
    if (st != GB_OK) return st;
    if (N < 0 || max_iters < 0) return fail(GB_ERR_INVALID_ARGUMENT, "N and max_iters must be non-negative");
    if (N == 0) return GB_OK;
    if (!x || !lb || !ub || !eps || !iterations || !status || !gap || !value || !best_vertex
            || !best_vertex_val)
        return fail(GB_ERR_INVALID_ARGUMENT, "null node buffer");
    if (block_threads == 0) {
        block_threads = auto_block_threads(o->n);
        // a plug-in objective's block_threads bound (device_resident_solver.cuh:1366-1393)
        if (plugin && plugin->objective && plugin->block_threads > 0)
            block_threads = std::min(block_threads, std::max(64, plugin->block_threads / 64 * 64));
    }
    if (block_threads < 64 || block_threads > 1024 || block_threads % 64 != 0)
        return fail(GB_ERR_INVALID_ARGUMENT,
                    "gpu_block_threads must be zero or a positive multiple of 64 up to 1024");
    const int n = o->n;
    out->L = lds_layout(o->kind, l->kind, n, l->order);
    if (lds && out->L.total > kMaxLdsBytes) return fail(GB_ERR_INVALID_ARGUMENT, "problem exceeds the LDS budget");
    out->threads = block_threads;
    out->work_bytes = (size_t)N * kWorkVectors * n * sizeof(double);
    RelaxParams& P = out->P;
    if (plugin && plugin->objective && plugin->objective_bytes > 0) {
        // the plug-in objective's workspace follows the work vectors
        P.obj_stride = (plugin->objective_bytes + 255) & ~size_t(255);
        out->work_bytes = ((out->work_bytes + 255) & ~size_t(255)) + (size_t)N * P.obj_stride;
    }
    P.obj_kind = o->kind;
    P.n = n;
    P.Qt = o->Qt;
    P.lin = o->linear;
    P.gs = o->grad_scale;
    P.ls = o->linear_scale;
    P.kappa = o->constant;
    P.order = o->order;
    P.A = o->A;
    P.B = o->B;
    P.lmo_kind = l->kind;
    P.rhs = l->right_side;
    P.equality = l->equality;
    P.is_int = l->is_int;
    P.lorder = l->order;
    P.pcap = l->cap;
    P.pcard = l->cardinality;
    P.dicg = a->algorithm == GB_ALG_DICG;
    P.ls_kind = a->line_search;
    P.secant_iters = a->secant_iters;
    P.secant_tol = a->secant_tol;
    P.secant_denom = a->secant_denom_eps;
    P.heuristic = a->heuristic;
    P.heuristic_offset = a->heuristic_offset;
    P.N = N;
    P.max_iters = max_iters;
    P.x = x;
    P.lb = lb;
    P.ub = ub;
    P.eps = eps;
    P.iters = iterations;
    P.status = status;
    P.gap = gap;
    P.value = value;
    P.best_v = best_vertex;
    P.best_val = best_vertex_val;
    P.flags = flags;
    P.ctol = ctol;
    return GB_OK;
}

gb_status relax_batch_fp64(const gb_objective_desc* o, const gb_algorithm_desc* a,
                           const gb_lmo_desc* l, int N, int max_iters, double* x, const double* lb,
                           const double* ub, const double* eps, int* iterations, uint8_t* status,
                           double* gap, double* value, double* best_vertex, double* best_vertex_val,
                           const uint8_t* flags, double ctol, int block_threads, void* stream) {
    Fp64Launch F;
    gb_status st = relax_fp64_setup(o, a, l, N, max_iters, x, lb, ub, eps, iterations, status, gap, value,
                                    best_vertex, best_vertex_val, flags, ctol, block_threads, &F);
    if (st != GB_OK || N == 0) return st;
    st = workspace_get(F.work_bytes, stream, &F.P.work);
    if (st != GB_OK) return st;
    const int e = launch_relax_fp64(F.P, F.threads, F.L.total, F.L, stream);
    if (e != 0) return hip_fail(static_cast<hipError_t>(e), __FILE__, __LINE__);
    return GB_OK;
}

// Device plug-ins: the batch is validated and laid out as for the built-in
// kernel, and the caller's launcher runs its own instantiation of it.
gb_status relax_batch_plugin(const gb_objective_desc* o, const gb_algorithm_desc* a, const gb_lmo_desc* l,
                             const gb_plugin_desc* plugin, int N, int max_iters, double* x, const double* lb,
                             const double* ub, const double* eps, int* iterations, uint8_t* status, double* gap,
                             double* value, double* best_vertex, double* best_vertex_val, const uint8_t* flags,
                             double ctol, int block_threads, void* stream) {
    if (!plugin) return fail(GB_ERR_INVALID_ARGUMENT, "null plug-in descriptor");
    Fp64Launch F;
    gb_status st = relax_fp64_setup(o, a, l, N, max_iters, x, lb, ub, eps, iterations, status, gap, value,
                                    best_vertex, best_vertex_val, flags, ctol, block_threads, &F, true, plugin);
    if (st != GB_OK || N == 0) return st;
    st = workspace_get(F.work_bytes, stream, &F.P.work);
    if (st != GB_OK) return st;
    if (F.P.obj_stride) {
        const size_t wb = ((size_t)N * kWorkVectors * o->n * sizeof(double) + 255) & ~size_t(255);
        F.P.obj_work = reinterpret_cast<unsigned char*>(F.P.work) + wb;
    }
    PluginLaunch pl{F.P, F.L, F.threads, F.L.total};
    const int e = plugin->launch(plugin->ctx, &pl, stream);
    if (e != 0) return hip_fail(static_cast<hipError_t>(e), __FILE__, __LINE__);
    return GB_OK;
}

// The allocation-free variant: the work vectors live in the library's static
// device array (relax_fp64.hip), whose address on the current device was
// looked up once per device; nothing is allocated or freed here.
static gb_status relax_batch_fp64_noalloc(const gb_objective_desc* o, const gb_algorithm_desc* a,
                                          const gb_lmo_desc* l, int N, int max_iters, double* x,
                                          const double* lb, const double* ub, const double* eps,
                                          int* iterations, uint8_t* status, double* gap, double* value,
                                          double* best_vertex, double* best_vertex_val, const uint8_t* flags,
                                          double ctol, int block_threads, void* stream) {
    Fp64Launch F;
    gb_status st = relax_fp64_setup(o, a, l, N, max_iters, x, lb, ub, eps, iterations, status, gap, value,
                                    best_vertex, best_vertex_val, flags, ctol, block_threads, &F);
    if (st != GB_OK || N == 0) return st;
    if (F.work_bytes > GB_NOALLOC_WORK_BYTES)
        return fail(GB_ERR_INVALID_ARGUMENT,
                    "batch too large for the allocation-free relaxation: N * n * 48 bytes of work vectors "
                    "exceed GB_NOALLOC_WORK_BYTES (" + std::to_string(GB_NOALLOC_WORK_BYTES) + ")");
    const int se = static_work_address(&F.P.work);
    if (se != 0) return hip_fail(static_cast<hipError_t>(se), __FILE__, __LINE__);
    const int e = launch_relax_fp64(F.P, F.threads, F.L.total, F.L, stream);
    if (e != 0) return hip_fail(static_cast<hipError_t>(e), __FILE__, __LINE__);
    return GB_OK;
}

// MultiBlock: K co-resident workgroups per node (relax_multiblock.hip); the
// nodes run in launches of as many complete groups as are resident at once.
gb_status relax_batch_multiblock(const gb_objective_desc* o, const gb_algorithm_desc* a, const gb_lmo_desc* l,
                                 int N, int max_iters, double* x, const double* lb, const double* ub,
                                 const double* eps, int* iterations, uint8_t* status, double* gap, double* value,
                                 double* best_vertex, double* best_vertex_val, const uint8_t* flags, double ctol,
                                 int block_threads, int blocks_per_node, void* stream) {
    gb_status st = validate_descriptors(o, a, l, false);
    if (st != GB_OK) return st;
    if (o->kind != GB_OBJ_DENSE_QUADRATIC || (l->kind != GB_LMO_BOX && l->kind != GB_LMO_SIMPLEX))
        return fail(GB_ERR_INVALID_ARGUMENT, "MultiBlock relaxes dense quadratics over a box or a simplex");
    const int K = blocks_per_node;
    if (K < 1 || K > kMbMaxK) return fail(GB_ERR_INVALID_ARGUMENT, "blocks_per_node must be 1..256");
    Fp64Launch F;
    st = relax_fp64_setup(o, a, l, N, max_iters, x, lb, ub, eps, iterations, status, gap, value, best_vertex,
                          best_vertex_val, flags, ctol, block_threads, &F, false);
    if (st != GB_OK || N == 0) return st;
    const int Tv = F.threads;  // the one-workgroup launch's block size: the order of every sum
    if (K * kMbThreads < Tv)
        return fail(GB_ERR_INVALID_ARGUMENT, "blocks_per_node * 256 must cover gpu_block_threads");
    const int cap = multiblock_capacity(l->kind);
    const int per_launch = std::min(N, cap / K);
    if (per_launch < 1) return fail(GB_ERR_INVALID_ARGUMENT, "blocks_per_node exceeds the resident workgroups");
    const int n = o->n;
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t work_b = up((size_t)per_launch * 6 * n * sizeof(double));
    const size_t red_b = up((size_t)per_launch * 3 * kMbSlots * sizeof(double));
    const size_t lstride = multiblock_lmo_bytes(l->kind, n);
    const size_t lwork_b = up((size_t)per_launch * lstride);
    const size_t bar_b = up((size_t)per_launch * sizeof(GroupBarrier));
    const size_t total = work_b + red_b + lwork_b + bar_b + sizeof(BarrierStatus) + 512;
    double* ws = nullptr;
    st = workspace_get(total, stream, &ws);
    if (st != GB_OK) return st;
    char* p = reinterpret_cast<char*>(ws);
    MbParams M{};
    M.K = K;
    M.Tv = Tv;
    M.work = reinterpret_cast<double*>(p);
    M.red = reinterpret_cast<double*>(p + work_b);
    M.lwork = reinterpret_cast<unsigned char*>(p + work_b + red_b);
    M.lwork_stride = lstride;
    char* bars = p + work_b + red_b + lwork_b;
    M.bars = reinterpret_cast<GroupBarrier*>(bars);
    M.bst = reinterpret_cast<BarrierStatus*>(bars + bar_b);
    M.spin = 1ull << 22;  // seconds of polling: a wait that long means the group is not resident
    {
        // the clear kernel's status is read with hipGetLastError, so an error
        // left by an earlier call would be reported as that launch's
        const hipError_t pending = hipGetLastError();
        if (pending != hipSuccess)
            return fail(GB_ERR_DEVICE, std::string("HIP error pending before the MultiBlock launch: ")
                                           + hipGetErrorString(pending));
    }
    for (int node0 = 0; node0 < N; node0 += per_launch) {
        const int nodes = std::min(per_launch, N - node0);
        M.node0 = node0;
        int e = launch_multiblock_clear(M.bars, nodes, M.bst, stream);
        if (e != 0) return hip_fail(static_cast<hipError_t>(e), __FILE__, __LINE__);
        e = launch_relax_multiblock(F.P, M, nodes, stream);
        if (e != 0) return hip_fail(static_cast<hipError_t>(e), __FILE__, __LINE__);
    }
    return GB_OK;
}

hipError_t create_dedicated_stream(hipStream_t* out) {
    // The runtime gives a stream with a CU mask a hardware queue of its own
    // instead of the least-used queue of its shared pool.
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    if (cus > 0) {
        std::vector<uint32_t> mask(static_cast<size_t>(cus + 31) / 32, 0u);
        for (int c = 0; c < cus; ++c) mask[c / 32] |= 1u << (c % 32);
        if (hipExtStreamCreateWithCUMask(out, static_cast<uint32_t>(mask.size()), mask.data()) == hipSuccess)
            return hipSuccess;
        (void)hipGetLastError();
    }
    return hipStreamCreate(out);
}

bool fast_simplex_ok(const gb_lmo_desc* l) {
    return l->kind == GB_LMO_SIMPLEX && std::isfinite(l->right_side) && l->right_side == std::floor(l->right_side)
        && std::fabs(l->right_side) < 1e15;
}

gb_status relax_batch_fast(const gb_objective_desc* o, const gb_algorithm_desc* a,
                           const gb_lmo_desc* l, int N, int max_iters, double* x, const double* lb,
                           const double* ub, const double* eps, int* iterations, uint8_t* status,
                           double* gap, double* value, double* best_vertex, double* best_vertex_val,
                           const uint8_t* flags, double ctol, const double* cutoff,
                           const gb_fast_options* opt, void* stream, int* counter, bool rescore,
                           const int* count) {
    gb_status st = validate_descriptors(o, a, l);
    if (st != GB_OK) return st;
    if (o->kind != GB_OBJ_DENSE_QUADRATIC || (l->kind != GB_LMO_BOX && l->kind != GB_LMO_SIMPLEX))
        return fail(GB_ERR_INVALID_ARGUMENT, "the batched form supports dense objectives over a box or a simplex");
    if (o->n > kMaxFastN) return fail(GB_ERR_INVALID_ARGUMENT, "the batched form supports n <= 512");
    if (l->kind == GB_LMO_SIMPLEX && !fast_simplex_ok(l))
        return fail(GB_ERR_INVALID_ARGUMENT, "the batched form's simplex needs a finite integral right_side");
    if (N < 0 || max_iters < 0) return fail(GB_ERR_INVALID_ARGUMENT, "N and max_iters must be non-negative");
    if (N == 0) return GB_OK;
    if (!x || !lb || !ub || !eps || !iterations || !status || !gap || !value || !best_vertex
            || !best_vertex_val)
        return fail(GB_ERR_INVALID_ARGUMENT, "null node buffer");
    gb_fast_options op{64, 0, 0, nullptr, nullptr};
    if (opt) op = *opt;
    if (op.waves_per_block < 0 || op.waves_per_block > 4 || op.refresh_interval < 0 || op.blocks < 0)
        return fail(GB_ERR_INVALID_ARGUMENT, "invalid batched-form options");
    ScopedCounter lent(stream);
    if (!counter) {
        st = lent.acquire();
        if (st != GB_OK) return st;
        counter = lent.get();
    }
    FastParams P{};
    P.n = o->n;
    P.n_pad = (o->n + 1) & ~1;
    P.Qt = o->Qt;
    P.lin = o->linear;
    P.gs = o->grad_scale;
    P.ls = o->linear_scale;
    P.kappa = o->constant;
    P.dicg = a->algorithm == GB_ALG_DICG;
    P.refresh = op.refresh_interval;
    P.N = N;
    P.max_iters = max_iters;
    P.x = x;
    P.lb = lb;
    P.ub = ub;
    P.eps = eps;
    P.iters = iterations;
    P.status = status;
    P.gap = gap;
    P.value = value;
    P.best_v = best_vertex;
    P.best_val = best_vertex_val;
    P.flags = flags;
    P.ctol = ctol;
    P.cutoff = cutoff;
    P.counter = counter;
    P.stats = op.stats;
    P.gout = op.gradient;
    P.rescore = rescore ? 1 : 0;
    P.count = count;
    P.simplex = l->kind == GB_LMO_SIMPLEX;
    P.equality = l->equality != 0;
    P.rhs = l->right_side;
    P.is_int = l->is_int;
    const int e = launch_relax_fast(P, op.waves_per_block, op.blocks, stream);
    if (e != 0) return hip_fail(static_cast<hipError_t>(e), __FILE__, __LINE__);
    return GB_OK;
}

}  // namespace gb

using namespace gb;

extern "C" gb_status gb_relax_batch_fast(const gb_objective_desc* objective,
                                         const gb_algorithm_desc* algorithm,
                                         const gb_lmo_desc* lmo, int N, int max_iters, double* x,
                                         const double* lb, const double* ub, const double* eps,
                                         int* iterations, uint8_t* status, double* gap,
                                         double* value, double* best_vertex,
                                         double* best_vertex_val, const uint8_t* integer_flags,
                                         double candidate_tol, const double* cutoff,
                                         const gb_fast_options* options, void* stream) {
    try {
        return relax_batch_fast(objective, algorithm, lmo, N, max_iters, x, lb, ub, eps,
                                iterations, status, gap, value, best_vertex, best_vertex_val,
                                integer_flags, candidate_tol, cutoff, options, stream);
    } catch (const std::exception& ex) {
        return fail(GB_ERR_RUNTIME, ex.what());
    }
}

// A captured batched-form launch: counter reset kernel + relaxation kernel,
// bound to fixed buffers and its own node counter.
struct gb_relax_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int* counter = nullptr;
    int device = 0;
};

extern "C" gb_status gb_relax_graph_destroy(gb_relax_graph* g) {
    if (!g) return GB_OK;
    int prev = 0;
    const bool have_prev = hipGetDevice(&prev) == hipSuccess;
    (void)hipSetDevice(g->device);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->counter) (void)hipFree(g->counter);
    if (have_prev) (void)hipSetDevice(prev);
    delete g;
    return GB_OK;
}

extern "C" gb_status gb_relax_graph_create(const gb_objective_desc* objective,
                                           const gb_algorithm_desc* algorithm,
                                           const gb_lmo_desc* lmo, int N, int max_iters, double* x,
                                           const double* lb, const double* ub, const double* eps,
                                           int* iterations, uint8_t* status, double* gap,
                                           double* value, double* best_vertex,
                                           double* best_vertex_val, const uint8_t* integer_flags,
                                           double candidate_tol, const double* cutoff,
                                           const gb_fast_options* options, gb_relax_graph** out) {
    if (!out) return fail(GB_ERR_INVALID_ARGUMENT, "null graph handle");
    *out = nullptr;
    if (N <= 0) return fail(GB_ERR_INVALID_ARGUMENT, "a captured launch needs N > 0");
    try {
        auto* g = new gb_relax_graph();
        hipStream_t s = nullptr;
        hipError_t e = hipGetDevice(&g->device);
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&g->counter), 64);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        if (e != hipSuccess) {
            gb_relax_graph_destroy(g);
            return hip_fail(e, __FILE__, __LINE__);
        }
        gb_status st = GB_OK;
        e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            st = relax_batch_fast(objective, algorithm, lmo, N, max_iters, x, lb, ub, eps, iterations,
                                  status, gap, value, best_vertex, best_vertex_val, integer_flags,
                                  candidate_tol, cutoff, options, s, g->counter);
            const hipError_t e2 = hipStreamEndCapture(s, &g->graph);
            if (st == GB_OK && e2 != hipSuccess) e = e2;
        }
        if (st == GB_OK && e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
        (void)hipStreamDestroy(s);
        if (st != GB_OK || e != hipSuccess) {
            const std::string msg = g_last_error;
            gb_relax_graph_destroy(g);
            if (st != GB_OK) return fail(st, msg);
            return hip_fail(e, __FILE__, __LINE__);
        }
        *out = g;
        return GB_OK;
    } catch (const std::exception& ex) {
        return fail(GB_ERR_RUNTIME, ex.what());
    }
}

extern "C" gb_status gb_relax_graph_launch(gb_relax_graph* g, void* stream) {
    if (!g || !g->exec) return fail(GB_ERR_INVALID_ARGUMENT, "null graph handle");
    GB_HIP(hipGraphLaunch(g->exec, reinterpret_cast<hipStream_t>(stream)));
    return GB_OK;
}

extern "C" gb_status gb_relax_batch_plugin(const gb_objective_desc* objective,
                                           const gb_algorithm_desc* algorithm, const gb_lmo_desc* lmo,
                                           const gb_plugin_desc* plugin, int N, int max_iters, double* x,
                                           const double* lb, const double* ub, const double* eps,
                                           int* iterations, uint8_t* status, double* gap, double* value,
                                           double* best_vertex, double* best_vertex_val,
                                           const uint8_t* integer_flags, double candidate_tol,
                                           int block_threads, void* stream) {
    try {
        return relax_batch_plugin(objective, algorithm, lmo, plugin, N, max_iters, x, lb, ub, eps, iterations,
                                  status, gap, value, best_vertex, best_vertex_val, integer_flags,
                                  candidate_tol, block_threads, stream);
    } catch (const std::exception& ex) {
        return fail(GB_ERR_RUNTIME, ex.what());
    }
}

extern "C" gb_status gb_relax_batch(const gb_objective_desc* objective,
                                    const gb_algorithm_desc* algorithm, const gb_lmo_desc* lmo,
                                    int N, int max_iters, double* x, const double* lb,
                                    const double* ub, const double* eps, int* iterations,
                                    uint8_t* status, double* gap, double* value,
                                    double* best_vertex, double* best_vertex_val,
                                    const uint8_t* integer_flags, double candidate_tol,
                                    int block_threads, void* stream) {
    try {
        return relax_batch_fp64(objective, algorithm, lmo, N, max_iters, x, lb, ub, eps,
                                iterations, status, gap, value, best_vertex, best_vertex_val,
                                integer_flags, candidate_tol, block_threads, stream);
    } catch (const std::exception& ex) {
        return fail(GB_ERR_RUNTIME, ex.what());
    }
}

extern "C" gb_status gb_relax_batch_multiblock(const gb_objective_desc* objective,
                                               const gb_algorithm_desc* algorithm, const gb_lmo_desc* lmo,
                                               int N, int max_iters, double* x, const double* lb,
                                               const double* ub, const double* eps, int* iterations,
                                               uint8_t* status, double* gap, double* value,
                                               double* best_vertex, double* best_vertex_val,
                                               const uint8_t* integer_flags, double candidate_tol,
                                               int block_threads, int blocks_per_node, void* stream) {
    try {
        return relax_batch_multiblock(objective, algorithm, lmo, N, max_iters, x, lb, ub, eps, iterations, status,
                                      gap, value, best_vertex, best_vertex_val, integer_flags, candidate_tol,
                                      block_threads, blocks_per_node, stream);
    } catch (const std::exception& ex) {
        return fail(GB_ERR_RUNTIME, ex.what());
    }
}

extern "C" gb_status gb_relax_batch_noalloc(const gb_objective_desc* objective,
                                            const gb_algorithm_desc* algorithm, const gb_lmo_desc* lmo,
                                            int N, int max_iters, double* x, const double* lb,
                                            const double* ub, const double* eps, int* iterations,
                                            uint8_t* status, double* gap, double* value,
                                            double* best_vertex, double* best_vertex_val,
                                            const uint8_t* integer_flags, double candidate_tol,
                                            int block_threads, void* stream) {
    try {
        return relax_batch_fp64_noalloc(objective, algorithm, lmo, N, max_iters, x, lb, ub, eps,
                                        iterations, status, gap, value, best_vertex, best_vertex_val,
                                        integer_flags, candidate_tol, block_threads, stream);
    } catch (const std::exception& ex) {
        return fail(GB_ERR_RUNTIME, ex.what());
    }
}

extern "C" gb_status gb_lmo_compute_v(const gb_lmo_desc* lmo, int n, int N, const double* direction,
                                      const double* lb, const double* ub, double* vertex, void* stream) {
    try {
        if (!lmo) return fail(GB_ERR_INVALID_ARGUMENT, "null descriptor");
        if (n <= 0 || N < 0) return fail(GB_ERR_INVALID_ARGUMENT, "n must be positive and N non-negative");
        if (N == 0) return GB_OK;
        if (!direction || !lb || !ub || !vertex) return fail(GB_ERR_INVALID_ARGUMENT, "null node buffer");
        if (lmo->kind == GB_LMO_BIRKHOFF && lmo->order * lmo->order != n)
            return fail(GB_ERR_INVALID_ARGUMENT, "Birkhoff LMO needs n == order^2");
        gb_status st = validate_lmo(lmo, n, GB_OBJ_DENSE_QUADRATIC, nullptr);
        if (st != GB_OK) return st;
        const LdsLayout L = lds_layout(-1, lmo->kind, n, lmo->order);
        if (L.total > kMaxLdsBytes) return fail(GB_ERR_INVALID_ARGUMENT, "problem exceeds the LDS budget");
        RelaxParams P{};
        P.obj_kind = -1;
        P.n = n;
        P.lmo_kind = lmo->kind;
        P.rhs = lmo->right_side;
        P.equality = lmo->equality;
        P.is_int = lmo->is_int;
        P.lorder = lmo->order;
        P.pcap = lmo->cap;
        P.pcard = lmo->cardinality;
        P.N = N;
        P.x = vertex;
        P.lb = lb;
        P.ub = ub;
        const int e = launch_lmo_probe(P, auto_block_threads(n), L.total, L, direction, stream);
        if (e != 0) return hip_fail(static_cast<hipError_t>(e), __FILE__, __LINE__);
        return GB_OK;
    } catch (const std::exception& ex) {
        return fail(GB_ERR_RUNTIME, ex.what());
    }
}

extern "C" const char* gb_last_error(void) { return g_last_error.c_str(); }

extern "C" const char* gb_version(void) { return "gpuboscia_amd 0.1.0 (gfx950)"; }

extern "C" int gb_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

extern "C" int gb_lmo_relaxation_feasible_cpu(const gb_lmo_desc* lmo, const uint8_t* is_int_host,
                                              const double* x, const double* lb, const double* ub,
                                              int n, double tol) {
    (void)is_int_host;
    if (!lmo || !x || n <= 0) return 0;
    switch (lmo->kind) {
    case GB_LMO_BOX:  // lmo.cuh:214-217: bounds are validated separately
        return 1;
    case GB_LMO_SIMPLEX: {  // lmo.cuh:476-485, lmo_dicg.h:156-178
        double mass = 0.0;
        if (lmo->equality) {
            for (int j = 0; j < n; ++j) {
                if (x[j] < -tol) return 0;
                mass += x[j];
            }
            return std::isfinite(mass) && std::fabs(mass - lmo->right_side) <= tol;
        }
        for (int j = 0; j < n; ++j) {
            if (x[j] < 0.0) return 0;
            mass += x[j];
        }
        return mass <= lmo->right_side + tol;
    }
    case GB_LMO_PORTFOLIO: {  // sum x = rhs, 0 <= x_i <= cap z_i, lb <= z <= ub, sum z <= card
        if (n % 2 != 0) return 0;
        const int m = n / 2;
        double xs = 0.0, zs = 0.0;
        for (int i = 0; i < m; ++i) {
            if (x[i] < -tol || x[i] > lmo->cap * x[m + i] + tol) return 0;
            if (lb && ub && (x[m + i] < lb[m + i] - tol || x[m + i] > ub[m + i] + tol)) return 0;
            xs += x[i];
            zs += x[m + i];
        }
        return std::fabs(xs - lmo->right_side) <= tol && zs <= lmo->cardinality + tol;
    }
    case GB_LMO_BIRKHOFF: {  // lmo.cuh:922-934
        const int m = lmo->order;
        if (n != m * m) return 0;
        for (int j = 0; j < n; ++j)
            if (x[j] < -tol || x[j] > 1 + tol) return 0;
        for (int i = 0; i < m; ++i) {
            double row = 0, col = 0;
            for (int j = 0; j < m; ++j) {
                row += x[i * m + j];
                col += x[j * m + i];
            }
            if (std::fabs(row - 1) > tol || std::fabs(col - 1) > tol) return 0;
        }
        return 1;
    }
    }
    return 0;
}
