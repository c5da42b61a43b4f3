/*
 * gpuboscia_amd — C-ABI of the MI355X (gfx950) Frank-Wolfe / DICG
 * node-relaxation library.
 *
 * The reference (LJS42/GPUBoscia) is a header-only C++/CUDA template
 * library; its plugin contract is compile-time (objective, algorithm and LMO
 * functors copied into kernels).  This C-ABI replaces that contract with
 * plain descriptors: every entry point takes plain pointers and sizes.  The
 * C++ drop-in header <gpuboscia/gpuboscia.hpp> maps the reference's types
 * (Settings, Result, BoxLMO, UnitSimplexLMO<N>, BirkhoffLMO<N>,
 * GlobalCachedDenseQuadraticObjective<N,T>, ...) onto these descriptors, and
 * compiles any other objective, LMO or heuristic type into the relaxation
 * kernel as a device plug-in (gb_plugin_desc, gb_relax_batch_plugin).
 *
 * Errors: every function returns a gb_status; on failure gb_last_error()
 * returns a thread-local message.  Device errors read "HIP <file>:<line> -
 * <message>" (the reference's DRS_CUDA_CHECK reads "CUDA ...").
 *
 * Device pointers are caller-owned, as in the reference (lmo.cuh:6-9,
 * heuristics.h:22).  Calls on one stream are ordered.  Library-owned
 * scratch is kept per (device, stream) (work vectors of gb_relax_batch) or
 * lent per call (node counters of the batched and mixed forms, returned
 * with an event recorded on the call's stream and lent again only after it
 * has signalled), so calls on different streams may run concurrently.  A
 * call captured into a HIP graph keeps its node counters for the graph's
 * replays; replays of one graph must not overlap.
 */
#ifndef GPUBOSCIA_AMD_H
#define GPUBOSCIA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    GB_OK = 0,
    GB_ERR_INVALID_ARGUMENT = 1, /* reference: std::invalid_argument   */
    GB_ERR_RUNTIME = 2,          /* reference: std::runtime_error      */
    GB_ERR_DEVICE = 3            /* reference: DRS_CUDA_CHECK failure  */
} gb_status;

/* FWAlgorithmKind (line_search_options.h:13-16). */
typedef enum { GB_ALG_FRANK_WOLFE = 0, GB_ALG_DICG = 1 } gb_algorithm_kind;

/* LineSearchKind (line_search_options.h:5-11). */
typedef enum {
    GB_LS_OPEN_LOOP = 0,
    GB_LS_SCALED_OPEN_LOOP = 1,
    GB_LS_SAFEGUARDED_SCALED_OPEN_LOOP = 2,
    GB_LS_SECANT = 3,
    GB_LS_QUADRATIC_EXACT = 4
} gb_line_search_kind;

/* Objective plugins. */
typedef enum {
    /* GlobalCachedDenseQuadraticObjective (dense_quadratic_objectives.cuh:18-112):
       f = 1/2 s_g x'Qx + s_g s_l c'x + kappa, Qt[k*n+j] == Q[j,k]. */
    GB_OBJ_DENSE_QUADRATIC = 0,
    /* Quadratic assignment over a Birkhoff polytope of order m (n = m*m):
       f(P) = ||A P - P B||_F^2, P row-major. No reference counterpart
       (analogue: examples/main.cu:131-181). */
    GB_OBJ_QAP = 1,
    /* A caller-compiled device objective (gb_plugin_desc): value_local, grad,
       optional prepare_block and quadratic_curvature_local, as the
       reference's objective contract (device_resident_solver.cuh:8-24). */
    GB_OBJ_PLUGIN = 2
} gb_objective_kind;

/* LMO plugins. */
typedef enum {
    GB_LMO_BOX = 0,       /* BoxLMO (lmo.cuh:212-321) */
    GB_LMO_SIMPLEX = 1,   /* UnitSimplexLMO / EqualitySimplexLMO (lmo.cuh:335-912) */
    GB_LMO_BIRKHOFF = 2,  /* BirkhoffLMO (lmo.cuh:918-1419) */
    /* Cardinality portfolio over v = (x, z), n = 2m: sum x = right_side,
       0 <= x_i <= cap z_i, lb_z <= z <= ub_z, sum z <= cardinality (no
       reference counterpart; analogue LinkedCardinalityLMO,
       examples/sparse_sharpe.cu:266-546).  Exact for gradients with gz >= 0;
       Frank-Wolfe only (no in-face oracle). */
    GB_LMO_PORTFOLIO = 3,
    /* A caller-compiled device LMO (gb_plugin_desc): compute_v,
       check_feasible_device, optional DICG trio and candidate hooks, as the
       reference's LMO contract (lmo.cuh:11-66, README.md:224-255). */
    GB_LMO_PLUGIN = 4
} gb_lmo_kind;

/* Candidate heuristics (heuristics.h:23-35). */
typedef enum {
    GB_HEUR_ROUND = 0,        /* RoundingHeuristic: round integer coordinates */
    GB_HEUR_FLOOR_OFFSET = 1, /* floor(x) + offset on integer coordinates    */
    GB_HEUR_PLUGIN = 2        /* a caller-compiled Heuristic::generate (gb_plugin_desc;
                                 heuristics.h:11-35, README.md:240-255) */
} gb_heuristic_kind;

/* Arithmetic of the relaxation. */
typedef enum {
    GB_PREC_FP64 = 0,      /* plain form: everything FP64 */
    GB_PREC_FP32 = 1,      /* iterates FP32, certificate FP64 */
    GB_PREC_BF16 = 2       /* BF16 MFMA gradient (FP32 accumulate), FP64 certificate */
} gb_precision;

typedef struct {
    int kind;                 /* gb_objective_kind */
    int n;                    /* number of variables */
    const double* Qt;         /* dense: device, n*n, column-major Q */
    const double* linear;     /* dense: device, n, may be NULL */
    double grad_scale;        /* s_g */
    double linear_scale;      /* s_l */
    double constant;          /* kappa */
    int order;                /* qap: m */
    const double* A;          /* qap: device, m*m row-major */
    const double* B;          /* qap: device, m*m row-major */
} gb_objective_desc;

typedef struct {
    int kind;                 /* gb_lmo_kind */
    double right_side;        /* simplex: rhs (default 1) */
    int equality;             /* simplex: 1 = EqualitySimplexLMO */
    const uint8_t* is_int;    /* simplex: device, n bytes (UnitSimplexLMO::is_int) */
    int order;                /* birkhoff: Order (n = Order*Order) */
    double cap;               /* portfolio: x_i <= cap z_i */
    double cardinality;       /* portfolio: sum z <= cardinality */
} gb_lmo_desc;

typedef struct {
    int algorithm;            /* gb_algorithm_kind */
    int line_search;          /* gb_line_search_kind */
    int secant_iters;         /* LineSearchOptions (line_search_options.h:18-23) */
    double secant_tol;
    double secant_denom_eps;
    int heuristic;            /* gb_heuristic_kind */
    double heuristic_offset;  /* GB_HEUR_FLOOR_OFFSET */
} gb_algorithm_desc;

/* Per-node status (device_resident_solver.cuh:192-202). */
enum {
    GB_NODE_FRACTIONAL_UNCONVERGED = 0,
    GB_NODE_FRACTIONAL_CONVERGED = 1,
    GB_NODE_INFEASIBLE = 2,
    GB_NODE_INTEGRAL_UNCONVERGED = 3,
    GB_NODE_INTEGRAL_CONVERGED = 4,
    GB_NODE_INITIAL_POINT_SUPPLIED = 0x80  /* input-only flag */
};

/*
 * Relax a batch of N nodes on one GPU, one workgroup per node, FP64.
 * Replaces drs_fw_relax_batch_kernel<Obj,Alg,LMO,Heur> and its launch
 * (device_resident_solver.cuh:631-660, :1834-1851); argument meaning as at
 * :634-654.  All arrays are device pointers, node-major [N][n]:
 *   x             in: warm start / ignored for fresh nodes; out: final iterate
 *   lb, ub        node bounds
 *   eps           [N] per-node FW gap tolerance
 *   iterations    [N] in: cumulative count (continuation); out: new count
 *   status        [N] in: 0 or GB_NODE_INITIAL_POINT_SUPPLIED; out: 0..4
 *   gap, value    [N] final FW gap and objective at the returned x
 *   best_vertex   [N][n], best_vertex_val [N]: best integer candidate (1e300 = none)
 *   integer_flags [n] u8, shared by all nodes (may be NULL)
 * block_threads: 0 = automatic (min(256, round_up(n, 64))), else a multiple
 * of 64 up to 1024.  stream: a hipStream_t (NULL = default stream).  The call
 * is asynchronous with respect to the host.
 */
gb_status gb_relax_batch(const gb_objective_desc* objective,
                         const gb_algorithm_desc* algorithm,
                         const gb_lmo_desc* lmo,
                         int N, int max_iters,
                         double* x, const double* lb, const double* ub,
                         const double* eps, int* iterations, uint8_t* status,
                         double* gap, double* value,
                         double* best_vertex, double* best_vertex_val,
                         const uint8_t* integer_flags, double candidate_tol,
                         int block_threads, void* stream);

/*
 * Device plug-ins: a caller-compiled objective, LMO and/or candidate
 * heuristic (the reference's template extension points: lmo.cuh:11-66,
 * heuristics.h:11-35, device_resident_solver.cuh:8-24 and :229-261,
 * README.md:224-255).  The caller builds the FP64 relaxation kernel with its
 * own types from include/gpuboscia/device/relax_fp64.hpp (hipcc, gfx950) and
 * registers a launcher; the library validates the batch, fills the launch
 * parameters and work vectors, and calls the launcher in place of its
 * built-in kernel.  include/gpuboscia/gpuboscia.hpp does all of this for
 * C++ types passed to gpuboscia::solve and DeviceResidentHybridSolver.
 *   launch     (ctx, launch, stream): launch = const gb::PluginLaunch*
 *              (relax_params.h); returns a hipError_t value (0 = success).
 *   ctx        passed back unchanged.
 *   objective  1: the objective descriptor has kind GB_OBJ_PLUGIN and gives
 *              only n; objective_bytes of global scratch per node.
 *   lmo        1: the LMO descriptor has kind GB_LMO_PLUGIN.
 *   heuristic  1: the algorithm's heuristic is GB_HEUR_PLUGIN.
 *   supports_dicg / supports_quadratic_exact: what the plug-in objective and
 *              LMO implement (DICG or the exact step are rejected otherwise,
 *              gpuboscia.cuh:135-149).
 *   block_threads  the objective's block_threads bound (0 = none).
 *   relaxation_feasible  host check of a root point (relaxation_feasible_cpu,
 *              lmo.cuh:30-35; tree.h:106-113), or NULL.
 */
typedef int (*gb_plugin_launch_fn)(void* ctx, const void* launch, void* stream);
typedef int (*gb_plugin_feasible_fn)(void* ctx, const double* x, const double* lb, const double* ub,
                                     int n, double tol);
typedef struct {
    gb_plugin_launch_fn launch;
    void* ctx;
    int objective;
    int lmo;
    int heuristic;
    int supports_dicg;
    int supports_quadratic_exact;
    int block_threads;
    size_t objective_bytes;
    gb_plugin_feasible_fn relaxation_feasible;
} gb_plugin_desc;

/*
 * gb_relax_batch with device plug-ins (the plain FP64 form; same arguments
 * and results).  Replaces drs_fw_relax_batch_kernel instantiated with user
 * Objective / LMO / Heuristic types (device_resident_solver.cuh:631-660).
 */
gb_status gb_relax_batch_plugin(const gb_objective_desc* objective,
                                const gb_algorithm_desc* algorithm,
                                const gb_lmo_desc* lmo,
                                const gb_plugin_desc* plugin,
                                int N, int max_iters,
                                double* x, const double* lb, const double* ub,
                                const double* eps, int* iterations, uint8_t* status,
                                double* gap, double* value,
                                double* best_vertex, double* best_vertex_val,
                                const uint8_t* integer_flags, double candidate_tol,
                                int block_threads, void* stream);

/*
 * Allocation-free variant of gb_relax_batch.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, semantics and results as gb_relax_batch, but the per-node
 * work vectors live in a static device array of GB_NOALLOC_WORK_BYTES per
 * GPU, held until the process ends, instead of a library allocation; the
 * call allocates and frees nothing.  Limit: N * n * 48 bytes must not exceed
 * GB_NOALLOC_WORK_BYTES (e.g. N = 1024 nodes up to n = 341), else
 * GB_ERR_INVALID_ARGUMENT.  Every call on one GPU shares that array: calls
 * that overlap on one GPU (different streams, or different host threads)
 * race and must be serialised by the caller.  The array keeps the last
 * call's contents; the kernel writes every work entry before reading it.
 */
#define GB_NOALLOC_WORK_BYTES (16u * 1024u * 1024u)
/*
 * MultiBlock relaxation (FWExecutionMode::MultiBlock with blocks_per_node,
 * device_resident_solver.cuh:1011-1013, multiblock.cuh, docs/multiblock.md):
 * gb_relax_batch with blocks_per_node co-resident workgroups of 256 threads
 * per node meeting at a device-atomic group barrier, for dense quadratics
 * over a box or a unit / equality simplex (n not limited by LDS; the
 * simplex ranks are taken over the whole group, lmo.cuh:620-784).  Every
 * product coordinate, sum, rank and scan follows the one-workgroup kernel's
 * order for block_threads (0 = automatic, as gb_relax_batch), so the outputs
 * equal gb_relax_batch's bit for bit where that kernel runs.  blocks_per_node * 256 >= block_threads; nodes run in launches of
 * as many complete groups as are resident.  A node whose group barrier timed
 * out (not resident together) gets status 0xFF.
 */
gb_status gb_relax_batch_multiblock(const gb_objective_desc* objective,
                                    const gb_algorithm_desc* algorithm,
                                    const gb_lmo_desc* lmo,
                                    int N, int max_iters,
                                    double* x, const double* lb, const double* ub,
                                    const double* eps, int* iterations, uint8_t* status,
                                    double* gap, double* value,
                                    double* best_vertex, double* best_vertex_val,
                                    const uint8_t* integer_flags, double candidate_tol,
                                    int block_threads, int blocks_per_node, void* stream);
gb_status gb_relax_batch_noalloc(const gb_objective_desc* objective,
                                 const gb_algorithm_desc* algorithm,
                                 const gb_lmo_desc* lmo,
                                 int N, int max_iters,
                                 double* x, const double* lb, const double* ub,
                                 const double* eps, int* iterations, uint8_t* status,
                                 double* gap, double* value,
                                 double* best_vertex, double* best_vertex_val,
                                 const uint8_t* integer_flags, double candidate_tol,
                                 int block_threads, void* stream);

/*
 * Extreme points of a built-in LMO for N given directions: vertex[i] =
 * argmin over the LMO's set with bounds lb[i], ub[i] of direction[i] . v,
 * one workgroup per node, by the same device code the relaxation kernels
 * call.  Replaces the LMO plugin's device compute_v (lmo.cuh:11-29) as the
 * reference's own probes launch it (benchmarks/benchmark_lmo.cu:16-38,
 * tests/test_pipeline.cu:193-299).  Device arrays [N][n]; an infeasible
 * Birkhoff node returns its forced edges only (all zeros when none),
 * like the reference's device LMO.  Asynchronous on `stream`.
 */
gb_status gb_lmo_compute_v(const gb_lmo_desc* lmo, int n, int N,
                           const double* direction, const double* lb, const double* ub,
                           double* vertex, void* stream);

/* Options of the batched-form relaxation (gb_relax_batch_fast). */
typedef struct {
    int refresh_interval;     /* recompute Q x directly every k iterations (0 = never) */
    int waves_per_block;      /* 1..4: nodes resident per workgroup, one wave each;
                                 0: one node per workgroup with its rows split over
                                 ceil(n/64) waves when the batch fits in flight
                                 (n > 64, over a simplex n > 128; else 4) */
    int blocks;               /* persistent workgroups (0 = automatic) */
    long long* stats;         /* optional device [N][2]: Q columns read by the incremental
                                 updates (a column serving Q v and Q a in one gather counts
                                 once) and direct n x n products per node (work
                                 accounting), or NULL */
    double* gradient;         /* optional device [N][n]: gradient at the returned x, or NULL */
} gb_fast_options;

/*
 * Batched-form relaxation of N nodes of a dense quadratic over a box or a
 * bounded simplex (no reference counterpart; it replaces the per-node block
 * of drs_fw_relax_batch_kernel, device_resident_solver.cuh:631-660, for
 * these LMOs).  One wave64 per node, persistent waves fed by a device node
 * counter, incremental FP64 products Q x / Q v / Q a, exact quadratic step,
 * early stop once f(x) - gap >= cutoff[i] (cutoff may be NULL), and a final
 * direct FP64 certificate (gap, value) at the returned x.  Buffers and
 * status codes as gb_relax_batch; the algorithm descriptor selects FW or
 * DICG (its line search is ignored: the step is always exact).  n <= 512.
 * Simplex (UnitSimplexLMO / EqualitySimplexLMO, lmo.cuh:335-912): the
 * right side must be integral and every node 0/1 (is_int set for every
 * variable, node bounds 0 or 1), where the greedy vertex is the first
 * rhs - sum lb eligible coordinates in (gradient, index) order, as the
 * reference's; a node outside that scope is not relaxed and gets status
 * 0xFF (value and gap NaN).  Over a simplex a supplied point is first
 * clipped to the node's bounds, coordinates within the face tolerance
 * (1e-6 + 1e-4 |bound|) are put on their bound, and the free ones are
 * scaled onto the mass limit if the point exceeds it (or, for an equality,
 * falls short of it).
 */
gb_status gb_relax_batch_fast(const gb_objective_desc* objective,
                              const gb_algorithm_desc* algorithm,
                              const gb_lmo_desc* lmo,
                              int N, int max_iters,
                              double* x, const double* lb, const double* ub,
                              const double* eps, int* iterations, uint8_t* status,
                              double* gap, double* value,
                              double* best_vertex, double* best_vertex_val,
                              const uint8_t* integer_flags, double candidate_tol,
                              const double* cutoff, const gb_fast_options* options,
                              void* stream);

/* ---------------------------------------------------------------------- */
/* Mixed-precision batched relaxation (no reference counterpart; the       */
/* reference computes every product in FP64, dense_quadratic_objectives.cuh:92-111). */
/* ---------------------------------------------------------------------- */

/* Q of a dense objective prepared for the BF16 MFMA products: split into
   `planes` bf16 planes (1: ~2^-8 relative per entry, 3: ~2^-24) in MFMA
   fragment order, plus s_g s_l c in FP32.  n <= 512.  Asynchronous on
   `stream`; the objective's Qt / linear must stay valid until it completes. */
typedef struct gb_mixed_q gb_mixed_q;
gb_status gb_mixed_prepare(const gb_objective_desc* objective, int planes, void* stream,
                           gb_mixed_q** out);
gb_status gb_mixed_destroy(gb_mixed_q* q);

/* Y[i] = Q V[i] (device [N][n] FP64 in and out) computed by the same MFMA
   fragments, V rounded to bf16 (exact for 0/1 vectors): the product error of
   the mixed form against FP64. */
gb_status gb_mixed_products(const gb_mixed_q* q, const double* V, int N, double* Y, void* stream);

/* The FP64 products that start pass 1 of gb_relax_batch_mixed at supplied
   points (diagnostic; no reference counterpart): out[i] = Q x[i], each entry
   one k-ordered FMA chain over Qt, for every node i with status
   GB_NODE_INITIAL_POINT_SUPPLIED and iterations[i] == 0; other rows of out
   are left as they are.  Dense objective, n <= 512; device arrays. */
gb_status gb_mixed_warm_products(const gb_objective_desc* objective, int N, const double* x,
                                 const uint8_t* status, const int* iterations, double* out,
                                 void* stream);

typedef struct {
    int row_tiles_per_wave;   /* 16-row tiles per wave (1..4; 0 = automatic); a tile
                                 (workgroup) runs ceil(n_pad / 16 / this) waves */
    int columns;              /* nodes per 16-column MFMA tile (1..16; 0 = automatic):
                                 fewer columns spread a small batch over more CUs */
    int blocks;               /* persistent tiles (0 = automatic) */
    float rel_floor;          /* the mixed pass stops at gap < max(eps, rel_floor (1 + |f|)) */
    int fp64_max_iters;       /* budget of the FP64 pass (0 = max(1, max_iters / 16), < 0 = skip it) */
    long long* mixed_iterations;  /* optional device [N]: iterations of the mixed pass */
    long long* fp64_stats;        /* optional device [N][2]: as gb_fast_options.stats */
    double* gradient;             /* optional device [N][n]: gradient at the returned x */
} gb_mixed_options;

/*
 * Mixed-precision relaxation of N nodes of a dense quadratic over a 0/1 box,
 * or (DICG only) over a bounded / equality simplex of 0/1 variables with an
 * integral right side, whose vertices pass 1 takes from the greedy on the
 * FP32 gradient (one wave per node column).
 * Pass 1 (BF16 MFMA): 16 nodes per wave64 in lockstep, products Q v / Q a
 * from the prepared planes with FP32 accumulation, FP32 iterates, FP64
 * step and reductions, DICG (or FW) with the exact step, per-node refill
 * from a device counter; it writes its iterate (status
 * GB_NODE_INITIAL_POINT_SUPPLIED) and its best vertex.  Pass 2: the FP64
 * batched kernel (gb_relax_batch_fast) continues from that iterate,
 * re-scores the vertex exactly, and returns FP64 gap, value and
 * certificate, so pruning decisions and incumbents are FP64.  Pass 1
 * starts a node with a supplied point (status GB_NODE_INITIAL_POINT_SUPPLIED)
 * at that point, with Q x computed in FP64 into the node's best_vertex row
 * first, when the tile shape allows it: row_tiles_per_wave != 3 and
 * n <= 256 (at most 16 row tiles of 16).  Other shapes start such a node at
 * the box vertex nearest to its point.  A node without a supplied point
 * starts at its upper bounds.  Nodes whose bounds are not all 0/1 and
 * continuations (iterations > 0 on entry) skip pass 1.
 * best_vertex rows hold a vertex on return only where best_vertex_val is
 * below 1e300.  fp64_max_iters < 0
 * runs pass 1 only (diagnostics).  Buffers as gb_relax_batch_fast;
 * best_vertex_val must be 1e300 on entry (as for the reference kernel).
 */
gb_status gb_relax_batch_mixed(const gb_mixed_q* q,
                               const gb_objective_desc* objective,
                               const gb_algorithm_desc* algorithm,
                               const gb_lmo_desc* lmo,
                               int N, int max_iters,
                               double* x, const double* lb, const double* ub,
                               const double* eps, int* iterations, uint8_t* status,
                               double* gap, double* value,
                               double* best_vertex, double* best_vertex_val,
                               const uint8_t* integer_flags, double candidate_tol,
                               const double* cutoff, const gb_mixed_options* options,
                               void* stream);

/*
 * Heterogeneous batches (no reference counterpart; the reference fixes n at
 * compile time, device_resident_solver.cuh:362, :1217-1218): independent
 * node batches of different problems and sizes relaxed in one call.  Each
 * entry is relaxed like gb_relax_batch_mixed (when `mixed` is set: dense
 * objective, n <= 512, a box LMO or, under DICG, a simplex with an integral
 * right side over 0/1 nodes), like gb_relax_batch_fast (other entries of
 * a dense objective with n <= 512 over a box, or over a simplex with an
 * integral right side, whose nodes must then be 0/1), or like
 * gb_relax_batch (every other entry: FP64, the algorithm's own line
 * search, cutoff unused), with the entry's own buffers, on one of
 * `concurrency` internal streams (1..8) forked from and joined back into
 * `stream`, so small batches of different instances share the GPU.  Every
 * entry's results equal those of its solo call; a mixed entry's solo call
 * being the one with gb_mixed_options.columns = min(16, max(c0, ceil(N k /
 * CUs))), where c0 is the automatic choice ceil(N / 256) clamped to 1..16
 * and k = min(concurrency, count), or k = 1 when `stream` is capturing (a
 * graph replay overlaps the streams less than eager calls do): the entries
 * share the chip, so their tiles widen towards the MFMA's 16 node columns,
 * which cuts the Q fragment reads per node (the same arithmetic as that
 * solo call).  A captured call thus equals an eager call with concurrency 1.
 * After one call on the same
 * buffers (which makes the library's allocations) a call can be captured in
 * a HIP graph on `stream` and replayed.
 */
typedef struct {
    const gb_objective_desc* objective;
    const gb_lmo_desc* lmo;
    const gb_mixed_q* mixed;      /* prepared Q for the mixed form, or NULL */
    int N;
    double* x;
    const double* lb;
    const double* ub;
    const double* eps;
    int* iterations;
    uint8_t* status;
    double* gap;
    double* value;
    double* best_vertex;
    double* best_vertex_val;
    const uint8_t* integer_flags;
    const double* cutoff;
} gb_instance_batch;

gb_status gb_relax_instances(int count, const gb_instance_batch* batches,
                             const gb_algorithm_desc* algorithm, int max_iters,
                             double candidate_tol, int concurrency, void* stream);

/*
 * FP32 form of the QAP relaxation (configuration "Birkhoff QAP: FP32
 * iterates, Hungarian on FP32 costs, FP64 bound"; no reference counterpart,
 * the reference runs BirkhoffLMO<Order> with an FP64 objective,
 * lmo.cuh:918-1419).  One workgroup per node, DICG with the exact quadratic
 * step: iterate and gradient in FP32 (pairwise steps update the gradient in
 * O(K m^2) for K rows moved), both linear minimisations exact assignments on
 * FP32 costs by a warm-started Hungarian with FP64 potentials, and every
 * returned value, gap and vertex value in FP64: gap = <G, P> - LB(v), LB the
 * Lagrangian bound of the solver's column potentials on the FP64 gradient,
 * so value - gap is a valid lower bound.  Orders up to 256; the m x m costs
 * sit in LDS up to m = 200.
 *   gb_qap32_prepare    FP32 copies of A, B and A'A, B B' (the objective's
 *                       A and B stay caller-owned and must outlive it).
 *   gb_relax_batch_qap32  gb_relax_batch's node buffers and meaning (n = m*m,
 *                       Birkhoff LMO of order m, node bounds forbid (ub < 1/2)
 *                       or force (lb > 1/2) edges); a node stops when a
 *                       vertex reaches stop_value (-inf: never); the FP32
 *                       gradient is recomputed from P every `refresh` steps
 *                       (0 = 64).  A supplied iterate that is not doubly
 *                       stochastic on the allowed edges is replaced by a
 *                       vertex.
 *   gb_qap32_assignments  T assignment problems solved in sequence by one
 *                       workgroup, each warm-started from the previous one:
 *                       costs [T][m][m] FP32 device (+inf = forbidden), perms
 *                       [T][m] row -> column (-1 if infeasible), bounds [T]
 *                       the Lagrangian bound of each solve.  Synchronous.
 */
typedef struct gb_qap32 gb_qap32;
gb_status gb_qap32_prepare(const gb_objective_desc* objective, void* stream, gb_qap32** out);
gb_status gb_qap32_destroy(gb_qap32* q);
gb_status gb_relax_batch_qap32(gb_qap32* q, const gb_lmo_desc* lmo, int N, int max_iters,
                               double* x, const double* lb, const double* ub,
                               const double* eps, int* iterations, uint8_t* status,
                               double* gap, double* value,
                               double* best_vertex, double* best_vertex_val,
                               const uint8_t* integer_flags, double candidate_tol,
                               double stop_value, int refresh, void* stream);
gb_status gb_qap32_assignments(int m, int T, const float* costs, int* perms, double* bounds, void* stream);

/*
 * Autotuner of the mixed-precision form (no reference counterpart): on a
 * scratch copy of a batch (the caller's buffers are read, never written) it
 * times the mixed pass for every valid tile shape (row tiles per wave 1..4,
 * node columns per tile 2..16) and keeps the fastest, then times the full
 * relaxation for FP64-pass budgets max_iters/32, /16, /8 and keeps the
 * lowest time per node that ends certified-converged.  Results are written
 * to *out (other fields of *out are kept); *seconds receives the winning
 * full-relaxation time.  Tuning changes speed only: every shape computes the
 * same algorithm, and the FP64 pass certifies whatever it returns.
 * Synchronous on `stream`; allocates its scratch buffers.
 */
gb_status gb_mixed_autotune(const gb_mixed_q* q, const gb_objective_desc* objective,
                            const gb_algorithm_desc* algorithm, const gb_lmo_desc* lmo,
                            int N, int max_iters, const double* x, const double* lb, const double* ub,
                            const double* eps, const int* iterations, const uint8_t* status,
                            const uint8_t* integer_flags, double candidate_tol, const double* cutoff,
                            gb_mixed_options* out, double* seconds, void* stream);

/*
 * Captured batched-form launch (hipGraph): the counter reset and the
 * persistent relaxation kernel of gb_relax_batch_fast recorded once for
 * fixed buffers and N, replayed with one gb_relax_graph_launch per batch.
 * Replays read the buffers' current contents, so a caller refills x / lb /
 * ub / eps / iterations / status in place between replays.  Each graph owns
 * its node counter; replays of one graph must not overlap.  No reference
 * counterpart: it replaces the per-batch launch path of
 * device_resident_solver.cuh:1775-1863 (run_slot).
 */
typedef struct gb_relax_graph gb_relax_graph;
gb_status gb_relax_graph_create(const gb_objective_desc* objective,
                                const gb_algorithm_desc* algorithm,
                                const gb_lmo_desc* lmo,
                                int N, int max_iters,
                                double* x, const double* lb, const double* ub,
                                const double* eps, int* iterations, uint8_t* status,
                                double* gap, double* value,
                                double* best_vertex, double* best_vertex_val,
                                const uint8_t* integer_flags, double candidate_tol,
                                const double* cutoff, const gb_fast_options* options,
                                gb_relax_graph** out);
gb_status gb_relax_graph_launch(gb_relax_graph* graph, void* stream);
gb_status gb_relax_graph_destroy(gb_relax_graph* graph);

/*
 * Self-test of the persistent-kernel group barrier on device atomics that
 * replaces the reference's cooperative grid syncs (drs_dicg_paired_batch_kernel,
 * device_resident_solver.cuh:764-1000; MultiBlockContext::node_sync,
 * multiblock.cuh:72-102).  Modelled on tests/test_multiblock.cu:8-31:
 * groups x group_blocks co-resident workgroups (cooperative launch) of
 * `threads` threads; group g runs phases_base + g*phase_step generations
 * of publish / sync / check / sync, and groups g > 0 start only after group
 * 0 has completed 100 generations.  Every wait is bounded by spin_limit
 * polls.  Outputs: *errors = visibility errors, *timeouts = waits that ran
 * out (both 0 on success), *seconds_per_sync = group 0's time per sync.
 * Synchronous; allocates its own buffers.
 */
gb_status gb_grid_barrier_selftest(int groups, int group_blocks, int threads,
                                   int phases_base, int phase_step, long long spin_limit,
                                   long long* errors, long long* timeouts,
                                   double* seconds_per_sync);

/* ---------------------------------------------------------------------- */
/* Learned branching scorer (no reference counterpart; the reference       */
/* branches on the most fractional variables, branching.h:24-118).          */
/* ---------------------------------------------------------------------- */

/* 8 -> 64 -> 64 -> 64 -> 1 ReLU MLP; parameters as floats
   W1[8][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] w4[64] b4. */
#define GB_MLP_FEATURES 8
#define GB_MLP_PARAMS 8961

/* Per-candidate features, in order:
   0 x_i, 1 fractionality min(x_i - floor, ceil - x_i), 2 g_i / max_j |g_j|,
   3 pseudocost-up gain, 4 pseudocost-down gain: mean bound gain per unit
   from the variable's relaxed children (or, without history, the
   quadratic estimate 1/2 s_g Q_ii d) times the distance d moved, divided
   by its mean over the node's candidates; 5 depth / n,
   6 FW gap / (1 + |f(x)|), 7 ub_i - lb_i. */

/* The feature rows the learned rule computes on the device for a relaxed
   batch (diagnostic; no reference counterpart): row i * n + j of F [N*n][8]
   holds the features of variable j of node i when j is a fractional
   candidate (integer flag set, distance to the nearest integer above tol),
   zeros otherwise.  Device arrays: X, G [N][n] iterate and gradient, LB, UB
   [N][n], depth [N] int, gap, value [N], pc_up, pc_down [n] pseudocost means
   (< 0: no history), Qt (the dense objective's, diagonal read in place) with
   grad_scale gs, flags [n]. */
gb_status gb_candidate_features(int n, int N, const double* X, const double* G, const double* LB,
                                const double* UB, const int* depth, const double* gap, const double* value,
                                const double* pc_up, const double* pc_down, const double* Qt, double gs,
                                const uint8_t* flags, double tol, float* F, void* stream);

typedef struct gb_brancher gb_brancher;

gb_status gb_brancher_create(unsigned long long seed, gb_brancher** out);
gb_status gb_brancher_destroy(gb_brancher* b);
gb_status gb_brancher_set_params(gb_brancher* b, const float* params, int count);
gb_status gb_brancher_get_params(gb_brancher* b, float* params, int count);
/* Feature standardization (x - mu) / sigma applied inside the network. */
gb_status gb_brancher_set_normalization(gb_brancher* b, const float* mu, const float* sigma);
/* Scores for M feature rows (device float [M][8] -> device float [M]). */
gb_status gb_brancher_predict(gb_brancher* b, const float* X, int M, float* Y, void* stream);
/* Mean squared error and its gradient (host float [GB_MLP_PARAMS], may be
   NULL) on device data; used for gradient checks. */
gb_status gb_brancher_loss_grad(gb_brancher* b, const float* X, const float* T, int M,
                                double* loss, float* grad_host);
/* Adam training on device data (X [M][8], T [M]), mini-batches in order. */
gb_status gb_brancher_train(gb_brancher* b, const float* X, const float* T, int M, int epochs,
                            int batch, float lr, double* final_loss);

/* ---------------------------------------------------------------------- */
/* Branch-and-bound (gpuboscia::solve, gpuboscia.cuh:56-170)               */
/* ---------------------------------------------------------------------- */

/* Problem (problem.h:4-10); host arrays. */
typedef struct {
    int n_vars;
    const double* lb;
    const double* ub;
    const int* int_vars;
    int n_int_vars;
} gb_problem;

/* progress(iter, open_nodes, bound, incumbent, new_incumbent)
   (gpuboscia.cuh:26, device_resident_solver.cuh:2078-2085). */
typedef void (*gb_progress_fn)(void* user, int iter, int open_nodes,
                               double bound, double incumbent,
                               int new_incumbent);
/* RootInitialPointValidator (tree.h:102-104): return nonzero if feasible. */
typedef int (*gb_root_validator_fn)(void* user, const double* x,
                                    const double* lb, const double* ub, int n,
                                    double tolerance);
/* Branching variable scorer: fills score[i] for each integer variable in
   candidates[0..count); larger is branched first.  Used by
   GB_BRANCH_CALLBACK. */
typedef void (*gb_branch_score_fn)(void* user, const double* x, int n,
                                   const int* candidates, int count,
                                   int depth, double* score);

/* A branch-and-bound node (node.h:9-33); host arrays of n_vars doubles. */
typedef struct {
    int id;
    int parent_id;
    int depth;
    const double* lb;
    const double* ub;
    const double* relaxation_x;   /* NULL = empty: the node starts fresh */
    int relaxation_iterations;    /* > 0: resume the open-loop schedule (needs relaxation_x) */
    double dual_bound;
} gb_node;

/* The child list a BranchStrategy callback fills. */
typedef struct gb_children gb_children;
/* Appends a copy of *child (id, parent_id, depth, dual_bound, lb, ub) to the
   list; relaxation_x and relaxation_iterations are ignored, since children
   start fresh (device_resident_solver.cuh:1990-1994). */
gb_status gb_children_push(gb_children* children, const gb_node* child);

/* Custom BranchStrategy (branching.h:13-18), called on the host after a
   node's relaxation, as device_resident_solver.cuh:1986-1995 calls it:
   `parent` is the node with its certified dual bound and the
   relaxation_x / relaxation_iterations it was launched with; *next_id is the
   first id reserved for its children (8 are reserved, as the reference's
   fetch_add(8); a callback that takes more moves later reservations past
   them); x is the node's final iterate.  Children are appended with
   gb_children_push; children whose bound reaches the incumbent are dropped.
   Return 0, or nonzero to stop the solve with GB_ERR_RUNTIME. */
typedef int (*gb_branch_fn)(void* user, const gb_node* parent, int* next_id, const double* x, int n,
                            const int* int_vars, int n_int_vars, gb_children* children);

typedef enum {
    GB_BRANCH_POW2 = 0,            /* branch_pow2(k) on top_k_infeasible (default k=3) */
    GB_BRANCH_MOST_FRACTIONAL = 1, /* single variable, most_infeasible */
    GB_BRANCH_LEARNED = 2,         /* on-GPU MLP scorer: the branch_k best-scored candidates,
                                      split 2^k ways; most-fractional without a model */
    GB_BRANCH_CALLBACK = 3,        /* host scorer callback, 2-way split */
    GB_BRANCH_CUSTOM = 4           /* host BranchStrategy callback (gb_branch_fn) */
} gb_branch_rule;

/* Settings + DeviceResidentConfig (gpuboscia.cuh:21-32,
   device_resident_solver.cuh:1004-1025). */
typedef struct {
    /* DeviceResidentConfig */
    int gpu_batch_size;       /* 1024 */
    int num_streams;          /* 8; with >= 2 the batched FP64 form keeps two node batches in
                                 flight and the mixed form min(3, num_streams) (1: one batch
                                 at a time); the plain form always keeps one */
    int gpu_block_threads;    /* 0 = automatic; else a multiple of 64 */
    int gpu_max_iters;        /* 20000 */
    double gpu_eps;           /* 1e-3 */
    double gpu_eps_decay;     /* 0.8 */
    double gpu_eps_min;       /* 1e-6 */
    double integer_tol;       /* 1e-6 */
    double abs_gap_tol;       /* 1e-6 */
    double rel_gap_tol;       /* 1e-2 */
    /* Settings */
    gb_algorithm_desc algorithm;   /* DICG + Secant(4, 1e-10, 1e-14) */
    int branch_rule;          /* gb_branch_rule */
    int branch_k;             /* 3 */
    double time_limit_seconds;     /* +inf */
    double natural_lower_bound;    /* -inf */
    const double* root_initial_point;  /* host, n_vars, or NULL */
    int root_initial_point_size;
    gb_root_validator_fn root_validator;
    void* root_validator_user;
    gb_progress_fn progress;
    void* progress_user;
    gb_branch_score_fn branch_score;
    void* branch_score_user;
    gb_branch_fn branch;      /* GB_BRANCH_CUSTOM */
    void* branch_user;
    int precision;            /* gb_precision (GB_PREC_FP64 = plain form) */
    int form;                 /* 0 = plain (reference schedule), 1 = batched FP64 form,
                                 2 = mixed-precision form (BF16/FP32 MFMA + FP64 certificate) */
    gb_brancher* learned_model;   /* GB_BRANCH_LEARNED scorer, or NULL (most fractional) */
    long long max_nodes;      /* safety cap; <= 0 = none (no reference counterpart) */
    int autotune;             /* form 2: tune tile shape and FP64-pass budget on the first
                                 batch of at least 64 nodes (gb_mixed_autotune); 0 = defaults */
    int device_frontier;      /* 1: device-resident search (no reference counterpart): the
                                 frontier, incumbent, pruning and branching run on the GPU and
                                 each round is one captured HIP graph; forms 1 and 2, dense
                                 quadratic over a 0/1 box, pow2 (branch_k <= 3) or most-
                                 fractional branching, gpu_batch_size <= 1024, no callbacks.
                                 Same tree, incumbent, bound and counts as num_streams = 1;
                                 the time limit is checked between groups of rounds */
    long long device_nodes;   /* node slots of the device-resident search (0 = automatic: a
                                 quarter of free device memory, at most 32 GiB) */
    int blocks_per_node;      /* plain form: > 1 relaxes every node with that many workgroups
                                 (MultiBlock, gb_relax_batch_multiblock; dense quadratic over a
                                 box or a simplex); 0 or 1: one workgroup per node (SingleBlock) */
    const gb_plugin_desc* plugin;  /* device plug-ins (gb_relax_batch_plugin), plain form only;
                                      NULL = built-in kinds */
} gb_settings;

/* Result (gpuboscia.cuh:34-42). x is caller-provided, n_vars doubles. */
typedef struct {
    double objective;
    double* x;
    double dual_bound;
    int nodes_processed;
    double elapsed_seconds;
    int reached_tolerance;
    int reached_time_limit;
    long long fw_iterations;  /* total FW iterations (no reference counterpart) */
    double relax_seconds;     /* time inside node relaxations */
} gb_result;

/* Fill *s with the reference defaults (create_default_settings,
   gpuboscia.cuh:44-47). */
void gb_default_settings(gb_settings* s);

/* Run branch-and-bound to the proven optimum (or a limit).
   Validation and errors follow gpuboscia.cuh:56-170 and tree.h:115-147. */
gb_status gb_solve(const gb_problem* problem,
                   const gb_objective_desc* objective,
                   const gb_lmo_desc* lmo,
                   const gb_settings* settings,
                   gb_result* result);

/* ---------------------------------------------------------------------- */
/* The low-level solver object (DeviceResidentHybridSolver,                 */
/* device_resident_solver.cuh:1184-1298): nodes pushed by the caller, a      */
/* scheduling observer, and the accessors of the last solve.                */
/* ---------------------------------------------------------------------- */

/* DRSSchedulingEventType / DRSSchedulingEvent (device_resident_solver.cuh:1027-1038). */
typedef enum {
    GB_SCHED_DISPATCHED = 0,
    GB_SCHED_COMPLETION_DETACHED = 1,
    GB_SCHED_POSTPROCESS_STARTED = 2
} gb_scheduling_event_type;

typedef struct {
    int type;               /* gb_scheduling_event_type */
    int slot_id;            /* the stream slot */
    unsigned long long generation;  /* dispatches of that slot so far, from 1 */
    int n_nodes;
} gb_scheduling_event;

/* Called on the solving thread; must not call back into the solver. */
typedef void (*gb_scheduling_fn)(void* user, const gb_scheduling_event* event);

typedef struct gb_solver gb_solver;

/* Constructor checks of the reference (device_resident_solver.cuh:1198-1218)
   and the descriptor checks of gb_solve.  The problem's arrays are copied;
   the objective's and LMO's device pointers must stay valid while the
   solver solves.  settings->time_limit_seconds is the constructor's time
   limit; settings->root_initial_point and natural_lower_bound are not used
   (the caller pushes its nodes). */
gb_status gb_solver_create(const gb_problem* problem, const gb_objective_desc* objective,
                           const gb_lmo_desc* lmo, const gb_settings* settings, gb_solver** out);
gb_status gb_solver_destroy(gb_solver* solver);
/* push (device_resident_solver.cuh:1247-1260): the same checks and messages. */
gb_status gb_solver_push(gb_solver* solver, const gb_node* node);
gb_status gb_solver_set_scheduling_callback(gb_solver* solver, gb_scheduling_fn fn, void* user);
/* set_progress_callback (device_resident_solver.cuh:1262-1266). */
gb_status gb_solver_set_progress_callback(gb_solver* solver, gb_progress_fn fn, void* user);
/* solve: relaxes the pushed nodes and their descendants to the proven
   optimum or a limit; the pushed nodes are consumed. */
gb_status gb_solver_solve(gb_solver* solver);

/* The accessors (device_resident_solver.cuh:1280-1298) after the last solve. */
typedef struct {
    double incumbent_value;
    double dual_bound;
    int nodes_processed;
    int last_gpu_iterations;
    double last_gpu_gap;
    int last_gpu_converged;
    int reached_tolerance;
    int reached_time_limit;
    int effective_gpu_batch_size;
    int effective_num_streams;
    int paired_gpu_batches;       /* always 0: no paired cooperative launches here */
    long long fw_iterations;      /* no reference counterpart */
    double elapsed_seconds;
    double relax_seconds;
} gb_solver_stats;
gb_status gb_solver_get_stats(const gb_solver* solver, gb_solver_stats* out);
/* incumbent_x: n_vars doubles into x; *has = 0 (x untouched) when there is no
   incumbent (the reference returns an empty vector). */
gb_status gb_solver_incumbent_x(const gb_solver* solver, double* x, int* has);

/* Strong-branching labels generated by the solver itself: runs the batched
   B&B with branch_pow2 on the settings' branch_k most fractional variables
   (the tree the scorer is used in) and, for the first max_nodes branched
   nodes, relaxes both children of every fractional candidate on the GPU.
   Each candidate yields a feature row (host float [8]) and the label
   score_i / max_j score_j with score = max(dn, 1e-6) * max(up, 1e-6) over
   the certified child-bound improvements.  Rows are appended up to
   capacity; *count receives the number written. */
gb_status gb_branching_dataset(const gb_problem* problem, const gb_objective_desc* objective,
                               const gb_lmo_desc* lmo, const gb_settings* settings,
                               int max_nodes, float* features, float* labels, int capacity,
                               int* count);

/* Host-side continuous feasibility check of the built-in LMOs
   (LMO::relaxation_feasible_cpu, lmo.cuh:30-35). is_int is a host array for
   the simplex kind (may be NULL). Returns 1 feasible, 0 not. */
int gb_lmo_relaxation_feasible_cpu(const gb_lmo_desc* lmo, const uint8_t* is_int_host,
                                   const double* x, const double* lb,
                                   const double* ub, int n, double tolerance);

/* Releases what the library holds on every device: pooled batch buffers and
   their streams, fork streams, node counters and work vectors (no reference
   counterpart; the reference frees its buffers in the solver's destructor,
   device_resident_solver.cuh:1242-1245).  Waits for the library's work to
   finish first.  It also runs by itself at process exit.  Call it only
   while no other call into the library is running; later calls re-create
   what they need.  Objects the caller owns (gb_mixed_q, gb_relax_graph,
   gb_brancher) are not touched. */
gb_status gb_shutdown(void);

/* Thread-local message of the last failure. */
const char* gb_last_error(void);

/* Library version string. */
const char* gb_version(void);

/* Number of visible HIP devices (0 when none); never throws. */
int gb_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* GPUBOSCIA_AMD_H */
