/*
 * CPU oracle for the Frank-Wolfe / DICG node-relaxation path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (gpuboscia_amd/, the
 * C-ABI library, the C++ drop-in header) links, loads or calls this code.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * it, and only as the checker or as the timed CPU baseline.
 *
 * It is a plain FP64 C++ restatement of the reference algorithm
 * (LJS42/GPUBoscia, headers under include/gpuboscia/).  Every function cites the
 * reference file:line it follows.  Block reductions use the canonical wave64
 * order documented in gb_oracle.cpp (the HIP kernels use the same order), so
 * per-node trajectories are reproducible bit for bit.
 *
 * Pinning status of each part is listed in the header of gb_oracle.cpp.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Objective kinds (mirrors the reference's objective plugins). */
enum { ORC_OBJ_DENSE = 0, ORC_OBJ_QAP = 1 };
/* LMO kinds. */
enum { ORC_LMO_BOX = 0, ORC_LMO_SIMPLEX = 1, ORC_LMO_BIRKHOFF = 2,
       ORC_LMO_PORTFOLIO = 3 };
/* Line searches (line_search_options.h:5-11). */
enum { ORC_LS_OPEN_LOOP = 0, ORC_LS_SCALED = 1, ORC_LS_SAFEGUARDED = 2,
       ORC_LS_SECANT = 3, ORC_LS_QUADRATIC_EXACT = 4 };

typedef struct {
    int kind;
    int n;
    /* dense: Qt[k*n+j] == Q[j,k] (dense_quadratic_objectives.cuh:25-27) */
    const double* Qt;
    const double* linear;   /* may be NULL */
    double grad_scale, linear_scale, constant;
    /* qap: f(P) = ||A P - P B||_F^2, P row-major order x order */
    int order;
    const double* A;
    const double* B;
} orc_objective;

typedef struct {
    int kind;
    double right_side;        /* simplex rhs */
    int equality;             /* simplex: sum == rhs */
    const uint8_t* is_int;    /* simplex: per-variable integer marks (n) */
    int order;                /* birkhoff order */
    /* portfolio: variables (x[0..m), z[m..2m)); sum x = 1, 0<=x_i<=cap*z_i,
       sum z <= card */
    int assets;
    double cap;
    double card;
} orc_lmo;

typedef struct {
    int dicg;                 /* 0 = Frank-Wolfe, 1 = DICG */
    int ls;                   /* ORC_LS_* */
    int secant_iters;
    double secant_tol;
    double secant_denom_eps;
    int heuristic;            /* 0 = round (RoundingHeuristic), 1 = floor + offset */
    double heuristic_offset;
} orc_alg;

/* Canonical block reductions (exposed for the reduction-order tests). */
double orc_block_sum(const double* lane_values, int lanes);

/* Objective primitives. */
void   orc_grad(const orc_objective* o, const double* x, double* g);
double orc_value(const orc_objective* o, const double* x, int lanes);
double orc_curvature(const orc_objective* o, const double* d, int lanes);

/* LMO primitives (device semantics, lmo.cuh). */
void orc_compute_v(const orc_lmo* l, int n, const double* g, const double* lb,
                   const double* ub, double* v, int lanes);
void orc_compute_inface(const orc_lmo* l, int n, const double* g,
                        const double* x, const double* lb, const double* ub,
                        double* a, int lanes);
double orc_max_step(const orc_lmo* l, int n, const double* d, const double* x,
                    const double* lb, const double* ub, int lanes);
int orc_node_feasible(const orc_lmo* l, int n, const double* lb,
                      const double* ub, const uint8_t* flags, int lanes);

/* Host helpers restated from the reference CPU headers. */
double orc_quadratic_exact_step(double gap, double curvature, double maximum);
double orc_certified_fw_lower_bound(double inherited, double primal, double gap);
double orc_gap_absolute(double bound, double incumbent);
double orc_gap_relative(double bound, double incumbent);
int    orc_gap_within_tolerance(double bound, double incumbent, double abs_tol,
                                double rel_tol);
double orc_depth_dependent_tolerance(double root, double decay, double minimum,
                                     int depth);
int orc_top_k_infeasible(const double* x, const int* int_vars, int n_int,
                         int k, double integer_tol, int* out);
int orc_most_infeasible(const double* x, const int* int_vars, int n_int);
void orc_box_inface_lmo_cpu(int n, const double* grad, const double* x,
                            const double* lb, const double* ub, double* a,
                            double atol, double rtol);
double orc_box_dicg_maximum_step_cpu(int n, const double* d, const double* x,
                                     const double* lb, const double* ub);
void orc_unit_simplex_inface_lmo_cpu(int n, double rhs, const uint8_t* is_int,
                                     const double* grad, const double* x,
                                     const double* lb, const double* ub,
                                     double* a, double atol, double rtol,
                                     int require_full_mass);
double orc_unit_simplex_dicg_maximum_step_cpu(int n, double rhs,
    const uint8_t* is_int, const double* d, const double* x, const double* lb,
    const double* ub, double tol, int require_full_mass);
/* Birkhoff exact assignment (lmo.cuh:1063-1171 == :1239-1417). Returns 1 on
   success. face may be NULL. */
int orc_birkhoff_solve(int order, const double* cost, const double* lb,
                       const double* ub, const double* face, double* out);

/* One batch of node relaxations (device_resident_solver.cuh:337-629).
   Arrays are node-major [N][n]; in/out semantics as the reference kernel. */
void orc_relax_batch(const orc_objective* o, const orc_alg* a,
                     const orc_lmo* l, int N, int max_iters, double* x,
                     const double* lb, const double* ub, const double* eps,
                     int* iterations, uint8_t* status, double* gap,
                     double* value, double* best_vertex,
                     double* best_vertex_val, const uint8_t* flags,
                     double candidate_tol, int lanes, int threads);

/* Batched-form relaxation (box LMO, dense objective), mirroring the wave64
   kernel operation for operation; cutoff may be NULL. */
void orc_relax_fast_batch(const orc_objective* o, int dicg, int N, int max_iters, double* x,
                          const double* lb, const double* ub, const double* eps,
                          int* iterations, uint8_t* status, double* gap, double* value,
                          double* best_vertex, double* best_vertex_val, const uint8_t* flags,
                          double candidate_tol, const double* cutoff, int refresh, int rescore,
                          int threads);

/* The batched form's simplex vertices on 0/1 data (lb, ub in {0,1},
   integral rhs): forward vertex for direction h, and in-face vertex at x
   for direction h (the away vertex passes h = -g). */
void orc_fast_simplex_vertex(int n, const double* h, const double* lb, const double* ub, double rhs,
                             int equality, double* out);
void orc_fast_simplex_inface(int n, const double* h, const double* x, const double* lb,
                             const double* ub, double rhs, int equality, double* out);

/* The same with an LMO: NULL or a box as above; a simplex over 0/1
   variables (every is_int set, 0/1 node bounds, integral rhs; other nodes
   get status 0xFF) with the batched form's greedy vertex. */
void orc_relax_fast_batch_lmo(const orc_objective* o, const orc_lmo* l, int dicg, int N, int max_iters,
                              double* x, const double* lb, const double* ub, const double* eps,
                              int* iterations, uint8_t* status, double* gap, double* value,
                              double* best_vertex, double* best_vertex_val, const uint8_t* flags,
                              double candidate_tol, const double* cutoff, int refresh, int rescore,
                              int threads);

/* Deterministic best-first B&B over orc_relax_batch. */
typedef struct {
    int batch_size;
    int max_iters;
    double eps, eps_decay, eps_min, integer_tol, abs_gap_tol, rel_gap_tol;
    int branch_k;             /* branch_pow2 k */
    int lanes;
    int threads;              /* CPU worker threads for the batch relax */
    double natural_lower_bound;
    double time_limit_seconds;
    long long max_nodes;      /* safety cap; <=0 = none */
    int form;                 /* 0 = plain schedule, 1 = batched form (fast relax, warm starts, cutoff) */
    int refresh;              /* batched form: direct Q x recompute interval */
    int num_streams;          /* form 1 with >= 2: two batches in flight, the next one popped
                                 before the current one is post-processed; otherwise the next
                                 one is popped first only when a full batch is waiting (the
                                 reference's refill before post-processing) */
} orc_bnb_settings;

typedef struct {
    double objective;
    double dual_bound;
    int nodes_processed;
    int reached_tolerance;
    int reached_time_limit;
    long long fw_iterations;
    double elapsed_seconds;
} orc_bnb_result;

int orc_bnb_solve(const orc_objective* o, const orc_alg* a, const orc_lmo* l,
                  int n, const double* lb, const double* ub,
                  const int* int_vars, int n_int, const orc_bnb_settings* s,
                  orc_bnb_result* r, double* x_out);

/* Exhaustive minimum of x^T Q x + c^T x over x in {0,1}^n (n <= 30);
   Q given row-major. Ties keep the lowest bit pattern. */
double orc_enumerate_binary_qp(int n, const double* Q, const double* c,
                               int threads, double* x_out);

/* Secant line search of one node (algorithms.cuh:121-258): the step along
   dir from x, given the FW gap (phi(0) = -gap) and the maximum step. */
double orc_secant_search(const orc_objective* o, const orc_alg* a, int n, int lanes,
                         double gap, double max_gamma, const double* x, const double* dir);

/* Least-squares instance 0.5 ||Ax - b||^2 of the reference's convergence
   matrix (tests/test_pipeline.cu:1068-1093), drawn with std::mt19937(seed)
   and std::normal_distribution<double> as that test does: Qt (n*n), linear
   (n), constant. */
void orc_least_squares_instance(unsigned seed, int m, int n, double* Qt, double* linear,
                                double* constant);

#ifdef __cplusplus
}
#endif
