// C++ drop-in for the reference's public API (include/gpuboscia/gpuboscia.cuh,
// problem.h, node.h, lmo.cuh, dense_quadratic_objectives.cuh,
// line_search_options.h, device_resident_solver.cuh:1004-1025), implemented
// over the gpuboscia_amd C-ABI (gpuboscia_amd.h).
//
// Same names, fields, defaults and exception types as the reference:
//   Problem, Node, BranchStrategy, LineSearchKind, FWAlgorithmKind,
//   LineSearchOptions, DeviceResidentConfig, BoxLMO, UnitSimplexLMO<MaxN,Eq>,
//   EqualitySimplexLMO<MaxN>, BirkhoffLMO<Order>,
//   GlobalCachedDenseQuadraticObjective<Vars,Threads>,
//   ILSGlobalCachedObjective<Vars,Threads>, gpuboscia::Settings,
//   gpuboscia::Result, gpuboscia::create_default_settings, gpuboscia::solve,
//   gpuboscia::RoundingHeuristic.
// Objectives hold caller-owned device pointers exactly as in the reference;
// the device computations are the library's gfx950 kernels, so objective and
// LMO kinds are the built-in ones (plus QuadraticAssignmentObjective).
// Host-side callbacks (progress, root validator, BranchStrategy) are plain
// std::function values as in the reference and run on the calling thread.
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <functional>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../gpuboscia_amd.h"

// ---- problem.h / node.h / branching.h / line_search_options.h ---------------
struct Problem {
    int n_vars;
    std::vector<double> lb;
    std::vector<double> ub;
    std::vector<int> int_vars;
};

enum class NodeStatus { Active, Solved };

struct Node {
    int id;
    int parent_id;
    int depth;
    std::vector<double> lb;
    std::vector<double> ub;
    std::vector<double> relaxation_x;
    int relaxation_iterations = 0;
    double dual_bound;
    NodeStatus status;
    Node(int id_, int parent_, int depth_, std::vector<double> lb_, std::vector<double> ub_)
        : id(id_), parent_id(parent_), depth(depth_), lb(std::move(lb_)), ub(std::move(ub_)),
          dual_bound(-1e300), status(NodeStatus::Active) {}
};

using BranchStrategy = std::function<std::vector<Node>(
    const Node& parent, int& next_id, const std::vector<double>& x, const std::vector<int>& int_vars)>;

enum class LineSearchKind { OpenLoop, ScaledOpenLoop, SafeguardedScaledOpenLoop, Secant, QuadraticExact };
enum class FWAlgorithmKind { FrankWolfe, DICG };

struct LineSearchOptions {
    LineSearchKind kind = LineSearchKind::Secant;
    int secant_iters = 4;
    double secant_tol = 1e-10;
    double secant_denom_eps = 1e-14;
};

enum class FWExecutionMode { SingleBlock, MultiBlock };

// device_resident_solver.cuh:1004-1025 (same fields and defaults).  Fields
// describing CUDA launch policies (paired DICG, MultiBlock) are accepted and
// validated; the gfx950 kernels choose their own launch shapes.
struct DeviceResidentConfig {
    enum class DICGOracleParallelism { Auto, Serial, Paired };
    int gpu_batch_size = 1024;
    int num_streams = 8;
    int gpu_block_threads = 0;      // 0 = automatic; else a multiple of 64 (wave64)
    int gpu_max_iters = 20000;
    double gpu_eps = 1e-3;
    double gpu_eps_decay = 0.8;
    double gpu_eps_min = 1e-6;
    double integer_tol = 1e-6;
    double abs_gap_tol = 1e-6;
    double rel_gap_tol = 1e-2;
    DICGOracleParallelism dicg_oracle_parallelism = DICGOracleParallelism::Auto;
    double dicg_paired_occupancy_fraction = 0.8;
    FWExecutionMode execution_mode = FWExecutionMode::SingleBlock;
    int blocks_per_node = 4;
};

// ---- LMOs (lmo.cuh) ----------------------------------------------------------
struct BoxLMO {
    static constexpr bool supports_dicg = true;
    bool relaxation_feasible_cpu(const std::vector<double>&, const std::vector<double>&,
                                 const std::vector<double>&, double) const {
        return true;
    }
    gb_lmo_desc desc(const uint8_t*) const { return gb_lmo_desc{GB_LMO_BOX, 1.0, 0, nullptr, 0}; }
    std::vector<uint8_t> integer_marks(int) const { return {}; }
};

template <int MaxN, bool RequireEquality = false>
struct UnitSimplexLMO {
    static constexpr bool supports_dicg = true;
    double right_side = 1.0;
    unsigned char is_int[MaxN]{};
    UnitSimplexLMO() = default;
    UnitSimplexLMO(double rhs, int n, const std::vector<int>& int_vars) : right_side(rhs) {
        set_integer_vars(n, int_vars);
    }
    void set_integer_vars(int n, const std::vector<int>& int_vars) {
        std::fill(is_int, is_int + MaxN, (unsigned char)0);
        const int capped = std::min(n, MaxN);
        for (int idx : int_vars)
            if (idx >= 0 && idx < capped) is_int[idx] = 1;
    }
    bool relaxation_feasible_cpu(const std::vector<double>& x, const std::vector<double>&,
                                 const std::vector<double>&, double tol) const {
        double mass = 0.0;
        if (RequireEquality) {
            for (double v : x) { if (v < -tol) return false; mass += v; }
            return std::isfinite(mass) && std::abs(mass - right_side) <= tol;
        }
        for (double v : x) { if (v < 0.0) return false; mass += v; }
        return mass <= right_side + tol;
    }
    gb_lmo_desc desc(const uint8_t* d_is_int) const {
        return gb_lmo_desc{GB_LMO_SIMPLEX, right_side, RequireEquality ? 1 : 0, d_is_int, 0};
    }
    std::vector<uint8_t> integer_marks(int n) const {
        std::vector<uint8_t> m(n, 0);
        for (int j = 0; j < n && j < MaxN; ++j) m[j] = is_int[j];
        return m;
    }
};

template <int MaxN>
using EqualitySimplexLMO = UnitSimplexLMO<MaxN, true>;

template <int Order>
struct BirkhoffLMO {
    static_assert(Order > 0, "Birkhoff order must be positive");
    static constexpr bool supports_dicg = true;
    static constexpr int variable_count = Order * Order;
    bool relaxation_feasible_cpu(const std::vector<double>& x, const std::vector<double>&,
                                 const std::vector<double>&, double tol) const {
        if (x.size() != static_cast<size_t>(variable_count)) return false;
        for (double v : x) if (v < -tol || v > 1 + tol) return false;
        for (int i = 0; i < Order; ++i) {
            double row = 0, col = 0;
            for (int j = 0; j < Order; ++j) { row += x[i * Order + j]; col += x[j * Order + i]; }
            if (std::abs(row - 1) > tol || std::abs(col - 1) > tol) return false;
        }
        return true;
    }
    gb_lmo_desc desc(const uint8_t*) const { return gb_lmo_desc{GB_LMO_BIRKHOFF, 1.0, 0, nullptr, Order}; }
    std::vector<uint8_t> integer_marks(int) const { return {}; }
};

// ---- objectives (dense_quadratic_objectives.cuh) ---------------------------
template <int Vars, int Threads = 64>
struct GlobalCachedDenseQuadraticObjective {
    static constexpr bool supports_quadratic_exact = true;
    static constexpr int n_vars = Vars;
    static constexpr int block_threads = Threads;
    const double* d_Qt;
    const double* d_linear;
    double grad_scale, linear_scale, constant;
    GlobalCachedDenseQuadraticObjective(const double* Q_transposed = nullptr,
                                        const double* linear = nullptr, double grad_scale_ = 1.0,
                                        double linear_scale_ = 1.0, double constant_ = 0.0)
        : d_Qt(Q_transposed), d_linear(linear), grad_scale(grad_scale_),
          linear_scale(linear_scale_), constant(constant_) {}
    gb_objective_desc desc() const {
        return gb_objective_desc{GB_OBJ_DENSE_QUADRATIC, Vars, d_Qt, d_linear, grad_scale,
                                 linear_scale, constant, 0, nullptr, nullptr};
    }
};

template <int Vars, int Threads = 256>
struct ILSGlobalCachedObjective : GlobalCachedDenseQuadraticObjective<Vars, Threads> {
    ILSGlobalCachedObjective(const double* d_AtA_transposed = nullptr, const double* d_Atb = nullptr,
                             double btb = 0.0)
        : GlobalCachedDenseQuadraticObjective<Vars, Threads>(d_AtA_transposed, d_Atb, 2.0, -1.0, btb) {}
};

// ||A P - P B||_F^2 over an Order x Order Birkhoff polytope (device A, B,
// row-major).  No reference counterpart (analogue: examples/main.cu:131-181).
template <int Order>
struct QuadraticAssignmentObjective {
    static constexpr bool supports_quadratic_exact = true;
    static constexpr int n_vars = Order * Order;
    const double* d_A;
    const double* d_B;
    gb_objective_desc desc() const {
        return gb_objective_desc{GB_OBJ_QAP, n_vars, nullptr, nullptr, 1.0, 1.0, 0.0, Order, d_A, d_B};
    }
};

namespace gpuboscia {

using RootInitialPointValidator = std::function<bool(
    const std::vector<double>&, const std::vector<double>&, const std::vector<double>&, double)>;

// heuristics.h:23-35 (default) and a floor-plus-offset variant.
struct RoundingHeuristic {
    int kind() const { return GB_HEUR_ROUND; }
    double offset() const { return 0.0; }
};
struct FloorHeuristic {
    double offset_ = 0.0;
    int kind() const { return GB_HEUR_FLOOR_OFFSET; }
    double offset() const { return offset_; }
};

struct Settings {
    DeviceResidentConfig device{};
    FWAlgorithmKind algorithm = FWAlgorithmKind::DICG;
    LineSearchOptions line_search{};
    BranchStrategy branch;
    std::function<void(int, int, double, double, bool)> progress;
    double time_limit_seconds = std::numeric_limits<double>::infinity();
    double natural_lower_bound = -std::numeric_limits<double>::infinity();
    std::vector<double> root_initial_point;
    RootInitialPointValidator root_initial_point_feasible;
    // Additions (no reference counterpart): the relaxation form, its
    // arithmetic, the branching rule of the search and the autotuner.
    int form = 0;                       // 0 = plain, 1 = batched FP64, 2 = mixed precision
    int precision = GB_PREC_FP64;       // form 2: GB_PREC_BF16 (1 bf16 plane) or FP32 (3 planes)
    int branch_rule = GB_BRANCH_POW2;   // used when `branch` is empty
    bool autotune = false;              // form 2: tune tile shape and FP64-pass budget
};

struct Result {
    double objective = std::numeric_limits<double>::infinity();
    std::vector<double> x;
    double dual_bound = -std::numeric_limits<double>::infinity();
    int nodes_processed = 0;
    double elapsed_seconds = 0.0;
    bool reached_tolerance = false;
    bool reached_time_limit = false;
};

inline Settings create_default_settings() { return Settings{}; }

namespace detail {

inline void throw_status(gb_status st) {
    if (st == GB_OK) return;
    const std::string msg = gb_last_error();
    if (st == GB_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

struct Callbacks {
    const Settings* s;
    const Problem* p;
};

inline void progress_tramp(void* u, int iter, int open, double bound, double inc, int new_inc) {
    auto* c = static_cast<Callbacks*>(u);
    c->s->progress(iter, open, bound, inc, new_inc != 0);
}

inline int validator_tramp(void* u, const double* x, const double* lb, const double* ub, int n,
                           double tol) {
    auto* c = static_cast<Callbacks*>(u);
    return c->s->root_initial_point_feasible(std::vector<double>(x, x + n),
                                             std::vector<double>(lb, lb + n),
                                             std::vector<double>(ub, ub + n), tol) ? 1 : 0;
}

inline int branch_tramp(void* u, const double* lb, const double* ub, const double* x, int n,
                        int depth, int node_id, double* clb, double* cub, int max_children) {
    auto* c = static_cast<Callbacks*>(u);
    Node parent(node_id, -1, depth, std::vector<double>(lb, lb + n), std::vector<double>(ub, ub + n));
    int next_id = 0;
    const auto children = c->s->branch(parent, next_id, std::vector<double>(x, x + n), c->p->int_vars);
    const int cnt = std::min<int>((int)children.size(), max_children);
    for (int i = 0; i < cnt; ++i) {
        std::copy(children[i].lb.begin(), children[i].lb.end(), clb + (size_t)i * n);
        std::copy(children[i].ub.begin(), children[i].ub.end(), cub + (size_t)i * n);
    }
    return cnt;
}

}  // namespace detail

// gpuboscia.cuh:108-170.  Objective and LMO are the built-in kinds above.
template <typename DeviceObjective, typename LMO, typename Heuristic = RoundingHeuristic>
Result solve(const Problem& problem, DeviceObjective objective, LMO lmo,
             const Settings& settings = create_default_settings(), Heuristic heuristic = {}) {
    const bool dicg = settings.algorithm == FWAlgorithmKind::DICG;
    if (settings.line_search.kind == LineSearchKind::QuadraticExact
            && !DeviceObjective::supports_quadratic_exact)
        throw std::invalid_argument("quadratic-exact requires an objective with quadratic curvature support");
    if (dicg && !LMO::supports_dicg)
        throw std::invalid_argument("the selected LMO does not support DICG; use Frank-Wolfe");
    // Root-point rules (tree.h:115-147) are applied by gb_solve with the same
    // messages; the LMO's relaxation_feasible_cpu has a C-ABI twin there.
    gb_settings s;
    gb_default_settings(&s);
    const DeviceResidentConfig& d = settings.device;
    s.gpu_batch_size = d.gpu_batch_size;
    s.num_streams = d.num_streams;
    s.gpu_block_threads = d.gpu_block_threads;
    s.gpu_max_iters = d.gpu_max_iters;
    s.gpu_eps = d.gpu_eps;
    s.gpu_eps_decay = d.gpu_eps_decay;
    s.gpu_eps_min = d.gpu_eps_min;
    s.integer_tol = d.integer_tol;
    s.abs_gap_tol = d.abs_gap_tol;
    s.rel_gap_tol = d.rel_gap_tol;
    s.algorithm.algorithm = dicg ? GB_ALG_DICG : GB_ALG_FRANK_WOLFE;
    s.algorithm.line_search = static_cast<int>(settings.line_search.kind);
    s.algorithm.secant_iters = settings.line_search.secant_iters;
    s.algorithm.secant_tol = settings.line_search.secant_tol;
    s.algorithm.secant_denom_eps = settings.line_search.secant_denom_eps;
    s.algorithm.heuristic = heuristic.kind();
    s.algorithm.heuristic_offset = heuristic.offset();
    s.time_limit_seconds = settings.time_limit_seconds;
    s.natural_lower_bound = settings.natural_lower_bound;
    s.form = settings.form;
    s.precision = settings.precision;
    s.branch_rule = settings.branch_rule;
    s.autotune = settings.autotune ? 1 : 0;
    detail::Callbacks cb{&settings, &problem};
    if (!settings.root_initial_point.empty()) {
        s.root_initial_point = settings.root_initial_point.data();
        s.root_initial_point_size = (int)settings.root_initial_point.size();
    }
    if (settings.root_initial_point_feasible) {
        s.root_validator = detail::validator_tramp;
        s.root_validator_user = &cb;
    }
    if (settings.progress) {
        s.progress = detail::progress_tramp;
        s.progress_user = &cb;
    }
    if (settings.branch) {
        s.branch_rule = GB_BRANCH_CUSTOM;
        s.branch = detail::branch_tramp;
        s.branch_user = &cb;
    }
    // Simplex integer marks live in the LMO functor (lmo.cuh:352); the
    // kernels read them from device memory for the duration of the solve.
    uint8_t* d_marks = nullptr;
    const std::vector<uint8_t> marks = lmo.integer_marks(problem.n_vars);
    if (!marks.empty()) {
        if (hipMalloc(reinterpret_cast<void**>(&d_marks), marks.size()) != hipSuccess
                || hipMemcpy(d_marks, marks.data(), marks.size(), hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error("HIP allocation of LMO integer marks failed");
    }
    const gb_lmo_desc ld = lmo.desc(d_marks);
    const gb_objective_desc od = objective.desc();
    gb_problem gp{problem.n_vars, problem.lb.data(), problem.ub.data(), problem.int_vars.data(),
                  (int)problem.int_vars.size()};
    Result r;
    std::vector<double> x(problem.n_vars > 0 ? problem.n_vars : 0);
    gb_result gr{};
    gr.x = x.data();
    const gb_status st = gb_solve(&gp, &od, &ld, &s, &gr);
    if (d_marks) (void)hipFree(d_marks);
    detail::throw_status(st);
    r.objective = gr.objective;
    if (std::isfinite(gr.objective)) r.x = x;
    r.dual_bound = gr.dual_bound;
    r.nodes_processed = gr.nodes_processed;
    r.elapsed_seconds = gr.elapsed_seconds;
    r.reached_tolerance = gr.reached_tolerance != 0;
    r.reached_time_limit = gr.reached_time_limit != 0;
    return r;
}

}  // namespace gpuboscia
