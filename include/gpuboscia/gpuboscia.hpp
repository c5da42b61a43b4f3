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
//   ILSGlobalCachedObjective<Vars,Threads>, the algorithm policies
//   (FrankWolfe<LS>, DecompositionInvariantFrankWolfe<LS> and their aliases,
//   algorithms.cuh:57-362), top_k_infeasible / most_infeasible / branch_pow2
//   (branching.h:24-118), DeviceResidentHybridSolver with push, solve, the
//   accessors and the scheduling observer (device_resident_solver.cuh:
//   1027-1038, 1184-1298), gpuboscia::Settings, gpuboscia::Result,
//   gpuboscia::create_default_settings, gpuboscia::solve_with_algorithm,
//   gpuboscia::solve, gpuboscia::RoundingHeuristic.
// Objectives hold caller-owned device pointers exactly as in the reference;
// the built-in kinds (and QuadraticAssignmentObjective) run the library's
// gfx950 kernels.  Any other objective, LMO or heuristic type is a device
// plug-in, as in the reference: its methods are compiled into the library's
// relaxation kernel template (gpuboscia/device/relax_fp64.hpp) in the
// caller's translation unit, which must then be built with hipcc for gfx950.
// Host-side callbacks (progress, root validator, BranchStrategy) are plain
// std::function values as in the reference and run on the calling thread.
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <exception>
#include <functional>
#include <limits>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../gpuboscia_amd.h"
#if defined(__HIPCC__)
#include "device/relax_fp64.hpp"
#endif

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

// ---- branching.h:24-118 --------------------------------------------------------
// The integer variable whose fractional part is closest to 1/2 (first on ties).
inline int most_infeasible(const std::vector<double>& x, const std::vector<int>& int_vars) {
    int best = -1;
    double best_dist = -1.0;
    for (int i : int_vars) {
        const double frac = x[i] - std::floor(x[i]);
        const double dist = std::min(frac, 1.0 - frac);
        if (dist > best_dist) {
            best_dist = dist;
            best = i;
        }
    }
    return best;
}

// Up to k fractional integer variables (distance to the nearest integer
// above integer_tol), most fractional first.  The candidates are sorted by
// the same std::sort call on the same sequence as the reference, so equal
// distances come out in the same order.
inline std::vector<int> top_k_infeasible(const std::vector<double>& x, const std::vector<int>& int_vars, int k,
                                         double integer_tol = 1e-6) {
    struct Cand {
        int idx;
        double dist;
    };
    std::vector<Cand> cands;
    cands.reserve(int_vars.size());
    for (int i : int_vars) {
        const double frac = x[i] - std::floor(x[i]);
        const double dist = std::min(frac, 1.0 - frac);
        if (dist > integer_tol) cands.push_back({i, dist});
    }
    std::sort(cands.begin(), cands.end(), [](const Cand& a, const Cand& b) { return a.dist > b.dist; });
    std::vector<int> out;
    for (int i = 0; i < k && i < (int)cands.size(); ++i) out.push_back(cands[i].idx);
    return out;
}

// 2^k children on the k most fractional variables: bit b of the child's
// index set means lb = ceil(x), clear means ub = floor(x).  Children take
// ids next_id, next_id + 1, ..., the parent's bound and depth + 1, and start
// fresh.
inline std::vector<Node> branch_pow2(const Node& parent, int& next_id, const std::vector<double>& x,
                                     const std::vector<int>& int_vars, int k = 3, double integer_tol = 1e-6) {
    const std::vector<int> vars = top_k_infeasible(x, int_vars, k, integer_tol);
    std::vector<Node> kids;
    if (vars.empty()) return kids;
    const int count = 1 << (int)vars.size();
    kids.reserve(count);
    for (int mask = 0; mask < count; ++mask) {
        Node c = parent;
        c.id = next_id++;
        c.parent_id = parent.id;
        c.depth = parent.depth + 1;
        c.dual_bound = parent.dual_bound;
        c.status = NodeStatus::Active;
        c.relaxation_x.clear();
        c.relaxation_iterations = 0;
        for (size_t b = 0; b < vars.size(); ++b) {
            const int v = vars[b];
            if (mask & (1 << b)) c.lb[v] = std::ceil(x[v]);
            else c.ub[v] = std::floor(x[v]);
        }
        kids.push_back(std::move(c));
    }
    return kids;
}

// ---- algorithm policies (algorithms.cuh:57-362) ---------------------------------
// Host-side tags: each maps to the library's algorithm descriptor (FW or
// DICG, and its step rule); the step rules themselves run in the kernels.
struct OpenLoopLS {
    static constexpr int kind = GB_LS_OPEN_LOOP;
};
struct ScaledOpenLoopLS {
    static constexpr int kind = GB_LS_SCALED_OPEN_LOOP;
};
struct SafeguardedScaledOpenLoopLS {
    static constexpr int kind = GB_LS_SAFEGUARDED_SCALED_OPEN_LOOP;
};
struct SecantLS {
    static constexpr int kind = GB_LS_SECANT;
    int max_iters;
    double tolerance;
    double denom_eps;
    SecantLS(int max_iters_ = 8, double tolerance_ = 1e-10, double denom_eps_ = 1e-14)
        : max_iters(max_iters_), tolerance(tolerance_), denom_eps(denom_eps_) {}
};
struct QuadraticExactLS {
    static constexpr int kind = GB_LS_QUADRATIC_EXACT;
};

namespace gpuboscia_detail {
template <class LS>
inline void fill_line_search(const LS&, gb_algorithm_desc& a) {
    a.line_search = LS::kind;
}
inline void fill_line_search(const SecantLS& ls, gb_algorithm_desc& a) {
    a.line_search = GB_LS_SECANT;
    a.secant_iters = ls.max_iters;
    a.secant_tol = ls.tolerance;
    a.secant_denom_eps = ls.denom_eps;
}
}  // namespace gpuboscia_detail

template <typename LS = OpenLoopLS>
struct FrankWolfe {
    static constexpr bool uses_dicg = false;
    LS ls{};
    void fill(gb_algorithm_desc& a) const {
        a.algorithm = GB_ALG_FRANK_WOLFE;
        gpuboscia_detail::fill_line_search(ls, a);
    }
};

template <typename LS = OpenLoopLS>
struct DecompositionInvariantFrankWolfe {
    static constexpr bool uses_dicg = true;
    LS ls{};
    void fill(gb_algorithm_desc& a) const {
        a.algorithm = GB_ALG_DICG;
        gpuboscia_detail::fill_line_search(ls, a);
    }
};

using FrankWolfeOpenLoop = FrankWolfe<OpenLoopLS>;
using FrankWolfeSecant = FrankWolfe<SecantLS>;
using FrankWolfePolicy = FrankWolfeOpenLoop;
using DICGOpenLoop = DecompositionInvariantFrankWolfe<OpenLoopLS>;
using DICGScaledOpenLoop = DecompositionInvariantFrankWolfe<ScaledOpenLoopLS>;
using DICGSafeguardedScaledOpenLoop = DecompositionInvariantFrankWolfe<SafeguardedScaledOpenLoopLS>;
using DICGSecant = DecompositionInvariantFrankWolfe<SecantLS>;
using DICGPolicy = DICGOpenLoop;
using FrankWolfeQuadraticExact = FrankWolfe<QuadraticExactLS>;
using DICGQuadraticExact = DecompositionInvariantFrankWolfe<QuadraticExactLS>;

enum class FWExecutionMode { SingleBlock, MultiBlock };

// device_resident_solver.cuh:1004-1025 (same fields and defaults).
// execution_mode = MultiBlock relaxes every node with blocks_per_node
// workgroups (dense quadratics over a box or a simplex); the paired-DICG
// fields are accepted and validated, and the gfx950 kernels choose their
// own launch shapes.
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
    static constexpr int device_kind = GB_LMO_BOX;
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
    static constexpr int device_kind = GB_LMO_SIMPLEX;
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
    static constexpr int device_kind = GB_LMO_BIRKHOFF;
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
    static constexpr int device_kind = GB_OBJ_DENSE_QUADRATIC;
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
    static constexpr int device_kind = GB_OBJ_QAP;
    static constexpr int n_vars = Order * Order;
    const double* d_A;
    const double* d_B;
    gb_objective_desc desc() const {
        return gb_objective_desc{GB_OBJ_QAP, n_vars, nullptr, nullptr, 1.0, 1.0, 0.0, Order, d_A, d_B};
    }
};

// The same objective and the Birkhoff LMO with the order chosen at run
// time (instances read from files); no reference counterpart, whose types
// fix it at compile time.
struct RuntimeQuadraticAssignmentObjective {
    static constexpr bool supports_quadratic_exact = true;
    static constexpr int device_kind = GB_OBJ_QAP;
    int order;
    const double* d_A;
    const double* d_B;
    gb_objective_desc desc() const {
        return gb_objective_desc{GB_OBJ_QAP, order * order, nullptr, nullptr, 1.0, 1.0, 0.0, order, d_A, d_B};
    }
};
struct RuntimeBirkhoffLMO {
    static constexpr bool supports_dicg = true;
    static constexpr int device_kind = GB_LMO_BIRKHOFF;
    int order;
    bool relaxation_feasible_cpu(const std::vector<double>& x, const std::vector<double>&,
                                 const std::vector<double>&, double tol) const {
        if (x.size() != static_cast<size_t>(order) * order) return false;
        for (double v : x) if (v < -tol || v > 1 + tol) return false;
        for (int i = 0; i < order; ++i) {
            double row = 0, col = 0;
            for (int j = 0; j < order; ++j) { row += x[i * order + j]; col += x[j * order + i]; }
            if (std::abs(row - 1) > tol || std::abs(col - 1) > tol) return false;
        }
        return true;
    }
    gb_lmo_desc desc(const uint8_t*) const { return gb_lmo_desc{GB_LMO_BIRKHOFF, 1.0, 0, nullptr, order}; }
    std::vector<uint8_t> integer_marks(int) const { return {}; }
};

namespace gpuboscia {

// tree.h:104-113, gpuboscia.cuh:49-54, algorithms.cuh:265-269
template <class LMO, class = void>
struct HasRelaxationPointCheck : std::false_type {};
template <class LMO>
struct HasRelaxationPointCheck<LMO, std::void_t<decltype(std::declval<const LMO&>().relaxation_feasible_cpu(
                                        std::declval<const std::vector<double>&>(),
                                        std::declval<const std::vector<double>&>(),
                                        std::declval<const std::vector<double>&>(), 0.0))>> : std::true_type {};
template <class LMO, class = void>
struct lmo_supports_dicg : std::false_type {};
template <class LMO>
struct lmo_supports_dicg<LMO, std::void_t<decltype(LMO::supports_dicg)>>
    : std::integral_constant<bool, LMO::supports_dicg> {};
template <class Objective, class = void>
struct SupportsQuadraticExact : std::false_type {};
template <class Objective>
struct SupportsQuadraticExact<Objective, std::void_t<decltype(Objective::supports_quadratic_exact)>>
    : std::integral_constant<bool, Objective::supports_quadratic_exact> {};

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

namespace detail {

inline void throw_status(gb_status st) {
    if (st == GB_OK) return;
    const std::string msg = gb_last_error();
    if (st == GB_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

// Host callbacks behind the C-ABI: the user's std::function values, the
// problem's integer variables, and the first exception a callback threw
// (rethrown once the solve returns, as the reference's solve() would let it
// propagate).
struct Callbacks {
    BranchStrategy branch;
    std::function<void(int, int, double, double, bool)> progress;
    RootInitialPointValidator validator;
    std::function<void(const gb_scheduling_event&)> scheduling;
    const std::vector<int>* int_vars = nullptr;
    std::exception_ptr error;
    void rethrow() {
        if (error) {
            std::exception_ptr e = error;
            error = nullptr;
            std::rethrow_exception(e);
        }
    }
};

inline void progress_tramp(void* u, int iter, int open, double bound, double inc, int new_inc) {
    auto* c = static_cast<Callbacks*>(u);
    if (c->error) return;
    try {
        c->progress(iter, open, bound, inc, new_inc != 0);
    } catch (...) {
        c->error = std::current_exception();
    }
}

inline void scheduling_tramp(void* u, const gb_scheduling_event* ev) {
    auto* c = static_cast<Callbacks*>(u);
    if (c->error) return;
    try {
        c->scheduling(*ev);
    } catch (...) {
        c->error = std::current_exception();
    }
}

inline int validator_tramp(void* u, const double* x, const double* lb, const double* ub, int n, double tol) {
    auto* c = static_cast<Callbacks*>(u);
    try {
        return c->validator(std::vector<double>(x, x + n), std::vector<double>(lb, lb + n),
                            std::vector<double>(ub, ub + n), tol) ? 1 : 0;
    } catch (...) {
        c->error = std::current_exception();
        return 0;
    }
}

// BranchStrategy behind gb_branch_fn: the parent as a Node (with its
// relaxation_x, relaxation_iterations and certified dual bound), next_id in
// and out, and every returned child handed back with its id, parent id,
// depth, bound and box.
inline int branch_tramp(void* u, const gb_node* p, int* next_id, const double* x, int n, const int*, int,
                        gb_children* out) {
    auto* c = static_cast<Callbacks*>(u);
    try {
        Node parent(p->id, p->parent_id, p->depth, std::vector<double>(p->lb, p->lb + n),
                    std::vector<double>(p->ub, p->ub + n));
        if (p->relaxation_x) parent.relaxation_x.assign(p->relaxation_x, p->relaxation_x + n);
        parent.relaxation_iterations = p->relaxation_iterations;
        parent.dual_bound = p->dual_bound;
        int nid = *next_id;
        const std::vector<Node> kids = c->branch(parent, nid, std::vector<double>(x, x + n), *c->int_vars);
        *next_id = nid;
        for (const Node& k : kids) {
            if ((int)k.lb.size() != n || (int)k.ub.size() != n)
                throw std::invalid_argument("invalid node buffer dimensions or resume state");
            const gb_node cn{k.id, k.parent_id, k.depth, k.lb.data(), k.ub.data(), nullptr, 0, k.dual_bound};
            throw_status(gb_children_push(out, &cn));
        }
        return 0;
    } catch (...) {
        c->error = std::current_exception();
        return 1;
    }
}

// Simplex integer marks live in the LMO functor (lmo.cuh:352); the kernels
// read them from device memory for as long as the descriptor is used.
struct DeviceMarks {
    uint8_t* ptr = nullptr;
    explicit DeviceMarks(const std::vector<uint8_t>& marks) {
        if (marks.empty()) return;
        if (hipMalloc(reinterpret_cast<void**>(&ptr), marks.size()) != hipSuccess
                || hipMemcpy(ptr, marks.data(), marks.size(), hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error("HIP allocation of LMO integer marks failed");
    }
    ~DeviceMarks() {
        if (ptr) (void)hipFree(ptr);
    }
    DeviceMarks(const DeviceMarks&) = delete;
    DeviceMarks& operator=(const DeviceMarks&) = delete;
};

// ---- device plug-ins (lmo.cuh:11-66, heuristics.h:11-35, README.md:224-255) --
// The built-in types describe themselves to the library (desc(), kind());
// any other objective, LMO or heuristic type is a device plug-in: its
// methods are compiled into the library's relaxation kernel template in the
// caller's translation unit (hipcc, gfx950) and launched through
// gb_plugin_desc.  Plug-in types are copied into the kernel by value, as in
// the reference, so they hold scalars, fixed arrays and device pointers.
template <class T, class = void>
struct has_objective_desc : std::false_type {};
template <class T>
struct has_objective_desc<T, std::void_t<decltype(std::declval<const T&>().desc())>> : std::true_type {};
template <class T, class = void>
struct has_lmo_desc : std::false_type {};
template <class T>
struct has_lmo_desc<T, std::void_t<decltype(std::declval<const T&>().desc((const uint8_t*)nullptr))>>
    : std::true_type {};
template <class T, class = void>
struct has_heuristic_kind : std::false_type {};
template <class T>
struct has_heuristic_kind<T, std::void_t<decltype(std::declval<const T&>().kind())>> : std::true_type {};
template <class T, class = void>
struct shared_bytes_of : std::integral_constant<size_t, 0> {};
template <class T>
struct shared_bytes_of<T, std::void_t<decltype(T::shared_bytes)>> : std::integral_constant<size_t, (size_t)T::shared_bytes> {};
template <class T, class = void>
struct block_threads_of : std::integral_constant<int, 0> {};
template <class T>
struct block_threads_of<T, std::void_t<decltype(T::block_threads)>> : std::integral_constant<int, (int)T::block_threads> {};
template <class T, class = void>
struct n_vars_of : std::integral_constant<int, 0> {};
template <class T>
struct n_vars_of<T, std::void_t<decltype(T::n_vars)>> : std::integral_constant<int, (int)T::n_vars> {};

#if defined(__HIPCC__)
template <class O>
auto device_objective(const O& o) {
    if constexpr (has_objective_desc<O>::value) {
        if constexpr (O::device_kind == GB_OBJ_DENSE_QUADRATIC) return gb::Obj<GB_OBJ_DENSE_QUADRATIC>{};
        else return gb::Obj<GB_OBJ_QAP>{};
    } else {
        return gb::PluginObjective<O>{o};
    }
}
template <class L>
auto device_lmo(const L& l) {
    if constexpr (has_lmo_desc<L>::value) {
        if constexpr (L::device_kind == GB_LMO_BOX) return gb::Lmo<GB_LMO_BOX>{};
        else if constexpr (L::device_kind == GB_LMO_SIMPLEX) return gb::Lmo<GB_LMO_SIMPLEX>{};
        else return gb::Lmo<GB_LMO_BIRKHOFF>{};
    } else {
        return gb::PluginLmo<L>{l};
    }
}
template <class H>
auto device_heuristic(const H& h) {
    if constexpr (has_heuristic_kind<H>::value) return gb::BuiltinHeuristic{};
    else return gb::PluginHeuristic<H>{h};
}
#endif

// Descriptors of one (objective, LMO, heuristic) triple for the C-ABI, with
// the plug-in launcher when any of them is not built in.  Holds copies of
// the objects for the length of a solve.
template <class O, class L, class H>
class Binding {
  public:
    static constexpr bool builtin =
        has_objective_desc<O>::value && has_lmo_desc<L>::value && has_heuristic_kind<H>::value;
#if !defined(__HIPCC__)
    static_assert(builtin, "custom device objective, LMO or heuristic types are compiled into the relaxation "
                           "kernel: build this translation unit with hipcc --offload-arch=gfx950");
#endif
    Binding(const O& o, const L& l, const H& h, int n)
        : o_(o), l_(l), h_(h), marks_(marks_of(l, n)) {
        if constexpr (has_objective_desc<O>::value) {
            od_ = o_.desc();
        } else {
            od_ = gb_objective_desc{GB_OBJ_PLUGIN, n_vars_of<O>::value > 0 ? n_vars_of<O>::value : n,
                                    nullptr, nullptr, 1.0, 1.0, 0.0, 0, nullptr, nullptr};
        }
        if constexpr (has_lmo_desc<L>::value) ld_ = l_.desc(marks_.ptr);
        else ld_ = gb_lmo_desc{GB_LMO_PLUGIN, 1.0, 0, nullptr, 0, 0.0, 0.0};
#if defined(__HIPCC__)
        if constexpr (!builtin) {
            pd_.launch = &Binding::launch;
            pd_.ctx = this;
            pd_.objective = has_objective_desc<O>::value ? 0 : 1;
            pd_.lmo = has_lmo_desc<L>::value ? 0 : 1;
            pd_.heuristic = has_heuristic_kind<H>::value ? 0 : 1;
            pd_.supports_dicg = gb::plugin_traits::has_inface<L>::value ? 1 : 0;
            pd_.supports_quadratic_exact = gb::plugin_traits::has_curvature<O>::value ? 1 : 0;
            pd_.block_threads = has_objective_desc<O>::value ? 0 : block_threads_of<O>::value;
            pd_.objective_bytes = has_objective_desc<O>::value ? 0 : shared_bytes_of<O>::value;
            pd_.relaxation_feasible = HasRelaxationPointCheck<L>::value ? &Binding::feasible : nullptr;
        }
#endif
    }
    Binding(const Binding&) = delete;
    Binding& operator=(const Binding&) = delete;
    const gb_objective_desc* objective() const { return &od_; }
    const gb_lmo_desc* lmo() const { return &ld_; }
    const gb_plugin_desc* plugin() const { return builtin ? nullptr : &pd_; }
    void fill_heuristic(gb_algorithm_desc& a) const {
        if constexpr (has_heuristic_kind<H>::value) {
            a.heuristic = h_.kind();
            a.heuristic_offset = h_.offset();
        } else {
            a.heuristic = GB_HEUR_PLUGIN;
            a.heuristic_offset = 0.0;
        }
    }

  private:
    static std::vector<uint8_t> marks_of(const L& l, int n) {
        if constexpr (has_lmo_desc<L>::value) return l.integer_marks(n);
        else return {};
    }
#if defined(__HIPCC__)
    static int launch(void* ctx, const void* pl, void* stream) {
        const Binding* b = static_cast<const Binding*>(ctx);
        return gb::launch_relax_plugin(*static_cast<const gb::PluginLaunch*>(pl), device_objective(b->o_),
                                       device_lmo(b->l_), device_heuristic(b->h_), stream);
    }
    // tree.h:106-113: the LMO's relaxation_feasible_cpu, when it has one
    static int feasible(void* ctx, const double* x, const double* lb, const double* ub, int n, double tol) {
        const Binding* b = static_cast<const Binding*>(ctx);
        if constexpr (HasRelaxationPointCheck<L>::value) {
            return b->l_.relaxation_feasible_cpu(std::vector<double>(x, x + n), std::vector<double>(lb, lb + n),
                                                 std::vector<double>(ub, ub + n), tol)
                       ? 1
                       : 0;
        } else {
            return 0;
        }
    }
#endif
    O o_;
    L l_;
    H h_;
    DeviceMarks marks_;
    gb_objective_desc od_{};
    gb_lmo_desc ld_{};
    gb_plugin_desc pd_{};
};

// MultiBlock (device_resident_solver.cuh:1395-1440): blocks_per_node
// workgroups per node on one stream; the library launches a batch in as many
// rounds of resident node groups as it needs, so gpu_batch_size is kept.
inline void fill_config(const DeviceResidentConfig& d, gb_settings& s) {
    if (d.execution_mode != FWExecutionMode::SingleBlock && d.execution_mode != FWExecutionMode::MultiBlock)
        throw std::invalid_argument("unknown FW execution mode");
    if (d.execution_mode == FWExecutionMode::MultiBlock) {
        if (d.blocks_per_node < 1) throw std::invalid_argument("MultiBlock blocks_per_node must be positive");
        if (d.dicg_oracle_parallelism == DeviceResidentConfig::DICGOracleParallelism::Paired)
            throw std::invalid_argument("MultiBlock cannot combine with paired DICG");
    }
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
    if (d.execution_mode == FWExecutionMode::MultiBlock) {
        s.blocks_per_node = d.blocks_per_node;
        s.num_streams = 1;
    }
}

}  // namespace detail
}  // namespace gpuboscia

// ---- device_resident_solver.cuh:1027-1038 ----------------------------------------
enum class DRSSchedulingEventType { Dispatched, CompletionDetached, PostprocessStarted };

struct DRSSchedulingEvent {
    DRSSchedulingEventType type;
    int slot_id;
    std::uint64_t generation;
    int n_nodes;
};

// ---- DeviceResidentHybridSolver (device_resident_solver.cuh:1184-1298) --------------
// The low-level solver: nodes are pushed by the caller and relaxed with the
// given algorithm policy; a BranchStrategy decides the children.  Over the
// library's gb_solver object.  The Problem is held by reference, as in the
// reference, and must outlive the solver.
template <typename DeviceObjective, typename Algorithm, typename LMO = BoxLMO,
          typename Heuristic = gpuboscia::RoundingHeuristic>
class DeviceResidentHybridSolver {
  public:
    DeviceResidentHybridSolver(const Problem& prob, DeviceObjective obj, Algorithm alg, BranchStrategy branch_sel,
                               DeviceResidentConfig cfg = {}, LMO lmo = {},
                               double time_limit_seconds = std::numeric_limits<double>::infinity(),
                               Heuristic heuristic = {})
        : prob_(prob), bind_(obj, lmo, heuristic, prob.n_vars) {
        if (!std::isfinite(cfg.dicg_paired_occupancy_fraction) || cfg.dicg_paired_occupancy_fraction <= 0.0
                || cfg.dicg_paired_occupancy_fraction > 1.0)
            throw std::invalid_argument("invalid tolerance configuration");
        gb_default_settings(&s_);
        gpuboscia::detail::fill_config(cfg, s_);
        alg.fill(s_.algorithm);
        bind_.fill_heuristic(s_.algorithm);
        s_.plugin = bind_.plugin();
        s_.time_limit_seconds = time_limit_seconds;
        cb_.int_vars = &prob_.int_vars;
        if (branch_sel) {
            cb_.branch = std::move(branch_sel);
            s_.branch_rule = GB_BRANCH_CUSTOM;
            s_.branch = gpuboscia::detail::branch_tramp;
            s_.branch_user = &cb_;
        }
        const gb_problem gp{prob_.n_vars, prob_.lb.data(), prob_.ub.data(), prob_.int_vars.data(),
                            (int)prob_.int_vars.size()};
        gpuboscia::detail::throw_status(gb_solver_create(&gp, bind_.objective(), bind_.lmo(), &s_, &h_));
    }
    ~DeviceResidentHybridSolver() { (void)gb_solver_destroy(h_); }
    DeviceResidentHybridSolver(const DeviceResidentHybridSolver&) = delete;
    DeviceResidentHybridSolver& operator=(const DeviceResidentHybridSolver&) = delete;

    void push(const Node& nd) {
        if (nd.lb.size() != (size_t)prob_.n_vars || nd.ub.size() != nd.lb.size()
                || (!nd.relaxation_x.empty() && nd.relaxation_x.size() != nd.lb.size()))
            throw std::invalid_argument("invalid node buffer dimensions or resume state");
        const gb_node gn{nd.id, nd.parent_id, nd.depth, nd.lb.data(), nd.ub.data(),
                         nd.relaxation_x.empty() ? nullptr : nd.relaxation_x.data(), nd.relaxation_iterations,
                         nd.dual_bound};
        gpuboscia::detail::throw_status(gb_solver_push(h_, &gn));
    }
    void set_progress_callback(std::function<void(int, int, double, double, bool)> cb) {
        cb_.progress = std::move(cb);
        if (cb_.progress)
            gpuboscia::detail::throw_status(
                gb_solver_set_progress_callback(h_, gpuboscia::detail::progress_tramp, &cb_));
        else
            gpuboscia::detail::throw_status(gb_solver_set_progress_callback(h_, nullptr, nullptr));
    }
    // Invoked on the solving thread; must not call back into the solver.
    void set_scheduling_callback(std::function<void(const DRSSchedulingEvent&)> cb) {
        if (cb) {
            auto f = std::move(cb);
            cb_.scheduling = [f](const gb_scheduling_event& e) {
                f(DRSSchedulingEvent{static_cast<DRSSchedulingEventType>(e.type), e.slot_id,
                                     (std::uint64_t)e.generation, e.n_nodes});
            };
            gpuboscia::detail::throw_status(
                gb_solver_set_scheduling_callback(h_, gpuboscia::detail::scheduling_tramp, &cb_));
        } else {
            cb_.scheduling = nullptr;
            gpuboscia::detail::throw_status(gb_solver_set_scheduling_callback(h_, nullptr, nullptr));
        }
    }
    void solve() {
        const gb_status st = gb_solver_solve(h_);
        cb_.rethrow();
        gpuboscia::detail::throw_status(st);
        gpuboscia::detail::throw_status(gb_solver_get_stats(h_, &stats_));
        std::vector<double> x((size_t)prob_.n_vars);
        int has = 0;
        gpuboscia::detail::throw_status(gb_solver_incumbent_x(h_, x.data(), &has));
        inc_x_ = has ? x : std::vector<double>{};
    }

    double incumbent_value() const { return stats_.incumbent_value; }
    std::vector<double> incumbent_x() const { return inc_x_; }
    int nodes_processed() const { return stats_.nodes_processed; }
    double dual_bound() const { return stats_.dual_bound; }
    int last_gpu_iterations() const { return stats_.last_gpu_iterations; }
    double last_gpu_gap() const { return stats_.last_gpu_gap; }
    bool last_gpu_converged() const { return stats_.last_gpu_converged != 0; }
    bool reached_tolerance() const { return stats_.reached_tolerance != 0; }
    bool reached_time_limit() const { return stats_.reached_time_limit != 0; }
    int effective_gpu_batch_size() const { return s_.gpu_batch_size; }
    int effective_num_streams() const { return s_.num_streams; }
    int paired_gpu_batches() const { return 0; }
    // Additions (no reference counterpart).
    long long fw_iterations() const { return stats_.fw_iterations; }
    double elapsed_seconds() const { return stats_.elapsed_seconds; }

  private:
    const Problem& prob_;
    gpuboscia::detail::Binding<DeviceObjective, LMO, Heuristic> bind_;
    gb_settings s_{};
    gpuboscia::detail::Callbacks cb_;
    gb_solver* h_ = nullptr;
    gb_solver_stats stats_{std::numeric_limits<double>::infinity(), -std::numeric_limits<double>::infinity()};
    std::vector<double> inc_x_;
};

namespace gpuboscia {

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
    bool device_frontier = false;       // forms 1, 2: frontier, pruning and branching on the GPU
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

// gpuboscia.cuh:56-104: one solve with an explicit algorithm policy.  The
// root node carries the natural lower bound and the root initial point
// (validated by tree.h:115-147's rules inside gb_solve); without a
// BranchStrategy the built-in branch_pow2(k = 3) rule runs in the library.
template <typename DeviceObjective, typename Algorithm, typename LMO, typename Heuristic = RoundingHeuristic>
Result solve_with_algorithm(const Problem& problem, DeviceObjective objective, LMO lmo, const Settings& settings,
                            Algorithm algorithm, Heuristic heuristic = {}) {
    gb_settings s;
    gb_default_settings(&s);
    detail::fill_config(settings.device, s);
    algorithm.fill(s.algorithm);
    detail::Binding<DeviceObjective, LMO, Heuristic> bind(objective, lmo, heuristic, problem.n_vars);
    bind.fill_heuristic(s.algorithm);
    s.plugin = bind.plugin();
    s.time_limit_seconds = settings.time_limit_seconds;
    s.natural_lower_bound = settings.natural_lower_bound;
    if (settings.form != 0 && settings.device.execution_mode == FWExecutionMode::MultiBlock)
        throw std::invalid_argument("MultiBlock relaxes nodes in the plain form (form = 0)");
    s.form = settings.form;
    s.precision = settings.precision;
    s.branch_rule = settings.branch_rule;
    s.autotune = settings.autotune ? 1 : 0;
    s.device_frontier = settings.device_frontier ? 1 : 0;
    detail::Callbacks cb;
    cb.int_vars = &problem.int_vars;
    if (!settings.root_initial_point.empty()) {
        s.root_initial_point = settings.root_initial_point.data();
        s.root_initial_point_size = (int)settings.root_initial_point.size();
    }
    if (settings.root_initial_point_feasible) {
        cb.validator = settings.root_initial_point_feasible;
        s.root_validator = detail::validator_tramp;
        s.root_validator_user = &cb;
    }
    if (settings.progress) {
        cb.progress = settings.progress;
        s.progress = detail::progress_tramp;
        s.progress_user = &cb;
    }
    if (settings.branch) {
        cb.branch = settings.branch;
        s.branch_rule = GB_BRANCH_CUSTOM;
        s.branch = detail::branch_tramp;
        s.branch_user = &cb;
    }
    const gb_problem gp{problem.n_vars, problem.lb.data(), problem.ub.data(), problem.int_vars.data(),
                        (int)problem.int_vars.size()};
    std::vector<double> x(problem.n_vars > 0 ? problem.n_vars : 0);
    gb_result gr{};
    gr.x = x.data();
    const gb_status st = gb_solve(&gp, bind.objective(), bind.lmo(), &s, &gr);
    cb.rethrow();
    detail::throw_status(st);
    Result r;
    r.objective = gr.objective;
    if (std::isfinite(gr.objective)) r.x = x;
    r.dual_bound = gr.dual_bound;
    r.nodes_processed = gr.nodes_processed;
    r.elapsed_seconds = gr.elapsed_seconds;
    r.reached_tolerance = gr.reached_tolerance != 0;
    r.reached_time_limit = gr.reached_time_limit != 0;
    return r;
}

// gpuboscia.cuh:108-170: the algorithm policy from the settings, with the
// reference's checks and messages.
template <typename DeviceObjective, typename LMO, typename Heuristic = RoundingHeuristic>
Result solve(const Problem& problem, DeviceObjective objective, LMO lmo,
             const Settings& settings = create_default_settings(), Heuristic heuristic = {}) {
    const LineSearchKind k = settings.line_search.kind;
    const SecantLS secant{settings.line_search.secant_iters, settings.line_search.secant_tol,
                          settings.line_search.secant_denom_eps};
    if (k == LineSearchKind::QuadraticExact) {
        if (!SupportsQuadraticExact<DeviceObjective>::value)
            throw std::invalid_argument("quadratic-exact requires an objective with quadratic curvature support");
        if (settings.algorithm == FWAlgorithmKind::DICG) {
            if (!lmo_supports_dicg<LMO>::value) throw std::invalid_argument("the selected LMO does not support DICG");
            return solve_with_algorithm(problem, objective, lmo, settings, DICGQuadraticExact{}, heuristic);
        }
        return solve_with_algorithm(problem, objective, lmo, settings, FrankWolfeQuadraticExact{}, heuristic);
    }
    if (settings.algorithm == FWAlgorithmKind::DICG) {
        if (!lmo_supports_dicg<LMO>::value)
            throw std::invalid_argument("the selected LMO does not support DICG; use Frank-Wolfe");
        if (k == LineSearchKind::Secant)
            return solve_with_algorithm(problem, objective, lmo, settings, DICGSecant{secant}, heuristic);
        if (k == LineSearchKind::ScaledOpenLoop)
            return solve_with_algorithm(problem, objective, lmo, settings, DICGScaledOpenLoop{}, heuristic);
        if (k == LineSearchKind::SafeguardedScaledOpenLoop)
            return solve_with_algorithm(problem, objective, lmo, settings, DICGSafeguardedScaledOpenLoop{},
                                        heuristic);
        return solve_with_algorithm(problem, objective, lmo, settings, DICGOpenLoop{}, heuristic);
    }
    if (k == LineSearchKind::ScaledOpenLoop || k == LineSearchKind::SafeguardedScaledOpenLoop)
        throw std::invalid_argument("scaled open-loop is available only for DICG");
    if (k == LineSearchKind::Secant)
        return solve_with_algorithm(problem, objective, lmo, settings, FrankWolfeSecant{secant}, heuristic);
    return solve_with_algorithm(problem, objective, lmo, settings, FrankWolfeOpenLoop{}, heuristic);
}

}  // namespace gpuboscia
