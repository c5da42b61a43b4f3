// The reference's solver-level known-answer tests (tests/test_pipeline.cu),
// restated through the C++ drop-in header and its DeviceResidentHybridSolver:
//   * uncertified integral node requeue (:1604-1630): the root, cut to one
//     iteration per batch, is requeued, resumed at its open-loop step 1 with
//     x = 2/3 and its earlier bound, and the search ends at 0.09;
//   * completed slot refilled before post-processing (:723-779);
//   * a one-iteration budget does not close an uncertified integral node
//     (:864-877);
//   * Birkhoff DICG finds the target permutation (:968-998);
//   * secant tolerance stall regression through the solver's last-GPU
//     accessors (:1133-1205);
//   * equality-simplex convergence matrix (:1062-1131);
//   * MultiBlock configuration errors and the linear-objective solves that
//     end at the LMO's own minimum, K = 1, 2, 4, 8 (test_multiblock.cu:287-406),
//     and the same search with K = 3 as with one workgroup per node.
// Exit code 0 = every check passed.
#include <gpuboscia/gpuboscia.hpp>

#include <cmath>
#include <cstdio>
#include <limits>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

int failures = 0;

void check(bool ok, const std::string& what) {
    std::printf("%s: %s\n", ok ? "PASS" : "FAIL", what.c_str());
    std::fflush(stdout);
    if (!ok) ++failures;
}

double* upload(const std::vector<double>& v) {
    double* d = nullptr;
    if (hipMalloc(reinterpret_cast<void**>(&d), v.size() * sizeof(double)) != hipSuccess
            || hipMemcpy(d, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        throw std::runtime_error("upload failed");
    return d;
}

struct DeviceVec {
    double* p;
    explicit DeviceVec(const std::vector<double>& v) : p(upload(v)) {}
    ~DeviceVec() { (void)hipFree(p); }
};

std::vector<double> identity(int n) {
    std::vector<double> q((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) q[(size_t)i * n + i] = 1.0;
    return q;
}

double sumsq(const std::vector<double>& v) {
    double s = 0.0;
    for (double x : v) s += x * x;
    return s;
}

Problem make_problem(int n, const std::vector<double>& lb, const std::vector<double>& ub) {
    Problem p{};
    p.n_vars = n;
    p.lb = lb;
    p.ub = ub;
    p.int_vars.resize(n);
    std::iota(p.int_vars.begin(), p.int_vars.end(), 0);
    return p;
}

DeviceResidentConfig device_cfg(int batch, int streams) {
    DeviceResidentConfig cfg;
    cfg.gpu_batch_size = batch;
    cfg.num_streams = streams;
    cfg.gpu_max_iters = 50000;
    cfg.gpu_eps = 1e-6;
    return cfg;
}

BranchStrategy pow2_branch() {
    return [](const Node& nd, int& nid, const std::vector<double>& x, const std::vector<int>& iv) {
        return branch_pow2(nd, nid, x, iv, 3);
    };
}

void requeue_kat() {
    std::printf("\n[Uncertified integral node requeue]\n");
    DeviceVec dQ({1.0}), dT({0.3});
    using Objective = ILSGlobalCachedObjective<1, 32>;
    Objective obj{dQ.p, dT.p, 0.09};
    Problem problem = make_problem(1, {0}, {1});
    bool resumed = false, x_ok = false, bound_ok = false;
    auto branch = [&](const Node& node, int& id, const std::vector<double>& x, const std::vector<int>& iv) {
        if (node.id == 0) {
            resumed = node.relaxation_iterations == 1;
            x_ok = std::abs(x[0] - 2.0 / 3.0) <= 1e-12;
            bound_ok = node.dual_bound >= -0.51 - 1e-12;
        }
        return branch_pow2(node, id, x, iv, 1);
    };
    auto cfg = device_cfg(1, 1);
    cfg.gpu_max_iters = 1;
    DeviceResidentHybridSolver<Objective, FrankWolfeOpenLoop> solver(problem, obj, FrankWolfeOpenLoop{}, branch,
                                                                     cfg, BoxLMO{}, 5.0);
    solver.push(Node(0, -1, 0, problem.lb, problem.ub));
    solver.solve();
    check(resumed, "same root is requeued and resumed before branching");
    check(x_ok, "requeued root continues its open-loop schedule (x = 2/3)");
    check(bound_ok, "requeued root preserves its prior bound");
    check(std::abs(solver.incumbent_value() - 0.09) <= 1e-12, "GPU-only refinement preserves optimum 0.09");
    check(!solver.reached_time_limit(), "GPU continuation completes the search");
}

void refill_ordering() {
    std::printf("\n[Completed slot immediate refill ordering]\n");
    const int n = 2;
    const std::vector<double> c{0.4, 0.6};
    DeviceVec dQ(identity(n)), dc(c);
    using Objective = ILSGlobalCachedObjective<2>;
    Objective obj{dQ.p, dc.p, sumsq(c)};
    Problem prob = make_problem(n, {0, 0}, {1, 1});
    std::vector<DRSSchedulingEvent> events;
    DeviceResidentHybridSolver<Objective, FrankWolfePolicy> solver(prob, obj, FrankWolfePolicy{}, pow2_branch(),
                                                                   device_cfg(2, 1));
    solver.set_scheduling_callback([&](const DRSSchedulingEvent& e) { events.push_back(e); });
    const std::vector<std::vector<double>> fixed = {{0, 0}, {0, 1}, {1, 0}, {1, 1}};
    for (int i = 0; i < 4; ++i) {
        Node node(i, -1, 0, fixed[i], fixed[i]);
        node.dual_bound = -1e300;
        solver.push(node);
    }
    solver.solve();
    size_t detached1 = events.size(), dispatched2 = events.size(), post1 = events.size();
    for (size_t i = 0; i < events.size(); ++i) {
        const auto& e = events[i];
        if (e.slot_id != 0) continue;
        if (e.type == DRSSchedulingEventType::CompletionDetached && e.generation == 1) detached1 = i;
        if (e.type == DRSSchedulingEventType::Dispatched && e.generation == 2) dispatched2 = i;
        if (e.type == DRSSchedulingEventType::PostprocessStarted && e.generation == 1) post1 = i;
    }
    check(detached1 < events.size(), "first completed batch was detached from slot 0");
    check(dispatched2 < events.size(), "slot 0 was refilled with the second full batch");
    check(post1 < events.size(), "CPU post-processing started for the first batch");
    check(detached1 < dispatched2 && dispatched2 < post1,
          "completed slot is refilled before its old batch is post-processed");
    // the integral optimum of (x - c)^2 over {0,1}^2 is x = (0, 1), value 0.16 + 0.16
    check(std::abs(solver.incumbent_value() - (0.16 + 0.16)) <= 1e-12, "fixed nodes give the integral optimum");
    check(solver.nodes_processed() == 4, "four nodes processed");
}

void one_step_budget() {
    std::printf("\n[One-step budget keeps uncertified integral nodes open]\n");
    const std::vector<double> c{0.5, 0.5};
    DeviceVec dQ(identity(2)), dc(c);
    using Objective = ILSGlobalCachedObjective<2>;
    Objective obj{dQ.p, dc.p, sumsq(c)};
    Problem prob = make_problem(2, {0, 0}, {1, 1});
    auto cfg = device_cfg(2, 1);
    cfg.gpu_max_iters = 1;
    DeviceResidentHybridSolver<Objective, FrankWolfePolicy> solver(prob, obj, FrankWolfePolicy{}, pow2_branch(),
                                                                   cfg);
    solver.push(Node(0, -1, 0, prob.lb, prob.ub));
    solver.solve();
    check(std::abs(solver.incumbent_value() - 0.5) <= 1e-6, "short GPU batches preserve optimum 0.5");
    const auto x = solver.incumbent_x();
    bool integral = x.size() == 2;
    for (double v : x) integral = integral && (v == 0.0 || v == 1.0);
    check(integral, "GPU candidate scoring returns exact integer incumbents");
}

void birkhoff_target() {
    std::printf("\n[BirkhoffLMO DICG finds the target permutation]\n");
    constexpr int order = 4, n = order * order;
    std::vector<double> target(n, 0.0);
    for (int r = 0; r < order; ++r) target[r * order + (r + 1) % order] = 1.0;
    DeviceVec dQ(identity(n)), dc(target);
    using Objective = ILSGlobalCachedObjective<n>;
    Objective obj{dQ.p, dc.p, sumsq(target)};
    Problem prob = make_problem(n, std::vector<double>(n, 0.0), std::vector<double>(n, 1.0));
    DeviceResidentHybridSolver<Objective, DICGPolicy, BirkhoffLMO<order>> solver(
        prob, obj, DICGPolicy{}, pow2_branch(), device_cfg(4, 1), BirkhoffLMO<order>{});
    Node root(0, -1, 0, prob.lb, prob.ub);
    root.dual_bound = -1e300;
    solver.push(root);
    solver.solve();
    check(std::abs(solver.incumbent_value()) <= 1e-10, "Birkhoff DICG finds the target permutation");
    check(solver.incumbent_x() == target, "incumbent is the target permutation matrix");
}

struct StallProbe {
    int iterations;
    double gap;
    bool converged;
};

StallProbe stall_probe(SecantLS ls) {
    DeviceVec dQ({4e-4}), dc({-2e-4});
    using Objective = GlobalCachedDenseQuadraticObjective<1, 32>;
    Objective objective{dQ.p, dc.p, 1.0, 1.0, 0.0};
    Problem problem = make_problem(1, {0.0}, {1.0});
    problem.int_vars.clear();
    DeviceResidentConfig cfg = device_cfg(1, 1);
    cfg.gpu_max_iters = 64;
    cfg.gpu_eps = 1e-6;
    BranchStrategy no_branch = [](const Node&, int&, const std::vector<double>&, const std::vector<int>&) {
        return std::vector<Node>{};
    };
    DeviceResidentHybridSolver<Objective, DICGSecant> solver(problem, objective, DICGSecant{ls}, no_branch, cfg,
                                                             BoxLMO{}, 0.25);
    Node root(0, -1, 0, problem.lb, problem.ub);
    root.dual_bound = -1e300;
    solver.push(root);
    solver.solve();
    return {solver.last_gpu_iterations(), solver.last_gpu_gap(), solver.last_gpu_converged()};
}

void secant_stall() {
    std::printf("\n[Secant tolerance vs FW convergence regression]\n");
    const StallProbe sparse = stall_probe(SecantLS{8, 1e-3, 1e-4});
    const StallProbe tight = stall_probe(SecantLS{16, 1e-10, 1e-14});
    std::printf("  sparse: iterations=%d gap=%.6e converged=%d; tight: iterations=%d gap=%.6e converged=%d\n",
                sparse.iterations, sparse.gap, sparse.converged, tight.iterations, tight.gap, tight.converged);
    check(!sparse.converged && sparse.iterations >= 64 && sparse.gap > 1e-6,
          "sparse-Sharpe secant defaults reproduce the zero-step stall");
    check(tight.converged && tight.iterations < 64 && tight.gap < 1e-6,
          "tight secant tolerance converges on the identical root node");
}

template <class Algorithm>
void convergence_case(const char* name, Algorithm alg, const Problem& prob, const std::vector<double>& Qt,
                      const std::vector<double>& linear, double constant, bool expect) {
    constexpr int n = 120;
    DeviceVec dQ(Qt), dl(linear);
    using Objective = GlobalCachedDenseQuadraticObjective<n, 256>;
    Objective obj{dQ.p, dl.p, 1.0, 1.0, constant};
    EqualitySimplexLMO<n> lmo(200.0, n, prob.int_vars);
    DeviceResidentConfig cfg = device_cfg(1, 1);
    cfg.gpu_max_iters = 20000;
    cfg.gpu_eps = 1e-4;
    cfg.dicg_oracle_parallelism = DeviceResidentConfig::DICGOracleParallelism::Serial;
    BranchStrategy no_branch = [](const Node&, int&, const std::vector<double>&, const std::vector<int>&) {
        return std::vector<Node>{};
    };
    DeviceResidentHybridSolver<Objective, Algorithm, EqualitySimplexLMO<n>> solver(prob, obj, alg, no_branch, cfg,
                                                                                   lmo);
    Node root(0, -1, 0, prob.lb, prob.ub);
    root.dual_bound = -1e300;
    solver.push(root);
    solver.solve();
    const int it = solver.last_gpu_iterations();
    const double gap = solver.last_gpu_gap();
    std::printf("  %-28s iterations=%6d gap=%.6e converged=%s\n", name, it, gap,
                solver.last_gpu_converged() ? "yes" : "NO");
    check(std::isfinite(gap) && gap >= 0.0, std::string(name) + " returns a finite nonnegative FW gap");
    check(it > 0 && it <= 20000, std::string(name) + " reports a valid iteration count");
    if (expect) check(solver.last_gpu_converged(), std::string(name) + " converges on equality-simplex quadratic");
}

void convergence_matrix() {
    std::printf("\n[EqualitySimplex convergence matrix - same quadratic]\n");
    constexpr int n = 120, m = 200;
    const std::vector<double> lb(n, 0.0), ub(n, 8.0);
    std::mt19937 rng(0x5eed1234u);
    std::normal_distribution<double> a_dist(0.0, 1.0 / std::sqrt(static_cast<double>(m)));
    std::normal_distribution<double> b_dist(0.0, 1.0);
    std::vector<double> A((size_t)m * n), b(m);
    for (double& v : A) v = a_dist(rng);
    for (double& v : b) v = b_dist(rng);
    std::vector<double> Q((size_t)n * n, 0.0), linear(n, 0.0);
    for (int j = 0; j < n; ++j) {
        for (int k = 0; k < n; ++k)
            for (int i = 0; i < m; ++i) Q[(size_t)j * n + k] += A[(size_t)i * n + j] * A[(size_t)i * n + k];
        for (int i = 0; i < m; ++i) linear[j] -= A[(size_t)i * n + j] * b[i];
    }
    std::vector<double> Qt((size_t)n * n);
    for (int j = 0; j < n; ++j)
        for (int k = 0; k < n; ++k) Qt[(size_t)k * n + j] = Q[(size_t)j * n + k];
    double constant = 0.0;
    for (double v : b) constant += 0.5 * v * v;
    Problem prob = make_problem(n, lb, ub);
    convergence_case("FW + open-loop", FrankWolfeOpenLoop{}, prob, Qt, linear, constant, false);
    convergence_case("FW + secant", FrankWolfeSecant{SecantLS{16, 1e-10, 1e-14}}, prob, Qt, linear, constant, false);
    convergence_case("DICG + open-loop", DICGOpenLoop{}, prob, Qt, linear, constant, false);
    convergence_case("DICG + scaled open-loop", DICGScaledOpenLoop{}, prob, Qt, linear, constant, false);
    convergence_case("DICG + safeguarded scaled", DICGSafeguardedScaledOpenLoop{}, prob, Qt, linear, constant, true);
    convergence_case("DICG + secant", DICGSecant{SecantLS{16, 1e-10, 1e-14}}, prob, Qt, linear, constant, true);
}

void solve_with_algorithm_and_push_errors() {
    std::printf("\n[solve_with_algorithm and push validation]\n");
    const std::vector<double> t{0.3, 0.7};
    DeviceVec dQ(identity(2)), dt(t);
    ILSGlobalCachedObjective<2> obj{dQ.p, dt.p, sumsq(t)};
    Problem prob = make_problem(2, {0, 0}, {1, 1});
    auto s = gpuboscia::create_default_settings();
    s.natural_lower_bound = 0;
    const auto r = gpuboscia::solve_with_algorithm(prob, obj, BoxLMO{}, s, DICGQuadraticExact{});
    check(r.x.size() == 2 && r.x[0] == 0 && r.x[1] == 1 && std::abs(r.objective - 0.18) < 1e-8,
          "solve_with_algorithm(DICGQuadraticExact) gives [0, 1], 0.18");
    DeviceResidentHybridSolver<ILSGlobalCachedObjective<2>, DICGPolicy> solver(prob, obj, DICGPolicy{},
                                                                              pow2_branch());
    bool threw = false;
    try {
        Node bad(0, -1, 0, {0, 0, 0}, {1, 1, 1});
        solver.push(bad);
    } catch (const std::invalid_argument& e) {
        threw = std::string(e.what()).find("invalid node buffer dimensions") != std::string::npos;
    }
    check(threw, "push rejects a node of the wrong size");
    threw = false;
    try {
        Node bad(0, -1, 0, {0, 0}, {1, 1});
        bad.relaxation_x = {0.5, 1.5};
        solver.push(bad);
    } catch (const std::invalid_argument& e) {
        threw = std::string(e.what()).find("violates finite bounds") != std::string::npos;
    }
    check(threw, "push rejects an iterate outside the node's bounds");
    threw = false;
    try {
        Node bad(0, -1, 0, {0, 0}, {1, 1});
        bad.relaxation_iterations = 3;
        solver.push(bad);
    } catch (const std::invalid_argument& e) {
        threw = std::string(e.what()).find("resume state") != std::string::npos;
    }
    check(threw, "push rejects a resume count without an iterate");
    threw = false;
    try {
        auto cfg = device_cfg(0, 1);
        DeviceResidentHybridSolver<ILSGlobalCachedObjective<2>, DICGPolicy> bad(prob, obj, DICGPolicy{},
                                                                                pow2_branch(), cfg);
    } catch (const std::invalid_argument& e) {
        threw = std::string(e.what()).find("batch size") != std::string::npos;
    }
    check(threw, "constructor rejects a zero batch size");
    // an exception thrown by a BranchStrategy propagates out of solve()
    DeviceResidentHybridSolver<ILSGlobalCachedObjective<2>, DICGPolicy> thrower(
        prob, obj, DICGPolicy{},
        [](const Node&, int&, const std::vector<double>&, const std::vector<int>&) -> std::vector<Node> {
            throw std::domain_error("branch strategy failed");
        });
    Node root(0, -1, 0, prob.lb, prob.ub);
    thrower.push(root);
    threw = false;
    try {
        thrower.solve();
    } catch (const std::domain_error&) {
        threw = true;
    }
    check(threw, "a BranchStrategy exception propagates out of solve()");
}


template <class F>
void must_throw(F fn, const char* fragment, const std::string& what) {
    bool ok = false;
    std::string msg;
    try {
        fn();
    } catch (const std::invalid_argument& e) {
        msg = e.what();
        ok = msg.find(fragment) != std::string::npos;
    }
    check(ok, what + " (" + (msg.empty() ? std::string("no exception") : msg) + ")");
}

BranchStrategy no_branch() {
    return [](const Node&, int&, const std::vector<double>&, const std::vector<int>&) { return std::vector<Node>{}; };
}

// oracle_solver_integration (test_multiblock.cu:369-406): f(x) = g'x with
// g_j = j mod 7 - 3 over [0, 1]^N, no integer variables, so the incumbent is
// the LMO's minimum of g'v.
template <int N, class LMO>
void multiblock_linear(int k, const LMO& lmo, const char* name, double expected) {
    std::vector<double> g(N);
    for (int j = 0; j < N; ++j) g[j] = double(j % 7) - 3.0;
    DeviceVec dQ(std::vector<double>((size_t)N * N, 0.0)), dg(g);
    using Obj = GlobalCachedDenseQuadraticObjective<N>;
    Obj obj(dQ.p, dg.p, 1.0, 1.0, 0.0);
    Problem prob{};
    prob.n_vars = N;
    prob.lb.assign(N, 0.0);
    prob.ub.assign(N, 1.0);
    DeviceResidentConfig cfg;
    cfg.execution_mode = FWExecutionMode::MultiBlock;
    cfg.blocks_per_node = k;
    cfg.gpu_batch_size = 4;
    cfg.gpu_max_iters = 10;
    cfg.gpu_eps = cfg.gpu_eps_min = 1e-8;
    DeviceResidentHybridSolver<Obj, FrankWolfeOpenLoop, LMO> solver(prob, obj, {}, no_branch(), cfg, lmo);
    solver.push(Node(0, -1, 0, prob.lb, prob.ub));
    solver.solve();
    check(std::abs(solver.incumbent_value() - expected) <= 1e-8 && solver.effective_num_streams() == 1,
          std::string("MultiBlock K=") + std::to_string(k) + " " + name + " n=" + std::to_string(N)
              + ": optimum " + std::to_string(solver.incumbent_value()));
}

void multiblock_solver() {
    std::printf("\n[MultiBlock solver (test_multiblock.cu:287-406)]\n");
    constexpr int n = 16;
    DeviceVec dQ(identity(n)), dc(std::vector<double>(n, -0.5));
    using Obj = GlobalCachedDenseQuadraticObjective<n>;
    Obj obj(dQ.p, dc.p);
    Problem prob = make_problem(n, std::vector<double>(n, 0.0), std::vector<double>(n, 1.0));
    DeviceResidentConfig cfg;
    cfg.execution_mode = FWExecutionMode::MultiBlock;
    cfg.blocks_per_node = 2;
    auto construct = [&](DeviceResidentConfig c) {
        DeviceResidentHybridSolver<Obj, FrankWolfeSecant> solver(prob, obj, {}, no_branch(), c);
    };
    auto bad = cfg;
    bad.blocks_per_node = 0;
    must_throw([&] { construct(bad); }, "positive", "blocks_per_node = 0 is rejected");
    bad = cfg;
    bad.blocks_per_node = -2;
    must_throw([&] { construct(bad); }, "positive", "blocks_per_node = -2 is rejected");
    bad = cfg;
    bad.execution_mode = static_cast<FWExecutionMode>(99);
    must_throw([&] { construct(bad); }, "unknown FW execution mode", "an unknown execution mode is rejected");
    bad = cfg;
    bad.dicg_oracle_parallelism = DeviceResidentConfig::DICGOracleParallelism::Paired;
    must_throw([&] { construct(bad); }, "cannot combine", "paired DICG with MultiBlock is rejected");
    bad = cfg;
    bad.blocks_per_node = std::numeric_limits<int>::max();
    must_throw([&] { construct(bad); }, "complete node group", "a group larger than the resident workgroups is rejected");
    bad = cfg;
    bad.gpu_block_threads = 33;
    must_throw([&] { construct(bad); }, "multiple of 64", "gpu_block_threads = 33 is rejected (wave64)");
    for (int k : {1, 2, 4, 8}) {
        multiblock_linear<16>(k, UnitSimplexLMO<16>{}, "unit simplex", -3.0);
        multiblock_linear<16>(k, EqualitySimplexLMO<16>{}, "equality simplex", -3.0);
        multiblock_linear<257>(k, UnitSimplexLMO<257>{}, "unit simplex", -3.0);
        multiblock_linear<257>(k, EqualitySimplexLMO<257>{}, "equality simplex", -3.0);
        multiblock_linear<257>(k, BoxLMO{}, "box", -3.0 * 37 - 2.0 * 37 - 1.0 * 37);
    }
    // the same search, node for node, with K = 3 as with one workgroup per node
    constexpr int m = 30;
    std::mt19937_64 rng(7);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::vector<double> A((size_t)m * m), Q((size_t)m * m, 0.0), c(m);
    for (double& a : A) a = nd(rng);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
            double acc = 0.0;
            for (int t = 0; t < m; ++t) acc += A[(size_t)t * m + i] * A[(size_t)t * m + j];
            Q[(size_t)i * m + j] = acc / m + (i == j ? 0.1 : 0.0);
        }
    for (int j = 0; j < m; ++j) c[j] = nd(rng);
    DeviceVec dQm(Q), dcm(c);
    using ObjM = GlobalCachedDenseQuadraticObjective<m>;
    ObjM objm(dQm.p, dcm.p, 2.0, 0.5, 0.0);
    Problem pm = make_problem(m, std::vector<double>(m, 0.0), std::vector<double>(m, 1.0));
    double val[2], bound[2];
    int nodes[2];
    std::vector<double> xs[2];
    for (int r = 0; r < 2; ++r) {
        DeviceResidentConfig dc = device_cfg(64, 1);
        dc.abs_gap_tol = dc.rel_gap_tol = 1e-6;
        if (r == 1) {
            dc.execution_mode = FWExecutionMode::MultiBlock;
            dc.blocks_per_node = 3;
        }
        DeviceResidentHybridSolver<ObjM, DICGPolicy> solver(pm, objm, {}, pow2_branch(), dc);
        solver.push(Node(0, -1, 0, pm.lb, pm.ub));
        solver.solve();
        val[r] = solver.incumbent_value();
        bound[r] = solver.dual_bound();
        nodes[r] = solver.nodes_processed();
        xs[r] = solver.incumbent_x();
    }
    check(val[0] == val[1] && bound[0] == bound[1] && nodes[0] == nodes[1] && xs[0] == xs[1],
          "MultiBlock K=3 search equals the one-workgroup search (" + std::to_string(nodes[0]) + " nodes)");
}
}  // namespace

int main() {
    try {
        requeue_kat();
        refill_ordering();
        one_step_budget();
        birkhoff_target();
        secant_stall();
        convergence_matrix();
        solve_with_algorithm_and_push_errors();
        multiblock_solver();
    } catch (const std::exception& e) {
        std::printf("FAIL: exception %s\n", e.what());
        return 1;
    }
    std::printf("\n%d failure(s)\n", failures);
    return failures ? 1 : 0;
}
