// The reference's examples/minimal.cu through the C++ drop-in header:
// min (x0-.3)^2 + (x1-.7)^2 over binary x; expected x=[0,1], objective .18.
// Also exercises the std::function hooks (progress, root validator, custom
// BranchStrategy) and the exception mapping.  Exit code 0 = all checks pass.
#include <gpuboscia/gpuboscia.hpp>

#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <vector>

static double* upload(const std::vector<double>& v) {
    double* d = nullptr;
    if (hipMalloc(reinterpret_cast<void**>(&d), v.size() * sizeof(double)) != hipSuccess
            || hipMemcpy(d, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        throw std::runtime_error("upload failed");
    return d;
}

int main() {
    int failures = 0;
    auto check = [&](bool ok, const char* what) {
        std::printf("%s: %s\n", ok ? "PASS" : "FAIL", what);
        if (!ok) ++failures;
    };
    try {
        double* dQ = upload({1, 0, 0, 1});
        double* dt = upload({0.3, 0.7});
        ILSGlobalCachedObjective<2> objective{dQ, dt, 0.3 * 0.3 + 0.7 * 0.7};
        Problem problem{};
        problem.n_vars = 2;
        problem.lb = {0, 0};
        problem.ub = {1, 1};
        problem.int_vars = {0, 1};
        auto settings = gpuboscia::create_default_settings();
        settings.line_search.kind = LineSearchKind::QuadraticExact;
        settings.natural_lower_bound = 0;
        settings.time_limit_seconds = 10;
        int progress_calls = 0;
        settings.progress = [&](int, int, double, double, bool) { ++progress_calls; };
        const auto result = gpuboscia::solve(problem, objective, BoxLMO{}, settings);
        std::printf("x=[%.0f, %.0f], objective=%.8f, lower_bound=%.8f\n", result.x[0], result.x[1],
                    result.objective, result.dual_bound);
        check(result.x.size() == 2 && result.x[0] == 0 && result.x[1] == 1
                  && std::abs(result.objective - .18) < 1e-8, "minimal example optimum [0,1], 0.18");
        check(progress_calls > 0, "progress callback invoked");

        // A BranchStrategy that never branches: one node, no incumbent from branching.
        auto s2 = settings;
        s2.branch = [](const Node&, int&, const std::vector<double>&, const std::vector<int>&) {
            return std::vector<Node>{};
        };
        const auto r2 = gpuboscia::solve(problem, objective, BoxLMO{}, s2);
        check(r2.nodes_processed == 1, "custom BranchStrategy returning no children");

        // Supplied root point plus a custom validator that rejects it.
        auto s3 = settings;
        s3.root_initial_point = {0.8, 0.8};
        s3.root_initial_point_feasible = [](const std::vector<double>& x, const std::vector<double>&,
                                            const std::vector<double>&, double) { return x[0] + x[1] <= 1; };
        bool threw = false;
        try { gpuboscia::solve(problem, objective, BoxLMO{}, s3); }
        catch (const std::invalid_argument&) { threw = true; }
        check(threw, "root validator rejection raises std::invalid_argument");

        // Equality simplex LMO with integer marks carried by the functor.
        auto s4 = gpuboscia::create_default_settings();
        s4.natural_lower_bound = 0;
        EqualitySimplexLMO<2> simplex(1.0, 2, problem.int_vars);
        const auto r4 = gpuboscia::solve(problem, objective, simplex, s4);
        check(r4.x.size() == 2 && r4.x[0] + r4.x[1] == 1.0 && std::abs(r4.objective - .18) < 1e-8,
              "equality simplex optimum");

        auto s5 = settings;
        s5.device.gpu_block_threads = 96;
        threw = false;
        try { gpuboscia::solve(problem, objective, BoxLMO{}, s5); }
        catch (const std::invalid_argument& e) { threw = std::string(e.what()).find("multiple of 64") != std::string::npos; }
        check(threw, "invalid block size raises std::invalid_argument");
        (void)hipFree(dQ);
        (void)hipFree(dt);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return failures == 0 ? 0 : 1;
}
