// Graph isomorphism over the Birkhoff polytope (the reference's
// examples/main.cu): A is the graph `A` of a MAT instance, B a seeded random
// relabelling of it, and the solver minimises ||B X - X A||_F^2 over
// permutation matrices X, whose optimum 0 is an isomorphism.  Any order up
// to 256; the relaxation runs in the QAP's FP32 form (FP32 iterates,
// Hungarian on FP32 costs, FP64-certified bound), the natural lower bound
// is 0.  Options as the reference's examples (--help); --seed N reproduces a
// relabelling.
#include <gpuboscia/gpuboscia.hpp>
#include <gpuboscia/logger.hpp>
#include <gpuboscia/mat.hpp>
#include <gpuboscia/options.hpp>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#ifndef GPUBOSCIA_SOURCE_DIR
#define GPUBOSCIA_SOURCE_DIR "."
#endif

namespace {

std::uint64_t splitmix64(std::uint64_t& s) {
    std::uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct DeviceCopy {
    double* p = nullptr;
    explicit DeviceCopy(const std::vector<double>& v) {
        if (hipMalloc(reinterpret_cast<void**>(&p), v.size() * sizeof(double)) != hipSuccess
                || hipMemcpy(p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error("device copy failed");
    }
    ~DeviceCopy() { (void)hipFree(p); }
};

}  // namespace

int main(int argc, char** argv) {
    GPUBosciaExampleOptions opts;
    const int cli = parse_gpuboscia_example_options(argc, argv, opts);
    if (cli >= 0) return cli;
    const std::string path = opts.instance_path.empty()
                                 ? std::string(GPUBOSCIA_SOURCE_DIR) + "/tests/golden/mat/Triangular_10_45.mat"
                                 : opts.instance_path;
    gpuboscia::mat::Matrix A;
    try {
        A = gpuboscia::mat::read(path, "A");
    } catch (const std::exception& e) {
        std::fprintf(stderr, "Instance error: %s\n", e.what());
        return 1;
    }
    const int order = A.rows;
    if (A.rows != A.cols || order < 1 || order > 256) {
        std::fprintf(stderr, "This example needs a square instance of order 1..256; got %dx%d.\n", A.rows, A.cols);
        return 1;
    }
    std::uint64_t seed = opts.random_seed;
    if (!opts.random_seed_explicit)
        seed = (std::uint64_t(std::random_device{}()) << 32) ^ std::uint64_t(std::random_device{}())
               ^ std::uint64_t(std::chrono::high_resolution_clock::now().time_since_epoch().count());
    std::vector<int> perm(order);
    std::iota(perm.begin(), perm.end(), 0);
    std::uint64_t st = seed;
    for (int i = order - 1; i > 0; --i) std::swap(perm[i], perm[splitmix64(st) % std::uint64_t(i + 1)]);
    std::vector<double> B((size_t)order * order);
    for (int r = 0; r < order; ++r)
        for (int c = 0; c < order; ++c) B[(size_t)r * order + c] = A.values[(size_t)perm[r] * order + perm[c]];

    try {
        // ||B X - X A||^2: the QAP objective ||A' P - P B'||^2 with A' = B, B' = A
        DeviceCopy dA(B), dB(A.values);
        const RuntimeQuadraticAssignmentObjective objective{order, dA.p, dB.p};
        const RuntimeBirkhoffLMO lmo{order};
        const int n = order * order;
        Problem problem;
        problem.n_vars = n;
        problem.lb.assign(n, 0.0);
        problem.ub.assign(n, 1.0);
        problem.int_vars.resize(n);
        std::iota(problem.int_vars.begin(), problem.int_vars.end(), 0);

        auto settings = gpuboscia::create_default_settings();
        settings.device.gpu_batch_size = 128;
        settings.device.num_streams = 2;
        settings.device.gpu_max_iters = 500;
        settings.device.gpu_eps = opts.gpu_eps;
        settings.device.gpu_eps_decay = opts.gpu_eps_decay;
        settings.device.gpu_eps_min = opts.gpu_eps_min;
        settings.device.integer_tol = 1e-5;
        settings.device.abs_gap_tol = opts.abs_gap_tol;
        settings.device.rel_gap_tol = opts.rel_gap_tol;
        settings.algorithm = FWAlgorithmKind::DICG;     // the FP32 form's algorithm
        settings.line_search.kind = LineSearchKind::QuadraticExact;
        settings.time_limit_seconds = opts.time_limit_s;
        settings.natural_lower_bound = std::isfinite(opts.natural_lower_bound) ? opts.natural_lower_bound : 0.0;
        settings.form = 2;
        settings.precision = GB_PREC_FP32;

        Logger logger(true, opts.log_interval_nodes);
        settings.progress = [&logger](int i, int o, double b, double inc, bool fresh) {
            logger.print_row(i, o, b, inc, fresh);
        };
        LoggerParams lp;
        lp.traversal_strategy = "best_first";
        lp.branching_strategy = "branch_pow2(k=3, binary)";
        lp.algorithm = fw_algorithm_name(settings.algorithm);
        lp.line_search_method = line_search_name(settings.line_search.kind);
        lp.n_vars = n;
        lp.n_int_vars = n;
        lp.abs_gap_tol = settings.device.abs_gap_tol;
        lp.rel_gap_tol = settings.device.rel_gap_tol;
        lp.time_limit_s = settings.time_limit_seconds;
        lp.natural_lower_bound = settings.natural_lower_bound;
        lp.random_seed = seed;
        lp.print_random_seed = true;
        logger.print_header(lp);

        const gpuboscia::Result result = gpuboscia::solve(problem, objective, lmo, settings);
        SolutionStats stats;
        stats.primal_objective = result.objective;
        stats.dual_bound = result.dual_bound;
        stats.total_nodes = result.nodes_processed;
        stats.total_time_s = result.elapsed_seconds;
        stats.solution = result.x;
        stats.solution_status = result.reached_time_limit ? "Time limit reached"
                                                          : (result.reached_tolerance ? "Reached tolerance" : "Optimal");
        logger.print_footer(stats);
        bool iso = result.x.size() == (size_t)n && result.objective <= 1e-6;
        for (int r = 0; r < order && iso; ++r)  // B X = X A entry by entry
            for (int c = 0; c < order && iso; ++c) {
                double bx = 0.0, xa = 0.0;
                for (int k = 0; k < order; ++k) {
                    bx += B[(size_t)r * order + k] * result.x[(size_t)k * order + c];
                    xa += result.x[(size_t)r * order + k] * A.values[(size_t)k * order + c];
                }
                iso = std::abs(bx - xa) < 1e-9;
            }
        std::printf("\nGraph isomorphism example:\n  Instance:             %s\n  Random seed:          %llu\n"
                    "  Vertices:             %d\n  Objective:            %.10g\n  Isomorphism found:    %s\n",
                    path.c_str(), static_cast<unsigned long long>(seed), order, result.objective,
                    iso ? "yes" : "no");
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "graph_iso: %s\n", e.what());
        return 1;
    }
}
