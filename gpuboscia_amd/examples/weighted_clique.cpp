// Maximum-weight induced subgraph of fixed cardinality (the reference's
// examples/weighted_clique.cu):
//
//   minimise  2 sum_{i<j} (1 - w_ij) x_i x_j + lambda sum_i x_i (x_i - 1)
//   over      sum_i x_i = Cardinality, x binary,
//
// with random weights w_ij in [0, 1), lambda = Vertices - 1 (the diagonal
// term vanishes at binary points and makes Q diagonally dominant, so the
// relaxation is convex), over the equality simplex with integer marks.
// Minimising the complement weight maximises the induced weight.  Options:
// --seed N (weights), --time-limit SEC.  The weights come from the
// project's SplitMix64 stream (the reference draws them with
// std::mt19937_64, whose distributions are implementation-defined).
#include <gpuboscia/gpuboscia.hpp>
#include <gpuboscia/logger.hpp>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

constexpr int Vertices = 20;
constexpr int Cardinality = 5;

struct Options {
    unsigned long long seed = 222222222ull;
    double time_limit = 3600.0;
};

Options parse_options(int argc, char** argv) {
    Options o;
    for (int i = 1; i < argc; ++i) {
        const std::string arg = argv[i];
        if (arg == "--seed" && i + 1 < argc) o.seed = std::strtoull(argv[++i], nullptr, 10);
        else if (arg == "--time-limit" && i + 1 < argc) o.time_limit = std::atof(argv[++i]);
        else if (arg == "--help" || arg == "-h") {
            std::printf("Usage: %s [--seed N] [--time-limit SEC]\n", argv[0]);
            std::exit(0);
        } else {
            throw std::runtime_error("unknown or incomplete option: " + arg);
        }
    }
    return o;
}

double uniform01(std::uint64_t& s) {
    std::uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (double)((z ^ (z >> 31)) >> 11) * 0x1.0p-53;
}

}  // namespace

int main(int argc, char** argv) {
    double *d_Qt = nullptr, *d_lin = nullptr;
    try {
        const Options options = parse_options(argc, argv);
        std::uint64_t state = options.seed;
        std::vector<double> w((size_t)Vertices * Vertices, 0.0), comp((size_t)Vertices * Vertices, 0.0);
        for (int i = 0; i < Vertices; ++i)
            for (int j = i + 1; j < Vertices; ++j) {
                w[i * Vertices + j] = w[j * Vertices + i] = uniform01(state);
                comp[i * Vertices + j] = comp[j * Vertices + i] = 1.0 - w[i * Vertices + j];
            }
        // 0.5 x'Qx + lin'x with Q_ij = 2 (1 - w_ij), Q_ii = 2 lambda, lin_i = -lambda
        constexpr double lambda = Vertices - 1.0;
        std::vector<double> Qt((size_t)Vertices * Vertices, 0.0), lin(Vertices, -lambda);
        for (int i = 0; i < Vertices; ++i)
            for (int j = 0; j < Vertices; ++j) Qt[j * Vertices + i] = i == j ? 2.0 * lambda : 2.0 * comp[i * Vertices + j];
        if (hipMalloc(reinterpret_cast<void**>(&d_Qt), Qt.size() * sizeof(double)) != hipSuccess
                || hipMalloc(reinterpret_cast<void**>(&d_lin), lin.size() * sizeof(double)) != hipSuccess
                || hipMemcpy(d_Qt, Qt.data(), Qt.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess
                || hipMemcpy(d_lin, lin.data(), lin.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error("device copy failed");

        GlobalCachedDenseQuadraticObjective<Vertices, 64> objective(d_Qt, d_lin, 1.0, 1.0, 0.0);
        Problem problem;
        problem.n_vars = Vertices;
        problem.lb.assign(Vertices, 0.0);
        problem.ub.assign(Vertices, 1.0);
        problem.int_vars.resize(Vertices);
        std::iota(problem.int_vars.begin(), problem.int_vars.end(), 0);
        EqualitySimplexLMO<Vertices> lmo(static_cast<double>(Cardinality), Vertices, problem.int_vars);

        auto settings = gpuboscia::create_default_settings();
        settings.algorithm = FWAlgorithmKind::DICG;
        settings.time_limit_seconds = options.time_limit;
        settings.device.gpu_batch_size = 1024;
        settings.device.num_streams = 2;
        settings.device.gpu_max_iters = 1000;
        settings.device.abs_gap_tol = 1e-7;
        settings.device.rel_gap_tol = 1e-3;
        // sum x = C gives sum x^2 >= C^2 / V, so lambda (C^2 / V - C) bounds
        // the relaxation from below
        settings.natural_lower_bound = lambda * (double(Cardinality * Cardinality) / Vertices - Cardinality);

        Logger logger(true, 20000);
        settings.progress = [&logger](int i, int o, double b, double inc, bool fresh) {
            logger.print_row(i, o, b, inc, fresh);
        };
        LoggerParams lp;
        lp.traversal_strategy = "best_first";
        lp.branching_strategy = "branch_pow2(k=3, binary)";
        lp.algorithm = fw_algorithm_name(settings.algorithm);
        lp.line_search_method = line_search_name(settings.line_search.kind);
        lp.n_vars = Vertices;
        lp.n_int_vars = Vertices;
        lp.abs_gap_tol = settings.device.abs_gap_tol;
        lp.rel_gap_tol = settings.device.rel_gap_tol;
        lp.time_limit_s = options.time_limit;
        lp.natural_lower_bound = settings.natural_lower_bound;
        logger.print_header(lp);

        const gpuboscia::Result result = gpuboscia::solve(problem, objective, lmo, settings);
        SolutionStats stats;
        stats.primal_objective = result.objective;
        stats.dual_bound = result.dual_bound;
        stats.total_nodes = result.nodes_processed;
        stats.total_time_s = result.elapsed_seconds;
        stats.solution = result.x;
        stats.solution_status = result.reached_time_limit ? "Time limit reached"
                                                          : (result.reached_tolerance ? "Reached tolerance" : "Finished");
        logger.print_footer(stats);
        if (result.x.size() != (size_t)Vertices) throw std::runtime_error("finished without an incumbent solution");
        std::vector<int> sel;
        std::printf("selected vertices:");
        for (int i = 0; i < Vertices; ++i)
            if (result.x[i] > 0.5) {
                sel.push_back(i);
                std::printf(" %d", i + 1);
            }
        std::printf("\n");
        double weight = 0.0, score = 0.0;
        for (size_t a = 0; a < sel.size(); ++a)
            for (size_t b = a + 1; b < sel.size(); ++b) {
                weight += w[sel[a] * Vertices + sel[b]];
                score += comp[sel[a] * Vertices + sel[b]];
            }
        std::printf("minimum complement score (without factor 2): %.12f\n", score);
        std::printf("corresponding original internal weight: %.12f\n", weight);
        std::printf("GPUBoscia minimization objective: %.12f\n", result.objective);
        std::printf("dual bound: %.12f\n", result.dual_bound);
        std::printf("nodes processed: %d\n", result.nodes_processed);
        std::printf("elapsed seconds: %.6f\n", result.elapsed_seconds);
        (void)hipFree(d_lin);
        (void)hipFree(d_Qt);
        return 0;
    } catch (const std::exception& e) {
        if (d_lin) (void)hipFree(d_lin);
        if (d_Qt) (void)hipFree(d_Qt);
        std::fprintf(stderr, "weighted_clique: %s\n", e.what());
        return 1;
    }
}
