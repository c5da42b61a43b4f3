// Discrete sparse Sharpe-ratio portfolio (the reference's
// examples/sparse_sharpe.cu) with the device plug-ins of
// <gpuboscia/device/sparse_sharpe.hpp>:
//
//   minimise 0.5 v'(Sigma + eps I) v - mu'v   over 0 <= v_i <= U_i y_i,
//            sum y <= m,  y binary;
//
// the portfolio is w = v / sum(v).  Returns are the MAT matrix `data` (gross
// returns per row and asset) minus 1 over the last T rows; U_i bounds v_i on
// the ellipsoid {v : 0.5 v'Hv - mu'v <= 0} and -0.5 mu'H^-1 mu is the
// natural lower bound.  Needs hipcc for gfx950 (plug-in types).
#include <gpuboscia/device/sparse_sharpe.hpp>
#include <gpuboscia/gpuboscia.hpp>
#include <gpuboscia/logger.hpp>
#include <gpuboscia/mat.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#ifndef GPUBOSCIA_SOURCE_DIR
#define GPUBOSCIA_SOURCE_DIR "."
#endif

namespace {

constexpr int MaxAssets = 128;
constexpr int Variables = 2 * MaxAssets;

struct Options {
    std::string data = std::string(GPUBOSCIA_SOURCE_DIR) + "/tests/golden/mat/FF25EUnew.mat";
    int cardinality = 10;
    int observations = 60;
    int assets = 128;
    double epsilon = 1e-3;
    double time_limit = 3600.0;
    int threads = 0;
};

void usage(const char* prog) {
    std::printf("Usage: %s [options]\n"
                "  --data PATH          MAT file containing 'data'\n"
                "  --cardinality M      optimize at most M of the candidates (default 10)\n"
                "  --observations T     use last T rows; 0 means all (default 60)\n"
                "  --assets N           use the first N asset columns (default and at most %d)\n"
                "  --epsilon VALUE      covariance ridge (default 1e-3)\n"
                "  --threads N          threads per workgroup; 0=auto, otherwise a multiple of 64 up to 256\n"
                "  --time-limit SEC     time limit (default 3600)\n",
                prog, MaxAssets);
}

Options parse(int argc, char** argv) {
    Options o;
    auto value = [&](int& i) {
        if (++i >= argc) throw std::runtime_error("missing option value");
        return argv[i];
    };
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--data") o.data = value(i);
        else if (a == "--cardinality") o.cardinality = std::atoi(value(i));
        else if (a == "--observations") o.observations = std::atoi(value(i));
        else if (a == "--assets") o.assets = std::atoi(value(i));
        else if (a == "--epsilon") o.epsilon = std::atof(value(i));
        else if (a == "--threads") o.threads = std::atoi(value(i));
        else if (a == "--time-limit") o.time_limit = std::atof(value(i));
        else if (a == "--help" || a == "-h") {
            usage(argv[0]);
            std::exit(0);
        } else {
            throw std::runtime_error("unknown option: " + a);
        }
    }
    return o;
}

// Dense Cholesky factor L (H = L L') and the two triangular solves.
std::vector<double> cholesky(const std::vector<double>& H, int n) {
    std::vector<double> L((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = H[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) s -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
            if (i == j) {
                if (!(s > 0.0)) throw std::runtime_error("covariance matrix is not positive definite");
                L[(size_t)i * n + i] = std::sqrt(s);
            } else {
                L[(size_t)i * n + j] = s / L[(size_t)j * n + j];
            }
        }
    return L;
}
std::vector<double> solve_chol(const std::vector<double>& L, const std::vector<double>& b, int n) {
    std::vector<double> y(n), x(n);
    for (int i = 0; i < n; ++i) {
        double s = b[i];
        for (int j = 0; j < i; ++j) s -= L[(size_t)i * n + j] * y[j];
        y[i] = s / L[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = y[i];
        for (int j = i + 1; j < n; ++j) s -= L[(size_t)j * n + i] * x[j];
        x[i] = s / L[(size_t)i * n + i];
    }
    return x;
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
    Options opts;
    try {
        opts = parse(argc, argv);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "Argument error: %s\n", e.what());
        usage(argv[0]);
        return 2;
    }
    gpuboscia::mat::Matrix prices;
    try {
        prices = gpuboscia::mat::read(opts.data, "data");
    } catch (const std::exception& e) {
        std::fprintf(stderr, "Dataset error: %s\n", e.what());
        return 1;
    }
    const int total_rows = prices.rows, available = prices.cols;
    const int assets = std::min(opts.assets, available);
    if (assets < 1 || assets > MaxAssets || opts.cardinality < 1 || opts.cardinality > assets || opts.epsilon <= 0.0
            || opts.threads < 0 || opts.threads > 256 || opts.threads % 64 != 0) {
        std::fprintf(stderr,
                     "Invalid options: need 1 <= cardinality <= assets <= %d, assets <= %d available columns, "
                     "and threads = 0 or a multiple of 64 up to 256.\n",
                     MaxAssets, available);
        return 2;
    }
    const int T = opts.observations <= 0 ? total_rows : std::min(opts.observations, total_rows);
    const int first = total_rows - T;
    std::vector<double> R((size_t)T * assets), mu(assets, 0.0);
    for (int t = 0; t < T; ++t)
        for (int i = 0; i < assets; ++i) {
            const double r = prices.values[(size_t)(first + t) * available + i] - 1.0;
            R[(size_t)t * assets + i] = r;
            mu[i] += r / T;
        }
    std::vector<double> H((size_t)assets * assets);
    const double cs = T > 1 ? 1.0 / (T - 1) : 1.0;
    for (int i = 0; i < assets; ++i)
        for (int j = 0; j < assets; ++j) {
            double c = 0.0;
            for (int t = 0; t < T; ++t) c += (R[(size_t)t * assets + i] - mu[i]) * (R[(size_t)t * assets + j] - mu[j]);
            H[(size_t)i * assets + j] = c * cs + (i == j ? opts.epsilon : 0.0);
        }
    std::vector<double> U(assets, 0.0);
    double radius_sq = 0.0;
    try {
        const std::vector<double> L = cholesky(H, assets);
        const std::vector<double> center = solve_chol(L, mu, assets);
        radius_sq = std::max(std::inner_product(mu.begin(), mu.end(), center.begin(), 0.0), 0.0);
        for (int i = 0; i < assets; ++i) {
            std::vector<double> e(assets, 0.0);
            e[i] = 1.0;
            const std::vector<double> col = solve_chol(L, e, assets);
            U[i] = std::max(center[i] + std::sqrt(std::max(col[i] * radius_sq, 0.0)), 0.0);
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "Model error: %s\n", e.what());
        return 1;
    }
    try {
        DeviceCopy dH(H), dmu(mu), dU(U);
        const gpuboscia::SparseSharpeObjective<MaxAssets> objective{dH.p, dmu.p, assets};
        const gpuboscia::LinkedCardinalityLMO<MaxAssets> lmo{assets, opts.cardinality, dU.p, dH.p, dmu.p};
        Problem problem;
        problem.n_vars = Variables;
        problem.lb.assign(Variables, 0.0);
        problem.ub.assign(Variables, 0.0);
        for (int i = 0; i < assets; ++i) {
            problem.ub[i] = U[i];
            problem.ub[MaxAssets + i] = 1.0;
            problem.int_vars.push_back(MaxAssets + i);
        }
        auto settings = gpuboscia::create_default_settings();
        settings.device.gpu_batch_size = 1024;
        settings.device.num_streams = 8;
        settings.device.gpu_block_threads = opts.threads;
        settings.device.gpu_max_iters = 10000;
        settings.device.integer_tol = 1e-5;
        settings.device.abs_gap_tol = 1e-6;
        settings.device.rel_gap_tol = 1e-2;
        settings.algorithm = FWAlgorithmKind::DICG;
        settings.line_search.kind = LineSearchKind::Secant;
        settings.line_search.secant_iters = 16;
        settings.line_search.secant_tol = 1e-10;
        settings.line_search.secant_denom_eps = 1e-14;
        settings.time_limit_seconds = opts.time_limit;
        settings.natural_lower_bound = -0.5 * radius_sq;

        Logger logger(true, 2000);
        settings.progress = [&logger](int i, int o, double b, double inc, bool fresh) {
            logger.print_row(i, o, b, inc, fresh);
        };
        LoggerParams lp;
        lp.traversal_strategy = "best_first";
        lp.branching_strategy = "branch_pow2(k=3, binary)";
        lp.algorithm = fw_algorithm_name(settings.algorithm);
        lp.line_search_method = line_search_name(settings.line_search.kind);
        lp.n_vars = Variables;
        lp.n_int_vars = assets;
        lp.abs_gap_tol = settings.device.abs_gap_tol;
        lp.rel_gap_tol = settings.device.rel_gap_tol;
        lp.time_limit_s = opts.time_limit;
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

        int selected = 0, selected_y = 0;
        double scale = 0.0;
        if (result.x.size() == (size_t)Variables)
            for (int i = 0; i < assets; ++i) {
                if (result.x[i] > 1e-7) {
                    ++selected;
                    scale += result.x[i];
                }
                selected_y += result.x[MaxAssets + i] > 0.5;
            }
        std::printf("\nContinuous-weight sparse Sharpe portfolio:\n  Dataset:      %s\n"
                    "  Window:       last %d of %d observations\n  Candidates:   first %d of %d asset columns\n"
                    "  Portfolio:    optimize at most %d candidates\n  Bounds U:     [%.10g, %.10g]\n"
                    "  Support/y:    %d/%d\n  Portfolio weights:\n",
                    opts.data.c_str(), T, total_rows, assets, available, opts.cardinality,
                    *std::min_element(U.begin(), U.end()), *std::max_element(U.begin(), U.end()), selected,
                    selected_y);
        if (selected == 0 || !(scale > 0.0)) {
            std::printf("    no incumbent portfolio found\n");
            return 1;
        }
        double pv = 0.0, vhv = 0.0;
        for (int i = 0; i < assets; ++i) {
            pv += mu[i] * result.x[i];
            for (int j = 0; j < assets; ++j) vhv += result.x[i] * H[(size_t)i * assets + j] * result.x[j];
        }
        for (int i = 0; i < assets; ++i)
            if (result.x[i] > 1e-7)
                std::printf("    asset %-3d v=%.10g  w=%.10g  y=%.0f\n", i + 1, result.x[i], result.x[i] / scale,
                            result.x[MaxAssets + i]);
        const double f = 0.5 * vhv - pv;
        std::printf("  Objective 0.5 v'Hv - mu'v:      %.12g (solver %.12g)\n"
                    "  Regularized Sharpe ratio:       %.10g\n",
                    f, result.objective, vhv > 0.0 ? pv / std::sqrt(vhv) : 0.0);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "sparse_sharpe: %s\n", e.what());
        return 1;
    }
}
