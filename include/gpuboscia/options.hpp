// Command-line options of the example programs: the reference's
// GPUBosciaExampleOptions and its parser (examples/common/options.h:12-213),
// with the same flags, defaults, messages and return convention (-1 to
// continue, else the process exit code).  Host-only C++17.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>

#include "gpuboscia.hpp"

struct GPUBosciaExampleOptions {
    FWAlgorithmKind algorithm = FWAlgorithmKind::DICG;
    LineSearchOptions line_search;
    char oed_criterion = 'd';
    std::uint64_t random_seed = 1;
    bool random_seed_explicit = false;
    double gpu_eps = 1e-3;
    double gpu_eps_decay = 0.8;
    double gpu_eps_min = 1e-6;
    double abs_gap_tol = 1e-6;
    double rel_gap_tol = 1e-2;
    double time_limit_s = 3600.0;
    double natural_lower_bound = -std::numeric_limits<double>::infinity();
    int log_interval_nodes = 100;
    std::string instance_path;
};

inline void print_gpuboscia_example_usage(const char* prog) {
    std::printf("Usage: %s [--algorithm fw|dicg]\n"
                "          [--criterion a|d] [--seed N]\n"
                "          [--line-search open-loop|scaled-open-loop|"
                "safeguarded-scaled-open-loop|secant|quadratic-exact]\n"
                "          [--secant-iters N] [--secant-tol VALUE]\n"
                "          [--secant-denom-eps VALUE]\n"
                "          [--gpu-eps VALUE] [--gpu-eps-decay VALUE]\n"
                "          [--gpu-eps-min VALUE] [--abs-gap-tol VALUE]\n"
                "          [--rel-gap-tol VALUE] [--time-limit SECONDS]\n"
                "          [--natural-lower-bound VALUE]\n"
                "          [--log-interval-nodes N] [--instance PATH]\n",
                prog);
}

namespace gpuboscia {
namespace cli {

inline bool fail(const char* msg) {
    std::fprintf(stderr, "%s\n", msg);
    return false;
}

// A flag's value, or a message and false when it is missing.
inline bool value_of(int argc, char** argv, int& i, const char*& out) {
    if (i + 1 < argc) {
        out = argv[++i];
        return true;
    }
    std::fprintf(stderr, "Missing value for argument: %s\n", argv[i]);
    return false;
}

inline bool to_double(const char* s, double& out) {
    char* end = nullptr;
    out = std::strtod(s, &end);
    return end && *end == '\0';
}

}  // namespace cli
}  // namespace gpuboscia

inline int parse_gpuboscia_example_options(int argc, char** argv, GPUBosciaExampleOptions& opts) {
    namespace cli = gpuboscia::cli;
    for (int i = 1; i < argc; ++i) {
        const std::string arg = argv[i];
        const char* v = nullptr;
        if (arg == "--help") {
            print_gpuboscia_example_usage(argv[0]);
            return 0;
        }
        if (arg == "--algorithm") {
            if (!cli::value_of(argc, argv, i, v)) return 1;
            if (!std::strcmp(v, "fw")) opts.algorithm = FWAlgorithmKind::FrankWolfe;
            else if (!std::strcmp(v, "dicg")) opts.algorithm = FWAlgorithmKind::DICG;
            else return cli::fail("--algorithm must be either 'fw' or 'dicg'."), 1;
        } else if (arg == "--criterion") {
            if (!cli::value_of(argc, argv, i, v)) return 1;
            if (!std::strcmp(v, "a") || !std::strcmp(v, "A")) opts.oed_criterion = 'a';
            else if (!std::strcmp(v, "d") || !std::strcmp(v, "D")) opts.oed_criterion = 'd';
            else return cli::fail("--criterion must be either 'a' or 'd'."), 1;
        } else if (arg == "--seed") {
            if (!cli::value_of(argc, argv, i, v)) return 1;
            char* end = nullptr;
            const unsigned long long s = std::strtoull(v, &end, 10);
            if (!end || *end != '\0') return cli::fail("--seed must be an unsigned integer."), 1;
            opts.random_seed = static_cast<std::uint64_t>(s);
            opts.random_seed_explicit = true;
        } else if (arg == "--log-interval-nodes") {
            if (!cli::value_of(argc, argv, i, v)) return 1;
            char* end = nullptr;
            const long k = std::strtol(v, &end, 10);
            if (!end || *end != '\0' || k <= 0 || k > std::numeric_limits<int>::max())
                return cli::fail("--log-interval-nodes must be a positive integer."), 1;
            opts.log_interval_nodes = static_cast<int>(k);
        } else if (arg == "--time-limit") {
            if (!cli::value_of(argc, argv, i, v)) return 1;
            double t = 0.0;
            if (!cli::to_double(v, t) || !std::isfinite(t) || t <= 0.0)
                return cli::fail("--time-limit must be a positive finite number of seconds."), 1;
            opts.time_limit_s = t;
        } else if (arg == "--instance") {
            if (!cli::value_of(argc, argv, i, v)) return 1;
            opts.instance_path = v;
        } else if (arg == "--natural-lower-bound") {
            if (!cli::value_of(argc, argv, i, v)) return 1;
            double b = 0.0;
            if (!cli::to_double(v, b) || !std::isfinite(b))
                return cli::fail("--natural-lower-bound must be a finite number."), 1;
            opts.natural_lower_bound = b;
        } else if (arg == "--line-search") {
            if (!cli::value_of(argc, argv, i, v)) return 1;
            const std::string k = v;
            if (k == "open-loop") opts.line_search.kind = LineSearchKind::OpenLoop;
            else if (k == "scaled-open-loop") opts.line_search.kind = LineSearchKind::ScaledOpenLoop;
            else if (k == "safeguarded-scaled-open-loop")
                opts.line_search.kind = LineSearchKind::SafeguardedScaledOpenLoop;
            else if (k == "quadratic-exact") opts.line_search.kind = LineSearchKind::QuadraticExact;
            else if (k == "secant") opts.line_search.kind = LineSearchKind::Secant;
            else
                return cli::fail("--line-search must be 'open-loop', 'scaled-open-loop', "
                                 "'safeguarded-scaled-open-loop', 'secant', or 'quadratic-exact'."),
                       1;
        } else if (arg == "--secant-iters") {
            if (!cli::value_of(argc, argv, i, v)) return 1;
            opts.line_search.secant_iters = std::atoi(v);
        } else if (arg == "--secant-tol") {
            if (!cli::value_of(argc, argv, i, v)) return 1;
            opts.line_search.secant_tol = std::strtod(v, nullptr);
        } else if (arg == "--secant-denom-eps") {
            if (!cli::value_of(argc, argv, i, v)) return 1;
            opts.line_search.secant_denom_eps = std::strtod(v, nullptr);
        } else if (arg == "--gpu-eps") {
            if (!cli::value_of(argc, argv, i, v)) return 1;
            opts.gpu_eps = std::strtod(v, nullptr);
        } else if (arg == "--gpu-eps-decay") {
            if (!cli::value_of(argc, argv, i, v)) return 1;
            opts.gpu_eps_decay = std::strtod(v, nullptr);
        } else if (arg == "--gpu-eps-min") {
            if (!cli::value_of(argc, argv, i, v)) return 1;
            opts.gpu_eps_min = std::strtod(v, nullptr);
        } else if (arg == "--abs-gap-tol") {
            if (!cli::value_of(argc, argv, i, v)) return 1;
            opts.abs_gap_tol = std::strtod(v, nullptr);
        } else if (arg == "--rel-gap-tol") {
            if (!cli::value_of(argc, argv, i, v)) return 1;
            opts.rel_gap_tol = std::strtod(v, nullptr);
        } else {
            std::fprintf(stderr, "Unknown argument: %s\n", argv[i]);
            print_gpuboscia_example_usage(argv[0]);
            return 1;
        }
    }
    const LineSearchOptions& ls = opts.line_search;
    if (opts.algorithm == FWAlgorithmKind::FrankWolfe
            && (ls.kind == LineSearchKind::ScaledOpenLoop || ls.kind == LineSearchKind::SafeguardedScaledOpenLoop)) {
        std::fprintf(stderr, "scaled-open-loop is currently available only for DICG.\n");
        return 1;
    }
    if (ls.secant_iters <= 0 || !std::isfinite(ls.secant_tol) || ls.secant_tol <= 0.0
            || !std::isfinite(ls.secant_denom_eps) || ls.secant_denom_eps <= 0.0) {
        std::fprintf(stderr, "Secant parameters must be positive finite values.\n");
        return 1;
    }
    if (!std::isfinite(opts.gpu_eps) || opts.gpu_eps <= 0.0 || !std::isfinite(opts.gpu_eps_min)
            || opts.gpu_eps_min <= 0.0 || !std::isfinite(opts.gpu_eps_decay) || opts.gpu_eps_decay <= 0.0
            || opts.gpu_eps_decay > 1.0 || !std::isfinite(opts.abs_gap_tol) || opts.abs_gap_tol < 0.0
            || !std::isfinite(opts.rel_gap_tol) || opts.rel_gap_tol < 0.0) {
        std::fprintf(stderr, "Tolerance parameters are invalid.\n");
        return 1;
    }
    return -1;
}
