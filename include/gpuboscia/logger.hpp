// Solver progress output: the reference's Logger (include/gpuboscia/logger.h:
// 22-177) with the same parameter and statistics records, the same calls and
// the same printed layout (greeting, settings block, one row per printed
// node with a '*' on new incumbents, solution and search statistics), so
// scripts that read the reference's logs read these.  Host-only C++17.
//
//   gpuboscia::Logger log(true, 100);       // enabled, a row every 100 nodes
//   log.print_header(params);
//   settings.progress = [&](int i, int open, double b, double inc, bool fresh) {
//       log.print_row(i, open, b, inc, fresh); };
//   log.print_footer(stats);
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <string>
#include <vector>

#include "gpuboscia.hpp"

namespace gpuboscia {

// gap_utils.h:23-42: |U - L|, and (U - L) / min(|L|, |U|) (infinite when
// either is zero or not finite).
inline double gap_absolute(double bound, double incumbent) {
    const double inf = std::numeric_limits<double>::infinity();
    if (incumbent >= inf || bound <= -inf) return inf;
    return std::abs(incumbent - bound);
}
inline double gap_relative(double bound, double incumbent) {
    const double inf = std::numeric_limits<double>::infinity();
    if (!std::isfinite(bound) || !std::isfinite(incumbent)) return inf;
    const double scale = std::min(std::abs(bound), std::abs(incumbent));
    return scale == 0.0 ? inf : (incumbent - bound) / scale;
}

// line_search_options.h:25-46
inline const char* fw_algorithm_name(FWAlgorithmKind kind) {
    return kind == FWAlgorithmKind::DICG ? "Decomposition-Invariant Frank-Wolfe" : "Frank-Wolfe";
}
inline const char* line_search_name(LineSearchKind kind) {
    switch (kind) {
        case LineSearchKind::ScaledOpenLoop: return "scaled open-loop gamma_max * 2/(t+2)";
        case LineSearchKind::SafeguardedScaledOpenLoop: return "scaled open-loop with full limiting steps";
        case LineSearchKind::QuadraticExact: return "quadratic exact";
        case LineSearchKind::Secant: return "secant on phi'(gamma)";
        default: return "open-loop min(gamma_max, 2/(t+2))";
    }
}

struct LoggerParams {
    std::string traversal_strategy = "best_first";
    std::string branching_strategy = "branch_pow2";
    std::string algorithm = "Frank-Wolfe";
    std::string line_search_method = "open-loop";
    int n_vars = 0;
    int n_int_vars = 0;
    double abs_gap_tol = 1e-6;
    double rel_gap_tol = 1e-2;
    double time_limit_s = std::numeric_limits<double>::infinity();
    double natural_lower_bound = -std::numeric_limits<double>::infinity();
    std::uint64_t random_seed = 0;
    bool print_random_seed = false;
};

struct SolutionStats {
    std::string solution_status = "Unknown";
    double primal_objective = std::numeric_limits<double>::infinity();
    double dual_bound = -std::numeric_limits<double>::infinity();
    int total_nodes = 0;
    double total_time_s = 0.0;
    std::vector<double> solution;
    int solution_preview_size = 12;
};

class Logger {
  public:
    bool enabled;

    explicit Logger(bool on = true, int node_interval = 1000, std::FILE* out = stdout)
        : enabled(on), interval_(node_interval > 0 ? node_interval : 1), out_(out) {}

    void print_header(const LoggerParams& p) {
        if (!enabled) return;
        start_ = Clock::now();
        last_row_ = 0;
        put("Hello:)\n");
        put("Parameter settings.\n");
        put("  Tree traversal strategy:     %s\n", p.traversal_strategy.c_str());
        put("  Branching strategy:          %s\n", p.branching_strategy.c_str());
        put("  Algorithm:                   %s\n", p.algorithm.c_str());
        put("  Line search method:          %s\n", p.line_search_method.c_str());
        if (p.print_random_seed)
            put("  Random seed:                 %llu\n", static_cast<unsigned long long>(p.random_seed));
        put("  Termination:                 abs gap <= %.3e or rel gap <= %.3e\n", p.abs_gap_tol, p.rel_gap_tol);
        if (std::isfinite(p.time_limit_s)) put("  Time limit (s):              %.3f\n", p.time_limit_s);
        if (std::isfinite(p.natural_lower_bound))
            put("  Natural lower bound:         %.6e\n", p.natural_lower_bound);
        put("  Total number of variables:   %d\n", p.n_vars);
        put("  Number of integer variables: %d\n", p.n_int_vars);
        rule();
        put("  Iter   Open           Bound       Incumbent       Gap (abs)       Gap (rel)        Time (s)\n");
        rule();
    }

    void print_stage(const char* stage) {
        if (!enabled) return;
        put("--- [%.3f s] %s ---\n", seconds(), stage);
    }

    // One call per processed node; printed for the first node, every
    // `node_interval` nodes and whenever the incumbent improves.
    void print_row(int iter, int open_nodes, double dual_bound, double incumbent, bool is_new_incumbent = false) {
        if (!enabled) return;
        if (iter != 1 && !is_new_incumbent && iter - last_row_ < interval_) return;
        last_row_ = iter;
        char b[5][32];
        put("%c%5d%7d  %s  %s  %s  %s  %s\n", is_new_incumbent ? '*' : ' ', iter, open_nodes, sci(b[0], dual_bound),
            sci(b[1], incumbent), sci(b[2], gap_absolute(dual_bound, incumbent)),
            sci(b[3], gap_relative(dual_bound, incumbent)), sci(b[4], seconds()));
    }

    void print_footer(const SolutionStats& s) {
        if (!enabled) return;
        char b[3][32];
        rule();
        put("Solution Statistics.\n");
        put("  Solution Status:              %s\n", s.solution_status.c_str());
        put("  Primal Objective:             %s\n", sci(b[0], s.primal_objective));
        put("  Dual Bound:                   %s\n", sci(b[1], s.dual_bound));
        put("  Dual Gap (relative):          %s\n", sci(b[2], gap_relative(s.dual_bound, s.primal_objective)));
        if (!s.solution.empty()) {
            const int preview = std::min((int)s.solution.size(), std::max(0, s.solution_preview_size));
            int nonzeros = 0;
            for (double v : s.solution) nonzeros += std::fabs(v) > 1e-12;
            put("  Solution dimension:           %zu\n", s.solution.size());
            put("  Solution nonzeros:             %d\n", nonzeros);
            put("  Solution preview x[0:%d]:      [", preview);
            for (int i = 0; i < preview; ++i) put("%s%.6g", i ? ", " : "", s.solution[i]);
            put("]%s\n", preview < (int)s.solution.size() ? " ..." : "");
        } else {
            put("  Solution:                      not found\n");
        }
        put("Search Statistics.\n");
        put("  Total number of nodes processed: %d\n", s.total_nodes);
        char t[32];
        put("  Total time (s):               %s\n", sci(t, s.total_time_s));
    }

  private:
    using Clock = std::chrono::steady_clock;
    Clock::time_point start_ = Clock::now();
    int interval_;
    int last_row_ = 0;
    std::FILE* out_;

    template <class... Args>
    void put(const char* fmt, Args... args) {
        std::fprintf(out_, fmt, args...);
    }
    void put(const char* s) { std::fputs(s, out_); }
    double seconds() const { return std::chrono::duration<double>(Clock::now() - start_).count(); }
    // 14-character right-aligned scientific notation, "Inf" / "-Inf" / "NaN"
    static const char* sci(char* buf, double v) {
        if (std::isnan(v)) std::snprintf(buf, 32, "%14s", "NaN");
        else if (std::isinf(v)) std::snprintf(buf, 32, "%14s", v > 0 ? "Inf" : "-Inf");
        else std::snprintf(buf, 32, "%14.6e", v);
        return buf;
    }
    void rule() {
        put("---------------------------------------------"
            "------------------------------------------------\n");
    }
};

}  // namespace gpuboscia

// The reference declares these at global scope (logger.h, line_search_options.h).
using gpuboscia::fw_algorithm_name;
using gpuboscia::line_search_name;
using gpuboscia::Logger;
using gpuboscia::LoggerParams;
using gpuboscia::SolutionStats;
