// Host-only drivers of the example tooling headers, for the CPU tests:
//   host_tools mat FILE          every variable of a MAT v5 file
//                                ("name rows cols" then the values, row-major)
//   host_tools log               a logger session with fixed numbers
//   host_tools cli ARGS...       the parsed example options, or the parser's
//                                exit code
// Built with the host compiler (no device code, no GPU).
#include <gpuboscia/logger.hpp>
#include <gpuboscia/mat.hpp>
#include <gpuboscia/options.hpp>

#include <cstdio>
#include <exception>
#include <limits>
#include <string>

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string cmd = argv[1];
    try {
        if (cmd == "mat" && argc == 3) {
            for (const gpuboscia::mat::Matrix& m : gpuboscia::mat::read_all(argv[2])) {
                std::printf("%s %d %d\n", m.name.c_str(), m.rows, m.cols);
                for (double v : m.values) std::printf("%.17g\n", v);
            }
            return 0;
        }
        if (cmd == "log") {
            Logger log(true, 2);
            LoggerParams p;
            p.algorithm = fw_algorithm_name(FWAlgorithmKind::DICG);
            p.line_search_method = line_search_name(LineSearchKind::Secant);
            p.n_vars = 4;
            p.n_int_vars = 2;
            p.time_limit_s = 60.0;
            p.natural_lower_bound = 0.0;
            p.random_seed = 42;
            p.print_random_seed = true;
            log.print_header(p);
            const double inf = std::numeric_limits<double>::infinity();
            log.print_row(1, 1, -inf, inf, false);
            log.print_row(2, 3, -1.5, inf, false);   // skipped: interval 2
            log.print_row(3, 2, -1.25, 2.0, true);
            log.print_row(4, 2, -1.0, 2.0, false);   // skipped
            log.print_row(5, 1, 0.0, 2.0, false);
            SolutionStats s;
            s.solution_status = "Optimal";
            s.primal_objective = 2.0;
            s.dual_bound = 2.0;
            s.total_nodes = 5;
            s.total_time_s = 0.5;
            s.solution = {1.0, 0.0, 0.5, 0.0};
            s.solution_preview_size = 3;
            log.print_footer(s);
            return 0;
        }
        if (cmd == "cli") {
            GPUBosciaExampleOptions o;
            const int rc = parse_gpuboscia_example_options(argc - 1, argv + 1, o);
            if (rc >= 0) {
                std::printf("exit %d\n", rc);
                return 0;
            }
            std::printf("algorithm %s\nline_search %d %d %.17g %.17g\nseed %llu %d\n",
                        o.algorithm == FWAlgorithmKind::DICG ? "dicg" : "fw", (int)o.line_search.kind,
                        o.line_search.secant_iters, o.line_search.secant_tol, o.line_search.secant_denom_eps,
                        (unsigned long long)o.random_seed, o.random_seed_explicit ? 1 : 0);
            std::printf("eps %.17g %.17g %.17g\ngaps %.17g %.17g\ntime %.17g\nnatural %.17g\ninterval %d\n",
                        o.gpu_eps, o.gpu_eps_decay, o.gpu_eps_min, o.abs_gap_tol, o.rel_gap_tol, o.time_limit_s,
                        o.natural_lower_bound, o.log_interval_nodes);
            std::printf("criterion %c\ninstance %s\n", o.oed_criterion, o.instance_path.c_str());
            return 0;
        }
    } catch (const std::exception& e) {
        std::printf("error: %s\n", e.what());
        return 3;
    }
    return 2;
}
