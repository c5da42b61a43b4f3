// Exposes the reference's CPU-only helpers (compiled from
// /root/reference/include/gpuboscia, never copied) with a C ABI so the
// tests can compare the oracle's restatement against the real code.
// TEST INFRASTRUCTURE ONLY; built into oracle/_ref/ by oracle/ref.mk.
#include <gpuboscia/branching.h>
#include <gpuboscia/gap_utils.h>
#include <gpuboscia/line_search_options.h>
#include <gpuboscia/lmo_dicg.h>

#include <cstdint>
#include <vector>

extern "C" {

double ref_certified_fw_lower_bound(double a, double b, double c) {
    return gpuboscia::certified_fw_lower_bound(a, b, c);
}
double ref_gap_absolute(double a, double b) { return gpuboscia::gap_absolute(a, b); }
double ref_gap_relative(double a, double b) { return gpuboscia::gap_relative(a, b); }
int ref_gap_within_tolerance(double a, double b, double c, double d) {
    return gpuboscia::gap_within_tolerance(a, b, c, d) ? 1 : 0;
}
double ref_depth_dependent_tolerance(double r, double d, double m, int depth) {
    return gpuboscia::depth_dependent_tolerance(r, d, m, depth);
}
double ref_quadratic_exact_step(double g, double c, double m) { return quadratic_exact_step(g, c, m); }

int ref_top_k_infeasible(const double* x, int n, const int* iv, int n_int, int k, double tol, int* out) {
    std::vector<double> xv(x, x + n);
    std::vector<int> ivv(iv, iv + n_int);
    auto r = top_k_infeasible(xv, ivv, k, tol);
    for (size_t i = 0; i < r.size(); ++i) out[i] = r[i];
    return (int)r.size();
}
int ref_most_infeasible(const double* x, int n, const int* iv, int n_int) {
    std::vector<double> xv(x, x + n);
    std::vector<int> ivv(iv, iv + n_int);
    return most_infeasible(xv, ivv);
}
// branch_pow2: writes children lb/ub [count][n] and ids; returns count.
int ref_branch_pow2(const double* lb, const double* ub, int n, const double* x, const int* iv,
                    int n_int, int k, double tol, int next_id, double* out_lb, double* out_ub,
                    int* out_id, int* out_next) {
    Node parent(5, -1, 0, std::vector<double>(lb, lb + n), std::vector<double>(ub, ub + n));
    parent.dual_bound = 0.25;
    std::vector<double> xv(x, x + n);
    std::vector<int> ivv(iv, iv + n_int);
    int nid = next_id;
    auto ch = branch_pow2(parent, nid, xv, ivv, k, tol);
    for (size_t c = 0; c < ch.size(); ++c) {
        for (int j = 0; j < n; ++j) {
            out_lb[c * n + j] = ch[c].lb[j];
            out_ub[c * n + j] = ch[c].ub[j];
        }
        out_id[c] = ch[c].id;
    }
    *out_next = nid;
    return (int)ch.size();
}
void ref_box_inface_lmo_cpu(int n, const double* g, const double* x, const double* lb,
                            const double* ub, double* a, double atol, double rtol) {
    std::vector<double> gv(g, g + n), xv(x, x + n), l(lb, lb + n), u(ub, ub + n), av(n);
    box_inface_lmo_cpu(gv, xv, l, u, av, atol, rtol);
    for (int j = 0; j < n; ++j) a[j] = av[j];
}
double ref_box_dicg_maximum_step_cpu(int n, const double* d, const double* x, const double* lb,
                                     const double* ub) {
    std::vector<double> dv(d, d + n), xv(x, x + n), l(lb, lb + n), u(ub, ub + n);
    return box_dicg_maximum_step_cpu(dv, xv, l, u);
}
void ref_unit_simplex_inface_lmo_cpu(int n, double rhs, const uint8_t* is_int, const double* g,
                                     const double* x, const double* lb, const double* ub,
                                     double* a, double atol, double rtol, int full) {
    std::vector<unsigned char> ii(is_int, is_int + n);
    std::vector<double> gv(g, g + n), xv(x, x + n), l(lb, lb + n), u(ub, ub + n), av(n);
    unit_simplex_inface_lmo_cpu(rhs, ii, gv, xv, l, u, av, atol, rtol, full != 0);
    for (int j = 0; j < n; ++j) a[j] = av[j];
}
double ref_unit_simplex_dicg_maximum_step_cpu(int n, double rhs, const uint8_t* is_int,
                                              const double* d, const double* x, const double* lb,
                                              const double* ub, double tol, int full) {
    std::vector<unsigned char> ii(is_int, is_int + n);
    std::vector<double> dv(d, d + n), xv(x, x + n), l(lb, lb + n), u(ub, ub + n);
    return unit_simplex_dicg_maximum_step_cpu(rhs, ii, dv, xv, l, u, tol, full != 0);
}
}
