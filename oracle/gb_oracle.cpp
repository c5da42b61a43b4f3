// CPU oracle for the FW / DICG node-relaxation path.  TEST INFRASTRUCTURE
// ONLY: the product never links, loads or calls this file (see gb_oracle.h).
//
// Canonical reduction order (shared with the HIP kernels):
//   A block has `lanes` threads (a multiple of 64, i.e. whole wave64s).
//   Lane l accumulates its strided coordinates j = l, l+lanes, ... in
//   increasing j, starting from 0.0.  Each wave then tree-reduces its 64 lane
//   values with offsets 32,16,8,4,2,1 (value[i] = value[i] + value[i+off]),
//   and the wave totals are added in wave order: ((w0 + w1) + w2) + ...
//   Sums of products use fma(a, b, acc) exactly where the reference writes
//   `acc += a * b`.  Both sides are compiled with -ffp-contract=off, so no
//   other contraction happens.
//
// Pinning (what fixes each function besides the kernel comparison):
//   gap / bound / tolerance helpers ... reference KATs tests/test_cpu.cpp:41-158
//                                       and oracle/_ref (compiled reference headers)
//   top_k / branch_pow2 ............... tests/test_cpu.cpp:311-370 + oracle/_ref
//   box / simplex CPU DICG helpers .... tests/test_cpu.cpp:399-517 + oracle/_ref
//   quadratic_exact_step .............. tests/test_cpu.cpp:543-552 + oracle/_ref
//   box / simplex / Birkhoff forward LMO closed form, integer enumeration,
//                                       brute-force permutations (tests/); the
//                                       reference's LMO benchmark cases
//                                       (benchmark_lmo.cu:173-214) against scipy LP /
//                                       assignment optima; order-7 Birkhoff vs all
//                                       5040 permutations (test_pipeline.cu:193-299)
//   Hungarian tie order ............... all-zero costs give the identity, derived by
//                                       hand from the strict-< column scan
//   secant_search ..................... the reference's stall regression
//                                       (test_pipeline.cu:1133-1205) and convergence
//                                       matrix (:1062-1131); equals the exact step on
//                                       quadratics (closed form)
//   node relaxation ................... continuation KAT x=2/3, candidate 0.45 /
//                                       0.09, exact-step 0.25, interior optimum
//                                       (tests/test_pipeline.cu:1211-1441)
//   B&B ............................... exhaustive 2^n enumeration, analytic
//                                       separable optima (test_pipeline.cu:449-708)
//   portfolio LMO ..................... LP optima of tiny instances (scipy, committed
//                                       fixtures tests/golden/portfolio_lmo.json)
//   QAP objective ..................... finite differences, optimum 0 at P=Pi^T
//   batched form's 0/1 simplex ........ vertices equal the plain greedy restatement
//   (orc_fast_simplex_vertex / _inface) (compute_v / compute_inface) on random 0/1
//                                       cases with ties and signed zeros; forward
//                                       vertex = enumerated minimum; supplied-point
//                                       handoff by closed forms; relaxation values =
//                                       the plain form's; search = enumeration
#include "gb_oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <queue>
#include <thread>
#include <vector>

namespace {

constexpr double kInf = std::numeric_limits<double>::infinity();
constexpr double kNone = 1e300;  // "no candidate" score (device_resident_solver.cuh:307)

// ---------------------------------------------------------------------------
// Canonical reductions
// ---------------------------------------------------------------------------
double wave_tree_sum(double* v) {
    for (int off = 32; off > 0; off >>= 1)
        for (int i = 0; i < off; ++i) v[i] = v[i] + v[i + off];
    return v[0];
}

double lanes_sum(std::vector<double>& lane) {
    const int waves = (int)lane.size() / 64;
    double total = 0.0;
    for (int w = 0; w < waves; ++w) {
        const double s = wave_tree_sum(lane.data() + 64 * w);
        total = (w == 0) ? s : total + s;
    }
    return total;
}

double lanes_min(std::vector<double>& lane) {
    const int waves = (int)lane.size() / 64;
    double total = 0.0;
    for (int w = 0; w < waves; ++w) {
        double* v = lane.data() + 64 * w;
        for (int off = 32; off > 0; off >>= 1)
            for (int i = 0; i < off; ++i) v[i] = std::fmin(v[i], v[i + off]);
        total = (w == 0) ? v[0] : std::fmin(total, v[0]);
    }
    return total;
}

// Strided per-lane accumulation: acc_l = op(acc_l, j) over j = l, l+T, ...
template <class Op>
double block_sum(int n, int T, Op op) {
    std::vector<double> lane(T, 0.0);
    for (int l = 0; l < T; ++l) {
        double acc = 0.0;
        for (int j = l; j < n; j += T) acc = op(acc, j);
        lane[l] = acc;
    }
    return lanes_sum(lane);
}

// ---------------------------------------------------------------------------
// Objectives
// ---------------------------------------------------------------------------
// Dense quadratic: f = 1/2 s_g x'Qx + s_g s_l c'x + kappa
// (dense_quadratic_objectives.cuh:43-63, gradient :92-111).
void dense_grad(const orc_objective& o, const double* x, double* g) {
    const int n = o.n;
    for (int j = 0; j < n; ++j) {
        double acc = 0.0;
        for (int k = 0; k < n; ++k) acc = std::fma(o.Qt[(size_t)k * n + j], x[k], acc);
        const double lin = o.linear ? o.linear_scale * o.linear[j] : 0.0;
        g[j] = o.grad_scale * (acc + lin);
    }
}

double dense_value(const orc_objective& o, const double* x, int T) {
    const int n = o.n;
    std::vector<double> lane(T, 0.0);
    const double hs = 0.5 * o.grad_scale;
    const double ls = o.grad_scale * o.linear_scale;
    for (int l = 0; l < T; ++l) {
        double quad = 0.0, lin = 0.0;
        for (int j = l; j < n; j += T) {
            double qx = 0.0;
            for (int k = 0; k < n; ++k) qx = std::fma(o.Qt[(size_t)k * n + j], x[k], qx);
            quad = std::fma(x[j], qx, quad);
            if (o.linear) lin = std::fma(o.linear[j], x[j], lin);
        }
        double val = std::fma(hs, quad, ls * lin);
        if (l == 0) val = val + o.constant;
        lane[l] = val;
    }
    return lanes_sum(lane);
}

// d' Hessian d = s_g d'Qd (dense_quadratic_objectives.cuh:65-79).
double dense_curvature(const orc_objective& o, const double* d, int T) {
    const int n = o.n;
    return block_sum(n, T, [&](double acc, int j) {
        double p = 0.0;
        for (int k = 0; k < n; ++k) p = std::fma(o.Qt[(size_t)k * n + j], d[k], p);
        return std::fma(o.grad_scale * d[j], p, acc);
    });
}

// QAP: f(P) = ||A P - P B||_F^2 with P row-major m x m.  R = AP - PB,
// grad = 2 (A'R - R B').  Products accumulate in k order with fma.
void qap_residual(const orc_objective& o, const double* P, double* R) {
    const int m = o.order;
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
            double ap = 0.0, pb = 0.0;
            for (int k = 0; k < m; ++k) {
                ap = std::fma(o.A[i * m + k], P[k * m + j], ap);
                pb = std::fma(P[i * m + k], o.B[k * m + j], pb);
            }
            R[i * m + j] = ap - pb;
        }
}

void qap_grad(const orc_objective& o, const double* P, double* g) {
    const int m = o.order;
    std::vector<double> R((size_t)m * m);
    qap_residual(o, P, R.data());
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
            double atr = 0.0, rbt = 0.0;
            for (int k = 0; k < m; ++k) {
                atr = std::fma(o.A[k * m + i], R[k * m + j], atr);
                rbt = std::fma(R[i * m + k], o.B[j * m + k], rbt);
            }
            g[i * m + j] = 2.0 * (atr - rbt);
        }
}

double qap_value(const orc_objective& o, const double* P, int T) {
    const int m = o.order;
    std::vector<double> R((size_t)m * m);
    qap_residual(o, P, R.data());
    return block_sum(m * m, T, [&](double acc, int e) { return std::fma(R[e], R[e], acc); });
}

// d' Hessian d = 2 ||A D - D B||^2.
double qap_curvature(const orc_objective& o, const double* d, int T) {
    const int m = o.order;
    std::vector<double> R((size_t)m * m);
    qap_residual(o, d, R.data());
    const double s = block_sum(m * m, T, [&](double acc, int e) { return std::fma(R[e], R[e], acc); });
    return 2.0 * s;
}

void obj_grad(const orc_objective& o, const double* x, double* g) {
    if (o.kind == ORC_OBJ_QAP) qap_grad(o, x, g);
    else dense_grad(o, x, g);
}
double obj_value(const orc_objective& o, const double* x, int T) {
    return o.kind == ORC_OBJ_QAP ? qap_value(o, x, T) : dense_value(o, x, T);
}
double obj_curvature(const orc_objective& o, const double* d, int T) {
    return o.kind == ORC_OBJ_QAP ? qap_curvature(o, d, T) : dense_curvature(o, d, T);
}

// ---------------------------------------------------------------------------
// Simplex greedy allocation (lmo.cuh:517-618).  Eligible coordinates sorted
// by (gradient, index); capacities exclusive-scanned in the canonical chunked
// order below; each coordinate then receives min(max(rem - prefix, 0), cap).
// Canonical scan: S = next_pow2(n) entries, C = ceil(S/64) per lane of one
// wave.  Lane l sums its chunk in order, an inclusive Hillis-Steele scan over
// the 64 chunk totals (offsets 1,2,..,32) gives lane bases, then each chunk
// is prefixed sequentially.
// ---------------------------------------------------------------------------
int next_pow2(int v) { int r = 1; while (r < v) r <<= 1; return r; }

void chunked_exclusive_scan(const std::vector<double>& cap, std::vector<double>& prefix) {
    const int S = (int)cap.size();
    const int C = (S + 63) / 64;
    double tot[64], inc[64];
    for (int l = 0; l < 64; ++l) {
        double t = 0.0;
        for (int i = 0; i < C; ++i) {
            const int p = l * C + i;
            if (p < S) t = t + cap[p];
        }
        tot[l] = t;
    }
    for (int l = 0; l < 64; ++l) inc[l] = tot[l];
    for (int off = 1; off < 64; off <<= 1) {
        double nxt[64];
        for (int l = 0; l < 64; ++l) nxt[l] = (l >= off) ? inc[l] + inc[l - off] : inc[l];
        std::memcpy(inc, nxt, sizeof(inc));
    }
    for (int l = 0; l < 64; ++l) {
        double run = (l == 0) ? 0.0 : inc[l - 1];
        for (int i = 0; i < C; ++i) {
            const int p = l * C + i;
            if (p >= S) break;
            prefix[p] = run;
            run = run + cap[p];
        }
    }
}

void greedy_allocate(int n, const uint8_t* is_int, const double* g, const double* ub,
                     double* out, const uint8_t* fixed, double remaining,
                     bool force_full, bool floor_int, bool equality) {
    const int S = next_pow2(n);
    std::vector<int> order;
    for (int j = 0; j < n; ++j) {
        const bool eligible = (!fixed || !fixed[j]) && (force_full || g[j] <= 0.0);
        if (eligible) order.push_back(j);
    }
    std::sort(order.begin(), order.end(), [&](int a, int b) {
        return g[a] < g[b] || (g[a] == g[b] && a < b);
    });
    std::vector<double> cap(S, 0.0), prefix(S, 0.0);
    for (size_t p = 0; p < order.size(); ++p) {
        const int j = order[p];
        cap[p] = is_int[j] ? std::fmax(ub[j] - out[j], 0.0) : remaining;
    }
    chunked_exclusive_scan(cap, prefix);
    for (size_t p = 0; p < order.size(); ++p) {
        const int j = order[p];
        double add = std::fmax(remaining - prefix[p], 0.0);
        const double c = is_int[j] ? std::fmax(ub[j] - out[j], 0.0) : remaining;
        add = std::fmin(add, c);
        if (floor_int && is_int[j]) add = std::floor(add + (equality ? 1e-9 : 0.0));
        out[j] = out[j] + add;
    }
}

// ---------------------------------------------------------------------------
// Birkhoff: forced edges then shortest-augmenting-path Hungarian
// (lmo.cuh:1239-1417; CPU twin :1063-1171).  Ties in the column choice go
// to the lowest index.
// ---------------------------------------------------------------------------
bool birkhoff_solve(int m, const double* cost, const double* lb, const double* ub,
                    const double* face, double* out) {
    std::vector<int> fcol(m, -1), frow(m, -1);
    for (int r = 0; r < m; ++r) {
        for (int c = 0; c < m; ++c) {
            const int e = r * m + c;
            if (lb[e] > ub[e] + 1e-9) return false;
            const bool face_forced = face && face[e] >= 1.0 - 1e-9;
            if (lb[e] > 0.5 || face_forced) {
                if (ub[e] < 0.5 || fcol[r] >= 0 || frow[c] >= 0) return false;
                fcol[r] = c;
                frow[c] = r;
                out[e] = 1.0;
            }
        }
    }
    std::vector<int> rows, cols;
    for (int i = 0; i < m; ++i) {
        if (fcol[i] < 0) rows.push_back(i);
        if (frow[i] < 0) cols.push_back(i);
    }
    if (rows.size() != cols.size()) return false;
    const int k = (int)rows.size();
    std::vector<double> u(k + 1, 0.0), pv(k + 1, 0.0), minv(k + 1);
    std::vector<int> p(k + 1, 0), way(k + 1, 0);
    std::vector<uint8_t> used(k + 1);
    for (int i = 1; i <= k; ++i) {
        p[0] = i;
        int j0 = 0;
        std::fill(minv.begin(), minv.end(), kNone);
        std::fill(used.begin(), used.end(), 0);
        while (true) {
            used[j0] = 1;
            const int i0 = p[j0];
            for (int j = 1; j <= k; ++j) {
                if (used[j]) continue;
                const int e = rows[i0 - 1] * m + cols[j - 1];
                if (ub[e] >= 0.5 && (!face || face[e] > 1e-9)) {
                    const double cur = cost[e] - u[i0] - pv[j];
                    if (cur < minv[j]) { minv[j] = cur; way[j] = j0; }
                }
            }
            double delta = kNone;
            int j1 = -1;
            for (int j = 1; j <= k; ++j)
                if (!used[j] && minv[j] < kNone && (j1 < 0 || minv[j] < delta)) {
                    delta = minv[j];
                    j1 = j;
                }
            if (j1 < 0 || !std::isfinite(delta)) return false;
            for (int j = 0; j <= k; ++j) {
                if (used[j]) { u[p[j]] = u[p[j]] + delta; pv[j] = pv[j] - delta; }
                else minv[j] = minv[j] - delta;
            }
            j0 = j1;
            if (p[j0] == 0) break;
        }
        do { const int j1 = way[j0]; p[j0] = p[j1]; j0 = j1; } while (j0 != 0);
    }
    for (int j = 1; j <= k; ++j) out[rows[p[j] - 1] * m + cols[j - 1]] = 1.0;
    return true;
}

// ---------------------------------------------------------------------------
// Portfolio LMO {sum x = 1, 0 <= x_i <= cap z_i, lbz_i <= z_i <= ubz_i,
// sum z <= card}.  Exact LP minimum of gx'x + gz'z; see lmo_portfolio below.
// ---------------------------------------------------------------------------
}  // namespace

// ===========================================================================
// Cardinality-portfolio LMO (no reference counterpart; the nearest analogue
// is LinkedCardinalityLMO, examples/sparse_sharpe.cu:266-546, which has no
// sum x = rhs).  Variables x = v[0..m), z = v[m..2m):
//   sum x = rhs, 0 <= x_i <= cap z_i, lz_i <= z_i <= uz_i, sum z <= card.
// For a gradient with gz >= 0 (the mean-variance objective has gz = 0), z_i
// takes its least value max(lz_i, x_i / cap), and asset i offers x in two
// segments: A, up to cap lz_i at unit cost gx_i; B, from cap lz_i to cap uz_i
// at unit cost gx_i + gz_i / cap, each unit using 1 / cap of the cardinality
// budget card - sum lz.  Segment, pool-of-B and total capacities form a
// laminar family, on which filling the cheapest units first is an optimal LP
// vertex.  Segments are ordered by (unit cost, asset, A before B); negative
// gz_i count as 0 (the vertex is then exact for max(g, 0) on z).
// ===========================================================================
namespace {

void portfolio_vertex(int m, double rhs, double cap, double card, const double* g, const double* lz,
                      const double* uz, double* v) {
    const int S = 2 * m;
    std::vector<double> key(S);
    for (int i = 0; i < m; ++i) {
        key[2 * i] = g[i];
        key[2 * i + 1] = g[i] + std::fmax(g[m + i], 0.0) / cap;
    }
    std::vector<int> order(S);
    for (int q = 0; q < S; ++q) order[q] = q;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b] || (key[a] == key[b] && a < b); });
    double zsum = 0.0;
    for (int i = 0; i < m; ++i) {
        v[i] = 0.0;
        v[m + i] = lz[i];
        zsum = zsum + lz[i];
    }
    double pool = cap * (card - zsum);
    double rem = rhs;
    for (int r = 0; r < S; ++r) {
        if (!(rem > 0.0)) break;
        const int q = order[r];
        const int i = q >> 1;
        double take;
        if ((q & 1) == 0) {
            take = std::fmin(rem, cap * lz[i]);
        } else {
            take = std::fmin(std::fmin(rem, cap * (uz[i] - lz[i])), pool);
            if (take > 0.0) {
                pool = pool - take;
                v[m + i] = v[m + i] + take / cap;
            }
        }
        if (take > 0.0) {
            v[i] = v[i] + take;
            rem = rem - take;
        }
    }
}

// Largest x mass the node's bounds allow (feasibility: >= rhs).
double portfolio_capacity(int m, double cap, double card, const double* lz, const double* uz, double* zsum_out) {
    double zsum = 0.0, a = 0.0, b = 0.0;
    for (int i = 0; i < m; ++i) {
        zsum = zsum + lz[i];
        a = a + cap * lz[i];
        b = b + cap * (uz[i] - lz[i]);
    }
    *zsum_out = zsum;
    return a + std::fmin(b, cap * (card - zsum));
}


}  // namespace

// ---------------------------------------------------------------------------
// Exposed primitives
// ---------------------------------------------------------------------------
extern "C" double orc_block_sum(const double* lane_values, int lanes) {
    std::vector<double> lane(lane_values, lane_values + lanes);
    return lanes_sum(lane);
}

extern "C" void orc_grad(const orc_objective* o, const double* x, double* g) { obj_grad(*o, x, g); }
extern "C" double orc_value(const orc_objective* o, const double* x, int lanes) {
    return obj_value(*o, x, lanes);
}
extern "C" double orc_curvature(const orc_objective* o, const double* d, int lanes) {
    return obj_curvature(*o, d, lanes);
}

namespace {

double lmo_mass_sum(int n, int T, const std::vector<double>& vals) {
    return block_sum(n, T, [&](double acc, int j) { return acc + vals[j]; });
}

void compute_v(const orc_lmo& l, int n, const double* g, const double* lb,
               const double* ub, double* v, int T) {
    switch (l.kind) {
    case ORC_LMO_BOX:  // lmo.cuh:270-276
        for (int j = 0; j < n; ++j) v[j] = (g[j] > 0.0) ? lb[j] : ub[j];
        break;
    case ORC_LMO_SIMPLEX: {  // lmo.cuh:787-812
        std::vector<double> base(n);
        for (int j = 0; j < n; ++j) { v[j] = l.is_int[j] ? lb[j] : 0.0; base[j] = v[j]; }
        const double mass = lmo_mass_sum(n, T, base);
        const double remaining = l.right_side - mass;
        greedy_allocate(n, l.is_int, g, ub, v, nullptr, remaining, l.equality != 0,
                        l.equality != 0, l.equality != 0);
        break;
    }
    case ORC_LMO_BIRKHOFF:  // lmo.cuh:1000-1010
        for (int j = 0; j < n; ++j) v[j] = 0.0;
        if (n == l.order * l.order) birkhoff_solve(l.order, g, lb, ub, nullptr, v);
        break;
    case ORC_LMO_PORTFOLIO: {
        const int m = l.assets;
        portfolio_vertex(m, l.right_side, l.cap, l.card, g, lb + m, ub + m, v);
        for (int j = 2 * m; j < n; ++j) v[j] = 0.0;
        break;
    }
    }
}

void compute_inface(const orc_lmo& l, int n, const double* g, const double* x,
                    const double* lb, const double* ub, double* a, int T) {
    const double atol = 1e-6, rtol = 1e-4;
    switch (l.kind) {
    case ORC_LMO_BOX:  // lmo.cuh:278-295
        for (int j = 0; j < n; ++j) {
            const double ut = atol + rtol * std::fabs(ub[j]);
            const double lt = atol + rtol * std::fabs(lb[j]);
            if (std::fabs(x[j] - ub[j]) <= ut) a[j] = ub[j];
            else if (std::fabs(x[j] - lb[j]) <= lt) a[j] = lb[j];
            else a[j] = (g[j] > 0.0) ? lb[j] : ub[j];
        }
        break;
    case ORC_LMO_SIMPLEX: {  // lmo.cuh:814-863
        std::vector<uint8_t> fixed(n, 0);
        std::vector<double> am(n), xm(n);
        for (int j = 0; j < n; ++j) {
            a[j] = l.is_int[j] ? lb[j] : 0.0;
            const double lower = l.is_int[j] ? lb[j] : 0.0;
            const double ltol = atol + rtol * std::fmax(std::fabs(x[j]), std::fabs(lower));
            if (std::fabs(x[j] - lower) <= ltol) { fixed[j] = 1; a[j] = lower; }
            else if (l.is_int[j]) {
                const double utol = atol + rtol * std::fmax(std::fabs(x[j]), std::fabs(ub[j]));
                if (std::fabs(x[j] - ub[j]) <= utol) { fixed[j] = 1; a[j] = ub[j]; }
            }
            am[j] = a[j];
            xm[j] = x[j];
        }
        const double mass = lmo_mass_sum(n, T, am);
        const double x_mass = lmo_mass_sum(n, T, xm);
        const double remaining = l.right_side - mass;
        const bool force = l.equality || std::fabs(x_mass - l.right_side)
            <= atol + rtol * std::fmax(std::fabs(x_mass), std::fabs(l.right_side));
        greedy_allocate(n, l.is_int, g, ub, a, fixed.data(), remaining, force, true,
                        l.equality != 0);
        break;
    }
    case ORC_LMO_BIRKHOFF:  // lmo.cuh:1012-1020
        for (int j = 0; j < n; ++j) a[j] = 0.0;
        if (n == l.order * l.order) birkhoff_solve(l.order, g, lb, ub, x, a);
        break;
    case ORC_LMO_PORTFOLIO: {
        // Minimal face of x: coordinates at a bound stay there; the face LP
        // is solved by the same fixed-multiplier search restricted to the
        // free coordinates (bounds collapsed to the fixed value).
        const int m = l.assets;
        std::vector<double> flb(lb, lb + n), fub(ub, ub + n);
        for (int i = 0; i < m; ++i) {
            const double z = x[m + i];
            if (std::fabs(z - lb[m + i]) <= atol + rtol * std::fabs(lb[m + i])) fub[m + i] = lb[m + i];
            else if (std::fabs(z - ub[m + i]) <= atol + rtol * std::fabs(ub[m + i])) flb[m + i] = ub[m + i];
        }
        compute_v(l, n, g, flb.data(), fub.data(), a, T);
        break;
    }
    }
}

// Per-coordinate DICG step limit, block-min reduced (lmo.cuh:299-313,
// :865-884, :1022-1036), combined with the aggregate limit (:315-320,
// :888-906) as device_resident_solver.cuh:533-545 does.
double max_step(const orc_lmo& l, int n, const double* d, const double* x,
                const double* lb, const double* ub, int T) {
    std::vector<double> lane(T, 1.0);
    for (int t = 0; t < T; ++t) {
        double r = 1.0;
        for (int j = t; j < n; j += T) {
            const double dj = d[j];
            if (l.kind == ORC_LMO_SIMPLEX) {
                const double tol = 1e-6;
                if (dj > tol) r = std::fmin(r, (x[j] - (l.is_int[j] ? lb[j] : 0.0)) / dj);
                else if (dj < -tol) r = std::fmin(r, ((l.is_int[j] ? ub[j] : l.right_side) - x[j]) / -dj);
            } else if (l.kind == ORC_LMO_PORTFOLIO) {
                const int m = l.assets;
                double lo = lb[j], hi = ub[j];
                if (j < m) { lo = 0.0; hi = l.cap; }
                if (dj > 0.0) r = std::fmin(r, (x[j] - lo) / dj);
                else if (dj < 0.0) r = std::fmin(r, (hi - x[j]) / -dj);
            } else {
                if (l.kind == ORC_LMO_BOX) {
                    if (dj < 0.0) r = std::fmin(r, (ub[j] - x[j]) / -dj);
                    else if (dj > 0.0) r = std::fmin(r, (x[j] - lb[j]) / dj);
                } else {
                    if (dj > 0.0) r = std::fmin(r, (x[j] - lb[j]) / dj);
                    else if (dj < 0.0) r = std::fmin(r, (ub[j] - x[j]) / -dj);
                }
            }
        }
        lane[t] = r;
    }
    const double coord = lanes_min(lane);
    double aggregate = 1.0;
    if (l.kind == ORC_LMO_SIMPLEX) {
        std::vector<double> xv(x, x + n), dv(d, d + n);
        const double sx = lmo_mass_sum(n, T, xv);
        const double sd = lmo_mass_sum(n, T, dv);
        if (l.equality && std::fabs(sd) > 1e-9) aggregate = 0.0;
        else aggregate = sd < 0.0 ? std::fmin(1.0, (sx - l.right_side) / sd) : 1.0;
    } else if (l.kind == ORC_LMO_PORTFOLIO) {
        // x - gamma d must keep x_i <= cap z_i and sum z <= card.
        const int m = l.assets;
        double lim = 1.0;
        for (int i = 0; i < m; ++i) {
            const double slack = l.cap * x[m + i] - x[i];
            const double rate = l.cap * d[m + i] - d[i];   // slack(gamma) = slack - gamma*rate
            if (rate > 0.0) lim = std::fmin(lim, slack / rate);
        }
        double sz = 0.0, sdz = 0.0;
        for (int i = 0; i < m; ++i) { sz = sz + x[m + i]; sdz = sdz + d[m + i]; }
        if (sdz < 0.0) lim = std::fmin(lim, (sz - l.card) / sdz);
        aggregate = lim;
    }
    return std::fmax(0.0, std::fmin(coord, aggregate));
}

bool node_feasible(const orc_lmo& l, int n, const double* lb, const double* ub,
                   const uint8_t* flags, int T) {
    // device_resident_solver.cuh:206-217
    for (int j = 0; j < n; ++j)
        if (std::isnan(lb[j]) || std::isnan(ub[j]) || lb[j] > ub[j]
                || (flags && flags[j] && std::ceil(lb[j]) > std::floor(ub[j])))
            return false;
    switch (l.kind) {
    case ORC_LMO_BOX: return true;
    case ORC_LMO_SIMPLEX: {  // lmo.cuh:373-397
        std::vector<double> lo(n), hi(n);
        bool valid = true, cont = false;
        for (int j = 0; j < n; ++j) {
            const bool integer = flags ? flags[j] != 0 : l.is_int[j] != 0;
            lo[j] = integer ? std::ceil(std::fmax(0.0, lb[j])) : std::fmax(0.0, lb[j]);
            hi[j] = integer ? std::floor(ub[j]) : ub[j];
            valid = valid && lo[j] <= hi[j];
            cont = cont || !integer;
        }
        const double mand = lmo_mass_sum(n, T, lo);
        bool feas = valid && mand <= l.right_side + 1e-9;
        if (l.equality) {
            const double capsum = lmo_mass_sum(n, T, hi);
            feas = feas && capsum >= l.right_side - 1e-9
                && (cont || std::fabs(l.right_side - std::round(l.right_side)) <= 1e-9);
        }
        return feas;
    }
    case ORC_LMO_BIRKHOFF: {  // lmo.cuh:947-956
        if (n != l.order * l.order) return false;
        std::vector<double> zero(n, 0.0), out(n, 0.0);
        return birkhoff_solve(l.order, zero.data(), lb, ub, nullptr, out.data());
    }
    case ORC_LMO_PORTFOLIO: {
        const int m = l.assets;
        if (n != 2 * m) return false;
        double zsum = 0.0;
        const double capacity = portfolio_capacity(m, l.cap, l.card, lb + m, ub + m, &zsum);
        return zsum <= l.card + 1e-9 && capacity >= l.right_side - 1e-9;
    }
    }
    return false;
}

// device_resident_solver.cuh:264-293
bool rounded_feasible(const orc_lmo& l, int n, const double* x, double tol, int T) {
    switch (l.kind) {
    case ORC_LMO_BOX: return true;
    case ORC_LMO_SIMPLEX: {
        bool valid = true;
        std::vector<double> xv(x, x + n);
        for (int j = 0; j < n; ++j) valid = valid && x[j] >= -tol;
        const double s = lmo_mass_sum(n, T, xv);
        return valid && (l.equality ? std::fabs(s - l.right_side) <= tol : s <= l.right_side + tol);
    }
    case ORC_LMO_BIRKHOFF: {
        const int m = l.order;
        if (n != m * m) return false;
        bool valid = true;
        for (int i = 0; i < m; ++i) {
            double row = 0.0, col = 0.0;
            for (int j = 0; j < m; ++j) {
                if (x[i * m + j] < -tol) valid = false;
                row = row + x[i * m + j];
                col = col + x[j * m + i];
            }
            if (std::fabs(row - 1.0) > tol || std::fabs(col - 1.0) > tol) valid = false;
        }
        return valid;
    }
    case ORC_LMO_PORTFOLIO: {
        const int m = l.assets;
        double xs = 0.0, zs = 0.0;
        for (int i = 0; i < m; ++i) {
            if (x[i] < -tol || x[i] > l.cap * x[m + i] + tol) return false;
            xs = xs + x[i];
            zs = zs + x[m + i];
        }
        return std::fabs(xs - l.right_side) <= tol && zs <= l.card + tol;
    }
    }
    return false;
}

// device_resident_solver.cuh:297-313
double integer_value(const orc_objective& o, int n, const double* v, const double* lb,
                     const double* ub, const uint8_t* flags, double tol, int T) {
    for (int j = 0; j < n; ++j)
        if (!std::isfinite(v[j]) || v[j] < lb[j] - tol || v[j] > ub[j] + tol
                || (flags && flags[j] && v[j] != std::round(v[j])))
            return kNone;
    const double val = obj_value(o, v, T);
    return std::isfinite(val) ? val : kNone;
}

struct NodeOut {
    double gap = 0.0, value = 0.0, best_val = kNone;
    int iterations = 0;
    uint8_t status = 0;
};

// Secant line search on phi'(gamma) = grad f(x + gamma d) . d
// (algorithms.cuh:121-258).
double secant_search(const orc_objective& o, const orc_alg& a, int n, int T, double gap,
                     double max_gamma, const double* x, const double* dir,
                     std::vector<double>& tx, std::vector<double>& tg) {
    const double tol = a.secant_tol;
    const double f_base = obj_value(o, x, T);
    double lo = 0.0, hi = max_gamma, best_f = f_base, best_gamma = 0.0;
    double phi_lo = -gap, phi_hi, gamma = 0.0;
    bool done = phi_lo >= -tol;
    auto trial = [&](double gm, double& phi, double& f) {
        for (int j = 0; j < n; ++j) tx[j] = std::fma(gm, dir[j], x[j]);
        obj_grad(o, tx.data(), tg.data());
        phi = block_sum(n, T, [&](double acc, int j) { return std::fma(tg[j], dir[j], acc); });
        f = obj_value(o, tx.data(), T);
    };
    double f_hi;
    trial(max_gamma, phi_hi, f_hi);
    {
        const double f_abs = f_base >= 0.0 ? f_base : -f_base;
        const double f_allow = tol * (1.0 + f_abs);
        if (f_hi < best_f || (best_gamma == 0.0 && f_hi <= f_base + f_allow)) {
            best_f = f_hi;
            best_gamma = max_gamma;
        }
        if (!done && phi_hi <= tol && f_hi <= f_base + f_allow) { gamma = max_gamma; done = true; }
    }
    for (int it = 0; it < a.secant_iters; ++it) {
        if (done) break;
        const double denom = phi_hi - phi_lo;
        double gm = (denom > a.secant_denom_eps || denom < -a.secant_denom_eps)
            ? hi - phi_hi * (hi - lo) / denom : 0.5 * (lo + hi);
        if (!(gm > lo && gm < hi)) gm = 0.5 * (lo + hi);
        gamma = gm;
        double phi, ft;
        trial(gamma, phi, ft);
        const double f_abs = f_base >= 0.0 ? f_base : -f_base;
        const double f_allow = tol * (1.0 + f_abs);
        if (ft < best_f || (best_gamma == 0.0 && ft <= f_base + f_allow)) { best_f = ft; best_gamma = gamma; }
        if (phi < 0.0) { lo = gamma; phi_lo = phi; } else { hi = gamma; phi_hi = phi; }
        if ((phi <= tol && phi >= -tol) || (hi - lo) <= tol) { gamma = 0.5 * (lo + hi); done = true; }
    }
    const double f_abs = f_base >= 0.0 ? f_base : -f_base;
    const double f_allow = tol * (1.0 + f_abs);
    if (best_gamma > 0.0 && best_f <= f_base + f_allow) gamma = best_gamma;
    else if (!done) gamma = 0.0;
    return gamma;
}

// line_search_options.h:48-63
double quad_exact_step(double gap, double curvature, double maximum) {
    if (maximum <= 0.0) return 0.0;
    if (curvature > 0.0) {
        if (gap <= 0.0) return 0.0;
        const double step = gap / curvature;
        return step < maximum ? step : maximum;
    }
    return -gap + 0.5 * curvature * maximum < 0.0 ? maximum : 0.0;
}

// One node: device_resident_solver.cuh:337-629.
void relax_node(const orc_objective& o, const orc_alg& a, const orc_lmo& l, int n,
                int max_iters, double* x, const double* lb, const double* ub, double eps,
                int& iterations, uint8_t& status, double& out_gap, double& out_value,
                double* best_v, double& best_val, const uint8_t* flags, double ctol, int T) {
    const int initial = iterations;
    const bool supplied = (status & 0x80) != 0;
    best_val = kNone;
    double best_s = kNone;
    if (!node_feasible(l, n, lb, ub, flags, T)) {
        status = 2;
        out_value = kInf;
        out_gap = 0.0;
        return;
    }
    std::vector<double> g(n), v(n), aw(n), tx(n), tg(n);
    auto track = [&](const double* vert) {
        const double fv = integer_value(o, n, vert, lb, ub, flags, ctol, T);
        if (fv < best_s) { best_s = fv; best_val = fv; }
        if (fv < 1e299 && fv == best_s) std::copy(vert, vert + n, best_v);
    };
    if (initial == 0 && !supplied) {
        for (int j = 0; j < n; ++j) g[j] = -1.0 - (double)j / (n + 1.0);
        compute_v(l, n, g.data(), lb, ub, v.data(), T);
        std::copy(v.begin(), v.end(), x);
    }
    const bool secant_like = a.ls == ORC_LS_SECANT || a.ls == ORC_LS_QUADRATIC_EXACT;
    for (int iter = 0; iter < max_iters; ++iter) {
        iterations = initial + iter + 1;
        obj_grad(o, x, g.data());
        compute_v(l, n, g.data(), lb, ub, v.data(), T);
        track(v.data());
        const double gap = block_sum(n, T, [&](double acc, int j) { return std::fma(g[j], x[j] - v[j], acc); });
        const bool done = gap >= 0.0 && gap < eps;
        out_gap = gap;
        status = done ? 1 : 0;
        if (done) break;
        if (!a.dicg) {
            if (secant_like) {
                for (int j = 0; j < n; ++j) v[j] = v[j] - x[j];
                double gamma;
                if (a.ls == ORC_LS_SECANT) gamma = secant_search(o, a, n, T, gap, 1.0, x, v.data(), tx, tg);
                else gamma = quad_exact_step(gap, obj_curvature(o, v.data(), T), 1.0);
                for (int j = 0; j < n; ++j) x[j] = std::fma(gamma, v[j], x[j]);
            } else {
                const double gamma = 2.0 / (double)(initial + iter + 2);
                for (int j = 0; j < n; ++j) x[j] = std::fma(gamma, v[j] - x[j], x[j]);
            }
            continue;
        }
        for (int j = 0; j < n; ++j) g[j] = -g[j];
        compute_inface(l, n, g.data(), x, lb, ub, aw.data(), T);
        track(aw.data());
        for (int j = 0; j < n; ++j) { g[j] = -g[j]; aw[j] = aw[j] - v[j]; }
        const double gmax = max_step(l, n, aw.data(), x, lb, ub, T);
        if (secant_like) {
            for (int j = 0; j < n; ++j) aw[j] = -aw[j];
            const double pgap = block_sum(n, T, [&](double acc, int j) { return std::fma(-g[j], aw[j], acc); });
            double gamma;
            if (a.ls == ORC_LS_SECANT) gamma = secant_search(o, a, n, T, pgap, gmax, x, aw.data(), tx, tg);
            else gamma = quad_exact_step(pgap, obj_curvature(o, aw.data(), T), gmax);
            for (int j = 0; j < n; ++j) x[j] = std::fma(gamma, aw[j], x[j]);
        } else {
            const double ol = 2.0 / (double)(initial + iter + 2);
            double gamma;
            if (a.ls == ORC_LS_SAFEGUARDED) gamma = gmax <= ol ? gmax : gmax * ol;
            else if (a.ls == ORC_LS_SCALED) gamma = gmax * ol;
            else gamma = std::fmin(gmax, ol);
            for (int j = 0; j < n; ++j) x[j] = std::fma(-gamma, aw[j], x[j]);
        }
    }
    // Final certificate at the returned iterate (device_resident_solver.cuh:588-621).
    obj_grad(o, x, g.data());
    compute_v(l, n, g.data(), lb, ub, v.data(), T);
    track(v.data());
    const double fgap = block_sum(n, T, [&](double acc, int j) { return std::fma(g[j], x[j] - v[j], acc); });
    const double fval = obj_value(o, x, T);
    out_value = fval;
    out_gap = fgap;
    status = (fgap >= 0.0 && fgap < eps) ? 1 : 0;
    // Rounding heuristic (heuristics.h:23-35; device_resident_solver.cuh:315-335).
    if (flags) {
        for (int j = 0; j < n; ++j)
            v[j] = !flags[j] ? x[j]
                 : (a.heuristic == 1 ? std::floor(x[j]) + a.heuristic_offset : std::round(x[j]));
        if (rounded_feasible(l, n, v.data(), ctol, T)) {
            const double cv = integer_value(o, n, v.data(), lb, ub, flags, ctol, T);
            if (cv < best_val) { std::copy(v.begin(), v.end(), best_v); best_val = cv; }
        }
    }
    bool integral = true;
    for (int j = 0; j < n; ++j)
        if (!std::isfinite(x[j]) || (flags && flags[j] && std::fabs(x[j] - std::round(x[j])) > ctol))
            integral = false;
    if (integral) status = status ? 4 : 3;
}

}  // namespace

extern "C" void orc_compute_v(const orc_lmo* l, int n, const double* g, const double* lb,
                              const double* ub, double* v, int lanes) {
    compute_v(*l, n, g, lb, ub, v, lanes);
}
extern "C" void orc_compute_inface(const orc_lmo* l, int n, const double* g, const double* x,
                                   const double* lb, const double* ub, double* a, int lanes) {
    compute_inface(*l, n, g, x, lb, ub, a, lanes);
}
extern "C" double orc_max_step(const orc_lmo* l, int n, const double* d, const double* x,
                               const double* lb, const double* ub, int lanes) {
    return max_step(*l, n, d, x, lb, ub, lanes);
}
extern "C" int orc_node_feasible(const orc_lmo* l, int n, const double* lb, const double* ub,
                                 const uint8_t* flags, int lanes) {
    return node_feasible(*l, n, lb, ub, flags, lanes) ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Host helpers (gap_utils.h:14-62, line_search_options.h:48-63,
// branching.h:24-118, lmo_dicg.h:8-154)
// ---------------------------------------------------------------------------
extern "C" double orc_quadratic_exact_step(double gap, double curvature, double maximum) {
    return quad_exact_step(gap, curvature, maximum);
}
extern "C" double orc_certified_fw_lower_bound(double inherited, double primal, double gap) {
    if (!std::isfinite(primal) || !std::isfinite(gap) || gap < 0.0) return inherited;
    return std::max(inherited, primal - gap);
}
extern "C" double orc_gap_absolute(double bound, double incumbent) {
    if (incumbent >= kInf || bound <= -kInf) return kInf;
    return std::fabs(incumbent - bound);
}
extern "C" double orc_gap_relative(double bound, double incumbent) {
    if (!std::isfinite(bound) || !std::isfinite(incumbent)) return kInf;
    const double scale = std::min(std::fabs(bound), std::fabs(incumbent));
    if (scale == 0.0) return kInf;
    return (incumbent - bound) / scale;
}
extern "C" int orc_gap_within_tolerance(double bound, double incumbent, double abs_tol, double rel_tol) {
    if (!std::isfinite(bound) || !std::isfinite(incumbent)) return 0;
    return (orc_gap_absolute(bound, incumbent) <= abs_tol
            || orc_gap_relative(bound, incumbent) <= rel_tol) ? 1 : 0;
}
extern "C" double orc_depth_dependent_tolerance(double root, double decay, double minimum, int depth) {
    if (depth <= 0) return std::max(root, minimum);
    return std::max(minimum, root * std::pow(decay, depth));
}

namespace {
std::vector<int> top_k(const double* x, const int* iv, int n_int, int k, double tol) {
    struct Cand { int idx; double dist; };
    std::vector<Cand> c;
    for (int t = 0; t < n_int; ++t) {
        const int i = iv[t];
        const double frac = x[i] - std::floor(x[i]);
        const double dist = std::min(frac, 1.0 - frac);
        if (dist > tol) c.push_back({i, dist});
    }
    // Same comparator and container as branching.h:58-60 so that the
    // (unstable) std::sort orders ties identically.
    std::sort(c.begin(), c.end(), [](const Cand& a, const Cand& b) { return a.dist > b.dist; });
    const int m = std::min(k, (int)c.size());
    std::vector<int> r(m);
    for (int i = 0; i < m; ++i) r[i] = c[i].idx;
    return r;
}
}  // namespace

extern "C" int orc_top_k_infeasible(const double* x, const int* int_vars, int n_int, int k,
                                    double integer_tol, int* out) {
    const auto r = top_k(x, int_vars, n_int, k, integer_tol);
    for (size_t i = 0; i < r.size(); ++i) out[i] = r[i];
    return (int)r.size();
}
extern "C" int orc_most_infeasible(const double* x, const int* int_vars, int n_int) {
    int best = -1;
    double bd = -1.0;
    for (int t = 0; t < n_int; ++t) {
        const int i = int_vars[t];
        const double frac = x[i] - std::floor(x[i]);
        const double dist = std::min(frac, 1.0 - frac);
        if (dist > bd) { bd = dist; best = i; }
    }
    return best;
}

namespace {
bool isapprox(double a, double b, double atol, double rtol) {
    return std::fabs(a - b) <= atol + rtol * std::max(std::fabs(a), std::fabs(b));
}
}  // namespace

extern "C" void orc_box_inface_lmo_cpu(int n, const double* grad, const double* x, const double* lb,
                                       const double* ub, double* a, double atol, double rtol) {
    for (int j = 0; j < n; ++j) {
        if (isapprox(x[j], ub[j], atol, rtol)) a[j] = ub[j];
        else if (isapprox(x[j], lb[j], atol, rtol)) a[j] = lb[j];
        else a[j] = grad[j] > 0.0 ? lb[j] : ub[j];
    }
}
extern "C" double orc_box_dicg_maximum_step_cpu(int n, const double* d, const double* x,
                                                const double* lb, const double* ub) {
    double gm = 1.0;
    for (int j = 0; j < n; ++j) {
        if (d[j] < 0.0) gm = std::min(gm, (ub[j] - x[j]) / -d[j]);
        else if (d[j] > 0.0) gm = std::min(gm, (x[j] - lb[j]) / d[j]);
    }
    return gm;
}
extern "C" void orc_unit_simplex_inface_lmo_cpu(int n, double rhs, const uint8_t* is_int,
        const double* grad, const double* x, const double* lb, const double* ub, double* a,
        double atol, double rtol, int require_full_mass) {
    std::vector<uint8_t> fixed(n, 0), used(n, 0);
    double x_mass = 0.0;
    for (int j = 0; j < n; ++j) x_mass += x[j];
    const bool force = require_full_mass || isapprox(x_mass, rhs, atol, rtol);
    double mass = 0.0;
    for (int j = 0; j < n; ++j) {
        a[j] = is_int[j] ? lb[j] : 0.0;
        if (is_int[j]) {
            if (isapprox(x[j], lb[j], atol, rtol)) { fixed[j] = 1; a[j] = lb[j]; }
            else if (isapprox(x[j], ub[j], atol, rtol)) { fixed[j] = 1; a[j] = ub[j]; }
        } else if (isapprox(x[j], 0.0, atol, rtol)) {
            fixed[j] = 1;
        }
        mass += a[j];
    }
    if (isapprox(mass, rhs, atol, rtol)) return;
    double remaining = rhs - mass;
    while (remaining > 0.0) {
        int best = -1;
        for (int j = 0; j < n; ++j) {
            if (fixed[j] || used[j] || (!force && grad[j] > 0.0)) continue;
            if (best < 0 || grad[j] < grad[best] || (grad[j] == grad[best] && j < best)) best = j;
        }
        if (best < 0) break;
        used[best] = 1;
        double add = remaining;
        if (is_int[best]) add = std::min(std::max(ub[best] - a[best], 0.0), std::floor(remaining));
        a[best] += add;
        double s = 0.0;
        for (int j = 0; j < n; ++j) s += a[j];
        remaining = rhs - s;
        if (isapprox(remaining, 0.0, atol, rtol)) return;
    }
}
extern "C" double orc_unit_simplex_dicg_maximum_step_cpu(int n, double rhs, const uint8_t* is_int,
        const double* d, const double* x, const double* lb, const double* ub, double tol,
        int require_full_mass) {
    double gm = 1.0, sx = 0.0, sd = 0.0;
    for (int j = 0; j < n; ++j) {
        if (d[j] > tol) gm = std::min(gm, (x[j] - (is_int[j] ? lb[j] : 0.0)) / d[j]);
        else if (d[j] < -tol) gm = std::min(gm, ((is_int[j] ? ub[j] : rhs) - x[j]) / -d[j]);
        if (std::fabs(gm) <= tol) return 0.0;
        sx += x[j];
        sd += d[j];
    }
    if (require_full_mass && std::fabs(sd) > tol) return 0.0;
    if (sd < 0.0) gm = std::min(gm, (sx - rhs) / sd);
    return gm;
}
extern "C" int orc_birkhoff_solve(int order, const double* cost, const double* lb,
                                  const double* ub, const double* face, double* out) {
    return birkhoff_solve(order, cost, lb, ub, face, out) ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Batch relaxation (node-parallel over `threads` host threads)
// ---------------------------------------------------------------------------
extern "C" void orc_relax_batch(const orc_objective* o, const orc_alg* a, const orc_lmo* l,
                                int N, int max_iters, double* x, const double* lb,
                                const double* ub, const double* eps, int* iterations,
                                uint8_t* status, double* gap, double* value,
                                double* best_vertex, double* best_vertex_val,
                                const uint8_t* flags, double candidate_tol, int lanes,
                                int threads) {
    const int n = o->n;
    auto work = [&](int i) {
        relax_node(*o, *a, *l, n, max_iters, x + (size_t)i * n, lb + (size_t)i * n,
                   ub + (size_t)i * n, eps[i], iterations[i], status[i], gap[i], value[i],
                   best_vertex + (size_t)i * n, best_vertex_val[i], flags, candidate_tol, lanes);
    };
    if (threads <= 1 || N <= 1) {
        for (int i = 0; i < N; ++i) work(i);
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < std::min(threads, N); ++t)
        pool.emplace_back([&] {
            for (int i = next.fetch_add(1); i < N; i = next.fetch_add(1)) work(i);
        });
    for (auto& th : pool) th.join();
}

// ---------------------------------------------------------------------------
// Batched-form relaxation (dense objective over a box, or over a bounded
// simplex of 0/1 variables), mirroring the wave64 kernel
// gpuboscia_amd/csrc/relax_fast.hip operation for operation: one wave (64
// lanes) per node, incremental FP64 products, exact quadratic step, cutoff
// stop, final direct certificate.  No reference counterpart: pinned by the
// same KATs as the plain form (optima, enumeration), by the plain form's
// greedy for the simplex vertices, and by the kernel.
// ---------------------------------------------------------------------------
namespace {

struct FastNode {
    const orc_objective& o;
    int n;
    std::vector<double> lsc;
    double hs;

    void gemv(const std::vector<double>& z, std::vector<double>& y) const {
        for (int j = 0; j < n; ++j) {
            double acc = 0.0;
            for (int k = 0; k < n; ++k) acc = std::fma(o.Qt[(size_t)k * n + j], z[k], acc);
            y[j] = acc;
        }
    }
    void update(const std::vector<double>& zn, const std::vector<double>& zo,
                std::vector<double>& y) const {
        for (int k = 0; k < n; ++k) {  // k = lane + 64 t, t-major then lane: ascending k
            if (zn[k] == zo[k]) continue;
            const double delta = zn[k] - zo[k];
            for (int j = 0; j < n; ++j) y[j] = std::fma(o.Qt[(size_t)k * n + j], delta, y[j]);
        }
    }
    double dot_value(const std::vector<double>& z, const std::vector<double>& qz) const {
        return block_sum(n, 64, [&](double acc, int j) {
            return std::fma(z[j], std::fma(hs, qz[j], lsc[j]), acc);
        });
    }
};

}  // namespace

// The batched form's simplex vertices (0/1 data: every capacity 0 or 1).
// The greedy of lmo.cuh:517-618 restated as a sort of the eligible
// coordinates by (direction, index); the first rem of them move to ub.
namespace {
void greedy01(int n, const double* h, const std::vector<char>& elig, double rem, const double* ub, double* out) {
    std::vector<int> c;
    for (int j = 0; j < n; ++j)
        if (elig[j]) c.push_back(j);
    std::sort(c.begin(), c.end(), [&](int a, int b) { return h[a] < h[b] || (h[a] == h[b] && a < b); });
    const size_t take = rem >= (double)c.size() ? c.size() : (rem >= 1.0 ? (size_t)rem : 0);
    for (size_t i = 0; i < take; ++i) out[c[i]] = ub[c[i]];
}
}  // namespace

// Forward vertex for direction h: v = lb, eligible = capacity 1 and h <= 0
// (every coordinate for an equality), rem = rhs - sum lb (block order).
extern "C" void orc_fast_simplex_vertex(int n, const double* h, const double* lb, const double* ub, double rhs,
                                        int equality, double* out) {
    std::vector<char> elig(n);
    for (int j = 0; j < n; ++j) {
        out[j] = lb[j];
        elig[j] = ub[j] > lb[j] && (equality || h[j] <= 0.0);
    }
    const double lbsum = block_sum(n, 64, [&](double acc, int j) { return acc + std::fmax(0.0, lb[j]); });
    greedy01(n, h, elig, rhs - lbsum, ub, out);
}

// In-face vertex for direction h (the away vertex takes h = -g) at x
// (lmo.cuh:814-863): coordinates at a bound (atol 1e-6, rtol 1e-4) stay,
// the others start at lb; the face keeps the full mass when sum x == rhs.
extern "C" void orc_fast_simplex_inface(int n, const double* h, const double* x, const double* lb,
                                        const double* ub, double rhs, int equality, double* out) {
    const double atol = 1e-6, rtol = 1e-4;
    std::vector<char> fixed(n, 0), elig(n);
    for (int j = 0; j < n; ++j) {
        out[j] = lb[j];
        if (std::fabs(x[j] - lb[j]) <= atol + rtol * std::fmax(std::fabs(x[j]), std::fabs(lb[j]))) {
            fixed[j] = 1;
        } else if (std::fabs(x[j] - ub[j]) <= atol + rtol * std::fmax(std::fabs(x[j]), std::fabs(ub[j]))) {
            fixed[j] = 1;
            out[j] = ub[j];
        }
    }
    const double mass = block_sum(n, 64, [&](double acc, int j) { return acc + out[j]; });
    const double xmass = block_sum(n, 64, [&](double acc, int j) { return acc + x[j]; });
    const bool force = equality || std::fabs(xmass - rhs) <= atol + rtol * std::fmax(std::fabs(xmass), std::fabs(rhs));
    for (int j = 0; j < n; ++j) elig[j] = !fixed[j] && ub[j] > out[j] && (force || h[j] <= 0.0);
    greedy01(n, h, elig, rhs - mass, ub, out);
}

extern "C" void orc_relax_fast_batch_lmo(const orc_objective* op, const orc_lmo* lm, int dicg, int N,
                                         int max_iters, double* X, const double* LB, const double* UB,
                                         const double* EPS, int* ITERS, uint8_t* STATUS, double* GAP,
                                         double* VALUE, double* BV, double* BVV, const uint8_t* flags,
                                         double ctol, const double* cutoff, int refresh, int rescore,
                                         int threads) {
    const orc_objective& o = *op;
    const int n = o.n;
    // Bounded simplex over 0/1 variables (the batched form's scope): the
    // greedy of lmo.cuh:517-618 with every capacity 0 or 1, restated as a
    // sort of the eligible coordinates by (direction, index).
    const bool sx = lm && lm->kind == ORC_LMO_SIMPLEX;
    const double rhs = sx ? lm->right_side : 0.0;
    const bool eq = sx && lm->equality;
    auto work = [&](int node) {
        FastNode F{o, n, std::vector<double>(n, 0.0), 0.5 * o.grad_scale};
        const double lsc_scale = o.grad_scale * o.linear_scale;
        for (int j = 0; j < n; ++j) F.lsc[j] = o.linear ? lsc_scale * o.linear[j] : 0.0;
        const double* lb = LB + (size_t)node * n;
        const double* ub = UB + (size_t)node * n;
        double* xn = X + (size_t)node * n;
        double* bvn = BV + (size_t)node * n;
        const double eps = EPS[node];
        const double cut = cutoff ? cutoff[node] : kInf;
        const int initial = ITERS[node];
        const bool supplied = (STATUS[node] & 0x80) != 0;
        for (int j = 0; j < n; ++j)
            if (std::isnan(lb[j]) || std::isnan(ub[j]) || lb[j] > ub[j]
                    || (flags && flags[j] && std::ceil(lb[j]) > std::floor(ub[j]))) {
                STATUS[node] = 2;
                VALUE[node] = kInf;
                GAP[node] = 0.0;
                BVV[node] = kNone;
                ITERS[node] = initial;
                return;
            }
        if (sx) {
            bool binary = true, cont = false;
            for (int j = 0; j < n; ++j) {
                if (!lm->is_int[j] || (lb[j] != 0.0 && lb[j] != 1.0) || (ub[j] != 0.0 && ub[j] != 1.0)) binary = false;
                if (flags && !flags[j]) cont = true;
            }
            if (!binary) {
                STATUS[node] = 0xFF;
                VALUE[node] = std::nan("");
                GAP[node] = std::nan("");
                BVV[node] = kNone;
                ITERS[node] = initial;
                return;
            }
            // Lmo<SIMPLEX>::feasible (lmo.cuh:373-397)
            const double ms = block_sum(n, 64, [&](double acc, int j) { return acc + std::fmax(0.0, lb[j]); });
            bool feas = ms <= rhs + 1e-9;
            if (eq) {
                const double cs = block_sum(n, 64, [&](double acc, int j) { return acc + ub[j]; });
                feas = feas && cs >= rhs - 1e-9 && (cont || std::fabs(rhs - std::round(rhs)) <= 1e-9);
            }
            if (!feas) {
                STATUS[node] = 2;
                VALUE[node] = kInf;
                GAP[node] = 0.0;
                BVV[node] = kNone;
                ITERS[node] = initial;
                return;
            }
        }
        auto sx_vertex = [&](const std::vector<double>& h, std::vector<double>& out) {
            orc_fast_simplex_vertex(n, h.data(), lb, ub, rhs, eq, out.data());
        };
        double best = kNone;
        auto score = [&](const std::vector<double>& z, const std::vector<double>& qz) {
            bool ok = true;
            for (int j = 0; j < n; ++j)
                if (!std::isfinite(z[j]) || z[j] < lb[j] - ctol || z[j] > ub[j] + ctol
                        || (flags && flags[j] && z[j] != std::round(z[j])))
                    ok = false;
            const double s = F.dot_value(z, qz);
            if (!ok) return;
            const double f = s + o.constant;
            if (!std::isfinite(f)) return;
            if (f < best) { best = f; std::copy(z.begin(), z.end(), bvn); }
        };
        std::vector<double> x(n), qx(n), v(n), qv(n), a(n), qa(n), g(n), vn(n), an(n), d(n), qd(n);
        if (rescore && BVV[node] < kNone) {
            // candidate of an earlier (mixed-precision) phase, re-scored exactly
            std::vector<double> z(bvn, bvn + n), qz(n);
            F.gemv(z, qz);
            score(z, qz);
        }
        if (sx && initial == 0 && !supplied) {
            std::vector<double> h(n);
            for (int j = 0; j < n; ++j) h[j] = -1.0 - (double)j / (n + 1.0);
            sx_vertex(h, x);
        } else {
            for (int j = 0; j < n; ++j) x[j] = (initial == 0 && !supplied) ? ub[j] : xn[j];
            if (sx) {  // clip, put near-bound coordinates on their bound, scale onto the mass limit
                const double atol = 1e-6, rtol = 1e-4;
                std::vector<char> fr(n, 0);
                for (int j = 0; j < n; ++j) {
                    double v = std::fmin(std::fmax(x[j], lb[j]), ub[j]);
                    if (std::fabs(v - lb[j]) <= atol + rtol * std::fmax(std::fabs(v), std::fabs(lb[j]))) v = lb[j];
                    else if (std::fabs(v - ub[j]) <= atol + rtol * std::fmax(std::fabs(v), std::fabs(ub[j]))) v = ub[j];
                    else fr[j] = 1;
                    x[j] = v;
                }
                const double tx = block_sum(n, 64, [&](double acc, int j) { return acc + x[j]; });
                const double tdn = block_sum(n, 64, [&](double acc, int j) { return fr[j] ? acc + (x[j] - lb[j]) : acc; });
                const double tup = block_sum(n, 64, [&](double acc, int j) { return fr[j] ? acc + (ub[j] - x[j]) : acc; });
                const double tl = block_sum(n, 64, [&](double acc, int j) { return acc + lb[j]; });
                const double tu = block_sum(n, 64, [&](double acc, int j) { return acc + ub[j]; });
                if (tx > rhs) {
                    const double e = tx - rhs;
                    if (tdn > e) {
                        const double f = e / tdn;
                        for (int j = 0; j < n; ++j)
                            if (fr[j]) x[j] = std::fma(lb[j] - x[j], f, x[j]);
                    } else {
                        const double f = (rhs - tl) / (tx - tl);
                        for (int j = 0; j < n; ++j) x[j] = std::fma(x[j] - lb[j], f, lb[j]);
                    }
                } else if (eq && tx < rhs) {
                    const double sh = rhs - tx;
                    if (tup > sh) {
                        const double f = sh / tup;
                        for (int j = 0; j < n; ++j)
                            if (fr[j]) x[j] = std::fma(ub[j] - x[j], f, x[j]);
                    } else {
                        const double f = sh / (tu - tx);
                        for (int j = 0; j < n; ++j) x[j] = std::fma(ub[j] - x[j], f, x[j]);
                    }
                }
            }
        }
        F.gemv(x, qx);
        auto grad = [&]() { for (int j = 0; j < n; ++j) g[j] = o.grad_scale * qx[j] + F.lsc[j]; };
        auto fwd = [&](std::vector<double>& out) {
            if (sx) sx_vertex(g, out);
            else for (int j = 0; j < n; ++j) out[j] = g[j] > 0.0 ? lb[j] : ub[j];
        };
        grad();
        fwd(v);
        F.gemv(v, qv);
        bool have_a = false;
        int iterations = initial;
        const double atol = 1e-6, rtol = 1e-4;
        for (int it = 0; it < max_iters; ++it) {
            iterations = initial + it + 1;
            if (it > 0) {
                grad();
                fwd(vn);
                F.update(vn, v, qv);
                v = vn;
            }
            score(v, qv);
            const double gap = block_sum(n, 64, [&](double acc, int j) { return std::fma(g[j], x[j] - v[j], acc); });
            const double fx = F.dot_value(x, qx) + o.constant;
            if (gap >= 0.0 && gap < eps) break;
            if (fx - gap >= cut) break;
            double gmax;
            if (dicg) {
                if (sx) {  // simplex in-face vertex for -g (lmo.cuh:814-863) on 0/1 data
                    std::vector<double> h(n);
                    for (int j = 0; j < n; ++j) h[j] = -g[j];
                    orc_fast_simplex_inface(n, h.data(), x.data(), lb, ub, rhs, eq, an.data());
                } else {
                    for (int j = 0; j < n; ++j) {
                        const double ut = atol + rtol * std::fabs(ub[j]);
                        const double lt = atol + rtol * std::fabs(lb[j]);
                        if (std::fabs(x[j] - ub[j]) <= ut) an[j] = ub[j];
                        else if (std::fabs(x[j] - lb[j]) <= lt) an[j] = lb[j];
                        else an[j] = (-g[j] > 0.0) ? lb[j] : ub[j];
                    }
                }
                if (have_a) F.update(an, a, qa);
                else F.gemv(an, qa);
                have_a = true;
                a = an;
                score(a, qa);
                std::vector<double> lane(64, 1.0);
                for (int l = 0; l < 64; ++l) {
                    double r = 1.0;
                    for (int j = l; j < n; j += 64) {
                        const double dv = a[j] - v[j];
                        if (sx) {
                            if (dv > 1e-6) r = std::fmin(r, (x[j] - lb[j]) / dv);
                            else if (dv < -1e-6) r = std::fmin(r, (ub[j] - x[j]) / -dv);
                        } else {
                            if (dv < 0.0) r = std::fmin(r, (ub[j] - x[j]) / -dv);
                            else if (dv > 0.0) r = std::fmin(r, (x[j] - lb[j]) / dv);
                        }
                    }
                    lane[l] = r;
                }
                if (sx) {  // dicg_maximum_step_local + the mass limit (lmo.cuh:865-906)
                    const double coord = lanes_min(lane);
                    const double tx = block_sum(n, 64, [&](double acc, int j) { return acc + x[j]; });
                    const double td = block_sum(n, 64, [&](double acc, int j) { return acc + (a[j] - v[j]); });
                    double agg;
                    if (eq && std::fabs(td) > 1e-9) agg = 0.0;
                    else agg = td < 0.0 ? std::fmin(1.0, (tx - rhs) / td) : 1.0;
                    gmax = std::fmax(0.0, std::fmin(coord, agg));
                } else {
                    gmax = std::fmax(0.0, std::fmin(lanes_min(lane), 1.0));
                }
                for (int j = 0; j < n; ++j) { d[j] = v[j] - a[j]; qd[j] = qv[j] - qa[j]; }
            } else {
                for (int j = 0; j < n; ++j) { d[j] = v[j] - x[j]; qd[j] = qv[j] - qx[j]; }
                gmax = 1.0;
            }
            const double pgap = block_sum(n, 64, [&](double acc, int j) { return std::fma(-g[j], d[j], acc); });
            const double curv = block_sum(n, 64, [&](double acc, int j) { return std::fma(o.grad_scale * d[j], qd[j], acc); });
            const double gamma = quad_exact_step(pgap, curv, gmax);
            for (int j = 0; j < n; ++j) {
                x[j] = std::fma(gamma, d[j], x[j]);
                qx[j] = std::fma(gamma, qd[j], qx[j]);
            }
            if (refresh > 0 && (it + 1) % refresh == 0) F.gemv(x, qx);
        }
        F.gemv(x, qx);
        grad();
        fwd(v);
        F.gemv(v, qv);
        score(v, qv);
        const double fgap = block_sum(n, 64, [&](double acc, int j) { return std::fma(g[j], x[j] - v[j], acc); });
        const double fval = F.dot_value(x, qx) + o.constant;
        uint8_t st = (fgap >= 0.0 && fgap < eps) ? 1 : 0;
        if (flags) {
            std::vector<double> cand(n), qc(n);
            for (int j = 0; j < n; ++j) cand[j] = flags[j] ? std::round(x[j]) : x[j];
            F.gemv(cand, qc);
            bool take = true;
            if (sx) {  // Lmo<SIMPLEX>::rounded_feasible
                bool nonneg = true;
                for (int j = 0; j < n; ++j) nonneg = nonneg && cand[j] >= -ctol;
                const double tot = block_sum(n, 64, [&](double acc, int j) { return acc + cand[j]; });
                take = nonneg && (eq ? std::fabs(tot - rhs) <= ctol : tot <= rhs + ctol);
            }
            if (take) score(cand, qc);
        }
        bool integral = true;
        for (int j = 0; j < n; ++j) {
            if (!std::isfinite(x[j]) || (flags && flags[j] && std::fabs(x[j] - std::round(x[j])) > ctol)) integral = false;
            xn[j] = x[j];
        }
        if (integral) st = st ? 4 : 3;
        VALUE[node] = fval;
        GAP[node] = fgap;
        STATUS[node] = st;
        ITERS[node] = iterations;
        BVV[node] = best;
    };
    if (threads <= 1 || N <= 1) {
        for (int i = 0; i < N; ++i) work(i);
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < std::min(threads, N); ++t)
        pool.emplace_back([&] { for (int i = next.fetch_add(1); i < N; i = next.fetch_add(1)) work(i); });
    for (auto& th : pool) th.join();
}

extern "C" void orc_relax_fast_batch(const orc_objective* op, int dicg, int N, int max_iters,
                                     double* X, const double* LB, const double* UB,
                                     const double* EPS, int* ITERS, uint8_t* STATUS, double* GAP,
                                     double* VALUE, double* BV, double* BVV, const uint8_t* flags,
                                     double ctol, const double* cutoff, int refresh, int rescore,
                                     int threads) {
    orc_relax_fast_batch_lmo(op, nullptr, dicg, N, max_iters, X, LB, UB, EPS, ITERS, STATUS, GAP, VALUE, BV, BVV,
                             flags, ctol, cutoff, refresh, rescore, threads);
}

// ---------------------------------------------------------------------------
// Deterministic best-first branch-and-bound (device_resident_solver.cuh:
// 1570-2085, single batch in flight; see the product's host driver for the
// same schedule).  Queue order: (dual_bound, id) ascending.
// ---------------------------------------------------------------------------
namespace {
struct BNode {
    int id, parent, depth;
    std::vector<double> lb, ub, rx;
    int riters = 0;
    double bound;
};
struct Cmp {
    bool operator()(const BNode& a, const BNode& b) const {
        return a.bound > b.bound || (a.bound == b.bound && a.id > b.id);
    }
};
}  // namespace

extern "C" int orc_bnb_solve(const orc_objective* o, const orc_alg* a, const orc_lmo* l, int n,
                             const double* lb, const double* ub, const int* int_vars, int n_int,
                             const orc_bnb_settings* s, orc_bnb_result* r, double* x_out) {
    const auto t0 = std::chrono::steady_clock::now();
    std::priority_queue<BNode, std::vector<BNode>, Cmp> q;
    q.push(BNode{0, -1, 0, std::vector<double>(lb, lb + n), std::vector<double>(ub, ub + n), {}, 0,
                 s->natural_lower_bound});
    std::vector<uint8_t> flags(n, 0);
    for (int t = 0; t < n_int; ++t) flags[int_vars[t]] = 1;
    double inc = kInf;
    std::vector<double> inc_x;
    int next_id = 1, done_nodes = 0;
    long long fw_iters = 0;
    double final_bound = -kInf;
    bool tol_hit = false, time_hit = false, finished = false;
    const int B = s->batch_size;
    // Two slots.  When the batched schedule keeps two batches in flight
    // (separate streams) the next batch is popped, full or not, before the
    // current one is post-processed.  Otherwise (the plain form, or one
    // stream) the next batch is popped before the post-processing only when
    // a full batch is waiting (the reference's refill before post-processing,
    // device_resident_solver.cuh:1585-1596), else after it; both as
    // gpuboscia_amd/csrc/bnb.cpp does.
    const bool pipelined = s->form != 0 && s->num_streams >= 2;
    struct Slot {
        std::vector<BNode> batch;
        std::vector<double> X, L, U, E, G, V, BV, BVV, open_min;
        std::vector<int> IT;
        std::vector<uint8_t> ST;
        size_t pos = 0;
        bool busy() const { return pos < open_min.size(); }
    };
    Slot slot[2];
    auto live_bound = [&]() {
        double b = inc;
        if (!q.empty()) b = std::min(b, q.top().bound);
        for (const Slot& sl : slot)
            if (sl.busy()) b = std::min(b, sl.open_min[sl.pos]);
        return b;
    };
    auto finish_check = [&](bool batch_remaining) {
        if (q.empty() && !batch_remaining) { final_bound = inc; finished = true; return true; }
        const double b = live_bound();
        if (orc_gap_within_tolerance(b, inc, s->abs_gap_tol, s->rel_gap_tol)) {
            final_bound = b; tol_hit = true; finished = true; return true;
        }
        return false;
    };
    // Pop up to B live nodes into a slot and relax them.
    auto launch = [&](Slot& sl) {
        sl.batch.clear();
        sl.open_min.clear();
        sl.pos = 0;
        while (!q.empty() && (int)sl.batch.size() < B) {
            BNode nd = q.top();
            q.pop();
            if (nd.bound >= inc) continue;
            sl.batch.push_back(std::move(nd));
        }
        if (sl.batch.empty()) return;
        const int N = (int)sl.batch.size();
        sl.X.assign((size_t)N * n, 0.0);
        sl.L.assign((size_t)N * n, 0.0);
        sl.U.assign((size_t)N * n, 0.0);
        sl.E.assign(N, 0.0);
        sl.G.assign(N, 0.0);
        sl.V.assign(N, 0.0);
        sl.BV.assign((size_t)N * n, 0.0);
        sl.BVV.assign(N, kNone);
        sl.IT.assign(N, 0);
        sl.ST.assign(N, 0);
        for (int i = 0; i < N; ++i) {
            const BNode& nd = sl.batch[i];
            for (int j = 0; j < n; ++j) {
                sl.X[(size_t)i * n + j] = nd.rx.empty() ? nd.lb[j] : nd.rx[j];
                sl.L[(size_t)i * n + j] = nd.lb[j];
                sl.U[(size_t)i * n + j] = nd.ub[j];
            }
            sl.E[i] = orc_depth_dependent_tolerance(s->eps, s->eps_decay, s->eps_min, nd.depth);
            sl.IT[i] = nd.riters;
            sl.ST[i] = nd.rx.empty() ? 0 : 0x80;
        }
        sl.open_min.resize(N);
        for (int i = N - 1; i >= 0; --i)
            sl.open_min[i] = i + 1 < N ? std::min(sl.batch[i].bound, sl.open_min[i + 1]) : sl.batch[i].bound;
        if (s->form == 1) {
            std::vector<double> CUT(N, inc);
            orc_relax_fast_batch_lmo(o, l, a->dicg, N, s->max_iters, sl.X.data(), sl.L.data(), sl.U.data(), sl.E.data(),
                                 sl.IT.data(), sl.ST.data(), sl.G.data(), sl.V.data(), sl.BV.data(),
                                 sl.BVV.data(), flags.data(), s->integer_tol, CUT.data(), s->refresh, 0,
                                 s->threads);
        } else {
            orc_relax_batch(o, a, l, N, s->max_iters, sl.X.data(), sl.L.data(), sl.U.data(), sl.E.data(),
                            sl.IT.data(), sl.ST.data(), sl.G.data(), sl.V.data(), sl.BV.data(), sl.BVV.data(),
                            flags.data(), s->integer_tol, s->lanes, s->threads);
        }
    };
    int head = 0;
    while (!finished) {
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (el >= s->time_limit_seconds || (s->max_nodes > 0 && done_nodes >= s->max_nodes)) {
            final_bound = live_bound(); time_hit = true; break;
        }
        if (!slot[head].busy()) {
            launch(slot[head]);
            if (!slot[head].busy()) { final_bound = inc; finished = true; break; }
        }
        const int other = 1 - head;
        if (!slot[other].busy() && (pipelined || (int)q.size() >= B)) launch(slot[other]);
        Slot& sl = slot[head];
        const int N = (int)sl.batch.size();
        const int elsewhere = slot[other].busy() ? (int)slot[other].batch.size() : 0;
        const std::vector<double>& X = sl.X;
        const std::vector<double>& G = sl.G;
        const std::vector<double>& V = sl.V;
        const std::vector<double>& BV = sl.BV;
        const std::vector<double>& BVV = sl.BVV;
        const std::vector<double>& E = sl.E;
        const std::vector<int>& IT = sl.IT;
        const std::vector<uint8_t>& ST = sl.ST;
        for (int i = 0; i < N; ++i) fw_iters += IT[i] - sl.batch[i].riters;
        for (int i = 0; i < N && !finished; ++i) {
            BNode nd = std::move(sl.batch[i]);
            sl.pos = (size_t)i + 1;
            const bool remaining = i + 1 < N || elsewhere > 0;
            if (ST[i] == 2) { ++done_nodes; finish_check(remaining); continue; }
            const double node_eps = E[i];
            nd.bound = orc_certified_fw_lower_bound(nd.bound, V[i], G[i]);
            if (BVV[i] < 1e299 && std::isfinite(BVV[i]) && BVV[i] < inc) {
                inc = BVV[i];
                inc_x.assign(BV.begin() + (size_t)i * n, BV.begin() + (size_t)(i + 1) * n);
            }
            const bool cert = std::isfinite(G[i]) && G[i] >= 0.0 && G[i] < node_eps;
            if (nd.bound >= inc) { ++done_nodes; finish_check(remaining); continue; }
            if (ST[i] >= 3 && !cert) {
                nd.rx.assign(X.begin() + (size_t)i * n, X.begin() + (size_t)(i + 1) * n);
                nd.riters = IT[i];
                q.push(std::move(nd));
                continue;
            }
            if (ST[i] < 3) {
                const int base = next_id;
                next_id += 8;
                int nid = base;
                const double* xi = X.data() + (size_t)i * n;
                const auto vars = top_k(xi, int_vars, n_int, s->branch_k, s->integer_tol);
                const int k = (int)vars.size();
                for (int mask = 0; k > 0 && mask < (1 << k); ++mask) {
                    BNode c{nid++, nd.id, nd.depth + 1, nd.lb, nd.ub, {}, 0, nd.bound};
                    for (int b = 0; b < k; ++b) {
                        const int vi = vars[b];
                        if (mask & (1 << b)) c.lb[vi] = std::ceil(xi[vi]);
                        else c.ub[vi] = std::floor(xi[vi]);
                    }
                    if (s->form == 1 && l->kind != ORC_LMO_SIMPLEX) {  // a clipped iterate may break sum x <= rhs
                        c.rx.resize(n);
                        for (int j = 0; j < n; ++j) c.rx[j] = std::min(std::max(xi[j], c.lb[j]), c.ub[j]);
                    }
                    if (c.bound < inc) q.push(std::move(c));
                }
            }
            ++done_nodes;
            finish_check(remaining);
        }
        if (!finished && q.empty() && !slot[other].busy()) { final_bound = inc; finished = true; }
        if (slot[other].busy()) head = other;
    }
    r->objective = inc;
    r->dual_bound = final_bound;
    r->nodes_processed = done_nodes;
    r->reached_tolerance = tol_hit ? 1 : 0;
    r->reached_time_limit = time_hit ? 1 : 0;
    r->fw_iterations = fw_iters;
    r->elapsed_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (x_out)
        for (int j = 0; j < n; ++j) x_out[j] = inc_x.empty() ? std::nan("") : inc_x[j];
    return 0;
}

// ---------------------------------------------------------------------------
// Exhaustive enumeration of x^T Q x + c^T x over {0,1}^n (config C0 check).
// Gray-code walk per thread chunk: f changes by an O(n) update per step.
// ---------------------------------------------------------------------------
extern "C" double orc_enumerate_binary_qp(int n, const double* Q, const double* c, int threads,
                                          double* x_out) {
    if (n <= 0 || n > 30) return std::nan("");
    if (threads < 1) threads = 1;
    const int hi_bits = std::min(n, 6);
    const int chunks = 1 << hi_bits;
    const int lo = n - hi_bits;
    struct Best { double f; unsigned long long mask; };
    std::vector<Best> best(chunks, {kInf, 0});
    std::atomic<int> next{0};
    auto run = [&]() {
        std::vector<double> x(n), qx(n);
        for (int ch = next.fetch_add(1); ch < chunks; ch = next.fetch_add(1)) {
            // Fixed high bits = ch; walk low bits by Gray code.
            for (int j = 0; j < n; ++j) x[j] = 0.0;
            for (int b = 0; b < hi_bits; ++b) x[lo + b] = (ch >> b) & 1;
            // Exact f by direct evaluation at the chunk start, then exact
            // re-evaluation of every candidate that ties or improves.
            auto eval = [&](const std::vector<double>& xv) {
                double f = 0.0;
                for (int i = 0; i < n; ++i) {
                    if (xv[i] == 0.0) continue;
                    double row = 0.0;
                    for (int j = 0; j < n; ++j) row += Q[(size_t)i * n + j] * xv[j];
                    f += row + c[i];
                }
                return f;
            };
            for (int i = 0; i < n; ++i) {
                double s = 0.0;
                for (int j = 0; j < n; ++j) s += Q[(size_t)i * n + j] * x[j];
                qx[i] = s;  // (Q x)_i, symmetric Q
            }
            double f = eval(x);
            Best b{kInf, 0};
            const unsigned long long steps = 1ull << lo;
            unsigned long long gray = 0;
            for (unsigned long long t = 0; t < steps; ++t) {
                if (t > 0) {
                    const int bit = __builtin_ctzll(t);
                    const double dx = x[bit] == 0.0 ? 1.0 : -1.0;
                    // f(x + dx e) = f + dx (2 (Qx)_bit + c_bit) + Q_bb
                    f += dx * (2.0 * qx[bit] + c[bit]) + Q[(size_t)bit * n + bit];
                    x[bit] += dx;
                    for (int i = 0; i < n; ++i) qx[i] += dx * Q[(size_t)i * n + bit];
                    gray ^= 1ull << bit;
                }
                if (f <= b.f + 1e-9) {
                    const double fe = eval(x);
                    unsigned long long mask = gray | ((unsigned long long)ch << lo);
                    if (fe < b.f || (fe == b.f && mask < b.mask)) { b.f = fe; b.mask = mask; }
                }
            }
            best[ch] = b;
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) pool.emplace_back(run);
    for (auto& th : pool) th.join();
    Best g{kInf, 0};
    for (const Best& b : best)
        if (b.f < g.f || (b.f == g.f && b.mask < g.mask)) g = b;
    if (x_out)
        for (int j = 0; j < n; ++j) x_out[j] = (double)((g.mask >> j) & 1ull);
    return g.f;
}

// ---------------------------------------------------------------------------
// Secant line search, exposed for its closed-form pins: on a quadratic the
// derivative along d is linear, so the secant root is the exact step.
// ---------------------------------------------------------------------------
extern "C" double orc_secant_search(const orc_objective* o, const orc_alg* a, int n, int lanes,
                                    double gap, double max_gamma, const double* x, const double* dir) {
    std::vector<double> tx(n), tg(n);
    return secant_search(*o, *a, n, lanes, gap, max_gamma, x, dir, tx, tg);
}

// ---------------------------------------------------------------------------
// The least-squares instance of the reference's equality-simplex convergence
// matrix (tests/test_pipeline.cu:1068-1093): A (m x n) ~ N(0, 1/sqrt(m)) and
// b (m) ~ N(0, 1) from std::mt19937(seed) through two
// std::normal_distribution<double> objects, as that test draws them; then
// Q = A^T A, linear = -A^T b, constant = b^T b / 2, accumulated in the
// test's loop order with plain multiply-add (no contraction).  Outputs Qt
// (Qt[k*n+j] = Q[j,k]), linear, constant.  The normal draws are those of the
// C++ standard library this oracle is built with (libstdc++).
// ---------------------------------------------------------------------------
#include <random>

extern "C" void orc_least_squares_instance(unsigned seed, int m, int n, double* Qt, double* linear,
                                           double* constant) {
    std::mt19937 rng(seed);
    std::normal_distribution<double> a_dist(0.0, 1.0 / std::sqrt(static_cast<double>(m)));
    std::normal_distribution<double> b_dist(0.0, 1.0);
    std::vector<double> A((size_t)m * n), b(m);
    for (double& v : A) v = a_dist(rng);
    for (double& v : b) v = b_dist(rng);
    std::vector<double> Q((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j) linear[j] = 0.0;
    for (int j = 0; j < n; ++j) {
        for (int k = 0; k < n; ++k)
            for (int i = 0; i < m; ++i) Q[(size_t)j * n + k] += A[(size_t)i * n + j] * A[(size_t)i * n + k];
        for (int i = 0; i < m; ++i) linear[j] -= A[(size_t)i * n + j] * b[i];
    }
    for (int j = 0; j < n; ++j)
        for (int k = 0; k < n; ++k) Qt[(size_t)k * n + j] = Q[(size_t)j * n + k];
    double c = 0.0;
    for (double v : b) c += 0.5 * v * v;
    *constant = c;
}
