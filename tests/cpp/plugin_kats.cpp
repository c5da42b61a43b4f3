// Device plug-ins through the C++ drop-in header: caller objective, LMO and
// heuristic types compiled into the relaxation kernel (the reference's
// template extension points, lmo.cuh:11-66, heuristics.h:11-35,
// README.md:224-255).
//   * a plug-in box LMO, a plug-in dense quadratic objective and a plug-in
//     floor heuristic, written as a user would against the reference's
//     contracts, give the built-in kernel's outputs bit for bit (batches
//     through gb_relax_batch_plugin, whole searches through solve());
//   * the sparse Sharpe model (examples/sparse_sharpe.cu): the
//     LinkedCardinalityLMO vertex is a minimiser of <g, .> over its set
//     (checked against enumeration of every asset subset) with the
//     (cost, index) selection order, and the search proves the optimum found
//     by enumerating subsets and active sets of the continuous QP;
//   * the reference's capability errors for plug-in types.
// Exit code 0 = every check passed.
#include <gpuboscia/gpuboscia.hpp>
#include <gpuboscia/device/sparse_sharpe.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
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

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
struct Dev {
    T* p = nullptr;
    size_t n = 0;
    explicit Dev(size_t count) : n(count) { hip_check(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T) + 16), "hipMalloc"); }
    explicit Dev(const std::vector<T>& v) : Dev(v.size()) { put(v); }
    ~Dev() { (void)hipFree(p); }
    Dev(const Dev&) = delete;
    Dev& operator=(const Dev&) = delete;
    void put(const std::vector<T>& v) { hip_check(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "put"); }
    std::vector<T> get() const {
        std::vector<T> v(n);
        hip_check(hipMemcpy(v.data(), p, n * sizeof(T), hipMemcpyDeviceToHost), "get");
        return v;
    }
};

template <class T>
bool same_bits(const std::vector<T>& a, const std::vector<T>& b) {
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(T)) == 0;
}

// ---- plug-in types, written against the reference's contracts ---------------

// BoxLMO's behaviour (lmo.cuh:212-321) as a user plug-in.
struct PlainBoxLMO {
    static constexpr bool supports_dicg = true;
    bool relaxation_feasible_cpu(const std::vector<double>&, const std::vector<double>&, const std::vector<double>&,
                                 double) const {
        return true;
    }
    __device__ bool check_feasible_device(int, int, const double*, const double*, const uint8_t*, double*,
                                          double*) const {
        return true;
    }
    __device__ void compute_v(int tid, int n, const double* g, const double* lb, const double* ub, double* v) const {
        for (int j = tid; j < n; j += blockDim.x) v[j] = g[j] > 0.0 ? lb[j] : ub[j];
    }
    __device__ void compute_inface_v(int tid, int n, const double* g, const double* x, const double* lb,
                                     const double* ub, double* a, double atol = 1e-6, double rtol = 1e-4) const {
        for (int j = tid; j < n; j += blockDim.x) {
            if (fabs(x[j] - ub[j]) <= atol + rtol * fabs(ub[j])) a[j] = ub[j];
            else if (fabs(x[j] - lb[j]) <= atol + rtol * fabs(lb[j])) a[j] = lb[j];
            else a[j] = g[j] > 0.0 ? lb[j] : ub[j];
        }
    }
    __device__ double dicg_maximum_step_local(int tid, int n, const double* d, const double* x, const double* lb,
                                              const double* ub) const {
        double r = 1.0;
        for (int j = tid; j < n; j += blockDim.x) {
            if (d[j] < 0.0) r = fmin(r, (ub[j] - x[j]) / -d[j]);
            else if (d[j] > 0.0) r = fmin(r, (x[j] - lb[j]) / d[j]);
        }
        return r;
    }
    __device__ double dicg_sum_maximum_step(int, int, const double*, const double*) const { return 1.0; }
    __device__ bool candidate_feasible_device(int, int, const double*, const double*, const double*, double) const {
        return true;
    }
};

// Frank-Wolfe only: no DICG trio, no host feasibility check.
struct FwOnlyBoxLMO {
    __device__ bool check_feasible_device(int, int, const double*, const double*, const uint8_t*, double*,
                                          double*) const {
        return true;
    }
    __device__ void compute_v(int tid, int n, const double* g, const double* lb, const double* ub, double* v) const {
        for (int j = tid; j < n; j += blockDim.x) v[j] = g[j] > 0.0 ? lb[j] : ub[j];
    }
};

// GlobalCachedDenseQuadraticObjective as a user plug-in: x staged in the
// node's objective workspace (global memory), each row an FMA chain in k
// order, f = 1/2 s_g x'Qx + s_g s_l c'x + kappa.
template <int N>
struct UserDenseQuadratic {
    static constexpr bool supports_quadratic_exact = true;
    static constexpr int n_vars = N;
    static constexpr int block_threads = 256;
    static constexpr int shared_bytes = N * (int)sizeof(double);
    const double* Qt;
    const double* lin;
    double gs, ls, kappa;

    __device__ void stage(int tid, const double* src, unsigned char* mem) const {
        double* xs = reinterpret_cast<double*>(mem);
        for (int k = tid; k < N; k += blockDim.x) xs[k] = src[k];
        __syncthreads();
    }
    __device__ double row(int j, const unsigned char* mem) const {
        const double* xs = reinterpret_cast<const double*>(mem);
        double acc = 0.0;
        for (int k = 0; k < N; ++k) acc = fma(Qt[(size_t)k * N + j], xs[k], acc);
        return acc;
    }
    __device__ void grad(int tid, const double* x, double* g, unsigned char* mem) const {
        stage(tid, x, mem);
        for (int j = tid; j < N; j += blockDim.x) g[j] = gs * (row(j, mem) + ls * lin[j]);
        __syncthreads();
    }
    __device__ double value_local(int tid, const double* x, unsigned char* mem) const {
        stage(tid, x, mem);
        const double* xs = reinterpret_cast<const double*>(mem);
        double quad = 0.0, l = 0.0;
        for (int j = tid; j < N; j += blockDim.x) {
            quad = fma(xs[j], row(j, mem), quad);
            l = fma(lin[j], xs[j], l);
        }
        double val = fma(0.5 * gs, quad, gs * ls * l);
        if (tid == 0) val = val + kappa;
        __syncthreads();
        return val;
    }
    __device__ double quadratic_curvature_local(int tid, const double* d, unsigned char* mem) const {
        stage(tid, d, mem);
        const double* ds = reinterpret_cast<const double*>(mem);
        double c = 0.0;
        for (int j = tid; j < N; j += blockDim.x) c = fma(gs * ds[j], row(j, mem), c);
        __syncthreads();
        return c;
    }
};

// The same objective without curvature support.
template <int N>
struct NoCurvatureQuadratic {
    static constexpr int n_vars = N;
    UserDenseQuadratic<N> q;
    static constexpr int shared_bytes = N * (int)sizeof(double);
    __device__ void grad(int tid, const double* x, double* g, unsigned char* mem) const { q.grad(tid, x, g, mem); }
    __device__ double value_local(int tid, const double* x, unsigned char* mem) const {
        return q.value_local(tid, x, mem);
    }
};

// The README's FloorHeuristic (README.md:240-251).
struct FloorHeuristic {
    template <typename LMO>
    __device__ void generate(const LMO&, int tid, int n, const double* x, const double*, const double*,
                             const std::uint8_t* integer_flags, double* candidate) const {
        for (int j = tid; j < n; j += blockDim.x) candidate[j] = integer_flags[j] ? floor(x[j]) : x[j];
    }
};

// ---- instances ---------------------------------------------------------------

// Q = A'A/rows + 0.1 I, c ~ U(-1, 1); column-major Qt (symmetric).
void binary_qp(int n, unsigned seed, std::vector<double>& Q, std::vector<double>& c) {
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::uniform_real_distribution<double> ud(-1.0, 1.0);
    const int rows = n + n / 2;
    std::vector<double> A((size_t)rows * n);
    for (double& a : A) a = nd(rng);
    Q.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int t = 0; t < rows; ++t) s += A[(size_t)t * n + i] * A[(size_t)t * n + j];
            Q[(size_t)i * n + j] = s / rows + (i == j ? 0.1 : 0.0);
        }
    c.resize(n);
    for (double& v : c) v = ud(rng);
}

struct BatchOut {
    std::vector<double> x, gap, value, bv, bvv;
    std::vector<int> it;
    std::vector<uint8_t> st;
    bool operator==(const BatchOut& o) const {
        return same_bits(x, o.x) && same_bits(gap, o.gap) && same_bits(value, o.value) && same_bits(bv, o.bv)
               && same_bits(bvv, o.bvv) && same_bits(it, o.it) && same_bits(st, o.st);
    }
};

// One batch of N nodes: fresh nodes with random fixings of some variables.
template <class Run>
BatchOut run_batch(int n, int N, unsigned seed, Run run) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    std::vector<double> lb((size_t)N * n, 0.0), ub((size_t)N * n, 1.0);
    for (int i = 1; i < N; ++i)
        for (int j = 0; j < n; ++j)
            if (u(rng) < 0.2) (u(rng) < 0.5 ? lb : ub)[(size_t)i * n + j] = u(rng) < 0.5 ? 0.0 : 1.0;
    for (size_t k = 0; k < lb.size(); ++k)
        if (lb[k] > ub[k]) std::swap(lb[k], ub[k]);
    Dev<double> x(lb), dlb(lb), dub(ub), eps(std::vector<double>(N, 1e-6)), gap(N), val(N), bv((size_t)N * n),
        bvv(std::vector<double>(N, 1e300));
    Dev<int> it(std::vector<int>(N, 0));
    Dev<uint8_t> st(std::vector<uint8_t>(N, 0)), flags(std::vector<uint8_t>(n, 1));
    gpuboscia::detail::throw_status(run(N, x.p, dlb.p, dub.p, eps.p, it.p, st.p, gap.p, val.p, bv.p, bvv.p, flags.p));
    hip_check(hipDeviceSynchronize(), "batch");
    return BatchOut{x.get(), gap.get(), val.get(), bv.get(), bvv.get(), it.get(), st.get()};
}

template <class O, class L, class H>
BatchOut relax_with(const O& obj, const L& lmo, const H& heur, const gb_algorithm_desc& alg0, int n, int N,
                    unsigned seed, int max_iters = 300) {
    gpuboscia::detail::Binding<O, L, H> b(obj, lmo, heur, n);
    gb_algorithm_desc alg = alg0;
    b.fill_heuristic(alg);
    return run_batch(n, N, seed, [&](int N_, double* x, double* lb, double* ub, double* eps, int* it, uint8_t* st,
                                     double* gap, double* val, double* bv, double* bvv, uint8_t* flags) {
        if (b.plugin())
            return gb_relax_batch_plugin(b.objective(), &alg, b.lmo(), b.plugin(), N_, max_iters, x, lb, ub, eps, it,
                                         st, gap, val, bv, bvv, flags, 1e-9, 0, nullptr);
        return gb_relax_batch(b.objective(), &alg, b.lmo(), N_, max_iters, x, lb, ub, eps, it, st, gap, val, bv, bvv,
                              flags, 1e-9, 0, nullptr);
    });
}

gb_algorithm_desc algorithm(int alg, int ls) {
    gb_algorithm_desc a{alg, ls, 4, 1e-10, 1e-14, GB_HEUR_ROUND, 0.0};
    return a;
}

void batch_parity() {
    std::printf("\n[plug-in types give the built-in kernel's bits]\n");
    constexpr int n = 40;
    std::vector<double> Q, c;
    binary_qp(n, 11, Q, c);
    Dev<double> dQ(Q), dc(c);
    const GlobalCachedDenseQuadraticObjective<n> builtin_obj(dQ.p, dc.p, 2.0, 0.5, 0.0);
    const UserDenseQuadratic<n> user_obj{dQ.p, dc.p, 2.0, 0.5, 0.0};
    const struct {
        int alg, ls;
        const char* name;
    } algs[] = {{GB_ALG_DICG, GB_LS_SECANT, "DICG + secant"},
                {GB_ALG_FRANK_WOLFE, GB_LS_OPEN_LOOP, "FW + open loop"},
                {GB_ALG_DICG, GB_LS_QUADRATIC_EXACT, "DICG + exact step"},
                {GB_ALG_DICG, GB_LS_SAFEGUARDED_SCALED_OPEN_LOOP, "DICG + safeguarded scaled"}};
    for (const auto& a : algs) {
        const gb_algorithm_desc alg = algorithm(a.alg, a.ls);
        const BatchOut ref = relax_with(builtin_obj, BoxLMO{}, gpuboscia::RoundingHeuristic{}, alg, n, 24, 5);
        check(relax_with(builtin_obj, PlainBoxLMO{}, gpuboscia::RoundingHeuristic{}, alg, n, 24, 5) == ref,
              std::string("plug-in box LMO, ") + a.name);
        check(relax_with(user_obj, BoxLMO{}, gpuboscia::RoundingHeuristic{}, alg, n, 24, 5) == ref,
              std::string("plug-in dense objective, ") + a.name);
        check(relax_with(user_obj, PlainBoxLMO{}, gpuboscia::RoundingHeuristic{}, alg, n, 24, 5) == ref,
              std::string("plug-in objective and LMO, ") + a.name);
        const BatchOut floor_ref = relax_with(builtin_obj, BoxLMO{}, gpuboscia::FloorHeuristic{0.0}, alg, n, 24, 5, 7);
        check(relax_with(builtin_obj, BoxLMO{}, FloorHeuristic{}, alg, n, 24, 5, 7) == floor_ref,
              std::string("plug-in floor heuristic, ") + a.name);
    }
}

template <class O, class L, class H = gpuboscia::RoundingHeuristic>
gpuboscia::Result solve_qp(const O& obj, const L& lmo, int n, H heur = {}) {
    Problem p;
    p.n_vars = n;
    p.lb.assign(n, 0.0);
    p.ub.assign(n, 1.0);
    for (int j = 0; j < n; ++j) p.int_vars.push_back(j);
    auto s = gpuboscia::create_default_settings();
    s.device.abs_gap_tol = s.device.rel_gap_tol = 1e-6;
    s.device.gpu_batch_size = 64;
    return gpuboscia::solve(p, obj, lmo, s, heur);
}

void solve_parity() {
    std::printf("\n[searches with plug-in types equal the built-in search]\n");
    constexpr int n = 30;
    std::vector<double> Q, c;
    binary_qp(n, 3, Q, c);
    Dev<double> dQ(Q), dc(c);
    const GlobalCachedDenseQuadraticObjective<n> builtin_obj(dQ.p, dc.p, 2.0, 0.5, 0.0);
    const UserDenseQuadratic<n> user_obj{dQ.p, dc.p, 2.0, 0.5, 0.0};
    const gpuboscia::Result ref = solve_qp(builtin_obj, BoxLMO{}, n);
    auto same = [&](const gpuboscia::Result& r) {
        return r.objective == ref.objective && r.dual_bound == ref.dual_bound && r.x == ref.x
               && r.nodes_processed == ref.nodes_processed && r.reached_tolerance == ref.reached_tolerance;
    };
    check(ref.objective - ref.dual_bound <= 1e-6 * (1.0 + std::abs(ref.objective)) && ref.nodes_processed > 1,
          "built-in search proves its optimum (" + std::to_string(ref.nodes_processed) + " nodes)");
    check(same(solve_qp(builtin_obj, PlainBoxLMO{}, n)), "plug-in box LMO: same search");
    check(same(solve_qp(user_obj, BoxLMO{}, n)), "plug-in dense objective: same search");
    check(same(solve_qp(user_obj, PlainBoxLMO{}, n)), "plug-in objective and LMO: same search");
    const gpuboscia::Result fl = solve_qp(builtin_obj, BoxLMO{}, n, gpuboscia::FloorHeuristic{0.0});
    const gpuboscia::Result fu = solve_qp(builtin_obj, BoxLMO{}, n, FloorHeuristic{});
    check(fl.objective == fu.objective && fl.dual_bound == fu.dual_bound && fl.x == fu.x
              && fl.nodes_processed == fu.nodes_processed && fu.objective == ref.objective,
          "plug-in floor heuristic: same search and optimum");
}

// ---- sparse Sharpe --------------------------------------------------------------

constexpr int kMax = 16;  // MaxAssets of the test model (n = 32)
using LCLMO = gpuboscia::LinkedCardinalityLMO<kMax>;

// The minimum of <g, .> over D and the node's y bounds, by enumerating every
// subset of at most m assets (the v part of a subset's best point is U_i
// where g_v < 0 and 0 elsewhere).
double brute_lmo_value(const std::vector<double>& g, const std::vector<double>& lb, const std::vector<double>& ub,
                       const std::vector<double>& U, int assets, int m) {
    double best = std::numeric_limits<double>::infinity();
    for (unsigned mask = 0; mask < (1u << assets); ++mask) {
        if (__builtin_popcount(mask) > m) continue;
        bool ok = true;
        double val = 0.0;
        for (int i = 0; i < assets && ok; ++i) {
            const bool in = (mask >> i) & 1u;
            if (in && ub[kMax + i] < 0.5) ok = false;
            if (!in && lb[kMax + i] > 0.5) ok = false;
            if (in) val += g[kMax + i] + std::min(0.0, U[i] * g[i]);
        }
        if (ok) best = std::min(best, val);
    }
    return best;
}

void linked_cardinality_vertices() {
    std::printf("\n[LinkedCardinalityLMO vertices (examples/sparse_sharpe.cu:259-421)]\n");
    const int assets = 12, m = 4, N = 64, n = 2 * kMax;
    std::mt19937_64 rng(21);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::uniform_real_distribution<double> ud(0.0, 1.0);
    std::vector<double> U(assets);
    for (double& u : U) u = 0.2 + ud(rng);
    std::vector<double> G((size_t)N * n, 0.0), lb((size_t)N * n, 0.0), ub((size_t)N * n, 0.0);
    for (int r = 0; r < N; ++r) {
        for (int i = 0; i < assets; ++i) {
            double* g = G.data() + (size_t)r * n;
            g[i] = nd(rng);
            g[kMax + i] = 0.3 * nd(rng);
            if (r % 8 == 7) g[kMax + i] = std::round(g[kMax + i] * 4.0) / 4.0;  // ties in the costs
            if (r % 8 == 7) g[i] = 0.0;
            ub[(size_t)r * n + i] = U[i];
            ub[(size_t)r * n + kMax + i] = 1.0;
            const double f = ud(rng);
            if (r % 2 == 1 && f < 0.12) lb[(size_t)r * n + kMax + i] = 1.0;       // mandatory
            else if (r % 2 == 1 && f < 0.3) ub[(size_t)r * n + kMax + i] = 0.0;  // forbidden
        }
        int mand = 0;
        for (int i = 0; i < assets; ++i) mand += lb[(size_t)r * n + kMax + i] > 0.5;
        for (int i = 0; i < assets && mand > m; ++i)
            if (lb[(size_t)r * n + kMax + i] > 0.5) { lb[(size_t)r * n + kMax + i] = 0.0; --mand; }
    }
    Dev<double> dU(U), dG(G), dlb(lb), dub(ub), dv((size_t)N * n);
    Dev<double> dH(std::vector<double>((size_t)assets * assets, 0.0)), dmu(std::vector<double>(assets, 0.0));
    const LCLMO lmo{assets, m, dU.p, dH.p, dmu.p};
    gb::RelaxParams P{};
    P.n = n;
    P.N = N;
    P.lb = dlb.p;
    P.ub = dub.p;
    P.x = dv.p;
    const gb::LdsLayout L = gb::lds_layout(GB_OBJ_PLUGIN, GB_LMO_PLUGIN, n, 0);
    for (int threads : {64, 128}) {
        hip_check(static_cast<hipError_t>(
                      gb::launch_lmo_probe_with(P, threads, L.total, L, dG.p, gb::PluginLmo<LCLMO>{lmo}, nullptr)),
                  "probe");
        hip_check(hipDeviceSynchronize(), "probe");
        const std::vector<double> V = dv.get();
        int bad_value = 0, bad_set = 0, bad_feas = 0;
        for (int r = 0; r < N; ++r) {
            const double* g = G.data() + (size_t)r * n;
            const double* v = V.data() + (size_t)r * n;
            const std::vector<double> gr(g, g + n), lr(lb.begin() + (size_t)r * n, lb.begin() + (size_t)(r + 1) * n),
                ur(ub.begin() + (size_t)r * n, ub.begin() + (size_t)(r + 1) * n);
            double val = 0.0, ysum = 0.0;
            for (int i = 0; i < assets; ++i) {
                val += g[kMax + i] * v[kMax + i];
                if (v[i] != 0.0) val += U[i] * g[i];
                ysum += v[kMax + i];
                const bool y01 = v[kMax + i] == 0.0 || v[kMax + i] == 1.0;
                if (!y01 || v[i] < 0.0 || v[i] > U[i] * v[kMax + i] || v[kMax + i] < lr[kMax + i]
                        || v[kMax + i] > ur[kMax + i] || (v[i] != 0.0 && v[i] != U[i]))
                    ++bad_feas;
            }
            for (int j = assets; j < kMax; ++j) bad_feas += (v[j] != 0.0) + (v[kMax + j] != 0.0);
            if (ysum > m) ++bad_feas;
            const double best = brute_lmo_value(gr, lr, ur, U, assets, m);
            if (std::abs(val - best) > 1e-12 * (1.0 + std::abs(best))) ++bad_value;
            // (cost, index) order of the optional assets
            std::vector<std::pair<double, int>> opt;
            int mand = 0;
            for (int i = 0; i < assets; ++i) {
                const double cst = g[kMax + i] + std::min(0.0, U[i] * g[i]);
                if (lr[kMax + i] > 0.5) ++mand;
                else if (cst < 0.0 && ur[kMax + i] > 0.5) opt.push_back({cst, i});
            }
            std::sort(opt.begin(), opt.end());
            std::vector<int> want(assets, 0);
            for (int i = 0; i < assets; ++i) want[i] = lr[kMax + i] > 0.5;
            for (int k = 0; k < (int)opt.size() && k < m - mand; ++k) want[opt[k].second] = 1;
            for (int i = 0; i < assets; ++i)
                if ((v[kMax + i] == 1.0) != (want[i] == 1)) { ++bad_set; break; }
        }
        check(bad_feas == 0, "vertices lie in D within the node bounds (" + std::to_string(threads) + " threads)");
        check(bad_value == 0, "vertex values equal the subset enumeration's minimum (" + std::to_string(N) + " directions)");
        check(bad_set == 0, "selected assets follow the (cost, index) order, ties included");
    }
}

// min 0.5 v'Hv - mu'v over 0 <= v <= U on a subset: enumerate each
// coordinate at 0, at U or free, solve the free block and keep the KKT
// point (H is positive definite, so it is the minimiser).
double subset_qp(const std::vector<double>& H, const std::vector<double>& mu, const std::vector<double>& U,
                 int assets, const std::vector<int>& S) {
    const int k = (int)S.size();
    if (k == 0) return 0.0;
    double best = std::numeric_limits<double>::infinity();
    int patterns = 1;
    for (int t = 0; t < k; ++t) patterns *= 3;
    for (int pat = 0; pat < patterns; ++pat) {
        std::vector<int> state(k);
        for (int t = 0, q = pat; t < k; ++t, q /= 3) state[t] = q % 3;  // 0 at 0, 1 at U, 2 free
        std::vector<double> v(k, 0.0);
        std::vector<int> F;
        for (int t = 0; t < k; ++t) {
            if (state[t] == 1) v[t] = U[S[t]];
            if (state[t] == 2) F.push_back(t);
        }
        const int f = (int)F.size();
        // H_FF v_F = mu_F - H_FB v_B (Gaussian elimination, f <= 4)
        std::vector<double> M((size_t)f * (f + 1));
        for (int a = 0; a < f; ++a) {
            double rhs = mu[S[F[a]]];
            for (int t = 0; t < k; ++t)
                if (state[t] != 2) rhs -= H[(size_t)S[F[a]] * assets + S[t]] * v[t];
            for (int b = 0; b < f; ++b) M[(size_t)a * (f + 1) + b] = H[(size_t)S[F[a]] * assets + S[F[b]]];
            M[(size_t)a * (f + 1) + f] = rhs;
        }
        for (int col = 0; col < f; ++col) {
            int piv = col;
            for (int r = col + 1; r < f; ++r)
                if (std::abs(M[(size_t)r * (f + 1) + col]) > std::abs(M[(size_t)piv * (f + 1) + col])) piv = r;
            for (int cc = 0; cc <= f; ++cc) std::swap(M[(size_t)col * (f + 1) + cc], M[(size_t)piv * (f + 1) + cc]);
            for (int r = 0; r < f; ++r) {
                if (r == col) continue;
                const double fac = M[(size_t)r * (f + 1) + col] / M[(size_t)col * (f + 1) + col];
                for (int cc = col; cc <= f; ++cc) M[(size_t)r * (f + 1) + cc] -= fac * M[(size_t)col * (f + 1) + cc];
            }
        }
        bool ok = true;
        for (int a = 0; a < f; ++a) {
            v[F[a]] = M[(size_t)a * (f + 1) + f] / M[(size_t)a * (f + 1) + a];
            if (v[F[a]] < -1e-12 || v[F[a]] > U[S[F[a]]] + 1e-12) ok = false;
        }
        if (!ok) continue;
        for (int t = 0; t < k && ok; ++t) {
            double gt = -mu[S[t]];
            for (int u = 0; u < k; ++u) gt += H[(size_t)S[t] * assets + S[u]] * v[u];
            if (state[t] == 0 && gt < -1e-10) ok = false;
            if (state[t] == 1 && gt > 1e-10) ok = false;
        }
        if (!ok) continue;
        double val = 0.0;
        for (int t = 0; t < k; ++t) {
            val -= mu[S[t]] * v[t];
            for (int u = 0; u < k; ++u) val += 0.5 * v[t] * H[(size_t)S[t] * assets + S[u]] * v[u];
        }
        best = std::min(best, val);
    }
    return best;
}

void sparse_sharpe_search() {
    std::printf("\n[sparse Sharpe search with LinkedCardinalityLMO (examples/sparse_sharpe.cu)]\n");
    const int assets = 14, m = 4, T = 60, n = 2 * kMax;
    std::mt19937_64 rng(5);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::vector<double> R((size_t)T * assets), mu(assets, 0.0);
    for (int t = 0; t < T; ++t) {
        const double f = 0.03 * nd(rng);
        for (int i = 0; i < assets; ++i) {
            R[(size_t)t * assets + i] = 0.004 * (i % 4) + f * (0.6 + 0.1 * i) + 0.05 * nd(rng);
            mu[i] += R[(size_t)t * assets + i] / T;
        }
    }
    std::vector<double> H((size_t)assets * assets);
    for (int i = 0; i < assets; ++i)
        for (int j = 0; j < assets; ++j) {
            double s = 0.0;
            for (int t = 0; t < T; ++t) s += (R[(size_t)t * assets + i] - mu[i]) * (R[(size_t)t * assets + j] - mu[j]);
            H[(size_t)i * assets + j] = s / (T - 1) + (i == j ? 1e-3 : 0.0);
        }
    std::vector<double> U(assets, 1.5);
    double best = std::numeric_limits<double>::infinity();
    for (unsigned mask = 0; mask < (1u << assets); ++mask) {
        if (__builtin_popcount(mask) > m) continue;
        std::vector<int> S;
        for (int i = 0; i < assets; ++i)
            if ((mask >> i) & 1u) S.push_back(i);
        best = std::min(best, subset_qp(H, mu, U, assets, S));
    }
    Dev<double> dH(H), dmu(mu), dU(U);
    const gpuboscia::SparseSharpeObjective<kMax> obj{dH.p, dmu.p, assets};
    const LCLMO lmo{assets, m, dU.p, dH.p, dmu.p};
    Problem p;
    p.n_vars = n;
    p.lb.assign(n, 0.0);
    p.ub.assign(n, 0.0);
    for (int i = 0; i < assets; ++i) {
        p.ub[i] = U[i];
        p.ub[kMax + i] = 1.0;
        p.int_vars.push_back(kMax + i);
    }
    // An integral node whose relaxation is certified within its tolerance is
    // closed with the heuristic's candidate (device_resident_solver.cuh:
    // 1970-1984), so the node tolerance is tightened until that candidate is
    // the subset's own optimum.
    for (auto alg : {FWAlgorithmKind::DICG, FWAlgorithmKind::FrankWolfe}) {
        auto s = gpuboscia::create_default_settings();
        s.device.gpu_batch_size = 64;
        s.device.gpu_max_iters = 10000;
        s.device.gpu_eps = s.device.gpu_eps_min = 1e-9;
        s.device.integer_tol = 1e-5;
        s.device.abs_gap_tol = 1e-7;
        s.device.rel_gap_tol = 1e-6;
        s.algorithm = alg;
        s.line_search.kind = alg == FWAlgorithmKind::DICG ? LineSearchKind::Secant : LineSearchKind::QuadraticExact;
        s.line_search.secant_iters = 16;
        const gpuboscia::Result r = gpuboscia::solve(p, obj, lmo, s);
        const char* name = alg == FWAlgorithmKind::DICG ? "DICG + secant" : "FW + exact step";
        char buf[256];
        std::snprintf(buf, sizeof buf, "%s: objective %.9f, bound %.9f, enumeration %.9f, %d nodes", name,
                      r.objective, r.dual_bound, best, r.nodes_processed);
        check(r.objective - r.dual_bound <= 1e-6 * (1.0 + std::abs(r.objective))
                  && std::abs(r.objective - best) <= 2e-6 * (1.0 + std::abs(best)) && r.dual_bound <= best + 1e-9,
              buf);
        int sel = 0;
        bool linked = r.x.size() == (size_t)n;
        for (int i = 0; i < assets && linked; ++i) {
            sel += r.x[kMax + i] > 0.5;
            linked = r.x[i] >= -1e-9 && r.x[i] <= U[i] * r.x[kMax + i] + 1e-9;
        }
        check(linked && sel <= m, std::string(name) + ": incumbent has linked positions and at most m assets");
    }
}

void capability_errors() {
    std::printf("\n[capability errors for plug-in types (gpuboscia.cuh:115-149, tree.h:132-146)]\n");
    constexpr int n = 8;
    std::vector<double> Q, c;
    binary_qp(n, 1, Q, c);
    Dev<double> dQ(Q), dc(c);
    const GlobalCachedDenseQuadraticObjective<n> obj(dQ.p, dc.p, 2.0, 0.5, 0.0);
    Problem p;
    p.n_vars = n;
    p.lb.assign(n, 0.0);
    p.ub.assign(n, 1.0);
    for (int j = 0; j < n; ++j) p.int_vars.push_back(j);
    auto expect = [&](const std::function<void()>& f, const std::string& frag, const std::string& what) {
        try {
            f();
            check(false, what + " (no exception)");
        } catch (const std::invalid_argument& e) {
            check(std::string(e.what()).find(frag) != std::string::npos, what + " (" + e.what() + ")");
        }
    };
    auto s = gpuboscia::create_default_settings();
    s.device.abs_gap_tol = s.device.rel_gap_tol = 1e-6;
    expect([&] { gpuboscia::solve(p, obj, FwOnlyBoxLMO{}, s); }, "does not support DICG",
           "DICG with an LMO without the in-face trio");
    auto fw = s;
    fw.algorithm = FWAlgorithmKind::FrankWolfe;
    const gpuboscia::Result r = gpuboscia::solve(p, obj, FwOnlyBoxLMO{}, fw);
    const gpuboscia::Result rb = gpuboscia::solve(p, obj, BoxLMO{}, fw);
    check(r.objective == rb.objective && r.x == rb.x, "Frank-Wolfe with that LMO solves as the box LMO");
    auto ex = fw;
    ex.line_search.kind = LineSearchKind::QuadraticExact;
    const NoCurvatureQuadratic<n> nc{UserDenseQuadratic<n>{dQ.p, dc.p, 2.0, 0.5, 0.0}};
    expect([&] { gpuboscia::solve(p, nc, BoxLMO{}, ex); }, "quadratic-exact requires",
           "the exact step with an objective without curvature");
    auto root = fw;
    root.root_initial_point.assign(n, 0.5);
    expect([&] { gpuboscia::solve(p, obj, FwOnlyBoxLMO{}, root); }, "relaxation_feasible_cpu",
           "a root point with an LMO that has no host check");
    root.root_initial_point_feasible = [](const std::vector<double>&, const std::vector<double>&,
                                          const std::vector<double>&, double) { return true; };
    const gpuboscia::Result rr = gpuboscia::solve(p, obj, FwOnlyBoxLMO{}, root);
    check(rr.objective == rb.objective, "... accepted with a validator; same optimum");
    auto rootp = s;
    rootp.root_initial_point.assign(n, 0.5);
    const gpuboscia::Result rp = gpuboscia::solve(p, obj, PlainBoxLMO{}, rootp);
    check(rp.objective == rb.objective, "a root point checked by the plug-in's relaxation_feasible_cpu");
}

}  // namespace

int main() {
    try {
        batch_parity();
        solve_parity();
        linked_cardinality_vertices();
        sparse_sharpe_search();
        capability_errors();
    } catch (const std::exception& e) {
        std::printf("FAIL: exception %s\n", e.what());
        ++failures;
    }
    std::printf("\n%d failure(s)\n", failures);
    return failures == 0 ? 0 : 1;
}
