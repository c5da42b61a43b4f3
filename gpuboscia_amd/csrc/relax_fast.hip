// Batched-form node relaxation for dense quadratics over boxes, gfx950.
//
// One wave64 owns one B&B node; waves are persistent and pull node indices
// from a device counter, so converged nodes hand their wave to the next node
// of the batch (no tail of idle blocks).  Lane l holds coordinates
// j = l + 64 t (t < NT) of x, Qx, v, Qv, a, Qa, lb, ub in registers: every
// reduction is a wave shuffle and no workgroup barrier is executed.
//
// Algorithm (FP64 throughout):
//   * products Q x, Q v, Q a are kept incrementally: a box vertex changes in
//     few coordinates per iteration, so Q v_new = Q v + sum_k Q[:,k] dv_k over
//     the changed k only; Q x follows x += gamma d as Qx += gamma Q d;
//     Q x is recomputed directly every `refresh` iterations;
//   * exact quadratic step gamma = clip(gap_pair / d'Hd, 0, gamma_max)
//     (line_search_options.h:48-63), d'Hd from Q d = Qv - Qa;
//   * the iteration stops at the per-node FW tolerance, or as soon as the
//     FW lower bound f(x) - gap reaches the node's cutoff (the incumbent:
//     the node will be pruned);
//   * the returned certificate is always recomputed directly in FP64:
//     Q x by a GEMV, then v, gap and f(x), exactly like the plain form's
//     final step (device_resident_solver.cuh:588-621).
// Vertex and rounded-candidate scoring follow drs_integer_value and the
// rounding heuristic (device_resident_solver.cuh:297-335).
#include "gb_device.h"
#include "gb_fast.h"

namespace gb {
namespace {

__device__ __forceinline__ double wave_sum_all(double v) {
    return __shfl(wave_tree_sum(v), 0, kWave);
}
__device__ __forceinline__ double wave_min_all(double v) {
    return __shfl(wave_tree_min(v), 0, kWave);
}
__device__ __forceinline__ bool wave_all(bool p) { return __ballot(!p) == 0ull; }

__device__ __forceinline__ double quad_exact(double gap, double curvature, double maximum) {
    if (maximum <= 0.0) return 0.0;
    if (curvature > 0.0) {
        if (gap <= 0.0) return 0.0;
        const double step = gap / curvature;
        return step < maximum ? step : maximum;
    }
    return -gap + 0.5 * curvature * maximum < 0.0 ? maximum : 0.0;
}

template <int NT, bool QL>
struct WaveNode {
    const FastParams& P;
    const double* Q;    // Qt in LDS (QL) or global memory
    int lane;
    double* zs;  // this wave's LDS staging vector (n doubles)
    long long columns;  // column updates performed (work accounting)
    long long gemvs;    // direct products performed

    // One sweep over Q for NV = 1 or 2 vectors staged in LDS (za, zb):
    // ya = Q za, yb = Q zb for the lane's rows.  k runs outermost in batches
    // of kBatch columns over all of the lane's rows, so kBatch * NT loads
    // are in flight per round trip.  Every output is the
    // FMA chain over k ascending, whatever NV and kBatch.
    static constexpr int kBatch = NT >= 8 ? 2 : (NT >= 4 ? 4 : 8);
    __device__ void load_batch(int k0, double (&qq)[kBatch][NT]) {
        const int n = P.n;
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const int j = lane + kWave * t;
                const int k = k0 + u;
                qq[u][t] = (j < n && k < n) ? Q[(size_t)k * n + j] : 0.0;
            }
    }
    template <int NV>
    __device__ void sweep(const double* za, const double* zb, double (&ya)[NT], double (&yb)[NT]) {
        const int n = P.n;
        double acc[NV][NT];
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[v][t] = 0.0;
        const int batches = (n + kBatch - 1) / kBatch;
        for (int b = 0; b < batches; ++b) {
            double qq[kBatch][NT];
            load_batch(b * kBatch, qq);
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int k = b * kBatch + u;
                if (k < n) {
                    const double z0 = za[k];
#pragma unroll
                    for (int t = 0; t < NT; ++t) acc[0][t] = fma(qq[u][t], z0, acc[0][t]);
                    if constexpr (NV == 2) {
                        const double z1 = zb[k];
#pragma unroll
                        for (int t = 0; t < NT; ++t) acc[NV - 1][t] = fma(qq[u][t], z1, acc[NV - 1][t]);
                    }
                }
            }
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const bool in = lane + kWave * t < n;
            ya[t] = in ? acc[0][t] : 0.0;
            if constexpr (NV == 2) yb[t] = in ? acc[NV - 1][t] : 0.0;
        }
    }

    __device__ void stage(double* dst, const double (&z)[NT]) {
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int j = lane + kWave * t;
            if (j < P.n) dst[j] = z[t];
        }
    }

    // y = Q z for the lane's rows; z given in registers.
    __device__ void gemv(const double (&z)[NT], double (&y)[NT]) {
        stage(zs, z);
        wave_sync();
        sweep<1>(zs, zs, y, y);
        wave_sync();
        ++gemvs;
    }

    // y1 = Q z1 and y2 = Q z2 in one sweep over Q (the change-list buffer
    // stages z2; no update is in progress).
    __device__ void gemv2(const double (&z1)[NT], double (&y1)[NT], const double (&z2)[NT], double (&y2)[NT]) {
        if constexpr (NT >= 8) {  // n > 256: two sweeps keep the state in registers
            gemv(z1, y1);
            gemv(z2, y2);
            return;
        }
        stage(zs, z1);
        stage(dlist, z2);
        wave_sync();
        sweep<2>(zs, dlist, y1, y2);
        wave_sync();
        gemvs += 2;
    }

    // Changed columns are compacted into this wave's LDS list (k ascending)
    // and their loads issued kGroup at a time before the FMAs: the vertex of
    // an interior iterate flips many coordinates per iteration, and one L2
    // round trip per column would leave the wave latency-bound.
    static constexpr int kGroup = NT >= 4 ? 4 : 8;
    int* klist;      // this wave's LDS list of changed columns (n ints)
    double* dlist;   // their coefficients (n doubles)

    // y += sum over k with znew_k != zold_k of Q[:,k] (znew_k - zold_k), k
    // ascending (the FMA order per output coordinate is k order).
    __device__ int update(const double (&znew)[NT], const double (&zold)[NT], double (&y)[NT]) {
        const int n = P.n;
        int count = 0;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int j = lane + kWave * t;
            const bool ch = j < n && znew[t] != zold[t];
            const unsigned long long mask = __ballot(ch);
            if (ch) {
                const int pos = count + __popcll(mask & ((1ull << lane) - 1ull));
                klist[pos] = j;
                dlist[pos] = znew[t] - zold[t];
            }
            count += __popcll(mask);
        }
        wave_sync();
        for (int g0 = 0; g0 < count; g0 += kGroup) {
            double c[kGroup][NT];
            double dd[kGroup];
#pragma unroll
            for (int g = 0; g < kGroup; ++g) {
                const bool live = g0 + g < count;
                const int k = live ? klist[g0 + g] : 0;
                dd[g] = live ? dlist[g0 + g] : 0.0;
#pragma unroll
                for (int tt = 0; tt < NT; ++tt) {
                    const int jj = lane + kWave * tt;
                    c[g][tt] = (live && jj < n) ? Q[(size_t)k * n + jj] : 0.0;
                }
            }
#pragma unroll
            for (int g = 0; g < kGroup; ++g) {
                if (g0 + g < count) {
#pragma unroll
                    for (int tt = 0; tt < NT; ++tt) {
                        const int jj = lane + kWave * tt;
                        if (jj < n) y[tt] = fma(c[g][tt], dd[g], y[tt]);
                    }
                }
            }
        }
        wave_sync();
        columns += count;
        return count;
    }
};

// QL: Q is copied once into LDS and shared by the block's resident waves
// (the block is persistent, so the copy is amortised over all its nodes).
template <int NT, bool QL>
__global__ void __launch_bounds__(256) relax_fast_kernel(FastParams P) {
    extern __shared__ __align__(16) double lds_fast[];
    const int lane = lane_id();
    const int n = P.n;
    const size_t qsize = QL ? (size_t)n * n : 0;
    if (QL) {
        for (size_t i = threadIdx.x; i < qsize; i += blockDim.x) lds_fast[i] = P.Qt[i];
        __syncthreads();
    }
    // Per-wave LDS: staging vector (n_pad doubles), changed-column
    // coefficients (n_pad doubles), changed-column indices (n_pad ints).
    double* wbase = lds_fast + qsize + (size_t)wave_id() * (size_t)P.n_pad * 5 / 2;
    WaveNode<NT, QL> W{P, QL ? lds_fast : P.Qt, lane, wbase, 0, 0,
                       reinterpret_cast<int*>(wbase + 2 * P.n_pad), wbase + P.n_pad};
    const double hs = 0.5 * P.gs;
    const double lsc_scale = P.gs * P.ls;
    const double atol = 1e-6, rtol = 1e-4;

    while (true) {
        int node = 0;
        if (lane == 0) node = atomicAdd(P.counter, 1);
        node = __shfl(node, 0, kWave);
        if (node >= P.N) break;
        W.columns = 0;
        W.gemvs = 0;

        const double* lbn = P.lb + (size_t)node * n;
        const double* ubn = P.ub + (size_t)node * n;
        double* xn = P.x + (size_t)node * n;
        double* bvn = P.best_v + (size_t)node * n;
        const double eps = P.eps[node];
        const double cutoff = P.cutoff ? P.cutoff[node] : __builtin_inf();
        const int initial = P.iters[node];
        const bool supplied = (P.status[node] & GB_NODE_INITIAL_POINT_SUPPLIED) != 0;

        double lb[NT], ub[NT], lsc[NT], x[NT], qx[NT], v[NT], qv[NT], a[NT], qa[NT], g[NT];
        bool fl[NT];
        bool valid = true;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int j = lane + kWave * t;
            const bool in = j < n;
            lb[t] = in ? lbn[j] : 0.0;
            ub[t] = in ? ubn[j] : 0.0;
            lsc[t] = (in && P.lin) ? lsc_scale * P.lin[j] : 0.0;
            fl[t] = in && P.flags && P.flags[j];
            if (in && (isnan(lb[t]) || isnan(ub[t]) || lb[t] > ub[t]
                       || (fl[t] && ceil(lb[t]) > floor(ub[t]))))
                valid = false;
        }
        if (!wave_all(valid)) {
            if (lane == 0) {
                P.status[node] = GB_NODE_INFEASIBLE;
                P.value[node] = __builtin_inf();
                P.gap[node] = 0.0;
                P.best_val[node] = kNone;
                P.iters[node] = initial;
            }
            continue;
        }

        double best = kNone;
        // Score an integral vertex z whose product Qz is known.
        auto score = [&](const double (&z)[NT], const double (&qz)[NT]) {
            bool ok = true;
            double acc = 0.0;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const int j = lane + kWave * t;
                if (j < n) {
                    if (!isfinite(z[t]) || z[t] < lb[t] - P.ctol || z[t] > ub[t] + P.ctol
                            || (fl[t] && z[t] != round(z[t])))
                        ok = false;
                    acc = fma(z[t], fma(hs, qz[t], lsc[t]), acc);
                }
            }
            const bool all_ok = wave_all(ok);
            const double s = wave_sum_all(acc);
            if (!all_ok) return;
            double f = s + P.kappa;
            if (!isfinite(f)) return;
            if (f < best) {
                best = f;
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const int j = lane + kWave * t;
                    if (j < n) bvn[j] = z[t];
                }
            }
        };

        // Starting point: LMO vertex of the artificial direction (ub for a
        // box, since it is negative) or the supplied / warm-start iterate.
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int j = lane + kWave * t;
            x[t] = (j < n) ? ((initial == 0 && !supplied) ? ub[t] : xn[j]) : 0.0;
        }
        // A candidate from an earlier phase (the mixed-precision pass) is
        // re-scored exactly before this node's own candidates; its product
        // shares the sweep over Q with Q x.
        if (P.rescore && P.best_val[node] < kNone) {
            double z[NT], qz[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const int j = lane + kWave * t;
                z[t] = (j < n) ? bvn[j] : 0.0;
            }
            W.gemv2(z, qz, x, qx);
            score(z, qz);
        } else {
            W.gemv(x, qx);
        }
        auto grad = [&]() {
#pragma unroll
            for (int t = 0; t < NT; ++t) g[t] = P.gs * qx[t] + lsc[t];
        };
        auto fwd = [&](double (&out)[NT]) {
#pragma unroll
            for (int t = 0; t < NT; ++t) out[t] = (g[t] > 0.0) ? lb[t] : ub[t];
        };
        // In-face away vertex at x (lmo.cuh:278-295 face tolerances).
        auto away = [&](double (&out)[NT]) {
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const double ut = atol + rtol * fabs(ub[t]);
                const double lt = atol + rtol * fabs(lb[t]);
                if (fabs(x[t] - ub[t]) <= ut) out[t] = ub[t];
                else if (fabs(x[t] - lb[t]) <= lt) out[t] = lb[t];
                else out[t] = (-g[t] > 0.0) ? lb[t] : ub[t];
            }
        };
        grad();
        fwd(v);
        // DICG: the first away vertex depends only on x and g, so its
        // product shares the sweep with Q v (it is used if the node takes a
        // first step).
        bool have_a = false;
        if (P.dicg) {
            away(a);
            W.gemv2(v, qv, a, qa);
        } else {
            W.gemv(v, qv);
        }
        int iterations = initial;
        double gap = 0.0;
        int it = 0;
        for (; it < P.max_iters; ++it) {
            iterations = initial + it + 1;
            if (it > 0) {
                grad();
                double vn[NT];
                fwd(vn);
                W.update(vn, v, qv);
#pragma unroll
                for (int t = 0; t < NT; ++t) v[t] = vn[t];
            }
            score(v, qv);
            double ga = 0.0, fa = 0.0;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const int j = lane + kWave * t;
                if (j < n) {
                    ga = fma(g[t], x[t] - v[t], ga);
                    fa = fma(x[t], fma(hs, qx[t], lsc[t]), fa);
                }
            }
            gap = wave_sum_all(ga);
            const double fx = wave_sum_all(fa) + P.kappa;
            if (gap >= 0.0 && gap < eps) break;
            if (fx - gap >= cutoff) break;

            double d[NT], qd[NT], gmax;
            if (P.dicg) {
                // First step: a and Q a are those computed with Q v (x and g
                // have not changed since).
                if (have_a) {
                    double an[NT];
                    away(an);
                    W.update(an, a, qa);
#pragma unroll
                    for (int t = 0; t < NT; ++t) a[t] = an[t];
                }
                have_a = true;
                score(a, qa);
                double r = 1.0;
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const int j = lane + kWave * t;
                    const double dv = a[t] - v[t];  // direction a - v as the reference
                    if (j < n) {
                        if (dv < 0.0) r = fmin(r, (ub[t] - x[t]) / -dv);
                        else if (dv > 0.0) r = fmin(r, (x[t] - lb[t]) / dv);
                    }
                    d[t] = v[t] - a[t];
                    qd[t] = qv[t] - qa[t];
                }
                gmax = fmax(0.0, fmin(wave_min_all(r), 1.0));
            } else {
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    d[t] = v[t] - x[t];
                    qd[t] = qv[t] - qx[t];
                }
                gmax = 1.0;
            }
            double pg = 0.0, cv = 0.0;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const int j = lane + kWave * t;
                if (j < n) {
                    pg = fma(-g[t], d[t], pg);
                    cv = fma(P.gs * d[t], qd[t], cv);
                }
            }
            const double pgap = wave_sum_all(pg);
            const double curv = wave_sum_all(cv);
            const double gamma = quad_exact(pgap, curv, gmax);
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                x[t] = fma(gamma, d[t], x[t]);
                qx[t] = fma(gamma, qd[t], qx[t]);
            }
            if (P.refresh > 0 && (it + 1) % P.refresh == 0) W.gemv(x, qx);
        }

        // Direct FP64 certificate at the returned iterate.  A node that
        // stopped before its first step still holds the direct products
        // Q x and Q v of this iterate and vertex (recomputing them gives the
        // same bits), so they are reused.
        const bool fresh = it == 0;
        // The rounded candidate depends only on x: its product shares the
        // sweep with the direct Q x.
        double qc[NT];
        auto rounded = [&](double (&out)[NT]) {
#pragma unroll
            for (int t = 0; t < NT; ++t) out[t] = fl[t] ? round(x[t]) : x[t];
        };
        if (P.flags) {
            double cand[NT];
            rounded(cand);
            if (fresh) W.gemv(cand, qc);
            else W.gemv2(x, qx, cand, qc);
        } else if (!fresh) {
            W.gemv(x, qx);
        }
        grad();
        if (P.gout) {
            double* go = P.gout + (size_t)node * n;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const int j = lane + kWave * t;
                if (j < n) go[j] = g[t];
            }
        }
        if (!fresh) {
            fwd(v);
            W.gemv(v, qv);
        }
        score(v, qv);
        double ga = 0.0, fa = 0.0;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int j = lane + kWave * t;
            if (j < n) {
                ga = fma(g[t], x[t] - v[t], ga);
                fa = fma(x[t], fma(hs, qx[t], lsc[t]), fa);
            }
        }
        const double fgap = wave_sum_all(ga);
        const double fval = wave_sum_all(fa) + P.kappa;
        uint8_t st = (fgap >= 0.0 && fgap < eps) ? 1 : 0;
        bool integral = true;
        if (P.flags) {
            double cand[NT];
            rounded(cand);
            score(cand, qc);
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int j = lane + kWave * t;
            if (j < n) {
                if (!isfinite(x[t]) || (fl[t] && fabs(x[t] - round(x[t])) > P.ctol)) integral = false;
                xn[j] = x[t];
            }
        }
        if (wave_all(integral)) st = st ? GB_NODE_INTEGRAL_CONVERGED : GB_NODE_INTEGRAL_UNCONVERGED;
        if (lane == 0) {
            P.value[node] = fval;
            P.gap[node] = fgap;
            P.status[node] = st;
            P.iters[node] = iterations;
            P.best_val[node] = best;
            if (P.stats) {
                P.stats[2 * (size_t)node] = W.columns;
                P.stats[2 * (size_t)node + 1] = W.gemvs;
            }
        }
    }
}

__global__ void reset_counter_kernel(int* counter) { *counter = 0; }

// G[i] = s_g (Q x_i) + s_g s_l c for a batch of iterates: one wave per node,
// lane rows j, k-ordered FMA chain (features of the learned scorer).
__global__ void __launch_bounds__(64) batched_gradient_kernel(const double* Qt, const double* lin,
                                                              double gs, double ls, int n, int N,
                                                              const double* X, double* G) {
    const int node = blockIdx.x;
    if (node >= N) return;
    const double* x = X + (size_t)node * n;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        double acc = 0.0;
        for (int k = 0; k < n; ++k) acc = fma(Qt[(size_t)k * n + j], x[k], acc);
        G[(size_t)node * n + j] = gs * acc + (lin ? (gs * ls) * lin[j] : 0.0);
    }
}

// Feature rows of the learned scorer for every (node, variable) of a batch,
// row node * n + j; rows of variables that are not fractional candidates
// are zero.  Same definitions as the host rows of the label generator:
// 0 x_j, 1 min(frac, 1 - frac), 2 g_j / max|g|, 3 / 4 up / down gain
// (pseudocost mean gain per unit, or 1/2 s_g Q_jj d without history for
// the variable, times the distance d moved) over its mean across the node's
// candidates, 5 depth / n, 6 gap / (1 + |f|), 7 ub_j - lb_j.  pcu / pcd < 0
// means no history.  One wave per node.
__global__ void __launch_bounds__(64) candidate_features_kernel(
    int n, int N, const double* X, const double* G, const double* LB, const double* UB, const int* depth,
    const double* gap, const double* val, const double* pcu, const double* pcd, const double* qdiag,
    double gs, const uint8_t* flags, double tol, float* F) {
    const int node = blockIdx.x;
    if (node >= N) return;
    const int lane = threadIdx.x;
    const size_t base = static_cast<size_t>(node) * n;
    const double* x = X + base;
    const double* g = G + base;
    double gmax = 0.0, su = 0.0, sd = 0.0;
    int cnt = 0;
    auto is_cand = [&](int j, double& frac) {
        frac = x[j] - floor(x[j]);
        return flags[j] != 0 && fmin(frac, 1.0 - frac) > tol;
    };
    auto gains = [&](int j, double& gu, double& gd) {
        const double du = ceil(x[j]) - x[j], dd = x[j] - floor(x[j]);
        const double curv = 0.5 * gs * qdiag[j];
        gu = (pcu[j] >= 0.0 ? pcu[j] : curv * du) * du;
        gd = (pcd[j] >= 0.0 ? pcd[j] : curv * dd) * dd;
    };
    for (int j = lane; j < n; j += kWave) {
        gmax = fmax(gmax, fabs(g[j]));
        double frac;
        if (is_cand(j, frac)) {
            double gu, gd;
            gains(j, gu, gd);
            su += gu;
            sd += gd;
            ++cnt;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        gmax = fmax(gmax, __shfl_xor(gmax, off, kWave));
        su += __shfl_xor(su, off, kWave);
        sd += __shfl_xor(sd, off, kWave);
        cnt += __shfl_xor(cnt, off, kWave);
    }
    const double mu = su / (cnt ? static_cast<double>(cnt) : 1.0);
    const double md = sd / (cnt ? static_cast<double>(cnt) : 1.0);
    const double rgap = gap[node] / (1.0 + fabs(val[node]));
    for (int j = lane; j < n; j += kWave) {
        float* f = F + (base + j) * 8;
        double frac;
        if (is_cand(j, frac)) {
            double gu, gd;
            gains(j, gu, gd);
            f[0] = static_cast<float>(x[j]);
            f[1] = static_cast<float>(fmin(frac, 1.0 - frac));
            f[2] = static_cast<float>(g[j] / (gmax + 1e-12));
            f[3] = static_cast<float>(gu / (mu + 1e-12));
            f[4] = static_cast<float>(gd / (md + 1e-12));
            f[5] = static_cast<float>(depth[node]) / static_cast<float>(n);
            f[6] = static_cast<float>(rgap);
            f[7] = static_cast<float>(UB[base + j] - LB[base + j]);
        } else {
#pragma unroll
            for (int k = 0; k < 8; ++k) f[k] = 0.0f;
        }
    }
}

// One wave per node: k rounds of a wave argmax over the candidates after
// the previous pick in the order (score descending, index ascending).
__global__ void __launch_bounds__(64) topk_candidates_kernel(int n, int N, const double* X, const uint8_t* flags,
                                                             double tol, const float* score,
                                                             const uint8_t* status, int k, int* out) {
    const int node = blockIdx.x;
    if (node >= N) return;
    const int lane = threadIdx.x;
    const size_t base = static_cast<size_t>(node) * n;
    const bool open = status[node] < GB_NODE_INTEGRAL_UNCONVERGED;
    float last_s = __builtin_inff();
    int last_j = -1;
    bool more = true;
    for (int r = 0; r < k; ++r) {
        float bs = -__builtin_inff();
        int bj = 0x7FFFFFFF;
        if (more && open) {
            for (int j = lane; j < n; j += kWave) {
                const double xj = X[base + j];
                const double frac = xj - floor(xj);
                if (!flags[j] || !(fmin(frac, 1.0 - frac) > tol)) continue;
                const float sj = score[base + j];
                const bool after = sj < last_s || (sj == last_s && j > last_j);
                if (after && (sj > bs || (sj == bs && j < bj))) {
                    bs = sj;
                    bj = j;
                }
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const float os = __shfl_xor(bs, off, kWave);
            const int oj = __shfl_xor(bj, off, kWave);
            if (os > bs || (os == bs && oj < bj)) {
                bs = os;
                bj = oj;
            }
        }
        if (bj == 0x7FFFFFFF) more = false;
        if (lane == 0) out[static_cast<size_t>(node) * k + r] = more ? bj : -1;
        last_s = bs;
        last_j = bj;
    }
}

constexpr size_t kFastLdsBudget = 160 * 1024;

}  // namespace

int launch_batched_gradient(const double* Qt, const double* lin, double gs, double ls, int n, int N,
                            const double* X, double* G, void* stream) {
    if (N <= 0) return 0;
    hipLaunchKernelGGL(batched_gradient_kernel, dim3(N), dim3(64), 0,
                       reinterpret_cast<hipStream_t>(stream), Qt, lin, gs, ls, n, N, X, G);
    return (int)hipGetLastError();
}

int launch_candidate_features(int n, int N, const double* X, const double* G, const double* LB,
                              const double* UB, const int* depth, const double* gap, const double* val,
                              const double* pcu, const double* pcd, const double* qdiag, double gs,
                              const uint8_t* flags, double tol, float* F, void* stream) {
    if (N <= 0) return 0;
    hipLaunchKernelGGL(candidate_features_kernel, dim3(N), dim3(kWave), 0, reinterpret_cast<hipStream_t>(stream),
                       n, N, X, G, LB, UB, depth, gap, val, pcu, pcd, qdiag, gs, flags, tol, F);
    return static_cast<int>(hipGetLastError());
}

int launch_topk_candidates(int n, int N, const double* X, const uint8_t* flags, double tol, const float* score,
                           const uint8_t* status, int k, int* out, void* stream) {
    if (N <= 0 || k <= 0) return 0;
    hipLaunchKernelGGL(topk_candidates_kernel, dim3(N), dim3(kWave), 0, reinterpret_cast<hipStream_t>(stream),
                       n, N, X, flags, tol, score, status, k, out);
    return static_cast<int>(hipGetLastError());
}

int launch_relax_fast(const FastParams& p, int waves_per_block, int blocks, void* stream) {
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, st, p.counter);
    // per wave: 2 n_pad doubles + n_pad ints (n_pad even keeps 8-byte alignment)
    const size_t wave_lds = (size_t)waves_per_block * p.n_pad * 5 / 2 * sizeof(double);
    const size_t q_lds = (size_t)p.n * p.n * sizeof(double);
    const bool ql = q_lds + wave_lds <= kFastLdsBudget;
    const size_t lds = wave_lds + (ql ? q_lds : 0);
    const dim3 grid(blocks), block(waves_per_block * kWave);
#define GB_FAST(NT, QL)                                                                   \
    do {                                                                                  \
        auto k = relax_fast_kernel<NT, QL>;                                               \
        if (lds > 64 * 1024) {                                                            \
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),          \
                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                    \
            if (e != hipSuccess) return (int)e;                                           \
        }                                                                                 \
        hipLaunchKernelGGL(k, grid, block, lds, st, p);                                   \
    } while (0)
    if (p.n <= 64) { if (ql) GB_FAST(1, true); else GB_FAST(1, false); }
    else if (p.n <= 128) { if (ql) GB_FAST(2, true); else GB_FAST(2, false); }
    else if (p.n <= 256) GB_FAST(4, false);
    else if (p.n <= 512) GB_FAST(8, false);
    else return (int)hipErrorInvalidValue;
#undef GB_FAST
    return (int)hipGetLastError();
}

}  // namespace gb
