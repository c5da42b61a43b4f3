// Batched-form node relaxation for dense quadratics over boxes and over
// bounded simplices of 0/1 variables, gfx950.
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
#include <map>
#include <mutex>
#include <utility>

#include "gb_device.h"
#include "gb_fast.h"
#include "gb_select.h"

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
    int n_;             // problem size (a copy: the kernel argument stays in registers)
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
    __device__ __forceinline__ void load_batch(int k0, double (&qq)[kBatch][NT]) {
        const int n = n_;
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
    __device__ __forceinline__ void sweep(const double* za, const double* zb, double (&ya)[NT], double (&yb)[NT]) {
        const int n = n_;
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

    __device__ __forceinline__ void stage(double* dst, const double (&z)[NT]) {
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int j = lane + kWave * t;
            if (j < n_) dst[j] = z[t];
        }
    }

    // y = Q z for the lane's rows; z given in registers.
    __device__ __forceinline__ void gemv(const double (&z)[NT], double (&y)[NT]) {
        stage(zs, z);
        wave_sync();
        sweep<1>(zs, zs, y, y);
        wave_sync();
        ++gemvs;
    }

    // y1 = Q z1 and y2 = Q z2 in one sweep over Q (the change-list buffer
    // stages z2; no update is in progress).
    __device__ __forceinline__ void gemv2(const double (&z1)[NT], double (&y1)[NT], const double (&z2)[NT], double (&y2)[NT]) {
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
    static constexpr int kGroup = NT >= 8 ? 8 : (NT >= 4 ? 4 : 8);
    int* klist;      // this wave's LDS list of changed columns (n ints)
    double* dlist;   // their coefficients (n doubles)

    // y += sum over k with znew_k != zold_k of Q[:,k] (znew_k - zold_k), k
    // ascending (the FMA order per output coordinate is k order).
    __device__ __forceinline__ int update(const double (&znew)[NT], const double (&zold)[NT], double (&y)[NT]) {
        const int n = n_;
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
// SX: the bounded-simplex LMO (sum x <= rhs, or == rhs) over 0/1 integer
// variables instead of the box.  With 0/1 bounds and an integral rhs every
// capacity of the greedy (lmo.cuh:517-618) is 0 or 1 and every prefix mass an
// exact integer, so its vertex is: v = lb, plus 1 on the first r = rhs - sum lb
// eligible coordinates of capacity 1 in (gradient, index) order.  The kernel
// selects them with wave_select (no sort); the in-face vertex
// (lmo.cuh:814-863) and the maximum step (lmo.cuh:865-906) follow the same
// formulas with the wave's sums.  A node whose bounds are not 0/1 or with a
// variable not marked integer is not relaxed: status 0xFF.
template <int NT, bool QL, bool SX>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, NT >= 8 ? 1 : 8))) relax_fast_kernel(FastParams P) {
    extern __shared__ __align__(16) double lds_fast[];
    const int lane = lane_id();
    const int n = P.n;
    const int N = P.count ? *P.count : P.N;
    const size_t qsize = QL ? (size_t)n * n : 0;
    if (QL) {
        for (size_t i = threadIdx.x; i < qsize; i += blockDim.x) lds_fast[i] = P.Qt[i];
        __syncthreads();
    }
    // Per-wave LDS: staging vector (n_pad doubles), changed-column
    // coefficients (n_pad doubles), changed-column indices (n_pad ints).
    double* wbase = lds_fast + qsize + (size_t)wave_id() * (size_t)P.n_pad * 5 / 2;
    WaveNode<NT, QL> W{P.n, QL ? lds_fast : P.Qt, lane, wbase, 0, 0,
                       reinterpret_cast<int*>(wbase + 2 * P.n_pad), wbase + P.n_pad};
    const double hs = 0.5 * P.gs;
    const double lsc_scale = P.gs * P.ls;
    const double atol = 1e-6, rtol = 1e-4;

    while (true) {
        int node = 0;
        if (lane == 0) node = atomicAdd(P.counter, 1);
        node = __shfl(node, 0, kWave);
        if (node >= N) break;
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
        // Simplex: the 0/1 scope, then the node feasibility of
        // Lmo<GB_LMO_SIMPLEX>::feasible (lmo.cuh:373-397) with wave sums.
        double lbsum = 0.0;
        bool node_ok = wave_all(valid);
        if constexpr (SX) if (node_ok) {
            bool binary = true;
            double mand = 0.0, capsum = 0.0;
            bool cont = false;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const int j = lane + kWave * t;
                if (j < n) {
                    if (!P.is_int[j] || (lb[t] != 0.0 && lb[t] != 1.0) || (ub[t] != 0.0 && ub[t] != 1.0))
                        binary = false;
                    const bool integer = P.flags ? fl[t] : true;
                    mand += fmax(0.0, lb[t]);
                    capsum += ub[t];
                    cont = cont || !integer;
                }
            }
            if (!wave_all(binary)) {
                if (lane == 0) {
                    P.status[node] = 0xFF;
                    P.value[node] = __builtin_nan("");
                    P.gap[node] = __builtin_nan("");
                    P.best_val[node] = kNone;
                    P.iters[node] = initial;
                }
                continue;
            }
            const double ms = wave_sum_all(mand);
            bool feas = ms <= P.rhs + 1e-9;
            if (P.equality) {
                const double cs = wave_sum_all(capsum);
                const bool hc = !wave_all(!cont);
                feas = feas && cs >= P.rhs - 1e-9 && (hc || fabs(P.rhs - round(P.rhs)) <= 1e-9);
            }
            node_ok = feas;
            lbsum = ms;  // lb is 0/1: sum of lb == sum of max(0, lb)
        }
        if (!node_ok) {
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

        // Simplex vertex for the direction h: lb plus the first rhs - sum lb
        // coordinates of capacity 1 among the eligible ones (h <= 0, or all
        // of them for an equality) in (h, index) order.
        auto sx_vertex = [&](const double (&h)[NT], double (&out)[NT]) {
            unsigned long long key[NT];
            bool c[NT], sel[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const int j = lane + kWave * t;
                key[t] = order_key(h[t]);
                c[t] = j < n && ub[t] > lb[t] && (P.equality || h[t] <= 0.0);
            }
            wave_select<NT>(key, c, P.rhs - lbsum, sel, lane);
#pragma unroll
            for (int t = 0; t < NT; ++t) out[t] = sel[t] ? ub[t] : lb[t];
        };
        // Starting point: LMO vertex of the artificial direction
        // g_j = -1 - j / (n + 1) (ub for a box, since it is negative) or the
        // supplied / warm-start iterate.
        if (SX && initial == 0 && !supplied) {
            double h[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) h[t] = -1.0 - (double)(lane + kWave * t) / (n + 1.0);
            sx_vertex(h, x);
        } else {
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const int j = lane + kWave * t;
                x[t] = (j < n) ? ((initial == 0 && !supplied) ? ub[t] : xn[j]) : 0.0;
            }
            if constexpr (SX) {
                // A supplied point (an FP32 iterate of the mixed pass, say) is
                // clipped to the node's bounds, coordinates within the face
                // tolerance of a bound are put on it (so that the in-face
                // steps see exact faces), and the free coordinates are scaled
                // onto the mass limit when the point exceeds it (or, for an
                // equality, falls short of it); all of them when the free ones
                // cannot absorb it.
                bool fr[NT];
                double sx = 0.0, dn = 0.0, up = 0.0, sl = 0.0, su = 0.0;
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    fr[t] = false;
                    if (lane + kWave * t < n) {
                        double v = fmin(fmax(x[t], lb[t]), ub[t]);
                        if (fabs(v - lb[t]) <= atol + rtol * fmax(fabs(v), fabs(lb[t]))) v = lb[t];
                        else if (fabs(v - ub[t]) <= atol + rtol * fmax(fabs(v), fabs(ub[t]))) v = ub[t];
                        else fr[t] = true;
                        x[t] = v;
                        sx += v;
                        sl += lb[t];
                        su += ub[t];
                        if (fr[t]) {
                            dn += v - lb[t];
                            up += ub[t] - v;
                        }
                    }
                }
                const double tx = wave_sum_all(sx), tdn = wave_sum_all(dn), tup = wave_sum_all(up);
                const double tl = wave_sum_all(sl), tu = wave_sum_all(su);
                if (tx > P.rhs) {
                    const double e = tx - P.rhs;
                    if (tdn > e) {
                        const double f = e / tdn;
#pragma unroll
                        for (int t = 0; t < NT; ++t)
                            if (fr[t]) x[t] = fma(lb[t] - x[t], f, x[t]);
                    } else {
                        const double f = (P.rhs - tl) / (tx - tl);
#pragma unroll
                        for (int t = 0; t < NT; ++t) x[t] = fma(x[t] - lb[t], f, lb[t]);
                    }
                } else if (P.equality && tx < P.rhs) {
                    const double sh = P.rhs - tx;
                    if (tup > sh) {
                        const double f = sh / tup;
#pragma unroll
                        for (int t = 0; t < NT; ++t)
                            if (fr[t]) x[t] = fma(ub[t] - x[t], f, x[t]);
                    } else {
                        const double f = sh / (tu - tx);
#pragma unroll
                        for (int t = 0; t < NT; ++t) x[t] = fma(ub[t] - x[t], f, x[t]);
                    }
                }
            }
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
            if constexpr (SX) {
                sx_vertex(g, out);
            } else {
#pragma unroll
                for (int t = 0; t < NT; ++t) out[t] = (g[t] > 0.0) ? lb[t] : ub[t];
            }
        };
        // In-face away vertex at x: the LMO over the minimal face of x for
        // the direction -g (box: lmo.cuh:278-295 face tolerances; simplex:
        // lmo.cuh:814-863, coordinates at a bound stay fixed and the face
        // keeps the full mass when sum x == rhs within tolerance).
        auto away = [&](double (&out)[NT]) {
            if constexpr (SX) {
                unsigned long long key[NT];
                bool c[NT], sel[NT];
                double m = 0.0, xm = 0.0;
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const int j = lane + kWave * t;
                    bool fixed = false;
                    out[t] = lb[t];
                    if (j < n) {
                        if (fabs(x[t] - lb[t]) <= atol + rtol * fmax(fabs(x[t]), fabs(lb[t]))) {
                            fixed = true;
                        } else if (fabs(x[t] - ub[t]) <= atol + rtol * fmax(fabs(x[t]), fabs(ub[t]))) {
                            fixed = true;
                            out[t] = ub[t];
                        }
                        m += out[t];
                        xm += x[t];
                    }
                    key[t] = order_key(-g[t]);
                    c[t] = j < n && !fixed && ub[t] > out[t];
                }
                const double mass = wave_sum_all(m);
                const double xmass = wave_sum_all(xm);
                const bool force = P.equality || fabs(xmass - P.rhs) <= atol + rtol * fmax(fabs(xmass), fabs(P.rhs));
#pragma unroll
                for (int t = 0; t < NT; ++t) c[t] = c[t] && (force || -g[t] <= 0.0);
                wave_select<NT>(key, c, P.rhs - mass, sel, lane);
#pragma unroll
                for (int t = 0; t < NT; ++t) out[t] = sel[t] ? ub[t] : out[t];
            } else {
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const double ut = atol + rtol * fabs(ub[t]);
                    const double lt = atol + rtol * fabs(lb[t]);
                    if (fabs(x[t] - ub[t]) <= ut) out[t] = ub[t];
                    else if (fabs(x[t] - lb[t]) <= lt) out[t] = lb[t];
                    else out[t] = (-g[t] > 0.0) ? lb[t] : ub[t];
                }
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
                double r = 1.0, sx = 0.0, sd = 0.0;
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const int j = lane + kWave * t;
                    const double dv = a[t] - v[t];  // direction a - v as the reference
                    if (j < n) {
                        if constexpr (SX) {
                            if (dv > 1e-6) r = fmin(r, (x[t] - lb[t]) / dv);
                            else if (dv < -1e-6) r = fmin(r, (ub[t] - x[t]) / -dv);
                            sx += x[t];
                            sd += dv;
                        } else {
                            if (dv < 0.0) r = fmin(r, (ub[t] - x[t]) / -dv);
                            else if (dv > 0.0) r = fmin(r, (x[t] - lb[t]) / dv);
                        }
                    }
                    d[t] = v[t] - a[t];
                    qd[t] = qv[t] - qa[t];
                }
                if constexpr (SX) {
                    // coordinate ratios, then the mass limit (lmo.cuh:865-906)
                    const double coord = wave_min_all(r);
                    const double tx = wave_sum_all(sx);
                    const double td = wave_sum_all(sd);
                    double agg;
                    if (P.equality && fabs(td) > 1e-9) agg = 0.0;
                    else agg = td < 0.0 ? fmin(1.0, (tx - P.rhs) / td) : 1.0;
                    gmax = fmax(0.0, fmin(coord, agg));
                } else {
                    gmax = fmax(0.0, fmin(wave_min_all(r), 1.0));
                }
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
            bool take = true;
            if constexpr (SX) {  // Lmo<GB_LMO_SIMPLEX>::rounded_feasible
                bool nonneg = true;
                double sum = 0.0;
#pragma unroll
                for (int t = 0; t < NT; ++t)
                    if (lane + kWave * t < n) {
                        nonneg = nonneg && cand[t] >= -P.ctol;
                        sum += cand[t];
                    }
                const double tot = wave_sum_all(sum);
                take = wave_all(nonneg)
                    && (P.equality ? fabs(tot - P.rhs) <= P.ctol : tot <= P.rhs + P.ctol);
            }
            if (take) score(cand, qc);
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

// The same relaxation with one node per workgroup and its rows split over
// NT waves: wave w holds row j = lane + 64 w (the t = w slice of
// relax_fast_kernel's lane).  Each row's operations and each product's FMA
// order over k are those of relax_fast_kernel.  A reduction gathers every
// wave's per-row factors through LDS, and every wave then evaluates the same
// lane chain over t and the same wave tree, so results are bit-identical to
// relax_fast_kernel's and all waves take the same branches.  With one row
// per lane a wave keeps 8 Q columns in flight per L2 round trip, and every
// node of a batch of up to a few thousand is in flight at once.
constexpr int kRowSums = 3;  // fma-chain sums per exchange

// Every node of a batch is meant to be in flight at once (the launcher uses
// this layout only then), so the kernel asks for four waves per SIMD.
template <int NT, bool SX = false>
__global__ void __launch_bounds__(64 * NT, 4) relax_fast_rows_kernel(FastParams P) {
    extern __shared__ __align__(16) double lds_fast[];
    const int lane = lane_id();
    const int w = __builtin_amdgcn_readfirstlane(wave_id());
    const int n = P.n;
    const int N = P.count ? *P.count : P.N;
    const int np = P.n_pad;
    const int j = lane + kWave * w;  // this lane's row
    const bool in = j < n;
    const double* Q = P.Qt;
    double* zs = lds_fast;             // staged vector (np)
    double* zs2 = zs + np;             // second staged vector (np)
    double* dlist = zs2 + np;          // changed-column coefficients (np)
    int* klist = reinterpret_cast<int*>(dlist + np);  // changed columns (np ints)
    double* xf = dlist + np + np / 2;  // 2 buffers x NT x 64 x (2 kRowSums + 1)
    constexpr int kSlot = 2 * kRowSums + 1;
    int* xok = reinterpret_cast<int*>(xf + 2 * NT * kWave * kSlot);  // 2 buffers x NT x 64 x 2
    int* misc = xok + 2 * NT * kWave * 2;                            // counts[NT], node
    const double hs = 0.5 * P.gs;
    const double lsc_scale = P.gs * P.ls;
    const double atol = 1e-6, rtol = 1e-4;
    int xbuf = 0;

    // K fma-chain sums of (a_k, b_k) over the rows j = lane + 64 t < n, t
    // ascending, each followed by the wave tree; optionally a min chain from
    // 1.0 (entries +inf where a row has none) and two AND flags.
    struct Ex {
        double a[kRowSums], b[kRowSums], mn;
        bool ok, ok2;
        double sum[kRowSums], min;
        bool all_ok, all_ok2;
    };
    auto exchange = [&](Ex& e, int K, bool with_min, bool with_ok) {
        double* X = xf + (size_t)xbuf * NT * kWave * kSlot;
        int* O = xok + xbuf * NT * kWave * 2;
        double* mine = X + ((size_t)w * kWave + lane) * kSlot;
#pragma unroll
        for (int k = 0; k < kRowSums; ++k)
            if (k < K) {
                mine[2 * k] = e.a[k];
                mine[2 * k + 1] = e.b[k];
            }
        if (with_min) mine[2 * kRowSums] = e.mn;
        if (with_ok) {
            O[(w * kWave + lane) * 2] = e.ok ? 1 : 0;
            O[(w * kWave + lane) * 2 + 1] = e.ok2 ? 1 : 0;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kRowSums; ++k)
            if (k < K) {
                double acc = 0.0;
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const double* s = X + ((size_t)t * kWave + lane) * kSlot;
                    if (lane + kWave * t < n) acc = fma(s[2 * k], s[2 * k + 1], acc);
                }
                e.sum[k] = wave_sum_all(acc);
            }
        if (with_min) {
            double r = 1.0;
#pragma unroll
            for (int t = 0; t < NT; ++t) r = fmin(r, X[((size_t)t * kWave + lane) * kSlot + 2 * kRowSums]);
            e.min = wave_min_all(r);
        }
        if (with_ok) {
            bool o = true, o2 = true;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                o = o && O[(t * kWave + lane) * 2] != 0;
                o2 = o2 && O[(t * kWave + lane) * 2 + 1] != 0;
            }
            e.all_ok = wave_all(o);
            e.all_ok2 = wave_all(o2);
        }
        xbuf ^= 1;
    };

    long long columns = 0, gemvs = 0;
    constexpr int kB = 8;  // Q columns in flight per round trip
    // ya = Q za (and yb = Q zb) for this lane's row; za, zb staged in LDS.
    auto sweep = [&](bool two, double& ya, double& yb) {
        double acc0 = 0.0, acc1 = 0.0;
        for (int k0 = 0; k0 < n; k0 += kB) {
            double qq[kB];
#pragma unroll
            for (int u = 0; u < kB; ++u) qq[u] = (in && k0 + u < n) ? Q[(size_t)(k0 + u) * n + j] : 0.0;
#pragma unroll
            for (int u = 0; u < kB; ++u) {
                const int k = k0 + u;
                if (k < n) {
                    acc0 = fma(qq[u], zs[k], acc0);
                    if (two) acc1 = fma(qq[u], zs2[k], acc1);
                }
            }
        }
        ya = in ? acc0 : 0.0;
        if (two) yb = in ? acc1 : 0.0;
    };
    auto gemv = [&](double z, double& y) {
        if (in) zs[j] = z;
        __syncthreads();
        double unused = 0.0;
        sweep(false, y, unused);
        __syncthreads();
        ++gemvs;
    };
    auto gemv2 = [&](double z1, double& y1, double z2, double& y2) {
        if (in) {
            zs[j] = z1;
            zs2[j] = z2;
        }
        __syncthreads();
        sweep(true, y1, y2);
        __syncthreads();
        gemvs += 2;
    };
    // y += sum over changed k (znew_k != zold_k), k ascending, of
    // Q[:,k] (znew_k - zold_k): the list is compacted in wave order.
    auto update = [&](double znew, double zold, double& y) {
        const bool ch = in && znew != zold;
        const unsigned long long mask = __ballot(ch);
        if (lane == 0) misc[w] = __popcll(mask);
        __syncthreads();
        int off = 0, count = 0;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int c = misc[t];
            if (t < w) off += c;
            count += c;
        }
        if (ch) {
            const int pos = off + __popcll(mask & ((1ull << lane) - 1ull));
            klist[pos] = j;
            dlist[pos] = znew - zold;
        }
        __syncthreads();
        for (int g0 = 0; g0 < count; g0 += kB) {
            double c[kB], dd[kB];
#pragma unroll
            for (int g = 0; g < kB; ++g) {
                const bool live = g0 + g < count;
                const int k = live ? klist[g0 + g] : 0;
                dd[g] = live ? dlist[g0 + g] : 0.0;
                c[g] = (live && in) ? Q[(size_t)k * n + j] : 0.0;
            }
#pragma unroll
            for (int g = 0; g < kB; ++g)
                if (g0 + g < count && in) y = fma(c[g], dd[g], y);
        }
        __syncthreads();
        columns += count;
    };
    // The updates of Q v and Q a in one gather: the union of their changed
    // columns, k ascending, each column read once and applied to the
    // accumulators whose coefficient is nonzero, so each keeps the FMA chain
    // of its own update (the second coefficient list reuses zs2).
    auto update2 = [&](double vnew, double vold, double& yv, double anew, double aold, double& ya) {
        const bool cv = in && vnew != vold;
        const bool ca = in && anew != aold;
        const unsigned long long mask = __ballot(cv || ca);
        if (lane == 0) misc[w] = __popcll(mask);
        __syncthreads();
        int off = 0, count = 0;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int c = misc[t];
            if (t < w) off += c;
            count += c;
        }
        if (cv || ca) {
            const int pos = off + __popcll(mask & ((1ull << lane) - 1ull));
            klist[pos] = j;
            dlist[pos] = cv ? vnew - vold : 0.0;
            zs2[pos] = ca ? anew - aold : 0.0;
        }
        __syncthreads();
        for (int g0 = 0; g0 < count; g0 += kB) {
            double c[kB], dv[kB], da[kB];
#pragma unroll
            for (int g = 0; g < kB; ++g) {
                const bool live = g0 + g < count;
                const int k = live ? klist[g0 + g] : 0;
                dv[g] = live ? dlist[g0 + g] : 0.0;
                da[g] = live ? zs2[g0 + g] : 0.0;
                c[g] = (live && in) ? Q[(size_t)k * n + j] : 0.0;
            }
#pragma unroll
            for (int g = 0; g < kB; ++g)
                if (g0 + g < count && in) {
                    if (dv[g] != 0.0) yv = fma(c[g], dv[g], yv);
                    if (da[g] != 0.0) ya = fma(c[g], da[g], ya);
                }
        }
        __syncthreads();
        columns += count;
    };

    // Simplex (SX): the greedy's selection over the node's rows.  Every wave
    // stages its rows' keys and flags in LDS (the spaces of the changed-column
    // lists, free between updates), then runs the same wave_select over all
    // rows, so every wave holds the same result; each lane keeps its row's.
    // want: rhs - sum lb (forward), or rhs - the mass of the rows flagged
    // `base` (in-face) when use_mass.
    unsigned long long* skey = reinterpret_cast<unsigned long long*>(dlist);
    uint8_t* sflag = reinterpret_cast<uint8_t*>(klist);
    auto select_rows = [&](double h, bool cand, bool base, double want, bool use_mass) -> bool {
        if (in) {
            skey[j] = order_key(h);
            sflag[j] = static_cast<uint8_t>((cand ? 1u : 0u) | (base ? 2u : 0u));
        }
        __syncthreads();
        unsigned long long key[NT];
        bool c[NT], sel[NT];
        int mass = 0;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int jj = lane + kWave * t;
            const bool ii = jj < n;
            key[t] = ii ? skey[jj] : 0ull;
            const uint32_t f = ii ? sflag[jj] : 0u;
            c[t] = (f & 1u) != 0u;
            mass += __popcll(__ballot((f & 2u) != 0u));
        }
        wave_select<NT>(key, c, use_mass ? want - mass : want, sel, lane);
        bool mine = false;
#pragma unroll
        for (int t = 0; t < NT; ++t)
            if (t == w) mine = sel[t];
        __syncthreads();
        return mine;
    };

    while (true) {
        if (threadIdx.x == 0) misc[NT] = atomicAdd(P.counter, 1);
        __syncthreads();
        const int node = misc[NT];
        if (node >= N) break;
        columns = 0;
        gemvs = 0;

        const double* lbn = P.lb + (size_t)node * n;
        const double* ubn = P.ub + (size_t)node * n;
        double* xn = P.x + (size_t)node * n;
        double* bvn = P.best_v + (size_t)node * n;
        const double eps = P.eps[node];
        const double cutoff = P.cutoff ? P.cutoff[node] : __builtin_inf();
        const int initial = P.iters[node];
        const bool supplied = (P.status[node] & GB_NODE_INITIAL_POINT_SUPPLIED) != 0;

        const double lb = in ? lbn[j] : 0.0;
        const double ub = in ? ubn[j] : 0.0;
        const double lsc = (in && P.lin) ? lsc_scale * P.lin[j] : 0.0;
        const bool fl = in && P.flags && P.flags[j];
        Ex e{};
        e.ok = !(in && (isnan(lb) || isnan(ub) || lb > ub || (fl && ceil(lb) > floor(ub))));
        e.ok2 = true;
        double lbsum = 0.0;
        if constexpr (SX) {
            // the 0/1 scope, and Lmo<GB_LMO_SIMPLEX>::feasible with the
            // one-wave kernel's sums (exact: the values are 0 and 1)
            e.ok2 = !(in && (!P.is_int[j] || (lb != 0.0 && lb != 1.0) || (ub != 0.0 && ub != 1.0)));
            e.a[0] = in ? fmax(0.0, lb) : 0.0;
            e.b[0] = 1.0;
            e.a[1] = in ? ub : 0.0;
            e.b[1] = 1.0;
            exchange(e, 2, false, true);
        } else {
            exchange(e, 0, false, true);
        }
        if (SX && e.all_ok && !e.all_ok2) {
            if (threadIdx.x == 0) {
                P.status[node] = 0xFF;
                P.value[node] = __builtin_nan("");
                P.gap[node] = __builtin_nan("");
                P.best_val[node] = kNone;
                P.iters[node] = initial;
            }
            continue;
        }
        if (SX && e.all_ok) {
            Ex h{};
            h.ok = !(in && P.flags && !fl);  // all integer (no continuous variable)
            h.ok2 = true;
            exchange(h, 0, false, true);
            const double ms = e.sum[0];
            bool feas = ms <= P.rhs + 1e-9;
            if (P.equality) feas = feas && e.sum[1] >= P.rhs - 1e-9 && (!h.all_ok || fabs(P.rhs - round(P.rhs)) <= 1e-9);
            e.all_ok = feas;
            lbsum = ms;
        }
        if (!e.all_ok) {
            if (threadIdx.x == 0) {
                P.status[node] = GB_NODE_INFEASIBLE;
                P.value[node] = __builtin_inf();
                P.gap[node] = 0.0;
                P.best_val[node] = kNone;
                P.iters[node] = initial;
            }
            continue;
        }

        double best = kNone;
        // Score factors of an integral vertex z with known Q z; the sum goes
        // in slot `k` of an exchange, the feasibility flag in ok.
        auto score_in = [&](Ex& ex, int k, double z, double qz) {
            ex.ok = !(in && (!isfinite(z) || z < lb - P.ctol || z > ub + P.ctol || (fl && z != round(z))));
            ex.a[k] = z;
            ex.b[k] = fma(hs, qz, lsc);
        };
        auto score_out = [&](const Ex& ex, int k, double z) {
            if (!ex.all_ok) return;
            const double f = ex.sum[k] + P.kappa;
            if (!isfinite(f)) return;
            if (f < best) {
                best = f;
                if (in) bvn[j] = z;
            }
        };

        double x = in ? ((initial == 0 && !supplied) ? ub : xn[j]) : 0.0;
        if constexpr (SX) {
            if (initial == 0 && !supplied) {
                // the greedy vertex of the artificial direction g_j = -1 - j / (n + 1)
                const double h = -1.0 - (double)j / (n + 1.0);
                const bool sel = select_rows(h, in && ub > lb && (P.equality || h <= 0.0), false, P.rhs - lbsum, false);
                x = in ? (sel ? ub : lb) : 0.0;
            } else {
                // a supplied point: clipped, put on its faces and on the mass
                // limit (the one-wave kernel's sums and formulas)
                bool fr = false;
                if (in) {
                    double vv = fmin(fmax(x, lb), ub);
                    if (fabs(vv - lb) <= atol + rtol * fmax(fabs(vv), fabs(lb))) vv = lb;
                    else if (fabs(vv - ub) <= atol + rtol * fmax(fabs(vv), fabs(ub))) vv = ub;
                    else fr = true;
                    x = vv;
                }
                Ex m1{}, m2{};
                m1.a[0] = in ? x : 0.0;
                m1.b[0] = 1.0;
                m1.a[1] = fr ? x - lb : 0.0;
                m1.b[1] = 1.0;
                m1.a[2] = fr ? ub - x : 0.0;
                m1.b[2] = 1.0;
                exchange(m1, 3, false, false);
                m2.a[0] = in ? lb : 0.0;
                m2.b[0] = 1.0;
                m2.a[1] = in ? ub : 0.0;
                m2.b[1] = 1.0;
                exchange(m2, 2, false, false);
                const double tx = m1.sum[0], tdn = m1.sum[1], tup = m1.sum[2], tl = m2.sum[0], tu = m2.sum[1];
                if (tx > P.rhs) {
                    const double e0 = tx - P.rhs;
                    if (tdn > e0) {
                        if (fr) x = fma(lb - x, e0 / tdn, x);
                    } else {
                        x = fma(x - lb, (P.rhs - tl) / (tx - tl), lb);
                    }
                } else if (P.equality && tx < P.rhs) {
                    const double sh = P.rhs - tx;
                    if (tup > sh) {
                        if (fr) x = fma(ub - x, sh / tup, x);
                    } else {
                        x = fma(ub - x, sh / (tu - tx), x);
                    }
                }
                if (!in) x = 0.0;
            }
        }
        double qx = 0.0, v, qv = 0.0, a = 0.0, qa = 0.0, g;
        if (P.rescore && P.best_val[node] < kNone) {
            const double z = in ? bvn[j] : 0.0;
            double qz = 0.0;
            gemv2(z, qz, x, qx);
            Ex s{};
            score_in(s, 0, z, qz);
            s.ok2 = true;
            exchange(s, 1, false, true);
            score_out(s, 0, z);
        } else {
            gemv(x, qx);
        }
        auto grad = [&]() { g = P.gs * qx + lsc; };
        double txm = 0.0;  // simplex: sum x at the last in-face vertex (the mass limit reuses it)
        auto fwd = [&]() {
            if constexpr (SX) {
                const bool sel = select_rows(g, in && ub > lb && (P.equality || g <= 0.0), false, P.rhs - lbsum, false);
                return sel ? ub : lb;
            } else {
                return (g > 0.0) ? lb : ub;
            }
        };
        auto away = [&]() {
            if constexpr (SX) {
                // in-face vertex for -g (lmo.cuh:814-863): rows at a bound stay,
                // the face keeps the full mass when sum x == rhs within tolerance
                bool fixed = false;
                double base = lb;
                if (in) {
                    if (fabs(x - lb) <= atol + rtol * fmax(fabs(x), fabs(lb))) {
                        fixed = true;
                    } else if (fabs(x - ub) <= atol + rtol * fmax(fabs(x), fabs(ub))) {
                        fixed = true;
                        base = ub;
                    }
                }
                Ex mx{};
                mx.a[0] = in ? x : 0.0;
                mx.b[0] = 1.0;
                exchange(mx, 1, false, false);
                txm = mx.sum[0];
                const bool force = P.equality || fabs(txm - P.rhs) <= atol + rtol * fmax(fabs(txm), fabs(P.rhs));
                const bool cand = in && !fixed && ub > base && (force || -g <= 0.0);
                const bool sel = select_rows(-g, cand, in && base != 0.0, P.rhs, true);
                return in ? (sel ? ub : base) : 0.0;
            } else {
                const double ut = atol + rtol * fabs(ub);
                const double lt = atol + rtol * fabs(lb);
                if (fabs(x - ub) <= ut) return ub;
                if (fabs(x - lb) <= lt) return lb;
                return (-g > 0.0) ? lb : ub;
            }
        };
        grad();
        v = fwd();
        bool have_a = false;
        if (P.dicg) {
            a = away();
            gemv2(v, qv, a, qa);
        } else {
            gemv(v, qv);
        }
        int iterations = initial;
        int it = 0;
        for (; it < P.max_iters; ++it) {
            iterations = initial + it + 1;
            bool a_done = false;
            if (it > 0) {
                grad();
                const double vn = fwd();
                if (P.dicg && have_a) {
                    // the away vertex depends on x and g only: its update
                    // joins v's (unused if the iteration stops below)
                    const double an = away();
                    update2(vn, v, qv, an, a, qa);
                    a = an;
                    a_done = true;
                } else {
                    update(vn, v, qv);
                }
                v = vn;
            }
            // score(v), gap and f(x) in one exchange
            Ex s{};
            score_in(s, 0, v, qv);
            s.ok2 = true;
            s.a[1] = g;
            s.b[1] = x - v;
            s.a[2] = x;
            s.b[2] = fma(hs, qx, lsc);
            exchange(s, 3, false, true);
            score_out(s, 0, v);
            const double gap = s.sum[1];
            const double fx = s.sum[2] + P.kappa;
            if (gap >= 0.0 && gap < eps) break;
            if (fx - gap >= cutoff) break;

            double d, qd, gmax;
            Ex r{};
            if (P.dicg) {
                if (have_a && !a_done) {
                    const double an = away();
                    update(an, a, qa);
                    a = an;
                }
                have_a = true;
                // score(a), the DICG max step and the step sums in one exchange
                score_in(r, 0, a, qa);
                r.ok2 = true;
                const double dv = a - v;  // direction a - v as the reference
                if constexpr (SX) {
                    r.mn = !in ? __builtin_inf()
                           : dv > 1e-6 ? (x - lb) / dv
                           : dv < -1e-6 ? (ub - x) / -dv
                                        : __builtin_inf();
                } else {
                    r.mn = !in ? __builtin_inf()
                           : dv < 0.0 ? (ub - x) / -dv
                           : dv > 0.0 ? (x - lb) / dv
                                      : __builtin_inf();
                }
                d = v - a;
                qd = qv - qa;
                r.a[1] = -g;
                r.b[1] = d;
                r.a[2] = P.gs * d;
                r.b[2] = qd;
                exchange(r, 3, true, true);
                score_out(r, 0, a);
                if constexpr (SX) {
                    // the mass limit (lmo.cuh:865-906): sum x is the in-face
                    // vertex's (x has not moved since), sum (a - v) an exact
                    // integer from the vertices' counts
                    const int cnt_a = __popcll(__ballot(in && a != 0.0));
                    const int cnt_v = __popcll(__ballot(in && v != 0.0));
                    misc[NT + 1 + 2 * w] = cnt_a;
                    misc[NT + 2 + 2 * w] = cnt_v;
                    __syncthreads();
                    int ta = 0, tv = 0;
#pragma unroll
                    for (int t = 0; t < NT; ++t) {
                        ta += misc[NT + 1 + 2 * t];
                        tv += misc[NT + 2 + 2 * t];
                    }
                    __syncthreads();
                    const double td = (double)(ta - tv);
                    double agg;
                    if (P.equality && fabs(td) > 1e-9) agg = 0.0;
                    else agg = td < 0.0 ? fmin(1.0, (txm - P.rhs) / td) : 1.0;
                    gmax = fmax(0.0, fmin(r.min, agg));
                } else {
                    gmax = fmax(0.0, fmin(r.min, 1.0));
                }
            } else {
                d = v - x;
                qd = qv - qx;
                r.a[1] = -g;
                r.b[1] = d;
                r.a[2] = P.gs * d;
                r.b[2] = qd;
                // slot 0 unused: a zero pair adds nothing to a chain that is not read
                r.a[0] = 0.0;
                r.b[0] = 0.0;
                exchange(r, 3, false, false);
                gmax = 1.0;
            }
            const double gamma = quad_exact(r.sum[1], r.sum[2], gmax);
            x = fma(gamma, d, x);
            qx = fma(gamma, qd, qx);
            if (P.refresh > 0 && (it + 1) % P.refresh == 0) gemv(x, qx);
        }

        // Direct FP64 certificate at the returned iterate (see relax_fast_kernel).
        const bool fresh = it == 0;
        double qc = 0.0;
        const double cand = fl ? round(x) : x;
        if (P.flags) {
            if (fresh) gemv(cand, qc);
            else gemv2(x, qx, cand, qc);
        } else if (!fresh) {
            gemv(x, qx);
        }
        grad();
        if (P.gout && in) P.gout[(size_t)node * n + j] = g;
        if (!fresh) {
            v = fwd();
            gemv(v, qv);
        }
        Ex s{};
        score_in(s, 0, v, qv);
        s.ok2 = true;
        s.a[1] = g;
        s.b[1] = x - v;
        s.a[2] = x;
        s.b[2] = fma(hs, qx, lsc);
        exchange(s, 3, false, true);
        score_out(s, 0, v);
        const double fgap = s.sum[1];
        const double fval = s.sum[2] + P.kappa;
        uint8_t st = (fgap >= 0.0 && fgap < eps) ? 1 : 0;
        bool take = true;
        if (SX && P.flags) {  // Lmo<GB_LMO_SIMPLEX>::rounded_feasible
            Ex rf{};
            rf.a[0] = in ? cand : 0.0;
            rf.b[0] = 1.0;
            rf.ok = !(in && !(cand >= -P.ctol));
            rf.ok2 = true;
            exchange(rf, 1, false, true);
            const double tot = rf.sum[0];
            take = rf.all_ok && (P.equality ? fabs(tot - P.rhs) <= P.ctol : tot <= P.rhs + P.ctol);
        }
        Ex c{};
        if (P.flags) score_in(c, 0, cand, qc);
        else c.ok = true;
        c.ok2 = !(in && (!isfinite(x) || (fl && fabs(x - round(x)) > P.ctol)));
        exchange(c, P.flags ? 1 : 0, false, true);
        if (P.flags && take) score_out(c, 0, cand);
        if (in) xn[j] = x;
        if (c.all_ok2) st = st ? GB_NODE_INTEGRAL_CONVERGED : GB_NODE_INTEGRAL_UNCONVERGED;
        if (threadIdx.x == 0) {
            P.value[node] = fval;
            P.gap[node] = fgap;
            P.status[node] = st;
            P.iters[node] = iterations;
            P.best_val[node] = best;
            if (P.stats) {
                P.stats[2 * (size_t)node] = columns;
                P.stats[2 * (size_t)node + 1] = gemvs;
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
// Q_jj is read as qdiag[j * qstride] (Qt itself with qstride = n + 1).
// 0 x_j, 1 min(frac, 1 - frac), 2 g_j / max|g|, 3 / 4 up / down gain
// (pseudocost mean gain per unit, or 1/2 s_g Q_jj d without history for
// the variable, times the distance d moved) over its mean across the node's
// candidates, 5 depth / n, 6 gap / (1 + |f|), 7 ub_j - lb_j.  pcu / pcd < 0
// means no history.  One wave per node.
__global__ void __launch_bounds__(64) candidate_features_kernel(
    int n, int N, const double* X, const double* G, const double* LB, const double* UB, const int* depth,
    const double* gap, const double* val, const double* pcu, const double* pcd, const double* qdiag,
    size_t qstride, double gs, const uint8_t* flags, double tol, float* F) {
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
        const double curv = 0.5 * gs * qdiag[static_cast<size_t>(j) * qstride];
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
                                                             const uint8_t* status, int k, int* out,
                                                             double* xout) {
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
        if (lane == 0) {
            out[static_cast<size_t>(node) * k + r] = more ? bj : -1;
            if (xout) xout[static_cast<size_t>(node) * k + r] = more ? X[base + bj] : 0.0;
        }
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
                              const double* pcu, const double* pcd, const double* qdiag, size_t qstride,
                              double gs, const uint8_t* flags, double tol, float* F, void* stream) {
    if (N <= 0) return 0;
    hipLaunchKernelGGL(candidate_features_kernel, dim3(N), dim3(kWave), 0, reinterpret_cast<hipStream_t>(stream),
                       n, N, X, G, LB, UB, depth, gap, val, pcu, pcd, qdiag, qstride, gs, flags, tol, F);
    return static_cast<int>(hipGetLastError());
}

// lb[i] = lb8[i], ub[i] = ub8[i] for i < count (0/1 node bounds sent as bytes).
__global__ void __launch_bounds__(256) expand_bounds_kernel(size_t count, const uint8_t* lb8, const uint8_t* ub8,
                                                            double* lb, double* ub) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) {
        lb[i] = static_cast<double>(lb8[i]);
        ub[i] = static_cast<double>(ub8[i]);
    }
}

int launch_expand_bounds(size_t count, const uint8_t* lb8, const uint8_t* ub8, double* lb, double* ub,
                         void* stream) {
    if (count == 0) return 0;
    const size_t want = (count + 255) / 256;
    const unsigned blocks = static_cast<unsigned>(want < 2048 ? want : 2048);
    hipLaunchKernelGGL(expand_bounds_kernel, dim3(blocks), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       count, lb8, ub8, lb, ub);
    return static_cast<int>(hipGetLastError());
}

int launch_topk_candidates(int n, int N, const double* X, const uint8_t* flags, double tol, const float* score,
                           const uint8_t* status, int k, int* out, double* xout, void* stream) {
    if (N <= 0 || k <= 0) return 0;
    hipLaunchKernelGGL(topk_candidates_kernel, dim3(N), dim3(kWave), 0, reinterpret_cast<hipStream_t>(stream),
                       n, N, X, flags, tol, score, status, k, out, xout);
    return static_cast<int>(hipGetLastError());
}

// Start points from the device iterate store.  src[i] is the address of a
// stored iterate of n doubles, with bit 0 set when it is to be clipped to
// node i's box (a parent's iterate) and clear when it is copied as is (the
// node's own iterate), or 0 for the node's lower bound.
__global__ void __launch_bounds__(128) gather_starts_kernel(int n, const unsigned long long* src, const double* lb,
                                                            const double* ub, double* x) {
    const int node = blockIdx.x;
    const unsigned long long a = src[node];
    const size_t base = static_cast<size_t>(node) * n;
    const double* p = reinterpret_cast<const double*>(a & ~1ull);
    const bool clip = (a & 1ull) != 0ull;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const double l = lb[base + j];
        double v = l;
        if (p) v = clip ? fmin(fmax(p[j], l), ub[base + j]) : p[j];
        x[base + j] = v;
    }
}

// Iterates into the store: dst[i] (or 0 to skip) receives node i's x.
__global__ void __launch_bounds__(128) scatter_iterates_kernel(int n, const unsigned long long* dst,
                                                               const double* x) {
    const int node = blockIdx.x;
    double* p = reinterpret_cast<double*>(dst[node]);
    if (!p) return;
    const size_t base = static_cast<size_t>(node) * n;
    for (int j = threadIdx.x; j < n; j += blockDim.x) p[j] = x[base + j];
}

int launch_gather_starts(int n, int N, const unsigned long long* src, const double* lb, const double* ub, double* x,
                         void* stream) {
    if (N <= 0 || n <= 0) return 0;
    hipLaunchKernelGGL(gather_starts_kernel, dim3(N), dim3(128), 0, reinterpret_cast<hipStream_t>(stream), n, src,
                       lb, ub, x);
    return static_cast<int>(hipGetLastError());
}

int launch_scatter_iterates(int n, int N, const unsigned long long* dst, const double* x, void* stream) {
    if (N <= 0 || n <= 0) return 0;
    hipLaunchKernelGGL(scatter_iterates_kernel, dim3(N), dim3(128), 0, reinterpret_cast<hipStream_t>(stream), n,
                       dst, x);
    return static_cast<int>(hipGetLastError());
}

// LDS of relax_fast_rows_kernel: two staged vectors, the changed-column
// list, the exchange buffers, counts and the node id.
static size_t rows_lds_bytes(int n_pad, int nt) {
    const size_t doubles = (size_t)n_pad * 7 / 2 + (size_t)2 * nt * kWave * (2 * kRowSums + 1);
    const size_t ints = (size_t)2 * nt * kWave * 2 + nt + 1 + 2 * nt;  // + the simplex's per-wave counts
    return doubles * 8 + ints * 4;
}

// Nodes relax_fast_rows_kernel<NT> keeps in flight on this device with
// `lds` bytes of LDS per workgroup (occupancy x CUs), cached per device and
// LDS size.
template <int NT, bool SX>
static int rows_capacity(size_t lds) {
    static std::mutex mu;
    static std::map<std::pair<int, size_t>, int> cached;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    std::lock_guard<std::mutex> lk(mu);
    const auto it = cached.find({dev, lds});
    if (it != cached.end()) return it->second;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, relax_fast_rows_kernel<NT, SX>, NT * kWave, lds) !=
            hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    return cached[{dev, lds}] = per_cu * cus;
}

int launch_relax_fast(const FastParams& p, int waves_per_block, int blocks, void* stream) {
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, st, p.counter);
    const bool sx = p.simplex != 0;
    // (over a simplex the rows layout pays a selection barrier per vertex,
    // which only more than two waves per node repay: measured on 1024 nodes,
    // 3.4 against 2.7 ms at n = 80, 5.4 against 8.0 ms at n = 200)
    if (waves_per_block == 0 && p.n > (sx ? 128 : 64)) {
        // One node per workgroup when the whole batch is in flight at once;
        // otherwise the one-wave-per-node layout keeps more nodes resident.
        const int nt = p.n <= 128 ? 2 : (p.n <= 256 ? 4 : 8);
        const size_t lds = rows_lds_bytes(p.n_pad, nt);
        int cap = 0;
        if (sx) cap = nt == 2 ? rows_capacity<2, true>(lds) : nt == 4 ? rows_capacity<4, true>(lds) : rows_capacity<8, true>(lds);
        else cap = nt == 2 ? rows_capacity<2, false>(lds) : nt == 4 ? rows_capacity<4, false>(lds) : rows_capacity<8, false>(lds);
        if (p.N <= cap) {
            const dim3 grid(p.N < blocks || blocks <= 0 ? p.N : blocks), block(nt * kWave);
            if (sx) {
                if (nt == 2) hipLaunchKernelGGL((relax_fast_rows_kernel<2, true>), grid, block, lds, st, p);
                else if (nt == 4) hipLaunchKernelGGL((relax_fast_rows_kernel<4, true>), grid, block, lds, st, p);
                else hipLaunchKernelGGL((relax_fast_rows_kernel<8, true>), grid, block, lds, st, p);
            } else {
                if (nt == 2) hipLaunchKernelGGL((relax_fast_rows_kernel<2, false>), grid, block, lds, st, p);
                else if (nt == 4) hipLaunchKernelGGL((relax_fast_rows_kernel<4, false>), grid, block, lds, st, p);
                else hipLaunchKernelGGL((relax_fast_rows_kernel<8, false>), grid, block, lds, st, p);
            }
            return (int)hipGetLastError();
        }
    }
    if (waves_per_block == 0) waves_per_block = 4;
    if (blocks <= 0) {
        const int needed = (p.N + waves_per_block - 1) / waves_per_block;
        blocks = needed < kMaxFastBlocks ? needed : kMaxFastBlocks;
    }
    // per wave: 2 n_pad doubles + n_pad ints (n_pad even keeps 8-byte alignment)
    const size_t wave_lds = (size_t)waves_per_block * p.n_pad * 5 / 2 * sizeof(double);
    const size_t q_lds = (size_t)p.n * p.n * sizeof(double);
    const bool ql = q_lds + wave_lds <= kFastLdsBudget;
    const size_t lds = wave_lds + (ql ? q_lds : 0);
    const dim3 grid(blocks), block(waves_per_block * kWave);
#define GB_FAST_SX(NT, QL, SXK)                                                           \
    do {                                                                                  \
        auto k = relax_fast_kernel<NT, QL, SXK>;                                          \
        if (lds > 64 * 1024) {                                                            \
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),          \
                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                    \
            if (e != hipSuccess) return (int)e;                                           \
        }                                                                                 \
        hipLaunchKernelGGL(k, grid, block, lds, st, p);                                   \
    } while (0)
#define GB_FAST(NT, QL)                                                                   \
    do {                                                                                  \
        if (sx) GB_FAST_SX(NT, QL, true);                                                 \
        else GB_FAST_SX(NT, QL, false);                                                   \
    } while (0)
    if (p.n <= 64) { if (ql) GB_FAST(1, true); else GB_FAST(1, false); }
    else if (p.n <= 128) { if (ql) GB_FAST(2, true); else GB_FAST(2, false); }
    else if (p.n <= 256) GB_FAST(4, false);
    else if (p.n <= 512) GB_FAST(8, false);
    else return (int)hipErrorInvalidValue;
#undef GB_FAST
#undef GB_FAST_SX
    return (int)hipGetLastError();
}

}  // namespace gb
