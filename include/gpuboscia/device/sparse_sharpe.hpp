// The reference's sparse Sharpe-ratio model as device plug-ins
// (examples/sparse_sharpe.cu:225-546):
//
//   minimise  0.5 v'Hv - mu'v   over (v, y),  0 <= v_i <= U_i y_i,
//             sum(y) <= m,  y_i in {0, 1},
//
// with the variables laid out as [v (MaxAssets) | y (MaxAssets)], so
// n = 2 MaxAssets and y_i is variable MaxAssets + i; entries of v and y past
// `assets` stay at zero.  H is assets x assets row-major (the covariance plus
// a ridge), mu has `assets` entries, U_i = device_upper_bounds[i]; all three
// are caller-owned device arrays.
//
// Both types follow the reference's plug-in contracts and are compiled into
// the relaxation kernel in the caller's translation unit
// (gpuboscia::solve / DeviceResidentHybridSolver with hipcc for gfx950).
// The LMO is written for wave64 workgroups: its selections are exact integer
// ranks over LDS keys, block sums go through feasibility_block_sum.
#pragma once
#include <hip/hip_runtime.h>

#include "relax_fp64.hpp"

namespace gpuboscia {

// f(v, y) = 0.5 v'Hv - mu'v, grad = (Hv - mu, 0)  (sparse_sharpe.cu:225-257).
template <int MaxAssets, int Threads = 256>
struct SparseSharpeObjective {
    static constexpr bool supports_quadratic_exact = true;  // quadratic_curvature_local below
    static constexpr int n_vars = 2 * MaxAssets;
    static constexpr int block_threads = Threads;
    static constexpr int shared_bytes = 16;
    const double* d_H;
    const double* d_mu;
    int assets;

    __device__ void prepare_block(int, unsigned char*) const {}

    __device__ void grad(int tid, const double* x, double* g, unsigned char*) const {
        for (int i = tid; i < n_vars; i += blockDim.x) {
            if (i >= assets) {
                g[i] = 0.0;
                continue;
            }
            const double* row = d_H + (size_t)i * assets;
            double acc = -d_mu[i];
            for (int j = 0; j < assets; ++j) acc = fma(row[j], x[j], acc);
            g[i] = acc;
        }
        __syncthreads();
    }

    __device__ double value_local(int tid, const double* x, unsigned char*) const {
        double val = 0.0;
        for (int i = tid; i < assets; i += blockDim.x) {
            const double* row = d_H + (size_t)i * assets;
            double hx = 0.0;
            for (int j = 0; j < assets; ++j) hx = fma(row[j], x[j], hx);
            val = fma(x[i], fma(0.5, hx, -d_mu[i]), val);
        }
        return val;
    }

    __device__ double quadratic_curvature_local(int tid, const double* d, unsigned char*) const {
        double c = 0.0;
        for (int i = tid; i < assets; i += blockDim.x) {
            const double* row = d_H + (size_t)i * assets;
            double hd = 0.0;
            for (int j = 0; j < assets; ++j) hd = fma(row[j], d[j], hd);
            c = fma(d[i], hd, c);
        }
        return c;
    }
};

// D = {(v, y): 0 <= v_i <= U_i y_i, 1'y <= m} within the node's bounds on y
// (sparse_sharpe.cu:259-546).  Taking asset i costs g_y[i] + min(0, U_i
// g_v[i]); the vertex takes every mandatory asset (lb_y > 1/2) and then the
// cheapest negative-cost optional ones, up to m in all, ranked by (cost,
// index).  With v_i = U_i where g_v[i] < 0 and 0 elsewhere on the taken
// assets, that is an exact minimiser of <g, .> over D.
template <int MaxAssets>
struct LinkedCardinalityLMO {
    static constexpr bool supports_dicg = true;
    static constexpr int variables = 2 * MaxAssets;
    int assets;
    int cardinality;
    const double* device_upper_bounds;
    const double* device_H;   // the objective's H and mu, for the candidate heuristic
    const double* device_mu;

    __device__ static bool before(double ca, int a, double cb, int b) { return ca < cb || (ca == cb && a < b); }

    // The mandatory assets must fit the cardinality (bounds and integer
    // intervals are checked by the kernel before this).
    __device__ bool check_feasible_device(int tid, int, const double* lb, const double*, const uint8_t*, double*,
                                          double*) const {
        double mandatory = 0.0;
        for (int i = tid; i < assets; i += blockDim.x) mandatory += lb[MaxAssets + i] > 0.5 ? 1.0 : 0.0;
        return feasibility_block_sum(mandatory) <= cardinality;
    }

    __device__ void compute_v(int tid, int n, const double* g, const double* lb, const double* ub, double* v) const {
        __shared__ double cost[MaxAssets];
        __shared__ unsigned char state[MaxAssets];  // 0 out, 1 optional candidate, 2 mandatory
        double mand = 0.0;
        for (int i = tid; i < assets; i += blockDim.x) {
            const double c = g[MaxAssets + i] + fmin(0.0, device_upper_bounds[i] * g[i]);
            const bool m = lb[MaxAssets + i] > 0.5;
            cost[i] = c;
            state[i] = m ? 2 : ((c < 0.0 && ub[MaxAssets + i] > 0.5) ? 1 : 0);
            mand += m ? 1.0 : 0.0;
        }
        for (int j = tid; j < n; j += blockDim.x) v[j] = 0.0;
        const double mandatory = feasibility_block_sum(mand);  // also orders the writes above
        // check_feasible_device rules this out; the zero vector marks it
        if (mandatory > cardinality) return;
        const int slots = cardinality - (int)mandatory;
        for (int i = tid; i < assets; i += blockDim.x) {
            bool take = state[i] == 2;
            if (state[i] == 1) {
                int rank = 0;
                for (int k = 0; k < assets; ++k) rank += (state[k] == 1 && before(cost[k], k, cost[i], i)) ? 1 : 0;
                take = rank < slots;
            }
            if (take) {
                v[MaxAssets + i] = 1.0;
                v[i] = g[i] < 0.0 ? device_upper_bounds[i] : 0.0;
            }
        }
        __syncthreads();
    }

    // Minimal face of x (sparse_sharpe.cu:422-505): y coordinates within the
    // tolerance of a bound stay there, zero v stay zero, active links
    // v_i = U_i y_i stay active; a full cardinality makes every free asset a
    // candidate.  `g` is the direction to minimise (the kernel passes -grad).
    __device__ void compute_inface_v(int tid, int n, const double* g, const double* x, const double* lb,
                                     const double* ub, double* a, double atol = 1e-6, double rtol = 1e-4) const {
        __shared__ double cost[MaxAssets];
        __shared__ unsigned char flags[MaxAssets];  // bit 0 y fixed, 1 v zero, 2 link active, 3 candidate
        double fixed_sel = 0.0, ysum = 0.0;
        for (int j = tid; j < n; j += blockDim.x) a[j] = 0.0;
        __syncthreads();
        for (int i = tid; i < assets; i += blockDim.x) {
            const double U = device_upper_bounds[i];
            const double vi = x[i], yi = x[MaxAssets + i];
            const double ylo = lb[MaxAssets + i], yhi = ub[MaxAssets + i];
            const bool at_lo = fabs(yi - ylo) <= atol + rtol * fmax(fabs(yi), fabs(ylo));
            const bool at_hi = fabs(yi - yhi) <= atol + rtol * fmax(fabs(yi), fabs(yhi));
            const bool yfix = at_lo || at_hi;
            const bool vzero = fabs(vi) <= atol + rtol * fabs(vi);
            const bool link = fabs(vi - U * yi) <= atol + rtol * fmax(fabs(vi), fabs(U * yi));
            unsigned char f = (yfix ? 1 : 0) | (vzero ? 2 : 0) | (link ? 4 : 0);
            if (yfix) {
                const double fy = at_lo ? ylo : yhi;
                a[MaxAssets + i] = fy;
                if (fy > 0.5) fixed_sel += 1.0;
                if (!vzero) a[i] = link ? U * fy : (g[i] < 0.0 ? U * fy : 0.0);
            }
            flags[i] = f;
            ysum += yi;
        }
        const double fixed_selected = feasibility_block_sum(fixed_sel);
        const double sum_y = feasibility_block_sum(ysum);
        const bool full = fabs(sum_y - cardinality) <= atol + rtol * fmax(fabs(sum_y), (double)cardinality);
        for (int i = tid; i < assets; i += blockDim.x) {
            const unsigned char f = flags[i];
            double c = __builtin_inf();
            bool cand = false;
            if (!(f & 1)) {
                const double U = device_upper_bounds[i];
                c = g[MaxAssets + i];
                if (f & 4) c += U * g[i];
                else if (!(f & 2)) c += fmin(0.0, U * g[i]);
                cand = full || c < 0.0;
            }
            cost[i] = c;
            flags[i] = f | (cand ? 8 : 0);
        }
        __syncthreads();
        const int slots = cardinality - (int)fixed_selected;
        for (int i = tid; i < assets; i += blockDim.x) {
            const unsigned char f = flags[i];
            if (!(f & 8)) continue;
            int rank = 0;
            for (int k = 0; k < assets; ++k) rank += ((flags[k] & 8) && before(cost[k], k, cost[i], i)) ? 1 : 0;
            if (rank >= slots) continue;
            const double U = device_upper_bounds[i];
            a[MaxAssets + i] = 1.0;
            if (f & 4) a[i] = U;
            else if (!(f & 2)) a[i] = g[i] < 0.0 ? U : 0.0;
        }
        __syncthreads();
    }

    // Largest gamma with x - gamma d in D, coordinate constraints.
    __device__ double dicg_maximum_step_local(int tid, int, const double* d, const double* x, const double* lb,
                                              const double* ub, double tol = 1e-9) const {
        double r = 1.0;
        for (int i = tid; i < assets; i += blockDim.x) {
            const double U = device_upper_bounds[i];
            const double dv = d[i], dy = d[MaxAssets + i];
            if (dv > tol) r = fmin(r, x[i] / dv);
            const double ld = dv - U * dy, lx = x[i] - U * x[MaxAssets + i];
            if (ld < -tol) r = fmin(r, lx / ld);
            if (dy > tol) r = fmin(r, (x[MaxAssets + i] - lb[MaxAssets + i]) / dy);
            else if (dy < -tol) r = fmin(r, (ub[MaxAssets + i] - x[MaxAssets + i]) / -dy);
        }
        return fmax(0.0, r);
    }

    // ... and the cardinality row.
    __device__ double dicg_sum_maximum_step(int tid, int, const double* d, const double* x) const {
        double sx = 0.0, sd = 0.0;
        for (int i = tid; i < assets; i += blockDim.x) {
            sx += x[MaxAssets + i];
            sd += d[MaxAssets + i];
        }
        const double sum_x = feasibility_block_sum(sx);
        const double sum_d = feasibility_block_sum(sd);
        return sum_d < -1e-9 ? fmax(0.0, fmin(1.0, (sum_x - cardinality) / sum_d)) : 1.0;
    }

    // Rounded candidate (sparse_sharpe.cu:282-335): the mandatory assets,
    // then the largest positions of x up to m; v keeps x on them and is
    // scaled to the objective's minimum along that direction, capped by U.
    // Sequential on thread 0 (m and the asset count are small).
    __device__ void integer_candidate_device(int tid, int n, const double* x, const double* lb, const double* ub,
                                             const uint8_t*, double* c) const {
        for (int j = tid; j < n; j += blockDim.x) c[j] = 0.0;
        __syncthreads();
        if (tid == 0) {
            int selected = 0;
            for (int i = 0; i < assets; ++i)
                if (lb[MaxAssets + i] > 0.5) {
                    c[MaxAssets + i] = 1.0;
                    ++selected;
                }
            while (selected < cardinality) {
                int best = -1;
                for (int i = 0; i < assets; ++i)
                    if (c[MaxAssets + i] == 0.0 && ub[MaxAssets + i] > 0.5 && x[i] > 1e-12
                            && (best < 0 || x[i] > x[best]))
                        best = i;
                if (best < 0) break;
                c[MaxAssets + best] = 1.0;
                ++selected;
            }
            bool direction = false;
            for (int i = 0; i < assets; ++i)
                if (c[MaxAssets + i] == 1.0) {
                    c[i] = fmax(0.0, x[i]);
                    direction = direction || c[i] > 0.0;
                }
            if (!direction) {
                int best = -1;
                for (int i = 0; i < assets; ++i)
                    if ((c[MaxAssets + i] == 1.0 || (selected < cardinality && ub[MaxAssets + i] > 0.5))
                            && (best < 0 || device_mu[i] > device_mu[best]))
                        best = i;
                if (best >= 0 && device_mu[best] > 0.0) {
                    c[MaxAssets + best] = 1.0;
                    c[best] = 1.0;
                }
            }
            double lin = 0.0, quad = 0.0, cap = __builtin_inf();
            for (int i = 0; i < assets; ++i) {
                lin += device_mu[i] * c[i];
                if (c[i] > 0.0) cap = fmin(cap, device_upper_bounds[i] / c[i]);
                for (int j = 0; j < assets; ++j) quad += c[i] * device_H[(size_t)i * assets + j] * c[j];
            }
            const double scale = (lin > 0.0 && quad > 0.0) ? fmin(lin / quad, cap) : 0.0;
            for (int i = 0; i < assets; ++i) c[i] *= scale;
        }
        __syncthreads();
    }

    __device__ bool candidate_feasible_device(int tid, int n, const double* x, const double*, const double*,
                                              double tol) const {
        if (n != variables) return false;
        double sel = 0.0;
        bool ok = true;
        for (int i = tid; i < assets; i += blockDim.x) {
            const double y = x[MaxAssets + i];
            if (x[i] < -tol || x[i] > device_upper_bounds[i] * y + tol) ok = false;
            sel += y;
        }
        const double valid = feasibility_block_sum(ok ? 0.0 : 1.0);
        const double total = feasibility_block_sum(sel);
        return valid == 0.0 && total <= cardinality + tol;
    }
};

}  // namespace gpuboscia
