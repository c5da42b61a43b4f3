// The FP64 node-relaxation kernel of gfx950 as a template over its
// objective, LMO and candidate heuristic: the library instantiates it with
// its built-in kinds (Obj<K>, Lmo<K>, BuiltinHeuristic), and a caller
// instantiates it with its own device types through PluginObjective<U>,
// PluginLmo<U> and PluginHeuristic<U> (compile with hipcc for gfx950, then
// register launch_relax_plugin through gb_plugin_desc, gpuboscia_amd.h).
//
// One workgroup (whole wave64s) per B&B node, Frank-Wolfe or DICG with the
// reference's five step rules; follows drs_fw_relax_node
// (device_resident_solver.cuh:337-629) with the CUDA warp-32 reductions
// replaced by the canonical wave64 order of gb_device.h.  Block-uniform
// scalars (gap, step, line-search state) are held redundantly in every
// lane's registers instead of shared-memory scalars, so each reduction needs
// two barriers and no broadcast.
//
// Plug-in code runs inside this kernel exactly as the reference runs it
// inside drs_fw_relax_node: every thread of the workgroup calls each method
// with tid = threadIdx.x, loops over coordinates j = tid, tid + blockDim.x,
// ..., and must be wave64-clean (blockDim.x is a multiple of 64; use
// feasibility_block_sum / lmo_block_sum at the end of this file, not
// 32-lane warp idioms).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

#include "gb_device.h"
#include "relax_params.h"

namespace gb {
namespace {

struct Lds {
    double* red;
    int* ired;
    double* vec;
    double* sg;
    int* sidx;
    double* scap;
    double* spre;
    uint8_t* elig;
    uint8_t* fixed;
    int* fcol;
    int* frow;
    int* rows;
    int* cols;
    double* hu;
    int* hp;
    int* way;
    int* misc;
    double* hcost;  // residual Hungarian costs (null: read from global)
};

__device__ __forceinline__ int next_pow2(int v) { int r = 1; while (r < v) r <<= 1; return r; }

// ---------------------------------------------------------------------------
// Objectives
// ---------------------------------------------------------------------------
template <int OBJ>
struct Obj;

// GlobalCachedDenseQuadraticObjective (dense_quadratic_objectives.cuh:18-112).
// The multiplied vector is staged in LDS; each lane owns output rows j and
// walks k with coalesced loads of Qt[k*n + j].
template <>
struct Obj<GB_OBJ_DENSE_QUADRATIC> {
    __device__ static void stage(const RelaxParams& P, const Lds& s, const double* src) {
        for (int k = threadIdx.x; k < P.n; k += blockDim.x) s.vec[k] = src[k];
        __syncthreads();
    }
    __device__ static double row(const RelaxParams& P, const Lds& s, int j) {
        const int n = P.n;
        const double* q = P.Qt + j;
        double acc = 0.0;
        int k = 0;
        for (; k + 4 <= n; k += 4) {
            const double q0 = q[(size_t)(k + 0) * n], q1 = q[(size_t)(k + 1) * n];
            const double q2 = q[(size_t)(k + 2) * n], q3 = q[(size_t)(k + 3) * n];
            acc = fma(q0, s.vec[k + 0], acc);
            acc = fma(q1, s.vec[k + 1], acc);
            acc = fma(q2, s.vec[k + 2], acc);
            acc = fma(q3, s.vec[k + 3], acc);
        }
        for (; k < n; ++k) acc = fma(q[(size_t)k * n], s.vec[k], acc);
        return acc;
    }
    __device__ static void grad(const RelaxParams& P, const Lds& s, double* /*R*/,
                                const double* x, double* g) {
        stage(P, s, x);
        for (int j = threadIdx.x; j < P.n; j += blockDim.x) {
            const double acc = row(P, s, j);
            const double lin = P.lin ? P.ls * P.lin[j] : 0.0;
            g[j] = P.gs * (acc + lin);
        }
        __syncthreads();
    }
    __device__ static double value(const RelaxParams& P, const Lds& s, double* /*R*/,
                                   const double* x) {
        stage(P, s, x);
        const double hs = 0.5 * P.gs, ls = P.gs * P.ls;
        double quad = 0.0, lin = 0.0;
        for (int j = threadIdx.x; j < P.n; j += blockDim.x) {
            const double qx = row(P, s, j);
            quad = fma(s.vec[j], qx, quad);
            if (P.lin) lin = fma(P.lin[j], s.vec[j], lin);
        }
        double val = fma(hs, quad, ls * lin);
        if (threadIdx.x == 0) val = val + P.kappa;
        return block_sum(val, s.red);
    }
    __device__ static double curvature(const RelaxParams& P, const Lds& s, double* /*R*/,
                                       const double* d) {
        stage(P, s, d);
        double c = 0.0;
        for (int j = threadIdx.x; j < P.n; j += blockDim.x) {
            const double p = row(P, s, j);
            c = fma(P.gs * s.vec[j], p, c);
        }
        return block_sum(c, s.red);
    }
};

// QAP f(P) = ||A P - P B||_F^2 on an order-m Birkhoff polytope.  R = AP - PB
// is written to the node's residual vector; grad = 2 (A'R - R B').  Each
// thread computes blocks of kBI x kBJ outputs, loading kBI + kBJ values of
// each operand per k for 2 kBI kBJ FMAs; every output still accumulates its
// products in k order with fma.  Block edges clamp the loads and skip the
// stores.
template <>
struct Obj<GB_OBJ_QAP> {
    static constexpr int kBI = 2, kBJ = 4;

    // out = 2 (L M' - M N') style products of one pass over the blocks:
    // RES: out = A X - X B (ap - pb);  GRAD: out = 2 (A' R - R B').
    // Out of line with operands by value, so the blocked loop has the
    // registers to itself.
    template <bool GRAD>
    __device__ __noinline__ static void blocks(int m, const double* A, const double* B, const double* X,
                                               double* out) {
        const int nbi = (m + kBI - 1) / kBI, nbj = (m + kBJ - 1) / kBJ;
        for (int blk = threadIdx.x; blk < nbi * nbj; blk += blockDim.x) {
            const int i0 = (blk / nbj) * kBI, j0 = (blk % nbj) * kBJ;
            int ii[kBI], jj[kBJ];
#pragma unroll
            for (int a = 0; a < kBI; ++a) ii[a] = min(i0 + a, m - 1);
#pragma unroll
            for (int b = 0; b < kBJ; ++b) jj[b] = min(j0 + b, m - 1);
            double s1[kBI][kBJ], s2[kBI][kBJ];
#pragma unroll
            for (int a = 0; a < kBI; ++a)
#pragma unroll
                for (int b = 0; b < kBJ; ++b) { s1[a][b] = 0.0; s2[a][b] = 0.0; }
#pragma unroll 2
            for (int k = 0; k < m; ++k) {
                double l1[kBI], l2[kBI], r1[kBJ], r2[kBJ];
#pragma unroll
                for (int a = 0; a < kBI; ++a) {
                    if (GRAD) { l1[a] = A[k * m + ii[a]]; l2[a] = X[ii[a] * m + k]; }   // A[k][i], R[i][k]
                    else { l1[a] = A[ii[a] * m + k]; l2[a] = X[ii[a] * m + k]; }        // A[i][k], X[i][k]
                }
#pragma unroll
                for (int b = 0; b < kBJ; ++b) {
                    if (GRAD) { r1[b] = X[k * m + jj[b]]; r2[b] = B[jj[b] * m + k]; }   // R[k][j], B[j][k]
                    else { r1[b] = X[k * m + jj[b]]; r2[b] = B[k * m + jj[b]]; }        // X[k][j], B[k][j]
                }
#pragma unroll
                for (int a = 0; a < kBI; ++a)
#pragma unroll
                    for (int b = 0; b < kBJ; ++b) {
                        s1[a][b] = fma(l1[a], r1[b], s1[a][b]);
                        s2[a][b] = fma(l2[a], r2[b], s2[a][b]);
                    }
            }
#pragma unroll
            for (int a = 0; a < kBI; ++a)
#pragma unroll
                for (int b = 0; b < kBJ; ++b)
                    if (i0 + a < m && j0 + b < m)
                        out[(i0 + a) * m + j0 + b] = GRAD ? 2.0 * (s1[a][b] - s2[a][b]) : s1[a][b] - s2[a][b];
        }
    }
    __device__ static void residual(const RelaxParams& P, double* R, const double* X) {
        blocks<false>(P.order, P.A, P.B, X, R);
        __syncthreads();
    }
    __device__ static void grad(const RelaxParams& P, const Lds&, double* R, const double* x,
                                double* g) {
        residual(P, R, x);
        blocks<true>(P.order, P.A, P.B, R, g);
        __syncthreads();
    }
    __device__ static double value(const RelaxParams& P, const Lds& s, double* R, const double* x) {
        residual(P, R, x);
        double acc = 0.0;
        for (int e = threadIdx.x; e < P.order * P.order; e += blockDim.x) acc = fma(R[e], R[e], acc);
        return block_sum(acc, s.red);
    }
    __device__ static double curvature(const RelaxParams& P, const Lds& s, double* R,
                                       const double* d) {
        residual(P, R, d);
        double acc = 0.0;
        for (int e = threadIdx.x; e < P.order * P.order; e += blockDim.x) acc = fma(R[e], R[e], acc);
        return 2.0 * block_sum(acc, s.red);
    }
};

// ---------------------------------------------------------------------------
// LMOs
// ---------------------------------------------------------------------------
template <int LMO>
struct Lmo;

// BoxLMO (lmo.cuh:212-321).
template <>
struct Lmo<GB_LMO_BOX> {
    __device__ static bool feasible(const RelaxParams&, const Lds&, const double*, const double*,
                                    double*) { return true; }
    __device__ static void compute_v(const RelaxParams& P, const Lds&, const double* g,
                                     const double* lb, const double* ub, double* v) {
        for (int j = threadIdx.x; j < P.n; j += blockDim.x) v[j] = (g[j] > 0.0) ? lb[j] : ub[j];
    }
    __device__ static void inface(const RelaxParams& P, const Lds&, const double* g, const double* x,
                                  const double* lb, const double* ub, double* a) {
        const double atol = 1e-6, rtol = 1e-4;
        for (int j = threadIdx.x; j < P.n; j += blockDim.x) {
            const double ut = atol + rtol * fabs(ub[j]);
            const double lt = atol + rtol * fabs(lb[j]);
            if (fabs(x[j] - ub[j]) <= ut) a[j] = ub[j];
            else if (fabs(x[j] - lb[j]) <= lt) a[j] = lb[j];
            else a[j] = (g[j] > 0.0) ? lb[j] : ub[j];
        }
    }
    __device__ static double max_step(const RelaxParams& P, const Lds& s, const double* d,
                                      const double* x, const double* lb, const double* ub) {
        double r = 1.0;
        for (int j = threadIdx.x; j < P.n; j += blockDim.x) {
            const double dj = d[j];
            if (dj < 0.0) r = fmin(r, (ub[j] - x[j]) / -dj);
            else if (dj > 0.0) r = fmin(r, (x[j] - lb[j]) / dj);
        }
        const double coord = block_min(r, s.red);
        return fmax(0.0, fmin(coord, 1.0));
    }
    __device__ static bool rounded_feasible(const RelaxParams&, const Lds&, const double*) {
        return true;
    }
};

// UnitSimplexLMO / EqualitySimplexLMO (lmo.cuh:335-912).  The greedy
// allocation ranks eligible coordinates by (gradient, index) with an
// all-pairs count in LDS (exact, O(n^2/T) per call), scans the capacities
// in the canonical chunked order on wave 0, then fills every coordinate from
// its prefix mass.
template <>
struct Lmo<GB_LMO_SIMPLEX> {
    __device__ static void greedy(const RelaxParams& P, const Lds& s, const double* g,
                                  const double* ub, double* out, bool use_fixed, double remaining,
                                  bool force_full, bool floor_int) {
        const int n = P.n, S = next_pow2(n);
        const uint8_t* is_int = P.is_int;
        for (int j = threadIdx.x; j < n; j += blockDim.x) {
            s.sg[j] = g[j];
            s.elig[j] = (!use_fixed || !s.fixed[j]) && (force_full || g[j] <= 0.0);
        }
        for (int p = threadIdx.x; p < S; p += blockDim.x) { s.sidx[p] = -1; s.scap[p] = 0.0; }
        __syncthreads();
        for (int j = threadIdx.x; j < n; j += blockDim.x) {
            if (!s.elig[j]) continue;
            const double gj = s.sg[j];
            int r = 0;
            for (int i = 0; i < n; ++i) {
                const double gi = s.sg[i];
                r += (s.elig[i] && (gi < gj || (gi == gj && i < j))) ? 1 : 0;
            }
            s.sidx[r] = j;
            s.scap[r] = is_int[j] ? fmax(ub[j] - out[j], 0.0) : remaining;
        }
        __syncthreads();
        if (threadIdx.x < kWave) {
            const int l = threadIdx.x;
            const int C = (S + kWave - 1) / kWave;
            double t = 0.0;
            for (int i = 0; i < C; ++i) {
                const int p = l * C + i;
                if (p < S) t = t + s.scap[p];
            }
            double inc = t;
#pragma unroll
            for (int off = 1; off < kWave; off <<= 1) {
                const double o = __shfl_up(inc, off, kWave);
                if (l >= off) inc = inc + o;
            }
            double run = __shfl_up(inc, 1, kWave);
            if (l == 0) run = 0.0;
            for (int i = 0; i < C; ++i) {
                const int p = l * C + i;
                if (p >= S) break;
                s.spre[p] = run;
                run = run + s.scap[p];
            }
        }
        __syncthreads();
        const bool eq = P.equality != 0;
        for (int p = threadIdx.x; p < S; p += blockDim.x) {
            const int j = s.sidx[p];
            if (j < 0) continue;
            double add = fmax(remaining - s.spre[p], 0.0);
            const double c = is_int[j] ? fmax(ub[j] - out[j], 0.0) : remaining;
            add = fmin(add, c);
            if (floor_int && is_int[j]) add = floor(add + (eq ? 1e-9 : 0.0));
            out[j] = out[j] + add;
        }
    }
    __device__ static bool feasible(const RelaxParams& P, const Lds& s, const double* lb,
                                    const double* ub, double*) {
        bool valid = true, cont = false;
        double mand = 0.0, capsum = 0.0;
        for (int j = threadIdx.x; j < P.n; j += blockDim.x) {
            const bool integer = P.flags ? P.flags[j] != 0 : P.is_int[j] != 0;
            const double lo = integer ? ceil(fmax(0.0, lb[j])) : fmax(0.0, lb[j]);
            const double hi = integer ? floor(ub[j]) : ub[j];
            valid = valid && lo <= hi;
            mand = mand + lo;
            capsum = capsum + hi;
            cont = cont || !integer;
        }
        const bool iv = block_and(valid, s.ired);
        const double ms = block_sum(mand, s.red);
        bool feas = iv && ms <= P.rhs + 1e-9;
        if (P.equality) {
            const double cs = block_sum(capsum, s.red);
            const bool hc = block_or(cont, s.ired);
            feas = feas && cs >= P.rhs - 1e-9 && (hc || fabs(P.rhs - round(P.rhs)) <= 1e-9);
        }
        return feas;
    }
    __device__ static void compute_v(const RelaxParams& P, const Lds& s, const double* g,
                                     const double* lb, const double* ub, double* v) {
        double mass = 0.0;
        for (int j = threadIdx.x; j < P.n; j += blockDim.x) {
            v[j] = P.is_int[j] ? lb[j] : 0.0;
            mass = mass + v[j];
        }
        const double total = block_sum(mass, s.red);
        const bool eq = P.equality != 0;
        greedy(P, s, g, ub, v, false, P.rhs - total, eq, eq);
    }
    __device__ static void inface(const RelaxParams& P, const Lds& s, const double* g,
                                  const double* x, const double* lb, const double* ub, double* a) {
        const double atol = 1e-6, rtol = 1e-4;
        double mass = 0.0, xmass = 0.0;
        for (int j = threadIdx.x; j < P.n; j += blockDim.x) {
            uint8_t fx = 0;
            const bool ii = P.is_int[j] != 0;
            double aj = ii ? lb[j] : 0.0;
            const double lower = ii ? lb[j] : 0.0;
            const double ltol = atol + rtol * fmax(fabs(x[j]), fabs(lower));
            if (fabs(x[j] - lower) <= ltol) { fx = 1; aj = lower; }
            else if (ii) {
                const double utol = atol + rtol * fmax(fabs(x[j]), fabs(ub[j]));
                if (fabs(x[j] - ub[j]) <= utol) { fx = 1; aj = ub[j]; }
            }
            s.fixed[j] = fx;
            a[j] = aj;
            mass = mass + aj;
            xmass = xmass + x[j];
        }
        const double m = block_sum(mass, s.red);
        const double xm = block_sum(xmass, s.red);
        const bool force = P.equality
            || fabs(xm - P.rhs) <= atol + rtol * fmax(fabs(xm), fabs(P.rhs));
        greedy(P, s, g, ub, a, true, P.rhs - m, force, true);
    }
    __device__ static double max_step(const RelaxParams& P, const Lds& s, const double* d,
                                      const double* x, const double* lb, const double* ub) {
        const double tol = 1e-6;
        double r = 1.0;
        for (int j = threadIdx.x; j < P.n; j += blockDim.x) {
            const double dj = d[j];
            const bool ii = P.is_int[j] != 0;
            if (dj > tol) r = fmin(r, (x[j] - (ii ? lb[j] : 0.0)) / dj);
            else if (dj < -tol) r = fmin(r, ((ii ? ub[j] : P.rhs) - x[j]) / -dj);
        }
        const double coord = block_min(r, s.red);
        double sx = 0.0, sd = 0.0;
        for (int j = threadIdx.x; j < P.n; j += blockDim.x) sx = sx + x[j];
        const double tx = block_sum(sx, s.red);
        for (int j = threadIdx.x; j < P.n; j += blockDim.x) sd = sd + d[j];
        const double td = block_sum(sd, s.red);
        double agg;
        if (P.equality && fabs(td) > 1e-9) agg = 0.0;
        else agg = td < 0.0 ? fmin(1.0, (tx - P.rhs) / td) : 1.0;
        return fmax(0.0, fmin(coord, agg));
    }
    __device__ static bool rounded_feasible(const RelaxParams& P, const Lds& s, const double* x) {
        bool valid = true;
        double sum = 0.0;
        for (int j = threadIdx.x; j < P.n; j += blockDim.x) {
            valid = valid && x[j] >= -P.ctol;
            sum = sum + x[j];
        }
        const bool nn = block_and(valid, s.ired);
        const double t = block_sum(sum, s.red);
        return nn && (P.equality ? fabs(t - P.rhs) <= P.ctol : t <= P.rhs + P.ctol);
    }
};

// Cardinality-portfolio LMO over v = (x, z), n = 2m (no reference
// counterpart): sum x = rhs, 0 <= x_i <= cap z_i, lb_z <= z <= ub_z,
// sum z <= card.  With z at its least value, asset i offers x in segment
// q = 2i (up to cap lz_i at cost gx_i) and q = 2i + 1 (up to cap uz_i at
// cost gx_i + max(gz_i, 0) / cap, drawing 1/cap per unit on the pool
// cap (card - sum lz)); segments, pool and total form a laminar family, on
// which the cheapest-first fill is an optimal vertex.  Segments are ranked
// by (cost, q) with an all-pairs count; thread 0 fills in rank order.
// Frank-Wolfe only: the solver rejects DICG for this LMO.
template <>
struct Lmo<GB_LMO_PORTFOLIO> {
    __device__ static bool feasible(const RelaxParams& P, const Lds& s, const double* lb, const double* ub,
                                    double*) {
        const int m = P.n / 2;
        if (threadIdx.x == 0) {
            double zsum = 0.0, a = 0.0, b = 0.0;
            for (int i = 0; i < m; ++i) {
                zsum = zsum + lb[m + i];
                a = a + P.pcap * lb[m + i];
                b = b + P.pcap * (ub[m + i] - lb[m + i]);
            }
            const double capacity = a + fmin(b, P.pcap * (P.pcard - zsum));
            s.ired[0] = (zsum <= P.pcard + 1e-9 && capacity >= P.rhs - 1e-9) ? 1 : 0;
        }
        __syncthreads();
        const bool ok = s.ired[0] != 0;
        __syncthreads();
        return ok;
    }
    __device__ static void compute_v(const RelaxParams& P, const Lds& s, const double* g, const double* lb,
                                     const double* ub, double* v) {
        const int m = P.n / 2, S2 = 2 * m;
        const double cap = P.pcap;
        for (int q = threadIdx.x; q < S2; q += blockDim.x) {
            const int i = q >> 1;
            s.sg[q] = (q & 1) ? g[i] + fmax(g[m + i], 0.0) / cap : g[i];
        }
        __syncthreads();
        for (int q = threadIdx.x; q < S2; q += blockDim.x) {
            const double kq = s.sg[q];
            int r = 0;
            for (int t = 0; t < S2; ++t) {
                const double kt = s.sg[t];
                r += (kt < kq || (kt == kq && t < q)) ? 1 : 0;
            }
            s.sidx[r] = q;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double zsum = 0.0;
            for (int i = 0; i < m; ++i) {
                v[i] = 0.0;
                v[m + i] = lb[m + i];
                zsum = zsum + lb[m + i];
            }
            double pool = cap * (P.pcard - zsum);
            double rem = P.rhs;
            for (int r = 0; r < S2; ++r) {
                if (!(rem > 0.0)) break;
                const int q = s.sidx[r];
                const int i = q >> 1;
                double take;
                if ((q & 1) == 0) {
                    take = fmin(rem, cap * lb[m + i]);
                } else {
                    take = fmin(fmin(rem, cap * (ub[m + i] - lb[m + i])), pool);
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
        __syncthreads();
    }
    // No in-face oracle (DICG is rejected for this LMO); present for the
    // common node loop.
    __device__ static void inface(const RelaxParams& P, const Lds& s, const double* g, const double*,
                                  const double* lb, const double* ub, double* a) {
        compute_v(P, s, g, lb, ub, a);
    }
    __device__ static double max_step(const RelaxParams&, const Lds&, const double*, const double*,
                                      const double*, const double*) {
        return 0.0;
    }
    __device__ static bool rounded_feasible(const RelaxParams& P, const Lds& s, const double* x) {
        const int m = P.n / 2;
        if (threadIdx.x == 0) {
            bool ok = true;
            double xs = 0.0, zs = 0.0;
            for (int i = 0; i < m && ok; ++i) {
                if (x[i] < -P.ctol || x[i] > P.pcap * x[m + i] + P.ctol) ok = false;
                xs = xs + x[i];
                zs = zs + x[m + i];
            }
            s.ired[0] = (ok && fabs(xs - P.rhs) <= P.ctol && zs <= P.pcard + P.ctol) ? 1 : 0;
        }
        __syncthreads();
        const bool ok = s.ired[0] != 0;
        __syncthreads();
        return ok;
    }
};

// (value, index) minimum as in `better`, without branches.
__device__ __forceinline__ MinPair better_sel(MinPair a, MinPair b) {
    const bool take = (b.i >= 0) & ((a.i < 0) | (b.v < a.v) | ((b.v == a.v) & (b.i < a.i)));
    return take ? b : a;
}

// Wave-wide reductions by DPP within rows of 16 lanes (quad permutes,
// half-row and row mirrors) and lane reads across the 4 rows; results are
// wave-uniform.
template <int CTRL>
__device__ __forceinline__ int dpp_int(int v) {
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ double dpp_double(double v) {
    return __hiloint2double(dpp_int<CTRL>(__double2hiint(v)), dpp_int<CTRL>(__double2loint(v)));
}
__device__ __forceinline__ double lane_double(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_min_double(double v) {
    v = fmin(v, dpp_double<0xB1>(v));   // lane ^ 1
    v = fmin(v, dpp_double<0x4E>(v));   // lane ^ 2
    v = fmin(v, dpp_double<0x141>(v));  // mirror within 8
    v = fmin(v, dpp_double<0x140>(v));  // mirror within 16
    return fmin(fmin(lane_double(v, 0), lane_double(v, 16)), fmin(lane_double(v, 32), lane_double(v, 48)));
}
__device__ __forceinline__ int wave_min_int(int v) {
    v = min(v, dpp_int<0xB1>(v));
    v = min(v, dpp_int<0x4E>(v));
    v = min(v, dpp_int<0x141>(v));
    v = min(v, dpp_int<0x140>(v));
    return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

// (value, index) argmin of wave 0 over one candidate per lane (index < 0:
// none; values are never NaN): the smallest value, then the lowest index
// among the lanes holding it (+0 and -0 tie, as `better` compares them),
// and that entry's own value.  Equal to wave_argmin.  Lane l holds indices
// congruent to l + 1 modulo 64.
__device__ __forceinline__ MinPair wave_argmin_2p(MinPair p) {
    const double inf = __builtin_inf();
    const double vm = wave_min_double(p.i >= 0 ? p.v : inf);
    if (!(vm < inf)) return MinPair{kNone, -1};
    const int im = wave_min_int(p.i >= 0 && p.v == vm ? p.i : 0x7fffffff);
    return MinPair{lane_double(p.v, (im - 1) & (kWave - 1)), im};
}

// BirkhoffLMO (lmo.cuh:918-1419).  Forced edges (lb > 1/2, or face entries
// >= 1 - 1e-9) are fixed; the residual assignment is solved by the
// shortest-augmenting-path Hungarian method (hungarian_cpu,
// lmo.cuh:1063-1171) on wave 0: lane l owns the residual columns
// j = l + 1 + 64 t and keeps their minv, v potential, used flag and path
// link in registers; the residual costs are staged in LDS by the whole
// workgroup when they fit (forbidden entries as +inf, which no comparison
// accepts, exactly as skipping them).  The column choice is an exact
// (value, index) argmin, so ties resolve to the lowest index as in
// hungarian_cpu (lmo.cuh:1143-1146).  No block barrier inside the solve.
template <>
struct Lmo<GB_LMO_BIRKHOFF> {
    // Wave 0: forced edges into `out` and the residual rows and columns
    // (s.rows, s.cols, k = s.misc[0]).  Returns a wave-uniform flag; false
    // for an invalid forced set (with the forced writes the reference makes
    // before it fails) or a non-square residual.
    __device__ static bool forced_edges(const RelaxParams& P, const Lds& s, const double* lb,
                                        const double* ub, const double* face, double* out) {
        const int m = P.lorder;
        const int lane = threadIdx.x;
        for (int i = lane; i < m; i += kWave) { s.fcol[i] = -1; s.frow[i] = -1; }
        wave_sync();
        bool bad = false;
        for (int r = lane; r < m; r += kWave) {
            for (int c = 0; c < m; ++c) {
                const int e = r * m + c;
                if (lb[e] > ub[e] + 1e-9) bad = true;
                const bool ff = face && face[e] >= 1.0 - 1e-9;
                if (lb[e] > 0.5 || ff) {
                    if (ub[e] < 0.5 || s.fcol[r] >= 0) bad = true;
                    else {
                        s.fcol[r] = c;
                        if (atomicCAS(&s.frow[c], -1, r) != -1) bad = true;
                    }
                }
            }
        }
        if (__ballot(bad) != 0ull) {
            // Replay the reference's serial scan so the partial forced
            // writes of an invalid set match (lmo.cuh:1269-1299).
            if (lane == 0) {
                for (int i = 0; i < m; ++i) { s.fcol[i] = -1; s.frow[i] = -1; }
                bool ok = true;
                for (int r = 0; r < m && ok; ++r)
                    for (int c = 0; c < m; ++c) {
                        const int e = r * m + c;
                        if (lb[e] > ub[e] + 1e-9) { ok = false; break; }
                        const bool ff = face && face[e] >= 1.0 - 1e-9;
                        if (lb[e] > 0.5 || ff) {
                            if (ub[e] < 0.5 || s.fcol[r] >= 0 || s.frow[c] >= 0) { ok = false; break; }
                            s.fcol[r] = c;
                            s.frow[c] = r;
                            out[e] = 1.0;
                        }
                    }
            }
            wave_sync();
            return false;
        }
        wave_sync();
        for (int r = lane; r < m; r += kWave)
            if (s.fcol[r] >= 0) out[r * m + s.fcol[r]] = 1.0;
        if (lane == 0) {
            int k = 0, kc = 0;
            for (int i = 0; i < m; ++i) {
                if (s.fcol[i] < 0) s.rows[k++] = i;
                if (s.frow[i] < 0) s.cols[kc++] = i;
            }
            s.misc[0] = k;
            s.misc[1] = (k == kc) ? 1 : 0;
        }
        wave_sync();
        return s.misc[1] != 0;
    }

    // Wave 0: assignment of the k residual rows (T = ceil(k / 64) columns
    // per lane; STAGED: costs in LDS at hcost_off).  Lane l keeps, for its
    // columns j, minv, the v potential, the used flag, the path link, the
    // matched row hp[j] and that row's u potential in registers; the row of
    // the next column and its potential are read from the owning lane, so a
    // search step touches LDS only for the cost row.  LDS arrays are
    // addressed from the kernel's dynamic LDS (offsets by value), and the
    // row potentials are written back to LDS once per row.  Returns a
    // wave-uniform validity flag.
    template <int T, bool STAGED>
    __device__ __noinline__ static bool assign(int m, int k, unsigned hu_off, unsigned hp_off, unsigned way_off,
                                               unsigned rows_off, unsigned cols_off, unsigned hcost_off,
                                               const double* cost, const double* ub, const double* face,
                                               double* out) {
        extern __shared__ __align__(16) unsigned char lds_raw[];
        double* hu = reinterpret_cast<double*>(lds_raw + hu_off);
        int* hp = reinterpret_cast<int*>(lds_raw + hp_off);
        int* wy = reinterpret_cast<int*>(lds_raw + way_off);
        const int* rows = reinterpret_cast<const int*>(lds_raw + rows_off);
        const int* cols = reinterpret_cast<const int*>(lds_raw + cols_off);
        const double* hcost = reinterpret_cast<const double*>(lds_raw + hcost_off);
        const int lane = threadIdx.x;
        const double inf = __builtin_inf();
        double minv[T], hv[T], uj[T];
        int way[T], hpj[T];
        bool used[T];
#pragma unroll
        for (int t = 0; t < T; ++t) { hv[t] = 0.0; way[t] = 0; }
        for (int j = lane; j <= k; j += kWave) { hu[j] = 0.0; hp[j] = 0; }
        wave_sync();
        for (int i = 1; i <= k; ++i) {
            if (lane == 0) hp[0] = i;
#pragma unroll
            for (int t = 0; t < T; ++t) {
                const int j = kWave * t + lane + 1;
                minv[t] = kNone;
                used[t] = false;
                hpj[t] = j <= k ? hp[j] : 0;
                uj[t] = hpj[t] > 0 ? hu[hpj[t]] : 0.0;
            }
            double ui = hu[i];  // u of row i (column 0, the root of the tree)
            int cur = 0, row = i;
            double urow = ui;
            while (true) {
#pragma unroll
                for (int t = 0; t < T; ++t)
                    if (kWave * t + lane + 1 == cur) used[t] = true;
                const double* crow = STAGED ? hcost + (size_t)(row - 1) * k : nullptr;
                const size_t rbase = STAGED ? 0 : (size_t)rows[row - 1] * m;
                MinPair best{kNone, -1};
#pragma unroll
                for (int t = 0; t < T; ++t) {
                    const int j = kWave * t + lane + 1;
                    if (j <= k && !used[t]) {
                        double cr;
                        if (STAGED) {
                            cr = crow[j - 1];
                        } else {
                            const size_t e = rbase + cols[j - 1];
                            cr = (ub[e] >= 0.5 && (!face || face[e] > 1e-9)) ? cost[e] : inf;
                        }
                        const double c = cr - urow - hv[t];
                        if (c < minv[t]) { minv[t] = c; way[t] = cur; }
                        if (minv[t] < kNone) best = better_sel(best, MinPair{minv[t], j});
                    }
                }
                best = wave_argmin_2p(best);
                if (best.i < 0 || !isfinite(best.v)) { wave_sync(); return false; }
                const double delta = best.v;
                ui = ui + delta;
#pragma unroll
                for (int t = 0; t < T; ++t) {
                    if (kWave * t + lane + 1 <= k) {
                        if (used[t]) {
                            uj[t] = uj[t] + delta;
                            hv[t] = hv[t] - delta;
                        } else {
                            minv[t] = minv[t] - delta;
                        }
                    }
                }
                cur = best.i;
                const int owner = (cur - 1) & (kWave - 1), tc = (cur - 1) / kWave;
                int hsel = hpj[0];
                double usel = uj[0];
#pragma unroll
                for (int t = 1; t < T; ++t)
                    if (tc == t) { hsel = hpj[t]; usel = uj[t]; }
                const int next = __builtin_amdgcn_readlane(hsel, owner);
                if (next == 0) break;
                row = next;
                urow = lane_double(usel, owner);
            }
            // row potentials back to LDS, then augment along the recorded path
#pragma unroll
            for (int t = 0; t < T; ++t) {
                const int j = kWave * t + lane + 1;
                if (j <= k) {
                    if (hpj[t] > 0) hu[hpj[t]] = uj[t];
                    wy[j] = way[t];
                }
            }
            if (lane == 0) hu[i] = ui;
            wave_sync();
            if (lane == 0) {
                do { const int prev = wy[cur]; hp[cur] = hp[prev]; cur = prev; } while (cur != 0);
            }
            wave_sync();
        }
        for (int j = lane + 1; j <= k; j += kWave) out[(size_t)rows[hp[j] - 1] * m + cols[j - 1]] = 1.0;
        wave_sync();
        return true;
    }

    __device__ static bool solve(const RelaxParams& P, const Lds& s, const double* cost,
                                 const double* lb, const double* ub, const double* face, double* out) {
        for (int e = threadIdx.x; e < P.n; e += blockDim.x) out[e] = 0.0;
        __syncthreads();
        if (P.n != P.lorder * P.lorder) return false;
        if (threadIdx.x < kWave) {
            const bool ok = forced_edges(P, s, lb, ub, face, out);
            if (threadIdx.x == 0) s.misc[2] = ok ? 1 : 0;
        }
        __syncthreads();
        bool ok = s.misc[2] != 0;
        if (ok) {
            const int m = P.lorder;
            const int k = s.misc[0];
            if (s.hcost) {
                const double inf = __builtin_inf();
                for (int e = threadIdx.x; e < k * k; e += blockDim.x) {
                    const int r = e / k, c = e - r * k;
                    const size_t E = (size_t)s.rows[r] * m + s.cols[c];
                    s.hcost[e] = (ub[E] >= 0.5 && (!face || face[E] > 1e-9)) ? cost[E] : inf;
                }
            }
            __syncthreads();
            if (threadIdx.x < kWave) {
                extern __shared__ __align__(16) unsigned char lds_raw[];
                auto off = [&](const void* p) {
                    return (unsigned)(reinterpret_cast<const unsigned char*>(p) - lds_raw);
                };
                const unsigned o_hu = off(s.hu), o_hp = off(s.hp), o_wy = off(s.way), o_r = off(s.rows),
                               o_c = off(s.cols), o_q = s.hcost ? off(s.hcost) : 0u;
#define GB_ASSIGN(TT)                                                                            \
    (s.hcost ? assign<TT, true>(m, k, o_hu, o_hp, o_wy, o_r, o_c, o_q, cost, ub, face, out)       \
             : assign<TT, false>(m, k, o_hu, o_hp, o_wy, o_r, o_c, o_q, cost, ub, face, out))
                bool r;
                if (k <= kWave) r = GB_ASSIGN(1);
                else if (k <= 2 * kWave) r = GB_ASSIGN(2);
                else if (k <= 3 * kWave) r = GB_ASSIGN(3);
                else r = GB_ASSIGN(4);
#undef GB_ASSIGN
                if (threadIdx.x == 0) s.misc[2] = r ? 1 : 0;
            }
            __syncthreads();
            ok = s.misc[2] != 0;
        }
        __syncthreads();
        return ok;
    }
    __device__ static bool feasible(const RelaxParams& P, const Lds& s, const double* lb,
                                    const double* ub, double* scratch) {
        if (P.n != P.lorder * P.lorder) return false;
        // Zero direction: scratch holds both the cost (zeros) and the output.
        double* cost = scratch;
        double* out = scratch + P.n;
        for (int e = threadIdx.x; e < P.n; e += blockDim.x) cost[e] = 0.0;
        __syncthreads();
        return solve(P, s, cost, lb, ub, nullptr, out);
    }
    __device__ static void compute_v(const RelaxParams& P, const Lds& s, const double* g,
                                     const double* lb, const double* ub, double* v) {
        solve(P, s, g, lb, ub, nullptr, v);
    }
    __device__ static void inface(const RelaxParams& P, const Lds& s, const double* g,
                                  const double* x, const double* lb, const double* ub, double* a) {
        solve(P, s, g, lb, ub, x, a);
    }
    __device__ static double max_step(const RelaxParams& P, const Lds& s, const double* d,
                                      const double* x, const double* lb, const double* ub) {
        double r = 1.0;
        for (int j = threadIdx.x; j < P.n; j += blockDim.x) {
            const double dj = d[j];
            if (dj > 0.0) r = fmin(r, (x[j] - lb[j]) / dj);
            else if (dj < 0.0) r = fmin(r, (ub[j] - x[j]) / -dj);
        }
        const double coord = block_min(r, s.red);
        return fmax(0.0, fmin(coord, 1.0));
    }
    __device__ static bool rounded_feasible(const RelaxParams& P, const Lds& s, const double* x) {
        const int m = P.lorder;
        if (P.n != m * m) return false;
        bool valid = true;
        for (int i = threadIdx.x; i < m; i += blockDim.x) {
            double row = 0.0, col = 0.0;
            for (int j = 0; j < m; ++j) {
                if (x[i * m + j] < -P.ctol) valid = false;
                row = row + x[i * m + j];
                col = col + x[j * m + i];
            }
            if (fabs(row - 1.0) > P.ctol || fabs(col - 1.0) > P.ctol) valid = false;
        }
        return block_and(valid, s.ired);
    }
};

// ---------------------------------------------------------------------------
// Node relaxation body
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Candidate heuristics and the hooks around the LMO
// ---------------------------------------------------------------------------

// Hooks with built-in defaults; the plug-in adapters below overload them.
template <class L>
__device__ __forceinline__ bool lmo_feasible(const L& l, const RelaxParams& P, const Lds& s, const double* lb,
                                             const double* ub, double* v, double* /*d*/) {
    return l.feasible(P, s, lb, ub, v);
}
template <class L>
__device__ __forceinline__ bool lmo_rounded_feasible(const L& l, const RelaxParams& P, const Lds& s,
                                                     const double* c, const double*, const double*) {
    return l.rounded_feasible(P, s, c);
}
template <class L>
__device__ __forceinline__ bool lmo_integer_candidate(const L&, const RelaxParams&, const double*, const double*,
                                                      const double*, double*) {
    return false;
}
template <class L>
__device__ __forceinline__ const L& lmo_object(const L& l) { return l; }
template <class O>
__device__ __forceinline__ void objective_prepare(const O&, const RelaxParams&, const Lds&) {}

// Built-in candidate generation (heuristics.h:23-35): integer coordinates
// rounded (or floor + offset), continuous ones kept.  A plug-in LMO's
// integer_candidate_device takes precedence under the rounding rule, as in
// the reference's RoundingHeuristic.
struct BuiltinHeuristic {
    template <class L>
    __device__ void generate(const RelaxParams& P, const Lds&, const L& lmo, const double* x,
                             const double* lb, const double* ub, double* c) const {
        if (P.heuristic == GB_HEUR_ROUND && lmo_integer_candidate(lmo, P, x, lb, ub, c)) return;
        for (int j = threadIdx.x; j < P.n; j += blockDim.x) {
            double cj = x[j];
            if (P.flags[j]) cj = P.heuristic == GB_HEUR_FLOOR_OFFSET ? floor(x[j]) + P.heuristic_offset
                                                                   : round(x[j]);
            c[j] = cj;
        }
    }
};

// ---------------------------------------------------------------------------
// Plug-in adapters: caller types with the reference's device contracts
// ---------------------------------------------------------------------------
namespace plugin_traits {
template <class T, class = void>
struct has_prepare : std::false_type {};
template <class T>
struct has_prepare<T, std::void_t<decltype(std::declval<const T&>().prepare_block(0, (unsigned char*)nullptr))>>
    : std::true_type {};
template <class T, class = void>
struct has_curvature : std::false_type {};
template <class T>
struct has_curvature<T, std::void_t<decltype(std::declval<const T&>().quadratic_curvature_local(
                            0, (const double*)nullptr, (unsigned char*)nullptr))>> : std::true_type {};
template <class T, class = void>
struct has_inface : std::false_type {};
template <class T>
struct has_inface<T, std::void_t<decltype(std::declval<const T&>().compute_inface_v(
                         0, 0, (const double*)nullptr, (const double*)nullptr, (const double*)nullptr,
                         (const double*)nullptr, (double*)nullptr)),
                     decltype(std::declval<const T&>().dicg_maximum_step_local(
                         0, 0, (const double*)nullptr, (const double*)nullptr, (const double*)nullptr,
                         (const double*)nullptr)),
                     decltype(std::declval<const T&>().dicg_sum_maximum_step(0, 0, (const double*)nullptr,
                                                                             (const double*)nullptr))>>
    : std::true_type {};
// heuristics.h:11-18
template <class T, class = void>
struct has_integer_candidate : std::false_type {};
template <class T>
struct has_integer_candidate<T, std::void_t<decltype(std::declval<const T&>().integer_candidate_device(
                                    0, 0, (const double*)nullptr, (const double*)nullptr, (const double*)nullptr,
                                    (const uint8_t*)nullptr, (double*)nullptr))>> : std::true_type {};
// device_resident_solver.cuh:229-245
template <class T, class = void>
struct has_candidate_feasible : std::false_type {};
template <class T>
struct has_candidate_feasible<T, std::void_t<decltype(std::declval<const T&>().candidate_feasible_device(
                                     0, (const double*)nullptr, (const double*)nullptr, (const double*)nullptr,
                                     0.0))>> : std::true_type {};
template <class T, class = void>
struct has_cooperative_candidate_feasible : std::false_type {};
template <class T>
struct has_cooperative_candidate_feasible<T, std::void_t<decltype(std::declval<const T&>().candidate_feasible_device(
                                                 0, 0, (const double*)nullptr, (const double*)nullptr,
                                                 (const double*)nullptr, 0.0))>> : std::true_type {};
}  // namespace plugin_traits

// A caller objective (device_resident_solver.cuh:8-24): value_local(tid, x,
// mem), grad(tid, x, g, mem), optional prepare_block(tid, mem) and
// quadratic_curvature_local(tid, d, mem).  mem is the node's slice of the
// objective workspace (global memory).  Partial values are added in the
// canonical wave64 order.
template <class U>
struct PluginObjective {
    U u;
    __device__ unsigned char* mem(const RelaxParams& P) const {
        return P.obj_work ? P.obj_work + (size_t)blockIdx.x * P.obj_stride : nullptr;
    }
    __device__ void grad(const RelaxParams& P, const Lds&, double*, const double* x, double* g) const {
        u.grad(threadIdx.x, x, g, mem(P));
        __syncthreads();
    }
    __device__ double value(const RelaxParams& P, const Lds& s, double*, const double* x) const {
        return block_sum(u.value_local(threadIdx.x, x, mem(P)), s.red);
    }
    __device__ double curvature(const RelaxParams& P, const Lds& s, double*, const double* d) const {
        if constexpr (plugin_traits::has_curvature<U>::value)
            return block_sum(u.quadratic_curvature_local(threadIdx.x, d, mem(P)), s.red);
        else
            return 0.0;  // the library rejects the exact step for such objectives
    }
};
template <class U>
__device__ __forceinline__ void objective_prepare(const PluginObjective<U>& o, const RelaxParams& P, const Lds&) {
    if constexpr (plugin_traits::has_prepare<U>::value) {
        o.u.prepare_block(threadIdx.x, o.mem(P));
        __syncthreads();
    }
}

// A caller LMO (lmo.cuh:11-66): compute_v(tid, n, g, lb, ub, v) and
// check_feasible_device(tid, n, lb, ub, flags, v, d); the DICG trio
// compute_inface_v / dicg_maximum_step_local / dicg_sum_maximum_step
// (combined as device_resident_solver.cuh:533-545); candidate hooks
// integer_candidate_device and candidate_feasible_device (cooperative, or
// the legacy form on thread 0).
template <class U>
struct PluginLmo {
    U u;
    __device__ void compute_v(const RelaxParams& P, const Lds&, const double* g, const double* lb,
                              const double* ub, double* v) const {
        u.compute_v(threadIdx.x, P.n, g, lb, ub, v);
    }
    __device__ void inface(const RelaxParams& P, const Lds&, const double* g, const double* x, const double* lb,
                           const double* ub, double* a) const {
        if constexpr (plugin_traits::has_inface<U>::value) {
            u.compute_inface_v(threadIdx.x, P.n, g, x, lb, ub, a);
        } else {  // the library rejects DICG for such LMOs
            for (int j = threadIdx.x; j < P.n; j += blockDim.x) a[j] = x[j];
        }
    }
    __device__ double max_step(const RelaxParams& P, const Lds& s, const double* d, const double* x,
                               const double* lb, const double* ub) const {
        if constexpr (plugin_traits::has_inface<U>::value) {
            const double coord = block_min(u.dicg_maximum_step_local(threadIdx.x, P.n, d, x, lb, ub), s.red);
            const double sum = u.dicg_sum_maximum_step(threadIdx.x, P.n, d, x);
            return fmax(0.0, fmin(coord, sum));
        } else {
            return 0.0;
        }
    }
};
template <class U>
__device__ __forceinline__ bool lmo_feasible(const PluginLmo<U>& l, const RelaxParams& P, const Lds&,
                                             const double* lb, const double* ub, double* v, double* d) {
    return l.u.check_feasible_device(threadIdx.x, P.n, lb, ub, P.flags, v, d);
}
template <class U>
__device__ __forceinline__ bool lmo_rounded_feasible(const PluginLmo<U>& l, const RelaxParams& P, const Lds& s,
                                                     const double* c, const double* lb, const double* ub) {
    if constexpr (plugin_traits::has_cooperative_candidate_feasible<U>::value) {
        return l.u.candidate_feasible_device(threadIdx.x, P.n, c, lb, ub, P.ctol);
    } else if constexpr (plugin_traits::has_candidate_feasible<U>::value) {
        if (threadIdx.x == 0) s.ired[0] = l.u.candidate_feasible_device(P.n, c, lb, ub, P.ctol) ? 1 : 0;
        __syncthreads();
        const bool ok = s.ired[0] != 0;
        __syncthreads();
        return ok;
    } else {
        return false;  // an unknown set skips rounding
    }
}
template <class U>
__device__ __forceinline__ bool lmo_integer_candidate(const PluginLmo<U>& l, const RelaxParams& P, const double* x,
                                                      const double* lb, const double* ub, double* c) {
    if constexpr (plugin_traits::has_integer_candidate<U>::value) {
        l.u.integer_candidate_device(threadIdx.x, P.n, x, lb, ub, P.flags, c);
        return true;
    } else {
        return false;
    }
}
template <class U>
__device__ __forceinline__ const U& lmo_object(const PluginLmo<U>& l) { return l.u; }

// A caller heuristic (heuristics.h:11-35): generate(lmo, tid, n, x, lb, ub,
// flags, candidate) fills every candidate coordinate.  A built-in LMO is
// passed as the library's own (empty) LMO type.
template <class UH>
struct PluginHeuristic {
    UH h;
    template <class L>
    __device__ void generate(const RelaxParams& P, const Lds&, const L& lmo, const double* x, const double* lb,
                             const double* ub, double* c) const {
        h.generate(lmo_object(lmo), threadIdx.x, P.n, x, lb, ub, P.flags, c);
    }
};

__device__ __forceinline__ double quad_exact_step(double gap, double curvature, double maximum) {
    // line_search_options.h:48-63
    if (maximum <= 0.0) return 0.0;
    if (curvature > 0.0) {
        if (gap <= 0.0) return 0.0;
        const double step = gap / curvature;
        return step < maximum ? step : maximum;
    }
    return -gap + 0.5 * curvature * maximum < 0.0 ? maximum : 0.0;
}

template <class O>
struct Node {
    const O& obj;
    const RelaxParams& P;
    const Lds& s;
    const double* lb;
    const double* ub;
    double* R;

    // drs_integer_value (device_resident_solver.cuh:297-313)
    __device__ double integer_value(const double* v) {
        bool valid = true;
        for (int j = threadIdx.x; j < P.n; j += blockDim.x) {
            const double vj = v[j];
            if (!isfinite(vj) || vj < lb[j] - P.ctol || vj > ub[j] + P.ctol
                    || (P.flags && P.flags[j] && vj != round(vj)))
                valid = false;
        }
        if (!block_and(valid, s.ired)) return kNone;
        const double val = obj.value(P, s, R, v);
        return isfinite(val) ? val : kNone;
    }

    // SecantLS::search (algorithms.cuh:121-258).  All scalars are uniform.
    __device__ double secant(double gap, double max_gamma, const double* x, const double* dir,
                             double* tx, double* tg) {
        const double tol = P.secant_tol;
        const double f_base = obj.value(P, s, R, x);
        double lo = 0.0, hi = max_gamma, best_f = f_base, best_gamma = 0.0;
        double phi_lo = -gap, phi_hi = 0.0, gamma = 0.0;
        bool done = phi_lo >= -tol;
        const double f_abs = f_base >= 0.0 ? f_base : -f_base;
        const double f_allow = tol * (1.0 + f_abs);
        auto trial = [&](double gm, double& phi, double& f) {
            for (int j = threadIdx.x; j < P.n; j += blockDim.x) tx[j] = fma(gm, dir[j], x[j]);
            __syncthreads();
            obj.grad(P, s, R, tx, tg);
            double acc = 0.0;
            for (int j = threadIdx.x; j < P.n; j += blockDim.x) acc = fma(tg[j], dir[j], acc);
            phi = block_sum(acc, s.red);
            f = obj.value(P, s, R, tx);
        };
        double f_hi;
        trial(max_gamma, phi_hi, f_hi);
        if (f_hi < best_f || (best_gamma == 0.0 && f_hi <= f_base + f_allow)) {
            best_f = f_hi;
            best_gamma = max_gamma;
        }
        if (!done && phi_hi <= tol && f_hi <= f_base + f_allow) { gamma = max_gamma; done = true; }
        for (int it = 0; it < P.secant_iters; ++it) {
            if (done) break;
            const double denom = phi_hi - phi_lo;
            double gm = (denom > P.secant_denom || denom < -P.secant_denom)
                ? hi - phi_hi * (hi - lo) / denom : 0.5 * (lo + hi);
            if (!(gm > lo && gm < hi)) gm = 0.5 * (lo + hi);
            gamma = gm;
            double phi, ft;
            trial(gamma, phi, ft);
            if (ft < best_f || (best_gamma == 0.0 && ft <= f_base + f_allow)) { best_f = ft; best_gamma = gamma; }
            if (phi < 0.0) { lo = gamma; phi_lo = phi; } else { hi = gamma; phi_hi = phi; }
            if ((phi <= tol && phi >= -tol) || (hi - lo) <= tol) { gamma = 0.5 * (lo + hi); done = true; }
        }
        if (best_gamma > 0.0 && best_f <= f_base + f_allow) gamma = best_gamma;
        else if (!done) gamma = 0.0;
        return gamma;
    }
};

__device__ __forceinline__ Lds lds_views(unsigned char* lds_raw, const LdsLayout& LL) {
    Lds s;
    s.red = reinterpret_cast<double*>(lds_raw + LL.red);
    s.ired = reinterpret_cast<int*>(lds_raw + LL.ired);
    s.vec = reinterpret_cast<double*>(lds_raw + LL.vec);
    s.sg = reinterpret_cast<double*>(lds_raw + LL.sg);
    s.sidx = reinterpret_cast<int*>(lds_raw + LL.sidx);
    s.scap = reinterpret_cast<double*>(lds_raw + LL.scap);
    s.spre = reinterpret_cast<double*>(lds_raw + LL.spre);
    s.elig = lds_raw + LL.elig;
    s.fixed = lds_raw + LL.fixed;
    s.fcol = reinterpret_cast<int*>(lds_raw + LL.fcol);
    s.frow = reinterpret_cast<int*>(lds_raw + LL.frow);
    s.rows = reinterpret_cast<int*>(lds_raw + LL.rows);
    s.cols = reinterpret_cast<int*>(lds_raw + LL.cols);
    s.hu = reinterpret_cast<double*>(lds_raw + LL.hu);
    s.hp = reinterpret_cast<int*>(lds_raw + LL.hp);
    s.way = reinterpret_cast<int*>(lds_raw + LL.way);
    s.misc = reinterpret_cast<int*>(lds_raw + LL.misc);
    s.hcost = LL.hcost_bytes ? reinterpret_cast<double*>(lds_raw + LL.hcost) : nullptr;
    return s;
}

// Extreme points of the built-in LMOs for given directions (one workgroup
// per node, the compute_v of the relaxation kernel).
template <class L>
__global__ void __launch_bounds__(1024) lmo_probe_kernel(RelaxParams P, LdsLayout LL, const double* G, L lmo) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const Lds s = lds_views(lds_raw, LL);
    const int node = blockIdx.x;
    if (node >= P.N) return;
    const size_t off = (size_t)node * P.n;
    lmo.compute_v(P, s, G + off, P.lb + off, P.ub + off, P.x + off);
}

template <class O, class L, class H>
__global__ void __launch_bounds__(1024) relax_fp64_kernel(RelaxParams P, LdsLayout LL, O obj, L lmo, H heur) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const Lds s = lds_views(lds_raw, LL);
    const int node = blockIdx.x;
    if (node >= P.N) return;
    const int n = P.n;
    const int tid = threadIdx.x;
    const int T = blockDim.x;
    const double eps = P.eps[node];
    const int initial = P.iters[node];
    const bool supplied = (P.status[node] & GB_NODE_INITIAL_POINT_SUPPLIED) != 0;

    double* x = P.x + (size_t)node * n;
    const double* lb = P.lb + (size_t)node * n;
    const double* ub = P.ub + (size_t)node * n;
    double* best_v = P.best_v + (size_t)node * n;
    double* g = P.work + (size_t)node * kWorkVectors * n;
    double* v = g + n;
    double* aw = v + n;
    double* tx = aw + n;
    double* tg = tx + n;
    double* R = tg + n;
    Node<O> nd{obj, P, s, lb, ub, R};

    // Node feasibility before any objective access (device_resident_solver.cuh:428-435).
    bool valid = true;
    for (int j = tid; j < n; j += T)
        if (isnan(lb[j]) || isnan(ub[j]) || lb[j] > ub[j]
                || (P.flags && P.flags[j] && ceil(lb[j]) > floor(ub[j])))
            valid = false;
    bool feasible = block_and(valid, s.ired);
    if (feasible) feasible = lmo_feasible(lmo, P, s, lb, ub, v, tx);
    if (!feasible) {
        if (tid == 0) {
            P.status[node] = GB_NODE_INFEASIBLE;
            P.value[node] = __builtin_inf();
            P.gap[node] = 0.0;
            P.best_val[node] = kNone;
            P.iters[node] = initial;
        }
        return;
    }
    objective_prepare(obj, P, s);

    double best_s = kNone;
    auto track = [&](const double* vert) {
        const double fv = nd.integer_value(vert);
        if (fv < best_s) best_s = fv;
        if (fv < 1e299 && fv == best_s)
            for (int j = tid; j < n; j += T) best_v[j] = vert[j];
        __syncthreads();
    };

    if (initial == 0 && !supplied) {
        for (int j = tid; j < n; j += T) g[j] = -1.0 - (double)j / (n + 1.0);
        __syncthreads();
        lmo.compute_v(P, s, g, lb, ub, v);
        __syncthreads();
        for (int j = tid; j < n; j += T) x[j] = v[j];
        __syncthreads();
    }

    const bool secant_like = P.ls_kind == GB_LS_SECANT || P.ls_kind == GB_LS_QUADRATIC_EXACT;
    int iterations = initial;
    for (int iter = 0; iter < P.max_iters; ++iter) {
        iterations = initial + iter + 1;
        obj.grad(P, s, R, x, g);
        lmo.compute_v(P, s, g, lb, ub, v);
        __syncthreads();
        track(v);
        double acc = 0.0;
        for (int j = tid; j < n; j += T) acc = fma(g[j], x[j] - v[j], acc);
        const double gap = block_sum(acc, s.red);
        if (gap >= 0.0 && gap < eps) break;
        if (!P.dicg) {
            if (secant_like) {
                for (int j = tid; j < n; j += T) v[j] = v[j] - x[j];
                __syncthreads();
                double gamma;
                if (P.ls_kind == GB_LS_SECANT) gamma = nd.secant(gap, 1.0, x, v, tx, tg);
                else gamma = quad_exact_step(gap, obj.curvature(P, s, R, v), 1.0);
                for (int j = tid; j < n; j += T) x[j] = fma(gamma, v[j], x[j]);
            } else {
                const double gamma = 2.0 / (double)(initial + iter + 2);
                for (int j = tid; j < n; j += T) x[j] = fma(gamma, v[j] - x[j], x[j]);
            }
            __syncthreads();
            continue;
        }
        for (int j = tid; j < n; j += T) g[j] = -g[j];
        __syncthreads();
        lmo.inface(P, s, g, x, lb, ub, aw);
        __syncthreads();
        track(aw);
        for (int j = tid; j < n; j += T) { g[j] = -g[j]; aw[j] = aw[j] - v[j]; }
        __syncthreads();
        const double gmax = lmo.max_step(P, s, aw, x, lb, ub);
        if (secant_like) {
            double pacc = 0.0;
            for (int j = tid; j < n; j += T) {
                aw[j] = -aw[j];
                pacc = fma(-g[j], aw[j], pacc);
            }
            const double pgap = block_sum(pacc, s.red);
            double gamma;
            if (P.ls_kind == GB_LS_SECANT) gamma = nd.secant(pgap, gmax, x, aw, tx, tg);
            else gamma = quad_exact_step(pgap, obj.curvature(P, s, R, aw), gmax);
            for (int j = tid; j < n; j += T) x[j] = fma(gamma, aw[j], x[j]);
        } else {
            const double ol = 2.0 / (double)(initial + iter + 2);
            double gamma;
            if (P.ls_kind == GB_LS_SAFEGUARDED_SCALED_OPEN_LOOP) gamma = gmax <= ol ? gmax : gmax * ol;
            else if (P.ls_kind == GB_LS_SCALED_OPEN_LOOP) gamma = gmax * ol;
            else gamma = fmin(gmax, ol);
            for (int j = tid; j < n; j += T) x[j] = fma(-gamma, aw[j], x[j]);
        }
        __syncthreads();
    }

    // Final certificate at the returned iterate (device_resident_solver.cuh:588-621).
    obj.grad(P, s, R, x, g);
    lmo.compute_v(P, s, g, lb, ub, v);
    __syncthreads();
    track(v);
    double acc = 0.0;
    for (int j = tid; j < n; j += T) acc = fma(g[j], x[j] - v[j], acc);
    const double fgap = block_sum(acc, s.red);
    const double fval = obj.value(P, s, R, x);
    uint8_t st = (fgap >= 0.0 && fgap < eps) ? 1 : 0;

    // Candidate heuristic (heuristics.h:23-35, device_resident_solver.cuh:315-335).
    if (P.flags) {
        heur.generate(P, s, lmo, x, lb, ub, v);
        __syncthreads();
        if (lmo_rounded_feasible(lmo, P, s, v, lb, ub)) {
            const double cv = nd.integer_value(v);
            if (cv < best_s) {
                for (int j = tid; j < n; j += T) best_v[j] = v[j];
                best_s = cv;
            }
        }
    }
    bool integral = true;
    for (int j = tid; j < n; j += T)
        if (!isfinite(x[j]) || (P.flags && P.flags[j] && fabs(x[j] - round(x[j])) > P.ctol))
            integral = false;
    const bool all_int = block_and(integral, s.ired);
    if (all_int) st = st ? GB_NODE_INTEGRAL_CONVERGED : GB_NODE_INTEGRAL_UNCONVERGED;
    if (tid == 0) {
        P.value[node] = fval;
        P.gap[node] = fgap;
        P.status[node] = st;
        P.iters[node] = iterations;
        P.best_val[node] = best_s;
    }
}

// Launch the relaxation kernel with caller objects on a batch the library
// prepared (gb_plugin_desc.launch receives the PluginLaunch).  Returns a
// hipError_t value.
template <class O, class L, class H>
inline int launch_relax_plugin(const PluginLaunch& pl, const O& obj, const L& lmo, const H& heur, void* stream) {
    const void* k = reinterpret_cast<const void*>(&relax_fp64_kernel<O, L, H>);
    if (pl.lds_bytes > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pl.lds_bytes);
        if (e != hipSuccess) return (int)e;
    }
    RelaxParams P = pl.P;
    LdsLayout LL = pl.L;
    O o = obj;
    L l = lmo;
    H h = heur;
    void* args[] = {&P, &LL, &o, &l, &h};
    return (int)hipLaunchKernel(k, dim3(P.N), dim3(pl.threads), args, pl.lds_bytes,
                                reinterpret_cast<hipStream_t>(stream));
}

// Extreme points of an LMO for directions G ([N][n]) into P.x, bounds P.lb /
// P.ub: one workgroup per row, as inside the relaxation.  Returns a
// hipError_t value.
template <class L>
inline int launch_lmo_probe_with(const RelaxParams& p, int block_threads, size_t lds_bytes, const LdsLayout& LL,
                                 const double* G, const L& lmo, void* stream) {
    const void* k = reinterpret_cast<const void*>(&lmo_probe_kernel<L>);
    if (lds_bytes > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
        if (e != hipSuccess) return (int)e;
    }
    RelaxParams P = p;
    LdsLayout L2 = LL;
    const double* g = G;
    L l = lmo;
    void* args[] = {&P, &L2, &g, &l};
    return (int)hipLaunchKernel(k, dim3(P.N), dim3(block_threads), args, lds_bytes,
                                reinterpret_cast<hipStream_t>(stream));
}

}  // namespace
}  // namespace gb

// Block-wide sums for plug-in code, under the reference's helper names
// (lmo.cuh:91-120 feasibility_block_sum, :188-210 lmo_block_sum): every
// thread of the workgroup calls them and receives the total.  Wave64
// shuffle trees, wave totals added in wave order.
__device__ inline double feasibility_block_sum(double value) {
    __shared__ double partial[gb::kMaxWaves];
    __shared__ double total;
    value = gb::wave_tree_sum(value);
    if (gb::lane_id() == 0) partial[gb::wave_id()] = value;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = partial[0];
        for (int w = 1; w < gb::wave_count(); ++w) t = t + partial[w];
        total = t;
    }
    __syncthreads();
    const double r = total;
    __syncthreads();
    return r;
}
__device__ inline double lmo_block_sum(double value, int /*tid*/) { return feasibility_block_sum(value); }
