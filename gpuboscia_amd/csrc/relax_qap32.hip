// FP32 form of the QAP relaxation over the Birkhoff polytope, for gfx950.
//
// min ||A P - P B||_F^2 over doubly stochastic P (the configuration's
// "FP32 iterates, Hungarian on FP32 costs, FP64 bound").  One workgroup of
// 16 waves per B&B node; DICG with the exact quadratic step.
//
//   * Iterate P and gradient G = 2 (A'R - R B'), R = A P - P B, in FP32.
//     A DICG step moves mass between two permutations V (forward vertex) and
//     W (away vertex), D = V - W, nonzero only on the K rows where they
//     differ, so G(D) = 2 (A'A D - A'D B - A D B' + D B B') costs O(K m^2)
//     instead of a dense m^3 product: A'A D and D B B' read one entry each
//     (A'A, B B' precomputed), A'D B and A D B' are rank-K products staged
//     through LDS in chunks of 32 rows.  G is recomputed from P every
//     `refresh` steps.
//   * Both linear minimisations are exact assignment problems on FP32 costs
//     (the staged m x m cost matrix fills the CU's LDS at m = 200), solved by
//     the shortest-augmenting-path Hungarian with FP64 potentials on wave 0,
//     warm-started from the previous solve of the same oracle: column
//     potentials are kept, row potentials re-reduced, and only the matched
//     edges that stay tight are kept, so a step that changes few costs
//     re-augments few rows.  Lane l holds columns l + 1 + 64 t.
//   * The node's certificate is FP64: R and G recomputed from P in FP64, the
//     vertex of FP32(G) by the same solver, and the Lagrangian bound of its
//     column potentials v against the FP64 costs,
//         LB(v) = sum_j v_j + sum_i min_j (G_ij - v_j) <= min_V <G, V>,
//     which holds for any v, so gap = <G, P> - LB(v) is an FP64-certified
//     Frank-Wolfe gap and value - gap a valid bound whatever FP32 did.
//     Vertex values are exact FP64 sums.
#include <hip/hip_runtime.h>

#include <climits>

#include "gb_device.h"
#include "gb_internal.h"
#include "gb_qap32.h"

namespace gb {
namespace {

constexpr float kQInf = __builtin_inff();
constexpr double kDInf = __builtin_inf();

struct QView {
    float* cost;           // staged costs, or the gradient pass's chunk
    double* vsh;           // [m + 1] column potentials
    double* urow;          // [m + 1] row potentials after reduction
    unsigned char* rowm;   // [m + 1] matched rows
    double* red;           // [16] reductions
    int* ired;             // [16]
};

__device__ __forceinline__ QView qview(unsigned char* raw, const QapLds& L) {
    QView v;
    v.cost = reinterpret_cast<float*>(raw + L.cost);
    v.vsh = reinterpret_cast<double*>(raw + L.vsh);
    v.urow = reinterpret_cast<double*>(raw + L.urow);
    v.rowm = raw + L.rowm;
    v.red = reinterpret_cast<double*>(raw + L.red);
    v.ired = reinterpret_cast<int*>(raw + L.red + 16 * sizeof(double));
    return v;
}

__device__ __forceinline__ int rl_i(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ __forceinline__ double rl_d(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

template <class T>
__device__ __forceinline__ T sel4(const T (&a)[4], int t) {
    return t == 0 ? a[0] : (t == 1 ? a[1] : (t == 2 ? a[2] : a[3]));
}
template <class T>
__device__ __forceinline__ void set4(T (&a)[4], int t, T v) {
    if (t == 0) a[0] = v;
    else if (t == 1) a[1] = v;
    else if (t == 2) a[2] = v;
    else a[3] = v;
}

// Wave-wide minima by DPP within rows of 16 lanes (quad permutes, half-row
// and row mirrors) and lane reads across the four rows; wave-uniform.
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    return __hiloint2double(dpp_i<CTRL>(__double2hiint(v)), dpp_i<CTRL>(__double2loint(v)));
}
__device__ __forceinline__ double wave_min_all(double v) {
    v = fmin(v, dpp_d<0xB1>(v));   // lane ^ 1
    v = fmin(v, dpp_d<0x4E>(v));   // lane ^ 2
    v = fmin(v, dpp_d<0x141>(v));  // mirror within 8
    v = fmin(v, dpp_d<0x140>(v));  // mirror within 16
    return fmin(fmin(rl_d(v, 0), rl_d(v, 16)), fmin(rl_d(v, 32), rl_d(v, 48)));
}
__device__ __forceinline__ int wave_min_int(int v) {
    v = min(v, dpp_i<0xB1>(v));
    v = min(v, dpp_i<0x4E>(v));
    v = min(v, dpp_i<0x141>(v));
    v = min(v, dpp_i<0x140>(v));
    return min(min(rl_i(v, 0), rl_i(v, 16)), min(rl_i(v, 32), rl_i(v, 48)));
}

// (value, column) minimum with ties to the lower column: the order of the
// reference's column scan.  Values are never NaN.
__device__ __forceinline__ void wave_argmin_dj(double& v, int& j) {
    const double vm = wave_min_all(v);
    j = wave_min_int(v == vm ? j : INT_MAX);
    v = vm;
}

// The staged costs are read through an LDS pointer (ds_read), the global
// copy through a global one.
typedef __attribute__((address_space(3))) const float lds_cfloat;
template <bool kL>
__device__ __forceinline__ float cost_at(const float* C, size_t e) {
    if constexpr (kL) return ((const lds_cfloat*)C)[e];
    else return C[e];
}

// Assignment of minimum cost on C (m x m row-major FP32, +inf forbidden),
// warm-started from (vcol, pcol) and leaving the new state there; sigma
// gets row -> column.  Every thread calls it; returns false (block-uniform)
// when some row cannot be matched.
template <bool kL>
__device__ bool assign(int m, const float* C, double* vcol, int* pcol, int* sigma, const QView& V) {
    const int tid = threadIdx.x, lane = lane_id(), w = wave_id(), W = wave_count();
    for (int j = tid; j <= m; j += blockDim.x) V.vsh[j] = vcol[j];
    __syncthreads();
    // row potentials: u_i = min_j (c_ij - v_j), dual feasible for every
    // row, and the lowest column attaining it (in sigma until the end)
    for (int i = w; i < m; i += W) {
        double r = kDInf;
        int rj = INT_MAX;
        for (int j = lane; j < m; j += kWave) {
            const float c = cost_at<kL>(C, (size_t)i * m + j);
            if (c < kQInf) {
                const double d = (double)c - V.vsh[j + 1];
                if (d < r) {
                    r = d;
                    rj = j;
                }
            }
        }
        wave_argmin_dj(r, rj);
        if (lane == 0) {
            V.urow[i + 1] = r;
            sigma[i] = rj;
        }
    }
    __syncthreads();
    if (w == 0) {
        double v[4], up[4], mv[4];
        int p[4], way[4];
        bool used[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int j = 1 + lane + 64 * t;
            const bool valid = j <= m;
            v[t] = valid ? V.vsh[j] : 0.0;
            int pi = valid ? pcol[j] : 0;
            if (pi != 0) {  // keep the edge only if it is still tight
                const float c = cost_at<kL>(C, (size_t)(pi - 1) * m + (j - 1));
                if (!(c < kQInf) || ((double)c - v[t]) != V.urow[pi]) pi = 0;
            }
            p[t] = pi;
            up[t] = pi ? V.urow[pi] : 0.0;
            way[t] = 0;
            used[t] = false;
            mv[t] = kDInf;
        }
        for (int i = lane; i <= m; i += kWave) V.rowm[i] = 0;
        wave_sync();
#pragma unroll
        for (int t = 0; t < 4; ++t)
            if (p[t]) V.rowm[p[t]] = 1;
        wave_sync();
        // Greedy start: a free row takes its row minimum's column when that
        // column is free (reduced cost 0, so the duals stay feasible and the
        // matched edges tight).
        for (int i = 1; i <= m; ++i) {
            if (V.rowm[i]) continue;
            const int c = sigma[i - 1];
            if (c == INT_MAX) continue;
            const int o = c & 63, t0 = c >> 6;
            if (rl_i(sel4(p, t0), o) != 0) continue;
            if (lane == o) {
                set4(p, t0, i);
                set4(up, t0, V.urow[i]);
            }
            if (lane == 0) V.rowm[i] = 1;
        }
        wave_sync();
        bool ok = true;
        for (int i = 1; i <= m && ok; ++i) {
            if (V.rowm[i]) continue;
            double ui = V.urow[i];
            if (!(ui < kDInf)) {
                ok = false;
                break;
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                mv[t] = kDInf;
                used[t] = false;
            }
            int j0 = 0;
            for (;;) {
                int i0;
                double ui0;
                if (j0 == 0) {
                    i0 = i;
                    ui0 = ui;
                } else {
                    const int o = (j0 - 1) & 63, t0 = (j0 - 1) >> 6;
                    i0 = rl_i(sel4(p, t0), o);
                    ui0 = rl_d(sel4(up, t0), o);
                    if (lane == o) set4(used, t0, true);
                }
                const size_t row = (size_t)(i0 - 1) * m;
                double best = kDInf;
                int bj = INT_MAX;
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int j = 1 + lane + 64 * t;
                    if (j <= m && !used[t]) {
                        const float c = cost_at<kL>(C, row + (j - 1));
                        if (c < kQInf) {
                            const double cur = ((double)c - v[t]) - ui0;
                            if (cur < mv[t]) {
                                mv[t] = cur;
                                way[t] = j0;
                            }
                        }
                        if (mv[t] < best) {
                            best = mv[t];
                            bj = j;
                        }
                    }
                }
                wave_argmin_dj(best, bj);
                if (!(best < kDInf)) {
                    ok = false;
                    break;
                }
                const double delta = best;
                ui += delta;
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    if (1 + lane + 64 * t > m) continue;
                    if (used[t]) {
                        up[t] += delta;
                        v[t] -= delta;
                    } else {
                        mv[t] -= delta;
                    }
                }
                j0 = uni(bj);
                const int o = (j0 - 1) & 63, t0 = (j0 - 1) >> 6;
                if (rl_i(sel4(p, t0), o) == 0) break;
            }
            if (!ok) break;
            while (j0 != 0) {  // flip the alternating path
                const int o0 = (j0 - 1) & 63, t0 = (j0 - 1) >> 6;
                const int j1 = uni(rl_i(sel4(way, t0), o0));
                int pr;
                double pu;
                if (j1 == 0) {
                    pr = i;
                    pu = ui;
                } else {
                    const int o1 = (j1 - 1) & 63, t1 = (j1 - 1) >> 6;
                    pr = rl_i(sel4(p, t1), o1);
                    pu = rl_d(sel4(up, t1), o1);
                }
                if (lane == o0) {
                    set4(p, t0, pr);
                    set4(up, t0, pu);
                }
                j0 = j1;
            }
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int j = 1 + lane + 64 * t;
            if (j > m) continue;
            vcol[j] = v[t];
            pcol[j] = p[t];
            if (p[t]) sigma[p[t] - 1] = j - 1;
        }
        if (lane == 0) {
            vcol[0] = 0.0;
            pcol[0] = 0;
            V.ired[8] = ok ? 1 : 0;
        }
    }
    __syncthreads();
    const bool ok = V.ired[8] != 0;
    __syncthreads();
    return ok;
}

// Lagrangian bound of the column potentials in vcol on FP64 costs G (masked):
// sum_j v_j + sum_i min_{j allowed} (G_ij - v_j).
__device__ double lagrange_bound(int m, const double* G, const unsigned char* mask, const double* vcol,
                                 const QView& V) {
    const int lane = lane_id(), w = wave_id(), W = wave_count();
    for (int j = threadIdx.x; j <= m; j += blockDim.x) V.vsh[j] = vcol[j];
    __syncthreads();
    double part = 0.0;
    for (int i = w; i < m; i += W) {
        double r = kDInf;
        for (int j = lane; j < m; j += kWave)
            if (!mask || mask[(size_t)i * m + j]) r = fmin(r, G[(size_t)i * m + j] - V.vsh[j + 1]);
        r = wave_min_all(r);
        if (lane == 0) part += r;
    }
    for (int j = threadIdx.x; j < m; j += blockDim.x) part += V.vsh[j + 1];
    return block_sum(part, V.red);
}

// out = s (op(X1) op(Y1) - op(X2) op(Y2)), m x m, op = transpose where the
// flag says so; register blocks of 2 x 4 outputs, operands from L2.
template <class TO, class TA1, class TB1, class TA2, class TB2, bool T1A, bool T1B, bool T2A, bool T2B>
__device__ void prod2(int m, TO* out, TO s, const TA1* X1, const TB1* Y1, const TA2* X2, const TB2* Y2) {
    constexpr int BI = 2, BJ = 4;
    const int nbi = (m + BI - 1) / BI, nbj = (m + BJ - 1) / BJ;
    for (int blk = threadIdx.x; blk < nbi * nbj; blk += blockDim.x) {
        const int i0 = (blk / nbj) * BI, j0 = (blk % nbj) * BJ;
        int ii[BI], jj[BJ];
#pragma unroll
        for (int a = 0; a < BI; ++a) ii[a] = min(i0 + a, m - 1);
#pragma unroll
        for (int b = 0; b < BJ; ++b) jj[b] = min(j0 + b, m - 1);
        TO acc[BI][BJ];
#pragma unroll
        for (int a = 0; a < BI; ++a)
#pragma unroll
            for (int b = 0; b < BJ; ++b) acc[a][b] = TO(0);
        for (int k = 0; k < m; ++k) {
            TO l1[BI], l2[BI], r1[BJ], r2[BJ];
#pragma unroll
            for (int a = 0; a < BI; ++a) {
                l1[a] = (TO)(T1A ? X1[(size_t)k * m + ii[a]] : X1[(size_t)ii[a] * m + k]);
                l2[a] = (TO)(T2A ? X2[(size_t)k * m + ii[a]] : X2[(size_t)ii[a] * m + k]);
            }
#pragma unroll
            for (int b = 0; b < BJ; ++b) {
                r1[b] = (TO)(T1B ? Y1[(size_t)jj[b] * m + k] : Y1[(size_t)k * m + jj[b]]);
                r2[b] = (TO)(T2B ? Y2[(size_t)jj[b] * m + k] : Y2[(size_t)k * m + jj[b]]);
            }
#pragma unroll
            for (int a = 0; a < BI; ++a)
#pragma unroll
                for (int b = 0; b < BJ; ++b) {
                    acc[a][b] = fma(l1[a], r1[b], acc[a][b]);
                    acc[a][b] = fma(-l2[a], r2[b], acc[a][b]);
                }
        }
#pragma unroll
        for (int a = 0; a < BI; ++a)
#pragma unroll
            for (int b = 0; b < BJ; ++b)
                if (i0 + a < m && j0 + b < m) out[(size_t)(i0 + a) * m + j0 + b] = s * acc[a][b];
    }
    __syncthreads();
}

// Exact FP64 value of a permutation vertex: sum_ij (A_{i,s^-1(j)} - B_{s(i),j})^2.
__device__ double vertex_value(int m, const double* A, const double* B, const int* s, const int* inv,
                               const QView& V) {
    double acc = 0.0;
    const int mm = m * m;
    for (int e = threadIdx.x; e < mm; e += blockDim.x) {
        const int i = e / m, j = e - i * m;
        const double d = A[(size_t)i * m + inv[j]] - B[(size_t)s[i] * m + j];
        acc = fma(d, d, acc);
    }
    return block_sum(acc, V.red);
}

__device__ void invert(int m, const int* s, int* inv) {
    for (int i = threadIdx.x; i < m; i += blockDim.x) inv[s[i]] = i;
    __syncthreads();
}

// G += coef * G(D) for D = V - W on the rows S (K of them), V = sv, W = sa
// (sa null: D = the permutation sv on every row, K = m).  iv / ia invert
// sv / sa.  Each thread owns 1 x 4 blocks of G; the rank-K terms read four
// LDS tiles per chunk of 32 rows of S (row stride m rounded up to 4, so a
// block's four columns are one 16-byte read).
typedef __attribute__((address_space(3))) float lds_float;
typedef __attribute__((address_space(3))) int lds_int;
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) v4f lds_v4f;

__device__ void grad_update(const QapParams& Q, float* G, double coef, const int* S, int K, const int* sv,
                            const int* sa, const int* iv, const int* ia, const QView& V) {
    const int m = Q.m, mp = (m + 3) & ~3, nb = mp / 4;
    const float cf = (float)coef;
    lds_float* acol = (lds_float*)V.cost;        // [kQapChunk][mp]: A[k][x]
    lds_float* arow = acol + kQapChunk * mp;     //                  A[x][k]
    lds_float* db = arow + kQapChunk * mp;       //                  B[sv(k)][x] - B[sa(k)][x]
    lds_float* dbt = db + kQapChunk * mp;        //                  B[x][sv(k)] - B[x][sa(k)]
    lds_int* isv = (lds_int*)(dbt + kQapChunk * mp);
    lds_int* isa = isv + m;
    lds_int* iiv = isa + m;
    lds_int* iia = iiv + m;
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        isv[i] = sv[i];
        iiv[i] = iv[i];
        isa[i] = sa ? sa[i] : -1;
        iia[i] = sa ? ia[i] : -1;
    }
    for (int c0 = 0; c0 < K; c0 += kQapChunk) {
        const int cnt = min(kQapChunk, K - c0);
        __syncthreads();
        for (int e = threadIdx.x; e < cnt * mp; e += blockDim.x) {
            const int r = e / mp, x = e - r * mp;
            float a1 = 0.0f, a2 = 0.0f, d1 = 0.0f, d2 = 0.0f;
            if (x < m) {
                const int k = S[c0 + r];
                const int a = sv[k];
                a1 = Q.A[(size_t)k * m + x];
                a2 = Q.A[(size_t)x * m + k];
                d1 = Q.B[(size_t)a * m + x];
                d2 = Q.B[(size_t)x * m + a];
                if (sa) {
                    d1 -= Q.B[(size_t)sa[k] * m + x];
                    d2 -= Q.B[(size_t)x * m + sa[k]];
                }
            }
            acol[e] = a1;
            arow[e] = a2;
            db[e] = d1;
            dbt[e] = d2;
        }
        __syncthreads();
        for (int b = threadIdx.x; b < m * nb; b += blockDim.x) {
            const int i = b / nb, j0 = (b - i * nb) * 4;
            float t[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            if (c0 == 0) {
                // A'A D and D B B' (one entry each)
                const bool rowS = !sa || isv[i] != isa[i];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int j = j0 + q;
                    if (j >= m) continue;
                    const int kv = iiv[j];
                    if (!sa || isa[kv] != j) t[q] += Q.AtA[(size_t)i * m + kv];
                    if (sa) {
                        const int ka = iia[j];
                        if (isv[ka] != j) t[q] -= Q.AtA[(size_t)i * m + ka];
                    }
                    if (rowS) {
                        t[q] += Q.BBt[(size_t)isv[i] * m + j];
                        if (sa) t[q] -= Q.BBt[(size_t)isa[i] * m + j];
                    }
                }
            }
            // - A'D B - A D B' over this chunk of S
            float u[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            for (int r = 0; r < cnt; ++r) {
                const float ac = acol[r * mp + i], ar = arow[r * mp + i];
                const v4f d = *(const lds_v4f*)(db + r * mp + j0);
                const v4f dt = *(const lds_v4f*)(dbt + r * mp + j0);
                u[0] = fmaf(ar, dt.x, fmaf(ac, d.x, u[0]));
                u[1] = fmaf(ar, dt.y, fmaf(ac, d.y, u[1]));
                u[2] = fmaf(ar, dt.z, fmaf(ac, d.z, u[2]));
                u[3] = fmaf(ar, dt.w, fmaf(ac, d.w, u[3]));
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (j0 + q < m) {
                    float* g = G + (size_t)i * m + j0 + q;
                    *g = fmaf(cf, t[q] - u[q], *g);
                }
        }
    }
    __syncthreads();
}

// P, G of a node; the kernel's view of its workspace.
struct QNode {
    float *P, *G, *C;
    double *R64, *G64;
    unsigned char* mask;
    int *sv, *sa, *iv, *ia, *S;
    double *vF, *vA;
    int *pF, *pA;
};

__device__ QNode qnode(const QapParams& Q, int node) {
    unsigned char* b = Q.work + (size_t)node * Q.node_bytes;
    QNode q;
    q.P = reinterpret_cast<float*>(b + Q.NL.P);
    q.G = reinterpret_cast<float*>(b + Q.NL.G);
    q.C = reinterpret_cast<float*>(b + Q.NL.C);
    q.R64 = reinterpret_cast<double*>(b + Q.NL.R64);
    q.G64 = reinterpret_cast<double*>(b + Q.NL.G64);
    q.mask = b + Q.NL.mask;
    q.sv = reinterpret_cast<int*>(b + Q.NL.sv);
    q.sa = reinterpret_cast<int*>(b + Q.NL.sa);
    q.iv = reinterpret_cast<int*>(b + Q.NL.iv);
    q.ia = reinterpret_cast<int*>(b + Q.NL.ia);
    q.S = reinterpret_cast<int*>(b + Q.NL.S);
    q.vF = reinterpret_cast<double*>(b + Q.NL.vF);
    q.pF = reinterpret_cast<int*>(b + Q.NL.pF);
    q.vA = reinterpret_cast<double*>(b + Q.NL.vA);
    q.pA = reinterpret_cast<int*>(b + Q.NL.pA);
    return q;
}

// Gradient from P: R = A P - P B, G = 2 (A'R - R B') in FP32 (R in q.C).
__device__ void refresh(const QapParams& Q, const QNode& q) {
    const int m = Q.m;
    prod2<float, float, float, float, float, false, false, false, false>(m, q.C, 1.0f, Q.A, q.P, q.P, Q.B);
    prod2<float, float, float, float, float, true, false, false, true>(m, q.G, 2.0f, Q.A, q.C, q.C, Q.B);
}

template <bool kStaged>
__device__ float* cost_area(const QView& V, const QNode& q) {
    return kStaged ? V.cost : q.C;
}

template <bool kStaged>
__global__ void __launch_bounds__(kQapThreads) qap32_kernel(QapParams Q) {
    extern __shared__ __align__(16) unsigned char raw[];
    const QView V = qview(raw, Q.LL);
    const int node = blockIdx.x;
    if (node >= Q.N) return;
    const int m = Q.m, n = m * m, tid = threadIdx.x, T = blockDim.x;
    const QNode q = qnode(Q, node);
    const double* lb = Q.lb + (size_t)node * n;
    const double* ub = Q.ub + (size_t)node * n;
    double* x = Q.x + (size_t)node * n;
    double* best_v = Q.best_v + (size_t)node * n;
    const double eps = Q.eps[node];
    const int initial = Q.iters[node];
    const bool supplied = (Q.status[node] & GB_NODE_INITIAL_POINT_SUPPLIED) != 0;
    float* C = cost_area<kStaged>(V, q);

    // Allowed edges: ub >= 1/2, and a forced edge (lb > 1/2) owns its row and
    // column.  q.iv / q.ia hold the forced column of a row / row of a column.
    for (int i = tid; i < m; i += T) {
        q.iv[i] = -1;
        q.ia[i] = -1;
    }
    __syncthreads();
    bool bad = false;
    for (int e = tid; e < n; e += T) {
        if (!(lb[e] <= ub[e] + 1e-9)) bad = true;
        if (lb[e] > 0.5) {
            const int i = e / m, j = e - i * m;
            if (ub[e] < 0.5 || atomicCAS(&q.iv[i], -1, j) != -1 || atomicCAS(&q.ia[j], -1, i) != -1) bad = true;
        }
    }
    __syncthreads();
    for (int e = tid; e < n; e += T) {
        const int i = e / m, j = e - i * m;
        const int fc = q.iv[i], fr = q.ia[j];
        q.mask[e] = (ub[e] >= 0.5 && (fc < 0 || fc == j) && (fr < 0 || fr == i)) ? 1 : 0;
    }
    for (int j = tid; j <= m; j += T) {
        q.vF[j] = 0.0;
        q.vA[j] = 0.0;
        q.pF[j] = 0;
        q.pA[j] = 0;
    }
    __syncthreads();
    auto infeasible = [&]() {
        if (tid == 0) {
            Q.status[node] = GB_NODE_INFEASIBLE;
            Q.value[node] = kDInf;
            Q.gap[node] = 0.0;
            Q.best_val[node] = kNone;
            Q.iters[node] = initial;
        }
    };
    if (block_or(bad, V.ired)) {
        infeasible();
        return;
    }

    double best_s = kNone;
    auto track = [&](const int* s, double val) {
        if (val < best_s) {
            best_s = val;
            for (int e = tid; e < n; e += T) best_v[e] = (s[e / m] == e % m) ? 1.0 : 0.0;
        }
        __syncthreads();
    };

    // Start: the supplied iterate when it is doubly stochastic on the allowed
    // edges, else the vertex of the reference's artificial direction.
    bool from_x = false;
    if (supplied || initial > 0) {
        bool okx = true;
        for (int e = tid; e < n; e += T) {
            const float p = (float)x[e];
            q.P[e] = p;
            if (!(p >= 0.0f) || (!q.mask[e] && p != 0.0f)) okx = false;
        }
        __syncthreads();
        for (int i = tid; i < 2 * m; i += T) {
            double s = 0.0;
            for (int k = 0; k < m; ++k) s += i < m ? (double)q.P[(size_t)i * m + k] : (double)q.P[(size_t)k * m + (i - m)];
            if (fabs(s - 1.0) > 1e-4) okx = false;
        }
        from_x = block_and(okx, V.ired);
    }
    if (!from_x) {
        for (int e = tid; e < n; e += T) C[e] = q.mask[e] ? (float)(-1.0 - (double)e / (n + 1.0)) : kQInf;
        __syncthreads();
        if (!assign<kStaged>(m, C, q.vF, q.pF, q.sv, V)) {
            infeasible();
            return;
        }
        for (int e = tid; e < n; e += T) q.P[e] = (q.sv[e / m] == e % m) ? 1.0f : 0.0f;
        for (int e = tid; e < n; e += T) q.G[e] = 0.0f;
        // the artificial direction's potentials are no warm start for G
        for (int j = tid; j <= m; j += T) {
            q.vF[j] = 0.0;
            q.pF[j] = 0;
        }
        __syncthreads();
        invert(m, q.sv, q.iv);
        track(q.sv, vertex_value(m, Q.A64, Q.B64, q.sv, q.iv, V));
        for (int i = tid; i < m; i += T) q.S[i] = i;
        __syncthreads();
        grad_update(Q, q.G, 2.0, q.S, m, q.sv, nullptr, q.iv, nullptr, V);
    } else {
        refresh(Q, q);
    }

    int iterations = initial, since = 0;
    for (int iter = 0; iter < Q.max_iters; ++iter) {
        iterations = initial + iter + 1;
        // forward vertex: assignment on G
        for (int e = tid; e < n; e += T) C[e] = q.mask[e] ? q.G[e] : kQInf;
        __syncthreads();
        if (!assign<kStaged>(m, C, q.vF, q.pF, q.sv, V)) break;
        invert(m, q.sv, q.iv);
        track(q.sv, vertex_value(m, Q.A64, Q.B64, q.sv, q.iv, V));
        if (best_s <= Q.stop_value) break;
        double ga = 0.0;
        for (int e = tid; e < n; e += T) ga = fma((double)q.G[e], (double)q.P[e], ga);
        for (int i = tid; i < m; i += T) ga -= (double)q.G[(size_t)i * m + q.sv[i]];
        const double gap = block_sum(ga, V.red);
        if (gap >= 0.0 && gap < eps) break;
        // away vertex: assignment on -G over the support of P
        for (int e = tid; e < n; e += T) C[e] = (q.mask[e] && q.P[e] > 1e-7f) ? -q.G[e] : kQInf;
        __syncthreads();
        if (!assign<kStaged>(m, C, q.vA, q.pA, q.sa, V)) break;
        invert(m, q.sa, q.ia);
        track(q.sa, vertex_value(m, Q.A64, Q.B64, q.sa, q.ia, V));
        // rows where the vertices differ (wave ballots, prefix in LDS)
        int K = 0;
        {
            const int w = wave_id();
            const bool d = tid < m && q.sv[tid] != q.sa[tid];
            const unsigned long long b = __ballot(d);
            if (lane_id() == 0) V.ired[w] = __popcll(b);
            __syncthreads();
            int off = 0;
            for (int u = 0; u < w; ++u) off += V.ired[u];
            for (int u = 0; u < wave_count(); ++u) K += V.ired[u];
            if (d) q.S[off + __popcll(b & ((1ull << lane_id()) - 1ull))] = tid;
            __syncthreads();
        }
        if (K == 0) break;
        // pairwise gap, largest step, curvature 2 ||A D - D B||^2
        double pg = 0.0, gm = kDInf, cv = 0.0;
        for (int r = tid; r < K; r += T) {
            const int k = q.S[r];
            pg += (double)q.G[(size_t)k * m + q.sa[k]] - (double)q.G[(size_t)k * m + q.sv[k]];
            gm = fmin(gm, (double)q.P[(size_t)k * m + q.sa[k]]);
        }
        for (int e = tid; e < n; e += T) {
            const int i = e / m, j = e - i * m;
            float r = 0.0f;
            const int kv = q.iv[j], ka = q.ia[j];
            if (q.sa[kv] != j) r += Q.A[(size_t)i * m + kv];
            if (q.sv[ka] != j) r -= Q.A[(size_t)i * m + ka];
            if (q.sv[i] != q.sa[i]) r -= Q.B[(size_t)q.sv[i] * m + j] - Q.B[(size_t)q.sa[i] * m + j];
            cv = fma((double)r, (double)r, cv);
        }
        const double pgap = block_sum(pg, V.red);
        const double gmax = fmin(block_min(gm, V.red), 1.0);
        const double curv = 2.0 * block_sum(cv, V.red);
        double gamma = curv > 0.0 ? fmin(pgap / curv, gmax) : (pgap > 0.0 ? gmax : 0.0);
        if (!(gamma > 0.0)) break;
        const float g32 = (float)gamma;
        for (int r = tid; r < K; r += T) {
            const int k = q.S[r];
            float* pr = q.P + (size_t)k * m;
            pr[q.sv[k]] += g32;
            const float left = pr[q.sa[k]] - g32;
            pr[q.sa[k]] = left > 0.0f ? left : 0.0f;
        }
        __syncthreads();
        grad_update(Q, q.G, 2.0 * gamma, q.S, K, q.sv, q.sa, q.iv, q.ia, V);
        if (++since >= Q.refresh) {
            refresh(Q, q);
            since = 0;
        }
    }

    // FP64 certificate at P.
    prod2<double, double, float, float, double, false, false, false, false>(m, q.R64, 1.0, Q.A64, q.P, q.P,
                                                                            Q.B64);
    double fv = 0.0;
    for (int e = tid; e < n; e += T) fv = fma(q.R64[e], q.R64[e], fv);
    const double fval = block_sum(fv, V.red);
    prod2<double, double, double, double, double, true, false, false, true>(m, q.G64, 2.0, Q.A64, q.R64, q.R64,
                                                                            Q.B64);
    for (int e = tid; e < n; e += T) C[e] = q.mask[e] ? (float)q.G64[e] : kQInf;
    __syncthreads();
    double lbnd = -kDInf;
    if (assign<kStaged>(m, C, q.vF, q.pF, q.sv, V)) {
        invert(m, q.sv, q.iv);
        track(q.sv, vertex_value(m, Q.A64, Q.B64, q.sv, q.iv, V));
        lbnd = lagrange_bound(m, q.G64, q.mask, q.vF, V);
    }
    double gp = 0.0;
    for (int e = tid; e < n; e += T) gp = fma(q.G64[e], (double)q.P[e], gp);
    const double fgap = block_sum(gp, V.red) - lbnd;
    uint8_t st = (fgap >= 0.0 && fgap < eps) ? 1 : 0;

    // rounded candidate: a permutation on the allowed edges
    if (Q.flags) {
        bool perm = true;
        for (int i = tid; i < m; i += T) {
            int c = -1, cnt = 0;
            for (int j = 0; j < m; ++j)
                if (q.P[(size_t)i * m + j] > 0.5f) {
                    c = j;
                    ++cnt;
                }
            if (cnt != 1 || !q.mask[(size_t)i * m + c]) perm = false;
            q.sa[i] = c;
        }
        perm = block_and(perm, V.ired);
        if (perm) {
            for (int j = tid; j < m; j += T) q.ia[j] = -1;
            __syncthreads();
            for (int i = tid; i < m; i += T) q.ia[q.sa[i]] = i;
            __syncthreads();
            bool full = true;
            for (int j = tid; j < m; j += T) full = full && q.ia[j] >= 0;
            if (block_and(full, V.ired)) track(q.sa, vertex_value(m, Q.A64, Q.B64, q.sa, q.ia, V));
        }
    }
    bool integral = true;
    for (int e = tid; e < n; e += T) {
        const double p = (double)q.P[e];
        if (Q.flags && Q.flags[e] && fabs(p - rint(p)) > Q.ctol) integral = false;
    }
    if (block_and(integral, V.ired)) st = st ? GB_NODE_INTEGRAL_CONVERGED : GB_NODE_INTEGRAL_UNCONVERGED;
    for (int e = tid; e < n; e += T) x[e] = (double)q.P[e];
    if (tid == 0) {
        Q.value[node] = fval;
        Q.gap[node] = fgap;
        Q.status[node] = st;
        Q.iters[node] = iterations;
        Q.best_val[node] = best_s;
    }
}

// FP32 copies and A'A, B B' (FP64 sums) of the instance.
__global__ void qap32_prepare_kernel(int m, const double* A, const double* B, float* A32, float* B32, float* AtA,
                                     float* BBt) {
    const int mm = m * m;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < mm; e += gridDim.x * blockDim.x) {
        const int i = e / m, j = e - i * m;
        A32[e] = (float)A[e];
        B32[e] = (float)B[e];
        double a = 0.0, b = 0.0;
        for (int k = 0; k < m; ++k) {
            a = fma(A[(size_t)k * m + i], A[(size_t)k * m + j], a);
            b = fma(B[(size_t)i * m + k], B[(size_t)j * m + k], b);
        }
        AtA[e] = (float)a;
        BBt[e] = (float)b;
    }
}

// T assignment problems in sequence, warm-started (one workgroup).
__global__ void __launch_bounds__(kQapThreads) qap32_assign_kernel(int m, int T, const float* costs, int* perms,
                                                                  double* bounds, double* vstate, int* pstate,
                                                                  QapLds L, double* gwork) {
    extern __shared__ __align__(16) unsigned char raw[];
    const QView V = qview(raw, L);
    for (int j = threadIdx.x; j <= m; j += blockDim.x) {
        vstate[j] = 0.0;
        pstate[j] = 0;
    }
    __syncthreads();
    const int mm = m * m;
    for (int t = 0; t < T; ++t) {
        const float* src = costs + (size_t)t * mm;
        float* C = L.staged ? V.cost : reinterpret_cast<float*>(gwork);
        for (int e = threadIdx.x; e < mm; e += blockDim.x) C[e] = src[e];
        __syncthreads();
        int* out = perms + (size_t)t * m;
        const bool ok = L.staged ? assign<true>(m, C, vstate, pstate, out, V)
                                 : assign<false>(m, C, vstate, pstate, out, V);
        double part = 0.0;
        {
            const int lane = lane_id(), w = wave_id(), W = wave_count();
            for (int j = threadIdx.x; j <= m; j += blockDim.x) V.vsh[j] = vstate[j];
            __syncthreads();
            for (int i = w; i < m; i += W) {
                double r = kDInf;
                for (int j = lane; j < m; j += kWave)
                    if (src[(size_t)i * m + j] < kQInf) r = fmin(r, (double)src[(size_t)i * m + j] - V.vsh[j + 1]);
                r = wave_min_all(r);
                if (lane == 0) part += r;
            }
            for (int j = threadIdx.x; j < m; j += blockDim.x) part += V.vsh[j + 1];
        }
        const double b = block_sum(part, V.red);
        if (threadIdx.x == 0) bounds[t] = ok ? b : -kDInf;
        if (!ok)
            for (int i = threadIdx.x; i < m; i += blockDim.x) out[i] = -1;
        __syncthreads();
    }
}

}  // namespace

int launch_qap32_prepare(int m, const double* A, const double* B, float* A32, float* B32, float* AtA, float* BBt,
                         void* stream) {
    const int mm = m * m;
    hipLaunchKernelGGL(qap32_prepare_kernel, dim3((mm + 255) / 256), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), m, A, B, A32, B32, AtA, BBt);
    return (int)hipGetLastError();
}

int launch_qap32_relax(const QapParams& p, void* stream) {
    const void* k = p.LL.staged ? reinterpret_cast<const void*>(&qap32_kernel<true>)
                                : reinterpret_cast<const void*>(&qap32_kernel<false>);
    hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p.LL.total);
    if (e != hipSuccess) return (int)e;
    QapParams pp = p;
    void* args[] = {&pp};
    return (int)hipLaunchKernel(k, dim3(p.N), dim3(kQapThreads), args, p.LL.total,
                                reinterpret_cast<hipStream_t>(stream));
}

int launch_qap32_assignments(int m, int T, const float* costs, int* perms, double* bounds, double* vstate,
                             int* pstate, size_t lds_limit, void* stream) {
    QapLds L = qap_lds(m, lds_limit);
    double* gwork = nullptr;
    hipError_t e = hipMalloc(&gwork, (size_t)m * m * 4 + 256);
    if (e != hipSuccess) return (int)e;
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(&qap32_assign_kernel),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.total);
    if (e == hipSuccess) {
        void* args[] = {&m, &T, &costs, &perms, &bounds, &vstate, &pstate, &L, &gwork};
        e = hipLaunchKernel(reinterpret_cast<const void*>(&qap32_assign_kernel), dim3(1), dim3(kQapThreads), args,
                            L.total, reinterpret_cast<hipStream_t>(stream));
    }
    if (e == hipSuccess) e = hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream));
    (void)hipFree(gwork);
    return (int)e;
}

}  // namespace gb
