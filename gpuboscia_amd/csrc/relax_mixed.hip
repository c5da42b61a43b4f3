// Mixed-precision batched relaxation of dense quadratics over 0/1 boxes,
// gfx950 (the precision form of the batched B&B: BF16 MFMA products, FP32
// iterates, FP64 steps and reductions; the bound that decides pruning and the
// incumbent value are certified afterwards in FP64 by the batched FP64
// kernel, which continues from the iterate written here).
//
// Layout: one wave64 carries 16 node columns in lockstep.  Lane l serves
// column c = l % 16 and owns rows 16 t + 4 (l / 16) + r (t < RT, r < 4) of
// that column's x, Q x (registers, FP32).  Each pass computes Q v and Q a
// for all 16 columns with v_mfma_f32_16x16x32_bf16: the A operand is Q
// (pre-split into 1..3 bf16 planes q = q0 + q1 + q2, FP32 accumulation over
// the planes), the B operand the vertices v, a.  Box vertices of a 0/1 box
// are 0/1 and exact in bf16, so the only rounding of the product is the
// plane split of Q (about 2^-8 relative for one plane, 2^-24 for three) and
// the FP32 accumulation.  The B-operand k-slots of lane l for k-step s are
// the rows the lane owns in tiles 2s and 2s+1, so vertices go from the
// lane's bit masks straight into the MFMA operand.
//
// Per column the pass is one DICG (or FW) iteration of the batched form:
// v = argmin over the box of g'v, the in-face away vertex a (face tolerance
// atol 1e-6 + rtol 1e-4, lmo.cuh:278-295), gap = g'(x - v), the exact
// quadratic step clip(-g'd / (s_g d'Qd), 0, gamma_max)
// (line_search_options.h:48-63), x += gamma d, Qx += gamma Q d.  Every
// vertex is an integral candidate (drs_integer_value,
// device_resident_solver.cuh:297-313); its value is tracked approximately
// and re-scored in FP64 by the certifying kernel.  A column stops at
// gap < max(eps, rel_floor (1 + |f|)), at f - gap >= cutoff, or at the
// iteration budget, writes its iterate and takes the next node from the
// device counter; the wave exits when the counter is exhausted and all of
// its columns are idle.  A fresh node starts at the box vertex nearest to
// its supplied point (the upper bounds when none is supplied, as the plain
// form's init vertex); nodes with bounds other than 0/1 and continuations
// are left to the FP64 kernel.
#include "gb_device.h"
#include "gb_mixed.h"

namespace gb {
namespace {

typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short bf16_rne(float f) {
    unsigned int u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<unsigned short>(u >> 16);
}

__device__ __forceinline__ float bf16_value(unsigned short h) {
    return __uint_as_float(static_cast<unsigned int>(h) << 16);
}

// Plane p of q: b_p = bf16(q - b_0 - ... - b_{p-1}), residuals in FP64.
__device__ unsigned short bf16_plane(double q, int p) {
    double r = q;
    unsigned short b = 0;
    for (int i = 0; i <= p; ++i) {
        b = bf16_rne(static_cast<float>(r));
        r -= static_cast<double>(bf16_value(b));
    }
    return b;
}

// 8 bits -> 8 bf16 values 0.0 / 1.0 (0x3F80), two per dword.
__device__ __forceinline__ bf16x8 expand_bits(unsigned int bits) {
    uint4 w;
    w.x = ((bits & 1u) * 0x3F80u) | (((bits >> 1) & 1u) * 0x3F800000u);
    w.y = (((bits >> 2) & 1u) * 0x3F80u) | (((bits >> 3) & 1u) * 0x3F800000u);
    w.z = (((bits >> 4) & 1u) * 0x3F80u) | (((bits >> 5) & 1u) * 0x3F800000u);
    w.w = (((bits >> 6) & 1u) * 0x3F80u) | (((bits >> 7) & 1u) * 0x3F800000u);
    return __builtin_bit_cast(bf16x8, w);
}

__device__ __forceinline__ bf16x8 load_frag(const uint4* p) {
    return __builtin_bit_cast(bf16x8, *p);
}

__device__ __forceinline__ double quad_exact_step(double gap, double curvature, double maximum) {
    if (maximum <= 0.0) return 0.0;
    if (curvature > 0.0) {
        if (gap <= 0.0) return 0.0;
        const double step = gap / curvature;
        return step < maximum ? step : maximum;
    }
    return -gap + 0.5 * curvature * maximum < 0.0 ? maximum : 0.0;
}

// Sum / min over the 4 lanes of a column (lanes c, c+16, c+32, c+48); every
// lane of the column receives the same value.
__device__ __forceinline__ double col_sum(double v) {
    v = v + __shfl_xor(v, 16, kWave);
    return v + __shfl_xor(v, 32, kWave);
}
__device__ __forceinline__ double col_min(double v) {
    v = fmin(v, __shfl_xor(v, 16, kWave));
    return fmin(v, __shfl_xor(v, 32, kWave));
}
__device__ __forceinline__ bool col_and(bool p) {
    int v = p ? 1 : 0;
    v &= __shfl_xor(v, 16, kWave);
    v &= __shfl_xor(v, 32, kWave);
    return v != 0;
}

__global__ void mixed_prepare_kernel(const double* Qt, const double* lin, double gs, double ls, int n,
                                     int rt, int planes, uint4* frag, float* lsc) {
    const int S = rt / 2;
    const size_t total = static_cast<size_t>(planes) * rt * S * kWave;
    const size_t idx = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx < static_cast<size_t>(rt) * 16) {
        const int j = static_cast<int>(idx);
        lsc[j] = (j < n && lin) ? static_cast<float>((gs * ls) * lin[j]) : 0.0f;
    }
    if (idx >= total) return;
    const int lane = static_cast<int>(idx % kWave);
    size_t rest = idx / kWave;
    const int s = static_cast<int>(rest % S);
    rest /= S;
    const int t = static_cast<int>(rest % rt);
    const int p = static_cast<int>(rest / rt);
    const int row = 16 * t + (lane & 15);
    const int q = lane >> 4;
    unsigned short h[8];
    for (int i = 0; i < 8; ++i) {
        const int k = 16 * (2 * s + (i >> 2)) + 4 * q + (i & 3);
        const double v = (row < n && k < n) ? Qt[static_cast<size_t>(k) * n + row] : 0.0;
        h[i] = bf16_plane(v, p);
    }
    uint4 w;
    w.x = h[0] | (static_cast<unsigned int>(h[1]) << 16);
    w.y = h[2] | (static_cast<unsigned int>(h[3]) << 16);
    w.z = h[4] | (static_cast<unsigned int>(h[5]) << 16);
    w.w = h[6] | (static_cast<unsigned int>(h[7]) << 16);
    frag[idx] = w;
}

constexpr int kPhaseDone = 0, kPhaseInit = 1, kPhaseActive = 2, kPhaseFetch = 3;
constexpr int kRed = 7;  // per-column partials exchanged between the waves of a tile

template <int RW, int PL, bool DICG>
__global__ void __launch_bounds__(RW == 1 ? 512 : 64 * ((16 + RW - 1) / RW)) relax_mixed_kernel(MixedParams P) {
    constexpr int EW = 4 * RW;  // elements per lane: rows 16 t + 4 q + r of the wave's tiles
    extern __shared__ __align__(16) unsigned char mixed_smem[];
    const int RT = P.rt;
    const int S = RT / 2;
    const int W = blockDim.x / kWave;
    const MixedLds L = mixed_lds_layout(RT, W);
    float* lsc_lds = reinterpret_cast<float*>(mixed_smem + L.lsc);
    uint8_t* bits = mixed_smem + L.bits;
    double* red = reinterpret_cast<double*>(mixed_smem + L.red);
    int* fetch = reinterpret_cast<int*>(mixed_smem + L.fetch);
    int* okw = reinterpret_cast<int*>(mixed_smem + L.ok);
    for (int i = threadIdx.x; i < 16 * RT; i += blockDim.x) lsc_lds[i] = P.lsc[i];
    __syncthreads();
    const int lane = lane_id();
    const int w = wave_id();
    const int c = lane & 15;
    const int q = lane >> 4;
    const int n = P.n;
    const float gs = P.gs;
    const float hs = 0.5f * P.gs;
    const size_t tstride = static_cast<size_t>(S) * kWave;
    const uint4* fbase = P.frag + lane;

    // Wave w owns row tiles t = w + W j (j < RW); element e = 4 j + r is row
    // 16 t + 4 q + r of column c.
    float x[EW], qx[EW];
#pragma unroll
    for (int e = 0; e < EW; ++e) {
        x[e] = 0.0f;
        qx[e] = 0.0f;
    }
    // The wave's Q fragments are the same in every pass: where they fit
    // beside the rest of the state (32..128 VGPRs for n <= 256) they are
    // loaded once and kept in registers for the kernel's lifetime.
    constexpr bool kCache = (PL == 1 && RW != 3) || (PL == 3 && RW == 1);
    constexpr int kSMax = 8;
    bf16x8 fc[kCache ? PL : 1][kCache ? RW : 1][kCache ? kSMax : 1];
    if (kCache) {
#pragma unroll
        for (int p = 0; p < PL; ++p)
#pragma unroll
            for (int j = 0; j < RW; ++j)
#pragma unroll
                for (int s = 0; s < kSMax; ++s) {
                    const int t = w + W * j;
                    if (t < RT && s < S)
                        fc[p][j][s] = load_frag(fbase + (static_cast<size_t>(p) * RT + t) * tstride +
                                                static_cast<size_t>(s) * kWave);
                }
    }
    uint32_t lbm = 0, ubm = 0, x0m = 0, bestm = 0;
    int phase = c < P.cols ? kPhaseFetch : kPhaseDone;
    int node = -1, it = 0, initial = 0;
    double eps = 0.0, cutoff = 0.0, best = kNone;

    while (true) {
        // Columns without a node take the next ones from the counter (wave 0
        // fetches, the tile's waves share the result); nodes outside the
        // 0/1-box scope and continuations are left to the FP64 kernel.  The
        // column state is replicated in every wave, so all waves take the
        // same branches and meet at the same barriers.
        while (__ballot(phase == kPhaseFetch) != 0ull) {
            if (w == 0 && q == 0 && phase == kPhaseFetch) fetch[c] = atomicAdd(P.counter, 1);
            __syncthreads();
            if (phase == kPhaseFetch) {
                node = fetch[c];
                if (node >= P.N) phase = kPhaseDone;
            }
            const bool loading = phase == kPhaseFetch;
            bool ok = true;
            uint32_t lm = 0, um = 0, xm = 0;
            int init_it = 0;
            if (loading) {
                const double* lbn = P.lb + static_cast<size_t>(node) * n;
                const double* ubn = P.ub + static_cast<size_t>(node) * n;
                const double* xn = P.x + static_cast<size_t>(node) * n;
                init_it = P.iters[node];
                const bool from_x = init_it != 0 || (P.status[node] & GB_NODE_INITIAL_POINT_SUPPLIED) != 0;
#pragma unroll 1
                for (int j = 0; j < RW; ++j) {
                    const int t = w + W * j;
#pragma unroll 1
                    for (int r = 0; r < 4; ++r) {
                        const int row = 16 * t + 4 * q + r;
                        const int e = 4 * j + r;
                        if (t < RT && row < n) {
                            const double l = lbn[row], u = ubn[row];
                            if (!(l <= u) || !(l == 0.0 || l == 1.0) || !(u == 0.0 || u == 1.0)) ok = false;
                            const double xv = from_x ? xn[row] : u;
                            const bool xb = (xv >= 0.5 * (l + u)) ? (u == 1.0) : (l == 1.0);
                            lm |= static_cast<uint32_t>(l == 1.0) << e;
                            um |= static_cast<uint32_t>(u == 1.0) << e;
                            xm |= static_cast<uint32_t>(xb) << e;
                        }
                    }
                }
            }
            ok = col_and(ok && init_it == 0);
            if (q == 0) okw[w * 16 + c] = ok ? 1 : 0;
            __syncthreads();
            bool all_ok = true;
            for (int v = 0; v < W; ++v) all_ok = all_ok && okw[v * 16 + c] != 0;
            if (loading && all_ok) {
                phase = kPhaseInit;
                lbm = lm;
                ubm = um;
                x0m = xm;
                it = 0;
                initial = init_it;
                eps = P.eps[node];
                cutoff = P.cutoff ? P.cutoff[node] : __builtin_inf();
                best = kNone;
                bestm = 0;
            }
            __syncthreads();
        }
        if (__ballot(phase != kPhaseDone) == 0ull) break;
        const bool act = phase == kPhaseActive;
        const bool ini = phase == kPhaseInit;

        // Vertices of the wave's rows, partial sums at x, vertex nibbles to LDS.
        uint32_t vm = 0, am = 0;
        float gp = 0.0f, fp = 0.0f;
#pragma unroll
        for (int e = 0; e < EW; ++e) {
            const int t = w + W * (e >> 2);
            const int row = 16 * t + 4 * q + (e & 3);
            const float ls = t < RT ? lsc_lds[row] : 0.0f;
            const bool lb1 = (lbm >> e) & 1u;
            const bool ub1 = (ubm >> e) & 1u;
            const float g = gs * qx[e] + ls;
            const bool vb = g > 0.0f ? lb1 : ub1;
            bool ab = false;
            if (DICG) {
                const float uf = ub1 ? 1.0f : 0.0f, lf = lb1 ? 1.0f : 0.0f;
                if (fabsf(x[e] - uf) <= 1e-6f + 1e-4f * uf) ab = ub1;
                else if (fabsf(x[e] - lf) <= 1e-6f + 1e-4f * lf) ab = lb1;
                else ab = (-g > 0.0f) ? lb1 : ub1;
            }
            vm |= static_cast<uint32_t>(vb) << e;
            am |= static_cast<uint32_t>(ab) << e;
            gp += g * (x[e] - (vb ? 1.0f : 0.0f));
            fp += x[e] * (hs * qx[e] + ls);
        }
        if (ini) {
            vm = x0m;
            am = 0;
        } else if (!act) {
            vm = 0;
            am = 0;
        }
#pragma unroll
        for (int j = 0; j < RW; ++j) {
            const int t = w + W * j;
            if (t < RT) bits[t * kWave + lane] = static_cast<uint8_t>(((vm >> (4 * j)) & 0xFu) | (((am >> (4 * j)) & 0xFu) << 4));
        }
        __syncthreads();

        // Q [v a] for the wave's row tiles over all k-steps; the next
        // k-step's fragments are loaded while the current one multiplies.
        f32x4 accv[RW], acca[RW];
#pragma unroll
        for (int j = 0; j < RW; ++j) {
            accv[j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            acca[j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        }
        if (kCache) {
#pragma unroll
            for (int s = 0; s < kSMax; ++s) {
                if (s < S) {
                    const unsigned int b0 = bits[(2 * s) * kWave + lane];
                    const unsigned int b1 = bits[(2 * s + 1) * kWave + lane];
                    const bf16x8 bv = expand_bits((b0 & 0xFu) | ((b1 & 0xFu) << 4));
                    const bf16x8 ba = expand_bits((b0 >> 4) | (b1 & 0xF0u));
#pragma unroll
                    for (int p = 0; p < PL; ++p)
#pragma unroll
                        for (int j = 0; j < RW; ++j) {
                            if (w + W * j < RT) {
                                accv[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fc[p][j][s], bv, accv[j], 0, 0, 0);
                                if (DICG)
                                    acca[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fc[p][j][s], ba, acca[j], 0, 0, 0);
                            }
                        }
                }
            }
        } else {
        bf16x8 cur[PL][RW];
#pragma unroll
        for (int p = 0; p < PL; ++p)
#pragma unroll
            for (int j = 0; j < RW; ++j) {
                const int t = w + W * j;
                if (t < RT) cur[p][j] = load_frag(fbase + (static_cast<size_t>(p) * RT + t) * tstride);
            }
        for (int s = 0; s < S; ++s) {
            bf16x8 nxt[PL][RW];
            if (s + 1 < S) {
#pragma unroll
                for (int p = 0; p < PL; ++p)
#pragma unroll
                    for (int j = 0; j < RW; ++j) {
                        const int t = w + W * j;
                        if (t < RT)
                            nxt[p][j] = load_frag(fbase + (static_cast<size_t>(p) * RT + t) * tstride +
                                                  static_cast<size_t>(s + 1) * kWave);
                    }
            }
            const unsigned int b0 = bits[(2 * s) * kWave + lane];
            const unsigned int b1 = bits[(2 * s + 1) * kWave + lane];
            const bf16x8 bv = expand_bits((b0 & 0xFu) | ((b1 & 0xFu) << 4));
            const bf16x8 ba = expand_bits((b0 >> 4) | (b1 & 0xF0u));
#pragma unroll
            for (int p = 0; p < PL; ++p)
#pragma unroll
                for (int j = 0; j < RW; ++j) {
                    const int t = w + W * j;
                    if (t < RT) {
                        accv[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur[p][j], bv, accv[j], 0, 0, 0);
                        if (DICG) acca[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur[p][j], ba, acca[j], 0, 0, 0);
                    }
                }
            if (s + 1 < S) {
#pragma unroll
                for (int p = 0; p < PL; ++p)
#pragma unroll
                    for (int j = 0; j < RW; ++j) cur[p][j] = nxt[p][j];
            }
        }
        }

        // Vertex values and step partials; Q d replaces Q v in accv (Q v for
        // a column that is being initialised).
        float fvp = 0.0f, fap = 0.0f, pgp = 0.0f, cvp = 0.0f, rr = 1.0f;
#pragma unroll
        for (int e = 0; e < EW; ++e) {
            const int t = w + W * (e >> 2);
            const int row = 16 * t + 4 * q + (e & 3);
            const float ls = t < RT ? lsc_lds[row] : 0.0f;
            const float qv = accv[e >> 2][e & 3];
            const float vf = ((vm >> e) & 1u) ? 1.0f : 0.0f;
            fvp += vf * (hs * qv + ls);
            const float g = gs * qx[e] + ls;
            float d, qd;
            if (DICG) {
                const float af = ((am >> e) & 1u) ? 1.0f : 0.0f;
                const float qa = acca[e >> 2][e & 3];
                fap += af * (hs * qa + ls);
                d = vf - af;
                qd = qv - qa;
                // a - v is -1, 0 or 1 on a 0/1 box: the ratio needs no division.
                const float dv = af - vf;
                const float uf = ((ubm >> e) & 1u) ? 1.0f : 0.0f;
                const float lf = ((lbm >> e) & 1u) ? 1.0f : 0.0f;
                if (dv < 0.0f) rr = fminf(rr, uf - x[e]);
                else if (dv > 0.0f) rr = fminf(rr, x[e] - lf);
            } else {
                d = vf - x[e];
                qd = qv - qx[e];
            }
            pgp += -g * d;
            cvp += gs * d * qd;
            accv[e >> 2][e & 3] = ini ? qv : qd;
        }
        // Column totals: 4 lanes of a column, then the tile's waves in order.
        double part[kRed] = {col_sum(gp), col_sum(fp), col_sum(fvp), col_sum(fap),
                             col_sum(pgp), col_sum(cvp), col_min(rr)};
        if (q == 0) {
#pragma unroll
            for (int k = 0; k < kRed; ++k) red[(w * 16 + c) * kRed + k] = part[k];
        }
        __syncthreads();
        double tot[kRed];
#pragma unroll
        for (int k = 0; k < kRed; ++k) tot[k] = red[c * kRed + k];
        for (int v = 1; v < W; ++v) {
#pragma unroll
            for (int k = 0; k < kRed - 1; ++k) tot[k] = tot[k] + red[(v * 16 + c) * kRed + k];
            tot[kRed - 1] = fmin(tot[kRed - 1], red[(v * 16 + c) * kRed + kRed - 1]);
        }
        const double gap = tot[0];
        const double fx = tot[1] + P.kappa;
        const double fv = tot[2] + P.kappa;
        const double fa = DICG ? tot[3] + P.kappa : kNone;
        const double pg = tot[4];
        const double cv = tot[5];
        const double gmax = DICG ? fmax(0.0, fmin(tot[6], 1.0)) : 1.0;

        bool retire = false;
        float gm = 0.0f;
        bool step = false;
        if (act) {
            ++it;
            if (fv < best) {
                best = fv;
                bestm = vm;
            }
            if (DICG && fa < best) {
                best = fa;
                bestm = am;
            }
            const double tol = fmax(eps, static_cast<double>(P.rel_floor) * (1.0 + fabs(fx)));
            const bool stop = (gap >= 0.0 && gap < tol) || (fx - gap >= cutoff);
            if (!stop) {
                gm = static_cast<float>(quad_exact_step(pg, cv, gmax));
                step = true;
            }
            retire = stop || it >= P.max_iters;
        }
        // x += gamma d, Q x += gamma Q d for stepping columns; Q x = Q x0 for
        // columns being initialised.
#pragma unroll
        for (int e = 0; e < EW; ++e) {
            const float qd = accv[e >> 2][e & 3];
            if (ini) {
                qx[e] = qd;
                x[e] = ((x0m >> e) & 1u) ? 1.0f : 0.0f;
            } else if (step) {
                const float vf = ((vm >> e) & 1u) ? 1.0f : 0.0f;
                const float d = DICG ? vf - (((am >> e) & 1u) ? 1.0f : 0.0f) : vf - x[e];
                const float uf = ((ubm >> e) & 1u) ? 1.0f : 0.0f;
                const float lf = ((lbm >> e) & 1u) ? 1.0f : 0.0f;
                x[e] = fminf(fmaxf(fmaf(gm, d, x[e]), lf), uf);
                qx[e] = fmaf(gm, qd, qx[e]);
            }
        }
        if (ini) phase = kPhaseActive;

        if (retire) {
            double* xn = P.x + static_cast<size_t>(node) * n;
            double* bvn = P.best_v + static_cast<size_t>(node) * n;
#pragma unroll
            for (int e = 0; e < EW; ++e) {
                const int t = w + W * (e >> 2);
                const int row = 16 * t + 4 * q + (e & 3);
                if (t < RT && row < n) {
                    xn[row] = static_cast<double>(x[e]);
                    if (best < kNone) bvn[row] = ((bestm >> e) & 1u) ? 1.0 : 0.0;
                }
            }
            if (w == 0 && q == 0) {
                P.iters[node] = initial + it;
                P.status[node] = GB_NODE_INITIAL_POINT_SUPPLIED;
                P.best_val[node] = best;
                if (P.stats) P.stats[node] = it;
            }
            phase = kPhaseFetch;
        }
    }
}

// Y[i] = Q V[i] through the same fragments (diagnostic of the product error).
template <int RT>
__global__ void __launch_bounds__(64) mixed_products_kernel(const uint4* frag, int planes, int n, int N,
                                                            const double* V, double* Y) {
    constexpr int S = RT / 2;
    const int lane = lane_id();
    const int c = lane & 15;
    const int q = lane >> 4;
    const int col = blockIdx.x * 16 + c;
    const size_t tstride = static_cast<size_t>(S) * kWave;
    f32x4 acc[RT];
#pragma unroll
    for (int t = 0; t < RT; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < S; ++s) {
        bf16x8 b;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int row = 16 * (2 * s + (i >> 2)) + 4 * q + (i & 3);
            const double v = (col < N && row < n) ? V[static_cast<size_t>(col) * n + row] : 0.0;
            b[i] = static_cast<short>(bf16_rne(static_cast<float>(v)));
        }
        for (int p = 0; p < planes; ++p) {
            const uint4* fr = frag + (static_cast<size_t>(p) * RT * S + s) * kWave + lane;
#pragma unroll
            for (int t = 0; t < RT; ++t)
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(load_frag(fr + t * tstride), b, acc[t], 0, 0, 0);
        }
    }
    if (col >= N) return;
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * t + 4 * q + r;
            if (row < n) Y[static_cast<size_t>(col) * n + row] = static_cast<double>(acc[t][r]);
        }
}

__global__ void mixed_reset_kernel(int* counter) { *counter = 0; }

}  // namespace

int launch_mixed_prepare(const double* Qt, const double* lin, double gs, double ls, int n, int rt,
                         int planes, uint4* frag, float* lsc, void* stream) {
    const size_t total = static_cast<size_t>(planes) * rt * (rt / 2) * kWave;
    const unsigned blocks = static_cast<unsigned>((total + 255) / 256);
    hipLaunchKernelGGL(mixed_prepare_kernel, dim3(blocks), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), Qt, lin, gs, ls, n, rt, planes, frag, lsc);
    return static_cast<int>(hipGetLastError());
}

int launch_relax_mixed(const MixedParams& p, bool dicg, int rw, int blocks, void* stream) {
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(mixed_reset_kernel, dim3(1), dim3(1), 0, st, p.counter);
    const int waves = (p.rt + rw - 1) / rw;
    const dim3 grid(blocks), block(waves * kWave);
    const size_t lds = mixed_lds_layout(p.rt, waves).total;
#define GB_MIXED_ONE(RW, PL, DI)                                                                \
    do {                                                                                        \
        auto k = relax_mixed_kernel<RW, PL, DI>;                                                \
        if (lds > 64 * 1024) {                                                                  \
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),                \
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
            if (e != hipSuccess) return (int)e;                                                 \
        }                                                                                       \
        hipLaunchKernelGGL(k, grid, block, lds, st, p);                                         \
    } while (0)
#define GB_MIXED(RW)                                                     \
    case RW:                                                             \
        if (p.planes == 1) {                                             \
            if (dicg) GB_MIXED_ONE(RW, 1, true);                         \
            else GB_MIXED_ONE(RW, 1, false);                             \
        } else {                                                         \
            if (dicg) GB_MIXED_ONE(RW, 3, true);                         \
            else GB_MIXED_ONE(RW, 3, false);                             \
        }                                                                \
        break;
    switch (rw) {
        GB_MIXED(1)
        GB_MIXED(2)
        GB_MIXED(3)
        GB_MIXED(4)
        default:
            return static_cast<int>(hipErrorInvalidValue);
    }
#undef GB_MIXED
#undef GB_MIXED_ONE
    return static_cast<int>(hipGetLastError());
}

int launch_mixed_products(const MixedQ& q, const double* V, int N, double* Y, void* stream) {
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const dim3 grid((N + 15) / 16), block(kWave);
#define GB_PROD(RT)                                                                                 \
    case RT:                                                                                        \
        hipLaunchKernelGGL(mixed_products_kernel<RT>, grid, block, 0, st, q.frag, q.planes, q.n, N, V, Y); \
        break;
    switch (q.rt) {
        GB_PROD(2)
        GB_PROD(4)
        GB_PROD(6)
        GB_PROD(8)
        GB_PROD(10)
        GB_PROD(12)
        GB_PROD(14)
        GB_PROD(16)
        default:
            return static_cast<int>(hipErrorInvalidValue);
    }
#undef GB_PROD
    return static_cast<int>(hipGetLastError());
}

}  // namespace gb
