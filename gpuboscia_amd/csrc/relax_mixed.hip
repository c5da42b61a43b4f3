// Mixed-precision batched relaxation of dense quadratics over 0/1 boxes and
// 0/1 bounded simplices, gfx950 (the precision form of the batched B&B:
// BF16 MFMA products, FP32 iterates, partial sums and step; the bound that
// decides pruning and the incumbent value are certified afterwards in FP64
// by the batched FP64 kernel, which continues from the iterate written
// here).
//
// Layout: a workgroup of W waves carries up to 16 node columns in lockstep;
// wave w owns the row tiles t = w + W j (j < RW, t < RT) of every column.
// Two lane layouts meet in each pass:
//  - MFMA layout: lane l holds rows 16 t + 4 (l / 16) + r (r < 4) of column
//    l % 16 in the accumulators of v_mfma_f32_16x16x32_bf16;
//  - element layout: column k's x, Q x (FP32 registers) live in the G =
//    16 / RPT adjacent lanes k G .. k G + G - 1, lane k G + m holding rows
//    16 t + RPT m + r (r < RPT) of each of the wave's tiles.  RPT = 1, 2, 4
//    serves up to 4, 8, 16 columns, so a tile with few columns still keeps
//    every lane busy in the elementwise work, which bounds the pass.
// One pass: the vertices v, a are chosen in the element layout and written
// to LDS as bf16 0/1 B operands in MFMA k-slot order (a tile of up to 8
// columns puts v in B columns 0..7 and a in 8..15 of one operand, 16 columns
// use one operand each); Q v and Q a come from the MFMA with the A operand Q (pre-split into 1..3 bf16 planes q = q0 +
// q1 + q2, FP32 accumulation over the planes, fragments held in registers
// where they fit); the accumulators return to the element layout through a
// per-wave LDS exchange; partial sums are reduced over the column's lanes
// (DPP) and over the waves (LDS, in wave order).  Box vertices of a 0/1 box
// are 0/1 and exact in bf16, so the only rounding of the product is the
// plane split of Q (about 2^-8 relative for one plane, 2^-24 for three) and
// the FP32 accumulation.
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
// device counter; the workgroup exits when the counter is exhausted and all
// of its columns are idle.  A fresh node starts at the box vertex nearest to
// its supplied point (the upper bounds when none is supplied, as the plain
// form's init vertex); nodes with bounds other than 0/1 and continuations
// are left to the FP64 kernel.  A node's trajectory depends only on the node
// and the launch shape (RW, RPT), not on the column or tile that serves it.
//
// Over a bounded simplex (SX, DICG) the vertices are not elementwise: each
// pass stages the columns' gradients and row flags in LDS, one wave per
// column picks v and another the in-face a with the greedy's selection
// (mixed_simplex_v, mixed_simplex_a), and the rows read them back (two
// barriers); the maximum step adds the mass limit from the column sums of x
// and v - a.
// A column starts at the greedy vertex of the artificial direction, and a
// node whose lower bounds exceed the right side is left to the FP64 kernel.
#include <type_traits>

#include "gb_device.h"
#include "gb_mixed.h"
#include "gb_select.h"

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

__device__ __forceinline__ bf16x8 load_frag(const uint4* p) {
    return __builtin_bit_cast(bf16x8, *p);
}

constexpr unsigned int kBf16One = 0x3F80u;

// RPT consecutive rows of a B-operand slot: bit r of `bits` -> bf16 1.0 or 0.0.
template <int RPT>
__device__ __forceinline__ void put_rows(unsigned short* p, uint32_t bits) {
    if constexpr (RPT == 1) {
        *p = static_cast<unsigned short>((bits & 1u) * kBf16One);
    } else if constexpr (RPT == 2) {
        *reinterpret_cast<uint32_t*>(p) = (bits & 1u) * kBf16One | ((bits >> 1) & 1u) * (kBf16One << 16);
    } else {
        uint2 v;
        v.x = (bits & 1u) * kBf16One | ((bits >> 1) & 1u) * (kBf16One << 16);
        v.y = ((bits >> 2) & 1u) * kBf16One | ((bits >> 3) & 1u) * (kBf16One << 16);
        *reinterpret_cast<uint2*>(p) = v;
    }
}

// Exact quadratic step on the segment [0, maximum] (FP32).
__device__ __forceinline__ float quad_exact_step(float gap, float curvature, float maximum) {
    if (maximum <= 0.0f) return 0.0f;
    if (curvature > 0.0f) {
        if (gap <= 0.0f) return 0.0f;
        const float step = gap / curvature;
        return step < maximum ? step : maximum;
    }
    return -gap + 0.5f * curvature * maximum < 0.0f ? maximum : 0.0f;
}

// In-row lane permutations (DPP); every lane of the wave must be active.
constexpr int kDppSwap1 = 0xB1;        // quad_perm [1, 0, 3, 2]
constexpr int kDppSwap2 = 0x4E;        // quad_perm [2, 3, 0, 1]
constexpr int kDppHalfMirror = 0x141;  // lane i <-> 7 - i within 8
constexpr int kDppMirror = 0x140;      // lane i <-> 15 - i within 16

template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}

// Sum / min over an aligned group of G = 4, 8, 16 lanes: a butterfly, so
// every lane of the group ends with the same value (a + b == b + a).
template <int G>
__device__ __forceinline__ float group_sum(float v) {
    v = v + dpp<kDppSwap1>(v);
    v = v + dpp<kDppSwap2>(v);
    if constexpr (G >= 8) v = v + dpp<kDppHalfMirror>(v);
    if constexpr (G >= 16) v = v + dpp<kDppMirror>(v);
    return v;
}
template <int G>
__device__ __forceinline__ float group_min(float v) {
    v = fminf(v, dpp<kDppSwap1>(v));
    v = fminf(v, dpp<kDppSwap2>(v));
    if constexpr (G >= 8) v = fminf(v, dpp<kDppHalfMirror>(v));
    if constexpr (G >= 16) v = fminf(v, dpp<kDppMirror>(v));
    return v;
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

// Simplex vertices of one node column, computed by a wave from the
// column's rows staged in LDS: g (float) and flags (bit 0: lb = 1, bit 1:
// ub = 1, bit 2: x at a bound, bit 3: x at its upper bound).  Two tasks per
// column, on different waves: the forward vertex v, the greedy's for g (for
// the artificial direction g_j = -1 - j / (n + 1) while the column
// initialises), and the in-face vertex a, the greedy's over the free rows of
// the face for -g, the face keeping the full mass when sum x (txm) equals rhs
// within tolerance (lmo.cuh:517-618, :814-863; 0/1 capacities, so
// wave_select gives their vertices exactly on the FP32 gradient).  Each
// task writes one byte per row (0 or 1); the v task also writes *ok = 0 for
// a node the set cannot hold (sum lb > rhs, or for an equality sum ub < rhs).
template <int NTS>
__device__ __forceinline__ void mixed_simplex_v(const float* sgc, const uint8_t* sfc, uint8_t* out, int n, double rhs,
                                                bool eq, bool ini, int lane, int* ok) {
    unsigned long long key[NTS];
    bool c[NTS], sel[NTS];
    uint32_t f[NTS];
    int lbc = 0, ubc = 0;
#pragma unroll
    for (int t = 0; t < NTS; ++t) {
        const int j = lane + kWave * t;
        const bool in = j < n;
        f[t] = in ? sfc[j] : 0u;
        const double h = !in ? 0.0 : (ini ? -1.0 - static_cast<double>(j) / (n + 1.0) : static_cast<double>(sgc[j]));
        key[t] = order_key(h);
        c[t] = in && (f[t] & 2u) && !(f[t] & 1u) && (eq || h <= 0.0);
        lbc += __popcll(__ballot(in && (f[t] & 1u)));
        ubc += __popcll(__ballot(in && (f[t] & 2u)));
    }
    wave_select<NTS>(key, c, rhs - lbc, sel, lane);
#pragma unroll
    for (int t = 0; t < NTS; ++t) {
        const int j = lane + kWave * t;
        if (j < n) out[j] = ((f[t] & 1u) || sel[t]) ? 1 : 0;
    }
    if (lane == 0) *ok = (lbc <= rhs + 1e-9 && (!eq || ubc >= rhs - 1e-9)) ? 1 : 0;
}

template <int NTS>
__device__ __forceinline__ void mixed_simplex_a(const float* sgc, const uint8_t* sfc, uint8_t* out, int n, double rhs,
                                                bool eq, float txm, int lane) {
    unsigned long long key[NTS];
    bool c[NTS], sel[NTS], a[NTS];
    int mass = 0;
#pragma unroll
    for (int t = 0; t < NTS; ++t) {
        const int j = lane + kWave * t;
        const bool in = j < n;
        const uint32_t f = in ? sfc[j] : 0u;
        a[t] = (f & 1u) || (f & 12u) == 12u;  // lb, or ub where x sits at its upper bound
        mass += __popcll(__ballot(in && a[t]));
        const double h = in ? -static_cast<double>(sgc[j]) : 0.0;
        key[t] = order_key(h);
        c[t] = in && !(f & 4u) && (f & 2u) && !a[t];
        sel[t] = h <= 0.0;  // eligibility without the full-mass face
    }
    const bool force = eq || fabs(static_cast<double>(txm) - rhs) <=
                                 1e-6 + 1e-4 * fmax(fabs(static_cast<double>(txm)), fabs(rhs));
#pragma unroll
    for (int t = 0; t < NTS; ++t) c[t] = c[t] && (force || sel[t]);
    wave_select<NTS>(key, c, rhs - mass, sel, lane);
#pragma unroll
    for (int t = 0; t < NTS; ++t) {
        const int j = lane + kWave * t;
        if (j < n) out[j] = (a[t] || sel[t]) ? 1 : 0;
    }
}

// The wave's Q fragments are the same in every pass: where they fit beside
// the rest of the state they are loaded once and kept in registers for the
// kernel's lifetime.  CS: the k-steps held (0: none): 8 (n <= 256) for one
// plane and 1, 2 or 4 row tiles per wave and for three planes and one row
// tile (32..128 registers); 7 (n <= 224) for three planes and 2 row tiles
// per wave with one row per lane (168 VGPRs at two waves per SIMD) or 4 row
// tiles per wave with up to 2 rows per lane (336 registers at one wave per
// SIMD, partly in the AGPR file).  The other three-plane shapes would spill.
constexpr int mixed_cached_steps(int rw, int pl, int rpt) {
    if (pl == 1) return rw == 3 ? 0 : 8;
    if (rw == 1) return 8;
    return ((rw == 2 && rpt == 1) || (rw == 4 && rpt <= 2)) ? 7 : 0;
}

// QL: otherwise the tile copies all fragments into LDS once (when they fit)
// and reads them from there each pass instead of from L2.  BIG: 17..32 row
// tiles (n up to 512), 4 row tiles per wave, up to 16 k-steps.
template <int RW, int PL, bool DICG, int RPT, bool QL, int CS, bool BIG = false, bool SX = false>
__global__ void __launch_bounds__(BIG ? 512 : (RW == 1 ? 512 : 64 * ((16 + RW - 1) / RW)))
    relax_mixed_kernel(MixedParams P) {
    static_assert(!BIG || (RW == 4 && !QL && CS == 0), "wide tiles run 4 row tiles per wave from L2");
    static_assert(!SX || (DICG && !QL), "the simplex runs DICG, without an LDS copy of Q");
    constexpr int kRed = mixed_red(SX);
    constexpr int G = 16 / RPT;  // lanes of a column in the element layout
    constexpr int E = RW * RPT;  // elements per lane: rows 16 t_j + RPT m + r, e = RPT j + r
    constexpr unsigned long long kGroupMask = (1ull << G) - 1ull;
    // Up to 8 columns (RPT <= 2): the away vertex a of column k is B column
    // k + 8 of the one operand, so a single MFMA chain gives Q v and Q a.
    constexpr bool kJoint = DICG && RPT <= 2;
    constexpr bool kTwo = DICG && !kJoint;  // 16 columns: separate chains for v and a
    // Starting at a supplied point needs registers that the 3-tile and the
    // wide (streamed) shapes do not have; their nodes start at a vertex.
    constexpr bool kWarmStart = mixed_warm_start(RW, BIG);
    extern __shared__ __align__(16) unsigned char mixed_smem[];
    const int RT = P.rt;
    const int S = RT / 2;
    const int W = blockDim.x / kWave;
    const MixedLds L = mixed_lds_layout(RT, W, RW, QL ? mixed_q_bytes(RT, PL) : 0, SX);
    float* sg = reinterpret_cast<float*>(mixed_smem + L.sg);
    uint8_t* sf = reinterpret_cast<uint8_t*>(mixed_smem + L.sf);
    uint8_t* sr = reinterpret_cast<uint8_t*>(mixed_smem + L.sr);
    int* sok = reinterpret_cast<int*>(mixed_smem + L.sok);
    uint4* vops = reinterpret_cast<uint4*>(mixed_smem + L.vops);
    uint4* aops = reinterpret_cast<uint4*>(mixed_smem + L.aops);
    float* red = reinterpret_cast<float*>(mixed_smem + L.red);
    int* fetch = reinterpret_cast<int*>(mixed_smem + L.fetch);
    int* okw = reinterpret_cast<int*>(mixed_smem + L.ok);
    const int lane = lane_id();
    const int w = __builtin_amdgcn_readfirstlane(wave_id());  // wave-uniform: tile tests branch on scalars
    const int N = P.count ? *P.count : P.N;
    float* xch = reinterpret_cast<float*>(mixed_smem + L.xch) + static_cast<size_t>(w) * (2 * RW * 256);
    // Columns that never hold a node multiply zeros.
    for (int i = threadIdx.x; i < S * 64; i += blockDim.x) {
        vops[i] = make_uint4(0u, 0u, 0u, 0u);
        aops[i] = make_uint4(0u, 0u, 0u, 0u);
    }
    const uint4* qf = reinterpret_cast<const uint4*>(mixed_smem + L.qf);
    if constexpr (QL) {
        uint4* dst = reinterpret_cast<uint4*>(mixed_smem + L.qf);
        for (int i = threadIdx.x; i < PL * RT * S * kWave; i += blockDim.x) dst[i] = P.frag[i];
    }
    const int c = lane & 15;   // MFMA layout: column c, rows 4 q + r
    const int q = lane >> 4;
    const int col = lane / G;  // element layout: column col, rows RPT m + r
    const int m = lane % G;
    const int n = P.n;
    const float gs = P.gs;
    const float hs = 0.5f * P.gs;
    const size_t tstride = static_cast<size_t>(S) * kWave;
    const uint4* fbase = P.frag + lane;

    float x[E], qx[E], ls[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int t = w + W * (e / RPT);
        const int row = 16 * t + RPT * m + e % RPT;
        ls[e] = t < RT ? P.lsc[row] : 0.0f;
        x[e] = 0.0f;
        qx[e] = 0.0f;
    }
    constexpr bool kCache = CS > 0;
    static_assert(!(kCache && QL), "register-cached fragments need no LDS copy");
    constexpr int kSMax = kCache ? CS : 8;
    bf16x8 fc[kCache ? PL : 1][kCache ? RW : 1][kCache ? kSMax : 1];
    if (kCache) {
#pragma unroll
        for (int p = 0; p < PL; ++p)
#pragma unroll
            for (int j = 0; j < RW; ++j)
#pragma unroll
                for (int s = 0; s < kSMax; ++s) {
                    const int t = w + W * j;
                    fc[p][j][s] = (t < RT && s < S) ? load_frag(fbase + (static_cast<size_t>(p) * RT + t) * tstride +
                                                                static_cast<size_t>(s) * kWave)
                                                    : bf16x8{};
                }
    }
    uint32_t lbm = 0, ubm = 0, x0m = 0, bestm = 0;
    float txm = 0.0f;  // simplex: sum x of the column's iterate
    int phase = col < P.cols ? kPhaseFetch : kPhaseDone;
    int node = -1, it = 0, initial = 0;
    double eps = 0.0, cutoff = 0.0, best = kNone;
    __syncthreads();

    while (true) {
        // Columns without a node take the next ones from the counter (wave 0
        // fetches, the tile's waves share the result); nodes outside the
        // 0/1-box scope and continuations are left to the FP64 kernel.  The
        // column state is replicated in every wave and in the column's lanes,
        // so all waves take the same branches and meet at the same barriers.
        while (__ballot(phase == kPhaseFetch) != 0ull) {
            if (w == 0 && m == 0 && phase == kPhaseFetch) fetch[col] = atomicAdd(P.counter, 1);
            __syncthreads();
            if (phase == kPhaseFetch) {
                node = fetch[col];
                if (node >= N) phase = kPhaseDone;
            }
            const bool loading = phase == kPhaseFetch;
            bool ok = true;
            uint32_t lm = 0, um = 0, xm = 0;
            int init_it = 0;
            bool warm = false;
            if (loading) {
                const double* lbn = P.lb + static_cast<size_t>(node) * n;
                const double* ubn = P.ub + static_cast<size_t>(node) * n;
                const double* xn = P.x + static_cast<size_t>(node) * n;
                init_it = P.iters[node];
                const bool from_x = init_it != 0 || (P.status[node] & GB_NODE_INITIAL_POINT_SUPPLIED) != 0;
                // a supplied point with Q x at hand: the column starts at x itself
                warm = kWarmStart && !SX && P.qx0 != nullptr && from_x && init_it == 0;
#pragma unroll 1
                for (int e = 0; e < E; ++e) {
                    const int t = w + W * (e / RPT);
                    const int row = 16 * t + RPT * m + e % RPT;
                    if (t < RT && row < n) {
                        const double l = lbn[row], u = ubn[row];
                        if (!(l <= u) || !(l == 0.0 || l == 1.0) || !(u == 0.0 || u == 1.0)) ok = false;
                        if (SX && !P.is_int[row]) ok = false;
                        const double xv = from_x ? xn[row] : u;
                        const bool xb = (xv >= 0.5 * (l + u)) ? (u == 1.0) : (l == 1.0);
                        lm |= static_cast<uint32_t>(l == 1.0) << e;
                        um |= static_cast<uint32_t>(u == 1.0) << e;
                        xm |= static_cast<uint32_t>(xb) << e;
                    }
                }
            }
            const unsigned long long bad = __ballot(!(ok && init_it == 0));
            if (m == 0) okw[w * 16 + col] = ((bad >> (col * G)) & kGroupMask) == 0ull ? 1 : 0;
            __syncthreads();
            bool all_ok = true;
            for (int v = 0; v < W; ++v) all_ok = all_ok && okw[v * 16 + col] != 0;
            if (loading && all_ok) {
                phase = warm ? kPhaseActive : kPhaseInit;
                lbm = lm;
                ubm = um;
                x0m = xm;
                it = 0;
                initial = init_it;
                eps = P.eps[node];
                cutoff = P.cutoff ? P.cutoff[node] : __builtin_inf();
                best = kNone;
                bestm = 0;
                if (kWarmStart && warm) {
                    // x clipped to the box and Q x (the column's registers are
                    // free while it fetches)
                    const double* xn = P.x + static_cast<size_t>(node) * n;
                    const double* qn = P.qx0 + static_cast<size_t>(node) * n;
#pragma unroll
                    for (int e = 0; e < E; ++e) {
                        const int t = w + W * (e / RPT);
                        const int row = 16 * t + RPT * m + e % RPT;
                        const float lf = ((lm >> e) & 1u) ? 1.0f : 0.0f;
                        const float uf = ((um >> e) & 1u) ? 1.0f : 0.0f;
                        x[e] = 0.0f;
                        qx[e] = 0.0f;
                        if (t < RT && row < n) {
                            x[e] = fminf(fmaxf(static_cast<float>(xn[row]), lf), uf);
                            qx[e] = static_cast<float>(qn[row]);
                        }
                    }
                }
            }
            __syncthreads();
        }
        if (__ballot(phase != kPhaseDone) == 0ull) break;
        bool act = phase == kPhaseActive;
        bool ini = phase == kPhaseInit;

        // Vertices of the lane's rows, partial sums at x.
        uint32_t vm = 0, am = 0;
        float gp = 0.0f, fp = 0.0f;
        float g[E];
        if constexpr (SX) {
            // The column's rows go to LDS, a wave per column and vertex picks
            // v or a, and the rows read them back.
            const int NP = 16 * RT;
            float* sgc = sg + col * NP;
            uint8_t* sfc = sf + col * NP;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int t = w + W * (e / RPT);
                const int row = 16 * t + RPT * m + e % RPT;
                const bool lb1 = (lbm >> e) & 1u;
                const bool ub1 = (ubm >> e) & 1u;
                g[e] = gs * qx[e] + ls[e];
                if (t < RT && row < n && (act || ini)) {
                    const float lf = lb1 ? 1.0f : 0.0f, uf = ub1 ? 1.0f : 0.0f;
                    const bool at_l = fabsf(x[e] - lf) <= 1e-6f + 1e-4f * fmaxf(fabsf(x[e]), lf);
                    const bool at_u = !at_l && fabsf(x[e] - uf) <= 1e-6f + 1e-4f * fmaxf(fabsf(x[e]), uf);
                    sgc[row] = g[e];
                    sfc[row] = static_cast<uint8_t>((lb1 ? 1u : 0u) | (ub1 ? 2u : 0u) | ((at_l || at_u) ? 4u : 0u) |
                                                    (at_u ? 8u : 0u));
                }
            }
            __syncthreads();
            const int nts = (n + kWave - 1) / kWave;
            uint8_t* sv = sr;                 // v of every column
            uint8_t* sa = sr + 16 * NP;       // a of every column
            // tasks 2k (v of column k) and 2k + 1 (a of column k) over the waves
            for (int task = w; task < 2 * (kWave / G); task += W) {
                const int k = task >> 1;
                const int kp = __builtin_amdgcn_readlane(phase, k * G);
                if (kp != kPhaseActive && kp != kPhaseInit) continue;
                const bool kini = kp == kPhaseInit;
                const float* gk = sg + k * NP;
                const uint8_t* fk = sf + k * NP;
                const bool eq = P.equality != 0;
                if ((task & 1) == 0) {
                    uint8_t* o = sv + k * NP;
                    if constexpr (BIG) {
                        mixed_simplex_v<8>(gk, fk, o, n, P.rhs, eq, kini, lane, sok + k);
                    } else {
                        if (nts <= 1) mixed_simplex_v<1>(gk, fk, o, n, P.rhs, eq, kini, lane, sok + k);
                        else if (nts <= 2) mixed_simplex_v<2>(gk, fk, o, n, P.rhs, eq, kini, lane, sok + k);
                        else mixed_simplex_v<4>(gk, fk, o, n, P.rhs, eq, kini, lane, sok + k);
                    }
                } else if (!kini) {
                    uint8_t* o = sa + k * NP;
                    const float kt = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, txm), k * G));
                    if constexpr (BIG) {
                        mixed_simplex_a<8>(gk, fk, o, n, P.rhs, eq, kt, lane);
                    } else {
                        if (nts <= 1) mixed_simplex_a<1>(gk, fk, o, n, P.rhs, eq, kt, lane);
                        else if (nts <= 2) mixed_simplex_a<2>(gk, fk, o, n, P.rhs, eq, kt, lane);
                        else mixed_simplex_a<4>(gk, fk, o, n, P.rhs, eq, kt, lane);
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int t = w + W * (e / RPT);
                const int row = 16 * t + RPT * m + e % RPT;
                if (t < RT && row < n) {
                    vm |= static_cast<uint32_t>(sr[col * NP + row] & 1u) << e;
                    am |= static_cast<uint32_t>(sr[(16 + col) * NP + row] & 1u) << e;
                }
            }
            if ((act || ini) && sok[col] == 0) {
                // a node the set cannot hold: left to the FP64 kernel
                act = false;
                ini = false;
                phase = kPhaseFetch;
            }
            if (ini) x0m = vm;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                gp += g[e] * (x[e] - (((vm >> e) & 1u) ? 1.0f : 0.0f));
                fp += x[e] * (hs * qx[e] + ls[e]);
            }
        } else
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const bool lb1 = (lbm >> e) & 1u;
            const bool ub1 = (ubm >> e) & 1u;
            g[e] = gs * qx[e] + ls[e];
            const bool vb = g[e] > 0.0f ? lb1 : ub1;
            bool ab = false;
            if (DICG) {
                const float uf = ub1 ? 1.0f : 0.0f, lf = lb1 ? 1.0f : 0.0f;
                const bool at_u = fabsf(x[e] - uf) <= 1e-6f + 1e-4f * uf;
                const bool at_l = fabsf(x[e] - lf) <= 1e-6f + 1e-4f * lf;
                ab = at_u ? ub1 : (at_l ? lb1 : ((-g[e] > 0.0f) ? lb1 : ub1));
            }
            vm |= static_cast<uint32_t>(vb) << e;
            am |= static_cast<uint32_t>(ab) << e;
            gp += g[e] * (x[e] - (vb ? 1.0f : 0.0f));
            fp += x[e] * (hs * qx[e] + ls[e]);
        }
        vm = ini ? x0m : (act ? vm : 0u);
        am = act ? am : 0u;
        // B operands: uint4 slot (s * 16 + k) * 4 + qq holds rows 4 qq .. 4 qq + 3
        // of tiles 2s, 2s+1 of column k as 8 bf16 (k-slot i = 4 h + r).
        const int qq = (RPT * m) >> 2;
        const int r0 = (RPT * m) & 3;
#pragma unroll
        for (int j = 0; j < RW; ++j) {
            const int t = w + W * j;
            if (t < RT) {
                const int slot = ((t >> 1) * 16 + col) * 4 + qq;
                const int off = (t & 1) * 4 + r0;
                constexpr uint32_t kRows = (1u << RPT) - 1u;
                put_rows<RPT>(reinterpret_cast<unsigned short*>(vops + slot) + off, (vm >> (RPT * j)) & kRows);
                if (kJoint)
                    put_rows<RPT>(reinterpret_cast<unsigned short*>(vops + slot + 32) + off, (am >> (RPT * j)) & kRows);
                else if (DICG)
                    put_rows<RPT>(reinterpret_cast<unsigned short*>(aops + slot) + off, (am >> (RPT * j)) & kRows);
            }
        }
        __syncthreads();

        // Q v, Q a for the wave's row tiles over all k-steps (tiles past RT
        // multiply zero fragments; the k-step count is a compile-time
        // constant of each case, so the MFMA chain has no branches).
        f32x4 accv[RW], acca[RW];
#pragma unroll
        for (int j = 0; j < RW; ++j) {
            accv[j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            acca[j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        }
        auto products = [&](auto steps) {
            // (the launcher runs a register-cached kernel only for S <= CS)
            constexpr int SS = decltype(steps)::value < kSMax ? decltype(steps)::value : kSMax;
            if constexpr (kCache) {
#pragma unroll
                for (int s = 0; s < SS; ++s) {
                    const bf16x8 bv = __builtin_bit_cast(bf16x8, vops[(s * 16 + c) * 4 + q]);
                    const bf16x8 ba = kTwo ? __builtin_bit_cast(bf16x8, aops[(s * 16 + c) * 4 + q]) : bv;
#pragma unroll
                    for (int p = 0; p < PL; ++p)
#pragma unroll
                        for (int j = 0; j < RW; ++j) {
                            accv[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fc[p][j][s], bv, accv[j], 0, 0, 0);
                            if (kTwo)
                                acca[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fc[p][j][s], ba, acca[j], 0, 0, 0);
                        }
                }
            } else {
                // Streaming: the next k-step's fragments load while the
                // current one multiplies.
                bf16x8 cur[PL][RW];
#pragma unroll
                for (int p = 0; p < PL; ++p)
#pragma unroll
                    for (int j = 0; j < RW; ++j) {
                        const int t = w + W * j;
                        cur[p][j] = t >= RT ? bf16x8{}
                                    : QL ? __builtin_bit_cast(bf16x8, qf[(static_cast<size_t>(p) * RT + t) * tstride + lane])
                                         : load_frag(fbase + (static_cast<size_t>(p) * RT + t) * tstride);
                    }
#pragma unroll
                for (int s = 0; s < SS; ++s) {
                    bf16x8 nxt[PL][RW];
                    if (s + 1 < SS) {
#pragma unroll
                        for (int p = 0; p < PL; ++p)
#pragma unroll
                            for (int j = 0; j < RW; ++j) {
                                const int t = w + W * j;
                                const size_t at = (static_cast<size_t>(p) * RT + t) * tstride +
                                                  static_cast<size_t>(s + 1) * kWave;
                                nxt[p][j] = t >= RT ? bf16x8{}
                                            : QL ? __builtin_bit_cast(bf16x8, qf[at + lane]) : load_frag(fbase + at);
                            }
                    }
                    const bf16x8 bv = __builtin_bit_cast(bf16x8, vops[(s * 16 + c) * 4 + q]);
                    const bf16x8 ba = kTwo ? __builtin_bit_cast(bf16x8, aops[(s * 16 + c) * 4 + q]) : bv;
#pragma unroll
                    for (int p = 0; p < PL; ++p)
#pragma unroll
                        for (int j = 0; j < RW; ++j) {
                            accv[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur[p][j], bv, accv[j], 0, 0, 0);
                            if (kTwo) acca[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur[p][j], ba, acca[j], 0, 0, 0);
                        }
                    if (s + 1 < SS) {
#pragma unroll
                        for (int p = 0; p < PL; ++p)
#pragma unroll
                            for (int j = 0; j < RW; ++j) cur[p][j] = nxt[p][j];
                    }
                }
            }
        };
        if constexpr (BIG) {
            // 9..16 k-steps: a runtime loop (one k-step of fragments in
            // flight ahead), so the unrolled chains do not hold every
            // step's addresses in registers.
            const uint4* fp[PL][RW];
#pragma unroll
            for (int p = 0; p < PL; ++p)
#pragma unroll
                for (int j = 0; j < RW; ++j) {
                    const int t = w + W * j;
                    fp[p][j] = fbase + (static_cast<size_t>(p) * RT + (t < RT ? t : 0)) * tstride;
                }
            bf16x8 cur[PL][RW];
#pragma unroll
            for (int p = 0; p < PL; ++p)
#pragma unroll
                for (int j = 0; j < RW; ++j) cur[p][j] = w + W * j < RT ? load_frag(fp[p][j]) : bf16x8{};
#pragma unroll 1
            for (int s = 0; s < S; ++s) {
                bf16x8 nxt[PL][RW];
                const bool more = s + 1 < S;
#pragma unroll
                for (int p = 0; p < PL; ++p)
#pragma unroll
                    for (int j = 0; j < RW; ++j)
                        nxt[p][j] = (more && w + W * j < RT) ? load_frag(fp[p][j] + static_cast<size_t>(s + 1) * kWave)
                                                             : bf16x8{};
                const bf16x8 bv = __builtin_bit_cast(bf16x8, vops[(s * 16 + c) * 4 + q]);
                const bf16x8 ba = kTwo ? __builtin_bit_cast(bf16x8, aops[(s * 16 + c) * 4 + q]) : bv;
#pragma unroll
                for (int p = 0; p < PL; ++p)
#pragma unroll
                    for (int j = 0; j < RW; ++j) {
                        accv[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur[p][j], bv, accv[j], 0, 0, 0);
                        if (kTwo) acca[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur[p][j], ba, acca[j], 0, 0, 0);
                    }
#pragma unroll
                for (int p = 0; p < PL; ++p)
#pragma unroll
                    for (int j = 0; j < RW; ++j) cur[p][j] = nxt[p][j];
            }
        } else {
            switch (S) {
                case 1: products(std::integral_constant<int, 1>{}); break;
                case 2: products(std::integral_constant<int, 2>{}); break;
                case 3: products(std::integral_constant<int, 3>{}); break;
                case 4: products(std::integral_constant<int, 4>{}); break;
                case 5: products(std::integral_constant<int, 5>{}); break;
                case 6: products(std::integral_constant<int, 6>{}); break;
                case 7: products(std::integral_constant<int, 7>{}); break;
                default: products(std::integral_constant<int, 8>{}); break;
            }
        }

        // Accumulators to the element layout: the wave's own LDS region,
        // [array][j][column][row] floats (no other wave touches it); with the
        // joint operand, B column c >= 8 holds Q a of column c - 8.
        float qv[E], qa[E];
#pragma unroll
        for (int j = 0; j < RW; ++j) {
            float* pv = kJoint ? xch + (((c >> 3) * RW + j) * 16 + (c & 7)) * 16 + 4 * q
                               : xch + ((0 * RW + j) * 16 + c) * 16 + 4 * q;
#pragma unroll
            for (int r = 0; r < 4; ++r) pv[r] = accv[j][r];
            if (kTwo) {
                float* pa = xch + ((1 * RW + j) * 16 + c) * 16 + 4 * q;
#pragma unroll
                for (int r = 0; r < 4; ++r) pa[r] = acca[j][r];
            }
        }
#pragma unroll
        for (int j = 0; j < RW; ++j) {
            const float* pv = xch + ((0 * RW + j) * 16 + col) * 16 + RPT * m;
            const float* pa = xch + ((1 * RW + j) * 16 + col) * 16 + RPT * m;
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                qv[RPT * j + r] = pv[r];
                qa[RPT * j + r] = DICG ? pa[r] : 0.0f;
            }
        }

        // Vertex values and step partials; Q d replaces Q v (Q v for a column
        // that is being initialised).
        float fvp = 0.0f, fap = 0.0f, pgp = 0.0f, cvp = 0.0f, rr = 1.0f;
        float sxp = 0.0f, sdp = 0.0f, svp = 0.0f;  // simplex: sum x, sum (v - a), sum v
        float qd[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const float vf = ((vm >> e) & 1u) ? 1.0f : 0.0f;
            fvp += vf * (hs * qv[e] + ls[e]);
            if (SX) {
                sxp += x[e];
                sdp += vf - (((am >> e) & 1u) ? 1.0f : 0.0f);
                svp += vf;
            }
            float d, qde;
            if (DICG) {
                const float af = ((am >> e) & 1u) ? 1.0f : 0.0f;
                fap += af * (hs * qa[e] + ls[e]);
                d = vf - af;
                qde = qv[e] - qa[e];
                // a - v is -1, 0 or 1 on a 0/1 box: the ratio needs no division.
                const float uf = ((ubm >> e) & 1u) ? 1.0f : 0.0f;
                const float lf = ((lbm >> e) & 1u) ? 1.0f : 0.0f;
                const float lim = d > 0.0f ? uf - x[e] : (d < 0.0f ? x[e] - lf : 1.0f);
                rr = fminf(rr, lim);
            } else {
                d = vf - x[e];
                qde = qv[e] - qx[e];
            }
            pgp += -g[e] * d;
            cvp += gs * d * qde;
            qd[e] = ini ? qv[e] : qde;
        }
        // Column totals: the column's lanes (DPP), then the waves in order.
        float part[kRed];
        part[0] = group_sum<G>(gp);
        part[1] = group_sum<G>(fp);
        part[2] = group_sum<G>(fvp);
        part[3] = DICG ? group_sum<G>(fap) : 0.0f;
        part[4] = group_sum<G>(pgp);
        part[5] = group_sum<G>(cvp);
        part[6] = DICG ? group_min<G>(rr) : 1.0f;
        part[7] = SX ? group_sum<G>(sxp) : 0.0f;
        if constexpr (SX) {
            part[8] = group_sum<G>(sdp);
            part[9] = group_sum<G>(svp);
            part[10] = 0.0f;
            part[11] = 0.0f;
        }
        if (m == 0) {
            float* o = red + (w * 16 + col) * kRed;
#pragma unroll
            for (int k = 0; k < kRed; ++k) o[k] = part[k];
        }
        __syncthreads();
        constexpr int kTot = SX ? 10 : 7;
        float tot[kTot];
        {
            const float* o = red + col * kRed;
#pragma unroll
            for (int k = 0; k < kTot; ++k) tot[k] = o[k];
        }
        for (int v = 1; v < W; ++v) {
            const float* o = red + (v * 16 + col) * kRed;
#pragma unroll
            for (int k = 0; k < kTot; ++k) {
                if (k == 6) tot[6] = fminf(tot[6], o[6]);
                else tot[k] = tot[k] + o[k];
            }
        }
        const float gap = tot[0];
        const double fx = static_cast<double>(tot[1]) + P.kappa;
        const double fv = static_cast<double>(tot[2]) + P.kappa;
        const double fa = DICG ? static_cast<double>(tot[3]) + P.kappa : kNone;
        float gmax = DICG ? fmaxf(0.0f, fminf(tot[6], 1.0f)) : 1.0f;
        if constexpr (SX) {
            // the mass limit of the maximum step (lmo.cuh:865-906), d = a - v
            const double tx = tot[7], td = -static_cast<double>(tot[8]);
            double agg;
            if (P.equality && fabs(td) > 1e-9) agg = 0.0;
            else agg = td < 0.0 ? fmin(1.0, (tx - P.rhs) / td) : 1.0;
            gmax = fmaxf(0.0f, fminf(gmax, static_cast<float>(agg)));
        }

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
            const bool stop = (gap >= 0.0f && static_cast<double>(gap) < tol) || (fx - gap >= cutoff);
            if (!stop) {
                gm = quad_exact_step(tot[4], tot[5], gmax);
                step = true;
            }
            retire = stop || it >= P.max_iters;
        }
        // x += gamma d, Q x += gamma Q d for stepping columns; Q x = Q x0 for
        // columns being initialised.
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const float vf = ((vm >> e) & 1u) ? 1.0f : 0.0f;
            const float d = DICG ? vf - (((am >> e) & 1u) ? 1.0f : 0.0f) : vf - x[e];
            const float uf = ((ubm >> e) & 1u) ? 1.0f : 0.0f;
            const float lf = ((lbm >> e) & 1u) ? 1.0f : 0.0f;
            const float nx = fminf(fmaxf(fmaf(gm, d, x[e]), lf), uf);
            const float nq = fmaf(gm, qd[e], qx[e]);
            const float x0 = ((x0m >> e) & 1u) ? 1.0f : 0.0f;
            x[e] = ini ? x0 : (step ? nx : x[e]);
            qx[e] = ini ? qd[e] : (step ? nq : qx[e]);
        }
        if constexpr (SX) txm = ini ? tot[9] : (step ? fmaf(gm, tot[8], tot[7]) : tot[7]);
        if (ini) phase = kPhaseActive;

        if (retire) {
            double* xn = P.x + static_cast<size_t>(node) * n;
            double* bvn = P.best_v + static_cast<size_t>(node) * n;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int t = w + W * (e / RPT);
                const int row = 16 * t + RPT * m + e % RPT;
                if (t < RT && row < n) {
                    xn[row] = static_cast<double>(x[e]);
                    if (best < kNone) bvn[row] = ((bestm >> e) & 1u) ? 1.0 : 0.0;
                }
            }
            if (w == 0 && m == 0) {
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

// Q x for the nodes with a supplied point and no iterations, in FP64.  A
// workgroup takes kWarmNodes nodes and one output row per thread, so Qt is
// read once per kWarmNodes nodes; every output is the k-ordered FMA chain
// sum_k Qt[k][j] x[k].
constexpr int kWarmNodes = 8;
__global__ void __launch_bounds__(256) mixed_warm_products_kernel(const double* Qt, int n, int N, const double* X,
                                                                  const uint8_t* status, const int* iters,
                                                                  double* out, const int* count) {
    if (count && *count < N) N = *count;
    __shared__ double xs[kWarmNodes][kMixedMaxN];
    __shared__ int warm[kWarmNodes];
    const int node0 = blockIdx.x * kWarmNodes;
    if (threadIdx.x < kWarmNodes) {
        const int i = node0 + static_cast<int>(threadIdx.x);
        warm[threadIdx.x] = (i < N && (status[i] & GB_NODE_INITIAL_POINT_SUPPLIED) != 0 && iters[i] == 0) ? 1 : 0;
    }
    __syncthreads();
    bool any = false;
    for (int v = 0; v < kWarmNodes; ++v) any = any || warm[v] != 0;
    if (!any) return;  // block-uniform: no node of this block starts at a supplied point
    for (int v = 0; v < kWarmNodes; ++v)
        for (int k = threadIdx.x; k < n; k += blockDim.x)
            xs[v][k] = warm[v] ? X[static_cast<size_t>(node0 + v) * n + k] : 0.0;
    __syncthreads();
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        double acc[kWarmNodes];
#pragma unroll
        for (int v = 0; v < kWarmNodes; ++v) acc[v] = 0.0;
        for (int k = 0; k < n; ++k) {
            const double q = Qt[static_cast<size_t>(k) * n + j];
#pragma unroll
            for (int v = 0; v < kWarmNodes; ++v) acc[v] = fma(q, xs[v][k], acc[v]);
        }
#pragma unroll
        for (int v = 0; v < kWarmNodes; ++v)
            if (warm[v]) out[static_cast<size_t>(node0 + v) * n + j] = acc[v];
    }
}

}  // namespace

int launch_mixed_warm_products(const double* Qt, int n, int N, const double* X, const uint8_t* status,
                               const int* iters, double* out, void* stream, const int* count) {
    if (N <= 0 || n <= 0) return 0;
    if (n > kMixedMaxN) return static_cast<int>(hipErrorInvalidValue);
    const unsigned blocks = static_cast<unsigned>((N + kWarmNodes - 1) / kWarmNodes);
    hipLaunchKernelGGL(mixed_warm_products_kernel, dim3(blocks), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       Qt, n, N, X, status, iters, out, count);
    return static_cast<int>(hipGetLastError());
}

int launch_mixed_prepare(const double* Qt, const double* lin, double gs, double ls, int n, int rt,
                         int planes, uint4* frag, float* lsc, void* stream) {
    const size_t total = static_cast<size_t>(planes) * rt * (rt / 2) * kWave;
    const unsigned blocks = static_cast<unsigned>((total + 255) / 256);
    hipLaunchKernelGGL(mixed_prepare_kernel, dim3(blocks), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), Qt, lin, gs, ls, n, rt, planes, frag, lsc);
    return static_cast<int>(hipGetLastError());
}

template <typename K>
int start_mixed(K kernel, dim3 grid, dim3 block, size_t lds, hipStream_t st, const MixedParams& p) {
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return (int)e;
    }
    hipLaunchKernelGGL(kernel, grid, block, lds, st, p);
    return 0;
}

template <int RW, int PL, bool DI, int RPT, bool SX = false>
int launch_mixed_shape(const MixedParams& p, dim3 grid, dim3 block, int waves, hipStream_t st) {
    const size_t base = mixed_lds_layout(p.rt, waves, RW, 0, SX).total;
    if constexpr (RW == 4) {
        if (p.rt > kMixedSmallRowTiles)
            return start_mixed(relax_mixed_kernel<4, PL, DI, RPT, false, 0, true, SX>, grid, block, base, st, p);
    }
    // (the simplex's selection stage leaves no room for three planes of
    // fragments in registers beside more than one row tile per wave)
    constexpr int kCS = (SX && PL == 3 && RW >= 2) ? 0 : mixed_cached_steps(RW, PL, RPT);
    if constexpr (kCS > 0) {
        if (p.rt / 2 <= kCS)
            return start_mixed(relax_mixed_kernel<RW, PL, DI, RPT, false, kCS, false, SX>, grid, block, base, st, p);
    }
    if constexpr (kCS < kMixedSmallRowTiles / 2) {
        if constexpr (!SX) {  // the simplex keeps its rows in LDS instead of a copy of Q
            const size_t with_q = mixed_lds_layout(p.rt, waves, RW, mixed_q_bytes(p.rt, PL)).total;
            if (with_q <= kMixedLdsBudget)
                return start_mixed(relax_mixed_kernel<RW, PL, DI, RPT, true, 0>, grid, block, with_q, st, p);
        }
        return start_mixed(relax_mixed_kernel<RW, PL, DI, RPT, false, 0, false, SX>, grid, block, base, st, p);
    } else {
        return static_cast<int>(hipErrorInvalidValue);  // unreachable: every k-step is held
    }
}

int launch_relax_mixed(const MixedParams& p, bool dicg, int rw, int blocks, void* stream) {
    if (p.cols < 1 || p.cols > 16 || rw < 1 || rw > kMixedMaxRowTilesPerWave || p.rt < 2 ||
        p.rt > 2 * kMixedSmallRowTiles || (p.rt > kMixedSmallRowTiles && rw != 4) || (p.simplex && !dicg))
        return static_cast<int>(hipErrorInvalidValue);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(mixed_reset_kernel, dim3(1), dim3(1), 0, st, p.counter);
    const int waves = (p.rt + rw - 1) / rw;
    const int rpt = mixed_rows_per_lane(p.cols);
    const dim3 grid(blocks), block(waves * kWave);
    int e = 0;
#define GB_MIXED_RPT_SX(RW, PL, DI, SX)                                                      \
    do {                                                                                     \
        if (rpt == 1) e = launch_mixed_shape<RW, PL, DI, 1, SX>(p, grid, block, waves, st);  \
        else if (rpt == 2) e = launch_mixed_shape<RW, PL, DI, 2, SX>(p, grid, block, waves, st); \
        else e = launch_mixed_shape<RW, PL, DI, 4, SX>(p, grid, block, waves, st);           \
    } while (0)
#define GB_MIXED_RPT(RW, PL, DI) GB_MIXED_RPT_SX(RW, PL, DI, false)
#define GB_MIXED(RW)                                                     \
    case RW:                                                             \
        if (p.simplex) {                                                 \
            if (p.planes == 1) GB_MIXED_RPT_SX(RW, 1, true, true);       \
            else GB_MIXED_RPT_SX(RW, 3, true, true);                     \
        } else if (p.planes == 1) {                                      \
            if (dicg) GB_MIXED_RPT(RW, 1, true);                         \
            else GB_MIXED_RPT(RW, 1, false);                             \
        } else {                                                         \
            if (dicg) GB_MIXED_RPT(RW, 3, true);                         \
            else GB_MIXED_RPT(RW, 3, false);                             \
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
#undef GB_MIXED_RPT
#undef GB_MIXED_RPT_SX
    if (e != 0) return e;
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
        GB_PROD(18)
        GB_PROD(20)
        GB_PROD(22)
        GB_PROD(24)
        GB_PROD(26)
        GB_PROD(28)
        GB_PROD(30)
        GB_PROD(32)
        default:
            return static_cast<int>(hipErrorInvalidValue);
    }
#undef GB_PROD
    return static_cast<int>(hipGetLastError());
}

}  // namespace gb
