// Wave-level selection of the greedy simplex vertex (0/1 capacities): the
// first `want` members of a set in (key, index) order, without a sort.
// Shared by the batched FP64 kernel and the mixed-precision pass.
#pragma once
#include <hip/hip_runtime.h>

namespace gb {

// Order-preserving unsigned image of a double, with -0 and +0 equal (the
// greedy's comparisons are on doubles, lmo.cuh:517-618).
__device__ __forceinline__ unsigned long long order_key(double g) {
    const unsigned long long u = static_cast<unsigned long long>(__double_as_longlong(g == 0.0 ? 0.0 : g));
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

// Largest X with #{c && w < X} < want (1 <= want <= #c): MSB-first
// bisection over the 32-bit words w; every count is a ballot popcount, so
// X is wave-uniform.
template <int NT>
__device__ __forceinline__ unsigned int wave_cut(const unsigned int (&w)[NT], const unsigned long long (&cm)[NT],
                                                 int want) {
    unsigned int X = 0;
    for (int b = 31; b >= 0; --b) {
        const unsigned int Y = X | (1u << b);
        int cnt = 0;
#pragma unroll
        for (int t = 0; t < NT; ++t) cnt += __popcll(__ballot(w[t] < Y) & cm[t]);
        if (cnt < want) X = Y;
    }
    return X;
}

// sel = the `want` smallest members of the set c by (key, index), index
// j = lane + 64 t (t-major order): the first `want` coordinates of the
// greedy's rank order.  The cut is found on the high words, then, within the
// high word it splits, on the low words; equal keys go by index.
template <int NT>
__device__ __forceinline__ void wave_select(const unsigned long long (&key)[NT], const bool (&c)[NT], double want_d,
                            bool (&sel)[NT], int lane) {
    unsigned long long cm[NT];
    int total = 0;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        cm[t] = __ballot(c[t]);
        total += __popcll(cm[t]);
    }
    if (want_d >= static_cast<double>(total) || !(want_d >= 1.0)) {
        const bool all = want_d >= static_cast<double>(total);
#pragma unroll
        for (int t = 0; t < NT; ++t) sel[t] = all && c[t];
        return;
    }
    const int want = static_cast<int>(want_d);
    unsigned int hi[NT], lo[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        hi[t] = static_cast<unsigned int>(key[t] >> 32);
        lo[t] = static_cast<unsigned int>(key[t]);
    }
    const unsigned int th = wave_cut<NT>(hi, cm, want);
    int below = 0, group = 0;
    unsigned long long gm[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        below += __popcll(__ballot(hi[t] < th) & cm[t]);
        gm[t] = __ballot(hi[t] == th) & cm[t];
        group += __popcll(gm[t]);
    }
    const int need = want - below;  // 1 <= need <= group
#pragma unroll
    for (int t = 0; t < NT; ++t) sel[t] = c[t] && hi[t] < th;
    if (need >= group) {
#pragma unroll
        for (int t = 0; t < NT; ++t) sel[t] = sel[t] || (c[t] && hi[t] == th);
        return;
    }
    const unsigned int tl = wave_cut<NT>(lo, gm, need);
    int below2 = 0;
#pragma unroll
    for (int t = 0; t < NT; ++t) below2 += __popcll(__ballot(lo[t] < tl) & gm[t]);
    const int need2 = need - below2;  // >= 1 members with key == (th, tl), taken by index
    const unsigned long long lt = (1ull << lane) - 1ull;
    int base = 0;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const bool grp = c[t] && hi[t] == th;
        const unsigned long long tm = __ballot(lo[t] == tl) & gm[t];
        const bool tie = grp && lo[t] == tl;
        const int r = base + __popcll(tm & lt);
        sel[t] = sel[t] || (grp && lo[t] < tl) || (tie && r < need2);
        base += __popcll(tm);
    }
}

}  // namespace gb
