// Parameters of the batched-form relaxation kernel (relax_fast.hip).
#pragma once
#include <stdint.h>

#include "../../include/gpuboscia_amd.h"

namespace gb {

struct FastParams {
    int n;
    int n_pad;              // per-wave LDS staging stride (doubles)
    const double* Qt;       // column-major Q (Qt[k*n + j] == Q[j,k])
    const double* lin;      // may be NULL
    double gs, ls, kappa;
    int dicg;
    int refresh;            // direct recompute of Q x every `refresh` iterations (0 = never)
    int N, max_iters;
    double* x;
    const double* lb;
    const double* ub;
    const double* eps;
    int* iters;
    uint8_t* status;
    double* gap;
    double* value;
    double* best_v;
    double* best_val;
    const uint8_t* flags;
    double ctol;
    const double* cutoff;   // per node, may be NULL (= +inf)
    int* counter;           // work queue head (reset by the launcher)
    long long* stats;       // optional [N][2]: column updates, direct GEMVs (may be NULL)
    double* gout;           // optional [N][n]: gradient at the returned iterate (may be NULL)
    int rescore;            // re-score best_v[i] in FP64 first when best_val[i] < 1e300
    const int* count;       // optional device batch size: nodes [0, *count) of the N rows are relaxed
    // Bounded-simplex LMO (simplex != 0): sum x <= rhs (== rhs when equality),
    // integral rhs, 0/1 variables all marked in is_int (device, n bytes).
    int simplex;
    int equality;
    double rhs;
    const uint8_t* is_int;
};

constexpr int kMaxFastBlocks = 2048;  // persistent workgroups of the batched form (blocks = 0: automatic)
int launch_relax_fast(const FastParams& p, int waves_per_block, int blocks, void* stream);
int launch_candidate_features(int n, int N, const double* X, const double* G, const double* LB,
                              const double* UB, const int* depth, const double* gap, const double* val,
                              const double* pcu, const double* pcd, const double* qdiag, size_t qstride,
                              double gs, const uint8_t* flags, double tol, float* F, void* stream);
int launch_batched_gradient(const double* Qt, const double* lin, double gs, double ls, int n, int N,
                            const double* X, double* G, void* stream);
// Widens byte node bounds (0/1) into double arrays: lb[i] = lb8[i], ub[i] = ub8[i].
int launch_expand_bounds(size_t count, const uint8_t* lb8, const uint8_t* ub8, double* lb, double* ub,
                         void* stream);
// Per node the k best-scored fractional candidates (flags[j] set and
// min(frac, 1 - frac) > tol), score descending with ties to the lower index,
// -1 after the last; nodes with status >= GB_NODE_INTEGRAL_UNCONVERGED get
// none.  out is [N][k]; xout (may be NULL) receives X at each candidate.
int launch_topk_candidates(int n, int N, const double* X, const uint8_t* flags, double tol, const float* score,
                           const uint8_t* status, int k, int* out, double* xout, void* stream);
// Device iterate store.  Gather: x[i] = the iterate at src[i] (bit 0 set:
// clipped to [lb[i], ub[i]]), or lb[i] when src[i] is 0.  Scatter: node i's
// x to the address dst[i] (0: skipped).  src, dst are [N] device addresses.
int launch_gather_starts(int n, int N, const unsigned long long* src, const double* lb, const double* ub, double* x,
                         void* stream);
int launch_scatter_iterates(int n, int N, const unsigned long long* dst, const double* x, void* stream);

}  // namespace gb
