// Device helpers shared by the gfx950 kernels: canonical wave64 reductions,
// wave-local synchronisation, block argmin.
//
// Canonical reduction order (identical in the CPU oracle): lane l of a block
// of T threads accumulates coordinates j = l, l+T, ... in increasing j from
// 0.0; each wave64 reduces with shuffle-down offsets 32,16,8,4,2,1; wave
// totals are added in wave order.  All kernels are compiled with
// -ffp-contract=off and use fma() exactly where a product is accumulated.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gb {

constexpr int kWave = 64;
constexpr int kMaxWaves = 16;          // up to 1024 threads per block
constexpr double kNone = 1e300;        // "no candidate" score

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }
__device__ __forceinline__ int wave_id() { return threadIdx.x / kWave; }
__device__ __forceinline__ int wave_count() { return blockDim.x / kWave; }

// Orders LDS traffic of the lanes of one wave (the wave runs in lockstep).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ double wave_tree_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = v + __shfl_down(v, off, kWave);
    return v;  // valid in lane 0
}

__device__ __forceinline__ double wave_tree_min(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_down(v, off, kWave));
    return v;  // valid in lane 0
}

// Block sum in canonical order; every thread receives the total.
// `red` is an LDS scratch of kMaxWaves doubles.
__device__ __forceinline__ double block_sum(double v, double* red) {
    v = wave_tree_sum(v);
    if (lane_id() == 0) red[wave_id()] = v;
    __syncthreads();
    double t = red[0];
    const int W = wave_count();
    for (int w = 1; w < W; ++w) t = t + red[w];
    __syncthreads();
    return t;
}

__device__ __forceinline__ double block_min(double v, double* red) {
    v = wave_tree_min(v);
    if (lane_id() == 0) red[wave_id()] = v;
    __syncthreads();
    double t = red[0];
    const int W = wave_count();
    for (int w = 1; w < W; ++w) t = fmin(t, red[w]);
    __syncthreads();
    return t;
}

__device__ __forceinline__ bool block_and(bool p, int* ired) {
    const unsigned long long b = __ballot(!p);
    if (lane_id() == 0) ired[wave_id()] = (b != 0ull);
    __syncthreads();
    bool all = true;
    const int W = wave_count();
    for (int w = 0; w < W; ++w) all = all && (ired[w] == 0);
    __syncthreads();
    return all;
}

__device__ __forceinline__ bool block_or(bool p, int* ired) {
    return !block_and(!p, ired);
}

// (value, index) minimum; ties and invalid entries (index < 0) resolved as
// the reference's better_pair (lmo.cuh:117-128): lower value, then lower
// index.  Exact, hence independent of the reduction shape.
struct MinPair { double v; int i; };

__device__ __forceinline__ MinPair better(MinPair a, MinPair b) {
    if (b.i < 0) return a;
    if (a.i < 0 || b.v < a.v || (b.v == a.v && b.i < a.i)) return b;
    return a;
}

__device__ __forceinline__ MinPair wave_argmin(MinPair p) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        MinPair o{__shfl_xor(p.v, off, kWave), __shfl_xor(p.i, off, kWave)};
        p = better(p, o);
    }
    return p;  // valid in every lane
}

}  // namespace gb
