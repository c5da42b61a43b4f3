// Internal declarations shared by the host library sources and kernels.
#pragma once
#include <stdint.h>
#include <stddef.h>

#include "../../include/gpuboscia_amd.h"

namespace gb {

// Flattened launch parameters of the FP64 node-relaxation kernel.  Mirrors
// the argument list of drs_fw_relax_batch_kernel
// (device_resident_solver.cuh:634-654) with the functors replaced by plain
// descriptors.
struct RelaxParams {
    // objective
    int obj_kind;
    int n;
    const double* Qt;
    const double* lin;
    double gs, ls, kappa;
    int order;
    const double* A;
    const double* B;
    // lmo
    int lmo_kind;
    double rhs;
    int equality;
    const uint8_t* is_int;
    int lorder;
    double pcap, pcard;     // portfolio LMO
    // algorithm
    int dicg;
    int ls_kind;
    int secant_iters;
    double secant_tol;
    double secant_denom;
    int heuristic;
    double heuristic_offset;
    // node buffers ([N][n] node-major)
    int N;
    int max_iters;
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
    double* work;          // [N][kWorkVectors][n]
};

constexpr int kWorkVectors = 6;  // g, v, a, trial_x, trial_g, qap residual

// LDS layout of the FP64 kernel (bytes; offsets 16-byte aligned).
struct LdsLayout {
    size_t red, ired, vec, sg, sidx, scap, spre, elig, fixed;
    size_t fcol, frow, rows, cols, hu, hp, way, misc;
    size_t hcost, hcost_bytes;  // staged residual Hungarian costs (0 bytes: not staged)
    size_t total;
};

inline size_t lds_align(size_t v) { return (v + 15) & ~size_t(15); }

inline int next_pow2_host(int v) { int r = 1; while (r < v) r <<= 1; return r; }

inline LdsLayout lds_layout(int obj_kind, int lmo_kind, int n, int order) {
    LdsLayout L{};
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off = lds_align(off + bytes); return o; };
    L.red = take(16 * sizeof(double));
    L.ired = take(16 * sizeof(int));
    L.vec = take(obj_kind == GB_OBJ_DENSE_QUADRATIC ? (size_t)n * sizeof(double) : 0);
    if (lmo_kind == GB_LMO_SIMPLEX || lmo_kind == GB_LMO_PORTFOLIO) {
        const int S = next_pow2_host(n);
        L.sg = take((size_t)n * sizeof(double));
        L.sidx = take((size_t)S * sizeof(int));
        L.scap = take((size_t)S * sizeof(double));
        L.spre = take((size_t)S * sizeof(double));
        L.elig = take((size_t)n);
        L.fixed = take((size_t)n);
    }
    if (lmo_kind == GB_LMO_BIRKHOFF) {
        const int m = order;
        L.fcol = take((size_t)m * sizeof(int));
        L.frow = take((size_t)m * sizeof(int));
        L.rows = take((size_t)m * sizeof(int));
        L.cols = take((size_t)m * sizeof(int));
        L.hu = take((size_t)(m + 1) * sizeof(double));
        L.hp = take((size_t)(m + 1) * sizeof(int));
        L.way = take((size_t)(m + 1) * sizeof(int));
        L.misc = take(8 * sizeof(int));
        // The residual cost matrix (at most m x m) in LDS when it fits
        // beside the rest of the layout in one CU's 160 KB.
        const size_t cb = (size_t)m * m * sizeof(double);
        if (off + cb <= 160 * 1024) {
            L.hcost = take(cb);
            L.hcost_bytes = cb;
        }
    }
    L.total = off;
    return L;
}

// Address, on the current device, of the static device array that holds the
// work vectors of the allocation-free relaxation (GB_NOALLOC_WORK_BYTES;
// relax_fp64.hip).  Looked up once per device.  Returns a hipError_t value.
int static_work_address(double** out);

// Extreme points of the built-in LMOs for directions G ([N][n]) into p.x,
// with bounds p.lb / p.ub (relax_fp64.hip).  Returns a hipError_t value.
int launch_lmo_probe(const RelaxParams& p, int block_threads, size_t lds_bytes,
                     const LdsLayout& L, const double* G, void* stream);

// Multi-workgroup relaxation (relax_multiblock.hip): K workgroups of
// kMbThreads threads per node, nodes [node0, node0 + nodes) per launch.
struct GroupBarrier;
struct BarrierStatus;
struct MbParams {
    int K;                   // workgroups per node
    int Tv;                  // lanes of the canonical sums (the one-workgroup block size)
    int node0;               // first node of this launch
    GroupBarrier* bars;      // [nodes per launch]
    BarrierStatus* bst;
    unsigned long long spin;
    double* red;             // [nodes per launch][3][kMbSlots]
    double* work;            // [nodes per launch][6][n]: g, v, a, trial x, trial g, product
    unsigned char* lwork;    // [nodes per launch][lwork_stride]: the LMO's own vectors
    size_t lwork_stride;
};
constexpr int kMbThreads = 256;
constexpr int kMbSlots = 16 + 256;   // wave totals (Tv <= 1024), then one slot per workgroup (K <= 256)
constexpr int kMbMaxK = 256;
int launch_multiblock_clear(GroupBarrier* bars, int count, BarrierStatus* bst, void* stream);
// Bytes per node of the LMO's vectors (multiblock_lmo_layout), 0 for a box.
size_t multiblock_lmo_bytes(int lmo_kind, int n);
int multiblock_capacity(int lmo_kind);
int launch_relax_multiblock(const RelaxParams& p, const MbParams& m, int nodes, void* stream);

// Launch the FP64 relaxation kernel (relax_fp64.hip).  Returns a hipError_t
// value as int.
int launch_relax_fp64(const RelaxParams& p, int block_threads, size_t lds_bytes,
                      const LdsLayout& L, void* stream);

}  // namespace gb
