// Launch parameters of the FP64 node-relaxation kernel, shared by the
// library and by plug-in translation units that instantiate the kernel
// (gpuboscia/device/relax_fp64.hpp).
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "../../gpuboscia_amd.h"

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
    // plug-in objective workspace: [N][obj_stride] bytes of global memory
    // (the reference's objective_workspace, device_resident_solver.cuh:396-400)
    unsigned char* obj_work;
    size_t obj_stride;
};

constexpr int kWorkVectors = 6;  // g, v, a, trial_x, trial_g, qap residual

// LDS layout of the FP64 kernel (bytes; offsets 16-byte aligned).
struct LdsLayout {
    size_t red, ired, vec, sg, sidx, scap, spre, elig, fixed;
    size_t fcol, frow, rows, cols, hu, hp, way, misc;
    size_t hcost, hcost_bytes;  // staged residual Hungarian costs (0 bytes: not staged)
    size_t total;
};

// What the library hands a plug-in launcher (gb_plugin_desc.launch): the
// filled parameters, the LDS layout and the launch shape of one batch.
struct PluginLaunch {
    RelaxParams P;
    LdsLayout L;
    int threads;      // block size (a multiple of 64)
    size_t lds_bytes; // dynamic LDS per workgroup
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

}  // namespace gb
