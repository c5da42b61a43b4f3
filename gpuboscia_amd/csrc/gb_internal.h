// Internal declarations shared by the host library sources and kernels.
#pragma once
#include <stdint.h>
#include <stddef.h>

#include "../../include/gpuboscia_amd.h"
#include "../../include/gpuboscia/device/relax_params.h"

namespace gb {

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
