// Mixed-precision batched relaxation (relax_mixed.hip): parameters shared by
// the host launcher and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gpuboscia_amd.h"

namespace gb {

constexpr int kMixedMaxN = 512;   // 32 row tiles of 16: x, Qx, Qv, Qa in registers
// Up to 16 row tiles any tile shape; 17..32 row tiles run 4 row tiles per
// wave (at most 8 waves) with the fragments streamed from L2.
constexpr int kMixedSmallRowTiles = 16;
constexpr int kMixedMaxPlanes = 3;

// Q in MFMA A-fragment order, split into bf16 planes q = q0 + q1 + q2:
// fragment ((p * RT + t) * S + s) * 64 + lane holds 8 bf16 values
// Q[16 t + lane % 16][16 (2 s + (i >> 2)) + 4 (lane / 16) + (i & 3)], i < 8,
// so lane l's B-operand slots of k-step s are exactly the rows it owns in
// the output tiles 2s and 2s+1 (no cross-lane exchange of the vertex).
struct MixedQ {
    int n = 0;
    int rt = 0;        // row tiles (even); n_pad = 16 rt
    int planes = 0;
    uint4* frag = nullptr;
    float* lsc = nullptr;   // s_g s_l c, n_pad floats (0 beyond n)
    int device = 0;
};

struct MixedParams {
    int n, rt, planes;
    int cols;               // node columns per 16-column tile (1..16)
    const uint4* frag;
    const float* lsc;
    float gs;
    double kappa;
    int N, max_iters;
    double* x;
    const double* lb;
    const double* ub;
    const double* eps;
    int* iters;
    uint8_t* status;
    double* best_v;
    double* best_val;
    const double* cutoff;
    int* counter;
    float rel_floor;        // stop once gap < max(eps, rel_floor (1 + |f|))
    long long* stats;       // optional [N]: iterations of the mixed phase
    // optional [N][n]: Q x of the nodes with a supplied point and no
    // iterations (launch_mixed_warm_products); such nodes start at x itself
    const double* qx0;
    const int* count;       // optional device batch size: nodes [0, *count) of the N rows are relaxed
    // Bounded simplex (simplex != 0): sum x <= rhs (== rhs when equality)
    // over 0/1 variables all marked in is_int (device, n bytes), DICG only.
    int simplex;
    int equality;
    double rhs;
    const uint8_t* is_int;
};

int launch_mixed_prepare(const double* Qt, const double* lin, double gs, double ls, int n, int rt,
                         int planes, uint4* frag, float* lsc, void* stream);
// LDS of one tile (W waves of rw row tiles each):
//  vops, aops  B operands of the vertices v and a: 16-byte slot
//              (s * 16 + column) * 4 + q holds rows 4q .. 4q+3 of tiles 2s
//              and 2s+1 as 8 bf16 (1 KB per k-step);
//  xch         per wave, Q v and Q a of its tiles as [array][j][column][row]
//              floats (2 rw KB per wave);
//  red         per-column partial sums of each wave, mixed_red(sx) floats;
//  fetch, ok   fetched node ids, per-wave scope flags;
//  sel         (simplex) per column: the gradient of every row (float), the
//              row flags (byte), the rows' v and a (a byte each) and the
//              column's feasibility (int);
//  qf          (optional) the Q fragments, planes x rt x rt/2 KB.
// At most 8 + 8 + 32 + 8 KB for n <= 256 and 16 + 16 + 64 + 4 KB for n <= 512,
// plus the fragments when copied.
struct MixedLds {
    size_t vops, aops, xch, red, fetch, ok, sg, sf, sr, sok, qf, total;
};
// reduction slots per (wave, column): 8, or 12 with the simplex's mass sums
__host__ __device__ constexpr int mixed_red(bool sx) { return sx ? 12 : 8; }
constexpr size_t kMixedLdsBudget = 160 * 1024;  // LDS of one CU
__host__ __device__ inline size_t mixed_q_bytes(int rt, int planes) {
    return static_cast<size_t>(planes) * rt * (rt / 2) * 1024;
}
__host__ __device__ inline MixedLds mixed_lds_layout(int rt, int waves, int rw, size_t q_bytes, bool sx = false) {
    MixedLds L{};
    size_t o = 0;
    L.vops = o;
    o += 1024 * static_cast<size_t>(rt / 2);
    L.aops = o;
    o += 1024 * static_cast<size_t>(rt / 2);
    L.xch = o;
    o += sizeof(float) * 2 * 256 * static_cast<size_t>(rw) * waves;
    L.red = o;
    o += sizeof(float) * mixed_red(sx) * 16 * static_cast<size_t>(waves);
    L.fetch = o;
    o += sizeof(int) * 16;
    L.ok = o;
    o += sizeof(int) * static_cast<size_t>(waves) * 16;
    const size_t np = 16 * static_cast<size_t>(rt);
    L.sg = o;
    o += sx ? sizeof(float) * 16 * np : 0;
    L.sf = o;
    o += sx ? 16 * np : 0;
    L.sr = o;  // selection results: v bytes of every column, then a bytes
    o += sx ? 2 * 16 * np : 0;
    L.sok = o;
    o += sx ? sizeof(int) * 16 : 0;
    o = (o + 15) & ~static_cast<size_t>(15);
    L.qf = o;
    o += q_bytes;
    L.total = o;
    return L;
}

// Rows per lane and tile in the element layout: 1, 2, 4 for up to 4, 8, 16
// node columns per tile.
inline int mixed_rows_per_lane(int cols) { return cols <= 4 ? 1 : (cols <= 8 ? 2 : 4); }

constexpr int kMixedMaxRowTilesPerWave = 4;

struct MixedShape {
    int rw = 2;       // row tiles per wave
    int cols = 16;    // node columns per tile
    int blocks = 0;   // tiles launched
};
MixedShape mixed_default_shape(int rt, int N);

// rw: row tiles per wave (1..4); the tile runs ceil(rt / rw) waves.
int launch_relax_mixed(const MixedParams& p, bool dicg, int rw, int blocks, void* stream);
// Shapes whose pass 1 starts a node with a supplied point at that point
// (the others start at the box vertex nearest to it).
constexpr bool mixed_warm_start(int rw, bool wide) { return !wide && rw != 3; }
// out[i] = Q x[i] in FP64 (k-ordered FMA chains over Qt) for every node i
// with status GB_NODE_INITIAL_POINT_SUPPLIED and iters[i] == 0; other rows
// of out are left as they are.  n <= kMixedMaxN.
int launch_mixed_warm_products(const double* Qt, int n, int N, const double* X, const uint8_t* status,
                               const int* iters, double* out, void* stream, const int* count = nullptr);
int launch_mixed_products(const MixedQ& q, const double* V, int N, double* Y, void* stream);

}  // namespace gb
