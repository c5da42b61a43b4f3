// FP32 form of the QAP relaxation (relax_qap32.hip): iterates, gradient and
// assignment costs in FP32, every certificate in FP64.  Shared by the
// kernels and the host entry points (qap_api.cpp).
#pragma once
#include <stddef.h>
#include <stdint.h>

namespace gb {

constexpr int kQapThreads = 1024;   // one workgroup of 16 waves per node
constexpr int kQapMaxOrder = 256;   // four columns per lane in the assignment solver
constexpr int kQapChunk = 32;       // rows of a pairwise direction staged per gradient pass

// Per-node workspace, in bytes from the node's base (256-byte aligned parts).
struct QapNodeLayout {
    size_t P, G, C, R64, G64, mask, sv, sa, iv, ia, S, vF, pF, vA, pA, total;
};

inline QapNodeLayout qap_node_layout(int m) {
    QapNodeLayout L{};
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = (o + bytes + 255) & ~size_t(255);
        return at;
    };
    const size_t mm = (size_t)m * m;
    L.P = take(mm * 4);       // iterate (FP32)
    L.G = take(mm * 4);       // gradient (FP32), kept by pairwise updates
    L.C = take(mm * 4);       // costs when they do not fit in LDS; FP32 residual of a refresh
    L.R64 = take(mm * 8);     // FP64 residual A P - P B of the certificate
    L.G64 = take(mm * 8);     // FP64 gradient of the certificate
    L.mask = take(mm);        // 1 = edge allowed by the node's bounds
    L.sv = take((size_t)m * 4);   // forward vertex, row -> column
    L.sa = take((size_t)m * 4);   // away vertex
    L.iv = take((size_t)m * 4);   // inverses (column -> row)
    L.ia = take((size_t)m * 4);
    L.S = take((size_t)m * 4);    // rows where the two vertices differ
    L.vF = take((size_t)(m + 1) * 8);  // warm state of the forward assignment: column potentials
    L.pF = take((size_t)(m + 1) * 4);  //   and column -> row (1-based, 0 = free)
    L.vA = take((size_t)(m + 1) * 8);  // warm state of the in-face assignment
    L.pA = take((size_t)(m + 1) * 4);
    L.total = o;
    return L;
}

// LDS of the node kernel (bytes): the staged costs (or the gradient pass's
// chunk), column potentials, row potentials, row marks and reductions.
struct QapLds {
    size_t cost, vsh, urow, rowm, red, total;
    int staged;  // 1: the m x m costs live in LDS
};

inline QapLds qap_lds(int m, size_t limit) {
    QapLds L{};
    const size_t mm4 = (size_t)m * m * 4;
    // gradient pass: four [kQapChunk][m rounded up to 4] float tiles, then
    // the two vertices and their inverses (4 m ints)
    const size_t chunk = (size_t)4 * kQapChunk * ((m + 3) & ~3) * 4 + (size_t)4 * m * 4;
    auto al = [](size_t v) { return (v + 15) & ~size_t(15); };
    const size_t aux = al((size_t)(m + 1) * 8) * 2 + al((size_t)m + 1) + al(16 * 8 + 16 * 4);
    L.staged = al(mm4) + aux <= limit ? 1 : 0;
    size_t o = 0;
    L.cost = o;
    o += al(L.staged ? (mm4 > chunk ? mm4 : chunk) : chunk);
    L.vsh = o;
    o += al((size_t)(m + 1) * 8);
    L.urow = o;
    o += al((size_t)(m + 1) * 8);
    L.rowm = o;
    o += al((size_t)m + 1);
    L.red = o;  // 16 doubles, then 16 ints
    o += al(16 * 8 + 16 * 4);
    L.total = o;
    return L;
}

struct QapParams {
    int m, N;
    const float* A;      // [m][m] FP32 copies of the instance
    const float* B;
    const float* AtA;    // A'A and B B' (FP32, from FP64 products)
    const float* BBt;
    const double* A64;   // the instance in FP64 (certificates, vertex values)
    const double* B64;
    // node buffers, FP64 interface as the other forms ([N][m*m] node-major)
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
    double stop_value;   // a vertex at or below this ends the node (its bound is known)
    int max_iters;
    int refresh;         // gradient recomputed from P every `refresh` pairwise steps
    unsigned char* work; // [N][node layout]
    size_t node_bytes;
    QapNodeLayout NL;
    QapLds LL;
};

// Launches (relax_qap32.hip); each returns a hipError_t value.
int launch_qap32_prepare(int m, const double* A, const double* B, float* A32, float* B32, float* AtA, float* BBt,
                         void* stream);
int launch_qap32_relax(const QapParams& p, void* stream);
// T assignment problems solved in sequence by one workgroup, each warm-started
// from the previous one's potentials and matching: costs [T][m][m] FP32
// (+inf = forbidden), perms [T][m] (row -> column, -1 if infeasible),
// bounds [T] (FP64 Lagrangian bound of each solve's potentials).
int launch_qap32_assignments(int m, int T, const float* costs, int* perms, double* bounds, double* vstate,
                             int* pstate, size_t lds_limit, void* stream);

}  // namespace gb
