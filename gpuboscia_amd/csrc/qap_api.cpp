// C-ABI of the FP32 QAP form (relax_qap32.hip): instance preparation, the
// batch relaxation and the warm-started assignment sequence.
#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>
#include <string>

#include "gb_host.h"
#include "gb_internal.h"
#include "gb_qap32.h"

struct gb_qap32 {
    int m = 0;
    int device = 0;
    const double* A64 = nullptr;  // caller-owned instance (FP64)
    const double* B64 = nullptr;
    float* A32 = nullptr;  // library-owned FP32 copies, A'A, B B'
    float* B32 = nullptr;
    float* AtA = nullptr;
    float* BBt = nullptr;
};

namespace gb {

static gb_status qap32_prepare(const gb_objective_desc* o, void* stream, gb_qap32** out) {
    if (!o || !out) return fail(GB_ERR_INVALID_ARGUMENT, "null argument");
    if (o->kind != GB_OBJ_QAP || o->order <= 0 || o->n != o->order * o->order || !o->A || !o->B)
        return fail(GB_ERR_INVALID_ARGUMENT, "the FP32 QAP form needs a QAP objective (order, A, B, n = order^2)");
    if (o->order > kQapMaxOrder)
        return fail(GB_ERR_INVALID_ARGUMENT, "the FP32 QAP form supports orders up to 256");
    const int m = o->order;
    gb_qap32* q = new gb_qap32();
    q->m = m;
    q->A64 = o->A;
    q->B64 = o->B;
    hipError_t e = hipGetDevice(&q->device);
    float* buf = nullptr;
    if (e == hipSuccess) e = hipMalloc(&buf, (size_t)4 * m * m * sizeof(float));
    if (e != hipSuccess) {
        delete q;
        return hip_fail(e, __FILE__, __LINE__);
    }
    q->A32 = buf;
    q->B32 = buf + (size_t)m * m;
    q->AtA = buf + (size_t)2 * m * m;
    q->BBt = buf + (size_t)3 * m * m;
    const int k = launch_qap32_prepare(m, o->A, o->B, q->A32, q->B32, q->AtA, q->BBt, stream);
    if (k != 0) {
        (void)hipFree(buf);
        delete q;
        return hip_fail(static_cast<hipError_t>(k), __FILE__, __LINE__);
    }
    *out = q;
    return GB_OK;
}

gb_status relax_batch_qap32(gb_qap32* q, const gb_lmo_desc* l, int N, int max_iters, double* x, const double* lb,
                            const double* ub, const double* eps, int* iterations, uint8_t* status, double* gap,
                            double* value, double* best_vertex, double* best_vertex_val, const uint8_t* flags,
                            double ctol, double stop_value, int refresh, void* stream) {
    if (!q || !l) return fail(GB_ERR_INVALID_ARGUMENT, "null argument");
    if (l->kind != GB_LMO_BIRKHOFF || l->order != q->m)
        return fail(GB_ERR_INVALID_ARGUMENT, "the FP32 QAP form relaxes over the Birkhoff polytope of the instance's order");
    if (N < 0 || max_iters < 0) return fail(GB_ERR_INVALID_ARGUMENT, "N and max_iters must be non-negative");
    if (N == 0) return GB_OK;
    if (!x || !lb || !ub || !eps || !iterations || !status || !gap || !value || !best_vertex || !best_vertex_val)
        return fail(GB_ERR_INVALID_ARGUMENT, "null node buffer");
    int dev = 0;
    GB_HIP(hipGetDevice(&dev));
    if (dev != q->device) return fail(GB_ERR_INVALID_ARGUMENT, "the prepared QAP instance lives on another device");
    QapParams P{};
    P.m = q->m;
    P.N = N;
    P.A = q->A32;
    P.B = q->B32;
    P.AtA = q->AtA;
    P.BBt = q->BBt;
    P.A64 = q->A64;
    P.B64 = q->B64;
    P.x = x;
    P.lb = lb;
    P.ub = ub;
    P.eps = eps;
    P.iters = iterations;
    P.status = status;
    P.gap = gap;
    P.value = value;
    P.best_v = best_vertex;
    P.best_val = best_vertex_val;
    P.flags = flags;
    P.ctol = ctol;
    P.stop_value = stop_value;
    P.max_iters = max_iters;
    P.refresh = refresh > 0 ? refresh : 64;
    P.NL = qap_node_layout(q->m);
    P.LL = qap_lds(q->m, kMaxLdsBytes);
    P.node_bytes = P.NL.total;
    double* ws = nullptr;
    gb_status st = workspace_get(P.node_bytes * (size_t)N, stream, &ws);
    if (st != GB_OK) return st;
    P.work = reinterpret_cast<unsigned char*>(ws);
    const int e = launch_qap32_relax(P, stream);
    if (e != 0) return hip_fail(static_cast<hipError_t>(e), __FILE__, __LINE__);
    return GB_OK;
}

}  // namespace gb

using namespace gb;

extern "C" gb_status gb_qap32_prepare(const gb_objective_desc* objective, void* stream, gb_qap32** out) {
    try {
        return qap32_prepare(objective, stream, out);
    } catch (const std::exception& ex) {
        return fail(GB_ERR_RUNTIME, ex.what());
    }
}

extern "C" gb_status gb_qap32_destroy(gb_qap32* q) {
    if (!q) return GB_OK;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(q->device);
    (void)hipDeviceSynchronize();
    (void)hipFree(q->A32);
    (void)hipSetDevice(cur);
    delete q;
    return GB_OK;
}

extern "C" gb_status gb_relax_batch_qap32(gb_qap32* q, const gb_lmo_desc* lmo, int N, int max_iters, double* x,
                                          const double* lb, const double* ub, const double* eps, int* iterations,
                                          uint8_t* status, double* gap, double* value, double* best_vertex,
                                          double* best_vertex_val, const uint8_t* integer_flags,
                                          double candidate_tol, double stop_value, int refresh, void* stream) {
    try {
        return relax_batch_qap32(q, lmo, N, max_iters, x, lb, ub, eps, iterations, status, gap, value, best_vertex,
                                 best_vertex_val, integer_flags, candidate_tol, stop_value, refresh, stream);
    } catch (const std::exception& ex) {
        return fail(GB_ERR_RUNTIME, ex.what());
    }
}

extern "C" gb_status gb_qap32_assignments(int m, int T, const float* costs, int* perms, double* bounds,
                                          void* stream) {
    try {
        if (m <= 0 || m > kQapMaxOrder || T < 0) return fail(GB_ERR_INVALID_ARGUMENT, "need 0 < m <= 256, T >= 0");
        if (T == 0) return GB_OK;
        if (!costs || !perms || !bounds) return fail(GB_ERR_INVALID_ARGUMENT, "null argument");
        double* vs = nullptr;
        GB_HIP(hipMalloc(&vs, (size_t)(m + 1) * (sizeof(double) + sizeof(int)) + 64));
        int* ps = reinterpret_cast<int*>(vs + (m + 1));
        const int e = launch_qap32_assignments(m, T, costs, perms, bounds, vs, ps, kMaxLdsBytes, stream);
        (void)hipFree(vs);
        if (e != 0) return hip_fail(static_cast<hipError_t>(e), __FILE__, __LINE__);
        return GB_OK;
    } catch (const std::exception& ex) {
        return fail(GB_ERR_RUNTIME, ex.what());
    }
}
