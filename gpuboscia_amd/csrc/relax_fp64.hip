// Launches of the plain FP64 node-relaxation kernel with the built-in
// objectives, LMOs and heuristics.  The kernel itself is the template of
// include/gpuboscia/device/relax_fp64.hpp, which device plug-ins
// instantiate with their own types.
#include <mutex>

#include "gb_internal.h"
#include "../../include/gpuboscia/device/relax_fp64.hpp"

namespace gb {

// Work vectors of the allocation-free entry point (gb_relax_batch_noalloc):
// one file-scope device array per GPU, held for the life of the process.
// Host code never passes its name to a kernel (on the host the name is a
// stand-in, not a device address): static_work_address looks the address up
// once per device and every launch uses that.
static __device__ __align__(256) double scratch[GB_NOALLOC_WORK_BYTES / sizeof(double)];

int static_work_address(double** out) {
    constexpr int kDevices = 64;
    static std::mutex mu;
    static double* addr[kDevices] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= kDevices) return (int)hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        addr[dev] = static_cast<double*>(p);
    }
    *out = addr[dev];
    return (int)hipSuccess;
}

int launch_lmo_probe(const RelaxParams& p, int block_threads, size_t lds_bytes, const LdsLayout& L,
                     const double* G, void* stream) {
#define GB_LAUNCH(LMO)                                                                       \
    do {                                                                                     \
        return launch_lmo_probe_with(p, block_threads, lds_bytes, L, G, Lmo<LMO>{}, stream); \
    } while (0)
    if (p.lmo_kind == GB_LMO_BOX) GB_LAUNCH(GB_LMO_BOX);
    if (p.lmo_kind == GB_LMO_SIMPLEX) GB_LAUNCH(GB_LMO_SIMPLEX);
    if (p.lmo_kind == GB_LMO_BIRKHOFF) GB_LAUNCH(GB_LMO_BIRKHOFF);
    if (p.lmo_kind == GB_LMO_PORTFOLIO) GB_LAUNCH(GB_LMO_PORTFOLIO);
#undef GB_LAUNCH
    return (int)hipErrorInvalidValue;
}

int launch_relax_fp64(const RelaxParams& p, int block_threads, size_t lds_bytes, const LdsLayout& L,
                      void* stream) {
    const dim3 grid(p.N), block(block_threads);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
#define GB_LAUNCH(OBJ, LMO)                                                                   \
    do {                                                                                      \
        auto k = relax_fp64_kernel<Obj<OBJ>, Lmo<LMO>, BuiltinHeuristic>;                     \
        if (lds_bytes > 64 * 1024) {                                                          \
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),              \
                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);                  \
            if (e != hipSuccess) return (int)e;                                               \
        }                                                                                     \
        hipLaunchKernelGGL(k, grid, block, lds_bytes, st, p, L, Obj<OBJ>{}, Lmo<LMO>{},       \
                           BuiltinHeuristic{});                                               \
        return (int)hipGetLastError();                                                        \
    } while (0)
    if (p.obj_kind == GB_OBJ_DENSE_QUADRATIC) {
        if (p.lmo_kind == GB_LMO_BOX) GB_LAUNCH(GB_OBJ_DENSE_QUADRATIC, GB_LMO_BOX);
        if (p.lmo_kind == GB_LMO_SIMPLEX) GB_LAUNCH(GB_OBJ_DENSE_QUADRATIC, GB_LMO_SIMPLEX);
        if (p.lmo_kind == GB_LMO_BIRKHOFF) GB_LAUNCH(GB_OBJ_DENSE_QUADRATIC, GB_LMO_BIRKHOFF);
        if (p.lmo_kind == GB_LMO_PORTFOLIO) GB_LAUNCH(GB_OBJ_DENSE_QUADRATIC, GB_LMO_PORTFOLIO);
    } else {
        if (p.lmo_kind == GB_LMO_BOX) GB_LAUNCH(GB_OBJ_QAP, GB_LMO_BOX);
        if (p.lmo_kind == GB_LMO_SIMPLEX) GB_LAUNCH(GB_OBJ_QAP, GB_LMO_SIMPLEX);
        if (p.lmo_kind == GB_LMO_BIRKHOFF) GB_LAUNCH(GB_OBJ_QAP, GB_LMO_BIRKHOFF);
    }
#undef GB_LAUNCH
    return (int)hipErrorInvalidValue;
}

}  // namespace gb
