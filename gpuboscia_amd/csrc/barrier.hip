// Self-test of the device-atomic group barrier (gb_barrier.h), modelled on
// the reference's barrier stress test (tests/test_multiblock.cu:8-31):
// groups of k co-resident workgroups sync independently, groups run
// different numbers of generations and exit independently, and every group
// but the first waits until group 0 has completed 100 generations before its
// first sync (a grid-wide phase barrier would deadlock there).  In each
// generation every thread publishes a value and then checks the sum of its
// group's values, so a barrier that published only the arriving thread's
// writes is caught.
#include <hip/hip_runtime.h>

#include <chrono>
#include <string>

#include "gb_barrier.h"
#include "gb_host.h"

namespace gb {
namespace {

constexpr unsigned int kHeadStart = 100u;

__global__ void barrier_stress_kernel(GroupBarrier* bars, unsigned int* slots,
                                      unsigned long long* errors, unsigned int* progress,
                                      BarrierStatus* st, int k, int phases_base, int phase_step,
                                      unsigned long long spin_limit) {
    __shared__ int flag;
    const int group = blockIdx.x / k;
    const int rank = blockIdx.x % k;
    GroupBarrier* b = bars + group;
    if (group != 0 && threadIdx.x == 0) wait_at_least(progress, kHeadStart, spin_limit, st);
    if (!group_sync(b, k, spin_limit, st, &flag)) return;
    const unsigned int phases = static_cast<unsigned int>(phases_base + group * phase_step);
    unsigned int* mine = slots + (static_cast<size_t>(group) * k + rank) * blockDim.x + threadIdx.x;
    for (unsigned int phase = 1; phase <= phases; ++phase) {
        *mine = phase + threadIdx.x;
        if (!group_sync(b, k, spin_limit, st, &flag)) return;
        unsigned int sum = 0;
        for (int r = 0; r < k; ++r)
            sum += slots[(static_cast<size_t>(group) * k + r) * blockDim.x + threadIdx.x];
        if (sum != static_cast<unsigned int>(k) * (phase + threadIdx.x))
            __hip_atomic_fetch_add(errors, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!group_sync(b, k, spin_limit, st, &flag)) return;
        if (group == 0 && rank == 0 && threadIdx.x == 0)
            __hip_atomic_store(progress, phase, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

}  // namespace
}  // namespace gb

using namespace gb;

extern "C" gb_status gb_grid_barrier_selftest(int groups, int group_blocks, int threads,
                                              int phases_base, int phase_step,
                                              long long spin_limit, long long* errors,
                                              long long* timeouts, double* seconds_per_sync) {
    if (groups < 1 || group_blocks < 1 || threads < 64 || threads > 1024 || threads % 64 != 0
            || phase_step < 0 || spin_limit < 1 || !errors || !timeouts)
        return fail(GB_ERR_INVALID_ARGUMENT, "grid barrier self-test: invalid arguments");
    if (groups > 1 && phases_base < static_cast<int>(kHeadStart))
        return fail(GB_ERR_INVALID_ARGUMENT,
                    "grid barrier self-test: phases_base must be >= 100 when groups > 1");
    const long long total = static_cast<long long>(groups) * group_blocks;
    int dev = 0, cus = 0, coop = 0, per_cu = 0;
    GB_HIP(hipGetDevice(&dev));
    GB_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    GB_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    GB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(barrier_stress_kernel), threads, 0));
    if (!coop) return fail(GB_ERR_RUNTIME, "grid barrier self-test: no cooperative launch support");
    if (total > static_cast<long long>(per_cu) * cus)
        return fail(GB_ERR_INVALID_ARGUMENT,
                    "grid barrier self-test: " + std::to_string(total) +
                        " workgroups cannot be co-resident (limit " +
                        std::to_string(static_cast<long long>(per_cu) * cus) + ")");
    GroupBarrier* bars = nullptr;
    unsigned int* slots = nullptr;
    unsigned long long* err = nullptr;
    unsigned int* progress = nullptr;
    BarrierStatus* st = nullptr;
    gb_status rc = GB_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        (void)hipFree(bars);
        (void)hipFree(slots);
        (void)hipFree(err);
        (void)hipFree(progress);
        (void)hipFree(st);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    auto check = [&](hipError_t e, int line) {
        if (e != hipSuccess && rc == GB_OK) rc = hip_fail(e, __FILE__, line);
        return e == hipSuccess;
    };
    const size_t slot_count = static_cast<size_t>(total) * threads;
    if (!check(hipMalloc(reinterpret_cast<void**>(&bars), sizeof(GroupBarrier) * groups), __LINE__)
            || !check(hipMalloc(reinterpret_cast<void**>(&slots), slot_count * 4), __LINE__)
            || !check(hipMalloc(reinterpret_cast<void**>(&err), 8), __LINE__)
            || !check(hipMalloc(reinterpret_cast<void**>(&progress), 4), __LINE__)
            || !check(hipMalloc(reinterpret_cast<void**>(&st), sizeof(BarrierStatus)), __LINE__)
            || !check(hipMemset(bars, 0, sizeof(GroupBarrier) * groups), __LINE__)
            || !check(hipMemset(slots, 0, slot_count * 4), __LINE__)
            || !check(hipMemset(err, 0, 8), __LINE__)
            || !check(hipMemset(progress, 0, 4), __LINE__)
            || !check(hipMemset(st, 0, sizeof(BarrierStatus)), __LINE__)
            || !check(hipEventCreate(&e0), __LINE__) || !check(hipEventCreate(&e1), __LINE__)) {
        cleanup();
        return rc;
    }
    int k = group_blocks;
    unsigned long long limit = static_cast<unsigned long long>(spin_limit);
    void* args[] = {&bars, &slots, &err, &progress, &st, &k, &phases_base, &phase_step, &limit};
    if (check(hipEventRecord(e0, nullptr), __LINE__)
            && check(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(barrier_stress_kernel),
                                                dim3(static_cast<unsigned>(total)), dim3(threads), args,
                                                0, nullptr),
                     __LINE__)
            && check(hipEventRecord(e1, nullptr), __LINE__) && check(hipEventSynchronize(e1), __LINE__)) {
        unsigned long long herr = 0;
        BarrierStatus hst{};
        float ms = 0.f;
        if (check(hipMemcpy(&herr, err, 8, hipMemcpyDeviceToHost), __LINE__)
                && check(hipMemcpy(&hst, st, sizeof(hst), hipMemcpyDeviceToHost), __LINE__)
                && check(hipEventElapsedTime(&ms, e0, e1), __LINE__)) {
            *errors = static_cast<long long>(herr);
            *timeouts = static_cast<long long>(hst.timeouts);
            // Group 0 performs 2 * phases_base + 1 syncs.
            if (seconds_per_sync) *seconds_per_sync = ms * 1e-3 / (2.0 * phases_base + 1.0);
        }
    }
    cleanup();
    return rc;
}
