// Device-atomic group barrier for persistent kernels.
//
// Replaces the cooperative grid synchronisation of the reference
// (drs_dicg_paired_batch_kernel, device_resident_solver.cuh:764-1000, and
// MultiBlockContext::node_sync, multiblock.cuh:72-102): a group of `count`
// co-resident workgroups meets at a generation counter.  Each arrival is an
// acq_rel fetch_add on `arrived`; the last arrival resets the counter and
// release-stores generation + 1; the others spin on an acquire load of the
// generation.  The workgroup barrier orders the other threads' writes before
// the arriving thread's device-scope release (and its acquire before their
// later reads), so a group sync publishes all threads' writes; the stress
// self-test checks every thread's value each generation.  Waiters poll with
// relaxed loads and acquire once, when the generation has moved.  A fast
// workgroup cannot complete the next generation before every member has
// arrived, so a waiter never sees a wrapped generation.
//
// Every wait is bounded: after `spin_limit` polls a waiter records a timeout,
// raises `abort` and returns false, and every other waiter of the same
// protocol returns false once it sees `abort`.  A kernel using the barrier
// therefore always terminates, even if a launch breaks the co-residency
// assumption.
#pragma once
#include <hip/hip_runtime.h>

namespace gb {

// One barrier per group, on its own 128-byte line.
struct alignas(128) GroupBarrier {
    unsigned int arrived;
    unsigned int generation;
};

struct BarrierStatus {
    unsigned int abort;          // raised on the first timeout
    unsigned int pad0[31];
    unsigned long long timeouts; // polls that ran out
};

__device__ inline bool abort_raised(BarrierStatus* st) {
    return __hip_atomic_load(&st->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
}

__device__ inline void raise_abort(BarrierStatus* st) {
    __hip_atomic_fetch_add(&st->timeouts, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&st->abort, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Bounded spin until *word reaches at least `target` (acquire).  Thread-level:
// call from one thread per workgroup.
__device__ inline bool wait_at_least(const unsigned int* word, unsigned int target,
                                     unsigned long long spin_limit, BarrierStatus* st) {
    unsigned long long polls = 0;
    while (__hip_atomic_load(word, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if (abort_raised(st)) return false;
        if (++polls > spin_limit) {
            raise_abort(st);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    return true;
}

// Group sync across `count` workgroups sharing `b`.  Must be reached by all
// threads of the workgroup (uniform control flow).  `flag` is a workgroup
// LDS word used to broadcast the outcome.  Returns false after a timeout or
// an abort; callers then leave their loops.
__device__ inline bool group_sync(GroupBarrier* b, unsigned int count, unsigned long long spin_limit,
                                  BarrierStatus* st, int* flag) {
    if (count <= 1u) {
        __syncthreads();
        return true;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        bool ok = true;
        const unsigned int gen =
            __hip_atomic_load(&b->generation, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned int prior =
            __hip_atomic_fetch_add(&b->arrived, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (prior == count - 1u) {
            __hip_atomic_store(&b->arrived, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&b->generation, gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            unsigned long long polls = 0;
            while (__hip_atomic_load(&b->generation, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                if (abort_raised(st)) {
                    ok = false;
                    break;
                }
                if (++polls > spin_limit) {
                    raise_abort(st);
                    ok = false;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        *flag = ok ? 1 : 0;
    }
    __syncthreads();
    return *flag != 0;
}

}  // namespace gb
