// Synchronisation of co-resident workgroups (resident kernels): a counting
// barrier in two halves, and per-slab step flags.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.h"

namespace ct {

// Two-level counting barrier over `members` workgroups, in two halves so
// that work independent of the other members can run between them:
// team_arrive counts this workgroup in (after its writes), team_wait returns
// once every member has arrived. Member m first counts into group m % 8, the
// last arrival of a group counts into the top counter, and the last group
// arrival publishes the next generation. Counters only grow within a launch
// (they are zeroed before it), so "last" is recognised from the running
// count and nothing has to be reset between steps. With `coherent` the
// members' rows are agent-coherent stores: each thread waits until its own
// have completed, and the arrival needs no release fence.
__device__ __forceinline__ void team_arrive(GridBarrier *gb, uint32_t members, uint32_t member,
                                            uint32_t &gen, bool coherent) {
    if (coherent) __builtin_amdgcn_s_waitcnt(0);  // this thread's stores have completed
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t next = gen + 1;
        const uint32_t grp = member % kBarrierGroups;
        const uint32_t gsize = (members - grp + kBarrierGroups - 1) / kBarrierGroups;
        const uint32_t ngroups = members < kBarrierGroups ? members : kBarrierGroups;
        // release: this workgroup's state rows reach device-coherent memory
        if (!coherent) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const uint32_t a1 = __hip_atomic_fetch_add(&gb->group[grp].count, 1u, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
        if (a1 == next * gsize - 1) {
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
            const uint32_t a2 = __hip_atomic_fetch_add(&gb->top.count, 1u, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT);
            if (a2 == next * ngroups - 1) {
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
                __hip_atomic_store(&gb->generation.count, next, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
    gen += 1;
}

// Waits (at most `wait_ticks` of wall clock) until generation `gen` is
// published; false if the team aborted or the wait timed out (which raises
// the sticky abort word: `abort`, or the barrier's own when null). With
// `coherent` the rows are read with agent-coherent loads and no acquire fence
// follows.
__device__ __forceinline__ bool team_wait(GridBarrier *gb, uint32_t gen, int *ok_lds,
                                          uint64_t wait_ticks, bool coherent,
                                          uint32_t *abort = nullptr) {
    if (threadIdx.x == 0) {
        uint32_t *const ab = abort ? abort : &gb->abort.count;
        int ok = 1;
        const uint64_t t0 = uint64_t(wall_clock64());
        uint32_t polls = 0;
        while (__hip_atomic_load(&gb->generation.count, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT) < gen) {
            if (__hip_atomic_load(ab, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                ok = 0;
                break;
            }
            if ((polls++ & 63u) == 0 && uint64_t(wall_clock64()) - t0 > wait_ticks) {
                __hip_atomic_store(ab, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        // acquire: drop stale cached copies of the other workgroups' rows
        if (!coherent) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        *ok_lds = ok;
    }
    __syncthreads();
    return *ok_lds != 0;
}

// Slab flags: in a single-job resident launch, flags[s] counts the steps
// slab s has completed. A slab waits only for the slabs that own its
// neighbours ([smin, smax] but itself), not for the whole grid.
// Publishing follows the slab's writes of the step. With `coherent` those
// writes are agent-coherent stores: each thread waits until its own have
// completed, then the flag goes out with no release fence (which would write
// back the whole L2). Otherwise an agent-scope release precedes the flag.
__device__ __forceinline__ void slab_publish(uint32_t *flags, int32_t me, uint32_t steps,
                                             bool coherent) {
    if (coherent) __builtin_amdgcn_s_waitcnt(0);  // this thread's stores have completed
    __syncthreads();
    if (threadIdx.x == 0) {
        // release: this slab's state rows reach device-coherent memory
        if (!coherent) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(&flags[me], steps, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Waits (at most `wait_ticks` of wall clock) until every slab in [smin, smax]
// but `me` has completed `steps` steps; wave 0 polls 64 flags per round and
// reads the clock on the first unsatisfied round and every 64th. False if the
// job aborted or the wait timed out (which raises the sticky abort). With
// `coherent` the other slabs' rows are read with agent-coherent loads, so no
// acquire fence (which would invalidate the whole L2) follows the wait.
__device__ __forceinline__ bool slab_wait(const uint32_t *flags, GridBarrier *gb, int32_t smin,
                                          int32_t smax, int32_t me, uint32_t steps, int *ok_lds,
                                          uint64_t wait_ticks, bool coherent) {
    if (threadIdx.x < 64) {
        int ok = 1;
        const uint64_t t0 = uint64_t(wall_clock64());
        uint32_t polls = 0;
        for (;;) {
            int done = 1;
            for (int32_t s = smin + int32_t(threadIdx.x); s <= smax; s += 64)
                if (s != me &&
                    __hip_atomic_load(&flags[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < steps)
                    done = 0;
            if (__all(done)) break;
            const uint32_t ab = __builtin_amdgcn_readfirstlane(
                __hip_atomic_load(&gb->abort.count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            if (ab) {
                ok = 0;
                break;
            }
            if ((polls++ & 63u) == 0 && uint64_t(wall_clock64()) - t0 > wait_ticks) {
                if (threadIdx.x == 0)
                    __hip_atomic_store(&gb->abort.count, 1u, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        // acquire: drop stale cached copies of the other slabs' rows
        if (!coherent) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (threadIdx.x == 0) *ok_lds = ok;
    }
    __syncthreads();
    return *ok_lds != 0;
}

}  // namespace ct
