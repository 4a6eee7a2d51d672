// Reverse-mode (adjoint) gradient of one beat, gfx950.
//
// Objective (s_mse = mse_strength / S, s_reg = maximum_regularization_strength):
//   J = sum_t [ s_mse/2 |H x_t - mhat_t|^2 + s_reg/2 sum_v relu(|x_{t,v}|_1 - theta)^2 ]
// Its derivative in x_t is the reference's per-step factor
// c_t = s_mse H^T r_t + s_reg mr_t (derivation.rs:315-326, 509-535). For the
// connection (v,d) that reads neighbour w with delay D and coefficient a,
//   xbar_t[3w+o] = c_t[3w+o] + sum over the connections reading w of q_{t+D}[.,o]
//   ybar_t[o]    = sum_i G[3v+i][3d+o] xbar_t[3v+i] - a ybar_{t+1}[o]
//   q_t[o]       = a ybar_t[o] + ybar_{t+1}[o]          (what (v,d) sends back to w)
//   dJ/dG[3v+i][3d+o] = sum_t xbar_t[3v+i] y_t[o]
//   dJ/da            = sum_t sum_o ybar_t[o] (x_{t-D}[3w+o] - y_{t-1}[o])
//
// adjoint_step_kernel runs one backward step for all voxels (xbar, then ybar
// and q); q lives in a ring of R >= Dmax + 1 time slots. adjoint_gradient_kernel
// then runs one lane per connection over the filter-output checkpoints the
// forward pass stored every K steps: it recomputes each K-step segment and
// walks it backwards, accumulating dG and da.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "barrier.h"
#include "device.h"
#include "lds_sort.h"
#include "tile_ring.h"
#include "kernels.h"

namespace ct {
namespace {

typedef float float3_a4 __attribute__((ext_vector_type(3), aligned(4)));

constexpr int kAdjVoxels = 64;
constexpr int kAdjThreads = 256;

__global__ void max_delay_kernel(const int32_t *__restrict__ nbr, const int32_t *__restrict__ delays,
                                 int64_t len, int32_t *__restrict__ out) {
    __shared__ int32_t red[256];
    int32_t m = 0;
    for (int64_t k = int64_t(blockIdx.x) * 256 + threadIdx.x; k < len; k += int64_t(gridDim.x) * 256)
        if (nbr[k] >= 0) m = max(m, delays[k]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, red[0]);
}

__global__ void consumer_delays_kernel(const int32_t *__restrict__ cons_idx,
                                       const int32_t *__restrict__ delays,
                                       int32_t *__restrict__ cons_delay, int64_t len) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < len) cons_delay[k] = delays[cons_idx[k]];
}

// c_t for one state from the state triple of its voxel and its mapped residual.
__device__ __forceinline__ float step_factor(const float xv[3], int o, float e,
                                             const AdjointParams &p) {
    const float sum = fabsf(xv[0]) + fabsf(xv[1]) + fabsf(xv[2]);
    const float mr = sum > p.threshold ? (sum - p.threshold) * signum(xv[o]) : 0.0f;
    return fmaf(e, p.mse_scaling, mr * p.reg_scaling);
}

// One backward step s. Phase A: one lane per voxel w computes xbar_s of its
// three states (c_s, then the q of every connection reading w, in ascending
// connection order). Phase B: one lane per connection computes ybar_s and q_s.
__global__ __launch_bounds__(kAdjThreads) void adjoint_step_kernel(
    const int32_t *__restrict__ nbr, const float *__restrict__ coefs,
    const float *__restrict__ gains_rows, const int32_t *__restrict__ cons_ptr,
    const int32_t *__restrict__ cons_idx, const int32_t *__restrict__ cons_delay,
    const float *__restrict__ x, const float *__restrict__ E, float *__restrict__ ybar,
    float *__restrict__ qring, float *__restrict__ xbar, int32_t N, int32_t T, int32_t s,
    uint32_t rmask, AdjointParams p) {
    __shared__ float xb[kAdjVoxels * 3];
    const int64_t n = 3LL * N;
    const int64_t M = int64_t(N) * 78;  // one ring slot
    const int32_t v0 = xcd_block(blockIdx.x, gridDim.x) * kAdjVoxels;
    const int32_t nv = min(kAdjVoxels, N - v0);

    if (int(threadIdx.x) < nv) {
        const int32_t w = v0 + threadIdx.x;
        const int64_t row = int64_t(s) * n + 3LL * w;
        const float xv[3] = {x[row], x[row + 1], x[row + 2]};
        float acc[3];
#pragma unroll
        for (int o = 0; o < 3; ++o) acc[o] = step_factor(xv, o, E[row + o], p);
        const int32_t j1 = cons_ptr[w + 1];
        for (int32_t j = cons_ptr[w]; j < j1; ++j) {
            const int32_t t = s + cons_delay[j];
            if (t < T) {
                const float *q = qring + int64_t(uint32_t(t) & rmask) * M + int64_t(cons_idx[j]) * 3;
                const float3_a4 v = *reinterpret_cast<const float3_a4 *>(q);  // 12-byte load
                acc[0] += v.x;
                acc[1] += v.y;
                acc[2] += v.z;
            }
        }
#pragma unroll
        for (int o = 0; o < 3; ++o) {
            xb[threadIdx.x * 3 + o] = acc[o];
            xbar[row + o] = acc[o];
        }
    }
    __syncthreads();
    for (int item = threadIdx.x; item < nv * 26; item += kAdjThreads) {
        const int64_t c = int64_t(v0) * 26 + item;
        if (nbr[c] < 0) continue;
        const int vv = item / 26, d = item - 26 * vv;
        const int64_t grow = 3LL * (v0 + vv);
        const float a = coefs[c];
#pragma unroll
        for (int oo = 0; oo < 3; ++oo) {
            const float next = ybar[c * 3 + oo];
            const float *g = gains_rows + grow * 78 + 3 * d + oo;
            const float yb =
                g[0] * xb[3 * vv] + g[78] * xb[3 * vv + 1] + g[156] * xb[3 * vv + 2] - a * next;
            qring[int64_t(uint32_t(s) & rmask) * M + c * 3 + oo] = a * yb + next;
            ybar[c * 3 + oo] = yb;
        }
    }
}

// Resident backward sweep: one launch runs steps T-1 .. 0 for all voxels,
// with each connection's ybar recurrence owned by the workgroup of the voxel
// that CONSUMES its q. Connection c = (v, d) reading w with delay D and
// coefficient a contributes q_tau = a ybar_tau + ybar_{tau+1} to xbar_s[w] at
// tau = s + D, and ybar_tau = G_c^T xbar_tau[v] - a ybar_{tau+1}. So at sweep
// step s the lane owning (w, j) (c = the j-th consumer of w) reads
// xbar_{s+D}[v] from the xbar history (written D steps earlier in the sweep),
// advances its ybar (kept in registers) and leaves q in LDS; then one lane
// per voxel sums its consumers' q in list order. There is no q ring: the
// gathers hit the last Dmax rows of xbar, a small footprint, instead of a
// ring of 78 values per voxel and step. The values and summation orders are
// those of adjoint_step_kernel, so the results are the same bits.
//
// A workgroup owns voxels [vb, ve) (at most kResMaxVoxels). Registers per
// item (w, j): ybar (3) and one packed word; per phase-2 voxel lane its
// consumers' delays, three per word. The set-up writes each item's consumer
// gains and coefficient (prediction.rs gains G_c, coefficient a_c) to gcons
// (N, 26, 10) at the item's position in the slab's phase-1 order, which the
// step loop then streams contiguously. When every delay of the slab is at
// most kSortedMaxDelay, phase 1 runs the items sorted by (delay, consumer):
// the lanes of one gather read one xbar row at nearby voxels (the packed
// word is producer voxel 17 bits | item 14 bits, the delay sits in an LDS
// byte table). Otherwise items keep (voxel, consumer) order and the word is
// consumer voxel 17 bits | delay 15 bits.
constexpr int kSweepThreads = 512;
constexpr int kSweepItems = (kResMaxVoxels * 26 + kSweepThreads - 1) / kSweepThreads;
constexpr int kSweepBatch = 1;  // items per load group: 1 measured fastest (102 ms vs 112 at 4)
constexpr int kConsGainStride = 10;  // gcons floats per (voxel, consumer) slot
// Items per lane whose consumer gains and coefficient stay in registers for
// the whole sweep instead of being re-read from gcons every step (the fenced
// exchange needs more registers and keeps fewer). Coherent sweep of the
// 120k-voxel shell (same-run A/B): 4 items 54.0 ms, 5 items 53.1 ms (2
// spilled registers, outside the step loop), 6 items 56.9 ms and 8 items
// 65.7 ms (10 and 40 spilled, inside it).
constexpr int kSweepCachedCoherent = 5;
constexpr int kSweepCachedFenced = 2;
// The batched sweep also holds its job's fields, so fewer items fit without
// spills (1: 8 spilled registers, 2: 22, 3: 32); the spills stay out of the
// step loop, and the cached gains win. Batched epoch of 64 uneven scenarios
// in reverse mode (same-run A/Bs): one item 652.8 ms, two 646.9 ms; two
// 653.0 ms, three 649.6 ms. It takes jobs with the coherent exchange only.
constexpr int kBatchSweepCached = 3;
constexpr uint32_t kSweepNoEntry = 0xFFFFFFFFu;
static_assert(kSortSize >= kSweepItems * kSweepThreads || kSortSize >= kResMaxVoxels * 26,
              "every phase-1 position has a key");
static_assert(kSortSize * 8 <= kResMaxVoxels * 78 * 4, "sort keys must fit over q");

// The backward sweep of slab `slab` of the `slabs` slabs of job j, barrier
// member `member` of its team; false if the job aborted (a wait timed out,
// here or in another member). kCoherent: the xbar rows are stored through
// the L2 and gathered past it (sc1), and the barrier needs no agent-scope
// fences. `abort`: the sticky word a timed-out wait raises and every wait
// polls (null: the job barrier's own). The first wait gets `first_wait`
// ticks, the others `step_wait`.
template <bool kCoherent, int kSweepCachedItems>
__device__ __forceinline__ bool sweep_slab(const SweepJob &j, int32_t slab, int32_t slabs,
                                           uint32_t member, uint32_t *abort,
                                           uint64_t first_wait, uint64_t step_wait) {
    extern __shared__ float lds[];
    const int32_t *__restrict__ cons_ptr = j.cons_ptr;
    const int32_t *__restrict__ cons_idx = j.cons_idx;
    const int32_t *__restrict__ delays = j.delays;
    const float *__restrict__ gains_rows = j.gains_rows;
    const float *__restrict__ coefs = j.coefs;
    float *__restrict__ gcons = j.gcons;
    const float *__restrict__ x = j.x;
    const float *__restrict__ E = j.E;
    float *__restrict__ xbar = j.xbar;
    GridBarrier *__restrict__ gb = j.gb;
    const int32_t N = j.N, T = j.T;
    const AdjointParams p = j.p;
    float *q_lds = lds;  // [voxel][26][3]
    // the xbar history as one buffer resource
    const __amdgpu_buffer_rsrc_t xbrs = __builtin_amdgcn_make_buffer_rsrc(xbar, 0, -1, kBufferWord3);
    int *ok_lds = reinterpret_cast<int *>(lds + kResMaxVoxels * 78);  // [4]
    uint8_t *dtab = reinterpret_cast<uint8_t *>(ok_lds + 4);          // [phase-1 position]
    uint64_t *keys = reinterpret_cast<uint64_t *>(lds);                 // set-up only, over q
    const int64_t n = 3LL * N;
    const int32_t vb = int32_t(int64_t(slab) * N / slabs);
    const int32_t ve = int32_t(int64_t(slab + 1) * N / slabs);
    const int32_t nv = ve - vb, items = nv * 26;
    const int tid = threadIdx.x;
    uint32_t gen = 0;

    // phase-2 lane (tid < nv): its voxel's consumer count and delays
    uint32_t dj[9];
    int32_t ncons = 0;
#pragma unroll
    for (int q = 0; q < 9; ++q) dj[q] = 0u;
    if (tid < nv) {
        const int32_t j0 = cons_ptr[vb + tid];
        ncons = cons_ptr[vb + tid + 1] - j0;
#pragma unroll
        for (int j = 0; j < 26; ++j)
            if (j < ncons) dj[j / 3] |= uint32_t(min(delays[cons_idx[j0 + j]], 1023)) << (10 * (j % 3));
    }
    // Keys D << 40 | c << 14 | item, c = the consuming connection; absent
    // items sort last. keys[position] is the item at that phase-1 position.
    if (tid == 0) ok_lds[1] = 0;
    __syncthreads();
    for (int it = tid; it < kSortSize; it += kSweepThreads) {
        uint64_t key = ~0ull;
        if (it < items) {
            const int32_t w = vb + it / 26, j = it % 26;
            const int32_t j0 = cons_ptr[w];
            if (j < cons_ptr[w + 1] - j0) {
                const int32_t c = cons_idx[j0 + j];
                const int32_t D = min(delays[c], 32767);
                key = (uint64_t(D) << 40) | (uint64_t(c) << 14) | uint64_t(it);
                atomicMax(&ok_lds[1], D);
            }
        }
        keys[it] = key;
    }
    __syncthreads();
    const bool sorted = ok_lds[1] <= kSortedMaxDelay;
    if (sorted) sort_keys_lds<kSweepThreads, kSortSize>(keys);
    // phase-1 items: packed word, ybar state; their gains go to gcons here
    uint32_t pk[kSweepItems];
    float yb[kSweepItems][3];
    // bit k: item k has delay 1, so it reads the xbar row of the sweep step
    // just before and must wait for that step's barrier
    uint32_t late = 0u;
    float3_a4 gkeep[kSweepCachedItems][3];
    float akeep[kSweepCachedItems];
#pragma unroll
    for (int k = 0; k < kSweepCachedItems; ++k) {
        akeep[k] = 0.0f;
#pragma unroll
        for (int i = 0; i < 3; ++i) gkeep[k][i] = float3_a4{0.0f, 0.0f, 0.0f};
    }
#pragma unroll
    for (int k = 0; k < kSweepItems; ++k) {
        const int pos = tid + k * kSweepThreads;
        pk[k] = kSweepNoEntry;
#pragma unroll
        for (int o = 0; o < 3; ++o) yb[k][o] = 0.0f;
        if (pos < items) {
            const uint64_t key = keys[pos];
            if (key != ~0ull) {
                const int32_t D = int32_t(key >> 40);
                const int32_t c = int32_t((key >> 14) & 0x3FFFFFFu);
                const uint32_t v = uint32_t(c / 26), d = uint32_t(c % 26);
                const uint32_t it = uint32_t(key & 0x3FFFu);
                if (D == 1) late |= 1u << k;
                if (sorted) {
                    dtab[pos] = uint8_t(D);
                    pk[k] = v | (it << 17);
                } else {
                    pk[k] = v | (uint32_t(D) << 17);
                }
                float *dst = gcons + (int64_t(vb) * 26 + pos) * kConsGainStride;
#pragma unroll
                for (int in = 0; in < 3; ++in)
#pragma unroll
                    for (int o = 0; o < 3; ++o)
                        dst[3 * in + o] = gains_rows[(3 * int64_t(v) + in) * 78 + 3 * d + o];
                dst[9] = coefs[c];
                if (k < kSweepCachedItems) {
                    const int kc = k < kSweepCachedItems ? k : 0;
#pragma unroll
                    for (int in = 0; in < 3; ++in)
                        gkeep[kc][in] = float3_a4{dst[3 * in], dst[3 * in + 1], dst[3 * in + 2]};
                    akeep[kc] = dst[9];
                }
            }
        }
    }
    __syncthreads();  // the keys are dead: q is written over them; gcons rows are this lane's own

    // Step s: phase 1 of the items with delay >= 2 (rows published by the
    // barrier of step s + 2), the wait for step s + 1's barrier, phase 1 of
    // the delay-1 items, phase 2, and the arrival at step s's barrier.
    for (int32_t s = T - 1; s >= 0; --s) {
#pragma unroll 1
        for (uint32_t pass = 0; pass < 2; ++pass) {
        if (pass == 1 && s < T - 1 &&
            !team_wait(gb, gen, ok_lds, s == T - 2 ? first_wait : step_wait, kCoherent, abort))
            return false;
        // phase 1: each item's q for step s (its ybar at tau = s + D). The
        // empty asm statement hides the lane index from the optimiser, so the
        // per-item addresses are recomputed each step instead of being kept
        // live across the time loop (which would not fit the register file).
        int32_t lane = tid;
        asm volatile("" : "+v"(lane));
#pragma unroll
        for (int k0 = 0; k0 < kSweepItems; k0 += kSweepBatch) {
            float3_a4 xv[kSweepBatch], g[kSweepBatch][3];
            float av[kSweepBatch];
            bool on[kSweepBatch];
#pragma unroll
            for (int kk = 0; kk < kSweepBatch; ++kk) {
                const int k = k0 + kk < kSweepItems ? k0 + kk : kSweepItems - 1;
                const int pos = lane + k * kSweepThreads;
                const bool present = k0 + kk < kSweepItems && pk[k] != kSweepNoEntry &&
                                     ((late >> k) & 1u) == pass;
                // both forms of the delay, selected (no branch on `sorted`); an
                // absent item reads position 0 (positions past the slab's
                // items lie beyond the byte table)
                const int32_t Ds = int32_t(dtab[present ? pos : 0]);
                const int32_t D = !present ? 0 : sorted ? Ds : int32_t(pk[k] >> 17);
                const int32_t tau = s + D;
                on[kk] = present && tau < T;
                if (!on[kk]) continue;  // not this pass, absent, or tau beyond the beat
                const int64_t xe = int64_t(tau) * n + 3LL * (pk[k] & 0x1FFFFu);
                if constexpr (kCoherent)
                    xv[kk] = __builtin_amdgcn_raw_buffer_load_b96(xbrs, uint32_t(xe) * 4u, 0, kCacheSC1);
                else
                    xv[kk] = *reinterpret_cast<const float3_a4 *>(xbar + xe);
                if (k0 + kk < kSweepCachedItems) {
                    const int kc = k0 + kk < kSweepCachedItems ? k0 + kk : 0;
#pragma unroll
                    for (int i = 0; i < 3; ++i) g[kk][i] = gkeep[kc][i];
                    av[kk] = akeep[kc];
                } else {
                    const float *gs = gcons + (int64_t(vb) * 26 + pos) * kConsGainStride;
#pragma unroll
                    for (int i = 0; i < 3; ++i)
                        g[kk][i] = *reinterpret_cast<const float3_a4 *>(gs + 3 * i);
                    av[kk] = gs[9];
                }
            }
#pragma unroll
            for (int kk = 0; kk < kSweepBatch; ++kk) {
                const int k = k0 + kk;
                if (!on[kk]) continue;
                const int it = sorted ? int(pk[k] >> 17) : lane + k * kSweepThreads;
#pragma unroll
                for (int o = 0; o < 3; ++o) {
                    const float next = yb[k][o];
                    const float ny = g[kk][0][o] * xv[kk].x + g[kk][1][o] * xv[kk].y +
                                     g[kk][2][o] * xv[kk].z - av[kk] * next;
                    q_lds[it * 3 + o] = av[kk] * ny + next;
                    yb[k][o] = ny;
                }
            }
        }
        }
        __syncthreads();
        // phase 2: xbar_s of the slab's states, consumers in list order
        if (tid < nv) {
            const int32_t w = vb + tid;
            const int64_t row = int64_t(s) * n + 3LL * w;
            const float xs3[3] = {x[row], x[row + 1], x[row + 2]};
            float acc[3];
#pragma unroll
            for (int o = 0; o < 3; ++o) acc[o] = step_factor(xs3, o, E[row + o], p);
#pragma unroll
            for (int j = 0; j < 26; ++j) {
                const int32_t D = int32_t((dj[j / 3] >> (10 * (j % 3))) & 1023u);
                if (j < ncons && s + D < T) {
#pragma unroll
                    for (int o = 0; o < 3; ++o) acc[o] += q_lds[(tid * 26 + j) * 3 + o];
                }
            }
#pragma unroll
            for (int o = 0; o < 3; ++o)
                __hip_atomic_store(xbar + row + o, acc[o], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);  // written through the L2
        }
        team_arrive(gb, uint32_t(slabs), member, gen, kCoherent);
    }
    return true;
}

// One scenario's sweep: a slab per workgroup (launch_adjoint_sweep_resident
// picks the instantiation).
template <bool kCoherent>
__global__ __launch_bounds__(kSweepThreads) void adjoint_sweep_resident_kernel(
    const int32_t *__restrict__ cons_ptr, const int32_t *__restrict__ cons_idx,
    const int32_t *__restrict__ delays, const float *__restrict__ gains_rows,
    const float *__restrict__ coefs, float *__restrict__ gcons,
    const float *__restrict__ x, const float *__restrict__ E,
    float *__restrict__ xbar, GridBarrier *__restrict__ gb, int32_t N, int32_t T, AdjointParams p,
    uint64_t wait_ticks) {
    SweepJob j;
    j.cons_ptr = cons_ptr;
    j.cons_idx = cons_idx;
    j.delays = delays;
    j.gains_rows = gains_rows;
    j.coefs = coefs;
    j.gcons = gcons;
    j.x = x;
    j.E = E;
    j.xbar = xbar;
    j.gb = gb;
    j.N = N;
    j.T = T;
    j.p = p;
    j.slot0 = 0;
    j.width = int32_t(gridDim.x);
    j.coherent = kCoherent ? 1 : 0;
    (void)sweep_slab<kCoherent, kCoherent ? kSweepCachedCoherent : kSweepCachedFenced>(
        j, xcd_block(blockIdx.x, gridDim.x), int32_t(gridDim.x), blockIdx.x, nullptr, wait_ticks,
        wait_ticks);
}

// Many scenarios' sweeps in one launch: slot q walks its task list, taking
// slab q - slot0 of each job's team in turn (the host's placement makes the
// lists deadlock-free, as for propagate_batch_kernel). Every job exchanges
// its x̄ rows through agent-coherent accesses.
__global__ __launch_bounds__(kSweepThreads) void adjoint_sweep_batch_kernel(
    const SweepJob *__restrict__ jobs, const int32_t *__restrict__ task_ptr,
    const int32_t *__restrict__ task_job, uint32_t *abort, uint64_t first_wait,
    uint64_t step_wait) {
    const int32_t slot = xcd_block(blockIdx.x, gridDim.x);
    const int32_t q1 = task_ptr[slot + 1];
    for (int32_t q = task_ptr[slot]; q < q1; ++q) {
        const SweepJob job = jobs[task_job[q]];
        const int32_t member = slot - job.slot0;
        if (!sweep_slab<true, kBatchSweepCached>(job, member, job.width, uint32_t(member), abort,
                                                 first_wait, step_wait))
            return;
        __syncthreads();  // LDS is reused by the next job
    }
}

// gcons[w][j] = (G[3v+in][3d+o] for in, o; a) of the j-th consumer
// c = (v, d) of w: 10 floats per slot.

// Gradient pass: one lane per connection. For each K-step segment, from the
// last to the first: run the filter forward from the checkpoint the forward
// pass stored (y_{t0-1}), keeping only its state at every kGradChunk-step chunk
// start; then, chunk by chunk from the last, recompute the chunk's filter
// outputs into registers (from the same neighbour rows the walk reads) and
// walk it backwards (own-voxel xbar staged in LDS), recomputing ybar and
// accumulating dJ/dG and dJ/da.
constexpr int kGradThreads = 256;
constexpr int kGradChunk = 16;

// Sixteen consecutive floats of three rows `stride` apart, in 16-byte loads
// that need only 4-byte alignment.
typedef float float4_a4 __attribute__((ext_vector_type(4), aligned(4)));
__device__ __forceinline__ void load_rows16(const float *src, int64_t stride,
                                            float out[3][kGradChunk]) {
#pragma unroll
    for (int o = 0; o < 3; ++o)
#pragma unroll
        for (int q = 0; q < kGradChunk / 4; ++q) {
            const float4_a4 v = *reinterpret_cast<const float4_a4 *>(src + o * stride + 4 * q);
            out[o][4 * q] = v.x;
            out[o][4 * q + 1] = v.y;
            out[o][4 * q + 2] = v.z;
            out[o][4 * q + 3] = v.w;
        }
}
constexpr int kGradMaxVoxels = kGradThreads / 26 + 2;

// One backward step of a connection's filter adjoint in the gradient pass:
// ybar_t[o] = sum_i G[i][o] xbar_t[i] - a ybar_{t+1}[o], as fused
// multiply-adds (both gradient kernels use this, so they agree bit for bit).
__device__ __forceinline__ float ybar_step(const float G[3][3], const float xb[3], float a,
                                           float ybn, int o) {
    return fmaf(-a, ybn, fmaf(G[2][o], xb[2], fmaf(G[1][o], xb[1], G[0][o] * xb[0])));
}

__global__ __launch_bounds__(kGradThreads) void adjoint_gradient_kernel(
    const int32_t *__restrict__ nbr, const float *__restrict__ coefs,
    const int32_t *__restrict__ delays, const float *__restrict__ gT,
    const float *__restrict__ xT, const float *__restrict__ xbarT, const float *__restrict__ ckpt,
    float *__restrict__ ychunk, float *__restrict__ dGT, float *__restrict__ dA, int32_t N,
    int32_t T, int32_t K) {
    __shared__ float xs[kGradMaxVoxels][3][kGradChunk];
    const int64_t total = int64_t(N) * 26;
    const int64_t item0 = int64_t(xcd_block(blockIdx.x, gridDim.x)) * kGradThreads;
    const int64_t item = item0 + threadIdx.x;
    const bool active = item < total;
    const int32_t vbeg = int32_t(item0 / 26);
    const int32_t vend = int32_t(min(total - 1, item0 + kGradThreads - 1) / 26) + 1;
    const int32_t nvox = vend - vbeg;
    const int64_t it = active ? item : item0;
    const int32_t w = active ? nbr[it] : -1;
    const bool has = w >= 0;
    const int64_t n = 3LL * N;
    const int64_t M = total * 3;
    const int32_t v = int32_t(it / 26), d = int32_t(it % 26), vl = v - vbeg;
    const float a = coefs[it];
    const int32_t D = delays[it];
    const float *xw = xT + (has ? 3LL * w : 0) * T;
    const int32_t nseg = (T + K - 1) / K;
    // per segment, the filter state (y, x[t-D-1]) at the start of each chunk
    float *ys = ychunk + it * 6;
    const int64_t ystride = total * 6;

    float G[3][3], g[3][3];
#pragma unroll
    for (int in = 0; in < 3; ++in)
#pragma unroll
        for (int o = 0; o < 3; ++o) {
            G[in][o] = gT[int64_t(3 * d + o) * n + 3LL * v + in];
            g[in][o] = 0.0f;
        }
    float A = 0.0f, ybn[3] = {0.0f, 0.0f, 0.0f};
    for (int32_t k = nseg - 1; k >= 0; --k) {
        const int32_t t0 = k * K, t1 = min(T, t0 + K);
        const int32_t nchunks = (t1 - t0 + kGradChunk - 1) / kGradChunk;
        // pass 1: from the segment's checkpoint forward, keeping only the
        // state at each chunk start
        if (has) {
            float ycur[3], pin[3];
#pragma unroll
            for (int o = 0; o < 3; ++o) {
                ycur[o] = k == 0 ? 0.0f : ckpt[int64_t(k - 1) * M + it * 3 + o];
                pin[o] = D <= t0 - 1 ? xw[int64_t(o) * T + t0 - 1 - D] : 0.0f;
            }
            for (int32_t c = 0; c < nchunks; ++c) {
                const int32_t tc = t0 + c * kGradChunk;
#pragma unroll
                for (int o = 0; o < 3; ++o) {
                    ys[c * ystride + o] = ycur[o];
                    ys[c * ystride + 3 + o] = pin[o];
                }
                if (c == nchunks - 1) break;  // the last chunk's steps are not needed here
                float xc[3][kGradChunk];
                if (tc >= D) {
                    load_rows16(xw + (tc - D), T, xc);  // a full chunk: tc + 16 <= t1
                } else {
#pragma unroll
                    for (int o = 0; o < 3; ++o)
#pragma unroll
                        for (int j = 0; j < kGradChunk; ++j)
                            xc[o][j] = D <= tc + j ? xw[int64_t(o) * T + tc + j - D] : 0.0f;
                }
#pragma unroll
                for (int j = 0; j < kGradChunk; ++j)
#pragma unroll
                    for (int o = 0; o < 3; ++o) {
                        ycur[o] = fmaf(a, xc[o][j] - ycur[o], pin[o]);
                        pin[o] = xc[o][j];
                    }
            }
        }
        // pass 2: chunks from the last to the first; each recomputes its
        // filter outputs into registers from its start state and walks back
        for (int32_t c = nchunks - 1; c >= 0; --c) {
            const int32_t cbeg = t0 + c * kGradChunk, len = min(kGradChunk, t1 - cbeg);
            for (int idx = threadIdx.x; idx < nvox * len; idx += kGradThreads) {
                const int vv = idx / len, j = idx - vv * len;
                const int64_t row = 3LL * (vbeg + vv) * T + cbeg + j;
#pragma unroll
                for (int q = 0; q < 3; ++q) xs[vv][q][j] = xbarT[row + int64_t(q) * T];
            }
            float xi[3][kGradChunk];
            if (has && len == kGradChunk && cbeg >= D) {
                load_rows16(xw + (cbeg - D), T, xi);
            } else {
#pragma unroll
                for (int o = 0; o < 3; ++o)
#pragma unroll
                    for (int j = 0; j < kGradChunk; ++j) {
                        const int32_t src = cbeg + j - D;
                        xi[o][j] = (has && j < len && src >= 0) ? xw[int64_t(o) * T + src] : 0.0f;
                    }
            }
            float y0[3], yv[3][kGradChunk];
            if (has) {
                float pin[3];
#pragma unroll
                for (int o = 0; o < 3; ++o) {
                    y0[o] = ys[c * ystride + o];
                    pin[o] = ys[c * ystride + 3 + o];
                }
#pragma unroll
                for (int o = 0; o < 3; ++o) {
                    float yc = y0[o];
#pragma unroll
                    for (int j = 0; j < kGradChunk; ++j) {
                        yc = fmaf(a, xi[o][j] - yc, j == 0 ? pin[o] : xi[o][j - 1]);
                        yv[o][j] = yc;
                    }
                }
            }
            __syncthreads();
            if (has) {
#pragma unroll
                for (int j = kGradChunk - 1; j >= 0; --j) {
                    if (j >= len) continue;
                    const float xb[3] = {xs[vl][0][j], xs[vl][1][j], xs[vl][2][j]};
#pragma unroll
                    for (int o = 0; o < 3; ++o) {
                        const float yb = ybar_step(G, xb, a, ybn[o], o);
                        const float ym = j > 0 ? yv[o][j - 1] : y0[o];
#pragma unroll
                        for (int in = 0; in < 3; ++in) g[in][o] = fmaf(xb[in], yv[o][j], g[in][o]);
                        A = fmaf(yb, xi[o][j] - ym, A);
                        ybn[o] = yb;
                    }
                }
            }
            __syncthreads();
        }
    }
    if (!active || !has) return;  // no connection: no dependence on these parameters
#pragma unroll
    for (int in = 0; in < 3; ++in)
#pragma unroll
        for (int o = 0; o < 3; ++o) dGT[int64_t(3 * d + o) * n + 3LL * v + in] += g[in][o];
    dA[item] += A;
}

// Gradient pass over voxel tiles, for checkpoints every kTileChunk steps
// (every segment one chunk, so no segment is replayed to find its chunk
// starts). A workgroup takes the connections of one tile (derive_tiled_kernel's
// tables) and walks the chunks from the last to the first; the delayed
// inputs x_w[t - D] of its lanes come from an LDS ring of the tile's halo
// rows that holds chunks c-3 .. c while chunk c runs, filled going backwards.
// Each lane's arithmetic is adjoint_gradient_kernel's at K = kTileChunk,
// operation for operation, so the two give the same bits.
constexpr int kGradTileMaxDelay = kTileRing - kTileChunk - 1;  // the chunk's start input too
constexpr size_t kGradTiledLdsBytes =
    sizeof(float) * (size_t(kTileRingFloats) + kTileZeroFloats + size_t(kTileStageFloats));
static_assert(kGradChunk == kTileChunk, "one gradient chunk per ring chunk");

__global__ __launch_bounds__(kTileThreads) void adjoint_gradient_tiled_kernel(
    DeriveTiles tl, const int32_t *__restrict__ nbr, const float *__restrict__ coefs,
    const int32_t *__restrict__ delays, const float *__restrict__ gT,
    const float *__restrict__ xT, const float *__restrict__ xbar, const float *__restrict__ ckpt,
    float *__restrict__ dGT, float *__restrict__ dA, int32_t N, int32_t T) {
    extern __shared__ __align__(16) float lds[];
    float *ring = lds;                       // [halo][3][kTileRowStride]
    float *zeros = ring + kTileRingFloats;   // [3][kTileRowStride], +0: missing connections
    float *xs = zeros + kTileZeroFloats;     // [voxel][3][chunk] adjoint states of the tile
    const int64_t n = 3LL * N;
    const int64_t M = int64_t(N) * 78;
    const int32_t tile = xcd_block(blockIdx.x, gridDim.x);
    const int32_t pbeg = tl.vox_ptr[tile], nvox = tl.vox_ptr[tile + 1] - pbeg;
    const int32_t hbeg = tl.halo_ptr[tile], nhalo = tl.halo_ptr[tile + 1] - hbeg;
    const int tid = threadIdx.x;

    const TileLane L = tile_lane(tl, nbr, coefs, delays, pbeg, nvox, kGradTileMaxDelay);
    for (int i = tid; i < kTileZeroFloats; i += kTileThreads) zeros[i] = 0.0f;
    const bool active = L.active, has = L.has, far = L.far;
    const int d = L.d;
    const int32_t v = L.v, D = L.D;
    const int64_t item = L.item;
    const float a = L.a;
    const float *xw = xT + (has ? 3LL * L.w : 0) * T;
    const float *rrow = L.slot >= 0 ? ring + L.slot * 3 * kTileRowStride : zeros;

    float G[3][3], g[3][3];
#pragma unroll
    for (int in = 0; in < 3; ++in)
#pragma unroll
        for (int o = 0; o < 3; ++o) {
            G[in][o] = active ? gT[int64_t(3 * d + o) * n + 3LL * v + in] : 0.0f;
            g[in][o] = 0.0f;
        }
    float A = 0.0f, ybn[3] = {0.0f, 0.0f, 0.0f};

    // this thread's share of the copies: quads of halo rows, and one (voxel,
    // step) item of the tile's adjoint states, read from x̄ by step (T, n):
    // its voxel's three states at one step, 12 contiguous bytes (the tile's
    // voxels come in runs along the grid's fastest axis)
    TileRingCopy rc;
    rc.init(tl, hbeg, nhalo, T);
    const int sv = tid / kTileChunk, sj = tid - kTileChunk * sv;
    const bool sown = sv < nvox;
    const int64_t scol = sown ? 3LL * tl.vox[pbeg + sv] : 0;
    float sb[3];
    const auto fetch_xbar = [&](int32_t cb) {
        const bool live = sown && cb + sj < T;
        const float *src = xbar + int64_t(live ? cb + sj : 0) * n + scol;
#pragma unroll
        for (int q = 0; q < 3; ++q) sb[q] = live ? src[q] : 0.0f;
    };
    const auto stage_xbar = [&]() {
        if (sown) {
#pragma unroll
            for (int q = 0; q < 3; ++q) xs[(sv * 3 + q) * kTileChunk + sj] = sb[q];
        }
    };
    const auto fetch_rows = [&](int32_t cb) { rc.fetch(xT, cb, T); };
    const auto stage_rows = [&](int32_t cb) { rc.stage(ring, cb); };
    // chunk c's starting filter outputs (checkpoint c - 1, +0 for chunk 0),
    // loaded one chunk ahead with the copies
    float y0n[3];
    const auto fetch_start = [&](int32_t c) {
#pragma unroll
        for (int o = 0; o < 3; ++o)
            y0n[o] = c == 0 ? 0.0f : ckpt[int64_t(c - 1) * M + item * 3 + o];
    };

    const int32_t nchunks = (T + kTileChunk - 1) / kTileChunk;
    // chunks nchunks-1 .. nchunks-3 first; each iteration c then adds c-3
    for (int32_t c = nchunks - 1; c >= max(nchunks - 3, 0); --c) {
        fetch_rows(c * kTileChunk);
        stage_rows(c * kTileChunk);
    }
    if (nchunks - 4 >= 0) fetch_rows((nchunks - 4) * kTileChunk);
    fetch_xbar((nchunks - 1) * kTileChunk);
    fetch_start(nchunks - 1);
    for (int32_t c = nchunks - 1; c >= 0; --c) {
        const int32_t cbeg = c * kTileChunk, len = min(kTileChunk, T - cbeg);
        if (c - 3 >= 0) stage_rows(cbeg - 3 * kTileChunk);
        stage_xbar();
        const float y0c[3] = {y0n[0], y0n[1], y0n[2]};
        // Every load issued so far has landed (used above). Stated
        // explicitly, since the compiler cannot see it through the per-thread
        // conditions of the copies: without this it waits for the next
        // chunk's loads right after issuing them, before this chunk's steps.
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (c - 4 >= 0) fetch_rows(cbeg - 4 * kTileChunk);
        if (c - 1 >= 0) {
            fetch_xbar(cbeg - kTileChunk);
            fetch_start(c - 1);
        }
        // inputs: rb[o * kTileRowStride + j] = x_w[cbeg + j - D]; the chunk's
        // start input x_w[cbeg - 1 - D] at its own ring position
        const float *rb = rrow + ((cbeg - D) & (kTileRing - 1));
        const float *rp = rrow + ((cbeg - 1 - D) & (kTileRing - 1));
        const float *xl = xs + L.lv * 3 * kTileChunk;
        // every lane of the wave in the ring, past its delay, and a whole chunk:
        // no per-input checks
        const bool fast = __all(!active || cbeg - 1 - D >= 0) && len == kTileChunk;
        // kMode 0: no checks; 1: inputs before step 0 or past T read +0;
        // 2: as 1, and lanes beyond the ring read memory
        const auto walk = [&](auto mode_tag) __attribute__((always_inline)) {
            constexpr int kMode = decltype(mode_tag)::value;
            const float *xwb = xw + (cbeg - D);  // x_w[cbeg + j - D] = xwb[o T + j]
            const auto input = [&](int o, int32_t j) -> float {
                if (kMode == 0) return rb[o * kTileRowStride + j];
                if (j >= len || cbeg + j - D < 0) return 0.0f;
                if (kMode == 2 && far) return xwb[int64_t(o) * T + j];
                return rb[o * kTileRowStride + j];
            };
            float y0[3], pin[3];
#pragma unroll
            for (int o = 0; o < 3; ++o) {
                y0[o] = y0c[o];
                if (kMode == 0) pin[o] = rp[o * kTileRowStride];
                else if (cbeg - 1 - D < 0) pin[o] = 0.0f;
                else pin[o] = (kMode == 2 && far) ? xwb[int64_t(o) * T - 1] : rp[o * kTileRowStride];
            }
            // The chunk in two halves, the later first: the filter is run
            // over the first half to its end state, then each half is
            // recomputed into registers and walked back (half the registers
            // of the whole chunk, for 8 more filter steps).
            constexpr int kHalf = kTileChunk / 2;
            float yh[3], ph[3];
#pragma unroll
            for (int o = 0; o < 3; ++o) {
                float yc = y0[o], prev = pin[o];
#pragma unroll
                for (int j = 0; j < kHalf; ++j) {
                    const float xin = input(o, j);
                    yc = fmaf(a, xin - yc, prev);
                    prev = xin;
                }
                yh[o] = yc;
                ph[o] = prev;
            }
            // Inputs are read from LDS again where they are used next, not
            // kept in registers from their first read: an empty asm with a
            // memory clobber keeps the compiler from reusing the loaded values.
            asm volatile("" ::: "memory");
#pragma unroll
            for (int half = 1; half >= 0; --half) {
                const int jb = half * kHalf;
                float yv[3][kHalf], ys[3];
#pragma unroll
                for (int o = 0; o < 3; ++o) {
                    ys[o] = half ? yh[o] : y0[o];  // y at step jb - 1
                    float yc = ys[o], prev = half ? ph[o] : pin[o];
#pragma unroll
                    for (int j = 0; j < kHalf; ++j) {
                        const float xin = input(o, jb + j);
                        yc = fmaf(a, xin - yc, prev);
                        yv[o][j] = yc;
                        prev = xin;
                    }
                }
                asm volatile("" ::: "memory");
#pragma unroll
                for (int j = kHalf - 1; j >= 0; --j) {
                    if (kMode != 0 && jb + j >= len) continue;
                    const float xb[3] = {xl[jb + j], xl[kTileChunk + jb + j],
                                         xl[2 * kTileChunk + jb + j]};
#pragma unroll
                    for (int o = 0; o < 3; ++o) {
                        const float yb = ybar_step(G, xb, a, ybn[o], o);
                        const float ym = j > 0 ? yv[o][j - 1] : ys[o];
#pragma unroll
                        for (int in = 0; in < 3; ++in) g[in][o] = fmaf(xb[in], yv[o][j], g[in][o]);
                        A = fmaf(yb, input(o, jb + j) - ym, A);
                        ybn[o] = yb;
                    }
                }
                asm volatile("" ::: "memory");
            }
        };
        if (has) {
            if (__any(far)) walk(std::integral_constant<int, 2>{});
            else if (fast) walk(std::integral_constant<int, 0>{});
            else walk(std::integral_constant<int, 1>{});
        }
        __syncthreads();
    }
    if (!active || !has) return;  // no connection: no dependence on these parameters
#pragma unroll
    for (int in = 0; in < 3; ++in)
#pragma unroll
        for (int o = 0; o < 3; ++o) dGT[int64_t(3 * d + o) * n + 3LL * v + in] += g[in][o];
    dA[item] += A;
}

}  // namespace

hipError_t launch_max_delay(hipStream_t s, const int32_t *nbr, const int32_t *delays, int64_t len,
                            int32_t *out) {
    hipError_t e = launch_fill(s, reinterpret_cast<float *>(out), 0.0f, 1);  // bits of int 0
    if (e != hipSuccess || len <= 0) return e;
    max_delay_kernel<<<256, 256, 0, s>>>(nbr, delays, len, out);
    return hipGetLastError();
}

hipError_t launch_consumer_delays(hipStream_t s, const int32_t *cons_idx, const int32_t *delays,
                                  int32_t *cons_delay, int64_t len) {
    if (len <= 0) return hipSuccess;
    consumer_delays_kernel<<<unsigned((len + 255) / 256), 256, 0, s>>>(cons_idx, delays, cons_delay,
                                                                     len);
    return hipGetLastError();
}

hipError_t launch_adjoint_step(hipStream_t s, const int32_t *nbr, const float *coefs,
                               const float *gains_rows, const int32_t *cons_ptr,
                               const int32_t *cons_idx, const int32_t *cons_delay, const float *x,
                               const float *E, float *ybar, float *qring, float *xbar, int32_t N,
                               int32_t T, int32_t step, uint32_t rmask, const AdjointParams &p) {
    if (N <= 0) return hipSuccess;
    const unsigned blocks = unsigned((N + kAdjVoxels - 1) / kAdjVoxels);
    adjoint_step_kernel<<<blocks, kAdjThreads, 0, s>>>(nbr, coefs, gains_rows, cons_ptr, cons_idx,
                                                       cons_delay, x, E, ybar, qring, xbar, N, T,
                                                       step, rmask, p);
    return hipGetLastError();
}

size_t sweep_lds_bytes() { return sizeof(float) * kResMaxVoxels * 78 + 16 + kResMaxVoxels * 26; }

// Workgroups of the sweep that fit on one CU at once: the smaller count of
// its two instantiations (the launcher picks one per launch).
hipError_t sweep_capacity(int32_t *blocks_per_cu) {
    int fit = 1 << 30;
    for (const void *k : {reinterpret_cast<const void *>(&adjoint_sweep_resident_kernel<true>),
                          reinterpret_cast<const void *>(&adjoint_sweep_resident_kernel<false>)}) {
        hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           int(sweep_lds_bytes()));
        if (e != hipSuccess) return e;
        int nb = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, kSweepThreads, sweep_lds_bytes());
        if (e != hipSuccess) return e;
        fit = nb < fit ? nb : fit;
    }
    *blocks_per_cu = fit;
    return hipSuccess;
}

hipError_t sweep_batch_capacity(int32_t *blocks_per_cu) {
    const void *k = reinterpret_cast<const void *>(&adjoint_sweep_batch_kernel);
    hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       int(sweep_lds_bytes()));
    if (e != hipSuccess) return e;
    int nb = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, kSweepThreads, sweep_lds_bytes());
    if (e != hipSuccess) return e;
    *blocks_per_cu = nb;
    return hipSuccess;
}

hipError_t launch_adjoint_sweep_batch(hipStream_t s, int32_t slots, const SweepJob *jobs,
                                      GridBarrier *job_barriers, int32_t num_jobs,
                                      const int32_t *task_ptr, const int32_t *task_job,
                                      uint32_t *abort, uint64_t first_wait, uint64_t step_wait) {
    if (slots <= 0 || num_jobs <= 0) return hipSuccess;
    hipError_t e = launch_barrier_reset(s, job_barriers, num_jobs);
    if (e != hipSuccess) return e;
    adjoint_sweep_batch_kernel<<<slots, kSweepThreads, sweep_lds_bytes(), s>>>(
        jobs, task_ptr, task_job, abort, first_wait, step_wait);
    return hipGetLastError();
}

hipError_t launch_adjoint_sweep_resident(hipStream_t s, int32_t blocks, const int32_t *cons_ptr,
                                         const int32_t *cons_idx, const int32_t *delays,
                                         const float *gains_rows, const float *coefs,
                                         float *gcons, const float *x, const float *E,
                                         float *xbar, GridBarrier *gb, int32_t N, int32_t T,
                                         const AdjointParams &p, uint64_t wait_ticks,
                                         bool allow_coherent) {
    if (N <= 0 || T <= 0) return hipSuccess;
    hipError_t e = launch_barrier_reset(s, gb, 1);
    if (e != hipSuccess) return e;
    if (allow_coherent && coherent_exchange(T, 3LL * N))
        adjoint_sweep_resident_kernel<true><<<blocks, kSweepThreads, sweep_lds_bytes(), s>>>(
            cons_ptr, cons_idx, delays, gains_rows, coefs, gcons, x, E, xbar, gb, N, T, p,
            wait_ticks);
    else
        adjoint_sweep_resident_kernel<false><<<blocks, kSweepThreads, sweep_lds_bytes(), s>>>(
            cons_ptr, cons_idx, delays, gains_rows, coefs, gcons, x, E, xbar, gb, N, T, p,
            wait_ticks);
    return hipGetLastError();
}

int64_t adjoint_gradient_state_len(int32_t N, int32_t T, int32_t K) {
    // a segment spans min(K, T) steps; 6 floats per connection and chunk start
    const int64_t steps = std::max<int64_t>(1, std::min(K, T));
    return (steps + kGradChunk - 1) / kGradChunk * int64_t(N) * 26 * 6;
}

hipError_t adjoint_gradient_tiled_prepare() {
    return hipFuncSetAttribute(reinterpret_cast<const void *>(&adjoint_gradient_tiled_kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, int(kGradTiledLdsBytes));
}

hipError_t launch_adjoint_gradient_tiled(hipStream_t s, const DeriveTiles &tiles,
                                         const int32_t *nbr, const float *coefs,
                                         const int32_t *delays, const float *gT, const float *xT,
                                         const float *xbar, const float *ckpt, float *dGT,
                                         float *dA, int32_t N, int32_t T) {
    if (N <= 0 || T <= 0 || tiles.tiles <= 0) return hipSuccess;
    adjoint_gradient_tiled_kernel<<<tiles.tiles, kTileThreads, kGradTiledLdsBytes, s>>>(
        tiles, nbr, coefs, delays, gT, xT, xbar, ckpt, dGT, dA, N, T);
    return hipGetLastError();
}

hipError_t launch_adjoint_gradient(hipStream_t s, const int32_t *nbr, const float *coefs,
                                   const int32_t *delays, const float *gT, const float *xT,
                                   const float *xbarT, const float *ckpt, float *ychunk,
                                   int64_t ychunk_len, float *dGT, float *dA, int32_t N, int32_t T,
                                   int32_t K) {
    if (N <= 0) return hipSuccess;
    if (K < 1 || ychunk_len < adjoint_gradient_state_len(N, T, K)) return hipErrorInvalidValue;
    const unsigned blocks = unsigned((int64_t(N) * 26 + kGradThreads - 1) / kGradThreads);
    adjoint_gradient_kernel<<<blocks, kGradThreads, 0, s>>>(nbr, coefs, delays, gT, xT, xbarT,
                                                            ckpt, ychunk, dGT, dA, N, T, K);
    return hipGetLastError();
}

}  // namespace ct
