// gfx950 kernels of the CardioTrust epoch hot path.
//
// Built with -ffp-contract=off: every fused multiply-add below is an explicit
// fmaf() standing where the reference writes `mul_add`; all other products and
// sums stay separate so propagation results match the reference's CPU epoch
// bit for bit. Reference citations are relative to the reference's src/core/.
#include "kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>

#include "barrier.h"
#include "device.h"
#include "lds_sort.h"
#include "tile_ring.h"

namespace ct {
namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float float4_a4 __attribute__((ext_vector_type(4), aligned(4)));
constexpr int kMapTile = 128;  // steps per map_residuals workgroup

// ---------------------------------------------------------------------------
// Propagation: one launch per time step (prediction.rs:40-105).
//
// A workgroup owns kPropVoxels consecutive voxels. Phase 1 runs one lane per
// connection (v, d): it advances the three all-pass filters of that
// connection (out = 0..2) from the neighbour's delayed states and stages the
// outputs in LDS. Phase 2 runs one lane per state 3v+in and accumulates
// gains * outputs in ascending (d, out) order, then adds the control term.
// The filter state y is stored once per (v, d, out); the reference keeps one
// identical copy per input component `in`.
// ---------------------------------------------------------------------------
constexpr int kPropVoxels = 64;
constexpr int kPropThreads = 3 * kPropVoxels;
constexpr int kPropItems = (kPropVoxels * 26 + kPropThreads - 1) / kPropThreads;

// Each lane first issues every load it needs for all of its connections (or
// all 78 gains in phase 2) and only then runs the dependent arithmetic, so a
// wave keeps dozens of memory requests in flight instead of one per step of
// the accumulation chain.
__global__ __launch_bounds__(kPropThreads) void propagate_step_kernel(
    const int32_t *__restrict__ nbr, const float *__restrict__ coefs,
    const int32_t *__restrict__ delays, const float *__restrict__ gT, float *__restrict__ y,
    float *__restrict__ y_last, float *__restrict__ x, const float *__restrict__ b, float u_t,
    int32_t N, int32_t t, int32_t T) {
    __shared__ float ys[kPropVoxels * 78];
    __shared__ uint32_t mask[kPropVoxels];
    const int64_t n = 3LL * N;
    const int32_t v0 = xcd_block(blockIdx.x, gridDim.x) * kPropVoxels;
    const int32_t nv = min(kPropVoxels, N - v0);
    const int32_t items = nv * 26;
    const bool last = (t == T - 1);
    if (threadIdx.x < kPropVoxels) mask[threadIdx.x] = 0u;

    int32_t w[kPropItems], D[kPropItems];
    float a[kPropItems];
#pragma unroll
    for (int k = 0; k < kPropItems; ++k) {
        const int item = threadIdx.x + k * kPropThreads;
        const int64_t c = int64_t(v0) * 26 + item;
        const bool ok = item < items;
        w[k] = ok ? nbr[c] : -1;
        D[k] = ok ? delays[c] : 1;
        a[k] = ok ? coefs[c] : 0.0f;
    }
    float in0[kPropItems][3], in1[kPropItems][3], prev[kPropItems][3];
#pragma unroll
    for (int k = 0; k < kPropItems; ++k) {
        const int item = threadIdx.x + k * kPropThreads;
        const int64_t c = int64_t(v0) * 26 + item;
        const bool on = w[k] >= 0;
        const int64_t col = 3LL * (on ? w[k] : 0);
        const bool l0 = on && D[k] <= t, l1 = on && D[k] < t;
        const float *r0 = x + int64_t(l0 ? t - D[k] : 0) * n + col;
        const float *r1 = x + int64_t(l1 ? t - D[k] - 1 : 0) * n + col;
#pragma unroll
        for (int o = 0; o < 3; ++o) {
            in0[k][o] = l0 ? r0[o] : 0.0f;
            in1[k][o] = l1 ? r1[o] : 0.0f;
            prev[k][o] = on ? y[c * 3 + o] : 0.0f;
        }
    }
    __syncthreads();  // mask cleared
#pragma unroll
    for (int k = 0; k < kPropItems; ++k) {
        const int item = threadIdx.x + k * kPropThreads;
        if (w[k] < 0) continue;
        const int64_t c = int64_t(v0) * 26 + item;
        const int vl = item / 26;
        atomicOr(&mask[vl], 1u << (item - vl * 26));
#pragma unroll
        for (int o = 0; o < 3; ++o) {
            if (last) y_last[c * 3 + o] = prev[k][o];
            const float yn = fmaf(a[k], in0[k][o] - prev[k][o], in1[k][o]);
            y[c * 3 + o] = yn;
            ys[item * 3 + o] = yn;
        }
    }
    __syncthreads();

    const int vl = threadIdx.x / 3;
    if (vl < nv) {
        const int64_t i = 3LL * v0 + threadIdx.x;
        float g[78];
#pragma unroll
        for (int k = 0; k < 78; ++k) g[k] = gT[int64_t(k) * n + i];
        const uint32_t m = mask[vl];
        float acc = 0.0f;
#pragma unroll
        for (int d = 0; d < 26; ++d) {
            if (!((m >> d) & 1u)) continue;
#pragma unroll
            for (int o = 0; o < 3; ++o) acc = acc + g[3 * d + o] * ys[(vl * 26 + d) * 3 + o];
        }
        x[int64_t(t) * n + i] = acc + u_t * b[i];
    }
}

// ---------------------------------------------------------------------------
// Resident propagation: one launch runs a range of time steps.
//
// A workgroup owns a contiguous slab of voxels of one job (one scenario's
// beat) for all of its steps. Its filter outputs y (78 per voxel) live in LDS
// and its connection tables (neighbour, delay, coefficient) in registers, so
// per step only the gains and the delayed neighbour states are read from
// memory. Phase 1 (one lane per connection) and phase 2 (one lane per state)
// are the same arithmetic as propagate_step_kernel. Between steps each slab
// publishes its progress in the job's flags and waits only for the slabs that
// own its neighbours.
//
// The host launches no more workgroups than fit on the device at once. Every
// wait is bounded in wall-clock time: a workgroup that waits longer raises the
// job's `abort` word, every workgroup of the job leaves at its next wait, and
// the host reports the failure instead of the kernel waiting forever.
// ---------------------------------------------------------------------------

typedef float float3_a4 __attribute__((ext_vector_type(3), aligned(4)));

constexpr int kResThreads = 512;
constexpr int kResItems = (kResMaxVoxels * 26 + kResThreads - 1) / kResThreads;
constexpr int kResStates = (kResMaxVoxels * 3 + kResThreads - 1) / kResThreads;
// Gathers issued per group in phase 1 (2, 3 or 4 time the same) and
// connections per gains group in phase 2. The gain rows of each lane's first
// state held in registers for the launch are whole groups (gains do not
// change within a beat); the measured traffic is mostly gains, so every
// cached row saves time. Per 2000-step launch at the 120k-voxel shell
// (same-run A/B): 9-connection groups with one cached (27 rows) 53.2 ms,
// none cached 55.7; 5-connection groups with two cached (30 rows) 52.2;
// 4-connection groups with three cached (36 rows) 51.4, at 4 spilled
// registers; two groups of 6 (36 rows) spill 45, of 3 or 1 connections more.
constexpr int kResBatch = 2;
constexpr int kResGainConns = 4;
constexpr int kResCachedRows = 36;
// The batched kernel also holds its job's fields: 3 gathers per group,
// 9-connection groups and one cached group of 27 rows fit without spills
// (propagation of the 64-scenario uneven batch: 230.3 ms without the cached
// rows, 215.6 ms with them). The single-scenario setting (4-connection groups,
// 36 rows) spills 29 registers here, inside the step loop: 229.8 against
// 221.8 ms (same-run A/B; 234.1 ms with 2 gathers per group).
constexpr int kBatchGather = 3;
constexpr int kBatchGainConns = 9;
constexpr int kBatchCachedRows = 27;
constexpr uint32_t kNoConnection = 0xFFFFFFFFu;
constexpr int kResNeighbourBits = 17;  // voxel index field; delay or connection in the upper 15 bits
static_assert(kSortSize >= kResMaxVoxels * 26 && kSortSize < (1 << 14) + 1, "sort size");
static_assert(kSortSize * 8 <= kResMaxVoxels * 78 * 4, "sort keys must fit over the filter outputs");

// Steps [t_begin, t_end) of the slab set up by resident_run. kCoherent: the
// job exchanges rows through agent-coherent accesses. `sorted`: the slab's
// connections run in (delay, neighbour) order with the delays in dtab; both
// forms are computed and selected, so the per-connection code has no branch.
template <int kGather, int kGainConns, int kCachedRows, bool kCoherent>
__device__ __forceinline__ bool resident_steps(
    const ResidentJob &j, int32_t vb, int32_t nv, int32_t t_begin, int32_t t_end, bool sorted,
    float *ys,
    int *ok_lds, const uint8_t *dtab, const uint32_t (&pk)[kResItems],
    const float (&ak)[kResItems], float (&pv)[kResItems][3],
    const float (&gc)[kCachedRows > 0 ? kCachedRows : 1], int32_t smin, int32_t smax,
    int32_t me, uint64_t first_wait, uint64_t step_wait) {
    const int32_t N = j.N, T = j.T;
    const float *__restrict__ gT = j.gT;
    const float *__restrict__ b = j.b;
    const float *__restrict__ u = j.u;
    float *__restrict__ x = j.x;
    // the states as one buffer resource (raw, byte offsets): with kCoherent the
    // gathers carry sc1 and do not hit a stale copy in this XCD's L2
    const __amdgpu_buffer_rsrc_t xrs = __builtin_amdgcn_make_buffer_rsrc(x, 0, -1, kBufferWord3);
    const int64_t n = 3LL * N;
    const int tid = threadIdx.x;
    for (int32_t t = t_begin; t < t_end; ++t) {
        if (t > t_begin &&
            !slab_wait(j.flags, j.gb, smin, smax, me, uint32_t(t - t_begin), ok_lds,
                       t == t_begin + 1 ? first_wait : step_wait, kCoherent))
            return false;  // host reports the abort
        if (t == T - 1) {
            // the filter outputs before the last step (y_last: [v][3d+out], the
            // layout of ys; a missing connection's +0 equals its cleared value)
            for (int idx = tid; idx < nv * 78; idx += kResThreads)
                j.y_last[int64_t(vb) * 78 + idx] = ys[idx];
            __syncthreads();
        }
#pragma unroll
        for (int k0 = 0; k0 < kResItems; k0 += kGather) {
            float in0[kGather][3];
#pragma unroll
            for (int q = 0; q < kGather; ++q) {
                const int k = k0 + q;
#pragma unroll
                for (int o = 0; o < 3; ++o) in0[q][o] = 0.0f;
                if (k >= kResItems || pk[k] == kNoConnection) continue;
                // An opaque copy: otherwise the compiler keeps every value derived
                // from pk (addresses, delays) live across the step loop and spills.
                uint32_t pko = pk[k];
                asm volatile("" : "+v"(pko));
                // 32-bit element offsets: the host only picks this kernel when T * n < 2^32
                const uint32_t w = pko & ((1u << kResNeighbourBits) - 1);
                const int32_t Ds = int32_t(dtab[tid + k * kResThreads]);
                const int32_t D = sorted ? Ds : int32_t(pko >> kResNeighbourBits);
                if (D <= t) {
                    // the neighbour's three states: one 12-byte load
                    const uint32_t r0 = uint32_t(t - D) * uint32_t(n) + 3u * w;
                    float3_a4 v;
                    if constexpr (kCoherent)
                        v = __builtin_amdgcn_raw_buffer_load_b96(xrs, r0 * 4u, 0, kCacheSC1);
                    else
                        v = *reinterpret_cast<const float3_a4 *>(x + r0);
                    in0[q][0] = v.x;
                    in0[q][1] = v.y;
                    in0[q][2] = v.z;
                }
            }
#pragma unroll
            for (int q = 0; q < kGather; ++q) {
                const int k = k0 + q;
                if (k >= kResItems || pk[k] == kNoConnection) continue;
                uint32_t pko = pk[k];
                asm volatile("" : "+v"(pko));
                const int it = sorted ? int(pko >> kResNeighbourBits) : tid + k * kResThreads;
#pragma unroll
                for (int o = 0; o < 3; ++o) {
                    const float prev = ys[it * 3 + o];
                    ys[it * 3 + o] = fmaf(ak[k], in0[q][o] - prev, pv[k][o]);
                    pv[k][o] = in0[q][o];
                }
            }
        }
        __syncthreads();
        // Phase 2 (a lane per state): x_t[i] = sum over the 78 (connection,
        // output) pairs in order of G[i][.] * y, then + u_t b[i]. A missing
        // connection's y is +0 (reset per beat, never written) and its term a
        // signed zero, which leaves the sum (started at +0) unchanged: the
        // sum over all 78 equals the reference's sum over the existing ones
        // bit for bit whenever the missing connections' gains are finite.
        const float ut = u[t];
#pragma unroll 1
        for (int k = 0; k < kResStates; ++k) {
            const int sidx = tid + k * kResThreads;
            if (sidx >= nv * 3) break;
            const float *yv = ys + (sidx / 3) * 78;
            const uint32_t i = uint32_t(3 * vb + sidx);
            float acc = 0.0f;
#pragma unroll
            for (int h = 0; h < 78; h += 3 * kGainConns) {
                float g[3 * kGainConns];
                if (h + 3 * kGainConns <= kCachedRows && k == 0) {
#pragma unroll
                    for (int q = 0; q < 3 * kGainConns; ++q)
                        g[q] = gc[h + q < kCachedRows ? h + q : 0];
                } else {
#pragma unroll
                    for (int q = 0; q < 3 * kGainConns; ++q)
                        if (h + q < 78) g[q] = (gT + size_t(h + q) * size_t(n))[i];  // row h+q, uniform
                }
#pragma unroll
                for (int q = 0; q < 3 * kGainConns; ++q)
                    if (h + q < 78) acc = acc + g[q] * yv[h + q];
            }
            // written through the L2: the other slabs read it coherently
            __hip_atomic_store(x + int64_t(t) * n + i, acc + ut * b[i], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        // filter outputs y_t for the reverse-mode gradient, every ckpt_interval steps
        const int32_t K = j.ckpt_interval;
        if (K > 0 && (t + 1) % K == 0 && t + 1 < T) {
            float *dst = j.ckpt + int64_t((t + 1) / K - 1) * N * 78 + int64_t(vb) * 78;
            for (int idx = tid; idx < nv * 78; idx += kResThreads) dst[idx] = ys[idx];
        }
        slab_publish(j.flags, me, uint32_t(t + 1 - t_begin), kCoherent);
    }
    for (int idx = tid; idx < nv * 78; idx += kResThreads) j.y[int64_t(vb) * 78 + idx] = ys[idx];
    return true;
}


// Steps [t_begin, t_end) of voxels [vb, ve), one slab of job `j`. After each
// step the slab publishes its progress in j.flags; before the next it waits
// for the slabs owning its neighbours, the first wait up to `first_wait`
// ticks (they may still be finishing earlier jobs), the others `step_wait`.
// Returns false if the job aborted (its sticky abort word is set).
template <int kGather, int kGainConns, int kCachedRows, bool kCoherent>
__device__ __forceinline__ bool resident_run(const ResidentJob &j, int32_t vb, int32_t ve,
                                             int32_t t_begin, int32_t t_end,
                                             float *lds, uint64_t first_wait, uint64_t step_wait) {
    float *ys = lds;                                                       // [voxel][78]
    int *ok_lds = reinterpret_cast<int *>(lds + kResMaxVoxels * 78);     // [4]
    uint8_t *dtab = reinterpret_cast<uint8_t *>(ok_lds + 4);                // [sorted position]
    uint64_t *keys = reinterpret_cast<uint64_t *>(lds);  // set-up only, over ys
    const int32_t N = j.N, T = j.T;
    const int32_t *__restrict__ nbr = j.nbr;
    const float *__restrict__ gT = j.gT;
    const float *__restrict__ x = j.x;
    const int64_t n = 3LL * N;
    const int32_t nv = ve - vb;
    const int32_t items = nv * 26;
    const int tid = threadIdx.x;

    // Phase 1 runs the slab's connections in (delay, neighbour) order when
    // every delay fits the byte table: the lanes of one gather instruction
    // then read one state row at nearby voxels, about 29 cache lines per wave
    // instead of about 61 (one per lane). Key = D << 40 | w << 14 | local
    // connection; a missing connection sorts last.
    if (tid == 0) {
        ok_lds[1] = 0;        // largest delay
        ok_lds[2] = N;        // smallest neighbour
        ok_lds[3] = -1;       // largest neighbour
    }
    __syncthreads();
    for (int p = tid; p < kSortSize; p += kResThreads) {
        uint64_t key = ~0ull;
        if (p < items) {
            const int64_t c = int64_t(vb) * 26 + p;
            const int32_t w = nbr[c];
            if (w >= 0) {
                const int32_t D = min(j.delays[c], 32767);
                key = (uint64_t(D) << 40) | (uint64_t(w) << 14) | uint64_t(p);
                atomicMax(&ok_lds[1], D);
                atomicMin(&ok_lds[2], w);
                atomicMax(&ok_lds[3], w);
            }
        }
        keys[p] = key;
    }
    __syncthreads();
    const bool sorted = ok_lds[1] <= kSortedMaxDelay;
    // With slab flags, the slabs this one waits for: those owning its
    // neighbours (slab s holds voxels [s N / G, (s+1) N / G)).
    const int32_t G = j.width;
    const auto slab_of = [N, G](int32_t w) { return int32_t((int64_t(w + 1) * G - 1) / N); };
    const int32_t me = slab_of(vb);
    const int32_t smin = ok_lds[3] >= 0 ? slab_of(ok_lds[2]) : 1;
    const int32_t smax = ok_lds[3] >= 0 ? slab_of(ok_lds[3]) : 0;
    if (sorted) sort_keys_lds<kResThreads, kSortSize>(keys);
    // pk = neighbour (17 bits) | delay (15 bits) in connection order, or
    // neighbour | local connection (14 bits) in sorted order with the delay in
    // dtab[position].
    uint32_t pk[kResItems];
    float ak[kResItems];
    // pv = the neighbour states this connection read one step earlier, i.e.
    // x[t-D-1] at step t: only x[t-D] is gathered per step.
    float pv[kResItems][3];
#pragma unroll
    for (int k = 0; k < kResItems; ++k) {
        const int it = tid + k * kResThreads;
        pk[k] = kNoConnection;
        ak[k] = 0.0f;
#pragma unroll
        for (int o = 0; o < 3; ++o) pv[k][o] = 0.0f;
        if (it < items) {
            int32_t w = -1, D = 0, cl = it;
            if (sorted) {
                const uint64_t key = keys[it];
                if (key != ~0ull) {
                    D = int32_t(key >> 40);
                    w = int32_t((key >> 14) & 0x3FFFFFFu);
                    cl = int32_t(key & 0x3FFFu);
                    dtab[it] = uint8_t(D);
                    pk[k] = uint32_t(w) | (uint32_t(cl) << kResNeighbourBits);
                }
            } else {
                const int64_t c = int64_t(vb) * 26 + it;
                w = nbr[c];
                if (w >= 0) {
                    D = min(j.delays[c], 32767);
                    pk[k] = uint32_t(w) | (uint32_t(D) << kResNeighbourBits);
                }
            }
            if (w >= 0) {
                ak[k] = j.coefs[int64_t(vb) * 26 + cl];
                if (D < t_begin) {
                    const int64_t r1 = int64_t(t_begin - D - 1) * n + 3LL * w;
#pragma unroll
                    for (int o = 0; o < 3; ++o) pv[k][o] = x[r1 + o];
                }
            }
        }
    }
    // The first kCachedRows gains of the lane's first state stay in registers
    // for the whole launch (gains do not change within a beat).
    float gc[kCachedRows > 0 ? kCachedRows : 1];
#pragma unroll
    for (int q = 0; q < kCachedRows; ++q)
        gc[q] = tid < nv * 3 ? gT[int64_t(q) * n + 3LL * vb + tid] : 0.0f;
    __syncthreads();  // the keys are dead: ys is loaded over them
    for (int idx = tid; idx < nv * 78; idx += kResThreads) ys[idx] = j.y[int64_t(vb) * 78 + idx];
    __syncthreads();

    return resident_steps<kGather, kGainConns, kCachedRows, kCoherent>(
        j, vb, nv, t_begin, t_end, sorted, ys, ok_lds, dtab, pk, ak, pv, gc, smin, smax, me,
        first_wait, step_wait);
}

// One job over the whole grid: workgroup slab = xcd_block(blockIdx.x).
// kCoherent = job.coherent != 0 (the launcher picks the instantiation).
template <bool kCoherent>
__global__ __launch_bounds__(kResThreads) void propagate_resident_kernel(
    ResidentJob job, int32_t t_begin, int32_t t_end, uint64_t wait_ticks) {
    extern __shared__ float lds[];
    const int64_t slab = xcd_block(blockIdx.x, gridDim.x);
    const int32_t vb = int32_t(slab * job.N / gridDim.x);
    const int32_t ve = int32_t((slab + 1) * job.N / gridDim.x);
    resident_run<kResBatch, kResGainConns, kResCachedRows, kCoherent>(
        job, vb, ve, t_begin, t_end, lds, wait_ticks, wait_ticks);
}

// Many jobs in one persistent launch. Logical slot = xcd_block(blockIdx.x), so
// consecutive slots share an XCD. Slot q runs the tasks
// task_job[task_ptr[q] .. task_ptr[q+1]) in that order; a job occupies the
// contiguous slots [slot0, slot0 + width) and member (q - slot0) owns voxels
// [m N / width, (m+1) N / width). Every slot lists its jobs in one global
// order, so the earliest unfinished job always has its whole team and the
// launch cannot deadlock. kCoherent: every job exchanges its rows coherently
// (job.coherent != 0); otherwise every job takes the fenced exchange, which
// is correct for any job.
template <bool kCoherent>
__global__ __launch_bounds__(kResThreads) void propagate_batch_kernel(
    const ResidentJob *__restrict__ jobs, const int32_t *__restrict__ task_ptr,
    const int32_t *__restrict__ task_job, uint64_t first_wait, uint64_t step_wait) {
    extern __shared__ float lds[];
    const int32_t slot = xcd_block(blockIdx.x, gridDim.x);
    const int32_t q1 = task_ptr[slot + 1];
    for (int32_t q = task_ptr[slot]; q < q1; ++q) {
        const ResidentJob job = jobs[task_job[q]];
        const int32_t member = slot - job.slot0;
        const int32_t vb = int32_t(int64_t(member) * job.N / job.width);
        const int32_t ve = int32_t(int64_t(member + 1) * job.N / job.width);
        if (!resident_run<kBatchGather, kBatchGainConns, kBatchCachedRows, kCoherent>(job, vb, ve, 0, job.T, lds,
                                                                   first_wait, step_wait))
            return;
        __syncthreads();  // LDS is reused by the next job
    }
}

// ---------------------------------------------------------------------------
// Projection H·Xᵀ over all steps (prediction.rs:112-130 for every t), FP32
// MFMA 16x16x4 over tiles of kProjRows = 112 sensors (seven blocks of 16), so
// 108 sensors waste 4 rows instead of the 20 a 32-row granularity would.
// Workgroup (t-tile, split, sensor tile) covers kProjTile steps x 112 sensors
// over its K range and writes that partial product; launch_residuals sums the
// splits in fixed order. Wave w computes steps [w kProjTile / 4, (w + 1)
// kProjTile / 4) of the tile for all 112 sensors: at 128 steps, 14
// accumulators and 9 LDS operand reads per 14 MFMAs. (64-step tiles time the
// same; 256-step tiles need 178 VGPRs and run 38% slower.)
// ---------------------------------------------------------------------------
constexpr int kProjRowsPad = kProjRows + 32;   // H tile row stride: rows k, k+1 on opposite bank halves
constexpr int kProjStepsPad = kProjTile + 16;  // X tile row stride, the same property
constexpr int kProjWaveSteps = kProjTile / 4;  // steps of the tile per wave
constexpr int kProjStepBlocks = kProjWaveSteps / 16;
static_assert(kProjTile % 64 == 0, "whole 16-step MFMA blocks per wave");
__global__ __launch_bounds__(256) void project_kernel(const float *__restrict__ x,
                                                      const float *__restrict__ H,
                                                      float *__restrict__ partial,
                                                      float *__restrict__ xT, int32_t T,
                                                      int64_t n, int32_t Sp, int64_t kps) {
    // k-major tiles: a 16x16x4 operand read takes 16 consecutive floats of
    // four k rows; strides = 16 mod 32 words put the two rows of each 32-lane
    // half on disjoint banks
    __shared__ float As[kProjTileK][kProjStepsPad];  // X[t][k], kProjTile steps
    __shared__ float Bs[kProjTileK][kProjRowsPad];   // H[s][k], kProjRows sensors
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // Workgroups are numbered (step tile fastest, split, sensor tile) and
    // remapped XCD-contiguously: the step tiles of one split, which read the
    // same H columns, run on one XCD and share its L2.
    const int32_t nb = gridDim.x * gridDim.y * gridDim.z;
    const int32_t lin =
        xcd_block(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), nb);
    const int32_t t0 = (lin % gridDim.x) * kProjTile;
    const int32_t split = (lin / gridDim.x) % gridDim.y;
    const int32_t s0 = (lin / (gridDim.x * gridDim.y)) * kProjRows;
    const int64_t k0 = int64_t(split) * kps;
    const int64_t k1 = min(n, k0 + kps);
    typedef float floatx4 __attribute__((ext_vector_type(4)));
    floatx4 acc[kProjRows / 16][kProjStepBlocks];
#pragma unroll
    for (int sb = 0; sb < kProjRows / 16; ++sb)
#pragma unroll
        for (int tb = 0; tb < kProjStepBlocks; ++tb) acc[sb][tb] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    // loader: kProjTileK / 4 lanes cover one row's k values (kProjTileK / 32
    // lines of 128 bytes). The next k tile is loaded into registers while the
    // MFMAs run on this one.
    constexpr int kRowLanes = kProjTileK / 4, kRowsPerPass = 256 / kRowLanes;
    const int lrow = tid / kRowLanes, lk = (tid % kRowLanes) * 4;
    constexpr int kXPasses = kProjTile / kRowsPerPass;
    constexpr int kHPasses = (kProjRows + kRowsPerPass - 1) / kRowsPerPass;
    float4_a4 xa[kXPasses], ha[kHPasses];
    auto load_tile = [&](int64_t kb) __attribute__((always_inline)) {
        const bool full = kb + kProjTileK <= k1;
#pragma unroll
        for (int pass = 0; pass < kXPasses; ++pass) {
            const int32_t t = t0 + pass * kRowsPerPass + lrow;
            float4_a4 v = {0.0f, 0.0f, 0.0f, 0.0f};
            if (t < T) {
                const float *src = x + int64_t(t) * n + kb + lk;
                if (full) {
                    v = *reinterpret_cast<const float4_a4 *>(src);
                } else {
#pragma unroll
                    for (int j = 0; j < 4; ++j) v[j] = (kb + lk + j < k1) ? src[j] : 0.0f;
                }
            }
            xa[pass] = v;
        }
#pragma unroll
        for (int pass = 0; pass < kHPasses; ++pass) {
            const int r = pass * kRowsPerPass + lrow, srow = s0 + r;
            float4_a4 v = {0.0f, 0.0f, 0.0f, 0.0f};
            if (r < kProjRows && srow < Sp) {
                const float *src = H + int64_t(srow) * n + kb + lk;
                if (full) {
                    v = *reinterpret_cast<const float4_a4 *>(src);
                } else {
#pragma unroll
                    for (int j = 0; j < 4; ++j) v[j] = (kb + lk + j < k1) ? src[j] : 0.0f;
                }
            }
            ha[pass] = v;
        }
    };
    if (k0 < k1) load_tile(k0);
    const int li = lane & 15, lq = lane >> 4;
    for (int64_t kb = k0; kb < k1; kb += kProjTileK) {
#pragma unroll
        for (int pass = 0; pass < kXPasses; ++pass)
#pragma unroll
            for (int j = 0; j < 4; ++j) As[lk + j][pass * kRowsPerPass + lrow] = xa[pass][j];
#pragma unroll
        for (int pass = 0; pass < kHPasses; ++pass)
            if (pass * kRowsPerPass + lrow < kProjRows)
#pragma unroll
                for (int j = 0; j < 4; ++j) Bs[lk + j][pass * kRowsPerPass + lrow] = ha[pass][j];
        __syncthreads();
        if (kb + kProjTileK < k1) load_tile(kb + kProjTileK);
        if (xT && s0 == 0) {
            // the k-major tile is the states-by-step layout: one pass over x
            // also writes xT (row k: kRun consecutive steps per thread)
            constexpr int kRowThreads = 256 / kProjTileK, kRun = kProjTile / kRowThreads;
            const int kr = tid / kRowThreads, c0 = (tid % kRowThreads) * kRun;
            const int64_t k = kb + kr;
            if (k < k1) {
                float *dst = xT + k * T + t0 + c0;
                if (t0 + c0 + kRun <= T) {
#pragma unroll
                    for (int q = 0; q < kRun / 4; ++q) {
                        float4_a4 v;
#pragma unroll
                        for (int j = 0; j < 4; ++j) v[j] = As[kr][c0 + 4 * q + j];
                        *reinterpret_cast<float4_a4 *>(dst + 4 * q) = v;
                    }
                } else {
                    for (int c = 0; c < kRun && t0 + c0 + c < T; ++c) dst[c] = As[kr][c0 + c];
                }
            }
        }
        // operands of one k step (4 k values): the step blocks' X values and
        // the sensor blocks' H values; those of step kq + 4 are read while
        // the MFMAs of step kq run (the same MFMAs in the same order)
        float bv[2][kProjStepBlocks], av[2][kProjRows / 16];
        const auto read_operands = [&](int kq, int slot) __attribute__((always_inline)) {
            const int kr = kq + lq;
#pragma unroll
            for (int tb = 0; tb < kProjStepBlocks; ++tb)
                bv[slot][tb] = As[kr][wave * kProjWaveSteps + tb * 16 + li];
#pragma unroll
            for (int sb = 0; sb < kProjRows / 16; ++sb) av[slot][sb] = Bs[kr][sb * 16 + li];
        };
        read_operands(0, 0);
        // interleave the operand reads with the MFMAs (2.25 -> 2.16 ms)
        __builtin_amdgcn_iglp_opt(1);
#pragma unroll
        for (int kq = 0; kq < kProjTileK; kq += 4) {
            const int cur = (kq / 4) & 1;
            if (kq + 4 < kProjTileK) read_operands(kq + 4, cur ^ 1);
#pragma unroll
            for (int sb = 0; sb < kProjRows / 16; ++sb)
#pragma unroll
                for (int tb = 0; tb < kProjStepBlocks; ++tb)
                    acc[sb][tb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[cur][sb], bv[cur][tb],
                                                                       acc[sb][tb], 0, 0, 0);
        }
        __syncthreads();
    }
    // lane l holds D[4 (l / 16) + r][l % 16]: sensor rows, step columns
#pragma unroll
    for (int sb = 0; sb < kProjRows / 16; ++sb)
#pragma unroll
        for (int tb = 0; tb < kProjStepBlocks; ++tb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int32_t s = s0 + sb * 16 + 4 * lq + r;
                const int32_t t = t0 + wave * kProjWaveSteps + tb * 16 + li;
                if (t < T && s < Sp) partial[(int64_t(split) * T + t) * Sp + s] = acc[sb][tb][r];
            }
}

// Residuals and per-step MSE (estimation.rs:146-152, metrics.rs:189-190).
// One workgroup per step; the squared residuals are summed in the 8-way
// interleaved order of ndarray's Array::sum.
__global__ void residuals_kernel(const float *__restrict__ partial, int32_t splits,
                                 const float *__restrict__ mhat, float *__restrict__ m,
                                 float *__restrict__ R, float *__restrict__ loss_mse, int32_t T,
                                 int32_t S, int32_t Sp) {
    extern __shared__ float sq[];
    const int32_t t = blockIdx.x;
    for (int32_t s = threadIdx.x; s < Sp; s += blockDim.x) {
        float acc = 0.0f;
        for (int32_t k = 0; k < splits; ++k) acc += partial[(int64_t(k) * T + t) * Sp + s];
        if (s < S) m[int64_t(t) * S + s] = acc;
        if (!mhat) continue;  // forward-only simulation: predictions only
        if (s < S) {
            const float r = acc - mhat[int64_t(t) * S + s];
            R[int64_t(t) * Sp + s] = r;
            sq[s] = r * r;
        } else {
            R[int64_t(t) * Sp + s] = 0.0f;
        }
    }
    __syncthreads();
    if (mhat && threadIdx.x == 0) {
        float p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        int32_t s = 0;
        for (; S - s >= 8; s += 8)
            for (int j = 0; j < 8; ++j) p[j] = p[j] + sq[s + j];
        float acc = 0.0f;
        acc = acc + (p[0] + p[4]);
        acc = acc + (p[1] + p[5]);
        acc = acc + (p[2] + p[6]);
        acc = acc + (p[3] + p[7]);
        for (; s < S; ++s) acc = acc + sq[s];
        loss_mse[t] = acc / float(S);
    }
}

// Maximum regularisation (derivation.rs:509-535): per step, the sum over
// voxels of (|x|_1 - threshold)^2 where positive; partial sums per workgroup
// of kMaxregVoxels voxels, four per thread 256 apart (each load instruction
// stays coalesced), summed in the thread in that order, then in a fixed tree.
__global__ __launch_bounds__(256) void maxreg_partials_kernel(const float *__restrict__ x,
                                                              float *__restrict__ partial,
                                                              float *__restrict__ mr_last,
                                                              int32_t N, int32_t T,
                                                              float threshold) {
    constexpr int kPerThread = kMaxregVoxels / 256;
    __shared__ float red[256];
    const int32_t t = blockIdx.y;
    const float *row = x + int64_t(t) * 3 * N;
    float xs[kPerThread][3];
#pragma unroll
    for (int k = 0; k < kPerThread; ++k) {
        const int32_t v = int32_t(blockIdx.x) * kMaxregVoxels + k * 256 + int32_t(threadIdx.x);
#pragma unroll
        for (int o = 0; o < 3; ++o) xs[k][o] = v < N ? row[3LL * v + o] : 0.0f;
    }
    float f2 = 0.0f;
#pragma unroll
    for (int k = 0; k < kPerThread; ++k) {
        const int32_t v = int32_t(blockIdx.x) * kMaxregVoxels + k * 256 + int32_t(threadIdx.x);
        if (v >= N) break;
        const float sum = fabsf(xs[k][0]) + fabsf(xs[k][1]) + fabsf(xs[k][2]);
        float f = 0.0f;
        const bool on = sum > threshold;
        if (on) {
            f = sum - threshold;
            f2 = f2 + f * f;
        }
        if (t == T - 1) {
            float *mv = mr_last + 3LL * v;
#pragma unroll
            for (int o = 0; o < 3; ++o) mv[o] = on ? f * signum(xs[k][o]) : 0.0f;
        }
    }
    red[threadIdx.x] = f2;
    __syncthreads();
    for (int stride = 128; stride > 0; stride >>= 1) {
        if (int(threadIdx.x) < stride) red[threadIdx.x] += red[threadIdx.x + stride];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[int64_t(t) * gridDim.x + blockIdx.x] = red[0];
}

// Per-step sum of the workgroup partials (fixed tree order, one workgroup per
// step).
__global__ __launch_bounds__(256) void maxreg_step_sums_kernel(const float *__restrict__ partial,
                                                               int32_t blocks,
                                                               float *__restrict__ step_mr) {
    __shared__ float red[256];
    const int32_t t = blockIdx.x;
    float s = 0.0f;
    for (int32_t k = threadIdx.x; k < blocks; k += 256) s += partial[int64_t(t) * blocks + k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int stride = 128; stride > 0; stride >>= 1) {
        if (int(threadIdx.x) < stride) red[threadIdx.x] += red[threadIdx.x + stride];
        __syncthreads();
    }
    if (threadIdx.x == 0) step_mr[t] = red[0];
}

// Running maximum-regularisation sum (cumulative over steps and beats,
// derivation.rs:523) and the per-step losses (metrics.rs:191-195).
// Also the objective of the reverse-mode gradient (in double):
// J = sum_t mse_strength/2 * loss_mse[t] + reg_strength/2 * step_mr[t].
// The running maximum-regularisation sum and the per-step losses
// (metrics.rs:189-198): a sequential scan in the reference's order, so one
// thread runs it; the inputs are staged in LDS by the workgroup chunk by
// chunk, so the scan does not wait on a global load per step.
constexpr int kScanChunk = 2048;
__global__ __launch_bounds__(256) void maxreg_prefix_kernel(
    const float *__restrict__ step_mr, float *__restrict__ mr_sum,
    const float *__restrict__ loss_mse, float *__restrict__ loss_mr, float *__restrict__ loss,
    double *__restrict__ objective, float mse_strength, float reg_strength, int32_t T) {
    __shared__ float smr[kScanChunk], smse[kScanChunk];
    const int tid = threadIdx.x;
    float cum = 0.0f;
    double J = 0.0;
    if (tid == 0) cum = *mr_sum;
    for (int32_t c0 = 0; c0 < T; c0 += kScanChunk) {
        const int32_t len = min(kScanChunk, T - c0);
        for (int32_t i = tid; i < len; i += blockDim.x) {
            smr[i] = step_mr[c0 + i];
            smse[i] = loss_mse[c0 + i];
        }
        __syncthreads();
        if (tid == 0)
            for (int32_t i = 0; i < len; ++i) {
                cum = cum + smr[i];
                loss_mr[c0 + i] = cum;
                loss[c0 + i] = fmaf(reg_strength, cum, smse[i]);
                J += 0.5 * double(mse_strength) * double(smse[i]) +
                     0.5 * double(reg_strength) * double(smr[i]);
            }
        __syncthreads();
    }
    if (tid == 0) {
        *mr_sum = cum;
        *objective = J;
    }
}

__global__ __launch_bounds__(256) void transpose_kernel(const float *__restrict__ in,
                                                        float *__restrict__ out, int64_t rows,
                                                        int64_t cols) {
    __shared__ float tile[32][33];
    const int64_t c0 = int64_t(blockIdx.x) * 32, r0 = int64_t(blockIdx.y) * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int j = ty; j < 32; j += 8) {
        const int64_t r = r0 + j, c = c0 + tx;
        if (r < rows && c < cols) tile[j][tx] = in[r * cols + c];
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        const int64_t c = c0 + j, r = r0 + tx;
        if (r < rows && c < cols) out[c * rows + r] = tile[tx][j];
    }
}

// ---------------------------------------------------------------------------
// Mapped residuals Hᵀ·Rᵀ for up to 112 sensors (derivation.rs:538-551 for
// every t), FP32 MFMA 32x32x2. A workgroup owns 128 states (32 per wave) and
// a span of steps. Each lane loads its A operands for the whole contraction
// once, H[2p + lane/32][its state] for every sensor pair p (56 registers), so
// H is read from memory once per span; the loop over the span's 64-step
// tiles then stages only R (L2-resident, 1 MB at T = 2000) in LDS, and the
// next tile's R is loaded while the MFMAs run. The k pairs go in ascending
// order, as in map_residuals_kernel, so both give the same bits.
// kByStep: the output is E = (Hᵀ·Rᵀ)ᵀ, (T, n) by step, for the reverse-mode
// sweep: the same MFMAs with the operands swapped (each output element sums
// the same products in the same order, so E[t][i] == eT[i][t] bit for bit),
// which puts 32 consecutive states of one step in a lane group.
// ---------------------------------------------------------------------------
constexpr int kMapRegPairs = kProjRows / 2;  // sensor pairs held in registers (S <= 112)
constexpr int kMapRegTile = 64;              // steps per tile: two 32-step MFMA blocks
constexpr int kMapRegSpan = 16;              // tiles per workgroup (1024 steps)
constexpr int kMapRegStride = kMapRegTile + 4;  // LDS row (one sensor) in floats
constexpr int kMapRegQuads = kProjRows / 16;    // float4 of R per thread and tile
static_assert(kMapRegTile == 64 && kProjRows % 16 == 0, "R tile: 64 steps x 4 threads per step");
template <bool kByStep>
__global__ __launch_bounds__(256) void map_residuals_regs_kernel(const float *__restrict__ H,
                                                                 const float *__restrict__ R,
                                                                 float *__restrict__ eT, int64_t n,
                                                                 int32_t T, int32_t Se,
                                                                 int32_t Sp, int32_t accumulate) {
    // R tile, k-major: sensor k, tile step u = 32 q + l at Rs[k][2 l + q], so a
    // lane's operands of both 32-step blocks are one 8-byte read
    __shared__ float Rs[kProjRows][kMapRegStride];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t i0 = int64_t(blockIdx.x) * 128;
    const int64_t iw = i0 + wave * 32 + (lane & 31);
    const int kh = lane >> 5;
    const int32_t tb = blockIdx.y * kMapRegSpan * kMapRegTile;
    const int32_t te = min(T, tb + kMapRegSpan * kMapRegTile);
    // unconditional loads from clamped addresses (all in flight at once), then
    // zeros where the pair or the state is out of range
    float a[kMapRegPairs];
    const float *hcol = H + min(iw, n - 1) + int64_t(kh) * n;
#pragma unroll
    for (int p = 0; p < kMapRegPairs; ++p) a[p] = hcol[int64_t(min(2 * p, Se - 2)) * n];
#pragma unroll
    for (int p = 0; p < kMapRegPairs; ++p) a[p] = (2 * p < Se && iw < n) ? a[p] : 0.0f;
    // eT rows [i0, min(n, i0 + 128)) as one buffer: stores past them are dropped
    // (by step: E from row t0 of the tile, columns from i0, bounds checked per lane)
    const int64_t rows = kByStep ? 0 : min(int64_t(128), n - i0);
    const __amdgpu_buffer_rsrc_t ers =
        __builtin_amdgcn_make_buffer_rsrc(eT + i0 * T, 0, int32_t(4 * rows * T), kBufferWord3);
    const int32_t lane_off = kByStep ? 4 * (kh * 4 * int32_t(n) + wave * 32 + (lane & 31))
                                     : 4 * ((wave * 32 + kh * 4) * T + (lane & 31));
    // R loads: thread = tile step tt (4 threads per step), sensors 16 m + 4 c .. + 3
    const int tt = tid >> 2, c4 = (tid & 3) * 4;
    float *const rs_w = &Rs[c4][2 * (tt & 31) + (tt >> 5)];
    float4_a4 rv[kMapRegQuads];
    auto load_r = [&](int32_t t0) __attribute__((always_inline)) {
        const bool in_t = t0 + tt < T;
        const float *src = R + int64_t(t0 + tt) * Sp + c4;
#pragma unroll
        for (int m = 0; m < kMapRegQuads; ++m) {
            float4_a4 v = {0.0f, 0.0f, 0.0f, 0.0f};
            if (in_t && 16 * m + c4 < Se) v = *reinterpret_cast<const float4_a4 *>(src + 16 * m);
            rv[m] = v;
        }
    };
    load_r(tb);
    const float *const rs_r = &Rs[kh][2 * (lane & 31)];
    for (int32_t t0 = tb; t0 < te; t0 += kMapRegTile) {
        __syncthreads();  // the previous tile's operand reads are done
#pragma unroll
        for (int m = 0; m < kMapRegQuads; ++m)
#pragma unroll
            for (int j = 0; j < 4; ++j) rs_w[(16 * m + j) * kMapRegStride] = rv[m][j];
        __syncthreads();
        if (t0 + kMapRegTile < te) load_r(t0 + kMapRegTile);
        floatx16 acc0 = floatx16{0}, acc1 = floatx16{0};
        // The operands of pair p + 2 are read (unconditionally: every row is
        // in the tile) while pair p's MFMAs run; the MFMAs stop at Se. The
        // loop has no break, so it unrolls and a[] stays in registers.
        // an opaque copy: otherwise the 56 pair conditions are hoisted out of
        // the tile loop and spill
        int32_t se = Se;
        asm volatile("" : "+s"(se));
        float2 bq[kMapRegPairs];
#pragma unroll
        for (int p = 0; p < 2; ++p)
            bq[p] = *reinterpret_cast<const float2 *>(rs_r + 2 * p * kMapRegStride);
#pragma unroll
        for (int p = 0; p < kMapRegPairs; ++p) {
            if (p + 2 < kMapRegPairs)
                bq[p + 2] = *reinterpret_cast<const float2 *>(rs_r + 2 * (p + 2) * kMapRegStride);
            if (2 * p < se) {
                if constexpr (kByStep) {
                    acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(bq[p].x, a[p], acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(bq[p].y, a[p], acc1, 0, 0, 0);
                } else {
                    acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a[p], bq[p].x, acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a[p], bq[p].y, acc1, 0, 0, 0);
                }
            }
        }
        if constexpr (kByStep) {
            // lane l holds D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32]: step rows,
            // state columns, stored from row t0 through a buffer resource with
            // the row part of each offset uniform
            const __amdgpu_buffer_rsrc_t trs =
                __builtin_amdgcn_make_buffer_rsrc(eT + int64_t(t0) * n + i0, 0, -1, kBufferWord3);
            const bool in_i = iw < n;
#pragma unroll
            for (int q = 0; q < 2; ++q)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int32_t row = q * 32 + (r >> 2) * 8 + (r & 3);  // + 4 kh in lane_off
                    if (!in_i || t0 + row + 4 * kh >= te) continue;
                    const int32_t so = 4 * row * int32_t(n);
                    float v = q == 0 ? acc0[r] : acc1[r];
                    if (accumulate)
                        v += __uint_as_float(
                            __builtin_amdgcn_raw_buffer_load_b32(trs, lane_off, so, 0));
                    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), trs, lane_off, so, 0);
                }
            continue;
        }
        // lane l holds D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32]: state rows, step
        // columns. Stored through the workgroup's buffer resource: the row part
        // of each offset is uniform (scalar), the lane part one register.
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            if (t0 + q * 32 + (lane & 31) >= te) continue;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int32_t so = 4 * (((r >> 2) * 8 + (r & 3)) * T + t0 + q * 32);
                float v = q == 0 ? acc0[r] : acc1[r];
                if (accumulate)
                    v += __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(ers, lane_off, so, 0));
                __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), ers, lane_off, so, 0);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// The reverse-mode forcing of an update batch that shares one propagation:
// E = sum over its beats b of H_bᵀ·R_bᵀ, by step (T, n), up to 112 sensors,
// in one pass. The per-beat kernel would add each beat's product into E in
// memory (a read and a write of 4 n T bytes per beat); here the products of
// all beats accumulate in the MFMA accumulators. A workgroup owns 128 states
// and 128 steps (two 64-step tiles, both accumulators live); per beat
// it loads its lanes' H operands once (56 registers) and runs the two tiles
// with R staged in LDS as in map_residuals_regs_kernel. The sensor pairs go
// in ascending order within a beat, the beats in the order given.
// ---------------------------------------------------------------------------
constexpr int kMapBeatTiles = 2;  // 64-step tiles per workgroup
__global__ __launch_bounds__(256) void map_residuals_beats_kernel(
    const float *__restrict__ H, MapBeats mb, const float *__restrict__ R, int64_t r_stride,
    float *__restrict__ E, int64_t n, int32_t T, int32_t Se, int32_t Sp, int32_t accumulate) {
    __shared__ float Rs[kProjRows][kMapRegStride];  // as in map_residuals_regs_kernel
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int kh = lane >> 5;
    // step spans fastest, remapped XCD-contiguously: the spans of one state
    // tile, which load the same H columns, share an L2
    const int32_t spans = (T + kMapBeatTiles * kMapRegTile - 1) / (kMapBeatTiles * kMapRegTile);
    const int32_t lin = xcd_block(blockIdx.x, gridDim.x);
    const int64_t i0 = int64_t(lin / spans) * 128;
    const int32_t tb = (lin % spans) * kMapBeatTiles * kMapRegTile;
    const int32_t te = min(T, tb + kMapBeatTiles * kMapRegTile);
    const int64_t iw = i0 + wave * 32 + (lane & 31);
    const float *hcol = H + min(iw, n - 1) + int64_t(kh) * n;
    const int tt = tid >> 2, c4 = (tid & 3) * 4;
    float *const rs_w = &Rs[c4][2 * (tt & 31) + (tt >> 5)];
    const float *const rs_r = &Rs[kh][2 * (lane & 31)];
    floatx16 acc[kMapBeatTiles][2];
#pragma unroll
    for (int j = 0; j < kMapBeatTiles; ++j) acc[j][0] = acc[j][1] = floatx16{0};
    for (int32_t b = 0; b < mb.count; ++b) {
        float a[kMapRegPairs];
        const float *hb = hcol + mb.h_off[b];
#pragma unroll
        for (int p = 0; p < kMapRegPairs; ++p) a[p] = hb[int64_t(min(2 * p, Se - 2)) * n];
#pragma unroll
        for (int p = 0; p < kMapRegPairs; ++p) a[p] = (2 * p < Se && iw < n) ? a[p] : 0.0f;
        const float *rb = R + int64_t(b) * r_stride;
#pragma unroll
        for (int j = 0; j < kMapBeatTiles; ++j) {
            const int32_t t0 = tb + j * kMapRegTile;
            if (t0 >= te) continue;  // (no break: acc[] must stay in registers)
            float4_a4 rv[kMapRegQuads];
            {
                const bool in_t = t0 + tt < T;
                const float *src = rb + int64_t(t0 + tt) * Sp + c4;
#pragma unroll
                for (int m = 0; m < kMapRegQuads; ++m) {
                    float4_a4 v = {0.0f, 0.0f, 0.0f, 0.0f};
                    if (in_t && 16 * m + c4 < Se)
                        v = *reinterpret_cast<const float4_a4 *>(src + 16 * m);
                    rv[m] = v;
                }
            }
            __syncthreads();  // the previous tile's operand reads are done
#pragma unroll
            for (int m = 0; m < kMapRegQuads; ++m)
#pragma unroll
                for (int q = 0; q < 4; ++q) rs_w[(16 * m + q) * kMapRegStride] = rv[m][q];
            __syncthreads();
            int32_t se = Se;
            asm volatile("" : "+s"(se));
            float2 bq[kMapRegPairs];
#pragma unroll
            for (int p = 0; p < 2; ++p)
                bq[p] = *reinterpret_cast<const float2 *>(rs_r + 2 * p * kMapRegStride);
#pragma unroll
            for (int p = 0; p < kMapRegPairs; ++p) {
                if (p + 2 < kMapRegPairs)
                    bq[p + 2] =
                        *reinterpret_cast<const float2 *>(rs_r + 2 * (p + 2) * kMapRegStride);
                if (2 * p < se) {
                    acc[j][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(bq[p].x, a[p], acc[j][0], 0, 0, 0);
                    acc[j][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(bq[p].y, a[p], acc[j][1], 0, 0, 0);
                }
            }
        }
    }
    // lane l holds D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32]: step rows, state
    // columns, stored from row tb through a buffer resource with the row part
    // of each offset uniform
    const __amdgpu_buffer_rsrc_t trs =
        __builtin_amdgcn_make_buffer_rsrc(E + int64_t(tb) * n + i0, 0, -1, kBufferWord3);
    const int32_t lane_off = 4 * (kh * 4 * int32_t(n) + wave * 32 + (lane & 31));
    const bool in_i = iw < n;
#pragma unroll
    for (int j = 0; j < kMapBeatTiles; ++j)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int32_t row = j * kMapRegTile + q * 32 + (r >> 2) * 8 + (r & 3);  // + 4 kh
                if (!in_i || tb + row + 4 * kh >= te) continue;
                const int32_t so = 4 * row * int32_t(n);
                float v = acc[j][q][r];
                if (accumulate)
                    v += __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(trs, lane_off, so, 0));
                __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), trs, lane_off, so, 0);
            }
}

// ---------------------------------------------------------------------------
// Mapped residuals Hᵀ·Rᵀ for all steps (derivation.rs:538-551 for every t),
// FP32 MFMA 32x32x2, K = the sensors, S rounded up to even (the padded rows
// of H and columns of R, all zero, are not multiplied). Output eT is (n, T).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void map_residuals_kernel(const float *__restrict__ H,
                                                            const float *__restrict__ R,
                                                            float *__restrict__ eT, int64_t n,
                                                            int32_t T, int32_t Se, int32_t Sp,
                                                            int32_t accumulate) {
    // tile: 128 states x kMapTile steps; k (sensors) in chunks of 32, k-major
    // LDS tiles (rows padded by 4 floats)
    __shared__ float Hs[kGemmTileK][128 + 4];       // H[s][i]
    __shared__ float Rs[kGemmTileK][kMapTile + 4];  // R[t][s]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t i0 = int64_t(blockIdx.x) * 128;
    const int32_t t0 = blockIdx.y * kMapTile;
    floatx16 acc[kMapTile / 32];
#pragma unroll
    for (int q = 0; q < kMapTile / 32; ++q) acc[q] = floatx16{0};
    const bool full_i = i0 + 128 <= n;
    // H rows of a k tile: 32 lanes cover one row; R rows: 8 lanes per row. The
    // next k tile is loaded into registers while the MFMAs run on this one.
    const int hrow = tid >> 5, hq = (tid & 31) * 4;
    const int rrow = tid >> 3, rq = (tid & 7) * 4;
    float4_a4 hv[4], rv[kMapTile / 32];
    auto load_tile = [&](int32_t sb) __attribute__((always_inline)) {
#pragma unroll
        for (int pass = 0; pass < 4; ++pass) {
            const float *src = H + int64_t(sb + pass * 8 + hrow) * n + i0 + hq;
            float4_a4 v = {0.0f, 0.0f, 0.0f, 0.0f};
            if (full_i) {
                v = *reinterpret_cast<const float4_a4 *>(src);
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) v[j] = (i0 + hq + j < n) ? src[j] : 0.0f;
            }
            hv[pass] = v;
        }
#pragma unroll
        for (int pass = 0; pass < kMapTile / 32; ++pass) {
            const int32_t t = t0 + pass * 32 + rrow;
            float4_a4 v = {0.0f, 0.0f, 0.0f, 0.0f};
            if (t < T) v = *reinterpret_cast<const float4_a4 *>(R + int64_t(t) * Sp + sb + rq);
            rv[pass] = v;
        }
    };
    if (Se > 0) load_tile(0);
    for (int32_t sb = 0; sb < Se; sb += kGemmTileK) {
#pragma unroll
        for (int pass = 0; pass < 4; ++pass)
#pragma unroll
            for (int j = 0; j < 4; ++j) Hs[pass * 8 + hrow][hq + j] = hv[pass][j];
#pragma unroll
        for (int pass = 0; pass < kMapTile / 32; ++pass)
#pragma unroll
            for (int j = 0; j < 4; ++j) Rs[rq + j][pass * 32 + rrow] = rv[pass][j];
        __syncthreads();
        if (sb + kGemmTileK < Se) load_tile(sb + kGemmTileK);
        const int kend = min(kGemmTileK, Se - sb);  // even: the k pairs of one MFMA
#pragma unroll
        for (int kk = 0; kk < kGemmTileK; kk += 2) {
            if (kk >= kend) break;
            const int kr = kk + (lane >> 5);
            const float a = Hs[kr][wave * 32 + (lane & 31)];
#pragma unroll
            for (int q = 0; q < kMapTile / 32; ++q) {
                const float b = Rs[kr][q * 32 + (lane & 31)];
                acc[q] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc[q], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < kMapTile / 32; ++q)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int64_t i = i0 + wave * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
            const int32_t t = t0 + q * 32 + (lane & 31);
            if (i < n && t < T) eT[i * T + t] = accumulate ? eT[i * T + t] + acc[q][r] : acc[q][r];
        }
}

// ---------------------------------------------------------------------------
// Derivatives over all steps of one beat (derivation.rs:173-229, 307-329,
// 338-398, 408-501). One lane per connection (v, d) runs the whole time loop:
// it recomputes the connection's three all-pass outputs exactly as the
// forward pass did, tracks the FIR/IIR sensitivities, and keeps the nine gain
// derivatives and the coefficient derivative in registers.
//
// Time advances in chunks of kDeriveChunk steps. Per chunk, the workgroup
// stages the mapped residuals and states of its own voxels (shared by the 26
// lanes of a voxel) in LDS, and every lane loads its neighbour's three state
// rows for the chunk into registers, so each cache line of xT / eT is
// fetched once per chunk rather than once per step.
// ---------------------------------------------------------------------------
constexpr int kDeriveThreads = 256;
constexpr int kDeriveChunk = 16;
constexpr int kDeriveMaxVoxels = kDeriveThreads / 26 + 2;

// Sixteen consecutive floats of three rows `stride` apart, in 16-byte loads
// that need only 4-byte alignment.
typedef float float4_a4 __attribute__((ext_vector_type(4), aligned(4)));
typedef float float2_a8 __attribute__((ext_vector_type(2), aligned(8)));
static_assert(kDeriveChunk % 2 == 0, "the steady loop reads LDS rows in pairs of steps");
__device__ __forceinline__ void load_row_chunk(const float *src, int64_t stride,
                                               float out[3][kDeriveChunk]) {
#pragma unroll
    for (int o = 0; o < 3; ++o)
#pragma unroll
        for (int q = 0; q < kDeriveChunk / 4; ++q) {
            const float4_a4 v = *reinterpret_cast<const float4_a4 *>(src + o * stride + 4 * q);
            out[o][4 * q] = v.x;
            out[o][4 * q + 1] = v.y;
            out[o][4 * q + 2] = v.z;
            out[o][4 * q + 3] = v.w;
        }
}

// One connection's running state in the derivation. The coefficient
// derivative is kept as the 3x3 moment M[in][o] = sum over steps of
// s[o] e[in] (s = dy/da: FIR - IIR, or the Simple form's input difference),
// contracted with the connection's gains once after the last step: the
// reference's per-step sum over (in, o) of s[o] G[in][o] e[in] reassociated,
// 9 multiply-adds per step instead of 36 operations.
struct ConnState {
    float g[3][3], M[3][3];
    float f[3], h[3], y[3], prev_in[3];
};

// Per-state factor of the gain derivative at one step (derivation.rs:315-326
// with the maximum regularisation of derivation.rs:509-535): shared by the
// 26 connections of a voxel, so it is computed once while staging.
__device__ __forceinline__ void gain_factors(const float xo[3], const float e[3],
                                             const DeriveParams &p, float cg[3]) {
    const float sum = fabsf(xo[0]) + fabsf(xo[1]) + fabsf(xo[2]);
    const bool on = sum > p.threshold;
    const float fac = sum - p.threshold;
#pragma unroll
    for (int in = 0; in < 3; ++in) {
        const float mr = on ? fac * signum(xo[in]) : 0.0f;
        cg[in] = fmaf(e[in], p.mse_scaling, mr * p.reg_scaling);
    }
}

// One step of one connection. A missing connection (has = false) reads
// in0 = +0 and starts from y = prev_in = fir = iir = +0, so updating its
// filter and FIR unconditionally keeps them +0 (for a finite coefficient),
// as the reference's guarded updates do. With kSteady the caller has checked
// t >= D for every lane of the wave. The filter outputs and FIR / IIR follow
// the reference's operations exactly; the gain derivative fuses its
// multiply-add (derivation.rs:325-326 multiplies, then adds).
// kGains / kCoefs: the gain / coefficient derivatives are computed (not
// frozen); kSimple: the Simple coefficient derivative instead of Textbook.
// As template parameters they leave the unrolled step loop without branches.
template <bool kSteady, bool kGains, bool kCoefs, bool kSimple>
__device__ __forceinline__ void derive_step(ConnState &c, const float in0[3], const float e[3],
                                            const float cg[3], bool has, float a, int32_t D,
                                            int32_t t) {
    float yl[3];
#pragma unroll
    for (int o = 0; o < 3; ++o) {
        yl[o] = c.y[o];
        c.y[o] = fmaf(a, in0[o] - c.y[o], c.prev_in[o]);
    }
    if (kGains) {
#pragma unroll
        for (int in = 0; in < 3; ++in)
#pragma unroll
            for (int o = 0; o < 3; ++o) c.g[in][o] = fmaf(c.y[o], cg[in], c.g[in][o]);
    }
    if (kCoefs) {
        const bool on = kSteady || t >= D;
        float s[3];
        if (!kSimple) {
            // derivation.rs:417-499: FIR / IIR advance from step D on; the
            // combine runs every step with the sensitivities they hold
            if (on) {
#pragma unroll
                for (int o = 0; o < 3; ++o) {
                    c.f[o] = fmaf(-a, c.f[o], in0[o]);
                    c.h[o] = fmaf(-a, c.h[o], yl[o]);
                }
            }
#pragma unroll
            for (int o = 0; o < 3; ++o) s[o] = c.f[o] - c.h[o];
        } else {
            // derivation.rs:338-398: only existing connections past their delay
            const bool use = has && on;
#pragma unroll
            for (int o = 0; o < 3; ++o) s[o] = use ? in0[o] - yl[o] : 0.0f;
        }
#pragma unroll
        for (int in = 0; in < 3; ++in)
#pragma unroll
            for (int o = 0; o < 3; ++o) c.M[in][o] = fmaf(s[o], e[in], c.M[in][o]);
    }
#pragma unroll
    for (int o = 0; o < 3; ++o) c.prev_in[o] = in0[o];
}

template <bool kGains, bool kCoefs, bool kSimple>
__global__ __launch_bounds__(kDeriveThreads, 4) void derive_kernel(
    const int32_t *__restrict__ nbr, const float *__restrict__ coefs,
    const int32_t *__restrict__ delays, const float *__restrict__ initial_delays,
    const float *__restrict__ gT, const float *__restrict__ xT, const float *__restrict__ eT,
    float *__restrict__ fir, float *__restrict__ iir, float *__restrict__ dGT,
    float *__restrict__ dA, int32_t N, int32_t T, DeriveParams p) {
    // rows of 16 steps, aligned: the steady loop reads two steps per 8-byte load (four
    // per 16-byte load measured no faster and spilled: 128 VGPRs against 126)
    __shared__ __align__(16) float es[kDeriveMaxVoxels][3][kDeriveChunk];  // mapped residuals
    __shared__ __align__(16) float cs[kDeriveMaxVoxels][3][kDeriveChunk];  // gain-derivative factors
    const int64_t n = 3LL * N;
    const int64_t total = int64_t(N) * 26;
    const int64_t item0 = int64_t(xcd_block(blockIdx.x, gridDim.x)) * kDeriveThreads;
    const int64_t item = item0 + threadIdx.x;
    const bool active = item < total;
    const int32_t vbeg = int32_t(item0 / 26);
    const int32_t vend = int32_t(min(total - 1, item0 + kDeriveThreads - 1) / 26) + 1;
    const int32_t nvox = vend - vbeg;

    const int64_t it = active ? item : item0;
    const int32_t v = int32_t(it / 26), d = int32_t(it % 26);
    const int32_t vl = v - vbeg;
    const int32_t w = active ? nbr[it] : -1;
    const bool has = w >= 0;
    const float a = coefs[it];
    const int32_t D = delays[it];

    ConnState c;
#pragma unroll
    for (int in = 0; in < 3; ++in)
#pragma unroll
        for (int o = 0; o < 3; ++o) {
            c.g[in][o] = dGT[int64_t(3 * d + o) * n + 3LL * v + in];
            c.M[in][o] = 0.0f;
        }
#pragma unroll
    for (int o = 0; o < 3; ++o) {
        c.f[o] = fir[it * 3 + o];
        c.h[o] = iir[it * 3 + o];
        c.y[o] = 0.0f;
        c.prev_in[o] = 0.0f;
    }
    const float *xw = xT + (has ? 3LL * w : 0) * T;

    for (int32_t t0 = 0; t0 < T; t0 += kDeriveChunk) {
        for (int idx = threadIdx.x; idx < nvox * kDeriveChunk; idx += kDeriveThreads) {
            const int vv = idx / kDeriveChunk, j = idx - vv * kDeriveChunk;
            const int32_t t = t0 + j;
            const int64_t row = 3LL * (vbeg + vv) * T + t;
            float e[3] = {0.0f, 0.0f, 0.0f}, xo[3] = {0.0f, 0.0f, 0.0f}, cg[3];
            if (t < T) {
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    e[q] = eT[row + int64_t(q) * T];
                    xo[q] = xT[row + int64_t(q) * T];
                }
            }
            gain_factors(xo, e, p, cg);
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                es[vv][q][j] = e[q];
                cs[vv][q][j] = cg[q];
            }
        }
        float xin[3][kDeriveChunk];
        if (has && t0 >= D && t0 + kDeriveChunk <= T) {
            // the chunk's 16 inputs of each neighbour row are contiguous: four
            // 16-byte loads per row instead of sixteen 4-byte gathers
            load_row_chunk(xw + (t0 - D), T, xin);
        } else {
#pragma unroll
            for (int o = 0; o < 3; ++o)
#pragma unroll
                for (int j = 0; j < kDeriveChunk; ++j) {
                    const int32_t s = t0 + j - D;
                    xin[o][j] = (has && s >= 0 && t0 + j < T) ? xw[int64_t(o) * T + s] : 0.0f;
                }
        }
        __syncthreads();
        // steady chunk: whole, and every lane of the wave past its delay
        const bool steady = __all(!active || t0 >= D) && t0 + kDeriveChunk <= T;
        if (active && steady) {
#pragma unroll
            for (int m = 0; m < kDeriveChunk; m += 2) {
                float2_a8 ev[3], cv[3];
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    ev[q] = *reinterpret_cast<const float2_a8 *>(&es[vl][q][m]);
                    cv[q] = *reinterpret_cast<const float2_a8 *>(&cs[vl][q][m]);
                }
#pragma unroll
                for (int jj = 0; jj < 2; ++jj) {
                    const int j = m + jj;
                    const float in0[3] = {xin[0][j], xin[1][j], xin[2][j]};
                    const float e[3] = {ev[0][jj], ev[1][jj], ev[2][jj]};
                    const float cg[3] = {cv[0][jj], cv[1][jj], cv[2][jj]};
                    derive_step<true, kGains, kCoefs, kSimple>(c, in0, e, cg, has, a, D, t0 + j);
                }
            }
        } else if (active) {
#pragma unroll
            for (int j = 0; j < kDeriveChunk; ++j) {
                if (t0 + j < T) {
                    const float in0[3] = {xin[0][j], xin[1][j], xin[2][j]};
                    const float e[3] = {es[vl][0][j], es[vl][1][j], es[vl][2][j]};
                    const float cg[3] = {cs[vl][0][j], cs[vl][1][j], cs[vl][2][j]};
                    derive_step<false, kGains, kCoefs, kSimple>(c, in0, e, cg, has, a, D,
                                                                t0 + j);
                }
            }
        }
        __syncthreads();
    }
    if (!active) return;

#pragma unroll
    for (int in = 0; in < 3; ++in)
#pragma unroll
        for (int o = 0; o < 3; ++o) dGT[int64_t(3 * d + o) * n + 3LL * v + in] = c.g[in][o];
#pragma unroll
    for (int o = 0; o < 3; ++o) {
        fir[item * 3 + o] = c.f[o];
        iir[item * 3 + o] = c.h[o];
    }
    if (!kCoefs) return;
    // the beat's coefficient derivative: the moment contracted with the gains,
    // plus the difference regulariser's term, which the reference adds for
    // each (in, o) of every step it combines (every step in the Textbook
    // form, the steps t >= D of an existing connection in the Simple form)
    float dot = 0.0f;
#pragma unroll
    for (int in = 0; in < 3; ++in)
#pragma unroll
        for (int o = 0; o < 3; ++o)
            dot = fmaf(gT[int64_t(3 * d + o) * n + 3LL * v + in], c.M[in][o], dot);
    const float delay = float(D) + (1.0f - a) / (a + 1.0f);
    const float reg = p.diff_strength * powi5(initial_delays[item] - delay);
    const int32_t combined = kSimple ? (has ? max(T - D, 0) : 0) : T;
    dA[item] = dA[item] + fmaf(dot, p.mse_scaling, float(9 * combined) * reg);
}

// ---------------------------------------------------------------------------
// Derivatives over voxel tiles (the same derivation.rs references as
// derive_kernel). Every connection runs derive_kernel's arithmetic step for
// step; only the source of its delayed inputs differs. Per chunk of
// kDeriveChunk steps the workgroup copies that chunk of each of its halo rows
// (three state rows per halo voxel) into an LDS ring holding the last
// kTileRing steps, and each lane reads its neighbour's inputs x[t - D] from
// there. Step t of a row sits at position t mod kTileRing; the chunk at
// positions [0, kDeriveChunk) is also written behind the ring, so the inputs
// of a lane's chunk, which start anywhere in the ring, are contiguous. The
// loads of chunk c + 1 are in flight while the steps of chunk c run.
// ---------------------------------------------------------------------------
static_assert(kDeriveChunk == kTileChunk, "the ring is filled one derivation chunk at a time");
constexpr int kTileMaxDelay = kTileRing - kDeriveChunk;  // the ring holds chunks c-3 .. c
constexpr size_t kTileLdsBytes =
    sizeof(float) * (size_t(kTileRingFloats) + kTileZeroFloats + 2 * size_t(kTileStageFloats));

template <bool kGains, bool kCoefs, bool kSimple>
__global__ __launch_bounds__(kTileThreads) void derive_tiled_kernel(
    DeriveTiles tl, const int32_t *__restrict__ nbr, const float *__restrict__ coefs,
    const int32_t *__restrict__ delays, const float *__restrict__ initial_delays,
    const float *__restrict__ gT, const float *__restrict__ xT, const float *__restrict__ eT,
    float *__restrict__ fir, float *__restrict__ iir, float *__restrict__ dGT,
    float *__restrict__ dA, int32_t N, int32_t T, DeriveParams p) {
    extern __shared__ __align__(16) float lds[];
    float *ring = lds;                       // [halo][3][kTileRowStride]
    float *zeros = ring + kTileRingFloats;   // [3][kTileRowStride], +0: missing connections
    float *es = zeros + kTileZeroFloats;     // [voxel][3][chunk] mapped residuals
    float *cs = es + kTileStageFloats;       // [voxel][3][chunk] gain-derivative factors
    const int64_t n = 3LL * N;
    const int32_t tile = xcd_block(blockIdx.x, gridDim.x);
    const int32_t pbeg = tl.vox_ptr[tile], nvox = tl.vox_ptr[tile + 1] - pbeg;
    const int32_t hbeg = tl.halo_ptr[tile], nhalo = tl.halo_ptr[tile + 1] - hbeg;
    const int tid = threadIdx.x;

    const TileLane L = tile_lane(tl, nbr, coefs, delays, pbeg, nvox, kTileMaxDelay);
    for (int i = tid; i < kTileZeroFloats; i += kTileThreads) zeros[i] = 0.0f;
    const bool active = L.active, has = L.has, far = L.far;
    const int d = L.d;
    const int32_t v = L.v, D = L.D;
    const int64_t item = L.item;
    const float a = L.a;
    const float *xw = xT + (has ? 3LL * L.w : 0) * T;
    const float *rrow = L.slot >= 0 ? ring + L.slot * 3 * kTileRowStride : zeros;

    ConnState c;
#pragma unroll
    for (int in = 0; in < 3; ++in)
#pragma unroll
        for (int o = 0; o < 3; ++o) {
            c.g[in][o] = active ? dGT[int64_t(3 * d + o) * n + 3LL * v + in] : 0.0f;
            c.M[in][o] = 0.0f;
        }
#pragma unroll
    for (int o = 0; o < 3; ++o) {
        c.f[o] = active ? fir[item * 3 + o] : 0.0f;
        c.h[o] = active ? iir[item * 3 + o] : 0.0f;
        c.y[o] = 0.0f;
        c.prev_in[o] = 0.0f;
    }

    // this thread's share of the chunk copies: quads of halo rows, and one
    // (voxel, step) item of the tile's own residuals and states
    TileRingCopy rc;
    rc.init(tl, hbeg, nhalo, T);
    const int sv = tid / kDeriveChunk, sj = tid - kDeriveChunk * sv;
    const bool sown = sv < nvox;
    const int64_t srow = sown ? 3LL * tl.vox[pbeg + sv] * T + sj : 0;
    float se[3], sx[3];
    const auto fetch = [&](int32_t t0) {
        rc.fetch(xT, t0, T);
        const bool live = sown && t0 + sj < T;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            se[q] = live ? eT[srow + int64_t(q) * T + t0] : 0.0f;
            sx[q] = live ? xT[srow + int64_t(q) * T + t0] : 0.0f;
        }
    };
    const auto stage = [&](int32_t t0) {
        rc.stage(ring, t0);
        if (sown) {
            float cg[3];
            gain_factors(sx, se, p, cg);
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                es[(sv * 3 + q) * kDeriveChunk + sj] = se[q];
                cs[(sv * 3 + q) * kDeriveChunk + sj] = cg[q];
            }
        }
    };

    fetch(0);
    for (int32_t t0 = 0; t0 < T; t0 += kDeriveChunk) {
        stage(t0);
        // Every load issued so far has landed (stage used them). Stated
        // explicitly, since the compiler cannot see it through the per-thread
        // conditions of the copies: without this it waits for the next
        // chunk's loads right after issuing them, before this chunk's steps.
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (t0 + kDeriveChunk < T) fetch(t0 + kDeriveChunk);
        // the lane's 16 inputs of this chunk: rb[o * kTileRowStride + j] = x[t0 + j - D]
        const float *rb = rrow + ((t0 - D) & (kTileRing - 1));
        const float *el = es + L.lv * 3 * kDeriveChunk;
        const float *cl = cs + L.lv * 3 * kDeriveChunk;
        const auto at = [](const float *base, int q, int j) { return base[q * kDeriveChunk + j]; };
        const bool steady = __all(!active || t0 >= D) && t0 + kDeriveChunk <= T;
        const bool any_far = __any(far);
        if (active && steady && !any_far) {
#pragma unroll
            for (int m = 0; m < kDeriveChunk; m += 2) {
                float2_a8 ev[3], cv[3];
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    ev[q] = *reinterpret_cast<const float2_a8 *>(el + q * kDeriveChunk + m);
                    cv[q] = *reinterpret_cast<const float2_a8 *>(cl + q * kDeriveChunk + m);
                }
#pragma unroll
                for (int jj = 0; jj < 2; ++jj) {
                    const int j = m + jj;
                    const float in0[3] = {rb[j], rb[kTileRowStride + j], rb[2 * kTileRowStride + j]};
                    const float e[3] = {ev[0][jj], ev[1][jj], ev[2][jj]};
                    const float cg[3] = {cv[0][jj], cv[1][jj], cv[2][jj]};
                    derive_step<true, kGains, kCoefs, kSimple>(c, in0, e, cg, has, a, D, t0 + j);
                }
            }
        } else if (active) {
#pragma unroll
            for (int j = 0; j < kDeriveChunk; ++j) {
                const int32_t t = t0 + j;
                if (t < T) {
                    const int32_t s = t - D;
                    float in0[3];
#pragma unroll
                    for (int o = 0; o < 3; ++o)
                        in0[o] = s < 0 ? 0.0f
                                       : (far ? xw[int64_t(o) * T + s] : rb[o * kTileRowStride + j]);
                    const float e[3] = {at(el, 0, j), at(el, 1, j), at(el, 2, j)};
                    const float cg[3] = {at(cl, 0, j), at(cl, 1, j), at(cl, 2, j)};
                    derive_step<false, kGains, kCoefs, kSimple>(c, in0, e, cg, has, a, D, t);
                }
            }
        }
        __syncthreads();
    }
    if (!active) return;

#pragma unroll
    for (int in = 0; in < 3; ++in)
#pragma unroll
        for (int o = 0; o < 3; ++o) dGT[int64_t(3 * d + o) * n + 3LL * v + in] = c.g[in][o];
#pragma unroll
    for (int o = 0; o < 3; ++o) {
        fir[item * 3 + o] = c.f[o];
        iir[item * 3 + o] = c.h[o];
    }
    if (!kCoefs) return;
    float dot = 0.0f;
#pragma unroll
    for (int in = 0; in < 3; ++in)
#pragma unroll
        for (int o = 0; o < 3; ++o)
            dot = fmaf(gT[int64_t(3 * d + o) * n + 3LL * v + in], c.M[in][o], dot);
    const float delay = float(D) + (1.0f - a) / (a + 1.0f);
    const float reg = p.diff_strength * powi5(initial_delays[item] - delay);
    const int32_t combined = kSimple ? (has ? max(T - D, 0) : 0) : T;
    dA[item] = dA[item] + fmaf(dot, p.mse_scaling, float(9 * combined) * reg);
}

// update.rs:106-114, 158-168
__device__ __forceinline__ bool guarded(const UpdateGuard &g) {
    const auto word = [](const uint32_t *w) {
        return w && __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
    };
    return word(g.word0) || word(g.word1);
}

__global__ void sgd_kernel(float *__restrict__ v, const float *__restrict__ d, int64_t len,
                           float scale, UpdateGuard guard) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= len || guarded(guard)) return;
    v[k] = v[k] - scale * d[k];
}

// update.rs:119-146, 173-200
__global__ void adam_kernel(float *__restrict__ v, const float *__restrict__ d,
                            float *__restrict__ m1, float *__restrict__ m2, int64_t len,
                            const UpdateScalars *__restrict__ sc, float scale,
                            UpdateGuard guard) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= len || guarded(guard)) return;
    const float c1 = sc->c1, c2 = sc->c2;
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    const float omb1 = 1.0f - beta1, omb2 = 1.0f - beta2;
    const float g = d[k];
    const float a = m1[k] * beta1 + omb1 * g;
    const float b = m2[k] * beta2 + omb2 * g * g;
    m1[k] = a;
    m2[k] = b;
    const float ac = a / c1;
    const float bc = b / c2;
    const float factor = ac / (sqrtf(bc) + eps);
    v[k] = v[k] - scale * factor;
}

// update.rs:206-228
__global__ void roll_delays_kernel(float *__restrict__ coefs, int32_t *__restrict__ delays,
                                   int64_t len, UpdateGuard guard) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= len || guarded(guard)) return;
    const float margin = 1e-4f;
    float a = coefs[k];
    int32_t D = delays[k];
    if (a > 1.0f - margin) {
        if (D > 1) {
            a = 2.0f * margin;
            D -= 1;
        } else {
            a = 1.0f - margin;
        }
    } else if (a < margin) {
        if (D < 1000) {
            a = fmaf(2.0f, -margin, 1.0f);
            D += 1;
        } else {
            a = margin;
        }
    }
    coefs[k] = a;
    delays[k] = D;
}

// derivation.rs:554-594
__global__ void average_delays_kernel(const float *__restrict__ gT, const float *__restrict__ coefs,
                                      const int32_t *__restrict__ delays, float *__restrict__ avg,
                                      uint8_t *__restrict__ valid, int32_t N) {
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const int64_t n = 3LL * N;
    float delay_sum = 0.0f, gain_sum = 0.0f;
    for (int d = 0; d < 26; ++d) {
        const int c = d > 12 ? d + 1 : d;
        const int xyz = abs(c % 3 - 1) + abs((c / 3) % 3 - 1) + abs((c / 9) % 3 - 1);
        const float delay = float(delays[int64_t(v) * 26 + d]) +
                            (1.0f - coefs[int64_t(v) * 26 + d]) / (coefs[int64_t(v) * 26 + d] + 1.0f);
        const float corrected = delay / sqrtf(float(xyz));
        for (int in = 0; in < 3; ++in)
            for (int o = 0; o < 3; ++o) {
                const float gain = gT[int64_t(3 * d + o) * n + 3LL * v + in];
                delay_sum += fabsf(gain) * corrected;
                gain_sum += fabsf(gain);
            }
    }
    valid[v] = gain_sum != 0.0f;
    avg[v] = gain_sum != 0.0f ? delay_sum / gain_sum : 0.0f;
}

// derivation.rs:260-302
__global__ void smoothness_kernel(const int32_t *__restrict__ nbr, const float *__restrict__ avg,
                                  const uint8_t *__restrict__ valid, float *__restrict__ dA,
                                  float strength, int32_t N) {
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N || !valid[v]) return;
    const float own = avg[v];
    float hood = own, divisor = 1.0f;
    for (int d = 0; d < 26; ++d) {
        const int32_t w = nbr[int64_t(v) * 26 + d];
        if (w < 0) continue;
        if (valid[w]) {
            hood += avg[w];
            divisor += 1.0f;
        }
    }
    hood /= divisor;
    const float diff = hood - own;
    for (int d = 0; d < 26; ++d) dA[int64_t(v) * 26 + d] += strength * diff;
}

// metrics.rs:204-219: the means of the beat's per-step losses, in ndarray's
// order for a 1-D f32 mean (eight interleaved partial sums, combined
// pairwise, then the tail, then / T). The three arrays go through LDS in
// chunks; eight lanes per array keep the eight partial sums, and thread q < 3
// combines array q's.
__global__ __launch_bounds__(256) void batch_metrics_kernel(const float *loss, const float *loss_mse,
                                                            const float *loss_mr, float *loss_b,
                                                            float *loss_mse_b, float *loss_mr_b,
                                                            const UpdateScalars *sc, int32_t T) {
    __shared__ float sx[3][kScanChunk];
    __shared__ float part[3][8];
    const int tid = threadIdx.x, arr = tid / 8, j = tid % 8;
    const auto src = [&](int q) { return q == 0 ? loss_mse : q == 1 ? loss_mr : loss; };
    float pj = 0.0f;
    // the whole-chunk part of the array: chunks are multiples of 8 elements,
    // so the interleaving continues across chunks; the tail (len % 8) is the
    // last chunk's remainder
    const int32_t body = T / 8 * 8;
    for (int32_t c0 = 0; c0 < body; c0 += kScanChunk) {
        const int32_t len = min(kScanChunk, body - c0);
        for (int32_t i = tid; i < 3 * len; i += blockDim.x) {
            const int32_t q = i / len, e = i - q * len;
            sx[q][e] = src(q)[c0 + e];
        }
        __syncthreads();
        if (arr < 3)
            for (int32_t i = 0; i < len; i += 8) pj = pj + sx[arr][i + j];
        __syncthreads();
    }
    if (arr < 3) part[arr][j] = pj;
    __syncthreads();
    if (tid < 3) {
        const float *p = part[tid];
        float acc = 0.0f;
        acc = acc + (p[0] + p[4]);
        acc = acc + (p[1] + p[5]);
        acc = acc + (p[2] + p[6]);
        acc = acc + (p[3] + p[7]);
        const float *xs = src(tid);
        for (int32_t i = body; i < T; ++i) acc = acc + xs[i];
        float *dst = tid == 0 ? loss_mse_b : tid == 1 ? loss_mr_b : loss_b;
        dst[sc->batch_index] = acc / float(T);
    }
}

// (N, 78) per-connection filter rows -> (3N, 78) per-state rows.
__global__ void expand_rows_kernel(const float *__restrict__ in, float *__restrict__ out,
                                   int64_t N) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= N * 234) return;
    const int64_t row = k / 78, g = k % 78;
    out[k] = in[(row / 3) * 78 + g];
}

__global__ void copy_kernel(float *__restrict__ dst, const float *__restrict__ src, int64_t len) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < len) dst[k] = src[k];
}

__global__ void fill_kernel(float *p, float value, int64_t len) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < len) p[k] = value;
}

// Zeroes the counters of `count` barriers and keeps their abort words, so an
// abort stays visible to the host until it reads it after the epoch.
__global__ void barrier_reset_kernel(GridBarrier *gb, int32_t count) {
    constexpr int kWords = int(sizeof(GridBarrier) / sizeof(BarrierWord));
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count * (kWords - 1)) return;
    const int32_t j = k / (kWords - 1), w = k % (kWords - 1);  // abort is the last word
    reinterpret_cast<BarrierWord *>(gb + j)[w].count = 0u;
}

// splitmix64 finaliser (Steele, Lea and Flood 2014).
__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Uniform in [0, 1) from the top 53 bits of splitmix64(counter ^ key).
__device__ __forceinline__ double counter_uniform(uint64_t counter, uint64_t key) {
    return double(splitmix64(counter ^ key) >> 11) * (1.0 / 9007199254740992.0);
}

// Measurement noise (the noise step of data/simulation.rs:152-169, with this
// library's generator): element (b, t, s) of m (B, T, S) gets
// sd[s] * z_k, k = (s B + b) T + t -- the reference's sensor, beat, step
// sampling order -- where z_k = sqrt(-2 ln(1 - U(2k))) cos(2 pi U(2k+1)) is the
// Box-Muller deviate of two counter uniforms. The sum is taken in FP32 after
// rounding sd[s] * z_k (FP64) to FP32.
__global__ void measurement_noise_kernel(float *__restrict__ m, const float *__restrict__ sd,
                                         int32_t B, int32_t T, int32_t S, uint64_t key) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= int64_t(B) * T * S) return;
    const int32_t s = int32_t(e % S);
    const int32_t t = int32_t((e / S) % T);
    const int32_t b = int32_t(e / (int64_t(S) * T));
    const uint64_t k = (uint64_t(s) * uint64_t(B) + uint64_t(b)) * uint64_t(T) + uint64_t(t);
    const double u1 = 1.0 - counter_uniform(2 * k, key);
    const double u2 = counter_uniform(2 * k + 1, key);
    const double z = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
    m[e] = m[e] + float(double(sd[s]) * z);
}

// Biot-Savart measurement matrix, measurement.rs:66-123, in the reference's
// f32 operation order: d = (sensor + offset - voxel) / 1000, |d|^3 = r (r r)
// with r = sqrt((0 + d0^2) + d1^2) + d2^2), entries common * o_a.mul_add(d_b,
// -o_b d_c) / |d|^3. A thread per (beat, sensor, voxel) writes its 3 entries.
__global__ void measurement_matrix_kernel(const float *__restrict__ vp, const float *__restrict__ sp,
                                          const float *__restrict__ so, const float *__restrict__ off,
                                          float common, int32_t N, int32_t S, int32_t B,
                                          float *__restrict__ out) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= int64_t(B) * S * N) return;
    const int32_t v = int32_t(e % N);
    const int64_t bs = e / N;
    const int32_t s = int32_t(bs % S), b = int32_t(bs / S);
    float d[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) d[k] = ((sp[3 * s + k] + off[3 * b + k]) - vp[3LL * v + k]) / 1000.0f;
    float sq = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) sq = sq + d[k] * d[k];
    const float r = sqrtf(sq);
    const float r3 = r * (r * r);
    const float *o = so + 3 * s;
    float *row = out + bs * (3LL * N) + 3LL * v;
    row[0] = common * fmaf(o[2], d[1], -o[1] * d[2]) / r3;
    row[1] = common * fmaf(o[0], d[2], -o[2] * d[0]) / r3;
    row[2] = common * fmaf(o[1], d[0], -o[0] * d[1]) / r3;
}

// Connections of a voxel grid (allpass.rs:212-289 and allpass/delay.rs:35-111),
// a thread per (cell, offset): for a connectable cell and a connectable
// neighbour inside the grid, the nine output state indices of the offset and
// its delay in samples, |(in - out) / 1000| / v_in * rate in the reference's
// f32 operations. A delay below one sample raises *too_small.
__global__ void model_connections_kernel(int32_t nx, int32_t ny, int32_t nz,
                                         const int32_t *__restrict__ numbers,
                                         const float *__restrict__ pos,
                                         const float *__restrict__ vel, float rate,
                                         int32_t *__restrict__ osi, float *__restrict__ delays,
                                         int32_t *__restrict__ too_small) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t cells = int64_t(nx) * ny * nz;
    if (e >= cells * 27) return;
    const int64_t ci = e / 27;
    const int32_t o = int32_t(e % 27);
    const int32_t ox = o / 9 - 1, oy = (o / 3) % 3 - 1, oz = o % 3 - 1;
    if (o == 13) return;  // the cell itself
    const int32_t in_num = numbers[ci];
    if (in_num < 0) return;
    const int32_t z = int32_t(ci % nz), y = int32_t((ci / nz) % ny), x = int32_t(ci / (int64_t(ny) * nz));
    const int32_t X = x + ox, Y = y + oy, Z = z + oz;
    if (X < 0 || X >= nx || Y < 0 || Y >= ny || Z < 0 || Z >= nz) return;
    const int64_t co = (int64_t(X) * ny + Y) * nz + Z;
    const int32_t out_num = numbers[co];
    if (out_num < 0) return;
    // offset_to_delay_index (allpass.rs:582-592); gain index = 3 * delay index + out
    const int32_t di = o > 13 ? o - 1 : o;
    for (int32_t in = 0; in < 3; ++in)
        for (int32_t od = 0; od < 3; ++od) osi[int64_t(in_num + in) * 78 + 3 * di + od] = out_num + od;
    float sq = 0.0f;
    for (int k = 0; k < 3; ++k) {
        const float d = (pos[3 * ci + k] - pos[3 * co + k]) / 1000.0f;
        sq = sq + d * d;
    }
    const float delay = sqrtf(sq) / vel[ci] * rate;
    if (delay < 1.0f) atomicOr(too_small, 1);  // the reference's test (delay.rs:94)
    delays[int64_t(in_num / 3) * 26 + di] = delay;
}

// Spherical states, their values at the maximum magnitude and the activation
// time (data/shapes.rs:171-191, 255-270, 303-322), a thread per voxel walking
// the states x (T, 3N) once. The maximum is the first index whose magnitude
// exceeds every earlier non-NaN one (argmax_skipnan); a voxel with no non-NaN
// magnitude raises *no_max. The (T, N) outputs are written when not null.
__global__ void spherical_kernel(const float *__restrict__ x, int32_t T, int32_t N, float rate,
                                 float *__restrict__ mag, float *__restrict__ theta,
                                 float *__restrict__ phi, float *__restrict__ mag_max,
                                 float *__restrict__ theta_max, float *__restrict__ phi_max,
                                 float *__restrict__ act_ms, int32_t *__restrict__ no_max) {
    const int32_t v = int32_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const int64_t n = 3LL * N;
    int32_t best = -1;
    float bm = 0.0f, bt = 0.0f, bp = 0.0f;
    for (int32_t t = 0; t < T; ++t) {
        const float *s = x + int64_t(t) * n + 3LL * v;
        const float s0 = s[0], s1 = s[1], s2 = s[2];
        const float m = fabsf(s0) + fabsf(s1) + fabsf(s2);
        const float th = acosf(s2 / m);
        const float ph = atan2f(s1, s0);
        if (mag) mag[int64_t(t) * N + v] = m;
        if (theta) theta[int64_t(t) * N + v] = th;
        if (phi) phi[int64_t(t) * N + v] = ph;
        if (!isnan(m) && (best < 0 || m > bm)) {
            best = t;
            bm = m;
            bt = th;
            bp = ph;
        }
    }
    if (best < 0) {
        atomicOr(no_max, 1);
        return;
    }
    mag_max[v] = bm;
    theta_max[v] = bt;
    phi_max[v] = bp;
    act_ms[v] = float(best) / rate * 1000.0f;
}

// Per voxel, the maximum over steps of |x0| + |x1| + |x2| with NaN values
// skipped, NaN when every step is NaN: predict_voxeltype's abs array and its
// max_skipnan (algorithm/metrics.rs:409-436). Thread (lane, slice) walks
// steps slice, slice + kMaxL1Slices, ... of voxel blockIdx.x * 64 + lane; the
// slices combine in LDS. The maximum of exact values does not depend on that
// order.
constexpr int kMaxL1Voxels = 64;
constexpr int kMaxL1Slices = 4;
__global__ __launch_bounds__(kMaxL1Voxels *kMaxL1Slices) void voxel_max_l1_kernel(
    const float *__restrict__ x, int32_t T, int32_t N, float *__restrict__ out) {
    __shared__ float part[kMaxL1Slices][kMaxL1Voxels];
    const int lane = threadIdx.x % kMaxL1Voxels, slice = threadIdx.x / kMaxL1Voxels;
    const int32_t v = int32_t(blockIdx.x) * kMaxL1Voxels + lane;
    float m = __builtin_nanf("");
    if (v < N) {
        const int64_t n = 3LL * N;
        // one 12-byte load per step: 6.0 TB/s at 120k voxels x 500 steps;
        // four steps in flight per lane measured slower (126-137 us vs 120)
        for (int32_t t = slice; t < T; t += kMaxL1Slices) {
            const float *s = x + int64_t(t) * n + 3LL * v;
            const float a = fabsf(s[0]) + fabsf(s[1]) + fabsf(s[2]);
            if (a == a && !(m >= a)) m = a;  // NaN skipped; a NaN maximum takes a
        }
    }
    part[slice][lane] = m;
    __syncthreads();
    if (slice != 0 || v >= N) return;
#pragma unroll
    for (int k = 1; k < kMaxL1Slices; ++k) {
        const float a = part[k][lane];
        if (a == a && !(m >= a)) m = a;
    }
    out[v] = m;
}

// Two histograms over the grid cells (algorithm/metrics.rs:225-441): bin k of
// hist[0] (ground truth not Pathological) or hist[1] (Pathological) counts the
// cells first predicted Pathological at threshold index k, i.e. the smallest
// k with max <= thresholds[k]; k = kSegmentationThresholds for cells without a
// voxel or with a NaN maximum, which are never predicted Pathological.
__global__ void segmentation_histogram_kernel(const int32_t *__restrict__ numbers,
                                              const uint32_t *__restrict__ types, int64_t cells,
                                              const float *__restrict__ max_l1,
                                              SegmentationThresholds th,
                                              uint32_t *__restrict__ hist) {
    constexpr int kBins = kSegmentationThresholds + 1;
    __shared__ uint32_t h[2 * kBins];
    for (int i = threadIdx.x; i < 2 * kBins; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; c < cells;
         c += int64_t(gridDim.x) * blockDim.x) {
        const int32_t num = numbers[c];
        int k = kSegmentationThresholds;
        if (num >= 0) {
            const float m = max_l1[num / 3];
            if (m == m) {  // lower bound over the increasing thresholds
                int lo = 0, hi = kSegmentationThresholds;
                while (lo < hi) {
                    const int mid = (lo + hi) / 2;
                    if (m <= th.value[mid])
                        hi = mid;
                    else
                        lo = mid + 1;
                }
                k = lo;
            }
        }
        atomicAdd(&h[(types[c] == kVoxelTypePathological ? kBins : 0) + k], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 2 * kBins; i += blockDim.x)
        if (h[i]) atomicAdd(&hist[i], h[i]);
}

__global__ void fill_words_kernel(uint32_t *p, uint32_t v, int64_t len) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < len) p[k] = v;
}

// Zeroes one job's barrier counters (not its abort word) and `nflags` slab flags.
__global__ void resident_reset_kernel(GridBarrier *gb, uint32_t *flags, int32_t nflags) {
    constexpr int kWords = int(sizeof(GridBarrier) / sizeof(BarrierWord));
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < kWords - 1) reinterpret_cast<BarrierWord *>(gb)[k].count = 0u;
    if (flags && k < nflags) flags[k] = 0u;
}

inline uint64_t splitmix64_host(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline unsigned blocks_for(int64_t len, int threads) {
    return unsigned((len + threads - 1) / threads);
}

}  // namespace

hipError_t launch_propagate_step(hipStream_t s, const int32_t *nbr, const float *coefs,
                                 const int32_t *delays, const float *gT, float *y, float *y_last,
                                 float *x, const float *b, float u_t, int32_t N, int32_t t,
                                 int32_t T) {
    if (N <= 0) return hipSuccess;
    propagate_step_kernel<<<blocks_for(N, kPropVoxels), kPropThreads, 0, s>>>(
        nbr, coefs, delays, gT, y, y_last, x, b, u_t, N, t, T);
    return hipGetLastError();
}

size_t resident_lds_bytes() {
    return sizeof(float) * kResMaxVoxels * 78 + 16 + kResMaxVoxels * 26;
}

// Workgroups of a resident kernel that fit on one CU at once: the smaller
// count of its two instantiations (the launchers pick one per launch).
template <typename K>
hipError_t resident_fit(K kernel0, K kernel1, int32_t *blocks_per_cu) {
    int fit = 1 << 30;
    for (K k : {kernel0, kernel1}) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k),
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           int(resident_lds_bytes()));
        if (e != hipSuccess) return e;
        int nb = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void *>(k),
                                                         kResThreads, resident_lds_bytes());
        if (e != hipSuccess) return e;
        fit = std::min(fit, nb);
    }
    *blocks_per_cu = fit;
    return hipSuccess;
}

hipError_t resident_capacity(int32_t *blocks_per_cu) {
    return resident_fit(&propagate_resident_kernel<true>, &propagate_resident_kernel<false>,
                        blocks_per_cu);
}

hipError_t launch_propagate_resident(hipStream_t s, int32_t blocks, const ResidentJob &job,
                                     int32_t t_begin, int32_t t_end, uint64_t wait_ticks) {
    // the abort word's neighbours and the slab flags start from zero: a
    // kernel rather than a memset node
    if (!job.flags || blocks > kMaxSlabFlags) return hipErrorInvalidValue;
    hipError_t e = launch_resident_reset(s, job.gb, job.flags, blocks);
    if (e != hipSuccess) return e;
    if (job.coherent)
        propagate_resident_kernel<true><<<blocks, kResThreads, resident_lds_bytes(), s>>>(
            job, t_begin, t_end, wait_ticks);
    else
        propagate_resident_kernel<false><<<blocks, kResThreads, resident_lds_bytes(), s>>>(
            job, t_begin, t_end, wait_ticks);
    return hipGetLastError();
}

hipError_t launch_resident_reset(hipStream_t s, GridBarrier *gb, uint32_t *flags, int32_t nflags) {
    if (flags && nflags > kMaxSlabFlags) return hipErrorInvalidValue;
    resident_reset_kernel<<<blocks_for(nflags > 64 ? nflags : 64, 256), 256, 0, s>>>(
        gb, flags, flags ? nflags : 0);
    return hipGetLastError();
}

hipError_t launch_measurement_noise(hipStream_t s, float *m, const float *sd, int32_t B,
                                    int32_t T, int32_t S, uint64_t seed) {
    const int64_t len = int64_t(B) * T * S;
    if (len <= 0) return hipSuccess;
    // key of stream 1000 of `seed`, as the host-side generator derives it
    const uint64_t key = splitmix64_host((seed << 20) ^ 1000u);
    measurement_noise_kernel<<<blocks_for(len, 256), 256, 0, s>>>(m, sd, B, T, S, key);
    return hipGetLastError();
}

hipError_t launch_measurement_matrix(hipStream_t s, const float *vp, const float *sp,
                                     const float *so, const float *off, float common, int32_t N,
                                     int32_t S, int32_t B, float *out) {
    const int64_t len = int64_t(B) * S * N;
    if (len <= 0) return hipSuccess;
    measurement_matrix_kernel<<<blocks_for(len, 256), 256, 0, s>>>(vp, sp, so, off, common, N, S,
                                                                   B, out);
    return hipGetLastError();
}

hipError_t launch_model_connections(hipStream_t s, int32_t nx, int32_t ny, int32_t nz,
                                   const int32_t *numbers, const float *pos, const float *vel,
                                   float rate, int32_t num_voxels, int32_t *osi, float *delays,
                                   int32_t *too_small) {
    const int64_t n_osi = int64_t(num_voxels) * 3 * 78, n_d = int64_t(num_voxels) * 26;
    if (n_osi > 0) {
        fill_words_kernel<<<blocks_for(n_osi, 256), 256, 0, s>>>(
            reinterpret_cast<uint32_t *>(osi), 0xFFFFFFFFu, n_osi);
        fill_kernel<<<blocks_for(n_d, 256), 256, 0, s>>>(delays, 0.0f, n_d);
    }
    fill_words_kernel<<<1, 64, 0, s>>>(reinterpret_cast<uint32_t *>(too_small), 0u, 1);
    const int64_t threads = int64_t(nx) * ny * nz * 27;
    if (threads > 0)
        model_connections_kernel<<<blocks_for(threads, 256), 256, 0, s>>>(
            nx, ny, nz, numbers, pos, vel, rate, osi, delays, too_small);
    return hipGetLastError();
}

hipError_t launch_spherical(hipStream_t s, const float *x, int32_t T, int32_t N, float rate,
                            float *mag, float *theta, float *phi, float *mag_max,
                            float *theta_max, float *phi_max, float *act_ms, int32_t *no_max) {
    fill_words_kernel<<<1, 64, 0, s>>>(reinterpret_cast<uint32_t *>(no_max), 0u, 1);
    if (N > 0)
        spherical_kernel<<<blocks_for(N, 256), 256, 0, s>>>(x, T, N, rate, mag, theta, phi, mag_max,
                                                            theta_max, phi_max, act_ms, no_max);
    return hipGetLastError();
}

hipError_t launch_segmentation_histogram(hipStream_t s, const float *x, int32_t T, int32_t N,
                                         const int32_t *numbers, const uint32_t *types,
                                         int64_t cells, const SegmentationThresholds &th,
                                         float *max_l1, uint32_t *hist) {
    constexpr int kWords = 2 * (kSegmentationThresholds + 1);
    fill_words_kernel<<<1, 256, 0, s>>>(hist, 0u, kWords);
    if (N > 0)
        voxel_max_l1_kernel<<<blocks_for(N, kMaxL1Voxels), kMaxL1Voxels * kMaxL1Slices, 0, s>>>(
            x, T, N, max_l1);
    if (cells > 0) {
        const unsigned blocks = blocks_for(cells, 256);  // grid-stride beyond 2048 blocks
        segmentation_histogram_kernel<<<blocks < 2048u ? blocks : 2048u, 256, 0, s>>>(
            numbers, types, cells, max_l1, th, hist);
    }
    return hipGetLastError();
}

hipError_t launch_barrier_reset(hipStream_t s, GridBarrier *gb, int32_t count) {
    if (count <= 0) return hipSuccess;
    barrier_reset_kernel<<<blocks_for(int64_t(count) * 10, 256), 256, 0, s>>>(gb, count);
    return hipGetLastError();
}

hipError_t batch_capacity(int32_t *blocks_per_cu) {
    return resident_fit(&propagate_batch_kernel<true>, &propagate_batch_kernel<false>,
                        blocks_per_cu);
}

hipError_t launch_propagate_batch(hipStream_t s, int32_t slots, const ResidentJob *jobs,
                                  const int32_t *task_ptr, const int32_t *task_job,
                                  GridBarrier *barriers, int32_t num_jobs, uint32_t *flags,
                                  int32_t num_flags, uint64_t first_wait, uint64_t step_wait,
                                  bool coherent) {
    if (slots <= 0 || num_jobs <= 0) return hipSuccess;
    barrier_reset_kernel<<<blocks_for(int64_t(num_jobs) * 10, 256), 256, 0, s>>>(barriers, num_jobs);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (flags && num_flags > 0) {
        fill_words_kernel<<<blocks_for(num_flags, 256), 256, 0, s>>>(flags, 0u, num_flags);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (coherent)
        propagate_batch_kernel<true><<<slots, kResThreads, resident_lds_bytes(), s>>>(
            jobs, task_ptr, task_job, first_wait, step_wait);
    else
        propagate_batch_kernel<false><<<slots, kResThreads, resident_lds_bytes(), s>>>(
            jobs, task_ptr, task_job, first_wait, step_wait);
    return hipGetLastError();
}

hipError_t launch_project(hipStream_t s, const float *x, const float *H, float *partial,
                          float *xT, int32_t T, int64_t n, int32_t S, int32_t Sp, int32_t splits,
                          int64_t kps) {
    if (S <= 0 || S > Sp) return hipErrorInvalidValue;
    dim3 grid(blocks_for(T, kProjTile), splits, project_sensor_tiles(S));
    project_kernel<<<grid, 256, 0, s>>>(x, H, partial, xT, T, n, Sp, kps);
    return hipGetLastError();
}

hipError_t launch_residuals(hipStream_t s, const float *partial, int32_t splits, const float *mhat,
                            float *m, float *R, float *loss_mse, int32_t T, int32_t S, int32_t Sp) {
    const int threads = Sp < 1024 ? Sp : 1024;
    residuals_kernel<<<T, threads, sizeof(float) * Sp, s>>>(partial, splits, mhat, m, R, loss_mse,
                                                            T, S, Sp);
    return hipGetLastError();
}

hipError_t launch_maxreg_partials(hipStream_t s, const float *x, float *partial, float *mr_last,
                                  int32_t N, int32_t T, float threshold, int32_t blocks) {
    dim3 grid(blocks, T);
    maxreg_partials_kernel<<<grid, 256, 0, s>>>(x, partial, mr_last, N, T, threshold);
    return hipGetLastError();
}

hipError_t launch_maxreg_finalize(hipStream_t s, const float *partial, int32_t blocks,
                                  float *step_mr, float *mr_sum, const float *loss_mse,
                                  float *loss_mr, float *loss, double *objective,
                                  float mse_strength, float reg_strength, int32_t T) {
    maxreg_step_sums_kernel<<<T, 256, 0, s>>>(partial, blocks, step_mr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    maxreg_prefix_kernel<<<1, 256, 0, s>>>(step_mr, mr_sum, loss_mse, loss_mr, loss, objective,
                                          mse_strength, reg_strength, T);
    return hipGetLastError();
}

hipError_t launch_transpose(hipStream_t s, const float *in, float *out, int64_t rows, int64_t cols) {
    if (rows <= 0 || cols <= 0) return hipSuccess;
    dim3 grid(blocks_for(cols, 32), blocks_for(rows, 32));
    transpose_kernel<<<grid, 256, 0, s>>>(in, out, rows, cols);
    return hipGetLastError();
}

hipError_t launch_map_residuals_beats(hipStream_t s, const float *H, const MapBeats &beats,
                                      const float *R, int64_t r_stride, float *E, int64_t n,
                                      int32_t T, int32_t S, int32_t Sp, bool accumulate) {
    if (n <= 0 || beats.count <= 0) return hipSuccess;
    if (beats.count > kMapMaxBeats || !map_residuals_beats_ok(S, T, n) || S > Sp)
        return hipErrorInvalidValue;
    const int64_t spans = (T + kMapBeatTiles * kMapRegTile - 1) / (kMapBeatTiles * kMapRegTile);
    const int64_t blocks = spans * ((n + 127) / 128);
    if (blocks > (int64_t(1) << 31) - 1) return hipErrorInvalidValue;
    map_residuals_beats_kernel<<<unsigned(blocks), 256, 0, s>>>(H, beats, R, r_stride, E, n, T,
                                                                (S + 1) / 2 * 2, Sp,
                                                                accumulate ? 1 : 0);
    return hipGetLastError();
}

bool map_residuals_beats_ok(int32_t S, int32_t T, int64_t n) {
    // the by-step kernel's limits, and the byte offset of a 128-row span
    return map_residuals_by_step_ok(S, T, n) &&
           n * kMapBeatTiles * kMapRegTile * 4 < (int64_t(1) << 31);
}

bool map_residuals_by_step_ok(int32_t S, int32_t T, int64_t n) {
    // the register-held kernel: at most 112 sensors; its buffer offsets
    // (4 B x 128 rows x T, and 4 B x 64 rows x n by step) fit in 31 bits
    return S > 0 && (S + 1) / 2 * 2 <= kProjRows && int64_t(T) * 128 * 4 < (int64_t(1) << 31) &&
           n * 64 * 4 < (int64_t(1) << 31);
}

hipError_t launch_map_residuals(hipStream_t s, const float *H, const float *R, float *out,
                                int64_t n, int32_t T, int32_t S, int32_t Sp, bool accumulate,
                                bool by_step) {
    if (n <= 0) return hipSuccess;
    if (S <= 0 || S > Sp) return hipErrorInvalidValue;
    if (by_step && !map_residuals_by_step_ok(S, T, n)) return hipErrorInvalidValue;
    float *eT = out;
    const int32_t Se = (S + 1) / 2 * 2;
    // the register-held kernel's buffer offsets (4 B x 128 rows x T) fit in 31 bits
    if (Se <= kProjRows && int64_t(T) * 128 * 4 < (int64_t(1) << 31)) {
        dim3 grid(blocks_for(n, 128), blocks_for(T, kMapRegSpan * kMapRegTile));
        if (by_step)
            map_residuals_regs_kernel<true><<<grid, 256, 0, s>>>(H, R, out, n, T, Se, Sp,
                                                                 accumulate ? 1 : 0);
        else
            map_residuals_regs_kernel<false><<<grid, 256, 0, s>>>(H, R, out, n, T, Se, Sp,
                                                                  accumulate ? 1 : 0);
        return hipGetLastError();
    }
    dim3 grid(blocks_for(n, 128), blocks_for(T, kMapTile));
    map_residuals_kernel<<<grid, 256, 0, s>>>(H, R, eT, n, T, Se, Sp, accumulate ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_derive(hipStream_t s, const int32_t *nbr, const float *coefs,
                         const int32_t *delays, const float *initial_delays, const float *gT,
                         const float *xT, const float *eT, float *fir, float *iir, float *dGT,
                         float *dA, int32_t N, int32_t T, const DeriveParams &p) {
    if (N <= 0) return hipSuccess;
    const unsigned blocks = blocks_for(int64_t(N) * 26, kDeriveThreads);
#define CT_DERIVE(G, C, S)                                                                  \
    derive_kernel<G, C, S><<<blocks, kDeriveThreads, 0, s>>>(                                \
        nbr, coefs, delays, initial_delays, gT, xT, eT, fir, iir, dGT, dA, N, T, p)
    const bool g = !p.freeze_gains, c = !p.freeze_delays;
    if (g && c && !p.simple) CT_DERIVE(true, true, false);
    else if (g && c) CT_DERIVE(true, true, true);
    else if (g) CT_DERIVE(true, false, false);
    else if (c && !p.simple) CT_DERIVE(false, true, false);
    else if (c) CT_DERIVE(false, true, true);
    else CT_DERIVE(false, false, false);
#undef CT_DERIVE
    return hipGetLastError();
}

#define CT_DERIVE_VARIANTS(X) \
    X(true, true, false) X(true, true, true) X(true, false, false) X(false, true, false) \
    X(false, true, true) X(false, false, false)

size_t derive_tiled_lds_bytes() { return kTileLdsBytes; }

hipError_t derive_tiled_prepare() {
    hipError_t e = hipSuccess;
#define CT_TILED_ATTR(G, C, S)                                                               \
    if (e == hipSuccess)                                                                     \
        e = hipFuncSetAttribute(reinterpret_cast<const void *>(&derive_tiled_kernel<G, C, S>), \
                                hipFuncAttributeMaxDynamicSharedMemorySize,                  \
                                int(derive_tiled_lds_bytes()));
    CT_DERIVE_VARIANTS(CT_TILED_ATTR)
#undef CT_TILED_ATTR
    return e;
}

hipError_t launch_derive_tiled(hipStream_t s, const DeriveTiles &tiles, const int32_t *nbr,
                               const float *coefs, const int32_t *delays,
                               const float *initial_delays, const float *gT, const float *xT,
                               const float *eT, float *fir, float *iir, float *dGT, float *dA,
                               int32_t N, int32_t T, const DeriveParams &p) {
    if (N <= 0 || tiles.tiles <= 0) return hipSuccess;
#define CT_DERIVE_TILED(G, C, S)                                                              \
    derive_tiled_kernel<G, C, S><<<tiles.tiles, kTileThreads, derive_tiled_lds_bytes(), s>>>( \
        tiles, nbr, coefs, delays, initial_delays, gT, xT, eT, fir, iir, dGT, dA, N, T, p)
    const bool g = !p.freeze_gains, c = !p.freeze_delays;
    if (g && c && !p.simple) CT_DERIVE_TILED(true, true, false);
    else if (g && c) CT_DERIVE_TILED(true, true, true);
    else if (g) CT_DERIVE_TILED(true, false, false);
    else if (c && !p.simple) CT_DERIVE_TILED(false, true, false);
    else if (c) CT_DERIVE_TILED(false, true, true);
    else CT_DERIVE_TILED(false, false, false);
#undef CT_DERIVE_TILED
    return hipGetLastError();
}

hipError_t launch_sgd(hipStream_t s, float *v, const float *d, int64_t len, float scale,
                      const UpdateGuard &guard) {
    if (len <= 0) return hipSuccess;
    sgd_kernel<<<blocks_for(len, 256), 256, 0, s>>>(v, d, len, scale, guard);
    return hipGetLastError();
}

hipError_t launch_adam(hipStream_t s, float *v, const float *d, float *m1, float *m2, int64_t len,
                       const UpdateScalars *sc, float scale, const UpdateGuard &guard) {
    if (len <= 0) return hipSuccess;
    adam_kernel<<<blocks_for(len, 256), 256, 0, s>>>(v, d, m1, m2, len, sc, scale, guard);
    return hipGetLastError();
}

hipError_t launch_roll_delays(hipStream_t s, float *coefs, int32_t *delays, int64_t len,
                              const UpdateGuard &guard) {
    if (len <= 0) return hipSuccess;
    roll_delays_kernel<<<blocks_for(len, 256), 256, 0, s>>>(coefs, delays, len, guard);
    return hipGetLastError();
}

hipError_t launch_average_delays(hipStream_t s, const float *gT, const float *coefs,
                                 const int32_t *delays, float *avg, uint8_t *valid, int32_t N) {
    if (N <= 0) return hipSuccess;
    average_delays_kernel<<<blocks_for(N, 256), 256, 0, s>>>(gT, coefs, delays, avg, valid, N);
    return hipGetLastError();
}

hipError_t launch_smoothness(hipStream_t s, const int32_t *nbr, const float *avg,
                             const uint8_t *valid, float *dA, float strength, int32_t N) {
    if (N <= 0) return hipSuccess;
    smoothness_kernel<<<blocks_for(N, 256), 256, 0, s>>>(nbr, avg, valid, dA, strength, N);
    return hipGetLastError();
}

hipError_t launch_batch_metrics(hipStream_t s, const float *loss, const float *loss_mse,
                                const float *loss_mr, float *loss_b, float *loss_mse_b,
                                float *loss_mr_b, const UpdateScalars *sc, int32_t T) {
    batch_metrics_kernel<<<1, 256, 0, s>>>(loss, loss_mse, loss_mr, loss_b, loss_mse_b, loss_mr_b,
                                          sc, T);
    return hipGetLastError();
}

hipError_t launch_expand_rows(hipStream_t s, const float *in, float *out, int64_t N) {
    if (N <= 0) return hipSuccess;
    expand_rows_kernel<<<blocks_for(N * 234, 256), 256, 0, s>>>(in, out, N);
    return hipGetLastError();
}

hipError_t launch_copy(hipStream_t s, float *dst, const float *src, int64_t len) {
    if (len <= 0) return hipSuccess;
    copy_kernel<<<blocks_for(len, 256), 256, 0, s>>>(dst, src, len);
    return hipGetLastError();
}

hipError_t launch_fill(hipStream_t s, float *p, float value, int64_t len) {
    if (len <= 0) return hipSuccess;
    fill_kernel<<<blocks_for(len, 256), 256, 0, s>>>(p, value, len);
    return hipGetLastError();
}

}  // namespace ct
