// Batched epochs: many scenarios (each a ct_epoch_kernel with its own results,
// measurements and configuration) advanced by one epoch together.
//
// The epoch of every scenario keeps its own order of operations
// (algorithm.rs:124-281), so each scenario's results equal those of its own
// ct_epoch_kernel_execute. What changes is how the work is laid out on the
// device: the beats are taken in rounds (round k = the k-th beat of every
// scenario that has one), and the propagation of all scenarios of a round runs
// in ONE persistent launch (propagate_batch_kernel): a job table holds each
// scenario's arrays, and every workgroup slot walks its own task list. The
// planner packs the jobs (team width = ceil(N / voxels per workgroup) slots,
// length ~ T steps) onto the slots longest first, and the autotuner picks the
// voxels per workgroup by timing the first round. Under the reverse-mode
// gradient the backward passes left pending by the scenarios' update batches
// are flushed together, their sweeps in ONE persistent launch
// (adjoint_sweep_batch_kernel) placed by the same planner. Optionally the
// whole epoch is captured once in a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>

#include "device.h"
#include "runtime.h"

using namespace ctrt;

struct ct_batch {
    GpuRef gpu;
    std::vector<ct_epoch_kernel *> kernels;
    int32_t slots = 0;              // resident workgroups of the batched launch
    int32_t voxels_per_group = 0;   // 0 = autotune on the first epoch
    std::vector<int32_t> tuned;     // candidates tried, for the profile
    std::vector<double> tuned_ms;

    // plan: per round, the jobs (scenario indices) and their device tables
    struct Round {
        std::vector<int32_t> scenarios;   // all scenarios with a beat in this round
        std::vector<int32_t> propagate;   // those that propagate in this round
        std::vector<int32_t> resident;    // of those, the ones in the batched launch (job order)
        int32_t job0 = 0;                 // first job / barrier of the round
        int32_t flag0 = 0, nflags = 0;    // the round's slab flags (one per team member)
        int32_t task0 = 0;                // first task of the round
        int32_t ptr0 = 0;                 // first task_ptr entry (slots + 1 per round)
        bool coherent = true;             // every job of the launch exchanges rows coherently
        double predicted_ms = 0.0;
    };
    std::vector<Round> rounds;
    std::vector<int32_t> plan_checkpoints;  // checkpoint interval per scenario the plan used
    std::vector<int32_t> plan_groups;       // beats per shared propagation (0 = not shared)
    int32_t plan_vpg = -1;
    ct::ResidentJob *jobs = nullptr;
    int32_t *task_ptr = nullptr, *task_job = nullptr;
    ct::GridBarrier *barriers = nullptr;
    uint32_t *flags = nullptr;  // slab flags of every job, job after job
    int32_t num_jobs = 0;

    // backward sweeps: per flush point (before round p's propagation, p =
    // rounds after the last round; see plan_sweeps), the pending passes of
    // reverse-mode scenarios whose sweeps run in that point's batched launch
    struct SweepRound {
        std::vector<int32_t> scenarios;         // job order
        std::vector<ct::AdjointParams> params;  // each job's sweep parameters
        int32_t job0 = 0;                       // first job / barrier of the point
        int32_t ptr0 = 0;                       // first task_ptr entry (sweep_slots + 1 per point)
    };
    std::vector<SweepRound> sweep_rounds;
    int32_t sweep_slots = 0;  // workgroups of the batched sweep launch (0 = unused)
    ct::SweepJob *sweep_jobs = nullptr;
    int32_t *sweep_task_ptr = nullptr, *sweep_task_job = nullptr;
    ct::GridBarrier *sweep_barriers = nullptr;
    int32_t num_sweep_jobs = 0;
    bool sweep_replan = false;  // a planned sweep job no longer matched its scenario

    // graph of one epoch
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool warmed = false;  // an eager epoch ran since the last plan (all buffers exist)
    int32_t graph_epochs_left = 0;  // replays before the adjoint rings are re-sized
    std::vector<std::vector<int32_t>> graph_orders;  // what the captured epoch was built for
    std::vector<ct_algorithm> graph_cfgs;
    std::vector<int32_t> graph_gradients;

    // profiling (whole-epoch and propagation-launch times)
    bool profiling = false;
    hipEvent_t ev[8] = {};
    std::vector<std::pair<hipEvent_t, hipEvent_t>> prop_events;
    double last_epoch_ms = 0.0, last_propagate_ms = 0.0;
    int64_t last_propagate_launches = 0, last_voxel_steps = 0;

    std::vector<void *> allocations;

    // Side streams (default priority) for the scenarios' work after each
    // round's propagation: scenario s runs on side[s % side.size()], after the
    // `fork` event recorded behind the propagation; the next persistent launch
    // (propagation or batched sweep) waits for every side stream's `done`
    // event, so no side work overlaps one. Used while every reverse-mode
    // member shares one backward pass per update batch: those passes are
    // deferred to the flushes, whose sweeps run on the GPU's stream. A member
    // that would sweep inside its beat (a persistent kernel whose workgroups
    // must all be resident at once, which two sweeps on two streams would
    // contend for) puts everything on the GPU's stream.
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> done;
    hipEvent_t fork = nullptr;
    bool side_active = false;  // this epoch uses the side streams

    ~ct_batch() {
        for (hipStream_t st : side) (void)hipStreamDestroy(st);
        for (hipEvent_t e : done) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto &p : prop_events) {
            (void)hipEventDestroy(p.first);
            (void)hipEventDestroy(p.second);
        }
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        for (void *p : allocations) (void)hipFree(p);
    }
};

namespace {

// Side streams per batch: with the GPU's own stream four, the hardware queues
// a process gets by default.
constexpr int kSideStreams = 3;

int batch_checkpoint(const ct_epoch_kernel &k) {
    return uses_adjoint(k) ? k.checkpoint_interval : 0;
}

bool scenario_resident(const ct_results_gpu &r) {
    return r.N > 0 && r.T > 0 && r.propagation_mode != 1 && r.resident_ok;
}

// Contiguous-window placement of the jobs of one round onto `slots`
// workgroup slots (longest first). Fills job slot0 / width and the per-slot
// task lists in one global order (start time, then placement), which keeps
// the launch deadlock-free. Returns the predicted makespan (ms).
template <typename Job>
double place_round(const std::vector<ct_results_gpu *> &rs, int32_t slots, int32_t vpg,
                   std::vector<Job> &jobs, std::vector<std::vector<int32_t>> &tasks) {
    const int32_t nj = int32_t(rs.size());
    std::vector<double> cost(nj);
    std::vector<int32_t> order(nj);
    for (int32_t j = 0; j < nj; ++j) {
        const int64_t N = rs[j]->N;
        int64_t width = std::min<int64_t>(slots, std::max<int64_t>(1, (N + vpg - 1) / vpg));
        // the LDS holds at most kResMaxVoxels voxels per workgroup
        width = std::max<int64_t>(width, (N + ct::kResMaxVoxels - 1) / ct::kResMaxVoxels);
        jobs[j].width = int32_t(width);
        const double per_group = double((N + width - 1) / width);
        // a step costs a barrier (~4 us) plus work proportional to the slab
        cost[j] = double(rs[j]->T) * (4.0e-3 + 1.0e-4 * per_group);
        order[j] = j;
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return cost[a] > cost[b]; });
    std::vector<double> avail(slots, 0.0), start(nj, 0.0);
    const int32_t per_xcd = std::max(1, slots / ct::kNumXcds);
    for (int32_t j : order) {
        const int32_t w = jobs[j].width;
        double best = 1e300;
        int32_t best_s0 = 0;
        bool best_cross = true;
        for (int32_t s0 = 0; s0 + w <= slots; ++s0) {
            double st = 0.0;
            for (int32_t q = s0; q < s0 + w; ++q) st = std::max(st, avail[q]);
            // prefer windows inside one XCD (shared L2) at equal start
            const bool cross = w <= per_xcd && (s0 / per_xcd) != ((s0 + w - 1) / per_xcd);
            if (st < best - 1e-9 || (st < best + 1e-9 && best_cross && !cross)) {
                best = st;
                best_s0 = s0;
                best_cross = cross;
            }
        }
        jobs[j].slot0 = best_s0;
        start[j] = best;
        for (int32_t q = best_s0; q < best_s0 + w; ++q) avail[q] = best + cost[j];
    }
    // one global order for every slot: by start time, ties by placement
    std::vector<int32_t> rank(nj);
    for (int32_t i = 0; i < nj; ++i) rank[order[i]] = i;
    std::vector<int32_t> global(nj);
    for (int32_t j = 0; j < nj; ++j) global[j] = j;
    std::stable_sort(global.begin(), global.end(), [&](int32_t a, int32_t b) {
        if (start[a] != start[b]) return start[a] < start[b];
        return rank[a] < rank[b];
    });
    tasks.assign(slots, {});
    for (int32_t j : global)
        for (int32_t q = jobs[j].slot0; q < jobs[j].slot0 + jobs[j].width; ++q)
            tasks[q].push_back(j);
    double makespan = 0.0;
    for (double a : avail) makespan = std::max(makespan, a);
    return makespan;
}

template <typename V>
int batch_alloc(ct_batch &b, V **p, int64_t count) {
    void *raw = nullptr;
    CT_HIP(hipMalloc(&raw, std::max<size_t>(sizeof(V) * size_t(std::max<int64_t>(count, 1)), 256)));
    b.allocations.push_back(raw);
    *p = static_cast<V *>(raw);
    return CT_OK;
}

void batch_free(ct_batch &b, void *p) {
    if (!p) return;
    auto it = std::find(b.allocations.begin(), b.allocations.end(), p);
    if (it != b.allocations.end()) b.allocations.erase(it);
    (void)hipFree(p);
}

void drop_graph(ct_batch &b) {
    if (b.exec) (void)hipGraphExecDestroy(b.exec);
    if (b.graph) (void)hipGraphDestroy(b.graph);
    b.exec = nullptr;
    b.graph = nullptr;
    b.warmed = false;
}

// The batched backward sweeps. A reverse-mode scenario that shares one
// backward pass per update batch leaves that pass (and the update after it)
// pending when its update batch ends; the pending passes are flushed together
// just before a round in which one of their scenarios propagates again (its
// states would be overwritten), and at the end of the epoch. Flush point p <
// rounds is the one before round p's propagation, p = rounds the last. The
// jobs of a point are its pending scenarios whose sweep the resident kernel
// with the coherent exchange runs (sizes as for the batched propagation: a
// scenario needing more than half the slots sweeps on its own). Allocates the
// scenarios' backward buffers now: the job table holds their addresses.
int plan_sweeps(ct_batch &b, int32_t vpg) {
    const int32_t S = int32_t(b.kernels.size());
    const int32_t rounds = int32_t(b.rounds.size());
    const int32_t slots = b.slots > 0 ? b.sweep_slots : 0;
    b.sweep_rounds.assign(rounds + 1, {});
    b.sweep_replan = false;
    std::vector<ct::SweepJob> all_jobs;
    std::vector<int32_t> all_ptr, all_task;
    std::vector<int32_t> pending(S, 0);  // beats of the pending pass
    for (int32_t pt = 0; pt <= rounds; ++pt) {
        ct_batch::SweepRound &W = b.sweep_rounds[pt];
        bool flush = pt == rounds;
        if (!flush)
            for (int32_t s : b.rounds[pt].propagate) flush = flush || pending[s] > 0;
        std::vector<ct_results_gpu *> rs;
        std::vector<ct::SweepJob> jobs;
        for (int32_t s = 0; s < S && flush; ++s) {
            if (pending[s] <= 0) continue;
            const int32_t beats = pending[s];
            pending[s] = 0;
            ct_epoch_kernel &k = *b.kernels[s];
            ct_results_gpu &r = *k.r;
            if (slots <= 0 || (r.N + ct::kResMaxVoxels - 1) / ct::kResMaxVoxels > slots ||
                int64_t(r.N) * 2 > int64_t(slots) * ct::kResMaxVoxels)
                continue;
            CT_CHECK(prepare_backward(r, k.checkpoint_interval));
            if (!sweep_batchable(r)) continue;
            const ct::AdjointParams p = adjoint_params(r, k.cfg, beats);
            ct::SweepJob j{};
            j.cons_ptr = r.cons_ptr;
            j.cons_idx = r.cons_idx;
            j.delays = r.delays;
            j.gains_rows = r.staging;
            j.coefs = r.coefs;
            j.gcons = r.gcons;
            j.x = r.x;
            j.E = r.E;
            j.xbar = r.xbar;
            j.N = r.N;
            j.T = r.T;
            j.p = p;
            j.coherent = 1;
            W.scenarios.push_back(s);
            W.params.push_back(p);
            jobs.push_back(j);
            rs.push_back(&r);
        }
        // the passes the round's beats leave pending
        for (int32_t s = 0; pt < rounds && s < S; ++s) {
            const ct_epoch_kernel &k = *b.kernels[s];
            const ct_results_gpu &r = *k.r;
            if (!uses_adjoint(k) || !shares_propagation(k) || r.N <= 0 || r.B <= pt) continue;
            const int32_t group = update_group(k), first = pt - pt % group;
            const int32_t last = std::min<int32_t>(first + group, r.B) - 1;
            if (pt == last) pending[s] = last - first + 1;
        }
        W.job0 = int32_t(all_jobs.size());
        W.ptr0 = int32_t(all_ptr.size());
        std::vector<std::vector<int32_t>> tasks;
        if (!rs.empty()) (void)place_round(rs, slots, vpg, jobs, tasks);
        all_jobs.insert(all_jobs.end(), jobs.begin(), jobs.end());
        int32_t acc = int32_t(all_task.size());
        for (int32_t q = 0; q < slots; ++q) {
            all_ptr.push_back(acc);
            if (!tasks.empty())
                for (int32_t j : tasks[q]) {
                    all_task.push_back(j);
                    ++acc;
                }
        }
        all_ptr.push_back(acc);
    }
    batch_free(b, b.sweep_jobs);
    batch_free(b, b.sweep_task_ptr);
    batch_free(b, b.sweep_task_job);
    batch_free(b, b.sweep_barriers);
    b.sweep_jobs = nullptr;
    b.sweep_task_ptr = b.sweep_task_job = nullptr;
    b.sweep_barriers = nullptr;
    b.num_sweep_jobs = int32_t(all_jobs.size());
    if (all_jobs.empty()) return CT_OK;
    CT_CHECK(batch_alloc(b, &b.sweep_jobs, b.num_sweep_jobs));
    CT_CHECK(batch_alloc(b, &b.sweep_task_ptr, int64_t(all_ptr.size())));
    CT_CHECK(batch_alloc(b, &b.sweep_task_job, int64_t(all_task.size())));
    CT_CHECK(batch_alloc(b, &b.sweep_barriers, b.num_sweep_jobs));
    for (size_t j = 0; j < all_jobs.size(); ++j) all_jobs[j].gb = b.sweep_barriers + j;
    CT_HIP(hipMemset(b.sweep_barriers, 0, sizeof(ct::GridBarrier) * all_jobs.size()));
    CT_CHECK(upload(b.sweep_jobs, all_jobs.data(), sizeof(ct::SweepJob) * all_jobs.size()));
    CT_CHECK(upload(b.sweep_task_ptr, all_ptr.data(), sizeof(int32_t) * all_ptr.size()));
    CT_CHECK(upload(b.sweep_task_job, all_task.data(), sizeof(int32_t) * all_task.size()));
    return CT_OK;
}

// Builds the rounds and the device job tables for `vpg` voxels per group.
int plan(ct_batch &b, int32_t vpg) {
    CT_HIP(hipStreamSynchronize(b.gpu->stream));
    drop_graph(b);
    const int32_t S = int32_t(b.kernels.size());
    int32_t rounds = 0;
    for (ct_epoch_kernel *k : b.kernels) rounds = std::max(rounds, k->r->B);
    b.rounds.assign(rounds, {});
    b.plan_checkpoints.resize(S);
    b.plan_groups.resize(S);
    std::vector<ct::ResidentJob> all_jobs;
    std::vector<int32_t> all_ptr, all_task;
    for (int32_t s = 0; s < S; ++s) {
        ct_epoch_kernel &k = *b.kernels[s];
        b.plan_checkpoints[s] = batch_checkpoint(k);
        b.plan_groups[s] = shares_propagation(k) ? update_group(k) : 0;
        // checkpoint buffers are allocated now: the job table holds their address
        if (b.plan_checkpoints[s] > 0 && k.r->N > 0)
            CT_CHECK(ensure_checkpoints(*k.r, b.plan_checkpoints[s]));
    }
    for (int32_t rd = 0; rd < rounds; ++rd) {
        ct_batch::Round &R = b.rounds[rd];
        std::vector<ct_results_gpu *> rs;
        for (int32_t s = 0; s < S; ++s) {
            ct_results_gpu &r = *b.kernels[s]->r;
            if (r.B <= rd) continue;
            R.scenarios.push_back(s);
            if (!propagates_at(*b.kernels[s], rd)) continue;
            R.propagate.push_back(s);
            const bool fits = scenario_resident(r) &&
                              (r.N + ct::kResMaxVoxels - 1) / ct::kResMaxVoxels <= b.slots;
            // a scenario that alone needs more than half the slots at the
            // largest slab gains nothing from sharing the launch: it runs its
            // own resident launch, whose kernel also keeps gain rows in
            // registers (8 shells of 115k voxels: 53 against 58 ms each)
            const bool alone = int64_t(r.N) * 2 > int64_t(b.slots) * ct::kResMaxVoxels;
            if (fits && !alone && b.slots > 0) {
                R.resident.push_back(s);
                rs.push_back(&r);
            }
        }
        R.job0 = int32_t(all_jobs.size());
        R.task0 = int32_t(all_task.size());
        R.ptr0 = int32_t(all_ptr.size());
        std::vector<ct::ResidentJob> jobs;
        for (int32_t s : R.resident) {
            ct_results_gpu &r = *b.kernels[s]->r;
            jobs.push_back(resident_job(r, b.plan_checkpoints[s]));
            R.coherent = R.coherent && jobs.back().coherent != 0;
        }
        std::vector<std::vector<int32_t>> tasks;
        R.predicted_ms = rs.empty() ? 0.0 : place_round(rs, b.slots, vpg, jobs, tasks);
        all_jobs.insert(all_jobs.end(), jobs.begin(), jobs.end());
        // task_ptr of this round: slots + 1 entries into task_job, whose
        // entries index the round's jobs (the launch gets jobs + job0)
        int32_t acc = R.task0;
        for (int32_t q = 0; q < b.slots; ++q) {
            all_ptr.push_back(acc);
            if (!tasks.empty())
                for (int32_t j : tasks[q]) {
                    all_task.push_back(j);
                    ++acc;
                }
        }
        all_ptr.push_back(acc);
    }
    // device tables (barrier addresses patched in now that the array exists)
    batch_free(b, b.jobs);
    batch_free(b, b.task_ptr);
    batch_free(b, b.task_job);
    batch_free(b, b.barriers);
    batch_free(b, b.flags);
    b.jobs = nullptr;
    b.task_ptr = b.task_job = nullptr;
    b.barriers = nullptr;
    b.flags = nullptr;
    b.num_jobs = int32_t(all_jobs.size());
    // each job's members publish their completed steps in its own flag range
    std::vector<int32_t> flag_of(all_jobs.size() + 1, 0);
    for (size_t j = 0; j < all_jobs.size(); ++j) flag_of[j + 1] = flag_of[j] + all_jobs[j].width;
    for (ct_batch::Round &R : b.rounds) {
        const size_t j1 = size_t(R.job0) + R.resident.size();
        R.flag0 = flag_of[size_t(R.job0)];
        R.nflags = flag_of[j1] - R.flag0;
    }
    CT_CHECK(batch_alloc(b, &b.jobs, b.num_jobs));
    CT_CHECK(batch_alloc(b, &b.task_ptr, int64_t(all_ptr.size())));
    CT_CHECK(batch_alloc(b, &b.task_job, int64_t(all_task.size())));
    CT_CHECK(batch_alloc(b, &b.barriers, 1));
    CT_CHECK(batch_alloc(b, &b.flags, std::max<int64_t>(flag_of.back(), 1)));
    // one abort word for every job: a timed-out wait ends the whole launch at
    // the other jobs' next waits instead of leaving them to wait out their own
    // limits for workgroups that have left
    for (size_t j = 0; j < all_jobs.size(); ++j) {
        all_jobs[j].gb = b.barriers;
        all_jobs[j].flags = b.flags + flag_of[j];
    }
    CT_HIP(hipMemset(b.barriers, 0, sizeof(ct::GridBarrier)));
    CT_CHECK(upload(b.jobs, all_jobs.data(), sizeof(ct::ResidentJob) * all_jobs.size()));
    CT_CHECK(upload(b.task_ptr, all_ptr.data(), sizeof(int32_t) * all_ptr.size()));
    CT_CHECK(upload(b.task_job, all_task.data(), sizeof(int32_t) * all_task.size()));
    CT_CHECK(plan_sweeps(b, vpg));
    b.plan_vpg = vpg;
    return CT_OK;
}

bool plan_current(const ct_batch &b) {
    if (b.plan_vpg != b.voxels_per_group || b.sweep_replan) return false;
    for (size_t s = 0; s < b.kernels.size(); ++s) {
        const ct_epoch_kernel &k = *b.kernels[s];
        if (b.plan_checkpoints[s] != batch_checkpoint(k) ||
            b.plan_groups[s] != (shares_propagation(k) ? update_group(k) : 0))
            return false;
    }
    return true;
}

// The propagation of every scenario of round `rd`: the batched launch for the
// resident jobs, the per-step path for the rest.
int propagate_round(ct_batch &b, int32_t rd, bool timed) {
    const ct_batch::Round &R = b.rounds[rd];
    hipStream_t s = b.gpu->stream;
    if (R.propagate.empty()) return CT_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (timed) {  // the round's whole propagation: own launches and the batched one
        CT_HIP(hipEventCreate(&e0));
        CT_HIP(hipEventCreate(&e1));
        b.prop_events.push_back({e0, e1});
        CT_HIP(hipEventRecord(e0, s));
    }
    for (int32_t sc : R.propagate) {
        ct_epoch_kernel &k = *b.kernels[sc];
        const bool in_launch =
            std::find(R.resident.begin(), R.resident.end(), sc) != R.resident.end();
        if (in_launch)
            CT_CHECK(prepare_propagation(*k.r, b.plan_checkpoints[sc]));
        else
            CT_CHECK(propagate_beat(*k.r, nullptr, b.plan_checkpoints[sc]));
    }
    if (R.resident.empty()) {
        if (timed) CT_HIP(hipEventRecord(e1, s));
        return CT_OK;
    }
    // the members' wait limits (ct_results_set_wait_limit): the shortest bounds
    // every step; a job's first wait, which may include earlier jobs on its
    // slots, gets up to 30 times as long (at most 60 s)
    double step_s = 1e300;
    for (const ct_epoch_kernel *k : b.kernels) step_s = std::min(step_s, k->r->wait_seconds);
    const double first_s = std::min(60.0, 30.0 * step_s);
    CT_HIP(ct::launch_propagate_batch(s, b.slots, b.jobs + R.job0, b.task_ptr + R.ptr0,
                                      b.task_job, b.barriers, 1, b.flags + R.flag0, R.nflags,
                                      wait_ticks(*b.gpu, first_s), wait_ticks(*b.gpu, step_s),
                                      R.coherent));
    if (timed) CT_HIP(hipEventRecord(e1, s));
    return CT_OK;
}

// One epoch of every scenario, enqueued on the stream (no synchronisation).
// Every side stream's work so far before the GPU stream's next work.
int join_side(ct_batch &b) {
    if (!b.side_active) return CT_OK;
    for (size_t j = 0; j < b.side.size(); ++j) {
        CT_HIP(hipEventRecord(b.done[j], b.side[j]));
        CT_HIP(hipStreamWaitEvent(b.gpu->stream, b.done[j], 0));
    }
    return CT_OK;
}

// The GPU stream's work so far before every side stream's next work.
int fork_side(ct_batch &b) {
    if (!b.side_active) return CT_OK;
    CT_HIP(hipEventRecord(b.fork, b.gpu->stream));
    for (hipStream_t st : b.side) CT_HIP(hipStreamWaitEvent(st, b.fork, 0));
    return CT_OK;
}

// Runs `fn` with scenario s's work on its side stream (if any).
template <typename F>
int on_side(ct_batch &b, int32_t s, F &&fn) {
    ct_results_gpu &r = *b.kernels[s]->r;
    const hipStream_t keep = r.stream;
    if (b.side_active) r.stream = b.side[size_t(s) % b.side.size()];
    const int rc = fn();
    r.stream = keep;
    return rc;
}

int enqueue_epoch_body(ct_batch &b, bool timed);

// The members' updates also check the batched launch's abort word.
int enqueue_epoch(ct_batch &b, bool timed) {
    for (ct_epoch_kernel *k : b.kernels) {
        k->r->batch_abort = &b.barriers->abort.count;
        k->defer_backward = true;
        k->deferred_beats = 0;
    }
    const int rc = enqueue_epoch_body(b, timed);
    for (ct_epoch_kernel *k : b.kernels) {
        k->r->batch_abort = nullptr;
        k->defer_backward = false;
        k->deferred_beats = 0;
    }
    return rc;
}

// Flush point `pt` (see plan_sweeps): if it is the last, or a scenario of
// round pt's propagation has a pending backward pass, every pending pass
// runs now. The planned jobs whose scenario still matches (pending pass,
// sweep parameters, resident sweep possible) get their sweep inputs, one
// batched sweep launch and their gradient passes; any other pending pass runs
// whole on its scenario's own, after the launch (overwriting what the launch
// wrote for it), and the next epoch plans again. Then the updates that waited
// for those passes: `upd[s]` beats of scenario s (update number update[s]).
int flush_backward(ct_batch &b, int32_t pt, std::vector<int64_t> &upd,
                   std::vector<int32_t> &update) {
    const int32_t rounds = int32_t(b.rounds.size());
    bool flush = pt >= rounds;
    if (!flush)
        for (int32_t s : b.rounds[pt].propagate) flush = flush || b.kernels[s]->deferred_beats > 0;
    if (!flush) return CT_OK;
    const ct_batch::SweepRound *W =
        pt < int32_t(b.sweep_rounds.size()) ? &b.sweep_rounds[pt] : nullptr;
    std::vector<int32_t> in_launch;
    CT_CHECK(join_side(b));  // the pending passes' forcing may be on side streams
    if (W && !W->scenarios.empty()) {
        for (size_t i = 0; i < W->scenarios.size(); ++i) {
            ct_epoch_kernel &k = *b.kernels[W->scenarios[i]];
            const ct::AdjointParams p = adjoint_params(*k.r, k.cfg, k.deferred_beats);
            const bool ok = k.deferred_beats > 0 && sweep_batchable(*k.r) &&
                            std::memcmp(&p, &W->params[i], sizeof(p)) == 0;
            if (!ok) {
                b.sweep_replan = true;
                continue;
            }
            CT_CHECK(adjoint_backward_begin(*k.r, &k));
            in_launch.push_back(W->scenarios[i]);
        }
        if (!in_launch.empty()) {
            double step_s = 1e300;
            for (const ct_epoch_kernel *k : b.kernels) step_s = std::min(step_s, k->r->wait_seconds);
            const double first_s = std::min(60.0, 30.0 * step_s);
            CT_HIP(ct::launch_adjoint_sweep_batch(
                b.gpu->stream, b.sweep_slots, b.sweep_jobs + W->job0, b.sweep_barriers + W->job0,
                int32_t(W->scenarios.size()), b.sweep_task_ptr + W->ptr0, b.sweep_task_job,
                &b.barriers->abort.count, wait_ticks(*b.gpu, first_s), wait_ticks(*b.gpu, step_s)));
        }
    }
    // the other pending passes whole, sweeps included, on the GPU's stream
    const int32_t S = int32_t(b.kernels.size());
    for (int32_t s = 0; s < S; ++s) {
        ct_epoch_kernel &k = *b.kernels[s];
        if (k.deferred_beats > 0 &&
            std::find(in_launch.begin(), in_launch.end(), s) == in_launch.end())
            CT_CHECK(adjoint_backward_deferred(k));
    }
    // gradient passes of the batched sweeps and the waiting updates on the
    // side streams, joined before the next persistent launch
    CT_CHECK(fork_side(b));
    for (int32_t s = 0; s < S; ++s) {
        ct_epoch_kernel &k = *b.kernels[s];
        CT_CHECK(on_side(b, s, [&] {
            if (k.deferred_beats > 0) {
                k.deferred_beats = 0;
                CT_CHECK(adjoint_backward_end(*k.r, k.checkpoint_interval, &k));
            }
            if (upd[s] > 0) {
                CT_CHECK(finish_batch(k, upd[s], true, update[s]++));
                upd[s] = 0;
            }
            return CT_OK;
        }));
    }
    return join_side(b);
}

int enqueue_epoch_body(ct_batch &b, bool timed) {
    const int32_t S = int32_t(b.kernels.size());
    std::vector<int64_t> in_batch(S, 0), upd(S, 0);
    std::vector<int32_t> update(S, 0);
    b.side_active = !b.side.empty() &&
                    std::none_of(b.kernels.begin(), b.kernels.end(), [](const ct_epoch_kernel *k) {
                        return uses_adjoint(*k) && !shares_propagation(*k);
                    });
    for (int32_t s = 0; s < S; ++s) CT_CHECK(reset_derivatives(*b.kernels[s]->r));
    for (int32_t rd = 0; rd < int32_t(b.rounds.size()); ++rd) {
        // the round's propagation overwrites states the previous round's
        // beats read on the side streams
        if (rd > 0) CT_CHECK(join_side(b));
        // pending backward passes (reverse-mode scenarios only, which leave
        // the side streams unused) before states they read are overwritten
        CT_CHECK(flush_backward(b, rd, upd, update));
        CT_CHECK(propagate_round(b, rd, timed));
        CT_CHECK(fork_side(b));
        for (int32_t s : b.rounds[rd].scenarios) {
            ct_epoch_kernel &k = *b.kernels[s];
            CT_CHECK(on_side(b, s, [&] {
                CT_CHECK(finish_beat(k, rd));
                if (k.cfg.batch_size != 0 && ++in_batch[s] == k.cfg.batch_size) {
                    // after the backward pass, if the beat left it pending
                    if (k.deferred_beats > 0)
                        upd[s] = in_batch[s];
                    else
                        CT_CHECK(finish_batch(k, in_batch[s], true, update[s]++));
                    in_batch[s] = 0;
                }
                return CT_OK;
            }));
        }
    }
    CT_CHECK(join_side(b));
    CT_CHECK(flush_backward(b, int32_t(b.rounds.size()), upd, update));
    for (int32_t s = 0; s < S; ++s) {
        ct_epoch_kernel &k = *b.kernels[s];
        CT_CHECK(on_side(b, s, [&] {
            if (k.cfg.batch_size != 0) {
                if (in_batch[s] > 0) CT_CHECK(finish_batch(k, in_batch[s], false, update[s]++));
                return CT_OK;
            }
            return finish_batch(k, k.r->B, false, update[s]++);
        }));
    }
    return join_side(b);
}

// Reads every abort word after an epoch: the batched launches' and each
// scenario's own (its backward sweep, or its resident propagation outside the
// batched launch; after a graph replay the host cannot tell which ran, so all
// are read). A scenario whose update was skipped on the device (either word
// set) keeps its epoch counters; the others advance theirs. Any abort drops
// the graph and fails the call; the paths that timed out are not used again.
// `epoch`: after an epoch (advance the counters), not after the autotuner's
// propagation-only launches.
int check_barriers(ct_batch &b, bool epoch) {
    CT_HIP(hipStreamSynchronize(b.gpu->stream));
    uint32_t batch_abort = 0;
    if (b.barriers)
        CT_CHECK(download(&batch_abort, &b.barriers->abort.count, sizeof(batch_abort)));
    if (batch_abort) {
        // stop using the batched launch; the next execute plans again, with
        // every scenario on its own propagation path
        b.slots = 0;
        b.plan_vpg = -1;
        const uint32_t zero = 0;
        CT_CHECK(upload(&b.barriers->abort.count, &zero, sizeof(zero)));
    }
    int rc = CT_OK;
    if (batch_abort)
        rc = fail(CT_ERR_DEVICE, "batched propagation: a wait timed out (workgroups not all "
                                 "resident, or a wait limit too short); the epoch's updates were "
                                 "not applied, and the per-scenario paths are used from now on");
    for (ct_epoch_kernel *k : b.kernels) {
        const int own = check_resident(*k->r, true);
        if (epoch && own == CT_OK && !batch_abort) end_epoch(*k);
        if (own != CT_OK && rc == CT_OK) rc = own;
    }
    if (rc != CT_OK) {
        // the scenarios whose resident paths timed out leave the batched
        // launch too: plan again at the next execute
        const std::string msg = g_error;
        drop_graph(b);
        b.plan_vpg = -1;
        return fail(rc, msg);
    }
    return CT_OK;
}

// Autotuner: times round 0's batched propagation for each candidate number of
// voxels per workgroup and keeps the fastest. The timed launches write the
// same arrays the epoch then recomputes from scratch.
int autotune(ct_batch &b) {
    static const int32_t kCandidates[] = {472, 384, 320, 256, 192, 160, 128, 96, 64};
    b.tuned.clear();
    b.tuned_ms.clear();
    int32_t best = ct::kResMaxVoxels;
    double best_ms = 1e300;
    std::vector<std::vector<int32_t>> seen;
    for (int32_t vpg : kCandidates) {
        CT_CHECK(plan(b, vpg));
        if (b.rounds.empty() || b.rounds[0].resident.empty()) {
            b.voxels_per_group = best;  // nothing to tune: no batched launch in round 0
            return CT_OK;
        }
        // skip candidates that give the same team widths as one already timed
        std::vector<ct::ResidentJob> j0(b.rounds[0].resident.size());
        CT_CHECK(download(j0.data(), b.jobs + b.rounds[0].job0, sizeof(ct::ResidentJob) * j0.size()));
        std::vector<int32_t> widths;
        for (auto &j : j0) widths.push_back(j.width);
        if (std::find(seen.begin(), seen.end(), widths) != seen.end()) continue;
        seen.push_back(widths);
        hipEvent_t e0, e1;
        CT_HIP(hipEventCreate(&e0));
        CT_HIP(hipEventCreate(&e1));
        int rc = propagate_round(b, 0, false);
        if (rc == CT_OK && hipEventRecord(e0, b.gpu->stream) != hipSuccess) rc = CT_ERR_DEVICE;
        if (rc == CT_OK) rc = propagate_round(b, 0, false);
        if (rc == CT_OK && hipEventRecord(e1, b.gpu->stream) != hipSuccess) rc = CT_ERR_DEVICE;
        float ms = 0.0f;
        if (rc == CT_OK && hipEventSynchronize(e1) == hipSuccess) (void)hipEventElapsedTime(&ms, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        CT_CHECK(rc);
        CT_CHECK(check_barriers(b, false));
        b.tuned.push_back(vpg);
        b.tuned_ms.push_back(ms);
        if (ms < best_ms) {
            best_ms = ms;
            best = vpg;
        }
    }
    b.voxels_per_group = best;
    return CT_OK;
}

constexpr int32_t kGraphEpochs = 16;

bool same_cfg(const ct_algorithm &a, const ct_algorithm &b) {
    return std::memcmp(&a, &b, sizeof(ct_algorithm)) == 0;
}

// Is the captured epoch still the epoch the scenarios would run now?
bool graph_current(const ct_batch &b) {
    if (!b.exec || b.graph_epochs_left <= 0) return false;
    for (size_t s = 0; s < b.kernels.size(); ++s) {
        const ct_epoch_kernel &k = *b.kernels[s];
        if (k.order != b.graph_orders[s] || !same_cfg(k.cfg, b.graph_cfgs[s]) ||
            k.gradient * 100000 + k.checkpoint_interval + (shares_propagation(k) ? 1 << 30 : 0) !=
                b.graph_gradients[s])
            return false;
    }
    return true;
}

// Captures one epoch of the batch into a graph. The adjoint's q rings get a
// fixed size with room for the delays to grow during kGraphEpochs replays.
int capture(ct_batch &b) {
    hipStream_t s = b.gpu->stream;
    drop_graph(b);
    b.warmed = true;
    b.graph_orders.clear();
    b.graph_cfgs.clear();
    b.graph_gradients.clear();
    for (ct_epoch_kernel *k : b.kernels) {
        if (uses_adjoint(*k))
            CT_CHECK(fix_adjoint_ring(*k->r, int64_t(kGraphEpochs) * updates_per_epoch(*k)));
        b.graph_orders.push_back(k->order);
        b.graph_cfgs.push_back(k->cfg);
        b.graph_gradients.push_back(k->gradient * 100000 + k->checkpoint_interval +
                                    (shares_propagation(*k) ? 1 << 30 : 0));
    }
    CT_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_epoch(b, false);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    for (ct_epoch_kernel *k : b.kernels) k->r->ring_slots_fixed = 0;
    if (rc != CT_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    CT_HIP(e);
    b.graph = g;
    CT_HIP(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
    b.graph_epochs_left = kGraphEpochs;
    return CT_OK;
}

int clear_prop_events(ct_batch &b) {
    for (auto &p : b.prop_events) {
        (void)hipEventDestroy(p.first);
        (void)hipEventDestroy(p.second);
    }
    b.prop_events.clear();
    return CT_OK;
}

}  // namespace

extern "C" {

int ct_batch_new(ct_gpu *gpu, ct_epoch_kernel *const *kernels, int32_t count, ct_batch **out) {
    if (!gpu || !out || (count > 0 && !kernels)) return fail(CT_ERR_INVALID, "null argument");
    if (count < 1) return fail(CT_ERR_INVALID, "a batch needs at least one scenario");
    CT_CHECK(check_gpu(gpu));
    for (int32_t i = 0; i < count; ++i) {
        if (!kernels[i]) return fail(CT_ERR_INVALID, "null epoch kernel in the batch");
        if (kernels[i]->gpu != gpu || kernels[i]->r->gpu != gpu)
            return fail(CT_ERR_INVALID, "every scenario of a batch must live on the batch's GPU");
        for (int32_t j = 0; j < i; ++j)
            if (kernels[j]->r == kernels[i]->r)
                return fail(CT_ERR_INVALID, "two scenarios of a batch share one results object");
    }
    ct_batch *b = new ct_batch;
    b->gpu = gpu;
    b->kernels.assign(kernels, kernels + count);
    int cus = 0;
    int32_t per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, gpu->device) ==
            hipSuccess &&
        ct::batch_capacity(&per_cu) == hipSuccess && cus > 0 && per_cu > 0)
        b->slots = cus * per_cu;
    int32_t sweep_per_cu = 0;
    if (cus > 0 && ct::sweep_batch_capacity(&sweep_per_cu) == hipSuccess && sweep_per_cu > 0)
        b->sweep_slots = cus * sweep_per_cu;
    (void)hipGetLastError();
    for (int j = 0; j < kSideStreams; ++j) {
        hipStream_t st = nullptr;
        hipEvent_t e = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) break;
        b->side.push_back(st);
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            delete b;
            return fail(CT_ERR_DEVICE, "hipEventCreateWithFlags failed");
        }
        b->done.push_back(e);
    }
    if (b->side.size() != size_t(kSideStreams) ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) {
        delete b;
        return fail(CT_ERR_DEVICE, "creating the batch's side streams failed");
    }
    *out = b;
    return CT_OK;
}

void ct_batch_free(ct_batch *batch) {
    if (!batch) return;
    (void)hipSetDevice(batch->gpu->device);
    (void)hipStreamSynchronize(batch->gpu->stream);
    delete batch;
}

int ct_batch_set_voxels_per_group(ct_batch *b, int32_t voxels) {
    if (!b) return fail(CT_ERR_INVALID, "null argument");
    if (voxels != 0 && (voxels < 1 || voxels > ct::kResMaxVoxels))
        return fail(CT_ERR_INVALID, "voxels per workgroup must be 0 (autotune) or 1.." +
                                        std::to_string(ct::kResMaxVoxels));
    b->voxels_per_group = voxels;
    return CT_OK;
}

int ct_batch_set_graph(ct_batch *b, int32_t enabled) {
    if (!b) return fail(CT_ERR_INVALID, "null argument");
    CT_CHECK(check_gpu(b->gpu));
    CT_HIP(hipStreamSynchronize(b->gpu->stream));
    drop_graph(*b);
    b->use_graph = enabled != 0;
    return CT_OK;
}

int ct_batch_set_profiling(ct_batch *b, int32_t enabled) {
    if (!b) return fail(CT_ERR_INVALID, "null argument");
    b->profiling = enabled != 0;
    return CT_OK;
}

int ct_batch_execute(ct_batch *b) {
    if (!b) return fail(CT_ERR_INVALID, "null argument");
    ct_batch &B = *b;
    CT_CHECK(check_gpu(B.gpu));
    hipStream_t s = B.gpu->stream;
    for (ct_epoch_kernel *k : B.kernels) {
        if (k->mhat->beats != k->r->B || k->mhat->steps != k->r->T)
            return fail(CT_ERR_INVALID, "measurements do not match a scenario of the batch");
        CT_CHECK(begin_epoch(*k));
    }
    if (B.voxels_per_group == 0 && B.slots > 0) CT_CHECK(autotune(B));
    if (!plan_current(B)) CT_CHECK(plan(B, B.voxels_per_group > 0 ? B.voxels_per_group
                                                                  : ct::kResMaxVoxels));
    CT_CHECK(clear_prop_events(B));
    for (int i = 0; i < 2; ++i)
        if (!B.ev[i]) CT_HIP(hipEventCreate(&B.ev[i]));
    // per-scenario phase profiling stays off inside a batch (one timeline)
    std::vector<bool> prof;
    for (ct_epoch_kernel *k : B.kernels) {
        prof.push_back(k->profiling);
        k->profiling = false;
    }
    int rc = CT_OK;
    const bool graph = B.use_graph && B.warmed;
    if (graph && !graph_current(B)) rc = capture(B);
    if (rc == CT_OK && hipEventRecord(B.ev[0], s) != hipSuccess) rc = fail(CT_ERR_DEVICE, "event");
    if (rc == CT_OK) {
        if (graph) {
            if (hipGraphLaunch(B.exec, s) != hipSuccess)
                rc = fail(CT_ERR_DEVICE, "hipGraphLaunch of the batched epoch failed");
            B.graph_epochs_left -= 1;
        } else {
            rc = enqueue_epoch(B, B.profiling);
            B.warmed = true;
        }
    }
    for (size_t i = 0; i < B.kernels.size(); ++i) B.kernels[i]->profiling = prof[i];
    CT_CHECK(rc);
    CT_HIP(hipEventRecord(B.ev[1], s));
    CT_CHECK(check_barriers(B, true));  // advances the members' epoch counters
    float ms = 0.0f;
    CT_HIP(hipEventElapsedTime(&ms, B.ev[0], B.ev[1]));
    B.last_epoch_ms = ms;
    B.last_propagate_ms = 0.0;
    for (auto &p : B.prop_events) {
        float pm = 0.0f;
        CT_HIP(hipEventElapsedTime(&pm, p.first, p.second));
        B.last_propagate_ms += pm;
    }
    B.last_propagate_launches = int64_t(B.prop_events.size());
    int64_t vs = 0;
    for (ct_epoch_kernel *k : B.kernels) vs += int64_t(k->r->N) * k->r->T * k->r->B;
    B.last_voxel_steps = vs;
    return CT_OK;
}

int ct_batch_get_info(ct_batch *b, ct_batch_info *out) {
    if (!b || !out) return fail(CT_ERR_INVALID, "null argument");
    std::memset(out, 0, sizeof(*out));
    out->scenarios = int32_t(b->kernels.size());
    out->rounds = int32_t(b->rounds.size());
    out->slots = b->slots;
    out->voxels_per_group = b->plan_vpg > 0 ? b->plan_vpg : b->voxels_per_group;
    out->jobs = b->num_jobs;
    for (const auto &R : b->rounds) out->predicted_propagate_ms += R.predicted_ms;
    out->epoch_ms = b->last_epoch_ms;
    out->propagate_ms = b->last_propagate_ms;
    out->propagate_launches = b->last_propagate_launches;
    out->voxel_steps = b->last_voxel_steps;
    out->graph = b->exec != nullptr ? 1 : 0;
    const int32_t nt = int32_t(std::min<size_t>(b->tuned.size(), 16));
    out->tuned_candidates = nt;
    for (int32_t i = 0; i < nt; ++i) {
        out->tuned_voxels[i] = b->tuned[i];
        out->tuned_ms[i] = b->tuned_ms[i];
    }
    out->side_streams = b->side_active ? int32_t(b->side.size()) : 0;
    out->sweep_jobs = b->num_sweep_jobs;
    return CT_OK;
}

}  // extern "C"
