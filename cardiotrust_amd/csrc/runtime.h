// Internal host-runtime interface shared by api.cpp (single scenario) and
// batch.cpp (many scenarios per launch): the opaque handle types of
// cardiotrust_hip.h and the epoch building blocks.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cardiotrust_hip.h"
#include "kernels.h"

namespace ctrt {
extern thread_local std::string g_error;
int fail(int code, const std::string &msg);
}  // namespace ctrt

#define CT_HIP(expr)                                                                  \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return ctrt::fail(CT_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define CT_CHECK(expr)          \
    do {                        \
        int rc_ = (expr);       \
        if (rc_ != CT_OK) return rc_; \
    } while (0)

// Reference-counted as the reference's GPU (gpu.rs: a Clone of the OpenCL
// context and queue that every buffer and kernel keeps alive): the creator
// holds one reference and every object built on it one more, so objects may
// be freed before or after ct_gpu_free.
struct ct_gpu {
    int32_t device = 0;
    hipStream_t stream = nullptr;
    uint64_t wall_ticks_per_second = 100000000;  // wall_clock64() rate
    std::atomic<int32_t> refs{1};
};

namespace ctrt {
// Drops one reference; the last one destroys the stream and the handle.
void gpu_release(ct_gpu *gpu);

// An object's reference to its ct_gpu: taken on assignment, dropped on
// destruction (after the owning object's destructor body has run).
class GpuRef {
  public:
    GpuRef() = default;
    GpuRef(const GpuRef &) = delete;
    GpuRef &operator=(const GpuRef &) = delete;
    ~GpuRef() { reset(nullptr); }
    GpuRef &operator=(ct_gpu *g) {
        reset(g);
        return *this;
    }
    ct_gpu *operator->() const { return g_; }
    ct_gpu &operator*() const { return *g_; }
    operator ct_gpu *() const { return g_; }

  private:
    void reset(ct_gpu *g) {
        if (g) g->refs.fetch_add(1, std::memory_order_relaxed);
        if (g_) gpu_release(g_);
        g_ = g;
    }
    ct_gpu *g_ = nullptr;
};
}  // namespace ctrt

struct ct_buffer {
    ctrt::GpuRef gpu;
    float *data = nullptr;
    int32_t beats = 0, steps = 0, sensors = 0;
};

struct ct_results_gpu {
    ctrt::GpuRef gpu;
    // the stream this scenario's work goes to: the GPU's, or one of a batch's
    // side streams while the batch runs the scenario's beats after a round's
    // propagation
    hipStream_t stream = nullptr;
    int32_t N = 0, S = 0, Sp = 0, T = 0, B = 0, nbatch = 0, optimizer = 0;
    int64_t n = 0;
    std::vector<void *> allocations;
    std::vector<float> control_function;  // host copy: u[t] is a launch argument

    // model
    int32_t *nbr = nullptr, *delays = nullptr;
    float *coefs = nullptr, *initial_delays = nullptr, *gT = nullptr, *b = nullptr, *H = nullptr;
    // estimations
    float *x = nullptr, *y = nullptr, *y_last = nullptr, *m = nullptr, *R = nullptr;
    float *xT = nullptr, *eT = nullptr;
    float *avg = nullptr;
    uint8_t *avg_valid = nullptr;
    // derivatives
    float *dGT = nullptr, *dA = nullptr, *fir = nullptr, *iir = nullptr, *mr_last = nullptr,
          *mr_sum = nullptr;
    float *gm1 = nullptr, *gm2 = nullptr, *cm1 = nullptr, *cm2 = nullptr;
    int64_t step = 1;
    // metrics
    float *loss = nullptr, *loss_mse = nullptr, *loss_mr = nullptr;
    float *loss_b = nullptr, *loss_mse_b = nullptr, *loss_mr_b = nullptr;
    float *u = nullptr;  // control function on the device (resident propagation)
    float *sensor_std = nullptr;  // (S) measurement noise standard deviations
    int64_t sensor_std_have = 0;
    // propagation path: resident (one launch per beat) when the
    // problem fits on chip, else one launch per step
    // 0 auto, 1 per-step launches, 2 resident, 3 resident exchanging the
    // state rows through agent-scope fences (what problems too large for the
    // coherent exchange take)
    int32_t propagation_mode = 0;
    double wait_seconds = 2.0;     // bound on every resident wait (ct_results_set_wait_limit)
    bool xT_fresh = false;         // xT holds the current states (written by the projection)
    // derivation path: over voxel tiles (derive_tiled_kernel) when their
    // tables were built, else a lane per connection (derive_kernel)
    int32_t derive_mode = 0;  // 0 automatic, 1 a lane per connection, 2 tiled
    ct::DeriveTiles tiles{};  // device tables; tiles.tiles == 0 when not built
    bool resident_ok = false;
    bool resident_used = false;  // a resident launch is in flight since the last check
    int32_t resident_blocks = 0;
    bool sweep_ok = false;  // the resident adjoint sweep fits with resident_blocks workgroups
    ct::GridBarrier *barrier = nullptr;
    // while a batch enqueues this scenario's epoch: the abort word of the
    // batch's propagation launches, which the scenario's updates also check
    const uint32_t *batch_abort = nullptr;
    // scratch
    float *partial = nullptr, *mr_partial = nullptr, *staging = nullptr, *step_mr = nullptr;
    double *objective = nullptr;  // J of the last forward (adjoint objective)
    // reverse-mode gradient: host neighbour table (for the consumer lists) and
    // buffers allocated on first use
    std::vector<int32_t> nbr_host;
    int32_t *cons_ptr = nullptr, *cons_idx = nullptr, *cons_delay = nullptr, *dmax = nullptr;
    int32_t max_consumers = 0;    // longest consumer list (26 on a grid)
    int64_t cons_len = 0;
    int32_t ckpt_interval = 0;  // checkpoint interval of the last forward pass (0 = none)
    float *E = nullptr, *xbar = nullptr, *xbarT = nullptr, *ybar = nullptr;
    // residuals of the beats of one update batch that shares a propagation
    // (reverse mode), one (T, Sp) slot per position of the batch
    float *Rbeats = nullptr;
    int64_t Rbeats_have = 0;
    // the latest mapped residuals are in E (by step, the reverse-mode forcing)
    // rather than eT
    bool mapped_by_step = false;
    float *qring = nullptr, *ckpt = nullptr, *yseg = nullptr;
    float *gcons = nullptr;  // (N, 26, 10) consumer-ordered gains and coefficients
    uint32_t *flags = nullptr;  // kMaxSlabFlags per-slab step counters (resident launches)
    int64_t qring_slots = 0, ckpt_segments = 0, yseg_have = 0;
    int64_t ring_slots_fixed = 0;  // > 0: q-ring size fixed for a captured epoch
    int64_t pending_slots = 0;     // q-ring size of the update batch being differentiated
    int64_t gcons_have = 0;
    // FP64 objective (gradient checking), allocated on first use
    double *y64 = nullptr, *x64 = nullptr, *p64 = nullptr, *terms64 = nullptr, *j64 = nullptr;
    int64_t f64_have = 0;
    int32_t forward_beat = -1;  // beat whose forward pass the device state holds
    ct::UpdateScalars *scal = nullptr;       // (B) device, per update of an epoch
    ct::UpdateScalars *scal_host = nullptr;  // (B) pinned host staging
    int32_t splits = 1, mr_blocks = 1;
    int64_t kps = ct::kProjTileK;

    ~ct_results_gpu() {
        for (void *p : allocations) (void)hipFree(p);
        if (scal_host) (void)hipHostFree(scal_host);
    }

    template <typename V>
    int alloc(V **p, int64_t count) {
        const size_t bytes = std::max<size_t>(sizeof(V) * size_t(std::max<int64_t>(count, 0)), 256);
        void *raw = nullptr;
        hipError_t e = hipMalloc(&raw, bytes);
        if (e != hipSuccess)
            return ctrt::fail(CT_ERR_DEVICE, "hipMalloc(" + std::to_string(bytes) +
                                           " bytes): " + hipGetErrorString(e));
        allocations.push_back(raw);
        *p = static_cast<V *>(raw);
        return CT_OK;
    }
};

struct ct_epoch_kernel {
    ctrt::GpuRef gpu;
    ct_results_gpu *r = nullptr;
    const ct_buffer *mhat = nullptr;
    ct_algorithm cfg{};
    std::vector<int32_t> order;
    int32_t batch_index = 0;
    bool profiling = false;
    int32_t gradient = 0;            // 0 = reference derivation, 1 = reverse-mode adjoint
    int32_t checkpoint_interval = 64;
    // beats of one update batch share one propagation (their states are
    // identical: same parameters and drive); with the reverse-mode gradient
    // also one backward pass for the batch's summed objective. -1 automatic
    // (shared under both gradients: bit-identical results under the reference
    // derivation, the same gradient up to FP32 rounding under reverse mode);
    // 0 never; 1 always.
    int32_t share_mode = -1;
    // set by a batch for its epoch: the backward pass due at the end of an
    // update batch (shared propagation) is left to the batch, which runs the
    // sweeps of a round's scenarios in one launch; deferred_beats > 0 marks
    // such a pending pass and the beats its forcing sums
    bool defer_backward = false;
    int32_t deferred_beats = 0;
    std::vector<hipEvent_t> events;
    size_t used_events = 0;
    struct SpanRec {
        int phase;
        int64_t steps;
        size_t start, end;
    };
    std::vector<SpanRec> spans;

    ~ct_epoch_kernel() {
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
    }
};

namespace ctrt {

enum Phase {
    kPropagate, kProject, kRegularize, kTranspose, kMapResiduals, kDerive, kUpdate, kTotal,
    kAdjointSweep
};

struct Span {
    ct_epoch_kernel *k;
    int phase;
    int64_t steps;
    size_t start = 0;
    bool on = false;
    Span(ct_epoch_kernel *kk, int ph, int64_t st = 0) : k(kk), phase(ph), steps(st) {
        if (!k || !k->profiling) return;
        if (k->used_events + 2 > k->events.size()) return;
        start = k->used_events;
        k->used_events += 2;
        on = hipEventRecord(k->events[start], k->r->stream) == hipSuccess;
    }
    ~Span() {
        if (!on) return;
        if (hipEventRecord(k->events[start + 1], k->r->stream) == hipSuccess)
            k->spans.push_back({phase, steps, start, start + 1});
    }
};

int check_gpu(ct_gpu *gpu);
int upload(void *dst, const void *src, size_t bytes);
int download(void *dst, const void *src, size_t bytes);
int reset_derivatives(ct_results_gpu &r);
int ensure_checkpoints(ct_results_gpu &r, int32_t K);
// Propagation of one beat on its own (resident or per-step launches).
int propagate_beat(ct_results_gpu &r, ct_epoch_kernel *k, int32_t ckpt_interval);
// Projection, residuals and per-step MSE of the beat whose states x holds.
int project_beat(ct_results_gpu &r, int32_t beat, const float *mhat, ct_epoch_kernel *k);
int forward_beat(ct_results_gpu &r, int32_t beat, const float *mhat, ct_epoch_kernel *k,
                 int32_t ckpt_interval = 0);
int losses_beat(ct_results_gpu &r, const ct_algorithm &c, ct_epoch_kernel *k);
int derive_reference(ct_results_gpu &r, int32_t beat, const ct_algorithm &c, ct_epoch_kernel *k);
int derive_adjoint(ct_results_gpu &r, int32_t beat, const ct_algorithm &c, int32_t K,
                   ct_epoch_kernel *k);
// Everything of the beat at position `pos` of the beat order after its
// propagation (run_epoch's per-beat body).
int finish_beat(ct_epoch_kernel &k, int32_t pos);
// Beats per update batch, and whether the beat at `pos` runs the propagation
// (always, unless the update batch shares one propagation).
int32_t update_group(const ct_epoch_kernel &k);
bool propagates_at(const ct_epoch_kernel &k, int32_t pos);
bool uses_adjoint(const ct_epoch_kernel &k);
// Whether the beats of one update batch share one propagation (share_mode).
bool shares_propagation(const ct_epoch_kernel &k);
int finish_batch(ct_epoch_kernel &k, int64_t num_beats, bool reset_after, int32_t update);
// Reverse-mode backward pass in parts, for a batch that sweeps many scenarios
// in one launch: the buffers and q-ring size (prepare_backward), the sweep's
// parameters for a forcing over `num_beats` beats, whether the resident
// sweep with the coherent exchange can run the problem, the inputs of the
// sweep (begin) and the gradient pass after it (end); and a deferred pass run
// whole on the scenario's own (adjoint_backward_deferred).
int prepare_backward(ct_results_gpu &r, int32_t K);
ct::AdjointParams adjoint_params(const ct_results_gpu &r, const ct_algorithm &c,
                                 int32_t num_beats);
bool sweep_batchable(const ct_results_gpu &r);
int adjoint_backward_begin(ct_results_gpu &r, ct_epoch_kernel *k);
int adjoint_backward_end(ct_results_gpu &r, int32_t K, ct_epoch_kernel *k);
int adjoint_backward_deferred(ct_epoch_kernel &k);
int32_t updates_per_epoch(const ct_epoch_kernel &k);
int begin_epoch(ct_epoch_kernel &k);
void end_epoch(ct_epoch_kernel &k);
ct::ResidentJob resident_job(ct_results_gpu &r, int32_t ckpt_interval);
int prepare_propagation(ct_results_gpu &r, int32_t ckpt_interval);
// After resident launches (or, with `always`, unconditionally): synchronises
// and reads the results' abort word. A timed-out wait disables the resident
// paths for the problem, clears the word (so that later updates, which check
// it, run again) and returns CT_ERR_DEVICE.
int check_resident(ct_results_gpu &r, bool always = false);
// Fixes the adjoint's q-ring size for graph capture: at least the current
// largest delay + 1 + `headroom` (delays grow by at most one per update).
// Reads the delays (synchronises); allocates the ring.
int fix_adjoint_ring(ct_results_gpu &r, int64_t headroom);
int begin_profile(ct_epoch_kernel &k);
// Seconds -> wall_clock64() ticks of the device.
uint64_t wait_ticks(const ct_gpu &g, double seconds);

}  // namespace ctrt
