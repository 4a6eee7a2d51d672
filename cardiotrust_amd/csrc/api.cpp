// Host runtime behind include/cardiotrust_hip.h.
//
// Owns device memory and the stream, validates the reference-layout inputs,
// converts them to the device layouts described in kernels.h, and drives the
// epoch in the order of the reference's CPU run_epoch
// (src/core/algorithm.rs:124-281): per beat, the forward propagation over all
// steps, then the deferred projection, residuals, losses and derivatives;
// per batch, the smoothness term, the update and the batch metrics.
#include <mutex>

#include "device.h"
#include "runtime.h"
#include "tiles.h"

namespace ctrt {

thread_local std::string g_error;

int fail(int code, const std::string &msg) {
    g_error = msg;
    return code;
}

void gpu_release(ct_gpu *gpu) {
    if (gpu->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
    (void)hipSetDevice(gpu->device);
    if (gpu->stream) (void)hipStreamDestroy(gpu->stream);
    delete gpu;
}

int check_gpu(ct_gpu *gpu) {
    if (!gpu) return fail(CT_ERR_INVALID, "null ct_gpu");
    CT_HIP(hipSetDevice(gpu->device));
    return CT_OK;
}

int upload(void *dst, const void *src, size_t bytes) {
    if (!bytes) return CT_OK;
    if (!src) return fail(CT_ERR_INVALID, "null input array");
    CT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return CT_OK;
}

int download(void *dst, const void *src, size_t bytes) {
    if (!dst || !bytes) return CT_OK;
    CT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return CT_OK;
}

// Checks the 26-neighbour structure of output_state_indices (allpass.rs:212-289)
// and extracts the neighbour table; checks delays >= 1 on connections.
int neighbours_from_indices(int32_t N, const int32_t *osi, const int32_t *delays,
                            std::vector<int32_t> &nbr) {
    nbr.assign(size_t(N) * 26, -1);
    for (int64_t v = 0; v < N; ++v) {
        for (int d = 0; d < 26; ++d) {
            const int32_t first = osi[(3 * v) * 78 + 3 * d];
            int32_t w = -1;
            if (first >= 0) {
                if (first % 3 != 0 || first / 3 >= N)
                    return fail(CT_ERR_INVALID, "output_state_indices[" + std::to_string(3 * v) +
                                                    "," + std::to_string(3 * d) +
                                                    "] is not 3*voxel of a valid voxel");
                w = first / 3;
            }
            for (int in = 0; in < 3; ++in)
                for (int o = 0; o < 3; ++o) {
                    const int32_t got = osi[(3 * v + in) * 78 + 3 * d + o];
                    const int32_t want = w < 0 ? -1 : 3 * w + o;
                    if (got != want && !(w < 0 && got < 0))
                        return fail(CT_ERR_INVALID,
                                    "output_state_indices[" + std::to_string(3 * v + in) + "," +
                                        std::to_string(3 * d + o) +
                                        "] breaks the 26-neighbour form (expected " +
                                        std::to_string(want) + ", got " + std::to_string(got) + ")");
                }
            const int32_t D = delays[v * 26 + d];
            if (D < 0)
                return fail(CT_ERR_INVALID, "negative delay at (" + std::to_string(v) + "," +
                                                std::to_string(d) + ")");
            if (w >= 0 && D < 1)
                return fail(CT_ERR_INVALID, "delay < 1 on connection (" + std::to_string(v) + "," +
                                                std::to_string(d) + ")");
            nbr[v * 26 + d] = w;
        }
    }
    return CT_OK;
}

int reset_derivatives(ct_results_gpu &r) {
    hipStream_t s = r.stream;
    // fill kernels, not memset nodes, so that the epoch can be captured in a graph
    CT_HIP(ct::launch_fill(s, r.dGT, 0.0f, r.n * 78));
    CT_HIP(ct::launch_fill(s, r.dA, 0.0f, int64_t(r.N) * 26));
    CT_HIP(ct::launch_fill(s, r.fir, 0.0f, int64_t(r.N) * 78));
    CT_HIP(ct::launch_fill(s, r.iir, 0.0f, int64_t(r.N) * 78));
    CT_HIP(ct::launch_fill(s, r.mr_last, 0.0f, r.n));
    CT_HIP(ct::launch_fill(s, r.mr_sum, 0.0f, 1));
    return CT_OK;
}

uint64_t wait_ticks(const ct_gpu &g, double seconds) {
    return uint64_t(seconds * double(g.wall_ticks_per_second));
}

// Clears y / y_last (estimation.rs:82-86: a beat starts from rest) and records
// the checkpoint interval of the coming propagation.
int prepare_propagation(ct_results_gpu &r, int32_t ckpt_interval) {
    hipStream_t s = r.stream;
    if (ckpt_interval > 0) CT_CHECK(ensure_checkpoints(r, ckpt_interval));
    r.ckpt_interval = ckpt_interval;
    r.xT_fresh = false;  // the states are about to change
    CT_HIP(ct::launch_fill(s, r.y, 0.0f, int64_t(r.N) * 78));
    CT_HIP(ct::launch_fill(s, r.y_last, 0.0f, int64_t(r.N) * 78));
    return CT_OK;
}

ct::ResidentJob resident_job(ct_results_gpu &r, int32_t ckpt_interval) {
    ct::ResidentJob j{};
    j.nbr = r.nbr;
    j.coefs = r.coefs;
    j.delays = r.delays;
    j.gT = r.gT;
    j.y = r.y;
    j.y_last = r.y_last;
    j.x = r.x;
    j.b = r.b;
    j.u = r.u;
    j.ckpt = r.ckpt;
    j.gb = r.barrier;
    j.flags = r.flags;
    j.N = r.N;
    j.T = r.T;
    j.ckpt_interval = ckpt_interval;
    j.slot0 = 0;
    j.width = r.resident_blocks;
    j.coherent = r.propagation_mode != 3 && ct::coherent_exchange(r.T, r.n) ? 1 : 0;
    return j;
}

// Forward propagation of one beat (prediction.rs:15-105 for every step). With
// ckpt_interval = K > 0 the pass also stores the filter outputs y_t at
// t = kK - 1 (k = 1 .. ceil(T/K) - 1) for the reverse-mode gradient.
int propagate_beat(ct_results_gpu &r, ct_epoch_kernel *k, int32_t ckpt_interval) {
    hipStream_t s = r.stream;
    CT_CHECK(prepare_propagation(r, ckpt_interval));
    const int64_t M = int64_t(r.N) * 78;
    const bool resident = r.resident_ok && r.propagation_mode != 1;
    if (resident && r.N > 0) {
        Span sp(k, kPropagate, r.T);
        CT_HIP(ct::launch_propagate_resident(s, r.resident_blocks, resident_job(r, ckpt_interval),
                                             0, r.T, wait_ticks(*r.gpu, r.wait_seconds)));
        r.resident_used = true;
    } else {
        for (int32_t t = 0; t < r.T; ++t) {
            Span sp(k, kPropagate, 1);
            CT_HIP(ct::launch_propagate_step(s, r.nbr, r.coefs, r.delays, r.gT, r.y, r.y_last,
                                             r.x, r.b, r.control_function[t], r.N, t, r.T));
            if (ckpt_interval > 0 && (t + 1) % ckpt_interval == 0 && t + 1 < r.T && M > 0)
                CT_HIP(hipMemcpyAsync(r.ckpt + int64_t((t + 1) / ckpt_interval - 1) * M, r.y,
                                      sizeof(float) * M, hipMemcpyDeviceToDevice, s));
        }
    }
    return CT_OK;
}

// The projected measurements of the beat (prediction.rs:112-130 for every
// step); with `mhat` also residuals and per-step MSE.
int project_beat(ct_results_gpu &r, int32_t beat, const float *mhat, ct_epoch_kernel *k) {
    hipStream_t s = r.stream;
    r.forward_beat = beat;
    Span sp(k, kProject);
    const float *Hb = r.H + int64_t(beat) * r.Sp * r.n;
    CT_HIP(ct::launch_project(s, r.x, Hb, r.partial, r.xT, r.T, r.n, r.S, r.Sp, r.splits, r.kps));
    r.xT_fresh = true;
    CT_HIP(ct::launch_residuals(s, r.partial, r.splits,
                                mhat ? mhat + int64_t(beat) * r.T * r.S : nullptr,
                                r.m + int64_t(beat) * r.T * r.S, r.R, r.loss_mse, r.T, r.S, r.Sp));
    return CT_OK;
}

int forward_beat(ct_results_gpu &r, int32_t beat, const float *mhat, ct_epoch_kernel *k,
                 int32_t ckpt_interval) {
    CT_CHECK(propagate_beat(r, k, ckpt_interval));
    return project_beat(r, beat, mhat, k);
}

// Per-step maximum regularisation sums, the running sum and the per-step
// losses (derivation.rs:509-535, metrics.rs:180-196), and the adjoint
// objective J of the beat.
int losses_beat(ct_results_gpu &r, const ct_algorithm &c, ct_epoch_kernel *k) {
    hipStream_t s = r.stream;
    Span sp(k, kRegularize);
    CT_HIP(ct::launch_maxreg_partials(s, r.x, r.mr_partial, r.mr_last, r.N, r.T,
                                      c.maximum_regularization_threshold, r.mr_blocks));
    CT_HIP(ct::launch_maxreg_finalize(s, r.mr_partial, r.mr_blocks, r.step_mr, r.mr_sum,
                                      r.loss_mse, r.loss_mr, r.loss, r.objective,
                                      c.mse_strength, c.maximum_regularization_strength, r.T));
    return CT_OK;
}

// The reference's per-step derivation over the whole beat (derivation.rs:173-229).
int derive_reference(ct_results_gpu &r, int32_t beat, const ct_algorithm &c, ct_epoch_kernel *k) {
    hipStream_t s = r.stream;
    const float *Hb = r.H + int64_t(beat) * r.Sp * r.n;
    {
        Span sp(k, kMapResiduals);
        CT_HIP(ct::launch_map_residuals(s, Hb, r.R, r.eT, r.n, r.T, r.S, r.Sp, false));
        r.mapped_by_step = false;
    }
    if (c.freeze_gains && c.freeze_delays) return CT_OK;
    if (!r.xT_fresh) {
        Span sp(k, kTranspose);
        CT_HIP(ct::launch_transpose(s, r.x, r.xT, r.T, r.n));
        r.xT_fresh = true;
    }
    Span sp(k, kDerive);
    ct::DeriveParams p;
    p.mse_scaling = 1.0f / float(r.S) * c.mse_strength;
    p.reg_scaling = c.maximum_regularization_strength;
    p.threshold = c.maximum_regularization_threshold;
    p.diff_strength = c.difference_regularization_strength;
    p.freeze_gains = c.freeze_gains;
    p.freeze_delays = c.freeze_delays;
    p.simple = c.ap_derivative == 1;
    if (r.tiles.tiles > 0 && r.derive_mode != 1)
        CT_HIP(ct::launch_derive_tiled(s, r.tiles, r.nbr, r.coefs, r.delays, r.initial_delays,
                                       r.gT, r.xT, r.eT, r.fir, r.iir, r.dGT, r.dA, r.N, r.T, p));
    else
        CT_HIP(ct::launch_derive(s, r.nbr, r.coefs, r.delays, r.initial_delays, r.gT, r.xT,
                                 r.eT, r.fir, r.iir, r.dGT, r.dA, r.N, r.T, p));
    return CT_OK;
}

bool uses_adjoint(const ct_epoch_kernel &k) {
    return k.gradient == 1 && !(k.cfg.freeze_gains && k.cfg.freeze_delays);
}

static int adjoint_prepare(ct_results_gpu &r, int32_t K, int64_t &slots);
template <typename V>
int ensure(ct_results_gpu &r, V **p, int64_t &have, int64_t count, int64_t unit);
static int adjoint_forcing(ct_results_gpu &r, int32_t beat, bool accumulate, ct_epoch_kernel *k);
static int adjoint_backward(ct_results_gpu &r, const ct_algorithm &c, int32_t K, int64_t slots,
                            int32_t num_beats, ct_epoch_kernel *k);
static bool resident_sweep(const ct_results_gpu &r, int64_t slots);

int32_t update_group(const ct_epoch_kernel &k) {
    return k.cfg.batch_size == 0 ? k.r->B : k.cfg.batch_size;
}

bool shares_propagation(const ct_epoch_kernel &k) { return k.share_mode != 0; }

bool propagates_at(const ct_epoch_kernel &k, int32_t pos) {
    return !shares_propagation(k) || pos % update_group(k) == 0;
}

int finish_beat(ct_epoch_kernel &k, int32_t pos) {
    ct_results_gpu &r = *k.r;
    const int32_t beat = k.order[pos];
    CT_CHECK(project_beat(r, beat, k.mhat->data, &k));
    CT_CHECK(losses_beat(r, k.cfg, &k));
    if (!uses_adjoint(k)) return derive_reference(r, beat, k.cfg, &k);
    if (!shares_propagation(k))
        return derive_adjoint(r, beat, k.cfg, k.checkpoint_interval, &k);
    // one backward pass per update batch, for the batch's summed objective
    if (r.N <= 0) return CT_OK;
    const int32_t group = update_group(k), first = pos - pos % group;
    const int32_t last = std::min<int32_t>(first + group, int32_t(k.order.size())) - 1;
    if (pos == first) {
        CT_CHECK(adjoint_prepare(r, k.checkpoint_interval, r.pending_slots));
        if (r.ckpt_interval != k.checkpoint_interval)
            return fail(CT_ERR_STATE, "shared propagation without the adjoint's checkpoints");
    }
    const int32_t count = last - first + 1;
    if (count > 1 && ct::map_residuals_beats_ok(r.S, r.T, r.n)) {
        // the batch's forcing in one pass over its beats: keep each beat's
        // residuals, sum the products in the MFMA accumulators at the last
        const int64_t slot = int64_t(r.T) * r.Sp;
        if (pos == first) CT_CHECK(ensure(r, &r.Rbeats, r.Rbeats_have, count * slot, 1));
        CT_HIP(ct::launch_copy(r.stream, r.Rbeats + (pos - first) * slot, r.R, slot));
        if (pos != last) return CT_OK;
        Span sp(&k, kMapResiduals);
        for (int32_t c0 = 0; c0 < count; c0 += ct::kMapMaxBeats) {
            ct::MapBeats mb{};
            mb.count = std::min<int32_t>(ct::kMapMaxBeats, count - c0);
            for (int32_t q = 0; q < mb.count; ++q)
                mb.h_off[q] = int64_t(k.order[first + c0 + q]) * r.Sp * r.n;
            CT_HIP(ct::launch_map_residuals_beats(r.stream, r.H, mb, r.Rbeats + c0 * slot, slot,
                                                  r.E, r.n, r.T, r.S, r.Sp, c0 > 0));
        }
        r.mapped_by_step = true;
    } else {
        CT_CHECK(adjoint_forcing(r, beat, pos != first, &k));
        if (pos != last) return CT_OK;
    }
    if (k.defer_backward) {
        // a batch runs the pass, with the sweeps of other scenarios in one launch
        k.deferred_beats = count;
        return CT_OK;
    }
    return adjoint_backward(r, k.cfg, k.checkpoint_interval, r.pending_slots, count, &k);
}

static int run_beat(ct_epoch_kernel &k, int32_t pos) {
    if (propagates_at(k, pos))
        CT_CHECK(propagate_beat(*k.r, &k, uses_adjoint(k) ? k.checkpoint_interval : 0));
    return finish_beat(k, pos);
}

// algorithm.rs:188-212 / 224-247 / 255-278: average delays, batch derivatives,
// parameter update (update.rs:27-97) and batch metrics. `update` numbers the
// updates within the epoch: its Adam step and batch-metric slot come from the
// device scalars written by begin_epoch.
int finish_batch(ct_epoch_kernel &k, int64_t num_beats, bool reset_after, int32_t update) {
    ct_results_gpu &r = *k.r;
    hipStream_t s = r.stream;
    const ct_algorithm &c = k.cfg;
    Span sp(&k, kUpdate);
    const ct::UpdateScalars *sc = r.scal + update;
    CT_HIP(ct::launch_average_delays(s, r.gT, r.coefs, r.delays, r.avg, r.avg_valid, r.N));
    if (!c.freeze_delays && std::fabs(c.smoothness_regularization_strength - 0.0f) > 1.1920929e-7f)
        CT_HIP(ct::launch_smoothness(s, r.nbr, r.avg, r.avg_valid, r.dA,
                                     c.smoothness_regularization_strength, r.N));
    const int64_t bs = c.batch_size == 0 ? int64_t(r.T) * num_beats : int64_t(r.T) * c.batch_size;
    const float scale = c.learning_rate / float(bs);
    // derivatives from a timed-out resident launch are not applied
    const ct::UpdateGuard guard{&r.barrier->abort.count, r.batch_abort};
    if (!c.freeze_gains) {
        if (c.optimizer == 1) {
            CT_HIP(ct::launch_adam(s, r.gT, r.dGT, r.gm1, r.gm2, r.n * 78, sc, scale, guard));
        } else {
            CT_HIP(ct::launch_sgd(s, r.gT, r.dGT, r.n * 78, scale, guard));
        }
    }
    if (!c.freeze_delays) {
        if (c.optimizer == 1) {
            CT_HIP(ct::launch_adam(s, r.coefs, r.dA, r.cm1, r.cm2, int64_t(r.N) * 26, sc, scale,
                                   guard));
        } else {
            CT_HIP(ct::launch_sgd(s, r.coefs, r.dA, int64_t(r.N) * 26, scale, guard));
        }
        CT_HIP(ct::launch_roll_delays(s, r.coefs, r.delays, int64_t(r.N) * 26, guard));
    }
    if (reset_after) CT_CHECK(reset_derivatives(r));
    CT_HIP(ct::launch_batch_metrics(s, r.loss, r.loss_mse, r.loss_mr, r.loss_b, r.loss_mse_b,
                                    r.loss_mr_b, sc, r.T));
    return CT_OK;
}

int32_t updates_per_epoch(const ct_epoch_kernel &k) {
    const int32_t B = k.r->B, bs = k.cfg.batch_size;
    return bs == 0 ? 1 : (B + bs - 1) / bs;
}

// Checks the epoch's preconditions and writes its update scalars (Adam step
// and batch-metric slot of every update) to the device. Waits for the stream
// first, so the previous epoch no longer reads them.
int begin_epoch(ct_epoch_kernel &k) {
    ct_results_gpu &r = *k.r;
    const ct_algorithm &c = k.cfg;
    const int32_t updates = updates_per_epoch(k);
    if (k.batch_index < 0 || int64_t(k.batch_index) + updates > r.nbatch)
        return fail(CT_ERR_STATE, "batch index " + std::to_string(k.batch_index + updates - 1) +
                                      " outside the batch metrics (" + std::to_string(r.nbatch) +
                                      " entries)");
    if (c.optimizer == 1 && ((!c.freeze_gains && !r.gm1) || (!c.freeze_delays && !r.cm1)))
        return fail(CT_ERR_STATE, "Adam update without Adam moments (results were uploaded with "
                                  "optimizer = Sgd)");
    CT_HIP(hipStreamSynchronize(r.stream));
    for (int32_t u = 0; u < updates; ++u) {
        const float step = float(r.step + u);
        r.scal_host[u].c1 = 1.0f - std::pow(0.9f, step);
        r.scal_host[u].c2 = 1.0f - std::pow(0.999f, step);
        r.scal_host[u].batch_index = k.batch_index + u;
        r.scal_host[u].pad = 0;
    }
    CT_HIP(hipMemcpyAsync(r.scal, r.scal_host, sizeof(ct::UpdateScalars) * updates,
                          hipMemcpyHostToDevice, r.stream));
    return CT_OK;
}

// Host mirror of the device counters after an epoch's updates were enqueued.
void end_epoch(ct_epoch_kernel &k) {
    const int32_t updates = updates_per_epoch(k);
    k.r->step += updates;
    k.batch_index += updates;
}

// Decides whether the resident (on-chip filter state) kernel can run this
// problem: at most kResMaxVoxels voxels per workgroup with every workgroup
// resident on the device at once, neighbour indices below 2^17 - 1, delays
// and step offsets that fit its 15-bit / 32-bit fields.
// Voxel tiles of the tiled derivation, built once per problem from the host
// neighbour table. Without them (no voxels, or the kernels' LDS limit cannot
// be set) the derivation takes a lane per connection.
int plan_tiles(ct_results_gpu &r) {
    static std::once_flag prepared;
    static hipError_t prepare_status = hipSuccess;
    std::call_once(prepared, [] {
        prepare_status = ct::derive_tiled_prepare();
        if (prepare_status == hipSuccess) prepare_status = ct::adjoint_gradient_tiled_prepare();
    });
    if (r.N <= 0 || prepare_status != hipSuccess) return CT_OK;
    ctrt::VoxelTiles t;
    ctrt::build_voxel_tiles(r.nbr_host, r.N, ct::kTileMaxVoxels, ct::kTileMaxHalo, ct::kNoSlot, t);
    ctrt::order_tiles_for_dispatch(t, ct::kNumXcds);
    const int32_t G = t.tiles();
    if (G <= 0) return CT_OK;
    int32_t *vox_ptr = nullptr, *vox = nullptr, *halo_ptr = nullptr, *halo = nullptr;
    uint8_t *slot = nullptr;
    CT_CHECK(r.alloc(&vox_ptr, G + 1));
    CT_CHECK(r.alloc(&vox, int64_t(t.vox.size())));
    CT_CHECK(r.alloc(&halo_ptr, G + 1));
    CT_CHECK(r.alloc(&halo, int64_t(t.halo.size())));
    CT_CHECK(r.alloc(&slot, int64_t(t.slot.size())));
    CT_CHECK(upload(vox_ptr, t.vox_ptr.data(), sizeof(int32_t) * t.vox_ptr.size()));
    CT_CHECK(upload(vox, t.vox.data(), sizeof(int32_t) * t.vox.size()));
    CT_CHECK(upload(halo_ptr, t.halo_ptr.data(), sizeof(int32_t) * t.halo_ptr.size()));
    CT_CHECK(upload(halo, t.halo.data(), sizeof(int32_t) * t.halo.size()));
    CT_CHECK(upload(slot, t.slot.data(), t.slot.size()));
    r.tiles = ct::DeriveTiles{vox_ptr, vox, halo_ptr, halo, slot, G};
    return CT_OK;
}

int plan_resident(ct_results_gpu &r) {
    r.resident_ok = false;
    r.resident_blocks = 0;
    if (r.N <= 0) return CT_OK;
    int cus = 0;
    CT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, r.gpu->device));
    if (cus <= 0) return CT_OK;
    int32_t per_cu = 0;
    if (ct::resident_capacity(&per_cu) != hipSuccess || per_cu < 1) {
        (void)hipGetLastError();
        return CT_OK;
    }
    const int64_t capacity = int64_t(cus) * per_cu;
    const int64_t blocks = std::min<int64_t>(capacity, std::max<int64_t>(1, (r.N + 127) / 128));
    if ((r.N + blocks - 1) / blocks > ct::kResMaxVoxels || blocks > ct::kMaxSlabFlags) return CT_OK;
    if (r.N >= (1 << 17) - 1 || r.T > 32767 || int64_t(r.T) * r.n >= (int64_t(1) << 32))
        return CT_OK;
    r.resident_blocks = int32_t(blocks);
    r.resident_ok = true;
    int32_t sweep_per_cu = 0;
    r.sweep_ok = ct::sweep_capacity(&sweep_per_cu) == hipSuccess &&
                 blocks <= int64_t(cus) * sweep_per_cu;
    (void)hipGetLastError();
    return CT_OK;
}

// After resident launches: waits for the stream and reports a timed-out wait
// (workgroups not all resident, or a wait limit too short), which leaves the
// results of that launch invalid; the updates queued behind it checked the
// abort word and left the parameters unchanged. The resident forward and
// backward paths are then disabled for this problem, and the word is cleared
// (nothing waits on it any more, and the next updates must run).
int check_resident(ct_results_gpu &r, bool always) {
    if (!r.resident_used && !always) return CT_OK;
    r.resident_used = false;
    uint32_t abort = 0;
    CT_HIP(hipStreamSynchronize(r.stream));
    CT_HIP(hipMemcpy(&abort, &r.barrier->abort.count, sizeof(abort), hipMemcpyDeviceToHost));
    if (abort) {
        r.resident_ok = false;
        r.sweep_ok = false;
        const uint32_t zero = 0;
        CT_HIP(hipMemcpy(&r.barrier->abort.count, &zero, sizeof(zero), hipMemcpyHostToDevice));
        return fail(CT_ERR_DEVICE, "resident propagation: a wait timed out (workgroups not all "
                                   "resident, or the wait limit too short); the epoch's update "
                                   "was not applied, and the per-step path is used from now on");
    }
    return CT_OK;
}

// Grows a lazily allocated adjoint buffer to `count` elements.
template <typename V>
int ensure(ct_results_gpu &r, V **p, int64_t &have, int64_t count, int64_t unit) {
    if (*p && have >= count) return CT_OK;
    if (*p) {
        auto it = std::find(r.allocations.begin(), r.allocations.end(), static_cast<void *>(*p));
        if (it != r.allocations.end()) r.allocations.erase(it);
        CT_HIP(hipStreamSynchronize(r.stream));
        CT_HIP(hipFree(*p));
        *p = nullptr;
    }
    CT_CHECK(r.alloc(p, std::max<int64_t>(count, 1) * unit));
    have = count;
    return CT_OK;
}

int ensure_checkpoints(ct_results_gpu &r, int32_t K) {
    const int64_t M = int64_t(r.N) * 78;
    const int64_t nseg = (int64_t(r.T) + K - 1) / K;
    const int64_t need = std::max<int64_t>(nseg - 1, 1);
    int64_t have = r.ckpt_segments * M;
    CT_CHECK(ensure(r, &r.ckpt, have, need * M, 1));
    r.ckpt_segments = std::max(r.ckpt_segments, need);
    return CT_OK;
}

int fix_adjoint_ring(ct_results_gpu &r, int64_t headroom) {
    hipStream_t s = r.stream;
    const int64_t M = int64_t(r.N) * 78;
    CT_HIP(ct::launch_max_delay(s, r.nbr, r.delays, int64_t(r.N) * 26, r.dmax));
    int32_t dmax = 0;
    CT_HIP(hipMemcpyAsync(&dmax, r.dmax, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    CT_HIP(hipStreamSynchronize(s));
    int64_t slots = 1;
    while (slots < int64_t(dmax) + 1 + headroom) slots <<= 1;
    int64_t have_q = r.qring_slots * M;
    CT_CHECK(ensure(r, &r.qring, have_q, slots * M, 1));
    r.qring_slots = std::max(r.qring_slots, slots);
    r.ring_slots_fixed = slots;
    return CT_OK;
}

// Consumer lists (CSR): for each voxel w, the connections (v, d) that read
// it, in ascending connection order.
int build_consumers(ct_results_gpu &r) {
    if (r.cons_ptr) return CT_OK;
    const int64_t N = r.N;
    std::vector<int32_t> ptr(N + 1, 0), idx;
    for (int64_t c = 0; c < N * 26; ++c)
        if (r.nbr_host[c] >= 0) ptr[r.nbr_host[c] + 1] += 1;
    for (int64_t w = 0; w < N; ++w) ptr[w + 1] += ptr[w];
    idx.resize(ptr[N]);
    std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
    for (int64_t c = 0; c < N * 26; ++c)
        if (r.nbr_host[c] >= 0) idx[fill[r.nbr_host[c]]++] = int32_t(c);
    CT_CHECK(r.alloc(&r.cons_ptr, N + 1));
    CT_CHECK(r.alloc(&r.cons_idx, std::max<int64_t>(int64_t(idx.size()), 1)));
    CT_CHECK(r.alloc(&r.cons_delay, std::max<int64_t>(int64_t(idx.size()), 1)));
    r.cons_len = int64_t(idx.size());
    r.max_consumers = 0;
    for (int64_t w = 0; w < N; ++w) r.max_consumers = std::max(r.max_consumers, ptr[w + 1] - ptr[w]);
    CT_CHECK(upload(r.cons_ptr, ptr.data(), sizeof(int32_t) * (N + 1)));
    CT_CHECK(upload(r.cons_idx, idx.data(), sizeof(int32_t) * idx.size()));
    return CT_OK;
}

// Reverse-mode gradient of the forward pass held on the device (adjoint.hip);
// adds dJ/dG and dJ/da of this beat to the derivative accumulators.
// The one-launch backward sweep runs when the resident propagation does and
// its tables fit: delays below 1024 (ring of `slots` >= Dmax + 1 entries)
// and at most 26 consumers per voxel.
static bool resident_sweep(const ct_results_gpu &r, int64_t slots) {
    return r.sweep_ok && r.propagation_mode != 1 && slots <= 1024 && r.max_consumers <= 26;
}

// Buffers and q-ring size of the reverse-mode gradient (checkpoints every K).
static int adjoint_prepare(ct_results_gpu &r, int32_t K, int64_t &slots) {
    if (K < 1) return fail(CT_ERR_INVALID, "checkpoint interval must be >= 1");
    hipStream_t s = r.stream;
    const int64_t n = r.n, T = r.T, M = int64_t(r.N) * 78;
    CT_CHECK(build_consumers(r));
    int64_t have_e = r.E ? n * T : 0, have_xb = r.xbar ? n * T : 0, have_yb = r.ybar ? M : 0;
    CT_CHECK(ensure(r, &r.E, have_e, n * T, 1));
    CT_CHECK(ensure(r, &r.xbar, have_xb, n * T, 1));
    CT_CHECK(ensure(r, &r.ybar, have_yb, M, 1));  // (x̄ᵀ only when the gradient pass needs it)
    // ring of q over R >= Dmax + 1 slots (R a power of two)
    slots = r.ring_slots_fixed;
    if (slots <= 0) {
        CT_HIP(ct::launch_max_delay(s, r.nbr, r.delays, int64_t(r.N) * 26, r.dmax));
        int32_t dmax = 0;
        CT_HIP(hipMemcpyAsync(&dmax, r.dmax, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        CT_HIP(hipStreamSynchronize(s));
        slots = 1;
        while (slots < int64_t(dmax) + 1) slots <<= 1;
    }
    if (resident_sweep(r, slots)) {
        CT_CHECK(ensure(r, &r.gcons, r.gcons_have, int64_t(r.N) * 26 * 10, 1));
    } else {
        int64_t have_q = r.qring_slots * M;
        CT_CHECK(ensure(r, &r.qring, have_q, slots * M, 1));
        r.qring_slots = std::max(r.qring_slots, slots);
    }
    // gradient pass: filter state (6 values) per connection at each chunk
    // start of a segment
    CT_CHECK(ensure(r, &r.yseg, r.yseg_have, ct::adjoint_gradient_state_len(r.N, r.T, K), 1));
    return CT_OK;
}

// Hᵀ r of `beat` for every step, by step into E (T, n) where the kernel
// allows it (the sweep's layout), else into eT (n, T); with `accumulate`
// added to what it holds (beats of one update batch, in beat order).
static int adjoint_forcing(ct_results_gpu &r, int32_t beat, bool accumulate, ct_epoch_kernel *k) {
    hipStream_t s = r.stream;
    const float *Hb = r.H + int64_t(beat) * r.Sp * r.n;
    Span sp(k, kMapResiduals);
    const bool by_step = ct::map_residuals_by_step_ok(r.S, r.T, r.n);
    CT_HIP(ct::launch_map_residuals(s, Hb, r.R, by_step ? r.E : r.eT, r.n, r.T, r.S, r.Sp,
                                    accumulate, by_step));
    r.mapped_by_step = by_step;
    return CT_OK;
}

// Parameters of the backward sweep for a forcing that sums `num_beats` beats
// sharing the forward states (each beat contributes the maximum-regularisation
// term once).
ct::AdjointParams adjoint_params(const ct_results_gpu &r, const ct_algorithm &c,
                                 int32_t num_beats) {
    ct::AdjointParams p;
    p.mse_scaling = 1.0f / float(r.S) * c.mse_strength;
    p.reg_scaling = c.maximum_regularization_strength * float(num_beats);
    p.threshold = c.maximum_regularization_threshold;
    return p;
}

bool sweep_batchable(const ct_results_gpu &r) {
    return resident_sweep(r, r.pending_slots) && r.propagation_mode != 3 &&
           ct::coherent_exchange(r.T, r.n);
}

int prepare_backward(ct_results_gpu &r, int32_t K) {
    return adjoint_prepare(r, K, r.pending_slots);
}

// Backward pass, part 1: the forcing by step (transposed from eT if it was
// not written so) and the gains by state row, the sweep's inputs.
int adjoint_backward_begin(ct_results_gpu &r, ct_epoch_kernel *k) {
    hipStream_t s = r.stream;
    Span sp(k, kTranspose);
    if (!r.mapped_by_step)
        CT_HIP(ct::launch_transpose(s, r.eT, r.E, r.n, r.T));  // mapped residuals by step
    CT_HIP(ct::launch_transpose(s, r.gT, r.staging, 78, r.n));  // gains by state row
    return CT_OK;
}

// Part 2 on the scenario's own: the resident sweep (all steps in one launch,
// recurrences owned by the consuming voxels) or the per-step launches.
static int adjoint_sweep_own(ct_results_gpu &r, const ct::AdjointParams &p, int64_t slots,
                             ct_epoch_kernel *k) {
    hipStream_t s = r.stream;
    const int64_t M = int64_t(r.N) * 78;
    Span sp(k, kAdjointSweep);
    if (resident_sweep(r, slots)) {
        CT_HIP(ct::launch_adjoint_sweep_resident(s, r.resident_blocks, r.cons_ptr, r.cons_idx,
                                                 r.delays, r.staging, r.coefs, r.gcons, r.x,
                                                 r.E, r.xbar, r.barrier, r.N, r.T, p,
                                                 wait_ticks(*r.gpu, r.wait_seconds),
                                                 r.propagation_mode != 3));
        r.resident_used = true;
    } else {
        CT_HIP(ct::launch_consumer_delays(s, r.cons_idx, r.delays, r.cons_delay, r.cons_len));
        CT_HIP(ct::launch_fill(s, r.ybar, 0.0f, M));
        for (int32_t step = r.T - 1; step >= 0; --step)
            CT_HIP(ct::launch_adjoint_step(s, r.nbr, r.coefs, r.staging, r.cons_ptr, r.cons_idx,
                                           r.cons_delay, r.x, r.E, r.ybar, r.qring, r.xbar, r.N,
                                           r.T, step, uint32_t(slots - 1), p));
    }
    return CT_OK;
}

// Part 3: the gradient pass over x̄ (checkpoints every K steps).
int adjoint_backward_end(ct_results_gpu &r, int32_t K, ct_epoch_kernel *k) {
    hipStream_t s = r.stream;
    const int64_t n = r.n, T = r.T;
    Span sp(k, kDerive);
    if (!r.xT_fresh) {
        CT_HIP(ct::launch_transpose(s, r.x, r.xT, T, n));
        r.xT_fresh = true;
    }
    if (K == ct::kTileChunk && r.tiles.tiles > 0 && r.derive_mode != 1) {
        // reads x̄ by step: no transpose
        CT_HIP(ct::launch_adjoint_gradient_tiled(s, r.tiles, r.nbr, r.coefs, r.delays, r.gT, r.xT,
                                                 r.xbar, r.ckpt, r.dGT, r.dA, r.N, r.T));
    } else {
        int64_t have_xbt = r.xbarT ? n * T : 0;
        CT_CHECK(ensure(r, &r.xbarT, have_xbt, n * T, 1));
        CT_HIP(ct::launch_transpose(s, r.xbar, r.xbarT, T, n));
        CT_HIP(ct::launch_adjoint_gradient(s, r.nbr, r.coefs, r.delays, r.gT, r.xT, r.xbarT,
                                           r.ckpt, r.yseg, r.yseg_have, r.dGT, r.dA, r.N, r.T, K));
    }
    return CT_OK;
}

// Backward sweep and gradient pass for the forcing in E (or eT), which sums
// `num_beats` beats that share the forward states.
static int adjoint_backward(ct_results_gpu &r, const ct_algorithm &c, int32_t K, int64_t slots,
                            int32_t num_beats, ct_epoch_kernel *k) {
    CT_CHECK(adjoint_backward_begin(r, k));
    CT_CHECK(adjoint_sweep_own(r, adjoint_params(r, c, num_beats), slots, k));
    return adjoint_backward_end(r, K, k);
}

int adjoint_backward_deferred(ct_epoch_kernel &k) {
    const int32_t beats = k.deferred_beats;
    k.deferred_beats = 0;
    if (beats <= 0) return CT_OK;
    return adjoint_backward(*k.r, k.cfg, k.checkpoint_interval, k.r->pending_slots, beats, &k);
}

int derive_adjoint(ct_results_gpu &r, int32_t beat, const ct_algorithm &c, int32_t K,
                   ct_epoch_kernel *k) {
    if (r.N <= 0) return CT_OK;
    int64_t slots = 0;
    CT_CHECK(adjoint_prepare(r, K, slots));
    if (r.ckpt_interval != K || r.forward_beat != beat)
        return fail(CT_ERR_STATE, "the reverse-mode gradient needs the forward pass of this beat "
                                  "with checkpoints every " + std::to_string(K) + " steps");
    CT_CHECK(adjoint_forcing(r, beat, false, k));
    return adjoint_backward(r, c, K, slots, 1, k);
}

// ---------------------------------------------------------------------------
// Device scratch of ct_results_calculate_final_metrics_noalloc: a fixed array
// in the library's code object (each device holds its own copy, from the
// first use of the library on that device until the process ends), so the
// entry point needs no allocation. Laid out per call as the histogram
// (2 x 102 words), each voxel's maximum (N floats) and the staged cell
// arrays (numbers, types: `cells` words each), every part 256-byte aligned.
// Calls that overlap on one device share it and race.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = size_t(8) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch_address[kMaxDevices] = {};  // device address of `scratch`, per device

// The device address of `scratch` on `device`, which must be current. In host
// code the array's name stands for a host-side shadow, never for device
// memory: kernels, copies and fills get this address.
int scratch_address(int32_t device, void **out) {
    if (device < 0 || device >= kMaxDevices)
        return fail(CT_ERR_DEVICE, "device " + std::to_string(device) + " beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_address[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(CT_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") +
                                           hipGetErrorString(e));
        g_scratch_address[device] = p;
    }
    *out = g_scratch_address[device];
    return CT_OK;
}

// Argument checks shared by the two final-metrics entry points.
int final_metrics_check(const ct_results_gpu *results, const uint32_t *voxel_types,
                        const int32_t *voxel_numbers, int64_t cells) {
    if (!results || (cells > 0 && (!voxel_types || !voxel_numbers)))
        return fail(CT_ERR_INVALID, "null argument");
    if (cells < 0 || cells > INT32_MAX)
        return fail(CT_ERR_INVALID, "cell count must be in [0, 2^31 - 1]");
    // predict_voxeltype indexes the states at number, +1, +2 (metrics.rs:429-433)
    const int64_t n = 3LL * results->N;
    for (int64_t c = 0; c < cells; ++c) {
        const int32_t v = voxel_numbers[c];
        if (v < -1 || (v >= 0 && (v % 3 != 0 || v >= n)))
            return fail(CT_ERR_INVALID, "voxel number " + std::to_string(v) + " of cell " +
                                            std::to_string(c) +
                                            " is neither -1 nor a multiple of 3 below 3N");
    }
    return CT_OK;
}

constexpr int kFinalBins = ct::kSegmentationThresholds + 1;

// The device part of the final metrics: stages the cell arrays into d_num /
// d_types, runs the maxima and the histograms, and copies the histograms to
// `hist` (2 * kFinalBins words). Blocks.
int final_metrics_device(ct_results_gpu &r, const uint32_t *voxel_types,
                         const int32_t *voxel_numbers, int64_t cells, uint32_t *d_hist,
                         float *d_max, int32_t *d_num, uint32_t *d_types, uint32_t *hist) {
    ct::SegmentationThresholds th;
    for (int i = 0; i < ct::kSegmentationThresholds; ++i)
        th.value[i] = float(i) / 100.0f;  // metrics.rs:233
    hipStream_t s = r.stream;
    int rc = CT_OK;
    auto step = [&](hipError_t e) {
        if (rc == CT_OK && e != hipSuccess)
            rc = fail(CT_ERR_DEVICE, std::string("final metrics: ") + hipGetErrorString(e));
    };
    if (cells > 0) {
        step(hipMemcpyAsync(d_num, voxel_numbers, sizeof(int32_t) * cells, hipMemcpyHostToDevice, s));
        step(hipMemcpyAsync(d_types, voxel_types, sizeof(uint32_t) * cells, hipMemcpyHostToDevice,
                            s));
    }
    if (rc == CT_OK)
        step(ct::launch_segmentation_histogram(s, r.x, r.T, r.N, d_num, d_types, cells, th, d_max,
                                               d_hist));
    if (rc == CT_OK)
        step(hipMemcpyAsync(hist, d_hist, sizeof(uint32_t) * 2 * kFinalBins, hipMemcpyDeviceToHost,
                            s));
    step(hipStreamSynchronize(s));
    return rc;
}

// The counts of calculate_dice / _iou / _precision / _recall
// (metrics.rs:253-404) at each threshold, from the histograms' prefix sums:
// a cell predicted Pathological at threshold i has bin <= i.
void final_metrics_from_histograms(const uint32_t *hist, float *dice, float *iou, float *precision,
                                   float *recall) {
    int64_t gt_pos = 0;
    for (int k = 0; k < kFinalBins; ++k) gt_pos += hist[kFinalBins + k];
    int64_t tp = 0, pred_pos = 0;
    for (int i = 0; i < ct::kSegmentationThresholds; ++i) {
        tp += hist[kFinalBins + i];
        pred_pos += int64_t(hist[i]) + hist[kFinalBins + i];
        const int64_t fp = pred_pos - tp, fn = gt_pos - tp;
        const int64_t den = 2 * tp + fp + fn, uni = tp + fp + fn;
        if (dice) dice[i] = den == 0 ? 0.0f : float(2 * tp) / float(den);
        if (iou) iou[i] = uni == 0 ? 0.0f : float(tp) / float(uni);
        if (precision) precision[i] = pred_pos == 0 ? 0.0f : float(tp) / float(pred_pos);
        if (recall) recall[i] = gt_pos == 0 ? 1.0f : float(tp) / float(gt_pos);
    }
}

}  // namespace ctrt

using namespace ctrt;

extern "C" {

const char *ct_last_error(void) { return g_error.c_str(); }

int ct_gpu_new(int32_t device, ct_gpu **out) {
    if (!out) return fail(CT_ERR_INVALID, "null output handle");
    int count = 0;
    CT_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count)
        return fail(CT_ERR_DEVICE, "no HIP device " + std::to_string(device) + " (" +
                                       std::to_string(count) + " present)");
    CT_HIP(hipSetDevice(device));
    void *scratch_ptr = nullptr;  // resolved once per device, with the device current
    CT_CHECK(scratch_address(device, &scratch_ptr));
    ct_gpu *g = new ct_gpu;
    g->device = device;
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess &&
        khz > 0)
        g->wall_ticks_per_second = uint64_t(khz) * 1000u;
    hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete g;
        return fail(CT_ERR_DEVICE, std::string("hipStreamCreateWithFlags: ") + hipGetErrorString(e));
    }
    *out = g;
    return CT_OK;
}

void ct_gpu_free(ct_gpu *gpu) {
    if (gpu) gpu_release(gpu);
}

int ct_gpu_synchronize(ct_gpu *gpu) {
    CT_CHECK(check_gpu(gpu));
    CT_HIP(hipStreamSynchronize(gpu->stream));
    return CT_OK;
}

int ct_measurements_to_gpu(ct_gpu *gpu, const float *measurements, int32_t num_beats,
                           int32_t num_steps, int32_t num_sensors, ct_buffer **out) {
    CT_CHECK(check_gpu(gpu));
    if (!out || !measurements) return fail(CT_ERR_INVALID, "null argument");
    if (num_beats <= 0 || num_steps <= 0 || num_sensors <= 0)
        return fail(CT_ERR_INVALID, "measurements need positive beats, steps and sensors");
    const size_t bytes = sizeof(float) * size_t(num_beats) * num_steps * num_sensors;
    ct_buffer *b = new ct_buffer;
    b->gpu = gpu;
    b->beats = num_beats;
    b->steps = num_steps;
    b->sensors = num_sensors;
    hipError_t e = hipMalloc(&b->data, bytes);
    if (e == hipSuccess) e = hipMemcpy(b->data, measurements, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (b->data) (void)hipFree(b->data);
        delete b;
        return fail(CT_ERR_DEVICE, std::string("measurements upload: ") + hipGetErrorString(e));
    }
    *out = b;
    return CT_OK;
}

void ct_buffer_free(ct_buffer *buffer) {
    if (!buffer) return;
    if (buffer->data) (void)hipFree(buffer->data);
    delete buffer;
}

int ct_results_to_gpu(ct_gpu *gpu, const ct_results_desc *desc, ct_results_gpu **out) {
    CT_CHECK(check_gpu(gpu));
    if (!desc || !out) return fail(CT_ERR_INVALID, "null argument");
    const ct_results_desc &d = *desc;
    if (d.num_voxels < 0 || d.num_sensors <= 0 || d.num_steps <= 0 || d.num_beats <= 0 ||
        d.num_batch_entries <= 0)
        return fail(CT_ERR_INVALID, "need N >= 0 and positive sensors, steps, beats and batch "
                                    "entries");
    if (d.num_steps > 65535)
        return fail(CT_ERR_INVALID, "at most 65535 steps per beat are supported");
    if (int64_t(d.num_voxels) * 3 > 0x7fffffffLL)
        return fail(CT_ERR_INVALID, "too many voxels for 32-bit state indices");
    const int32_t N = d.num_voxels;
    std::vector<int32_t> nbr;
    if (N > 0) {
        if (!d.output_state_indices || !d.delays || !d.gains || !d.coefs || !d.initial_delays ||
            !d.control_matrix)
            return fail(CT_ERR_INVALID, "null parameter array");
        CT_CHECK(neighbours_from_indices(N, d.output_state_indices, d.delays, nbr));
    }
    if (!d.control_function || !d.measurement_matrix)
        return fail(CT_ERR_INVALID, "null control function or measurement matrix");

    ct_results_gpu *r = new ct_results_gpu;
    auto cleanup = [&](int rc) {
        delete r;
        return rc;
    };
    r->gpu = gpu;
    r->stream = gpu->stream;
    r->N = N;
    r->n = 3LL * N;
    r->S = d.num_sensors;
    r->Sp = (d.num_sensors + 127) / 128 * 128;
    r->T = d.num_steps;
    r->B = d.num_beats;
    r->nbatch = d.num_batch_entries;
    r->optimizer = d.optimizer;
    r->control_function.assign(d.control_function, d.control_function + d.num_steps);
    r->nbr_host = nbr;
    const int64_t n = r->n, T = r->T, S = r->S, Sp = r->Sp, B = r->B;

    // GEMM split-K plan: about 2048 workgroups for the projection, each
    // split at least kMinSplitStates states long. (A small scenario split
    // down to single 32-state k tiles writes partial sums larger than its
    // states: 93 splits x T x 128 floats for 3000 states.)
    constexpr int64_t kMinSplitStates = 512;
    const int64_t tiles = (T + ct::kProjTile - 1) / ct::kProjTile * ct::project_sensor_tiles(int32_t(S));
    int64_t splits = std::max<int64_t>(
        1, std::min<int64_t>((n + kMinSplitStates - 1) / kMinSplitStates,
                             2048 / std::max<int64_t>(tiles, 1)));
    int64_t kps = ((n + splits - 1) / splits + ct::kProjTileK - 1) / ct::kProjTileK * ct::kProjTileK;
    kps = std::max<int64_t>(kps, ct::kProjTileK);
    splits = std::max<int64_t>(1, (n + kps - 1) / kps);
    r->splits = int32_t(splits);
    r->kps = kps;
    r->mr_blocks = std::max<int32_t>(1, (N + ct::kMaxregVoxels - 1) / ct::kMaxregVoxels);

    int rc = CT_OK;
#define ALLOC(ptr, count)                                  \
    do {                                                   \
        if ((rc = r->alloc(&(ptr), (count))) != CT_OK) return cleanup(rc); \
    } while (0)
    ALLOC(r->nbr, int64_t(N) * 26);
    ALLOC(r->delays, int64_t(N) * 26);
    ALLOC(r->coefs, int64_t(N) * 26);
    ALLOC(r->initial_delays, int64_t(N) * 26);
    ALLOC(r->gT, n * 78);
    ALLOC(r->b, n);
    ALLOC(r->H, B * Sp * n);
    ALLOC(r->x, T * n);
    ALLOC(r->y, int64_t(N) * 78);
    ALLOC(r->y_last, int64_t(N) * 78);
    ALLOC(r->m, B * T * S);
    ALLOC(r->R, T * Sp);
    ALLOC(r->xT, n * T);
    ALLOC(r->eT, n * T);
    ALLOC(r->avg, N);
    ALLOC(r->avg_valid, N);
    ALLOC(r->dGT, n * 78);
    ALLOC(r->dA, int64_t(N) * 26);
    ALLOC(r->fir, int64_t(N) * 78);
    ALLOC(r->iir, int64_t(N) * 78);
    ALLOC(r->mr_last, n);
    ALLOC(r->mr_sum, 1);
    if (d.optimizer == 1) {
        ALLOC(r->gm1, n * 78);
        ALLOC(r->gm2, n * 78);
        ALLOC(r->cm1, int64_t(N) * 26);
        ALLOC(r->cm2, int64_t(N) * 26);
    }
    ALLOC(r->loss, T);
    ALLOC(r->loss_mse, T);
    ALLOC(r->loss_mr, T);
    ALLOC(r->loss_b, r->nbatch);
    ALLOC(r->loss_mse_b, r->nbatch);
    ALLOC(r->loss_mr_b, r->nbatch);
    ALLOC(r->partial, splits * T * Sp);
    ALLOC(r->mr_partial, T * r->mr_blocks);
    ALLOC(r->staging, n * 78);
    ALLOC(r->step_mr, T);
    ALLOC(r->objective, 1);
    ALLOC(r->dmax, 1);
    ALLOC(r->u, T);
    ALLOC(r->barrier, 1);
    ALLOC(r->flags, ct::kMaxSlabFlags);
    ALLOC(r->scal, std::max<int64_t>(B, 1));
#undef ALLOC
    if (hipHostMalloc(reinterpret_cast<void **>(&r->scal_host),
                      sizeof(ct::UpdateScalars) * std::max<int64_t>(B, 1)) != hipSuccess) {
        r->scal_host = nullptr;
        return cleanup(fail(CT_ERR_DEVICE, "hipHostMalloc of the update scalars failed"));
    }

    hipStream_t s = gpu->stream;
    auto zero = [&](void *p, size_t bytes) -> int {
        CT_HIP(hipMemsetAsync(p, 0, bytes, s));
        return CT_OK;
    };
    if ((rc = zero(r->H, sizeof(float) * B * Sp * n)) != CT_OK) return cleanup(rc);
    if ((rc = zero(r->m, sizeof(float) * B * T * S)) != CT_OK) return cleanup(rc);
    if ((rc = zero(r->x, sizeof(float) * T * n)) != CT_OK) return cleanup(rc);
    if ((rc = zero(r->eT, sizeof(float) * T * n)) != CT_OK) return cleanup(rc);
    if ((rc = zero(r->R, sizeof(float) * T * Sp)) != CT_OK) return cleanup(rc);
    for (float *p : {r->loss, r->loss_mse, r->loss_mr})
        if ((rc = zero(p, sizeof(float) * T)) != CT_OK) return cleanup(rc);
    for (float *p : {r->loss_b, r->loss_mse_b, r->loss_mr_b})
        if ((rc = zero(p, sizeof(float) * r->nbatch)) != CT_OK) return cleanup(rc);
    if ((rc = zero(r->barrier, sizeof(ct::GridBarrier))) != CT_OK) return cleanup(rc);
    if ((rc = zero(r->flags, sizeof(uint32_t) * ct::kMaxSlabFlags)) != CT_OK) return cleanup(rc);
    if ((rc = zero(r->y, sizeof(float) * N * 78)) != CT_OK) return cleanup(rc);
    if ((rc = zero(r->y_last, sizeof(float) * N * 78)) != CT_OK) return cleanup(rc);
    if ((rc = zero(r->avg, sizeof(float) * N)) != CT_OK) return cleanup(rc);
    if ((rc = zero(r->avg_valid, N)) != CT_OK) return cleanup(rc);
    if (r->gm1) {
        for (float *p : {r->gm1, r->gm2})
            if ((rc = zero(p, sizeof(float) * n * 78)) != CT_OK) return cleanup(rc);
        for (float *p : {r->cm1, r->cm2})
            if ((rc = zero(p, sizeof(float) * N * 26)) != CT_OK) return cleanup(rc);
    }
    if ((rc = reset_derivatives(*r)) != CT_OK) return cleanup(rc);
    if (hipStreamSynchronize(s) != hipSuccess) return cleanup(fail(CT_ERR_DEVICE, "memset failed"));

    if (N > 0) {
        if ((rc = upload(r->nbr, nbr.data(), sizeof(int32_t) * N * 26)) != CT_OK) return cleanup(rc);
        if ((rc = upload(r->delays, d.delays, sizeof(int32_t) * N * 26)) != CT_OK) return cleanup(rc);
        if ((rc = upload(r->coefs, d.coefs, sizeof(float) * N * 26)) != CT_OK) return cleanup(rc);
        if ((rc = upload(r->initial_delays, d.initial_delays, sizeof(float) * N * 26)) != CT_OK)
            return cleanup(rc);
        if ((rc = upload(r->b, d.control_matrix, sizeof(float) * n)) != CT_OK) return cleanup(rc);
        if ((rc = upload(r->u, d.control_function, sizeof(float) * T)) != CT_OK) return cleanup(rc);
        // gains (n, 78) -> gT (78, n) through the staging buffer
        if ((rc = upload(r->staging, d.gains, sizeof(float) * n * 78)) != CT_OK) return cleanup(rc);
        if (ct::launch_transpose(s, r->staging, r->gT, n, 78) != hipSuccess)
            return cleanup(fail(CT_ERR_DEVICE, "gain transpose launch failed"));
        for (int64_t bb = 0; bb < B; ++bb) {
            hipError_t e = hipMemcpy2DAsync(r->H + bb * Sp * n, sizeof(float) * n,
                                            d.measurement_matrix + bb * S * n, sizeof(float) * n,
                                            sizeof(float) * n, S, hipMemcpyHostToDevice, s);
            if (e != hipSuccess)
                return cleanup(fail(CT_ERR_DEVICE, std::string("H upload: ") + hipGetErrorString(e)));
        }
        if (hipStreamSynchronize(s) != hipSuccess)
            return cleanup(fail(CT_ERR_DEVICE, "upload failed"));
    }
    if ((rc = plan_resident(*r)) != CT_OK) return cleanup(rc);
    if ((rc = plan_tiles(*r)) != CT_OK) return cleanup(rc);
    *out = r;
    return CT_OK;
}

void ct_results_gpu_free(ct_results_gpu *results) {
    if (!results) return;
    (void)hipSetDevice(results->gpu->device);
    (void)hipStreamSynchronize(results->stream);
    delete results;
}

int ct_results_update_from_gpu(ct_results_gpu *results, const ct_results_host *o) {
    if (!results || !o) return fail(CT_ERR_INVALID, "null argument");
    ct_results_gpu &r = *results;
    CT_CHECK(check_gpu(r.gpu));
    hipStream_t s = r.stream;
    CT_HIP(hipStreamSynchronize(s));
    const int64_t n = r.n, N = r.N, T = r.T, S = r.S;
    auto via_staging = [&](float *dst, auto &&produce) -> int {
        if (!dst || n == 0) return CT_OK;
        CT_HIP(produce());
        CT_HIP(hipStreamSynchronize(s));
        return download(dst, r.staging, sizeof(float) * n * 78);
    };
    CT_CHECK(via_staging(o->gains, [&] { return ct::launch_transpose(s, r.gT, r.staging, 78, n); }));
    CT_CHECK(via_staging(o->d_gains, [&] { return ct::launch_transpose(s, r.dGT, r.staging, 78, n); }));
    CT_CHECK(via_staging(o->ap_outputs_now, [&] { return ct::launch_expand_rows(s, r.y, r.staging, N); }));
    CT_CHECK(via_staging(o->ap_outputs_last, [&] { return ct::launch_expand_rows(s, r.y_last, r.staging, N); }));
    CT_CHECK(via_staging(o->d_coefs_fir, [&] { return ct::launch_expand_rows(s, r.fir, r.staging, N); }));
    CT_CHECK(via_staging(o->d_coefs_iir, [&] { return ct::launch_expand_rows(s, r.iir, r.staging, N); }));
    CT_CHECK(download(o->coefs, r.coefs, sizeof(float) * N * 26));
    CT_CHECK(download(o->delays, r.delays, sizeof(int32_t) * N * 26));
    CT_CHECK(download(o->d_coefs, r.dA, sizeof(float) * N * 26));
    CT_CHECK(download(o->system_states, r.x, sizeof(float) * T * n));
    CT_CHECK(download(o->measurements, r.m, sizeof(float) * r.B * T * S));
    CT_CHECK(download(o->residuals, r.R + (T - 1) * r.Sp, sizeof(float) * S));
    if (o->mapped_residuals && n > 0) {
        if (r.mapped_by_step)  // the reverse-mode forcing: row T - 1 of E
            CT_CHECK(download(o->mapped_residuals, r.E + (T - 1) * n, sizeof(float) * n));
        else
            CT_HIP(hipMemcpy2D(o->mapped_residuals, sizeof(float), r.eT + (T - 1),
                               sizeof(float) * T, sizeof(float), n, hipMemcpyDeviceToHost));
    }
    CT_CHECK(download(o->maximum_regularization, r.mr_last, sizeof(float) * n));
    CT_CHECK(download(o->maximum_regularization_sum, r.mr_sum, sizeof(float)));
    if (o->average_delays && N > 0) {
        std::vector<uint8_t> valid(N);
        CT_CHECK(download(o->average_delays, r.avg, sizeof(float) * N));
        CT_CHECK(download(valid.data(), r.avg_valid, N));
        for (int64_t v = 0; v < N; ++v)
            if (!valid[v]) o->average_delays[v] = std::nanf("");
    }
    CT_CHECK(download(o->loss, r.loss, sizeof(float) * T));
    CT_CHECK(download(o->loss_mse, r.loss_mse, sizeof(float) * T));
    CT_CHECK(download(o->loss_maximum_regularization, r.loss_mr, sizeof(float) * T));
    CT_CHECK(download(o->loss_batch, r.loss_b, sizeof(float) * r.nbatch));
    CT_CHECK(download(o->loss_mse_batch, r.loss_mse_b, sizeof(float) * r.nbatch));
    CT_CHECK(download(o->loss_maximum_regularization_batch, r.loss_mr_b, sizeof(float) * r.nbatch));
    if (o->derivatives_step) *o->derivatives_step = r.step;
    return CT_OK;
}

int ct_results_simulate(ct_results_gpu *results, float *measurements_out) {
    if (!results) return fail(CT_ERR_INVALID, "null argument");
    ct_results_gpu &r = *results;
    CT_CHECK(check_gpu(r.gpu));
    for (int32_t beat = 0; beat < r.B; ++beat) CT_CHECK(forward_beat(r, beat, nullptr, nullptr));
    CT_CHECK(check_resident(r));
    CT_HIP(hipStreamSynchronize(r.stream));
    if (measurements_out)
        CT_CHECK(download(measurements_out, r.m, sizeof(float) * r.B * r.T * r.S));
    return CT_OK;
}

int ct_measurement_matrix(ct_gpu *gpu, int32_t num_voxels, const float *voxel_positions_mm,
                          int32_t num_sensors, const float *sensor_positions_mm,
                          const float *sensor_orientations, int32_t num_beats,
                          const float *sensor_offsets_mm, float voxel_size_mm, float *out) {
    CT_CHECK(check_gpu(gpu));
    if (num_voxels < 0 || num_sensors < 0 || num_beats < 0)
        return fail(CT_ERR_INVALID, "negative voxel, sensor or beat count");
    if (!out || (num_voxels > 0 && !voxel_positions_mm) ||
        (num_sensors > 0 && (!sensor_positions_mm || !sensor_orientations)) ||
        (num_beats > 0 && !sensor_offsets_mm))
        return fail(CT_ERR_INVALID, "null array");
    if (!(voxel_size_mm > 0.0f) || !std::isfinite(voxel_size_mm))
        return fail(CT_ERR_INVALID, "voxel size must be positive and finite");
    const int64_t len = int64_t(num_beats) * num_sensors * 3 * num_voxels;
    if (len == 0) return CT_OK;
    // the reference's common factor, in its f32 operations
    const float pi = 3.14159265358979323846f;
    const float side = voxel_size_mm / 1000.0f;
    const float volume = side * (side * side);
    const float common = (float(1.25663706212e-6) * volume) / (4.0f * pi) * 1e12f;
    hipStream_t s = gpu->stream;
    float *dev = nullptr;
    const size_t in_floats = 3 * (size_t(num_voxels) + 2 * size_t(num_sensors) + size_t(num_beats));
    CT_HIP(hipMalloc(&dev, sizeof(float) * (in_floats + size_t(len))));
    float *vp = dev, *sp = vp + 3 * size_t(num_voxels), *so = sp + 3 * size_t(num_sensors);
    float *off = so + 3 * size_t(num_sensors), *h = off + 3 * size_t(num_beats);
    int rc = CT_OK;
    auto step = [&](hipError_t e) {
        if (rc == CT_OK && e != hipSuccess)
            rc = fail(CT_ERR_DEVICE, std::string("measurement matrix: ") + hipGetErrorString(e));
    };
    step(hipMemcpyAsync(vp, voxel_positions_mm, sizeof(float) * 3 * num_voxels,
                        hipMemcpyHostToDevice, s));
    step(hipMemcpyAsync(sp, sensor_positions_mm, sizeof(float) * 3 * num_sensors,
                        hipMemcpyHostToDevice, s));
    step(hipMemcpyAsync(so, sensor_orientations, sizeof(float) * 3 * num_sensors,
                        hipMemcpyHostToDevice, s));
    step(hipMemcpyAsync(off, sensor_offsets_mm, sizeof(float) * 3 * num_beats,
                        hipMemcpyHostToDevice, s));
    if (rc == CT_OK)
        step(ct::launch_measurement_matrix(s, vp, sp, so, off, common, num_voxels, num_sensors,
                                           num_beats, h));
    if (rc == CT_OK) step(hipMemcpyAsync(out, h, sizeof(float) * len, hipMemcpyDeviceToHost, s));
    step(hipStreamSynchronize(s));
    (void)hipFree(dev);
    return rc;
}

int ct_model_connections(ct_gpu *gpu, int32_t nx, int32_t ny, int32_t nz,
                         const int32_t *state_numbers, const float *positions_mm,
                         const float *velocities_m_per_s, float sample_rate_hz,
                         int32_t num_voxels, int32_t *output_state_indices,
                         float *delay_samples) {
    CT_CHECK(check_gpu(gpu));
    if (nx < 0 || ny < 0 || nz < 0 || num_voxels < 0)
        return fail(CT_ERR_INVALID, "negative grid size or voxel count");
    const int64_t cells = int64_t(nx) * ny * nz;
    if (cells > 0 && (!state_numbers || !positions_mm || !velocities_m_per_s))
        return fail(CT_ERR_INVALID, "null grid array");
    if (num_voxels > 0 && (!output_state_indices || !delay_samples))
        return fail(CT_ERR_INVALID, "null output array");
    if (!(sample_rate_hz > 0.0f) || !std::isfinite(sample_rate_hz))
        return fail(CT_ERR_INVALID, "sample rate must be positive and finite");
    if (cells * 27 >= (int64_t(1) << 62)) return fail(CT_ERR_INVALID, "grid too large");
    // every number is a distinct first state index 3v, v < num_voxels
    std::vector<uint8_t> seen(size_t(num_voxels), 0);
    for (int64_t c = 0; c < cells; ++c) {
        const int32_t v = state_numbers[c];
        if (v < 0) continue;
        if (v % 3 != 0 || v / 3 >= num_voxels || seen[size_t(v / 3)])
            return fail(CT_ERR_INVALID, "state numbers must be distinct multiples of 3 below 3N");
        seen[size_t(v / 3)] = 1;
    }
    hipStream_t s = gpu->stream;
    const size_t n_osi = size_t(num_voxels) * 3 * 78, n_d = size_t(num_voxels) * 26;
    const size_t bytes = sizeof(int32_t) * size_t(cells) + sizeof(float) * size_t(cells) * 4 +
                         sizeof(int32_t) * n_osi + sizeof(float) * n_d + 256;
    char *dev = nullptr;
    CT_HIP(hipMalloc(&dev, bytes));
    int32_t *num = reinterpret_cast<int32_t *>(dev);
    float *pos = reinterpret_cast<float *>(num + cells);
    float *vel = pos + 3 * cells;
    int32_t *osi = reinterpret_cast<int32_t *>(vel + cells);
    float *dly = reinterpret_cast<float *>(osi + n_osi);
    int32_t *flag = reinterpret_cast<int32_t *>(dly + n_d);
    int rc = CT_OK;
    auto step = [&](hipError_t e) {
        if (rc == CT_OK && e != hipSuccess)
            rc = fail(CT_ERR_DEVICE, std::string("model connections: ") + hipGetErrorString(e));
    };
    int32_t too_small = 0;
    step(hipMemcpyAsync(num, state_numbers, sizeof(int32_t) * cells, hipMemcpyHostToDevice, s));
    step(hipMemcpyAsync(pos, positions_mm, sizeof(float) * 3 * cells, hipMemcpyHostToDevice, s));
    step(hipMemcpyAsync(vel, velocities_m_per_s, sizeof(float) * cells, hipMemcpyHostToDevice, s));
    if (rc == CT_OK)
        step(ct::launch_model_connections(s, nx, ny, nz, num, pos, vel, sample_rate_hz, num_voxels,
                                          osi, dly, flag));
    if (rc == CT_OK && num_voxels > 0) {
        step(hipMemcpyAsync(output_state_indices, osi, sizeof(int32_t) * n_osi,
                            hipMemcpyDeviceToHost, s));
        step(hipMemcpyAsync(delay_samples, dly, sizeof(float) * n_d, hipMemcpyDeviceToHost, s));
    }
    if (rc == CT_OK) step(hipMemcpyAsync(&too_small, flag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    step(hipStreamSynchronize(s));
    (void)hipFree(dev);
    if (rc == CT_OK && too_small)
        rc = fail(CT_ERR_INVALID, "can not configure delays below 1 sample");
    return rc;
}

int ct_results_spherical(ct_results_gpu *results, float sample_rate_hz, float *magnitude,
                         float *theta, float *phi, float *magnitude_max, float *theta_max,
                         float *phi_max, float *activation_time_ms) {
    if (!results) return fail(CT_ERR_INVALID, "null argument");
    ct_results_gpu &r = *results;
    CT_CHECK(check_gpu(r.gpu));
    if (!(sample_rate_hz > 0.0f) || !std::isfinite(sample_rate_hz))
        return fail(CT_ERR_INVALID, "sample rate must be positive and finite");
    if (r.N == 0) return CT_OK;
    if (r.T == 0) return fail(CT_ERR_INVALID, "no steps: the maxima are undefined");
    hipStream_t s = r.stream;
    const size_t full = size_t(r.T) * r.N, N = size_t(r.N);
    const size_t nfull = size_t(magnitude != nullptr) + (theta != nullptr) + (phi != nullptr);
    char *dev = nullptr;
    CT_HIP(hipMalloc(&dev, sizeof(float) * (nfull * full + 4 * N) + 256));
    float *p = reinterpret_cast<float *>(dev);
    float *d_mag = magnitude ? p : nullptr;
    if (magnitude) p += full;
    float *d_theta = theta ? p : nullptr;
    if (theta) p += full;
    float *d_phi = phi ? p : nullptr;
    if (phi) p += full;
    float *d_mm = p, *d_tm = p + N, *d_pm = p + 2 * N, *d_act = p + 3 * N;
    int32_t *d_flag = reinterpret_cast<int32_t *>(p + 4 * N);
    int rc = CT_OK;
    auto step = [&](hipError_t e) {
        if (rc == CT_OK && e != hipSuccess)
            rc = fail(CT_ERR_DEVICE, std::string("spherical states: ") + hipGetErrorString(e));
    };
    int32_t no_max = 0;
    step(ct::launch_spherical(s, r.x, r.T, r.N, sample_rate_hz, d_mag, d_theta, d_phi, d_mm, d_tm,
                              d_pm, d_act, d_flag));
    const std::pair<float *, float *> copies[] = {
        {magnitude, d_mag}, {theta, d_theta}, {phi, d_phi}};
    for (const auto &c : copies)
        if (c.first && rc == CT_OK)
            step(hipMemcpyAsync(c.first, c.second, sizeof(float) * full, hipMemcpyDeviceToHost, s));
    const std::pair<float *, float *> small[] = {
        {magnitude_max, d_mm}, {theta_max, d_tm}, {phi_max, d_pm}, {activation_time_ms, d_act}};
    for (const auto &c : small)
        if (c.first && rc == CT_OK)
            step(hipMemcpyAsync(c.first, c.second, sizeof(float) * N, hipMemcpyDeviceToHost, s));
    if (rc == CT_OK) step(hipMemcpyAsync(&no_max, d_flag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    step(hipStreamSynchronize(s));
    (void)hipFree(dev);
    if (rc == CT_OK && no_max)
        rc = fail(CT_ERR_INVALID, "a voxel has no non-NaN magnitude (argmax_skipnan fails)");
    return rc;
}

int ct_results_calculate_final_metrics(ct_results_gpu *results, const uint32_t *voxel_types,
                                       const int32_t *voxel_numbers, int64_t cells, float *dice,
                                       float *iou, float *precision, float *recall) {
    CT_CHECK(final_metrics_check(results, voxel_types, voxel_numbers, cells));
    ct_results_gpu &r = *results;
    CT_CHECK(check_gpu(r.gpu));
    const int64_t nvox = std::max<int32_t>(r.N, 1), ncell = std::max<int64_t>(cells, 1);
    char *dev = nullptr;
    CT_HIP(hipMalloc(&dev, sizeof(uint32_t) * 2 * kFinalBins + sizeof(float) * nvox +
                               (sizeof(int32_t) + sizeof(uint32_t)) * ncell));
    uint32_t *d_hist = reinterpret_cast<uint32_t *>(dev);
    float *d_max = reinterpret_cast<float *>(d_hist + 2 * kFinalBins);
    int32_t *d_num = reinterpret_cast<int32_t *>(d_max + nvox);
    uint32_t *d_types = reinterpret_cast<uint32_t *>(d_num + ncell);
    uint32_t hist[2 * kFinalBins] = {};
    const int rc = final_metrics_device(r, voxel_types, voxel_numbers, cells, d_hist, d_max, d_num,
                                        d_types, hist);
    (void)hipFree(dev);
    CT_CHECK(rc);
    final_metrics_from_histograms(hist, dice, iou, precision, recall);
    return CT_OK;
}

int ct_results_calculate_final_metrics_noalloc(ct_results_gpu *results,
                                               const uint32_t *voxel_types,
                                               const int32_t *voxel_numbers, int64_t cells,
                                               float *dice, float *iou, float *precision,
                                               float *recall) {
    CT_CHECK(final_metrics_check(results, voxel_types, voxel_numbers, cells));
    ct_results_gpu &r = *results;
    const auto aligned = [](size_t bytes) { return (bytes + 255) / 256 * 256; };
    const size_t o_max = aligned(sizeof(uint32_t) * 2 * kFinalBins);
    const size_t o_num = o_max + aligned(sizeof(float) * size_t(r.N));
    const size_t o_types = o_num + aligned(sizeof(int32_t) * size_t(cells));
    const size_t need = o_types + sizeof(uint32_t) * size_t(cells);
    if (need > kNoAllocScratchBytes)
        return fail(CT_ERR_INVALID,
                    "final metrics without allocation: " + std::to_string(r.N) + " voxels and " +
                        std::to_string(cells) + " cells need " + std::to_string(need) +
                        " bytes of scratch, more than its " +
                        std::to_string(kNoAllocScratchBytes) +
                        " (4 B per voxel + 8 B per cell + 1 KiB); use "
                        "ct_results_calculate_final_metrics");
    CT_CHECK(check_gpu(r.gpu));
    void *base = nullptr;
    CT_CHECK(scratch_address(r.gpu->device, &base));
    char *sp = static_cast<char *>(base);
    uint32_t hist[2 * kFinalBins] = {};
    CT_CHECK(final_metrics_device(r, voxel_types, voxel_numbers, cells,
                                  reinterpret_cast<uint32_t *>(sp),
                                  reinterpret_cast<float *>(sp + o_max),
                                  reinterpret_cast<int32_t *>(sp + o_num),
                                  reinterpret_cast<uint32_t *>(sp + o_types), hist));
    final_metrics_from_histograms(hist, dice, iou, precision, recall);
    return CT_OK;
}

int ct_results_simulate_noisy(ct_results_gpu *results, const float *sensor_std, uint64_t seed,
                              float *measurements_out) {
    if (!results || !sensor_std) return fail(CT_ERR_INVALID, "null argument");
    ct_results_gpu &r = *results;
    CT_CHECK(check_gpu(r.gpu));
    // Normal::new rejects a negative or non-finite standard deviation
    for (int32_t s = 0; s < r.S; ++s)
        if (!(sensor_std[s] >= 0.0f) || !std::isfinite(sensor_std[s]))
            return fail(CT_ERR_INVALID, "measurement noise standard deviation of sensor " +
                                            std::to_string(s) + " is negative or not finite");
    CT_CHECK(ensure(r, &r.sensor_std, r.sensor_std_have, std::max<int64_t>(r.S, 1), 1));
    CT_CHECK(upload(r.sensor_std, sensor_std, sizeof(float) * r.S));
    CT_CHECK(ct_results_simulate(results, nullptr));
    CT_HIP(ct::launch_measurement_noise(r.stream, r.m, r.sensor_std, r.B, r.T, r.S, seed));
    CT_HIP(hipStreamSynchronize(r.stream));
    if (measurements_out)
        CT_CHECK(download(measurements_out, r.m, sizeof(float) * r.B * r.T * r.S));
    return CT_OK;
}

int ct_epoch_kernel_new(ct_gpu *gpu, ct_results_gpu *results, const ct_buffer *actual,
                        const ct_algorithm *config, int32_t number_of_states,
                        int32_t number_of_sensors, int32_t number_of_steps, ct_epoch_kernel **out) {
    CT_CHECK(check_gpu(gpu));
    if (!results || !actual || !config || !out) return fail(CT_ERR_INVALID, "null argument");
    if (results->gpu != gpu) return fail(CT_ERR_INVALID, "results belong to another ct_gpu");
    if (int64_t(number_of_states) != results->n || number_of_sensors != results->S ||
        number_of_steps != results->T)
        return fail(CT_ERR_INVALID, "states/sensors/steps do not match the results");
    if (actual->beats != results->B || actual->steps != results->T || actual->sensors != results->S)
        return fail(CT_ERR_INVALID, "actual measurements are not (B, T, S) of the results");
    if (config->optimizer == 1 && !results->gm1)
        return fail(CT_ERR_INVALID, "Adam needs results uploaded with optimizer = 1");
    if (config->batch_size < 0) return fail(CT_ERR_INVALID, "negative batch size");
    ct_epoch_kernel *k = new ct_epoch_kernel;
    k->gpu = gpu;
    k->r = results;
    k->mhat = actual;
    k->cfg = *config;
    for (int32_t b = 0; b < results->B; ++b) k->order.push_back(b);
    *out = k;
    return CT_OK;
}

void ct_epoch_kernel_free(ct_epoch_kernel *kernel) {
    if (!kernel) return;
    (void)hipSetDevice(kernel->gpu->device);
    (void)hipStreamSynchronize(kernel->gpu->stream);
    delete kernel;
}

int ct_epoch_kernel_execute(ct_epoch_kernel *k) {
    if (!k) return fail(CT_ERR_INVALID, "null argument");
    ct_results_gpu &r = *k->r;
    CT_CHECK(check_gpu(r.gpu));
    const ct_algorithm &c = k->cfg;
    k->used_events = 0;
    k->spans.clear();
    if (k->profiling) {
        const size_t need = 2 * size_t(r.T) * r.B + 64 * size_t(r.B) + 64;
        while (k->events.size() < need) {
            hipEvent_t e;
            CT_HIP(hipEventCreate(&e));
            k->events.push_back(e);
        }
    }
    CT_CHECK(begin_epoch(*k));
    Span total(k, kTotal);
    CT_CHECK(reset_derivatives(r));
    const bool batched = c.batch_size != 0;
    int64_t in_batch = 0;
    int32_t update = 0;
    for (int32_t pos = 0; pos < int32_t(k->order.size()); ++pos) {
        CT_CHECK(run_beat(*k, pos));
        if (batched) {
            in_batch += 1;
            if (in_batch == c.batch_size) {
                CT_CHECK(finish_batch(*k, in_batch, true, update++));
                in_batch = 0;
            }
        }
    }
    if (batched) {
        if (in_batch > 0) CT_CHECK(finish_batch(*k, in_batch, false, update++));
    } else {
        CT_CHECK(finish_batch(*k, r.B, false, update++));
    }
    // a timed-out resident wait: the update was skipped on the device, so the
    // epoch's counters (Adam step, batch index) stay where they were
    CT_CHECK(check_resident(r));
    end_epoch(*k);
    return CT_OK;
}

int ct_results_set_propagation(ct_results_gpu *results, int32_t mode, int32_t *resident) {
    if (!results) return fail(CT_ERR_INVALID, "null argument");
    if (mode < 0 || mode > 3)
        return fail(CT_ERR_INVALID, "propagation mode must be 0, 1, 2 or 3");
    if (mode >= 2 && !results->resident_ok && results->N > 0)
        return fail(CT_ERR_INVALID, "the resident propagation kernel cannot hold this problem");
    results->propagation_mode = mode;
    if (resident) *resident = (results->resident_ok && mode != 1) ? 1 : 0;
    return CT_OK;
}

int ct_results_set_derivation(ct_results_gpu *results, int32_t mode, int32_t *tiled) {
    if (!results) return fail(CT_ERR_INVALID, "null argument");
    if (mode < 0 || mode > 2) return fail(CT_ERR_INVALID, "derivation mode must be 0, 1 or 2");
    if (mode == 2 && results->tiles.tiles <= 0 && results->N > 0)
        return fail(CT_ERR_INVALID, "no voxel tiles were built for this problem");
    results->derive_mode = mode;
    if (tiled) *tiled = (results->tiles.tiles > 0 && mode != 1) ? 1 : 0;
    return CT_OK;
}

int ct_results_set_wait_limit(ct_results_gpu *results, double seconds) {
    if (!results) return fail(CT_ERR_INVALID, "null argument");
    if (!(seconds >= 0.0) || !std::isfinite(seconds))
        return fail(CT_ERR_INVALID, "wait limit must be finite and >= 0 seconds");
    results->wait_seconds = seconds;
    return CT_OK;
}

int ct_epoch_kernel_set_gradient(ct_epoch_kernel *k, int32_t mode, int32_t checkpoint_interval) {
    if (!k) return fail(CT_ERR_INVALID, "null argument");
    if (mode < 0 || mode > 1) return fail(CT_ERR_INVALID, "gradient mode must be 0 or 1");
    if (checkpoint_interval < 1) return fail(CT_ERR_INVALID, "checkpoint interval must be >= 1");
    k->gradient = mode;
    k->checkpoint_interval = checkpoint_interval;
    return CT_OK;
}

int ct_results_forward(ct_results_gpu *results, const ct_buffer *actual, int32_t beat,
                       const ct_algorithm *config, int32_t checkpoint_interval,
                       double *objective) {
    if (!results || !actual || !config) return fail(CT_ERR_INVALID, "null argument");
    ct_results_gpu &r = *results;
    CT_CHECK(check_gpu(r.gpu));
    if (beat < 0 || beat >= r.B) return fail(CT_ERR_INVALID, "beat out of range");
    if (actual->beats != r.B || actual->steps != r.T || actual->sensors != r.S)
        return fail(CT_ERR_INVALID, "actual measurements are not (B, T, S) of the results");
    CT_HIP(hipMemsetAsync(r.mr_sum, 0, sizeof(float), r.stream));
    if (checkpoint_interval < 0) return fail(CT_ERR_INVALID, "negative checkpoint interval");
    CT_CHECK(forward_beat(r, beat, actual->data, nullptr, checkpoint_interval));
    CT_CHECK(losses_beat(r, *config, nullptr));
    CT_CHECK(check_resident(r));
    CT_HIP(hipStreamSynchronize(r.stream));
    if (objective) CT_CHECK(download(objective, r.objective, sizeof(double)));
    return CT_OK;
}

int ct_results_objective_f64(ct_results_gpu *results, const ct_buffer *actual, int32_t beat,
                             const ct_algorithm *config, int32_t param, int64_t index,
                             double delta, double *objective) {
    if (!results || !actual || !config || !objective) return fail(CT_ERR_INVALID, "null argument");
    ct_results_gpu &r = *results;
    CT_CHECK(check_gpu(r.gpu));
    if (beat < 0 || beat >= r.B) return fail(CT_ERR_INVALID, "beat out of range");
    if (actual->beats != r.B || actual->steps != r.T || actual->sensors != r.S)
        return fail(CT_ERR_INVALID, "actual measurements are not (B, T, S) of the results");
    if (param < 0 || param > 2) return fail(CT_ERR_INVALID, "param must be 0, 1 (gains) or 2 (coefs)");
    if ((param == 1 && (index < 0 || index >= r.n * 78)) ||
        (param == 2 && (index < 0 || index >= int64_t(r.N) * 26)))
        return fail(CT_ERR_INVALID, "parameter index out of range");
    hipStream_t s = r.stream;
    const int64_t n = r.n, T = r.T;
    const int64_t tiles = (r.Sp / 64) * ((T + 63) / 64);
    const int32_t splits = int32_t(std::min<int64_t>(64, std::max<int64_t>(1, (2048 + tiles - 1) /
                                                                              std::max<int64_t>(tiles, 1))));
    const int64_t need = int64_t(r.N) * 78 + T * n + int64_t(splits) * r.Sp * T + T + 1;
    if (r.f64_have < need) {
        for (double *p : {r.y64}) {
            if (!p) continue;
            auto it = std::find(r.allocations.begin(), r.allocations.end(), static_cast<void *>(p));
            if (it != r.allocations.end()) r.allocations.erase(it);
            CT_HIP(hipStreamSynchronize(s));
            CT_HIP(hipFree(p));
        }
        r.y64 = nullptr;
        CT_CHECK(r.alloc(&r.y64, need));
        r.f64_have = need;
    }
    r.x64 = r.y64 + int64_t(r.N) * 78;
    r.p64 = r.x64 + T * n;
    r.terms64 = r.p64 + int64_t(splits) * r.Sp * T;
    r.j64 = r.terms64 + T;
    ct::Perturbation pert{};
    pert.kind = param;
    pert.index = index;
    pert.delta = delta;
    CT_HIP(hipMemsetAsync(r.y64, 0, sizeof(double) * r.N * 78, s));
    for (int32_t t = 0; t < r.T; ++t)
        CT_HIP(ct::launch_propagate_f64(s, r.nbr, r.coefs, r.delays, r.gT, r.y64, r.x64, r.b,
                                        r.control_function[t], r.N, t, pert));
    const double s_mse = double(config->mse_strength) / double(r.S);
    CT_HIP(ct::launch_objective_f64(s, r.H + int64_t(beat) * r.Sp * r.n, r.x64,
                                    actual->data + int64_t(beat) * r.T * r.S, r.p64, splits,
                                    r.terms64, r.j64, r.S, r.Sp, r.T, r.N, s_mse,
                                    double(config->maximum_regularization_strength),
                                    double(config->maximum_regularization_threshold)));
    CT_HIP(hipStreamSynchronize(s));
    return download(objective, r.j64, sizeof(double));
}

int ct_results_backward(ct_results_gpu *results, const ct_buffer *actual, int32_t beat,
                        const ct_algorithm *config, int32_t truncate,
                        int32_t checkpoint_interval) {
    if (!results || !actual || !config) return fail(CT_ERR_INVALID, "null argument");
    ct_results_gpu &r = *results;
    CT_CHECK(check_gpu(r.gpu));
    if (r.forward_beat != beat)
        return fail(CT_ERR_STATE, "backward needs the forward pass of the same beat first");
    CT_CHECK(reset_derivatives(r));
    if (truncate) {
        ct_algorithm c = *config;
        CT_CHECK(derive_reference(r, beat, c, nullptr));
    } else {
        CT_CHECK(derive_adjoint(r, beat, *config, checkpoint_interval, nullptr));
    }
    CT_CHECK(check_resident(r));
    CT_HIP(hipStreamSynchronize(r.stream));
    return CT_OK;
}

int ct_epoch_kernel_set_share_propagation(ct_epoch_kernel *k, int32_t enabled) {
    if (!k) return fail(CT_ERR_INVALID, "null argument");
    if (enabled < -1 || enabled > 1) return fail(CT_ERR_INVALID, "share mode must be -1, 0 or 1");
    k->share_mode = enabled;
    return CT_OK;
}

int ct_epoch_kernel_set_freeze_delays(ct_epoch_kernel *k, int32_t value) {
    if (!k) return fail(CT_ERR_INVALID, "null argument");
    k->cfg.freeze_delays = value != 0;
    return CT_OK;
}

int ct_epoch_kernel_set_freeze_gains(ct_epoch_kernel *k, int32_t value) {
    if (!k) return fail(CT_ERR_INVALID, "null argument");
    k->cfg.freeze_gains = value != 0;
    return CT_OK;
}

int ct_epoch_kernel_set_beat_order(ct_epoch_kernel *k, const int32_t *order, int32_t len) {
    if (!k) return fail(CT_ERR_INVALID, "null argument");
    const int32_t B = k->r->B;
    std::vector<int32_t> o;
    if (!order) {
        for (int32_t b = 0; b < B; ++b) o.push_back(b);
    } else {
        if (len != B) return fail(CT_ERR_INVALID, "beat order must list every beat once");
        std::vector<char> seen(B, 0);
        for (int32_t i = 0; i < len; ++i) {
            if (order[i] < 0 || order[i] >= B || seen[order[i]])
                return fail(CT_ERR_INVALID, "beat order is not a permutation of 0..B-1");
            seen[order[i]] = 1;
            o.push_back(order[i]);
        }
    }
    k->order = o;
    return CT_OK;
}

int ct_epoch_kernel_set_batch_index(ct_epoch_kernel *k, int32_t batch_index) {
    if (!k) return fail(CT_ERR_INVALID, "null argument");
    if (batch_index < 0) return fail(CT_ERR_INVALID, "negative batch index");
    k->batch_index = batch_index;
    return CT_OK;
}

int ct_epoch_kernel_get_batch_index(const ct_epoch_kernel *k, int32_t *batch_index) {
    if (!k || !batch_index) return fail(CT_ERR_INVALID, "null argument");
    *batch_index = k->batch_index;
    return CT_OK;
}

int ct_epoch_kernel_last_loss(ct_epoch_kernel *k, float *loss, float *loss_mse,
                              float *loss_maximum_regularization, int32_t *is_normal) {
    if (!k) return fail(CT_ERR_INVALID, "null argument");
    CT_CHECK(check_gpu(k->gpu));
    if (k->batch_index < 1) return fail(CT_ERR_INVALID, "no batch has been recorded yet");
    ct_results_gpu &r = *k->r;
    const int64_t i = int64_t(k->batch_index) - 1;
    if (i >= r.nbatch) return fail(CT_ERR_INVALID, "batch index beyond the batch arrays");
    float v[3] = {0.0f, 0.0f, 0.0f};
    CT_HIP(hipStreamSynchronize(r.stream));
    CT_CHECK(download(&v[0], r.loss_b + i, sizeof(float)));
    CT_CHECK(download(&v[1], r.loss_mse_b + i, sizeof(float)));
    CT_CHECK(download(&v[2], r.loss_mr_b + i, sizeof(float)));
    if (loss) *loss = v[0];
    if (loss_mse) *loss_mse = v[1];
    if (loss_maximum_regularization) *loss_maximum_regularization = v[2];
    if (is_normal) *is_normal = std::isnormal(v[0]) ? 1 : 0;
    return CT_OK;
}

int ct_epoch_kernel_set_profiling(ct_epoch_kernel *k, int32_t enabled) {
    if (!k) return fail(CT_ERR_INVALID, "null argument");
    k->profiling = enabled != 0;
    return CT_OK;
}

int ct_epoch_kernel_get_profile(ct_epoch_kernel *k, ct_profile *out) {
    if (!k || !out) return fail(CT_ERR_INVALID, "null argument");
    CT_CHECK(check_gpu(k->gpu));
    CT_HIP(hipStreamSynchronize(k->gpu->stream));
    std::memset(out, 0, sizeof(*out));
    for (auto &sp : k->spans) {
        float ms = 0.0f;
        CT_HIP(hipEventElapsedTime(&ms, k->events[sp.start], k->events[sp.end]));
        switch (sp.phase) {
            case kPropagate:
                out->propagate_ms += ms;
                out->propagate_launches += 1;
                out->propagate_steps += sp.steps;
                break;
            case kProject: out->project_ms += ms; break;
            case kRegularize: out->regularize_ms += ms; break;
            case kTranspose: out->transpose_ms += ms; break;
            case kMapResiduals: out->map_residuals_ms += ms; break;
            case kDerive: out->derive_ms += ms; break;
            case kUpdate: out->update_ms += ms; break;
            case kTotal: out->total_ms += ms; break;
            case kAdjointSweep: out->adjoint_sweep_ms += ms; break;
        }
    }
    return CT_OK;
}

}  // extern "C"
