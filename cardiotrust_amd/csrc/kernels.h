// Internal launch interface of the gfx950 kernels (kernels.hip).
// Device layouts (N voxels, n = 3N states, T steps, Sp = padded sensors):
//   nbr, coefs, delays, initial_delays   [v][d]          (N, 26)
//   gT, dGT                              [3d+out][3v+in] (78, n)  gains transposed
//   y, y_last, fir, iir                  [v][3d+out]     (N, 78)  one filter per (v,d,out)
//   x                                    [t][i]          (T, n)
//   xT, eT                               [i][t]          (n, T)
//   H                                    [b][s][i]       (B, Sp, n), rows s >= S are zero
//   R (residuals)                        [t][s]          (T, Sp), columns s >= S are zero
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace ct {

constexpr int kGemmTileK = 32;
constexpr int kProjTileK = 32;  // states per k tile of the projection H·Xᵀ (its split-K unit)
constexpr int kProjTile = 128;  // steps per project workgroup (H·Xᵀ)
constexpr int kProjRows = 112;  // sensors per project workgroup (seven 16-row MFMA blocks)
// Sensor tiles of the projection for S sensors (the tiles cover rows up to
// S rounded up to 112; H holds Sp >= S rows, those >= S zero).
inline int32_t project_sensor_tiles(int32_t S) { return (S + kProjRows - 1) / kProjRows; }

struct DeriveParams {
    float mse_scaling;
    float reg_scaling;
    float threshold;
    float diff_strength;
    int32_t freeze_gains;
    int32_t freeze_delays;
    int32_t simple;
};

hipError_t launch_propagate_step(hipStream_t s, const int32_t *nbr, const float *coefs,
                                 const int32_t *delays, const float *gT, float *y, float *y_last,
                                 float *x, const float *b, float u_t, int32_t N, int32_t t,
                                 int32_t T);

// Resident propagation (one launch for steps [t_begin, t_end)). A workgroup
// holds at most kResMaxVoxels voxels; requires N < 2^17 - 1,
// blocks * kResMaxVoxels >= N and blocks <= what fits on the device at once.
// `abort` != 0 after the launch means a wait timed out. The counting barrier
// words serve the resident adjoint sweep; the propagation kernels wait on
// per-slab flags and use only `abort`.
constexpr int kResMaxVoxels = 472;
constexpr int kBarrierGroups = 8;
struct alignas(128) BarrierWord {
    uint32_t count;
    uint32_t pad[31];
};
struct GridBarrier {
    BarrierWord group[kBarrierGroups];
    BarrierWord top;
    BarrierWord generation;
    BarrierWord abort;  // sticky: set on a timed-out wait, cleared only by the host
};
static_assert(sizeof(GridBarrier) == 11 * sizeof(BarrierWord), "barrier layout");
// One job of the resident propagation: one scenario's beat. slot0 / width
// place its team in a batched launch.
struct ResidentJob {
    const int32_t *nbr;
    const float *coefs;
    const int32_t *delays;
    const float *gT;
    float *y;
    float *y_last;
    float *x;
    const float *b;
    const float *u;
    float *ckpt;          // filter outputs every ckpt_interval steps (or unused)
    GridBarrier *gb;        // the job's sticky abort word
    uint32_t *flags;        // completed steps per team member (width words)
    int32_t N;
    int32_t T;
    int32_t ckpt_interval;  // 0 = no checkpoints
    int32_t slot0;
    int32_t width;          // workgroups in the team
    // 1: the states are exchanged between workgroups through agent-coherent
    // accesses (stores written through the L2, gathers that bypass it), so a
    // step needs no agent-scope release / acquire; requires 4 T n < 2^32
    // (byte offsets of one buffer resource). 0: normal accesses with the
    // agent-scope fences around every wait.
    int32_t coherent;
};
// Most workgroups of a resident launch (one flag word each).
constexpr int kMaxSlabFlags = 4096;
size_t resident_lds_bytes();
hipError_t resident_capacity(int32_t *blocks_per_cu);
hipError_t launch_propagate_resident(hipStream_t s, int32_t blocks, const ResidentJob &job,
                                     int32_t t_begin, int32_t t_end, uint64_t wait_ticks);
// Batched resident propagation: `slots` workgroups (all resident at once) run
// the task lists task_job[task_ptr[q] .. task_ptr[q+1]) of slot q; `jobs`, the
// `num_jobs` abort words and the jobs' `num_flags` slab flags are device
// arrays, cleared here before the launch. `coherent`: every job has
// coherent != 0 (otherwise all jobs take the fenced exchange).
hipError_t batch_capacity(int32_t *blocks_per_cu);
hipError_t launch_propagate_batch(hipStream_t s, int32_t slots, const ResidentJob *jobs,
                                  const int32_t *task_ptr, const int32_t *task_job,
                                  GridBarrier *barriers, int32_t num_jobs, uint32_t *flags,
                                  int32_t num_flags, uint64_t first_wait, uint64_t step_wait,
                                  bool coherent);

// H·Xᵀ split-K partial sums; with xT not null the same pass also writes the
// states transposed, xT (n, T).
hipError_t launch_project(hipStream_t s, const float *x, const float *H, float *partial,
                          float *xT, int32_t T, int64_t n, int32_t S, int32_t Sp, int32_t splits,
                          int64_t k_per_split);

hipError_t launch_residuals(hipStream_t s, const float *partial, int32_t splits, const float *mhat,
                            float *m, float *R, float *loss_mse, int32_t T, int32_t S, int32_t Sp);

// Voxels per maxreg_partials workgroup (four per thread): the partial sums
// per step number ceil(N / kMaxregVoxels).
constexpr int kMaxregVoxels = 1024;
hipError_t launch_maxreg_partials(hipStream_t s, const float *x, float *partial, float *mr_last,
                                  int32_t N, int32_t T, float threshold, int32_t blocks);

hipError_t launch_maxreg_finalize(hipStream_t s, const float *partial, int32_t blocks,
                                  float *step_mr, float *mr_sum, const float *loss_mse,
                                  float *loss_mr, float *loss, double *objective,
                                  float mse_strength, float reg_strength, int32_t T);

hipError_t launch_transpose(hipStream_t s, const float *in, float *out, int64_t rows, int64_t cols);

// eT = Hᵀ R, or with `accumulate` eT = eT + Hᵀ R (each element one FP32 add).
// With `by_step` the output is E = eT transposed, (T, n) (the same values bit
// for bit); only where map_residuals_by_step_ok(S, T, n).
bool map_residuals_by_step_ok(int32_t S, int32_t T, int64_t n);
// E (by step, (T, n)) = the sum over up to kMapMaxBeats beats of H_bᵀ·R_bᵀ, with
// `accumulate` added to what E holds; H_b = H + h_off[b] (elements), R_b = R +
// b r_stride. Only where map_residuals_beats_ok(S, T, n).
constexpr int kMapMaxBeats = 8;
bool map_residuals_beats_ok(int32_t S, int32_t T, int64_t n);
struct MapBeats {
    int64_t h_off[kMapMaxBeats];
    int32_t count;
};
hipError_t launch_map_residuals_beats(hipStream_t s, const float *H, const MapBeats &beats,
                                      const float *R, int64_t r_stride, float *E, int64_t n,
                                      int32_t T, int32_t S, int32_t Sp, bool accumulate);
hipError_t launch_map_residuals(hipStream_t s, const float *H, const float *R, float *out,
                                int64_t n, int32_t T, int32_t S, int32_t Sp, bool accumulate,
                                bool by_step = false);

hipError_t launch_derive(hipStream_t s, const int32_t *nbr, const float *coefs,
                         const int32_t *delays, const float *initial_delays, const float *gT,
                         const float *xT, const float *eT, float *fir, float *iir, float *dGT,
                         float *dA, int32_t N, int32_t T, const DeriveParams &p);

// The same derivation over voxel tiles: a workgroup takes the connections of
// one tile (at most kTileMaxVoxels voxels) and stages the state rows of the
// tile's halo (its voxels and all their neighbours, at most kTileMaxHalo) in
// an LDS ring of kTileRing steps, so each row is read from memory once per
// tile instead of once per connection that reads it. Connections with a
// delay above kTileMaxDelay read their rows from memory. Tables (host-built
// by build_voxel_tiles): tile g owns the tile-order positions
// [vox_ptr[g], vox_ptr[g+1]) holding voxels vox[.], and halo voxels
// halo[halo_ptr[g] .. halo_ptr[g+1]); slot[26 p + d] is the halo index of the
// neighbour of connection d of the voxel at position p (kNoSlot if missing).
// The LDS ring holds kTileRing steps of each halo row; the chunk at ring
// positions [0, kTileChunk) is mirrored behind it.
constexpr int kTileThreads = 1024;
constexpr int kTileChunk = 16;
constexpr int kTileRing = 64;
constexpr int kTileRowStride = kTileRing + kTileChunk;
constexpr int kTileMaxVoxels = kTileThreads / 26;  // 39
constexpr int kTileMaxHalo = 152;
constexpr uint8_t kNoSlot = 0xFF;
struct DeriveTiles {
    const int32_t *vox_ptr;
    const int32_t *vox;
    const int32_t *halo_ptr;
    const int32_t *halo;
    const uint8_t *slot;
    int32_t tiles;
};
size_t derive_tiled_lds_bytes();
// Sets the tiled kernels' LDS limit; once per process before their first launch.
hipError_t derive_tiled_prepare();
hipError_t launch_derive_tiled(hipStream_t s, const DeriveTiles &tiles, const int32_t *nbr,
                               const float *coefs, const int32_t *delays,
                               const float *initial_delays, const float *gT, const float *xT,
                               const float *eT, float *fir, float *iir, float *dGT, float *dA,
                               int32_t N, int32_t T, const DeriveParams &p);

// Reverse-mode gradient (adjoint.hip).
struct AdjointParams {
    float mse_scaling;
    float reg_scaling;
    float threshold;
};
// One backward sweep of a batched sweep launch: a scenario's arrays as the
// resident sweep takes them (gains by state row, consumer table, forcing E
// and states x by step, x̄ written by step), the team placement slot0 / width
// (slab m of the team holds voxels [m N / width, (m + 1) N / width)), and the
// job's own barrier; a timed-out wait raises the launch's shared abort word.
struct SweepJob {
    const int32_t *cons_ptr;
    const int32_t *cons_idx;
    const int32_t *delays;
    const float *gains_rows;
    const float *coefs;
    float *gcons;
    const float *x;
    const float *E;
    float *xbar;
    GridBarrier *gb;
    int32_t N;
    int32_t T;
    AdjointParams p;
    int32_t slot0;
    int32_t width;
    int32_t coherent;  // 1: the job exchanges x̄ rows through agent-coherent accesses
};
// Workgroups of the batched sweep that fit on one CU at once.
hipError_t sweep_batch_capacity(int32_t *blocks_per_cu);
// Batched backward sweeps: `slots` workgroups (all resident at once) run the
// task lists task_job[task_ptr[q] .. task_ptr[q+1]) of slot q over `jobs`;
// the jobs' barriers (job_barriers[0 .. num_jobs)) are cleared here before
// the launch, `abort` (sticky) is not. A job's first wait gets `first_wait`
// ticks (its members may still be finishing earlier jobs), the others
// `step_wait`. Every job must have coherent != 0.
hipError_t launch_adjoint_sweep_batch(hipStream_t s, int32_t slots, const SweepJob *jobs,
                                      GridBarrier *job_barriers, int32_t num_jobs,
                                      const int32_t *task_ptr, const int32_t *task_job,
                                      uint32_t *abort, uint64_t first_wait, uint64_t step_wait);
hipError_t launch_max_delay(hipStream_t s, const int32_t *nbr, const int32_t *delays, int64_t len,
                            int32_t *out);
hipError_t launch_consumer_delays(hipStream_t s, const int32_t *cons_idx, const int32_t *delays,
                                  int32_t *cons_delay, int64_t len);
hipError_t launch_adjoint_step(hipStream_t s, const int32_t *nbr, const float *coefs,
                               const float *gains_rows, const int32_t *cons_ptr,
                               const int32_t *cons_idx, const int32_t *cons_delay, const float *x,
                               const float *E, float *ybar, float *qring, float *xbar, int32_t N,
                               int32_t T, int32_t step, uint32_t rmask, const AdjointParams &p);
// `ckpt` holds y at steps kK - 1 (k = 1 .. ceil(T/K) - 1), written by the
// forward propagation.
// Resident backward sweep (all steps in one launch; the workgroups stay
// resident and meet at `gb` after every step). Each connection's ybar
// recurrence runs in the workgroup of the voxel that consumes it. The launch
// first copies each connection's gains (from the gains rows, (n, 78)) and
// coefficient into gcons (N, 26, 10 floats of scratch) in its phase-1 order.
// Needs the resident propagation's limits (N < 2^17 - 1, at most
// kResMaxVoxels voxels per workgroup), delays below 1024 and at most 26
// consumers per voxel.
size_t sweep_lds_bytes();
hipError_t sweep_capacity(int32_t *blocks_per_cu);
hipError_t launch_adjoint_sweep_resident(hipStream_t s, int32_t blocks, const int32_t *cons_ptr,
                                         const int32_t *cons_idx, const int32_t *delays,
                                         const float *gains_rows, const float *coefs,
                                         float *gcons, const float *x, const float *E,
                                         float *xbar, GridBarrier *gb, int32_t N, int32_t T,
                                         const AdjointParams &p, uint64_t wait_ticks,
                                         bool allow_coherent);
// Whether a resident launch over T steps of n states exchanges the rows
// coherently (byte offsets of one buffer resource below 2^32) when allowed.
inline bool coherent_exchange(int32_t T, int64_t n) {
    return int64_t(T) * n * 4 < (int64_t(1) << 32);
}
hipError_t launch_barrier_reset(hipStream_t s, GridBarrier *gb, int32_t count);
// One job's barrier counters and its first `nflags` slab flags (flags may be null).
hipError_t launch_resident_reset(hipStream_t s, GridBarrier *gb, uint32_t *flags, int32_t nflags);
// Floats of the gradient pass's chunk-state buffer (`ychunk`) for N voxels,
// T steps and checkpoint interval K.
int64_t adjoint_gradient_state_len(int32_t N, int32_t T, int32_t K);
// ychunk_len: floats the caller allocated at ychunk; fewer than
// adjoint_gradient_state_len(N, T, K) gives hipErrorInvalidValue.
hipError_t launch_adjoint_gradient(hipStream_t s, const int32_t *nbr, const float *coefs,
                                   const int32_t *delays, const float *gT, const float *xT,
                                   const float *xbarT, const float *ckpt, float *ychunk,
                                   int64_t ychunk_len, float *dGT, float *dA, int32_t N, int32_t T,
                                   int32_t K);
// The gradient pass over voxel tiles (the derivation's tables) for
// checkpoints every kTileChunk steps; the same bits as launch_adjoint_gradient
// with K = kTileChunk. Connections with delays above 47 read their rows from
// memory. It reads x̄ by step, (T, n), as the sweep wrote it (the
// lane-per-connection pass reads x̄ᵀ).
hipError_t adjoint_gradient_tiled_prepare();
hipError_t launch_adjoint_gradient_tiled(hipStream_t s, const DeriveTiles &tiles,
                                         const int32_t *nbr, const float *coefs,
                                         const int32_t *delays, const float *gT, const float *xT,
                                         const float *xbar, const float *ckpt, float *dGT,
                                         float *dA, int32_t N, int32_t T);

// Sticky abort words that a parameter update checks before it writes: when
// either is nonzero (a resident launch whose states the update's derivatives
// came from timed out), the update leaves the parameters and moments as they
// are. Null words are not read.
struct UpdateGuard {
    const uint32_t *word0;
    const uint32_t *word1;
};
hipError_t launch_sgd(hipStream_t s, float *v, const float *d, int64_t len, float scale,
                      const UpdateGuard &guard);
// Per-update scalars of the parameter update, read by the kernels from
// device memory so that a captured epoch replays with each epoch's values:
// Adam bias corrections c1 = 1 - 0.9^step, c2 = 1 - 0.999^step and the
// batch-metric slot.
struct UpdateScalars {
    float c1;
    float c2;
    int32_t batch_index;
    int32_t pad;
};
hipError_t launch_adam(hipStream_t s, float *v, const float *d, float *m1, float *m2, int64_t len,
                       const UpdateScalars *sc, float scale, const UpdateGuard &guard);
hipError_t launch_roll_delays(hipStream_t s, float *coefs, int32_t *delays, int64_t len,
                              const UpdateGuard &guard);

hipError_t launch_average_delays(hipStream_t s, const float *gT, const float *coefs,
                                 const int32_t *delays, float *avg, uint8_t *valid, int32_t N);
hipError_t launch_smoothness(hipStream_t s, const int32_t *nbr, const float *avg,
                             const uint8_t *valid, float *dA, float strength, int32_t N);

hipError_t launch_batch_metrics(hipStream_t s, const float *loss, const float *loss_mse,
                                const float *loss_mr, float *loss_b, float *loss_mse_b,
                                float *loss_mr_b, const UpdateScalars *sc, int32_t T);

hipError_t launch_expand_rows(hipStream_t s, const float *in, float *out, int64_t N);

// dst[0 .. len) = src[0 .. len) (a kernel, so it may be captured in a graph).
hipError_t launch_copy(hipStream_t s, float *dst, const float *src, int64_t len);
hipError_t launch_fill(hipStream_t s, float *p, float value, int64_t len);
// Spherical states of x (T, 3N): (T, N) magnitude/theta/phi when not null,
// their values at each voxel's maximum magnitude and the activation time (N);
// *no_max = 1 if a voxel has no non-NaN magnitude.
hipError_t launch_spherical(hipStream_t s, const float *x, int32_t T, int32_t N, float rate,
                            float *mag, float *theta, float *phi, float *mag_max,
                            float *theta_max, float *phi_max, float *act_ms, int32_t *no_max);
// Segmentation metrics over thresholds i / 100, i = 0..100
// (algorithm/metrics.rs:225-241); VoxelType::Pathological's ordinal, as the
// reference's voxel_types.npy stores it (model/spatial/voxels.rs:318, 664-676).
constexpr int kSegmentationThresholds = 101;
constexpr uint32_t kVoxelTypePathological = 6;
struct SegmentationThresholds {
    float value[kSegmentationThresholds];
};
// max_l1 (N): each voxel's maximum over steps of |x0| + |x1| + |x2| (NaN
// skipped; NaN if none). hist (2, kSegmentationThresholds + 1), zeroed here:
// per ground truth class (not Pathological, Pathological) the number of cells
// first predicted Pathological at each threshold index, the last bin for
// cells never predicted Pathological. numbers (cells): -1 or a multiple of 3
// below 3N (checked by the caller).
hipError_t launch_segmentation_histogram(hipStream_t s, const float *x, int32_t T, int32_t N,
                                         const int32_t *numbers, const uint32_t *types,
                                         int64_t cells, const SegmentationThresholds &th,
                                         float *max_l1, uint32_t *hist);
// Connections of a grid of cells: output state indices (3N, 78; -1 = None)
// and delays in samples (N, 26; 0 without a connection); *too_small = 1 if a
// delay is below one sample.
hipError_t launch_model_connections(hipStream_t s, int32_t nx, int32_t ny, int32_t nz,
                                   const int32_t *numbers, const float *pos, const float *vel,
                                   float rate, int32_t num_voxels, int32_t *osi, float *delays,
                                   int32_t *too_small);
// Biot-Savart matrix (B, S, 3N) from voxel positions (N,3), sensor positions
// and orientations (S,3), beat offsets (B,3); `common` = mu0 V / (4 pi) 1e12.
hipError_t launch_measurement_matrix(hipStream_t s, const float *vp, const float *sp,
                                     const float *so, const float *off, float common, int32_t N,
                                     int32_t S, int32_t B, float *out);
// Measurement noise: m[b][t][s] += sd[s] * z_k with k = (s B + b) T + t and
// z_k the k-th standard normal of the counter-based generator keyed by `seed`.
hipError_t launch_measurement_noise(hipStream_t s, float *m, const float *sd, int32_t B,
                                    int32_t T, int32_t S, uint64_t seed);

// FP64 objective for gradient checking (gradcheck.hip). kind 0 = none,
// 1 = gains[index] in the (3N, 78) row-major layout, 2 = coefs[index] (N, 26).
struct Perturbation {
    int32_t kind;
    int32_t pad;
    int64_t index;
    double delta;
};
hipError_t launch_propagate_f64(hipStream_t s, const int32_t *nbr, const float *coefs,
                                const int32_t *delays, const float *gT, double *y, double *x,
                                const float *b, float u_t, int32_t N, int32_t t,
                                const Perturbation &pert);
hipError_t launch_objective_f64(hipStream_t s, const float *H, const double *X,
                                const float *mhat, double *P, int32_t splits, double *terms,
                                double *J, int32_t S, int32_t Sp, int32_t T, int32_t N,
                                double s_mse, double s_reg, double theta);

}  // namespace ct
