/*
 * cardiotrust_hip.h -- C ABI of the MI355X (gfx950) CardioTrust epoch library.
 *
 * Drop-in replacement for the reference's OpenCL backend
 * (the src/core/algorithm/gpu module). Each entry point names the reference
 * interface it replaces (paths relative to the reference root). All pointers
 * are plain host pointers; every array is in the reference's row-major layout:
 *
 *   gains, output_state_indices, ap_outputs_*, coefs_fir/iir   (3N, 78)
 *   coefs, delays, initial_delays                              (N, 26)
 *   control_matrix (3N)   control_function (T)
 *   measurement_matrix (B, S, 3N)   measurements (B, T, S)   system_states (T, 3N)
 *
 * The library copies what it is given and never frees caller memory.
 * Functions return CT_OK (0) or a negative status; ct_last_error() returns the
 * message of the calling thread's last failure. Handles are not thread-safe:
 * use one ct_epoch_kernel per thread, as the reference uses one OpenCL queue
 * per scenario thread (src/scheduler.rs:93, src/core/scenario.rs:807).
 */
#ifndef CARDIOTRUST_HIP_H
#define CARDIOTRUST_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CT_OK 0
#define CT_ERR_INVALID (-1)   /* bad argument, shape or parameter table */
#define CT_ERR_DEVICE (-2)    /* HIP runtime / device failure */
#define CT_ERR_STATE (-3)     /* call out of order, index out of range */

typedef struct ct_gpu ct_gpu;
typedef struct ct_buffer ct_buffer;
typedef struct ct_results_gpu ct_results_gpu;
typedef struct ct_epoch_kernel ct_epoch_kernel;
typedef struct ct_batch ct_batch;

/* Hot-path fields of config::algorithm::Algorithm
 * (src/core/config/algorithm.rs:25-58; defaults 59-85). */
typedef struct {
    float learning_rate;
    int32_t batch_size;                   /* 0 = all beats per update */
    int32_t optimizer;                    /* 0 = Sgd, 1 = Adam (refinement.rs:7-12) */
    float mse_strength;
    float maximum_regularization_strength;
    float maximum_regularization_threshold;
    float difference_regularization_strength;
    float smoothness_regularization_strength;
    int32_t freeze_gains;
    int32_t freeze_delays;
    int32_t ap_derivative;                /* 0 = Textbook, 1 = Simple */
} ct_algorithm;

/* Host arrays uploaded by ct_results_to_gpu: the model part of Results
 * (APParameters allpass.rs:29-36, ControlMatrix/ControlFunction control.rs,
 * MeasurementMatrix measurement.rs:22). */
typedef struct {
    int32_t num_voxels;                   /* N (states 3N); 0 is allowed */
    int32_t num_sensors;                  /* S */
    int32_t num_steps;                    /* T */
    int32_t num_beats;                    /* B */
    int32_t num_batch_entries;            /* length of the *_batch metrics (epochs x batches) */
    int32_t optimizer;                    /* 1 allocates the Adam moments */
    const float *gains;                   /* (3N,78) */
    const int32_t *output_state_indices;  /* (3N,78), -1 = None */
    const float *coefs;                   /* (N,26) */
    const int32_t *delays;                /* (N,26) */
    const float *initial_delays;          /* (N,26) */
    const float *control_matrix;          /* (3N) */
    const float *control_function;        /* (T) */
    const float *measurement_matrix;      /* (B,S,3N) */
} ct_results_desc;

/* Host destinations for ct_results_update_from_gpu; any pointer may be NULL.
 * Mirrors Results::update_from_gpu (scenario/results.rs:136-145) and the
 * per-struct update_from_gpu it calls. */
typedef struct {
    float *gains;                         /* (3N,78) */
    float *coefs;                         /* (N,26) */
    int32_t *delays;                      /* (N,26) */
    float *ap_outputs_now;                /* (3N,78) */
    float *ap_outputs_last;               /* (3N,78) */
    float *system_states;                 /* (T,3N), last processed beat */
    float *measurements;                  /* (B,T,S) */
    float *residuals;                     /* (S), last step */
    float *d_gains;                       /* (3N,78) */
    float *d_coefs;                       /* (N,26) */
    float *d_coefs_iir;                   /* (3N,78) */
    float *d_coefs_fir;                   /* (3N,78) */
    float *mapped_residuals;              /* (3N), last step */
    float *maximum_regularization;        /* (3N), last step */
    float *maximum_regularization_sum;    /* (1) */
    float *average_delays;                /* (N); NaN where None */
    float *loss;                          /* (T) */
    float *loss_mse;                      /* (T) */
    float *loss_maximum_regularization;   /* (T) */
    float *loss_batch;                    /* (num_batch_entries) */
    float *loss_mse_batch;
    float *loss_maximum_regularization_batch;
    int64_t *derivatives_step;            /* (1) Derivatives::step */
} ct_results_host;

/* Per-phase device time of the last ct_epoch_kernel_execute, from HIP events
 * on the library's stream (filled only while profiling is enabled). */
typedef struct {
    double propagate_ms;                  /* sum over propagation launches */
    int64_t propagate_launches;
    int64_t propagate_steps;              /* time steps those launches advanced */
    double project_ms;                    /* H·Xᵀ contraction + residuals + step losses */
    double regularize_ms;                 /* maximum-regularisation sums */
    double transpose_ms;
    double map_residuals_ms;              /* Hᵀ·R contraction */
    double derive_ms;                     /* gain / coefficient derivatives */
    double update_ms;                     /* batch derivatives, update, batch metrics */
    double total_ms;                      /* whole epoch, first to last event */
    double adjoint_sweep_ms;              /* reverse-mode sweep (gradient mode 1) */
} ct_profile;

const char *ct_last_error(void);

/* GPU::new (src/core/algorithm/gpu.rs:31-63): open `device` and create the
 * library's stream (default priority). The handle is reference-counted like
 * the reference's GPU (a Clone of the context that every buffer and kernel
 * holds): every object created on it keeps it alive, so ct_gpu_free may come
 * before or after the frees of those objects; the stream is destroyed with
 * the last of them. */
int ct_gpu_new(int32_t device, ct_gpu **out);
void ct_gpu_free(ct_gpu *gpu);
int ct_gpu_synchronize(ct_gpu *gpu);

/* Measurements::to_gpu (src/core/data/shapes.rs:451-463): (B,T,S) f32. */
int ct_measurements_to_gpu(ct_gpu *gpu, const float *measurements, int32_t num_beats,
                           int32_t num_steps, int32_t num_sensors, ct_buffer **out);
void ct_buffer_free(ct_buffer *buffer);

/* Results::to_gpu (src/core/scenario/results.rs:121-132). Validates that
 * output_state_indices has the 26-neighbour form of allpass.rs:212-289
 * (row 3v+in, column 3d+out holds 3w+out for every in, or -1) and that every
 * connected delay is >= 1. */
int ct_results_to_gpu(ct_gpu *gpu, const ct_results_desc *desc, ct_results_gpu **out);
void ct_results_gpu_free(ct_results_gpu *results);

/* Results::update_from_gpu (src/core/scenario/results.rs:136-145). Blocks. */
int ct_results_update_from_gpu(ct_results_gpu *results, const ct_results_host *out);

/* Simulation::run without noise (src/core/data/simulation.rs:129-150): the
 * forward propagation and H·x of every beat with the uploaded parameters.
 * Fills the device measurements (download with ct_results_update_from_gpu)
 * and, when `measurements_out` is not NULL, copies them there (B,T,S). */
int ct_results_simulate(ct_results_gpu *results, float *measurements_out);

/* MeasurementMatrix::from_model_spatial_description (src/core/model/
 * functional/measurement.rs:66-123) on the device, in the reference's f32
 * operations: out (B, S, 3N) with out[b][s][3v+k] the Biot-Savart factor of
 * voxel v's moment component k at sensor s of beat b. voxel_positions_mm
 * (N, 3) by voxel number, sensor_positions_mm and sensor_orientations (S, 3),
 * sensor_offsets_mm (B, 3) (SensorArrayMotion offsets), voxel_size_mm > 0.
 * Blocks until `out` is written. */
int ct_measurement_matrix(ct_gpu *gpu, int32_t num_voxels, const float *voxel_positions_mm,
                          int32_t num_sensors, const float *sensor_positions_mm,
                          const float *sensor_orientations, int32_t num_beats,
                          const float *sensor_offsets_mm, float voxel_size_mm, float *out);

/* init_output_state_indicies (src/core/model/functional/allpass.rs:212-289)
 * and calculate_delay_samples_array (allpass/delay.rs:35-111) on the device,
 * in the reference's f32 operations. The grid has nx*ny*nz cells in C order;
 * per cell, state_numbers holds the reference's voxel number (the first state
 * index 3v) or -1 where the cell is not connectable, positions_mm its centre
 * (cells, 3) and velocities_m_per_s the propagation velocity of its voxel type.
 * Fills output_state_indices (3N, 78; -1 = None) and delay_samples (N, 26; 0
 * where there is no connection). CT_ERR_INVALID for numbers that are not
 * distinct multiples of 3 below 3N, and, as the reference errors, for a delay
 * below one sample. Blocks until the outputs are written. */
int ct_model_connections(ct_gpu *gpu, int32_t nx, int32_t ny, int32_t nz,
                         const int32_t *state_numbers, const float *positions_mm,
                         const float *velocities_m_per_s, float sample_rate_hz,
                         int32_t num_voxels, int32_t *output_state_indices,
                         float *delay_samples);

/* SystemStatesSpherical, SystemStatesSphericalMax and
 * ActivationTimePerStateMs ::calculate (src/core/data/shapes.rs:150-330) over
 * the device states of the results (the last propagated beat, as the
 * reference's Simulation holds them), without copying the states back:
 * magnitude = |x0| + |x1| + |x2|, theta = acos(x2 / magnitude),
 * phi = atan2(x1, x0), each (T, N); at each voxel's first maximum magnitude
 * (NaN skipped) the three values and the activation time index / rate * 1000
 * ms, each (N). Any output may be NULL (not copied; the (T, N) arrays are then
 * not formed). CT_ERR_INVALID where the reference errors: a voxel whose
 * magnitudes are all NaN, or T = 0. */
int ct_results_spherical(ct_results_gpu *results, float sample_rate_hz, float *magnitude,
                         float *theta, float *phi, float *magnitude_max, float *theta_max,
                         float *phi_max, float *activation_time_ms);

/* metrics::calculate_final (src/core/algorithm/metrics.rs:225-441) over the
 * device states of the results (the last propagated beat, as the reference's
 * Estimations hold them): at threshold i / 100 (i = 0..100) a cell is
 * predicted Pathological when it holds a voxel whose maximum over steps of
 * |x0| + |x1| + |x2| (NaN values skipped, as ndarray-stats max_skipnan) is at
 * most the threshold, and the Dice score, IoU, precision and recall of those
 * predictions against the ground truth are written to dice, iou, precision
 * and recall (101 each; any may be NULL), with the reference's values for
 * empty denominators (0, 0, 0 and 1). The grid arrays (`cells` each) take the
 * reference's voxel_types.npy / voxel_numbers.npy forms: VoxelType ordinals
 * (Pathological = 6) and each cell's first state index, -1 for None. One
 * pass over the states replaces the reference's 101. CT_ERR_INVALID for a
 * voxel number that is neither -1 nor a multiple of 3 below 3N (the reference
 * indexes out of bounds). Blocks until the outputs are written. */
int ct_results_calculate_final_metrics(ct_results_gpu *results, const uint32_t *voxel_types,
                                       const int32_t *voxel_numbers, int64_t cells, float *dice,
                                       float *iou, float *precision, float *recall);

/* ct_results_calculate_final_metrics without any device allocation (no
 * reference counterpart: for callers that cannot allocate device memory,
 * such as code under HIP graph capture, tight serving loops or harnesses
 * that forbid hipMalloc). Same arguments, checks and results, bit for bit.
 * Its device memory is a fixed 8 MiB array that each device holds from the
 * library's first use there until the process ends: N voxels and `cells`
 * cells must satisfy 4 N + 8 cells + 1 KiB (each part rounded up to 256 B)
 * <= 8 MiB, else CT_ERR_INVALID. Calls on one device share that array:
 * calls that overlap on one device race. Blocks. */
int ct_results_calculate_final_metrics_noalloc(ct_results_gpu *results,
                                               const uint32_t *voxel_types,
                                               const int32_t *voxel_numbers, int64_t cells,
                                               float *dice, float *iou, float *precision,
                                               float *recall);

/* Simulation::run with its measurement noise (src/core/data/simulation.rs:
 * 129-173): ct_results_simulate, then m[b][t][s] += sensor_std[s] * z with z
 * standard normal, drawn in the reference's order (sensor, then beat, then
 * step). The reference passes the covariance diagonal entry of each sensor as
 * the standard deviation (simulation.rs:156-160); pass the same S values.
 * CT_ERR_INVALID for a negative or non-finite value, as Normal::new fails.
 * The reference draws z from ChaCha8Rng::seed_from_u64(42) through rand_distr;
 * this library draws it from its own counter-based generator keyed by `seed`
 * (splitmix64 counters, Box-Muller), so the noise is reproducible for a seed
 * but is not the reference's stream. The device measurements keep the noise;
 * `measurements_out` (may be NULL) receives them (B,T,S). */
int ct_results_simulate_noisy(ct_results_gpu *results, const float *sensor_std, uint64_t seed,
                              float *measurements_out);

/* Propagation path (no reference counterpart; a tuning switch of this
 * library): 0 = automatic (resident when the problem fits on chip), 1 = one
 * launch per time step, 2 = resident (one launch per beat whose workgroups
 * all stay resident, keep the filter outputs in LDS and wait after every step
 * for the workgroups owning their neighbours; CT_ERR_INVALID if the problem
 * does not fit), 3 = resident with the state rows exchanged through
 * agent-scope release / acquire fences at every wait, the path that
 * problems with 12 T N bytes of states >= 2^32 take (the others exchange
 * them through agent-coherent stores and loads; both give the same bits).
 * The resident backward sweep of the reverse-mode gradient follows the same
 * setting.
 * `resident` (may be NULL) receives 1 if the resident path will run. */
int ct_results_set_propagation(ct_results_gpu *results, int32_t mode, int32_t *resident);

/* Derivation path (no reference counterpart; a tuning switch of this
 * library): 0 = automatic (over voxel tiles when their tables were built at
 * set-up), 1 = a lane per connection reading its neighbour's states from
 * memory, 2 = over voxel tiles (a workgroup per spatially compact tile stages
 * the states of the tile and its neighbours once and its connections read
 * them from LDS; CT_ERR_INVALID if no tiles were built). Both paths compute
 * the same values in the same order, bit for bit.
 * `tiled` (may be NULL) receives 1 if the tiled path will run. */
int ct_results_set_derivation(ct_results_gpu *results, int32_t mode, int32_t *tiled);

/* Bound on every wait of the resident kernels, in seconds of wall clock
 * (default 2; no reference counterpart: a diagnostic switch of this library).
 * A wait that exceeds it sets the launch's sticky abort word, every workgroup
 * leaves at its next wait, the call returns CT_ERR_DEVICE, and the results
 * object uses the per-step kernels from then on. */
int ct_results_set_wait_limit(ct_results_gpu *results, double seconds);

/* EpochKernel::new (src/core/algorithm/gpu/epoch.rs:29-93). The sizes must
 * match the results; `actual_measurements` must be (B,T,S). */
int ct_epoch_kernel_new(ct_gpu *gpu, ct_results_gpu *results, const ct_buffer *actual_measurements,
                        const ct_algorithm *config, int32_t number_of_states,
                        int32_t number_of_sensors, int32_t number_of_steps,
                        ct_epoch_kernel **out);
void ct_epoch_kernel_free(ct_epoch_kernel *kernel);

/* EpochKernel::execute (gpu/epoch.rs:96-117) with the semantics of the CPU
 * epoch algorithm::run_epoch (src/core/algorithm.rs:124-281): every beat,
 * batch_size, SGD or Adam, all regularisers. Enqueues on the library stream
 * and returns; results are read with ct_results_update_from_gpu. */
int ct_epoch_kernel_execute(ct_epoch_kernel *kernel);

/* Gradient used by ct_epoch_kernel_execute (no reference counterpart):
 * mode 0 = the reference's per-step derivation (derivation.rs:173-229, the
 * default; truncated in time), mode 1 = the exact reverse-mode gradient of
 *   J = sum_t [ mse_strength/(2S) |H x_t - mhat_t|^2
 *             + maximum_regularization_strength/2 sum_v relu(|x_{t,v}|_1 - threshold)^2 ]
 * through the all-pass recursion and the gains, recomputing filter outputs
 * from checkpoints taken every `checkpoint_interval` steps. The difference and
 * smoothness regularisers are not part of J. */
int ct_epoch_kernel_set_gradient(ct_epoch_kernel *kernel, int32_t mode,
                                 int32_t checkpoint_interval);

/* Shared propagation (no reference counterpart). The beats of one update
 * batch (all beats, or batch_size of them) use the same parameters and
 * drive, and the propagation does not depend on the beat
 * (prediction.rs:15-105), so their states are identical: shared, the epoch
 * propagates once per update batch instead of once per beat, with
 * bit-identical results under gradient mode 0. Under gradient mode 1 the
 * batch's summed objective sum_b J_b is differentiated with one backward pass
 * (forcing summed over the beats), instead of one pass per beat: the same
 * gradient up to FP32 rounding. `enabled`: 1 shared, 0 one propagation per
 * beat, -1 automatic (the default): shared under both gradient modes.
 * CT_ERR_INVALID for other values. */
int ct_epoch_kernel_set_share_propagation(ct_epoch_kernel *kernel, int32_t enabled);

/* Forward / backward pair of one beat (the counterpart of an autograd
 * Function over the epoch's parameters; no reference counterpart).
 * ct_results_forward runs the propagation, projection, residuals and per-step
 * losses of `beat` with the current parameters and returns J (above); with
 * checkpoint_interval = K > 0 it also stores the filter outputs every K steps.
 * ct_results_backward then writes dJ/dgains and dJ/dcoefs of that beat into
 * the derivative arrays (read with ct_results_update_from_gpu: d_gains,
 * d_coefs); it needs the same K as the forward pass, or truncate = 1 for the
 * reference's truncated derivation. Neither updates the parameters. */
int ct_results_forward(ct_results_gpu *results, const ct_buffer *actual_measurements, int32_t beat,
                       const ct_algorithm *config, int32_t checkpoint_interval,
                       double *objective);
int ct_results_backward(ct_results_gpu *results, const ct_buffer *actual_measurements,
                        int32_t beat, const ct_algorithm *config, int32_t truncate,
                        int32_t checkpoint_interval);

/* Gradient check (no reference counterpart; the counterpart of
 * torch.autograd.gradcheck): the objective J of ct_results_forward for
 * `beat`, evaluated in FP64 (propagation, projection and sums in double) from
 * the FP32 parameters widened to double, with one parameter shifted by
 * `delta` in FP64: param 0 = none, 1 = gains[index] ((3N, 78) row-major),
 * 2 = coefs[index] ((N, 26) row-major). Central differences of it check
 * ct_results_backward at step sizes below FP32 resolution. Blocks. */
int ct_results_objective_f64(ct_results_gpu *results, const ct_buffer *actual_measurements,
                             int32_t beat, const ct_algorithm *config, int32_t param,
                             int64_t index, double delta, double *objective);

/* EpochKernel::set_freeze_delays / set_freeze_gains (gpu/epoch.rs:118-125). */
int ct_epoch_kernel_set_freeze_delays(ct_epoch_kernel *kernel, int32_t value);
int ct_epoch_kernel_set_freeze_gains(ct_epoch_kernel *kernel, int32_t value);

/* Replaces the unseeded beat shuffle of run_epoch (algorithm.rs:139-141):
 * the beats of the next epochs are visited in `order` (a permutation of
 * 0..B-1). NULL restores 0..B-1. */
int ct_epoch_kernel_set_beat_order(ct_epoch_kernel *kernel, const int32_t *order, int32_t len);

/* The `batch_index` argument of run_epoch (algorithm.rs:126): index of the
 * next *_batch metric entry. */
int ct_epoch_kernel_set_batch_index(ct_epoch_kernel *kernel, int32_t batch_index);
int ct_epoch_kernel_get_batch_index(const ct_epoch_kernel *kernel, int32_t *batch_index);

/* The divergence check of the reference's epoch loop (src/core/scenario.rs:
 * 750-780 and 846-884): the batch losses of the last recorded batch,
 * loss_batch[batch_index - 1] and its MSE and maximum-regularisation parts,
 * read back after the kernel's work; *is_normal = 0 when the loss is not a
 * normal number (NaN, infinite, zero or subnormal: the reference stops its
 * epochs). Any output may be NULL. CT_ERR_INVALID before the first batch. */
int ct_epoch_kernel_last_loss(ct_epoch_kernel *kernel, float *loss, float *loss_mse,
                              float *loss_maximum_regularization, int32_t *is_normal);

/* Per-phase HIP-event timing of ct_epoch_kernel_execute (no reference
 * counterpart; the reference times epochs with Criterion,
 * benches/gpu_epoch_benches.rs:40-53). Reading the profile synchronises. */
int ct_epoch_kernel_set_profiling(ct_epoch_kernel *kernel, int32_t enabled);
int ct_epoch_kernel_get_profile(ct_epoch_kernel *kernel, ct_profile *out);

/* ---- Batching many scenarios (no reference counterpart: the reference runs
 * one scenario per thread, each with its own OpenCL queue, src/scheduler.rs:93).
 *
 * A batch advances every member scenario by one epoch per ct_batch_execute,
 * with exactly the results of calling ct_epoch_kernel_execute on each member
 * (same order of operations per scenario; scenarios are independent). The
 * beats run in rounds (round k = every scenario's k-th beat in its beat
 * order); the propagation of a whole round is one persistent launch driven by
 * a job table (one job per scenario: team of workgroups, voxel slab per
 * workgroup, its arrays). Scenarios whose propagation does not fit the
 * resident kernel use their per-step path inside the same epoch. */
typedef struct {
    int32_t scenarios;
    int32_t rounds;                       /* = largest number of beats */
    int32_t slots;                        /* workgroups of the batched launch (0 = unused) */
    int32_t voxels_per_group;             /* planned voxels per workgroup */
    int32_t jobs;                         /* jobs of the batched launches over all rounds (a
                                             scenario needing more than half the slots runs
                                             its own resident launch instead) */
    int32_t graph;                        /* 1 if the last epoch replayed a captured graph */
    double predicted_propagate_ms;        /* planner's makespan estimate, all rounds */
    double epoch_ms;                      /* device time of the last epoch (HIP events) */
    double propagate_ms;                  /* the last epoch's propagation, all rounds: the
                                             batched launch and the scenarios' own ones
                                             (profiling on, eager epochs only) */
    int64_t propagate_launches;           /* rounds with propagation in that epoch */
    int64_t voxel_steps;                  /* sum over scenarios of N * T * B */
    int32_t tuned_candidates;             /* autotuner: voxels-per-workgroup values timed */
    int32_t tuned_voxels[16];
    double tuned_ms[16];                  /* time of round 0's batched propagation for each */
    int32_t side_streams;                 /* streams the last epoch ran the scenarios' beats on
                                             besides the GPU's own (0 while a reverse-mode
                                             member does not share its backward pass over an
                                             update batch: its sweep would run inside a beat,
                                             and a sweep must have the device to itself) */
    int32_t sweep_jobs;                   /* backward sweeps of the batched sweep launches over
                                             all rounds (reverse-mode members that share one
                                             backward pass per update batch; a scenario needing
                                             more than half the slots sweeps on its own) */
} ct_batch_info;

/* All kernels must live on `gpu`, with distinct results objects. The batch
 * keeps pointers to them: free the batch before its kernels. */
int ct_batch_new(ct_gpu *gpu, ct_epoch_kernel *const *kernels, int32_t count, ct_batch **out);
void ct_batch_free(ct_batch *batch);
/* One epoch of every scenario. Blocks until the epoch is done. */
int ct_batch_execute(ct_batch *batch);
/* Voxels per workgroup of the batched propagation: 0 (default) = autotune
 * by timing round 0 on the first execute; else 1..472. */
int ct_batch_set_voxels_per_group(ct_batch *batch, int32_t voxels);
/* 1 = capture the epoch in a hipGraph after one eager epoch and replay it
 * (re-captured when a member's beat order, configuration or gradient mode
 * changes, and every 16 epochs for reverse-mode members). */
int ct_batch_set_graph(ct_batch *batch, int32_t enabled);
/* Times the batched propagation launches of eager epochs. */
int ct_batch_set_profiling(ct_batch *batch, int32_t enabled);
int ct_batch_get_info(ct_batch *batch, ct_batch_info *out);

#ifdef __cplusplus
}
#endif
#endif
