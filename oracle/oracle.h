/*
 * CPU oracle for the CardioTrust epoch hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A from-source C++ restatement of the reference's CPU epoch (`run_epoch` and
 * everything it calls), in two precisions: an FP32 variant that follows the
 * reference's operation order (fused multiply-add exactly where the Rust code
 * writes `mul_add`, plain multiply-then-add elsewhere, built with
 * -ffp-contract=off), and an FP64 twin of the same algorithm used to report the
 * absolute error of every FP32 result.
 *
 * Only the test suite, the smoke check and the CPU-baseline leg of the
 * benchmark may load this library, and only as the checker. The product path
 * (libcardiotrust_hip.so) never links or calls it.
 *
 * Array layouts are the reference's row-major ndarray layouts:
 *   gains, output_state_indices, ap_outputs_*, coefs_fir/iir : (3N, 78)
 *   coefs, delays, initial_delays                            : (N, 26)
 *   system_states                                            : (T, 3N)
 *   measurement_matrix                                       : (B, S, 3N)
 *   measurements, actual_measurements                        : (B, T, S)
 * `real` pointers are float* for the *_f32 entry points and double* for *_f64.
 */
#ifndef CT_ORACLE_H
#define CT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t num_voxels;   /* N; states n = 3N */
    int32_t num_sensors;  /* S */
    int32_t num_steps;    /* T */
    int32_t num_beats;    /* B */

    /* model (APParameters, ControlMatrix, ControlFunction, MeasurementMatrix) */
    void *gains;                           /* (n,78)  in/out */
    const int32_t *output_state_indices;   /* (n,78)  -1 = None */
    void *coefs;                           /* (N,26)  in/out */
    int32_t *delays;                       /* (N,26)  in/out */
    const void *initial_delays;            /* (N,26) */
    const void *control_matrix;            /* (n) */
    const void *control_function;          /* (T) */
    const void *measurement_matrix;        /* (B,S,n) */
    const void *actual_measurements;       /* (B,T,S) */

    /* estimations */
    void *ap_outputs_now;                  /* (n,78) */
    void *ap_outputs_last;                 /* (n,78) */
    void *system_states;                   /* (T,n) */
    void *measurements;                    /* (B,T,S) */
    void *residuals;                       /* (S) */
    void *average_delays;                  /* (N) */
    uint8_t *average_delays_valid;         /* (N) 0 = None */

    /* derivatives */
    void *d_gains;                         /* (n,78) */
    void *d_coefs;                         /* (N,26) */
    void *d_coefs_iir;                     /* (n,78) */
    void *d_coefs_fir;                     /* (n,78) */
    void *mapped_residuals;                /* (n) */
    void *maximum_regularization;          /* (n) */
    void *maximum_regularization_sum;      /* (1) */
    int64_t *d_step;                       /* (1) Adam step, starts at 1 */
    void *gains_first_moment;              /* (n,78) Adam only, else NULL */
    void *gains_second_moment;
    void *coefs_first_moment;              /* (N,26) Adam only, else NULL */
    void *coefs_second_moment;

    /* metrics */
    void *loss;                            /* (T) */
    void *loss_mse;                        /* (T) */
    void *loss_maximum_regularization;     /* (T) */
    void *loss_batch;                      /* (num_batch_entries) */
    void *loss_mse_batch;
    void *loss_maximum_regularization_batch;
    int32_t num_batch_entries;
} ct_oracle_problem;

typedef struct {
    float learning_rate;
    int32_t batch_size;                    /* 0 = all beats */
    int32_t optimizer;                     /* 0 = SGD, 1 = Adam */
    float mse_strength;
    float maximum_regularization_strength;
    float maximum_regularization_threshold;
    float difference_regularization_strength;
    float smoothness_regularization_strength;
    int32_t freeze_gains;
    int32_t freeze_delays;
    int32_t ap_derivative;                 /* 0 = Textbook, 1 = Simple */
} ct_oracle_algorithm;

/* One epoch (algorithm.rs:124-281). `beat_order` replaces the reference's
 * unseeded shuffle (NULL = 0..B-1). Returns 0, or <0 on error. */
int ct_oracle_run_epoch_f32(ct_oracle_problem *p, const ct_oracle_algorithm *cfg,
                            const int32_t *beat_order, int32_t *batch_index);
int ct_oracle_run_epoch_f64(ct_oracle_problem *p, const ct_oracle_algorithm *cfg,
                            const int32_t *beat_order, int32_t *batch_index);

/* Forward-only simulation of every beat (simulation.rs:129-150 without noise):
 * fills system_states (last beat) and measurements for all beats. */
int ct_oracle_simulate_f32(ct_oracle_problem *p);
int ct_oracle_simulate_f64(ct_oracle_problem *p);

/* Full reverse-mode gradient of one beat (no reference counterpart): the
 * objective J (double) and dJ/dgains (3N,78), dJ/dcoefs (N,26) in the
 * problem's precision. J = sum_t [ s_mse/2 |H x_t - mhat_t|^2
 * + s_reg/2 sum_v relu(|x_{t,v}|_1 - threshold)^2 ], s_mse = mse_strength/S,
 * s_reg = maximum_regularization_strength. Uses the problem's parameters,
 * control and measurements; state arrays are not touched. */
int ct_oracle_adjoint_f64(ct_oracle_problem *p, const ct_oracle_algorithm *cfg, int32_t beat,
                          double *J, double *d_gains, double *d_coefs);
int ct_oracle_adjoint_f32(ct_oracle_problem *p, const ct_oracle_algorithm *cfg, int32_t beat,
                          double *J, float *d_gains, float *d_coefs);

/* Scalar helpers (allpass.rs:532-668). Index helpers return -1 for None. */
int32_t ct_oracle_offset_to_gain_index(int32_t x, int32_t y, int32_t z, int32_t out);
int32_t ct_oracle_gain_index_to_offset(int32_t gain_index, int32_t *xyz_out4);
int32_t ct_oracle_offset_to_delay_index(int32_t x, int32_t y, int32_t z);
int32_t ct_oracle_delay_index_to_offset(int32_t delay_index, int32_t *xyz3);
float ct_oracle_from_samples_to_coef(float samples);
int64_t ct_oracle_from_samples_to_usize(float samples);
float ct_oracle_from_coef_to_samples(float coef);

/* Piecewise kernels exposed for golden-vector tests. */
int ct_oracle_average_delays_f32(int32_t num_voxels, const float *gains, const float *coefs,
                                 const int32_t *delays, float *average_delays,
                                 uint8_t *valid);
void ct_oracle_update_sgd_f32(float *values, const float *derivatives, int64_t len,
                              float learning_rate, int64_t batch_size);
void ct_oracle_roll_delays_f32(float *coefs, int32_t *delays, int64_t len);
void ct_oracle_mapped_residuals_f32(const float *measurement_matrix, const float *residuals,
                                    int32_t num_sensors, int32_t num_states, float *out);
float ct_oracle_unrolled_dot_f32(const float *a, const float *b, int64_t len);
float ct_oracle_unrolled_sum_f32(const float *a, int64_t len);

/* MeasurementMatrix::from_model_spatial_description (measurement.rs:66-123)
 * in the reference's f32 operation order: out (B, S, 3N), voxel positions
 * (N, 3) by voxel number, sensor positions and orientations (S, 3), per-beat
 * sensor-array offsets (B, 3), all in mm. */
void ct_oracle_measurement_matrix_f32(int32_t num_voxels, const float *voxel_positions_mm,
                                      int32_t num_sensors, const float *sensor_positions_mm,
                                      const float *sensor_orientations, int32_t num_beats,
                                      const float *sensor_offsets_mm, float voxel_size_mm,
                                      float *out);

/* init_output_state_indicies (allpass.rs:212-289) and
 * calculate_delay_samples_array (allpass/delay.rs:35-111) over a grid of
 * nx*ny*nz cells in C order: numbers[c] is the cell's first state index (a
 * multiple of 3) or -1 where it is not connectable, positions_mm (cells, 3),
 * velocities_m_per_s the propagation velocity of the cell's voxel type.
 * Fills output_state_indices (3N, 78; -1 = None) and delay_samples (N, 26;
 * 0 without a connection). Returns -1, as the reference errors, when a delay
 * is below one sample. */
int ct_oracle_model_connections_f32(int32_t nx, int32_t ny, int32_t nz, const int32_t *numbers,
                                    const float *positions_mm, const float *velocities_m_per_s,
                                    float sample_rate_hz, int32_t num_voxels,
                                    int32_t *output_state_indices, float *delay_samples);
/* SystemStatesSpherical / SystemStatesSphericalMax / ActivationTimePerStateMs
 * ::calculate (data/shapes.rs:150-330) over states x (T, 3N): magnitude,
 * theta, phi (T, N); at each voxel's first maximum magnitude (NaN skipped)
 * the three values (N) and the activation time index / rate * 1000 ms (N).
 * Returns -1 where the reference errors (a voxel whose magnitudes are all NaN
 * or T = 0). */
int ct_oracle_spherical_f32(const float *x, int32_t num_steps, int32_t num_voxels,
                            float sample_rate_hz, float *magnitude, float *theta, float *phi,
                            float *magnitude_max, float *theta_max, float *phi_max,
                            float *activation_time_ms);
/* metrics::calculate_final (algorithm/metrics.rs:225-441) over states x
 * (T, n): Dice, IoU, precision and recall (101 each) of the voxel-type
 * predictions at thresholds i / 100 against the ground truth. types: VoxelType
 * ordinals per cell (Pathological = 6); numbers: each cell's first state
 * index, -1 for None (a multiple of 3 below n). */
void ct_oracle_calculate_final_f32(const float *x, int32_t num_steps, int64_t num_states,
                                   const uint32_t *types, const int32_t *numbers, int64_t cells,
                                   float *dice, float *iou, float *precision, float *recall);
/* calculate_delay_s (allpass/delay.rs:14-23). */
float ct_oracle_delay_s_f32(const float *input_position_mm, const float *output_position_mm,
                            float propagation_velocity_m_per_s);

#ifdef __cplusplus
}
#endif
#endif
