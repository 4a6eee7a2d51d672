// CPU oracle for the CardioTrust epoch hot path -- TEST INFRASTRUCTURE ONLY.
//
// This file restates the reference's CPU epoch from its source. It is loaded
// by the test suite, the smoke check and the CPU-baseline leg of the
// benchmark, and only as the checker: the product library never links it.
//
// Every function names the reference file:line it follows (paths relative to
// the reference repository root, under src/core/). The FP32 instantiation
// keeps the reference's operation order: std::fma exactly where the Rust code
// writes `mul_add`, separate multiply and add elsewhere. It must be compiled
// with -ffp-contract=off (see Makefile) so the compiler does not fuse them.
//
// Three reductions in the reference are performed by third-party crates that
// are not part of the reference repository (pinned in its Cargo.lock):
//   * ndarray 0.16.1 `numeric_util::unrolled_dot` (Array2.dot(Array1), used
//     by predict_measurements) and `numeric_util::unrolled_fold` (Array.sum /
//     mean, used by the metrics): eight interleaved partial sums p0..p7 over
//     the leading multiple of 8, combined as acc+(p0+p4), +(p1+p5), +(p2+p6),
//     +(p3+p7), then the remaining (<8) elements added in order;
//   * matrixmultiply 0.3.10 `sgemm` (general_mat_mul, used for the mapped
//     residuals Hᵀr): with K = S <= 256 a single packed K-panel; each output
//     is one accumulator updated in k order by fused multiply-add (the FMA
//     microkernel the crate selects on x86-64 hosts with FMA), then scaled by
//     alpha = 1 with beta = 0.
// These orders are restated from the crates' published algorithms; no
// reference build exists here to compare them bit for bit, so results that
// pass through them are compared with a tolerance (parity unpinned bitwise).

#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

template <typename R>
struct Problem {
    int64_t N, n, S, T, B;
    R *gains;
    const int32_t *osi;
    R *coefs;
    int32_t *delays;
    const R *initial_delays;
    const R *control_matrix;
    const R *control_function;
    const R *H;
    const R *mhat;
    R *y_now, *y_last, *x, *m, *r, *avg;
    uint8_t *avg_valid;
    R *dG, *dA, *iir, *fir, *e, *mr, *mr_sum;
    int64_t *step;
    R *gm1, *gm2, *cm1, *cm2;
    R *loss, *loss_mse, *loss_mr, *loss_b, *loss_mse_b, *loss_mr_b;
    int64_t nbatch;
};

template <typename R>
Problem<R> view(const ct_oracle_problem *p) {
    Problem<R> q;
    q.N = p->num_voxels;
    q.n = 3 * q.N;
    q.S = p->num_sensors;
    q.T = p->num_steps;
    q.B = p->num_beats;
    q.gains = static_cast<R *>(p->gains);
    q.osi = p->output_state_indices;
    q.coefs = static_cast<R *>(p->coefs);
    q.delays = p->delays;
    q.initial_delays = static_cast<const R *>(p->initial_delays);
    q.control_matrix = static_cast<const R *>(p->control_matrix);
    q.control_function = static_cast<const R *>(p->control_function);
    q.H = static_cast<const R *>(p->measurement_matrix);
    q.mhat = static_cast<const R *>(p->actual_measurements);
    q.y_now = static_cast<R *>(p->ap_outputs_now);
    q.y_last = static_cast<R *>(p->ap_outputs_last);
    q.x = static_cast<R *>(p->system_states);
    q.m = static_cast<R *>(p->measurements);
    q.r = static_cast<R *>(p->residuals);
    q.avg = static_cast<R *>(p->average_delays);
    q.avg_valid = p->average_delays_valid;
    q.dG = static_cast<R *>(p->d_gains);
    q.dA = static_cast<R *>(p->d_coefs);
    q.iir = static_cast<R *>(p->d_coefs_iir);
    q.fir = static_cast<R *>(p->d_coefs_fir);
    q.e = static_cast<R *>(p->mapped_residuals);
    q.mr = static_cast<R *>(p->maximum_regularization);
    q.mr_sum = static_cast<R *>(p->maximum_regularization_sum);
    q.step = p->d_step;
    q.gm1 = static_cast<R *>(p->gains_first_moment);
    q.gm2 = static_cast<R *>(p->gains_second_moment);
    q.cm1 = static_cast<R *>(p->coefs_first_moment);
    q.cm2 = static_cast<R *>(p->coefs_second_moment);
    q.loss = static_cast<R *>(p->loss);
    q.loss_mse = static_cast<R *>(p->loss_mse);
    q.loss_mr = static_cast<R *>(p->loss_maximum_regularization);
    q.loss_b = static_cast<R *>(p->loss_batch);
    q.loss_mse_b = static_cast<R *>(p->loss_mse_batch);
    q.loss_mr_b = static_cast<R *>(p->loss_maximum_regularization_batch);
    q.nbatch = p->num_batch_entries;
    return q;
}

// ndarray 0.16.1 numeric_util::unrolled_dot (see file header).
template <typename R>
R unrolled_dot(const R *xs, const R *ys, int64_t len) {
    R sum = 0;
    R p0 = 0, p1 = 0, p2 = 0, p3 = 0, p4 = 0, p5 = 0, p6 = 0, p7 = 0;
    int64_t i = 0;
    for (; len - i >= 8; i += 8) {
        p0 = p0 + xs[i + 0] * ys[i + 0];
        p1 = p1 + xs[i + 1] * ys[i + 1];
        p2 = p2 + xs[i + 2] * ys[i + 2];
        p3 = p3 + xs[i + 3] * ys[i + 3];
        p4 = p4 + xs[i + 4] * ys[i + 4];
        p5 = p5 + xs[i + 5] * ys[i + 5];
        p6 = p6 + xs[i + 6] * ys[i + 6];
        p7 = p7 + xs[i + 7] * ys[i + 7];
    }
    sum = sum + (p0 + p4);
    sum = sum + (p1 + p5);
    sum = sum + (p2 + p6);
    sum = sum + (p3 + p7);
    for (int64_t j = 0; i < len && j < 7; ++i, ++j) sum = sum + xs[i] * ys[i];
    return sum;
}

// ndarray 0.16.1 numeric_util::unrolled_fold with f = add (Array::sum).
template <typename R>
R unrolled_sum(const R *xs, int64_t len) {
    R acc = 0;
    R p0 = 0, p1 = 0, p2 = 0, p3 = 0, p4 = 0, p5 = 0, p6 = 0, p7 = 0;
    int64_t i = 0;
    for (; len - i >= 8; i += 8) {
        p0 = p0 + xs[i + 0];
        p1 = p1 + xs[i + 1];
        p2 = p2 + xs[i + 2];
        p3 = p3 + xs[i + 3];
        p4 = p4 + xs[i + 4];
        p5 = p5 + xs[i + 5];
        p6 = p6 + xs[i + 6];
        p7 = p7 + xs[i + 7];
    }
    acc = acc + (p0 + p4);
    acc = acc + (p1 + p5);
    acc = acc + (p2 + p6);
    acc = acc + (p3 + p7);
    for (int64_t j = 0; i < len && j < 7; ++i, ++j) acc = acc + xs[i];
    return acc;
}

// Rust f32::signum: NaN stays NaN, otherwise copysign(1, x) (so -0 -> -1).
template <typename R>
R signum(R v) {
    if (std::isnan(v)) return v;
    return std::copysign(R(1), v);
}

// Integer power as LLVM expands powi with a constant exponent: binary
// exponentiation, Res = x at the lowest set bit, then Res *= x^(2^k).
template <typename R>
R powi5(R v) {
    R sq = v * v;
    R q = sq * sq;
    return v * q;
}

// allpass.rs:665-668
template <typename R>
R coef_to_samples(R coef) {
    return (R(1) - coef) / (coef + R(1));
}

// estimation.rs:82-86
template <typename R>
void estimations_reset(Problem<R> &q) {
    std::memset(q.x, 0, sizeof(R) * q.T * q.n);
    std::memset(q.y_now, 0, sizeof(R) * q.n * 78);
}

// derivation.rs:114-122
template <typename R>
void derivatives_reset(Problem<R> &q) {
    std::memset(q.dG, 0, sizeof(R) * q.n * 78);
    std::memset(q.dA, 0, sizeof(R) * q.N * 26);
    std::memset(q.iir, 0, sizeof(R) * q.n * 78);
    std::memset(q.fir, 0, sizeof(R) * q.n * 78);
    std::memset(q.mr, 0, sizeof(R) * q.n);
    *q.mr_sum = 0;
}

// prediction.rs:40-87 (innovate_system_states_v1)
template <typename R>
void innovate(Problem<R> &q, int64_t t) {
    std::memcpy(q.y_last, q.y_now, sizeof(R) * q.n * 78);
    R *xt = q.x + t * q.n;
    for (int64_t i = 0; i < q.n; ++i) {
        for (int64_t g = 0; g < 78; ++g) {
            const int32_t o = q.osi[i * 78 + g];
            if (o < 0) continue;
            const int64_t ci = (i / 3) * 26 + g / 3;
            const R a = q.coefs[ci];
            const int64_t D = q.delays[ci];
            const R in0 = (D <= t) ? q.x[(t - D) * q.n + o] : R(0);
            const R in1 = (D < t) ? q.x[(t - D - 1) * q.n + o] : R(0);
            R &y = q.y_now[i * 78 + g];
            y = std::fma(a, in0 - y, in1);
            xt[i] += q.gains[i * 78 + g] * y;
        }
    }
}

// prediction.rs:94-105 (ndarray scaled_add: x = x + alpha * b)
template <typename R>
void add_control(Problem<R> &q, int64_t t) {
    R *xt = q.x + t * q.n;
    const R u = q.control_function[t];
    for (int64_t i = 0; i < q.n; ++i) xt[i] = xt[i] + u * q.control_matrix[i];
}

// prediction.rs:112-130 (Array2.dot(Array1) with alpha = 1, beta = 0)
template <typename R>
void predict_measurements(Problem<R> &q, int64_t beat, int64_t t) {
    const R *Hb = q.H + beat * q.S * q.n;
    const R *xt = q.x + t * q.n;
    R *mt = q.m + (beat * q.T + t) * q.S;
    for (int64_t s = 0; s < q.S; ++s) mt[s] = unrolled_dot(Hb + s * q.n, xt, q.n) * R(1);
}

// prediction.rs:15-26
template <typename R>
void system_prediction(Problem<R> &q, int64_t beat, int64_t t) {
    innovate(q, t);
    add_control(q, t);
    predict_measurements(q, beat, t);
}

// estimation.rs:146-152
template <typename R>
void residuals(Problem<R> &q, int64_t beat, int64_t t) {
    const R *mt = q.m + (beat * q.T + t) * q.S;
    const R *at = q.mhat + (beat * q.T + t) * q.S;
    for (int64_t s = 0; s < q.S; ++s) q.r[s] = mt[s] - at[s];
}

// derivation.rs:538-551 (general_mat_mul(1, Hᵀ, r, 0, e); see file header)
template <typename R>
void mapped_residuals(const R *Hb, const R *r, int64_t S, int64_t n, R *e) {
    for (int64_t i = 0; i < n; ++i) {
        R acc = 0;
        for (int64_t s = 0; s < S; ++s) acc = std::fma(Hb[s * n + i], r[s], acc);
        e[i] = acc;
    }
}

// derivation.rs:509-535
template <typename R>
void maximum_regularization(Problem<R> &q, int64_t t, R threshold) {
    const R *xt = q.x + t * q.n;
    for (int64_t i = 0; i < q.n; i += 3) {
        const R sum = std::fabs(xt[i]) + std::fabs(xt[i + 1]) + std::fabs(xt[i + 2]);
        if (sum > threshold) {
            const R f = sum - threshold;
            *q.mr_sum += f * f;
            q.mr[i] = f * signum(xt[i]);
            q.mr[i + 1] = f * signum(xt[i + 1]);
            q.mr[i + 2] = f * signum(xt[i + 2]);
        } else {
            q.mr[i] = 0;
            q.mr[i + 1] = 0;
            q.mr[i + 2] = 0;
        }
    }
}

// derivation.rs:307-329
template <typename R>
void derivatives_gains(Problem<R> &q, R mse_scaling, R reg_scaling) {
    for (int64_t i = 0; i < q.n; ++i) {
        const R c = std::fma(q.e[i], mse_scaling, q.mr[i] * reg_scaling);
        for (int64_t g = 0; g < 78; ++g) q.dG[i * 78 + g] += q.y_now[i * 78 + g] * c;
    }
}

// derivation.rs:408-501 (Textbook)
template <typename R>
void derivatives_coefs_textbook(Problem<R> &q, int64_t t, R mse_scaling, R diff_strength) {
    // FIR (418-445)
    for (int64_t i = 0; i < q.n; ++i) {
        for (int64_t g = 0; g < 78; ++g) {
            const int32_t o = q.osi[i * 78 + g];
            if (o < 0) continue;
            const int64_t ci = (i / 3) * 26 + g / 3;
            const int64_t D = q.delays[ci];
            const R a = q.coefs[ci];
            if (t >= D) {
                R &f = q.fir[i * 78 + g];
                f = std::fma(-a, f, q.x[(t - D) * q.n + o]);
            }
        }
    }
    // IIR (448-465)
    for (int64_t i = 0; i < q.n; ++i) {
        for (int64_t g = 0; g < 78; ++g) {
            const int64_t ci = (i / 3) * 26 + g / 3;
            const int64_t D = q.delays[ci];
            const R a = q.coefs[ci];
            if (t >= D) {
                R &h = q.iir[i * 78 + g];
                h = std::fma(-a, h, q.y_last[i * 78 + g]);
            }
        }
    }
    // Combine (468-499)
    for (int64_t i = 0; i < q.n; ++i) {
        for (int64_t g = 0; g < 78; ++g) {
            const int64_t ci = (i / 3) * 26 + g / 3;
            const R delay = R(q.delays[ci]) + coef_to_samples(q.coefs[ci]);
            const R delta = powi5(q.initial_delays[ci] - delay);
            const R iir = q.iir[i * 78 + g];
            const R fir = q.fir[i * 78 + g];
            const R G = q.gains[i * 78 + g];
            const R e = q.e[i];
            q.dA[ci] += std::fma((fir - iir) * G * e, mse_scaling, diff_strength * delta);
        }
    }
}

// derivation.rs:338-398 (Simple)
template <typename R>
void derivatives_coefs_simple(Problem<R> &q, int64_t t, R mse_scaling, R diff_strength) {
    for (int64_t i = 0; i < q.n; ++i) {
        for (int64_t g = 0; g < 78; ++g) {
            const int64_t ci = (i / 3) * 26 + g / 3;
            const R delay_f = R(q.delays[ci]) + coef_to_samples(q.coefs[ci]);
            const R delta = powi5(q.initial_delays[ci] - delay_f);
            const int64_t D = q.delays[ci];
            const int32_t o = q.osi[i * 78 + g];
            if (o < 0) continue;
            if (t >= D) {
                const R yl = q.y_last[i * 78 + g];
                const R xv = q.x[(t - D) * q.n + o];
                const R G = q.gains[i * 78 + g];
                const R e = q.e[i];
                q.dA[ci] += std::fma((xv - yl) * G * e, mse_scaling, diff_strength * delta);
            }
        }
    }
}

// derivation.rs:173-229
template <typename R>
void step_derivatives(Problem<R> &q, const ct_oracle_algorithm &c, int64_t t, int64_t beat) {
    mapped_residuals(q.H + beat * q.S * q.n, q.r, q.S, q.n, q.e);
    maximum_regularization(q, t, R(c.maximum_regularization_threshold));
    const R mse_scaling = R(1) / R(q.S) * R(c.mse_strength);
    if (!c.freeze_gains)
        derivatives_gains(q, mse_scaling, R(c.maximum_regularization_strength));
    if (!c.freeze_delays) {
        if (c.ap_derivative == 1)
            derivatives_coefs_simple(q, t, mse_scaling, R(c.difference_regularization_strength));
        else
            derivatives_coefs_textbook(q, t, mse_scaling, R(c.difference_regularization_strength));
    }
}

// derivation.rs:554-594
template <typename R>
void average_delays(int64_t N, const R *gains, const R *coefs, const int32_t *delays, R *avg,
                    uint8_t *valid) {
    for (int64_t v = 0; v < N; ++v) {
        R delay_sum = 0, gain_sum = 0;
        for (int64_t d = 0; d < 26; ++d) {
            int32_t off[3];
            ct_oracle_delay_index_to_offset(int32_t(d), off);
            const R xyz = R(std::abs(off[0]) + std::abs(off[1]) + std::abs(off[2]));
            const R delay = R(delays[v * 26 + d]) + coef_to_samples(coefs[v * 26 + d]);
            const R corrected = delay / std::sqrt(xyz);
            for (int64_t in = 0; in < 3; ++in)
                for (int64_t out = 0; out < 3; ++out) {
                    const R gain = gains[(v * 3 + in) * 78 + d * 3 + out];
                    delay_sum += std::fabs(gain) * corrected;
                    gain_sum += std::fabs(gain);
                }
        }
        if (gain_sum == R(0)) {
            valid[v] = 0;
            avg[v] = 0;
        } else {
            valid[v] = 1;
            avg[v] = delay_sum / gain_sum;
        }
    }
}

// derivation.rs:260-302
template <typename R>
void smoothness_derivatives(Problem<R> &q, R strength) {
    for (int64_t v = 0; v < q.N; ++v) {
        for (int64_t od = 0; od < 26; ++od) {
            if (!q.avg_valid[v]) continue;
            const R own = q.avg[v];
            R hood = own;
            R divisor = 1;
            for (int64_t d = 0; d < 26; ++d) {
                const int32_t ni = q.osi[(v * 3) * 78 + d * 3];
                if (ni < 0) continue;
                const int64_t w = ni / 3;
                if (q.avg_valid[w]) {
                    hood += q.avg[w];
                    divisor += 1;
                }
            }
            hood /= divisor;
            const R diff = hood - own;
            q.dA[v * 26 + od] += strength * diff;
        }
    }
}

// derivation.rs:241-256 (abs_diff_ne(0, f32::EPSILON))
template <typename R>
void batch_derivatives(Problem<R> &q, const ct_oracle_algorithm &c) {
    if (!c.freeze_delays &&
        std::fabs(c.smoothness_regularization_strength - 0.0f) > 1.1920929e-7f)
        smoothness_derivatives(q, R(c.smoothness_regularization_strength));
}

// update.rs:106-114 / 158-168: v = v - (lr / bs) * d
template <typename R>
void update_sgd(R *v, const R *d, int64_t len, R lr, int64_t bs) {
    const R s = lr / R(bs);
    for (int64_t k = 0; k < len; ++k) v[k] = v[k] - s * d[k];
}

// update.rs:119-146 / 173-200
template <typename R>
void update_adam(R *v, const R *d, R *m1, R *m2, int64_t len, int64_t step, R lr, int64_t bs) {
    const R beta1 = R(0.9f), beta2 = R(0.999f), eps = R(1e-8f);
    const R omb1 = R(1) - beta1, omb2 = R(1) - beta2;
    const R c1 = R(1) - std::pow(beta1, R(step));
    const R c2 = R(1) - std::pow(beta2, R(step));
    const R s = lr / R(bs);
    for (int64_t k = 0; k < len; ++k) {
        m1[k] = m1[k] * beta1 + omb1 * d[k];
        m2[k] = m2[k] * beta2 + omb2 * d[k] * d[k];
        const R m1c = m1[k] / c1;
        const R m2c = m2[k] / c2;
        const R factor = m1c / (std::sqrt(m2c) + eps);
        v[k] = v[k] - s * factor;
    }
}

// update.rs:206-228
template <typename R>
void roll_delays(R *coefs, int32_t *delays, int64_t len) {
    const R margin = R(1e-4f);
    for (int64_t k = 0; k < len; ++k) {
        R &a = coefs[k];
        int32_t &D = delays[k];
        if (a > R(1) - margin) {
            if (D > 1) {
                a = R(2) * margin;
                D -= 1;
            } else {
                a = R(1) - margin;
            }
        } else if (a < margin) {
            if (D < 1000) {
                a = std::fma(R(2), -margin, R(1));
                D += 1;
            } else {
                a = margin;
            }
        }
    }
}

// update.rs:27-97
template <typename R>
int update(Problem<R> &q, const ct_oracle_algorithm &c, int64_t num_beats) {
    const int64_t bs = c.batch_size == 0 ? q.T * num_beats : q.T * int64_t(c.batch_size);
    const R lr = R(c.learning_rate);
    if (!c.freeze_gains) {
        if (c.optimizer == 1) {
            if (!q.gm1 || !q.gm2) return -2;
            update_adam(q.gains, q.dG, q.gm1, q.gm2, q.n * 78, *q.step, lr, bs);
        } else {
            update_sgd(q.gains, q.dG, q.n * 78, lr, bs);
        }
    }
    if (!c.freeze_delays) {
        if (c.optimizer == 1) {
            if (!q.cm1 || !q.cm2) return -2;
            update_adam(q.coefs, q.dA, q.cm1, q.cm2, q.N * 26, *q.step, lr, bs);
        } else {
            update_sgd(q.coefs, q.dA, q.N * 26, lr, bs);
        }
        roll_delays(q.coefs, q.delays, q.N * 26);
    }
    *q.step += 1;
    return 0;
}

// metrics.rs:180-196
template <typename R>
void metrics_step(Problem<R> &q, R reg_strength, int64_t t) {
    std::vector<R> sq(q.S);
    for (int64_t s = 0; s < q.S; ++s) sq[s] = q.r[s] * q.r[s];
    q.loss_mse[t] = unrolled_sum(sq.data(), q.S) / R(q.S);
    q.loss_mr[t] = *q.mr_sum;
    q.loss[t] = std::fma(reg_strength, q.loss_mr[t], q.loss_mse[t]);
}

// metrics.rs:204-219
template <typename R>
int metrics_batch(Problem<R> &q, int64_t idx) {
    if (idx < 0 || idx >= q.nbatch) return -3;
    q.loss_mse_b[idx] = unrolled_sum(q.loss_mse, q.T) / R(q.T);
    q.loss_mr_b[idx] = unrolled_sum(q.loss_mr, q.T) / R(q.T);
    q.loss_b[idx] = unrolled_sum(q.loss, q.T) / R(q.T);
    return 0;
}

template <typename R>
int check(const Problem<R> &q) {
    if (q.N < 0 || q.S <= 0 || q.T <= 0 || q.B <= 0) return -1;
    for (int64_t k = 0; k < q.N * 26; ++k)
        if (q.delays[k] < 0) return -1;
    return 0;
}

// algorithm.rs:124-281
template <typename R>
int run_epoch(const ct_oracle_problem *pp, const ct_oracle_algorithm *cfg, const int32_t *order,
              int32_t *batch_index) {
    Problem<R> q = view<R>(pp);
    if (check(q)) return -1;
    const ct_oracle_algorithm &c = *cfg;
    derivatives_reset(q);
    const bool batched = c.batch_size != 0;
    int64_t n_in_batch = 0;
    for (int64_t k = 0; k < q.B; ++k) {
        const int64_t beat = order ? order[k] : k;
        if (beat < 0 || beat >= q.B) return -1;
        estimations_reset(q);
        for (int64_t t = 0; t < q.T; ++t) {
            system_prediction(q, beat, t);
            residuals(q, beat, t);
            step_derivatives(q, c, t, beat);
            metrics_step(q, R(c.maximum_regularization_strength), t);
        }
        if (batched) {
            n_in_batch += 1;
            if (n_in_batch == c.batch_size) {
                average_delays(q.N, q.gains, q.coefs, q.delays, q.avg, q.avg_valid);
                batch_derivatives(q, c);
                if (int rc = update(q, c, n_in_batch)) return rc;
                derivatives_reset(q);
                n_in_batch = 0;
                if (int rc = metrics_batch(q, *batch_index)) return rc;
                *batch_index += 1;
            }
        }
    }
    if (batched) {
        if (n_in_batch > 0) {
            average_delays(q.N, q.gains, q.coefs, q.delays, q.avg, q.avg_valid);
            batch_derivatives(q, c);
            if (int rc = update(q, c, n_in_batch)) return rc;
            if (int rc = metrics_batch(q, *batch_index)) return rc;
            *batch_index += 1;
        }
    } else {
        average_delays(q.N, q.gains, q.coefs, q.delays, q.avg, q.avg_valid);
        batch_derivatives(q, c);
        if (int rc = update(q, c, q.B)) return rc;
        if (int rc = metrics_batch(q, *batch_index)) return rc;
        *batch_index += 1;
    }
    return 0;
}

// simulation.rs:129-150 (noise-free)
template <typename R>
int simulate(const ct_oracle_problem *pp) {
    Problem<R> q = view<R>(pp);
    if (check(q)) return -1;
    for (int64_t beat = 0; beat < q.B; ++beat) {
        estimations_reset(q);
        for (int64_t t = 0; t < q.T; ++t) system_prediction(q, beat, t);
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Full reverse-mode gradient of one beat (the build's own derivation; no
// reference counterpart). Objective, with s_mse = mse_strength / S and
// s_reg = maximum_regularization_strength:
//   J = sum_t [ s_mse/2 |H x_t - mhat_t|^2 + s_reg/2 sum_v relu(|x_{t,v}|_1 - theta)^2 ]
// whose partial derivative in x_t is the reference's per-step factor
// c_t = s_mse H^T r_t + s_reg mr_t (derivation.rs:315-326). Adjoint recursion,
// for connection (v,d) reading neighbour w with delay D and coefficient a:
//   xbar_t[3w+o] = c_t[3w+o] + sum_{(v,d): nbr(v,d)=w} (a ybar_{t+D}[v,d,o] + ybar_{t+D+1}[v,d,o])
//   ybar_t[v,d,o] = sum_i G[3v+i,3d+o] xbar_t[3v+i] - a ybar_{t+1}[v,d,o]
//   dJ/dG[3v+i,3d+o] = sum_t xbar_t[3v+i] y_t[v,d,o]
//   dJ/da[v,d] = sum_t sum_o ybar_t[v,d,o] (xi_{t-D}[o] - y_{t-1}[v,d,o])
// ---------------------------------------------------------------------------
template <typename R>
int adjoint(const ct_oracle_problem *pp, const ct_oracle_algorithm *c, int32_t beat, double *J_out,
            R *dG, R *dA) {
    Problem<R> q = view<R>(pp);
    if (check(q) || beat < 0 || beat >= q.B) return -1;
    const int64_t N = q.N, n = q.n, T = q.T, S = q.S;
    const R s_mse = R(1) / R(S) * R(c->mse_strength);
    const R s_reg = R(c->maximum_regularization_strength);
    const R theta = R(c->maximum_regularization_threshold);
    // forward, keeping every filter output (one per (v, d, out))
    std::vector<R> x(T * n, R(0)), y((T + 1) * N * 78, R(0));  // y[t+1] = y_t, y[0] = y_{-1} = 0
    for (int64_t t = 0; t < T; ++t) {
        R *yt = y.data() + (t + 1) * N * 78;
        const R *yp = y.data() + t * N * 78;
        for (int64_t v = 0; v < N; ++v)
            for (int64_t d = 0; d < 26; ++d) {
                const int32_t first = q.osi[(3 * v) * 78 + 3 * d];
                if (first < 0) continue;
                const R a = q.coefs[v * 26 + d];
                const int64_t D = q.delays[v * 26 + d];
                for (int o = 0; o < 3; ++o) {
                    const R i0 = D <= t ? x[(t - D) * n + first + o] : R(0);
                    const R i1 = D < t ? x[(t - D - 1) * n + first + o] : R(0);
                    yt[(v * 26 + d) * 3 + o] = std::fma(a, i0 - yp[(v * 26 + d) * 3 + o], i1);
                }
            }
        for (int64_t i = 0; i < n; ++i) {
            const int64_t v = i / 3;
            R acc = 0;
            for (int64_t g = 0; g < 78; ++g) {
                if (q.osi[i * 78 + g] < 0) continue;
                acc += q.gains[i * 78 + g] * yt[v * 78 + g];
            }
            x[t * n + i] = acc + q.control_function[t] * q.control_matrix[i];
        }
    }
    // objective and per-step factors c_t
    const R *Hb = q.H + beat * S * n;
    const R *mh = q.mhat + beat * T * S;
    std::vector<R> cst(T * n), r(S);
    double J = 0;
    for (int64_t t = 0; t < T; ++t) {
        for (int64_t s = 0; s < S; ++s) {
            R m = 0;
            for (int64_t i = 0; i < n; ++i) m += Hb[s * n + i] * x[t * n + i];
            r[s] = m - mh[t * S + s];
            J += 0.5 * double(s_mse) * double(r[s]) * double(r[s]);
        }
        for (int64_t i = 0; i < n; ++i) {
            R e = 0;
            for (int64_t s = 0; s < S; ++s) e += Hb[s * n + i] * r[s];
            cst[t * n + i] = e * s_mse;
        }
        for (int64_t v = 0; v < N; ++v) {
            const R *xv = x.data() + t * n + 3 * v;
            const R sum = std::fabs(xv[0]) + std::fabs(xv[1]) + std::fabs(xv[2]);
            if (sum > theta) {
                const R f = sum - theta;
                J += 0.5 * double(s_reg) * double(f) * double(f);
                for (int k = 0; k < 3; ++k) cst[t * n + 3 * v + k] += s_reg * f * signum(xv[k]);
            }
        }
    }
    // backward
    std::vector<R> xbar(n), ybar((T + 1) * N * 78, R(0));  // ybar[t] = ybar_t, ybar[T] = 0
    for (int64_t k = 0; k < n * 78; ++k) dG[k] = 0;
    for (int64_t k = 0; k < N * 26; ++k) dA[k] = 0;
    for (int64_t t = T - 1; t >= 0; --t) {
        for (int64_t i = 0; i < n; ++i) xbar[i] = cst[t * n + i];
        for (int64_t v = 0; v < N; ++v)
            for (int64_t d = 0; d < 26; ++d) {
                const int32_t first = q.osi[(3 * v) * 78 + 3 * d];
                if (first < 0) continue;
                const R a = q.coefs[v * 26 + d];
                const int64_t D = q.delays[v * 26 + d];
                for (int o = 0; o < 3; ++o) {
                    R contrib = 0;
                    if (t + D < T) contrib += a * ybar[(t + D) * N * 78 + (v * 26 + d) * 3 + o];
                    if (t + D + 1 < T) contrib += ybar[(t + D + 1) * N * 78 + (v * 26 + d) * 3 + o];
                    xbar[first + o] += contrib;
                }
            }
        for (int64_t v = 0; v < N; ++v)
            for (int64_t d = 0; d < 26; ++d) {
                const int32_t first = q.osi[(3 * v) * 78 + 3 * d];
                if (first < 0) continue;
                const R a = q.coefs[v * 26 + d];
                const int64_t D = q.delays[v * 26 + d];
                for (int o = 0; o < 3; ++o) {
                    R yb = -a * ybar[(t + 1) * N * 78 + (v * 26 + d) * 3 + o];
                    for (int in = 0; in < 3; ++in)
                        yb += q.gains[(3 * v + in) * 78 + 3 * d + o] * xbar[3 * v + in];
                    ybar[t * N * 78 + (v * 26 + d) * 3 + o] = yb;
                    const R yt = y[(t + 1) * N * 78 + (v * 26 + d) * 3 + o];
                    const R ytm1 = y[t * N * 78 + (v * 26 + d) * 3 + o];
                    const R xi = D <= t ? x[(t - D) * n + first + o] : R(0);
                    for (int in = 0; in < 3; ++in)
                        dG[(3 * v + in) * 78 + 3 * d + o] += xbar[3 * v + in] * yt;
                    dA[v * 26 + d] += yb * (xi - ytm1);
                }
            }
    }
    *J_out = J;
    return 0;
}

}  // namespace

extern "C" {

int ct_oracle_adjoint_f64(ct_oracle_problem *p, const ct_oracle_algorithm *cfg, int32_t beat,
                          double *J, double *d_gains, double *d_coefs) {
    return adjoint<double>(p, cfg, beat, J, d_gains, d_coefs);
}
int ct_oracle_adjoint_f32(ct_oracle_problem *p, const ct_oracle_algorithm *cfg, int32_t beat,
                          double *J, float *d_gains, float *d_coefs) {
    return adjoint<float>(p, cfg, beat, J, d_gains, d_coefs);
}

int ct_oracle_run_epoch_f32(ct_oracle_problem *p, const ct_oracle_algorithm *cfg,
                            const int32_t *beat_order, int32_t *batch_index) {
    return run_epoch<float>(p, cfg, beat_order, batch_index);
}
int ct_oracle_run_epoch_f64(ct_oracle_problem *p, const ct_oracle_algorithm *cfg,
                            const int32_t *beat_order, int32_t *batch_index) {
    return run_epoch<double>(p, cfg, beat_order, batch_index);
}
int ct_oracle_simulate_f32(ct_oracle_problem *p) { return simulate<float>(p); }
int ct_oracle_simulate_f64(ct_oracle_problem *p) { return simulate<double>(p); }

// allpass.rs:532-549
int32_t ct_oracle_offset_to_gain_index(int32_t x, int32_t y, int32_t z, int32_t out) {
    if (x == 0 && y == 0 && z == 0) return -1;
    int32_t index = out + (z + 1) * 3 + (y + 1) * 9 + (x + 1) * 27;
    if (index > 27 + 9 + 3) index -= 3;
    return index;
}

// allpass.rs:560-576
int32_t ct_oracle_gain_index_to_offset(int32_t g, int32_t *xyz_out4) {
    if (g < 0 || g > 77) return -1;
    const int32_t c = g >= 27 + 9 + 3 ? g + 3 : g;
    xyz_out4[3] = c % 3;
    xyz_out4[2] = (c / 3) % 3 - 1;
    xyz_out4[1] = (c / 9) % 3 - 1;
    xyz_out4[0] = (c / 27) % 3 - 1;
    return 0;
}

// allpass.rs:582-592
int32_t ct_oracle_offset_to_delay_index(int32_t x, int32_t y, int32_t z) {
    if (x == 0 && y == 0 && z == 0) return -1;
    int32_t index = (z + 1) + (y + 1) * 3 + (x + 1) * 9;
    if (index > 9 + 3 + 1) index -= 1;
    return index;
}

// allpass.rs:603-618
int32_t ct_oracle_delay_index_to_offset(int32_t d, int32_t *xyz3) {
    if (d < 0 || d > 26) return -1;
    const int32_t c = d > 9 + 3 ? d + 1 : d;
    xyz3[2] = c % 3 - 1;
    xyz3[1] = (c / 3) % 3 - 1;
    xyz3[0] = (c / 9) % 3 - 1;
    return 0;
}

// allpass.rs:644-649 (Rust `%` on f32 is fmod; clamp to [1e-4, 1-1e-4])
float ct_oracle_from_samples_to_coef(float samples) {
    const float fractional = std::fmod(samples, 1.0f);
    const float coef = (1.0f - fractional) / (1.0f + fractional);
    const float lo = 1e-4f, hi = 1.0f - 1e-4f;
    return coef < lo ? lo : (coef > hi ? hi : coef);
}

// allpass.rs:658-660 (saturating float->usize cast)
int64_t ct_oracle_from_samples_to_usize(float samples) {
    if (!(samples > 0.0f)) return 0;
    return int64_t(samples);
}

float ct_oracle_from_coef_to_samples(float coef) { return coef_to_samples(coef); }

int ct_oracle_average_delays_f32(int32_t num_voxels, const float *gains, const float *coefs,
                                 const int32_t *delays, float *avg, uint8_t *valid) {
    average_delays<float>(num_voxels, gains, coefs, delays, avg, valid);
    return 0;
}

void ct_oracle_update_sgd_f32(float *values, const float *derivatives, int64_t len,
                              float learning_rate, int64_t batch_size) {
    update_sgd<float>(values, derivatives, len, learning_rate, batch_size);
}

void ct_oracle_roll_delays_f32(float *coefs, int32_t *delays, int64_t len) {
    roll_delays<float>(coefs, delays, len);
}

void ct_oracle_mapped_residuals_f32(const float *H, const float *r, int32_t S, int32_t n,
                                    float *out) {
    mapped_residuals<float>(H, r, S, n, out);
}

float ct_oracle_unrolled_dot_f32(const float *a, const float *b, int64_t len) {
    return unrolled_dot<float>(a, b, len);
}

float ct_oracle_unrolled_sum_f32(const float *a, int64_t len) { return unrolled_sum<float>(a, len); }

// measurement.rs:66-123. Every operation is the reference's f32 operation:
// voxel volume (size/1000).powi(3); common factor (mu0 as f32 * V) / (4 pi)
// * 1e12 with mu0 = 1.25663706212e-6 (physical_constants 0.5.0, CODATA 2018);
// distance (sensor + offset - voxel) / 1000; |d|^3 = sqrt of the ndarray sum
// ((0 + d0^2) + d1^2) + d2^2, then powi(3) = |d| * (|d| * |d|); each entry
// common_factor * o_a.mul_add(d_b, -o_b * d_c) / |d|^3.
void ct_oracle_measurement_matrix_f32(int32_t num_voxels, const float *voxel_positions_mm,
                                      int32_t num_sensors, const float *sensor_positions_mm,
                                      const float *sensor_orientations, int32_t num_beats,
                                      const float *sensor_offsets_mm, float voxel_size_mm,
                                      float *out) {
    const float pi = 3.14159265358979323846f;
    const float side = voxel_size_mm / 1000.0f;
    const float volume = side * (side * side);
    const float common = (float(1.25663706212e-6) * volume) / (4.0f * pi) * 1e12f;
    const int64_t n = 3LL * num_voxels;
    for (int32_t b = 0; b < num_beats; ++b)
        for (int32_t v = 0; v < num_voxels; ++v)
            for (int32_t s = 0; s < num_sensors; ++s) {
                float d[3];
                for (int k = 0; k < 3; ++k) {
                    const float sp = sensor_positions_mm[3 * s + k] + sensor_offsets_mm[3 * b + k];
                    d[k] = (sp - voxel_positions_mm[3 * int64_t(v) + k]) / 1000.0f;
                }
                float sq = 0.0f;
                for (int k = 0; k < 3; ++k) sq = sq + d[k] * d[k];
                const float r = std::sqrt(sq);
                const float r3 = r * (r * r);
                const float *o = sensor_orientations + 3 * s;
                float *row = out + (int64_t(b) * num_sensors + s) * n + 3LL * v;
                row[0] = common * std::fma(o[2], d[1], -o[1] * d[2]) / r3;
                row[1] = common * std::fma(o[0], d[2], -o[2] * d[0]) / r3;
                row[2] = common * std::fma(o[1], d[0], -o[0] * d[1]) / r3;
            }
}

// data/shapes.rs:171-191 (spherical), 255-270 (max), 303-322 (activation
// time). acos and atan2 are the C library's float functions, which Rust's
// f32::acos / f32::atan2 call. argmax_skipnan keeps the first index whose
// value is greater than every earlier non-NaN value.
int ct_oracle_spherical_f32(const float *x, int32_t T, int32_t N, float rate, float *mag,
                            float *theta, float *phi, float *mag_max, float *theta_max,
                            float *phi_max, float *act_ms) {
    const int64_t n = 3LL * N;
    for (int32_t t = 0; t < T; ++t)
        for (int32_t v = 0; v < N; ++v) {
            const float *s = x + t * n + 3LL * v;
            const float m = std::fabs(s[0]) + std::fabs(s[1]) + std::fabs(s[2]);
            mag[int64_t(t) * N + v] = m;
            theta[int64_t(t) * N + v] = std::acos(s[2] / m);
            phi[int64_t(t) * N + v] = std::atan2(s[1], s[0]);
        }
    for (int32_t v = 0; v < N; ++v) {
        int32_t best = -1;
        for (int32_t t = 0; t < T; ++t) {
            const float m = mag[int64_t(t) * N + v];
            if (std::isnan(m)) continue;
            if (best < 0 || m > mag[int64_t(best) * N + v]) best = t;
        }
        if (best < 0) return -1;
        mag_max[v] = mag[int64_t(best) * N + v];
        theta_max[v] = theta[int64_t(best) * N + v];
        phi_max[v] = phi[int64_t(best) * N + v];
        act_ms[v] = float(best) / rate * 1000.0f;
    }
    return 0;
}

// algorithm/metrics.rs:225-441 (calculate_final, calculate_for_threshold,
// the four metric functions, predict_voxeltype), loop for loop: for each
// threshold i as f32 / 100.0, each cell holding a voxel number k gets the
// series |x[t][k]| + |x[t][k+1]| + |x[t][k+2]| and is Pathological when its
// max_skipnan is <= threshold, else Ventricle; cells without a number stay
// None. max_skipnan is ndarray-stats 0.6.0's (pinned in the reference's
// Cargo.lock; not part of the reference): a fold over the non-NaN values
// keeping the larger, NaN when there is none, so such a cell is never
// Pathological. Each metric counts over all cells as its function does and
// converts the usize counts to f32 before dividing.
namespace {
constexpr uint8_t kPredNone = 0, kPredVentricle = 1, kPredPathological = 2;
constexpr uint32_t kTypePathological = 6;  // VoxelType ordinal (model/spatial/voxels.rs:664-676)
}  // namespace

void ct_oracle_calculate_final_f32(const float *x, int32_t T, int64_t n, const uint32_t *types,
                                   const int32_t *numbers, int64_t cells, float *dice, float *iou,
                                   float *precision, float *recall) {
    std::vector<uint8_t> pred(size_t(std::max<int64_t>(cells, 1)));
    std::vector<float> abs_series(size_t(std::max<int32_t>(T, 1)));
    for (int i = 0; i <= 100; ++i) {
        const float threshold = float(i) / 100.0f;
        for (int64_t c = 0; c < cells; ++c) {
            pred[c] = kPredNone;
            if (numbers[c] < 0) continue;
            const int64_t k = numbers[c];
            for (int32_t t = 0; t < T; ++t)
                abs_series[t] = std::fabs(x[t * n + k]) + std::fabs(x[t * n + k + 1]) +
                                std::fabs(x[t * n + k + 2]);
            bool have = false;
            float mx = 0.0f;
            for (int32_t t = 0; t < T; ++t) {
                if (std::isnan(abs_series[t])) continue;
                if (!have || abs_series[t] > mx) mx = abs_series[t];
                have = true;
            }
            const float max_skipnan = have ? mx : std::nanf("");
            pred[c] = max_skipnan <= threshold ? kPredPathological : kPredVentricle;
        }
        auto count = [&](auto keep) {
            int64_t m = 0;
            for (int64_t c = 0; c < cells; ++c)
                m += keep(pred[c] == kPredPathological, types[c] == kTypePathological) ? 1 : 0;
            return m;
        };
        // calculate_dice (metrics.rs:373-404)
        {
            const int64_t tp = count([](bool p, bool g) { return g && p; });
            const int64_t fp = count([](bool p, bool g) { return !g && p; });
            const int64_t fn = count([](bool p, bool g) { return g && !p; });
            const int64_t den = 2 * tp + fp + fn;
            dice[i] = den == 0 ? 0.0f : float(2 * tp) / float(den);
        }
        // calculate_iou (metrics.rs:335-364)
        {
            const int64_t inter = count([](bool p, bool g) { return g && p; });
            const int64_t uni = count([](bool p, bool g) { return g || p; });
            iou[i] = uni == 0 ? 0.0f : float(inter) / float(uni);
        }
        // calculate_precision (metrics.rs:300-326)
        {
            const int64_t pp = count([](bool p, bool) { return p; });
            const int64_t tp = count([](bool p, bool g) { return g && p; });
            precision[i] = pp == 0 ? 0.0f : float(tp) / float(pp);
        }
        // calculate_recall (metrics.rs:266-292)
        {
            const int64_t gp = count([](bool, bool g) { return g; });
            const int64_t tp = count([](bool p, bool g) { return g && p; });
            recall[i] = gp == 0 ? 1.0f : float(tp) / float(gp);
        }
    }
}

// allpass/delay.rs:14-23: |(in - out) / 1000| / v, the norm the sqrt of the
// ndarray sum ((0 + d0^2) + d1^2) + d2^2 in f32.
float ct_oracle_delay_s_f32(const float *in, const float *out, float velocity) {
    float sq = 0.0f;
    for (int k = 0; k < 3; ++k) {
        const float d = (in[k] - out[k]) / 1000.0f;
        sq = sq + d * d;
    }
    return std::sqrt(sq) / velocity;
}

// allpass.rs:212-289 and allpass/delay.rs:35-111: the same loops (cells in C
// order, offsets x, then y, then z, in -1..=1), the same validity test
// (voxels.rs:117-150: inside the grid and connectable) and the same indices.
int ct_oracle_model_connections_f32(int32_t nx, int32_t ny, int32_t nz, const int32_t *numbers,
                                    const float *positions_mm, const float *velocities,
                                    float sample_rate_hz, int32_t num_voxels, int32_t *osi,
                                    float *delay_samples) {
    std::fill(osi, osi + int64_t(num_voxels) * 3 * 78, -1);
    std::fill(delay_samples, delay_samples + int64_t(num_voxels) * 26, 0.0f);
    auto cell = [&](int32_t x, int32_t y, int32_t z) { return (int64_t(x) * ny + y) * nz + z; };
    for (int32_t x = 0; x < nx; ++x)
        for (int32_t y = 0; y < ny; ++y)
            for (int32_t z = 0; z < nz; ++z) {
                const int64_t ci = cell(x, y, z);
                const int32_t in_num = numbers[ci];
                if (in_num < 0) continue;
                for (int32_t ox = -1; ox <= 1; ++ox)
                    for (int32_t oy = -1; oy <= 1; ++oy)
                        for (int32_t oz = -1; oz <= 1; ++oz) {
                            if (ox == 0 && oy == 0 && oz == 0) continue;
                            const int32_t X = x + ox, Y = y + oy, Z = z + oz;
                            if (X < 0 || X >= nx || Y < 0 || Y >= ny || Z < 0 || Z >= nz) continue;
                            const int64_t co = cell(X, Y, Z);
                            const int32_t out_num = numbers[co];
                            if (out_num < 0) continue;
                            for (int32_t in = 0; in < 3; ++in)
                                for (int32_t od = 0; od < 3; ++od)
                                    osi[int64_t(in_num + in) * 78 +
                                        ct_oracle_offset_to_gain_index(ox, oy, oz, od)] = out_num + od;
                            const float delay = ct_oracle_delay_s_f32(positions_mm + 3 * ci,
                                                                      positions_mm + 3 * co,
                                                                      velocities[ci]) *
                                                sample_rate_hz;
                            if (delay < 1.0f) return -1;
                            delay_samples[int64_t(in_num / 3) * 26 +
                                          ct_oracle_offset_to_delay_index(ox, oy, oz)] = delay;
                        }
            }
    return 0;
}

}  // extern "C"
