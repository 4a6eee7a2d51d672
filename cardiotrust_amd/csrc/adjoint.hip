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
// then runs one lane per connection: a forward sweep that stores the filter
// outputs every K steps, and a backward sweep that recomputes each K-step
// segment from its checkpoint and accumulates dG and da.
#include <hip/hip_runtime.h>

#include "device.h"
#include "kernels.h"

namespace ct {
namespace {

constexpr int kAdjVoxels = 64;
constexpr int kAdjThreads = 3 * kAdjVoxels;

__global__ void max_delay_kernel(const int32_t *__restrict__ nbr, const int32_t *__restrict__ delays,
                                 int64_t len, int32_t *__restrict__ out) {
    __shared__ int32_t red[256];
    int32_t m = 0;
    for (int64_t k = int64_t(blockIdx.x) * 256 + threadIdx.x; k < len; k += int64_t(gridDim.x) * 256)
        if (nbr[k] >= 0) m = max(m, delays[k]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, red[0]);
}

// c_t for one state from the state triple of its voxel and its mapped residual.
__device__ __forceinline__ float step_factor(const float xv[3], int o, float e,
                                             const AdjointParams &p) {
    const float sum = fabsf(xv[0]) + fabsf(xv[1]) + fabsf(xv[2]);
    const float mr = sum > p.threshold ? (sum - p.threshold) * signum(xv[o]) : 0.0f;
    return fmaf(e, p.mse_scaling, mr * p.reg_scaling);
}

__global__ __launch_bounds__(kAdjThreads) void adjoint_step_kernel(
    const int32_t *__restrict__ nbr, const float *__restrict__ coefs,
    const int32_t *__restrict__ delays, const float *__restrict__ gains_rows,
    const int32_t *__restrict__ cons_ptr, const int32_t *__restrict__ cons_idx,
    const float *__restrict__ x, const float *__restrict__ E, float *__restrict__ ybar,
    float *__restrict__ qring, float *__restrict__ xbar, int32_t N, int32_t T, int32_t s,
    uint32_t rmask, AdjointParams p) {
    __shared__ float xb[kAdjVoxels * 3];
    const int64_t n = 3LL * N;
    const int64_t M = int64_t(N) * 78;  // one ring slot
    const int32_t v0 = xcd_block(blockIdx.x, gridDim.x) * kAdjVoxels;
    const int32_t nv = min(kAdjVoxels, N - v0);

    const int vl = threadIdx.x / 3, o = threadIdx.x - 3 * vl;
    if (vl < nv) {
        const int32_t w = v0 + vl;
        const int64_t i = 3LL * w + o;
        const float *xs = x + int64_t(s) * n + 3LL * w;
        const float xv[3] = {xs[0], xs[1], xs[2]};
        float acc = step_factor(xv, o, E[int64_t(s) * n + i], p);
        for (int32_t j = cons_ptr[w]; j < cons_ptr[w + 1]; ++j) {
            const int32_t c = cons_idx[j];
            const int32_t t = s + delays[c];
            if (t < T) acc += qring[int64_t(uint32_t(t) & rmask) * M + int64_t(c) * 3 + o];
        }
        xb[threadIdx.x] = acc;
        xbar[int64_t(s) * n + i] = acc;
    }
    __syncthreads();
    for (int item = threadIdx.x; item < nv * 26; item += kAdjThreads) {
        const int64_t c = int64_t(v0) * 26 + item;
        if (nbr[c] < 0) continue;
        const int vv = item / 26, d = item - 26 * vv;
        const int64_t row = 3LL * (v0 + vv);
        const float a = coefs[c];
#pragma unroll
        for (int oo = 0; oo < 3; ++oo) {
            const float next = ybar[c * 3 + oo];
            const float *g = gains_rows + row * 78 + 3 * d + oo;
            const float yb =
                g[0] * xb[3 * vv] + g[78] * xb[3 * vv + 1] + g[156] * xb[3 * vv + 2] - a * next;
            qring[int64_t(uint32_t(s) & rmask) * M + c * 3 + oo] = a * yb + next;
            ybar[c * 3 + oo] = yb;
        }
    }
}

constexpr int kGradThreads = 256;

__global__ __launch_bounds__(kGradThreads) void adjoint_gradient_kernel(
    const int32_t *__restrict__ nbr, const float *__restrict__ coefs,
    const int32_t *__restrict__ delays, const float *__restrict__ gT,
    const float *__restrict__ xT, const float *__restrict__ xbarT, float *__restrict__ ckpt,
    float *__restrict__ yseg, float *__restrict__ dGT, float *__restrict__ dA, int32_t N,
    int32_t T, int32_t K) {
    const int64_t item = int64_t(xcd_block(blockIdx.x, gridDim.x)) * kGradThreads + threadIdx.x;
    const int64_t total = int64_t(N) * 26;
    if (item >= total) return;
    const int32_t w = nbr[item];
    if (w < 0) return;  // no connection: no dependence on these parameters
    const int64_t n = 3LL * N;
    const int64_t M = total * 3;
    const int32_t v = int32_t(item / 26), d = int32_t(item % 26);
    const float a = coefs[item];
    const int32_t D = delays[item];
    const float *xw = xT + 3LL * w * T;
    const float *xbv = xbarT + 3LL * v * T;
    const int32_t nseg = (T + K - 1) / K;

    // forward sweep: filter outputs, stored at the end of every segment but the last
    float y[3] = {0.0f, 0.0f, 0.0f}, prev_in[3] = {0.0f, 0.0f, 0.0f};
    for (int32_t t = 0; t < T; ++t) {
#pragma unroll
        for (int o = 0; o < 3; ++o) {
            const float in0 = D <= t ? xw[int64_t(o) * T + t - D] : 0.0f;
            y[o] = fmaf(a, in0 - y[o], prev_in[o]);
            prev_in[o] = in0;
        }
        if ((t + 1) % K == 0 && t + 1 < T) {
            const int64_t k = (t + 1) / K;
#pragma unroll
            for (int o = 0; o < 3; ++o) ckpt[(k - 1) * M + item * 3 + o] = y[o];
        }
    }

    float G[3][3], g[3][3];
#pragma unroll
    for (int in = 0; in < 3; ++in)
#pragma unroll
        for (int o = 0; o < 3; ++o) {
            G[in][o] = gT[int64_t(3 * d + o) * n + 3LL * v + in];
            g[in][o] = 0.0f;
        }
    float A = 0.0f;
    float ybn[3] = {0.0f, 0.0f, 0.0f};
    for (int32_t k = nseg - 1; k >= 0; --k) {
        const int32_t t0 = k * K, t1 = min(T, t0 + K);
        // recompute the segment from its checkpoint (y_{t0-1})
        float yc[3], ys[3], pin[3];
#pragma unroll
        for (int o = 0; o < 3; ++o) {
            yc[o] = k == 0 ? 0.0f : ckpt[int64_t(k - 1) * M + item * 3 + o];
            ys[o] = yc[o];
            pin[o] = D <= t0 - 1 ? xw[int64_t(o) * T + t0 - 1 - D] : 0.0f;
        }
        for (int32_t t = t0; t < t1; ++t) {
#pragma unroll
            for (int o = 0; o < 3; ++o) {
                const float in0 = D <= t ? xw[int64_t(o) * T + t - D] : 0.0f;
                ys[o] = fmaf(a, in0 - ys[o], pin[o]);
                pin[o] = in0;
                yseg[int64_t(t - t0) * M + item * 3 + o] = ys[o];
            }
        }
        for (int32_t s = t1 - 1; s >= t0; --s) {
            float xb[3], yb[3];
#pragma unroll
            for (int in = 0; in < 3; ++in) xb[in] = xbv[int64_t(in) * T + s];
#pragma unroll
            for (int o = 0; o < 3; ++o) {
                yb[o] = G[0][o] * xb[0] + G[1][o] * xb[1] + G[2][o] * xb[2] - a * ybn[o];
                const float yt = yseg[int64_t(s - t0) * M + item * 3 + o];
                const float ym = s > t0 ? yseg[int64_t(s - 1 - t0) * M + item * 3 + o] : yc[o];
                const float xi = D <= s ? xw[int64_t(o) * T + s - D] : 0.0f;
#pragma unroll
                for (int in = 0; in < 3; ++in) g[in][o] += xb[in] * yt;
                A += yb[o] * (xi - ym);
                ybn[o] = yb[o];
            }
        }
    }
#pragma unroll
    for (int in = 0; in < 3; ++in)
#pragma unroll
        for (int o = 0; o < 3; ++o) dGT[int64_t(3 * d + o) * n + 3LL * v + in] += g[in][o];
    dA[item] += A;
}

}  // namespace

hipError_t launch_max_delay(hipStream_t s, const int32_t *nbr, const int32_t *delays, int64_t len,
                            int32_t *out) {
    hipError_t e = hipMemsetAsync(out, 0, sizeof(int32_t), s);
    if (e != hipSuccess || len <= 0) return e;
    max_delay_kernel<<<256, 256, 0, s>>>(nbr, delays, len, out);
    return hipGetLastError();
}

hipError_t launch_adjoint_step(hipStream_t s, const int32_t *nbr, const float *coefs,
                               const int32_t *delays, const float *gains_rows,
                               const int32_t *cons_ptr, const int32_t *cons_idx, const float *x,
                               const float *E, float *ybar, float *qring, float *xbar, int32_t N,
                               int32_t T, int32_t step, uint32_t rmask, const AdjointParams &p) {
    if (N <= 0) return hipSuccess;
    const unsigned blocks = unsigned((N + kAdjVoxels - 1) / kAdjVoxels);
    adjoint_step_kernel<<<blocks, kAdjThreads, 0, s>>>(nbr, coefs, delays, gains_rows, cons_ptr,
                                                       cons_idx, x, E, ybar, qring, xbar, N, T,
                                                       step, rmask, p);
    return hipGetLastError();
}

hipError_t launch_adjoint_gradient(hipStream_t s, const int32_t *nbr, const float *coefs,
                                   const int32_t *delays, const float *gT, const float *xT,
                                   const float *xbarT, float *ckpt, float *yseg, float *dGT,
                                   float *dA, int32_t N, int32_t T, int32_t K) {
    if (N <= 0) return hipSuccess;
    const unsigned blocks = unsigned((int64_t(N) * 26 + kGradThreads - 1) / kGradThreads);
    adjoint_gradient_kernel<<<blocks, kGradThreads, 0, s>>>(nbr, coefs, delays, gT, xT, xbarT,
                                                            ckpt, yseg, dGT, dA, N, T, K);
    return hipGetLastError();
}

}  // namespace ct
