// FP64 evaluation of the reverse-mode objective for gradient checking, gfx950.
//
//   J = sum_t [ s_mse/2 |H x_t - mhat_t|^2 + s_reg/2 sum_v relu(|x_{t,v}|_1 - theta)^2 ]
//
// with the FP32 parameters widened to double and at most one parameter
// shifted by an FP64 `delta`, so that central differences (J(p+h) - J(p-h))/2h
// can be taken at h far below FP32 resolution. The arithmetic is the
// propagation of prediction.rs:40-105 (filter y = a (x[t-D] - y) + x[t-D-1],
// state = sum over connections of gain x y, plus u_t b) carried out in double.
//
// propagate_f64_kernel: per workgroup of 8 voxels, one lane per connection
// advances its 3 filters, then one lane per state sums its 78 gain terms
// (delays >= 1, so a step reads only earlier rows). project_f64_kernel: tiled split-K contraction P[split][s][t] of
// H (Sp, n) with X (T, n). objective_terms_f64_kernel: per step, the split
// sums, the squared residuals and the maximum-regularisation term.
// objective_sum_f64_kernel: the sum over steps in one fixed order.
#include <hip/hip_runtime.h>

#include "device.h"
#include "kernels.h"

namespace ct {
namespace {

constexpr int kF64Voxels = 8;  // voxels per workgroup: 208 filter lanes, 24 state lanes

__global__ __launch_bounds__(256) void propagate_f64_kernel(
    const int32_t *__restrict__ nbr, const float *__restrict__ coefs,
    const int32_t *__restrict__ delays, const float *__restrict__ gT, double *__restrict__ y,
    double *__restrict__ x, const float *__restrict__ b, float u_t, int32_t N, int32_t t,
    Perturbation pert) {
    __shared__ double ys[kF64Voxels * 78];
    const int32_t v0 = blockIdx.x * kF64Voxels;
    const int32_t nv = min(kF64Voxels, N - v0);
    const int64_t n = 3LL * N;
    const int tid = threadIdx.x;
    // phase 1: lane (voxel, connection) advances its three filters
    if (tid < nv * 26) {
        const int64_t c = int64_t(v0) * 26 + tid;
        const int32_t w = nbr[c];
        double yo[3] = {0.0, 0.0, 0.0};
        if (w >= 0) {
            double a = double(coefs[c]);
            if (pert.kind == 2 && pert.index == c) a += pert.delta;
            const int32_t D = delays[c];
            for (int o = 0; o < 3; ++o) {
                const double in0 = D <= t ? x[int64_t(t - D) * n + 3LL * w + o] : 0.0;
                const double in1 = D < t ? x[int64_t(t - D - 1) * n + 3LL * w + o] : 0.0;
                yo[o] = fma(a, in0 - y[c * 3 + o], in1);
                y[c * 3 + o] = yo[o];
            }
        }
        for (int o = 0; o < 3; ++o) ys[tid * 3 + o] = yo[o];
    }
    __syncthreads();
    // phase 2: lane (voxel, in) sums its 78 gain terms, connections ascending
    if (tid < nv * 3) {
        const int vl = tid / 3;
        const int64_t i = 3LL * v0 + tid;
        double acc = 0.0;
        for (int g = 0; g < 78; ++g) {
            double gain = double(gT[int64_t(g) * n + i]);
            if (pert.kind == 1 && pert.index == i * 78 + g) gain += pert.delta;
            acc += gain * ys[vl * 78 + g];
        }
        x[int64_t(t) * n + i] = acc + double(u_t) * double(b[i]);
    }
}

// P[split][s][t] = sum over the split's k range of H[s][k] X[t][k].
// 64 x 64 output tile per workgroup, 4 x 4 per thread, k in chunks of 32.
constexpr int kF64Tile = 64, kF64K = 32;
__global__ __launch_bounds__(256) void project_f64_kernel(const float *__restrict__ H,
                                                          const double *__restrict__ X,
                                                          double *__restrict__ P, int32_t Sp,
                                                          int32_t T, int64_t n, int64_t kps) {
    __shared__ double hs[kF64K][kF64Tile + 1];
    __shared__ double xs[kF64K][kF64Tile + 1];
    const int tid = threadIdx.x, ts = tid / 16, tt = tid % 16;
    const int32_t s0 = blockIdx.x * kF64Tile, t0 = blockIdx.y * kF64Tile;
    const int64_t k0 = int64_t(blockIdx.z) * kps, k1 = min(n, k0 + kps);
    double acc[4][4] = {};
    for (int64_t kb = k0; kb < k1; kb += kF64K) {
        for (int e = tid; e < kF64Tile * kF64K; e += 256) {
            const int r = e / kF64K, kk = e % kF64K;
            const int64_t k = kb + kk;
            hs[kk][r] = (k < k1) ? double(H[int64_t(s0 + r) * n + k]) : 0.0;
            const int32_t t = t0 + r;
            xs[kk][r] = (k < k1 && t < T) ? X[int64_t(t) * n + k] : 0.0;
        }
        __syncthreads();
        for (int kk = 0; kk < kF64K; ++kk) {
            double hv[4], xv[4];
            for (int q = 0; q < 4; ++q) {
                hv[q] = hs[kk][ts * 4 + q];
                xv[q] = xs[kk][tt * 4 + q];
            }
            for (int p = 0; p < 4; ++p)
                for (int q = 0; q < 4; ++q) acc[p][q] = fma(hv[p], xv[q], acc[p][q]);
        }
        __syncthreads();
    }
    for (int p = 0; p < 4; ++p)
        for (int q = 0; q < 4; ++q) {
            const int32_t s = s0 + ts * 4 + p, t = t0 + tt * 4 + q;
            if (t < T) P[(int64_t(blockIdx.z) * Sp + s) * T + t] = acc[p][q];
        }
}

// One workgroup per step t: terms[t] = s_mse/2 sum_s r^2 + s_reg/2 sum_v f^2.
__global__ __launch_bounds__(256) void objective_terms_f64_kernel(
    const double *__restrict__ P, int32_t splits, const float *__restrict__ mhat,
    const double *__restrict__ X, double *__restrict__ terms, int32_t S, int32_t Sp, int32_t T,
    int32_t N, double s_mse, double s_reg, double theta) {
    __shared__ double red[256];
    const int32_t t = blockIdx.x;
    const int64_t n = 3LL * N;
    double mse = 0.0, reg = 0.0;
    for (int32_t s = threadIdx.x; s < S; s += 256) {
        double m = 0.0;
        for (int32_t k = 0; k < splits; ++k) m += P[(int64_t(k) * Sp + s) * T + t];
        const double r = m - double(mhat[int64_t(t) * S + s]);
        mse += r * r;
    }
    for (int32_t v = threadIdx.x; v < N; v += 256) {
        const double *xv = X + int64_t(t) * n + 3LL * v;
        const double f = fabs(xv[0]) + fabs(xv[1]) + fabs(xv[2]) - theta;
        if (f > 0.0) reg += f * f;
    }
    red[threadIdx.x] = 0.5 * s_mse * mse + 0.5 * s_reg * reg;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) terms[t] = red[0];
}

__global__ void objective_sum_f64_kernel(const double *__restrict__ terms, int32_t T,
                                         double *__restrict__ J) {
    if (threadIdx.x != 0) return;
    double acc = 0.0;
    for (int32_t t = 0; t < T; ++t) acc += terms[t];
    *J = acc;
}

inline unsigned grid_for(int64_t len, int threads) { return unsigned((len + threads - 1) / threads); }

}  // namespace

hipError_t launch_propagate_f64(hipStream_t s, const int32_t *nbr, const float *coefs,
                                const int32_t *delays, const float *gT, double *y, double *x,
                                const float *b, float u_t, int32_t N, int32_t t,
                                const Perturbation &pert) {
    if (N <= 0) return hipSuccess;
    propagate_f64_kernel<<<grid_for(N, kF64Voxels), 256, 0, s>>>(nbr, coefs, delays, gT, y, x, b,
                                                                 u_t, N, t, pert);
    return hipGetLastError();
}

hipError_t launch_objective_f64(hipStream_t s, const float *H, const double *X,
                                const float *mhat, double *P, int32_t splits, double *terms,
                                double *J, int32_t S, int32_t Sp, int32_t T, int32_t N,
                                double s_mse, double s_reg, double theta) {
    const int64_t n = 3LL * N;
    if (T <= 0) {
        objective_sum_f64_kernel<<<1, 64, 0, s>>>(terms, 0, J);
        return hipGetLastError();
    }
    int64_t kps = (n + splits - 1) / splits;
    kps = (kps + kF64K - 1) / kF64K * kF64K;
    dim3 grid(Sp / kF64Tile, grid_for(T, kF64Tile), splits);
    project_f64_kernel<<<grid, 256, 0, s>>>(H, X, P, Sp, T, n, kps);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    objective_terms_f64_kernel<<<T, 256, 0, s>>>(P, splits, mhat, X, terms, S, Sp, T, N, s_mse,
                                                 s_reg, theta);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    objective_sum_f64_kernel<<<1, 64, 0, s>>>(terms, T, J);
    return hipGetLastError();
}

}  // namespace ct
