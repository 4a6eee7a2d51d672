// Calibration probe for HBM byte counters: copies a buffer of known size with
// a plain float4 kernel so that rocprofv3 FETCH_SIZE / WRITE_SIZE readings of
// the epoch kernels can be scaled by what the counters report for exactly
// `bytes` read and `bytes` written.
//   usage: copy_probe [MiB=1024] [repeats=3]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void copy_kernel(const float4 *__restrict__ in, float4 *__restrict__ out, size_t n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += size_t(gridDim.x) * blockDim.x)
        out[i] = in[i];
}

int main(int argc, char **argv) {
    const size_t mib = argc > 1 ? strtoull(argv[1], nullptr, 10) : 1024;
    const int repeats = argc > 2 ? atoi(argv[2]) : 3;
    const size_t bytes = mib << 20;
    const size_t n = bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        fprintf(stderr, "hipMalloc failed\n");
        return 1;
    }
    (void)hipMemset(a, 1, bytes);
    (void)hipMemset(b, 0, bytes);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    for (int r = 0; r < repeats; ++r) {
        (void)hipEventRecord(e0, nullptr);
        copy_kernel<<<8192, 256>>>(a, b, n);
        (void)hipEventRecord(e1, nullptr);
        if (hipEventSynchronize(e1) != hipSuccess) {
            fprintf(stderr, "copy failed\n");
            return 1;
        }
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        printf("copy %zu MiB: %.3f ms, %.1f GB/s (read+write)\n", mib, ms, 2.0 * bytes / ms / 1e6);
    }
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
