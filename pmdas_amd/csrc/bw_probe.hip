// HBM probe: a plain float4 copy of a known size, used to measure the
// achievable copy bandwidth and to scale the FETCH_SIZE / WRITE_SIZE counters
// (the copy reads and writes exactly `bytes` each).
//   usage: bw_probe [MiB]      (default 1024)
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>

__global__ __launch_bounds__(256) void copy_probe(const float4* __restrict__ a, float4* __restrict__ b,
                                                  size_t n) {
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) b[i] = a[i];
}

__global__ __launch_bounds__(256) void fill_probe(float4* a, size_t n) {
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        a[i] = make_float4(1.0f, 2.0f, 3.0f, 4.0f);
}

int main(int argc, char** argv) {
    const size_t mib = argc > 1 ? (size_t)atoll(argv[1]) : 1024;
    const size_t bytes = mib << 20, n = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        printf("allocation failed\n");
        return 1;
    }
    hipLaunchKernelGGL(fill_probe, dim3(8192), dim3(256), 0, 0, a, n);
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    float best = 1e30f;
    for (int r = 0; r < 5; ++r) {
        (void)hipEventRecord(e0, 0);
        hipLaunchKernelGGL(copy_probe, dim3(8192), dim3(256), 0, 0, a, b, n);
        (void)hipEventRecord(e1, 0);
        if (hipEventSynchronize(e1) != hipSuccess) return 1;
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) best = ms;
        printf("copy %zu MiB: %.3f ms  %.1f GB/s (read + write)\n", mib, ms, 2.0 * bytes / ms / 1e6);
    }
    printf("best: %.1f GB/s\n", 2.0 * bytes / best / 1e6);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
