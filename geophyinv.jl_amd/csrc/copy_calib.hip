// Counter calibration: a plain device-to-device copy of known size.
//
// Profiling aid for the FDTD kernels: run under `rocprofv3 --pmc FETCH_SIZE` (and, in a
// separate run, WRITE_SIZE) to read what the counters report for exactly `bytes` read and
// `bytes` written per dispatch; the ratio scales the counters of the stencil kernels.
//   usage: copy_calib [MiB=512] [reps=5]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void copy_f4(const float4* __restrict__ a, float4* __restrict__ b, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) b[i] = a[i];
}

int main(int argc, char** argv)
{
    const size_t mib = argc > 1 ? strtoul(argv[1], nullptr, 10) : 512;
    const int reps = argc > 2 ? atoi(argv[2]) : 5;
    const size_t bytes = mib << 20;
    const size_t n = bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        fprintf(stderr, "hipMalloc failed\n");
        return 1;
    }
    if (hipMemset(a, 1, bytes) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const int blocks = 256 * 8 * 4;
    float best = 1e30f;
    for (int r = 0; r < reps; ++r) {
        (void)hipEventRecord(e0, 0);
        hipLaunchKernelGGL(copy_f4, dim3(blocks), dim3(256), 0, 0, a, b, n);
        (void)hipEventRecord(e1, 0);
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) best = ms;
    }
    if (hipGetLastError() != hipSuccess) return 1;
    printf("copy_calib: %zu bytes read + %zu bytes written per dispatch, best %.4f ms, %.1f GB/s\n",
           bytes, bytes, best, 2.0 * bytes / (best * 1e-3) / 1e9);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
