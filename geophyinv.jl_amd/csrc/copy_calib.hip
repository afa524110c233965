// Counter calibration and copy bandwidth: device-to-device copies of known size.
//
// Profiling aid for the FDTD kernels: run under `rocprofv3 --pmc FETCH_SIZE` (and, in a
// separate run, WRITE_SIZE) to read what the counters report for exactly `bytes` read and
// `bytes` written per dispatch; the ratio scales the counters of the stencil kernels.  It
// also times several copy kernels, so the best of them is what a plain stream of reads and
// writes of that size achieves on the device:
//   gridstride  one float4 per thread and iteration, 8192 workgroups looping over the array
//   unroll4     4 float4 per thread (all loads, then all stores), one workgroup per 16 KiB
//   unroll8     8 float4 per thread, one workgroup per 32 KiB
//   unroll8_nt  unroll8 with non-temporal loads and stores
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

typedef float v4f __attribute__((ext_vector_type(4)));

template <int U, bool NT>
__global__ __launch_bounds__(256) void copy_unrolled(const v4f* __restrict__ a, v4f* __restrict__ b, size_t n)
{
    const size_t base = (size_t)blockIdx.x * 256 * U + threadIdx.x;
    v4f v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const size_t i = base + (size_t)u * 256;
        if (i < n) v[u] = NT ? __builtin_nontemporal_load(a + i) : a[i];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const size_t i = base + (size_t)u * 256;
        if (i < n) {
            if (NT) __builtin_nontemporal_store(v[u], b + i);
            else b[i] = v[u];
        }
    }
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
    const char* names[4] = {"gridstride", "unroll4", "unroll8", "unroll8_nt"};
    const v4f* va = reinterpret_cast<const v4f*>(a);
    v4f* vb = reinterpret_cast<v4f*>(b);
    for (int kind = 0; kind < 4; ++kind) {
        float best = 1e30f;
        for (int r = 0; r < reps; ++r) {
            (void)hipEventRecord(e0, 0);
            switch (kind) {
            case 0: hipLaunchKernelGGL(copy_f4, dim3(256 * 8 * 4), dim3(256), 0, 0, a, b, n); break;
            case 1: hipLaunchKernelGGL((copy_unrolled<4, false>), dim3((unsigned)((n + 1023) / 1024)), dim3(256), 0, 0,
                                       va, vb, n); break;
            case 2: hipLaunchKernelGGL((copy_unrolled<8, false>), dim3((unsigned)((n + 2047) / 2048)), dim3(256), 0, 0,
                                       va, vb, n); break;
            default: hipLaunchKernelGGL((copy_unrolled<8, true>), dim3((unsigned)((n + 2047) / 2048)), dim3(256), 0,
                                        0, va, vb, n); break;
            }
            (void)hipEventRecord(e1, 0);
            (void)hipEventSynchronize(e1);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (ms < best) best = ms;
        }
        if (hipGetLastError() != hipSuccess) return 1;
        printf("copy_calib %-10s: %zu bytes read + %zu bytes written per dispatch, best %.4f ms, %.1f GB/s\n",
               names[kind], bytes, bytes, best, 2.0 * bytes / (best * 1e-3) / 1e9);
    }
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
