// Checks that the CPU build's runtime catches what it is meant to catch (run on the host
// only, by tests/test_cpu_emu.py): each case is one process that must end with status 86
// and a message naming the access.
//   oob        a kernel writes one float past the end of a hipMalloc'd buffer
//   after-free a kernel queued behind a hipFreeAsync of its buffer on another stream (the
//              free ordered first by an event) reads the freed buffer
//   ordered    the same two streams with the free ordered after the kernel: no fault
//   barrier    workgroup memory and __syncthreads: each of 4 steps writes, waits, reads
//              another thread's value and waits again (no fault, and the sums are checked)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>

__global__ void touch(float* p, int n)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i <= n) p[i] += 1.0f;   // i == n is out of bounds
}

__global__ void read_all(const float* p, float* out, int n)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) out[i] = p[i];
}

// out[g*256 + t] = sum over steps r of the value thread (t + 1 + r) % 256 stored in step r
__global__ void rotate(float* out)
{
    __shared__ float tile[256];
    const int t = (int)threadIdx.x;
    float acc = 0.0f;
    for (int r = 0; r < 4; ++r) {
        tile[t] = (float)(1000 * r + t + 256 * (int)blockIdx.x);
        __syncthreads();
        acc += tile[(t + 1 + r) % 256];
        __syncthreads();
    }
    out[blockIdx.x * 256 + t] = acc;
}

int main(int argc, char** argv)
{
    const char* which = argc > 1 ? argv[1] : "";
    const int n = 1024;
    float *a = nullptr, *b = nullptr;
    hipMalloc(&a, n * sizeof(float));
    hipMalloc(&b, n * sizeof(float));
    hipStream_t s1, s2;
    hipStreamCreateWithFlags(&s1, hipStreamNonBlocking);
    hipStreamCreateWithFlags(&s2, hipStreamNonBlocking);
    hipEvent_t e;
    hipEventCreate(&e);
    if (!strcmp(which, "oob")) {
        hipLaunchKernelGGL(touch, dim3(n / 256 + 1), dim3(256), 0, s1, a, n);
        hipStreamSynchronize(s1);
    } else if (!strcmp(which, "after-free")) {
        hipFreeAsync(a, s1);
        hipEventRecord(e, s1);
        hipStreamWaitEvent(s2, e, 0);
        hipLaunchKernelGGL(read_all, dim3(n / 256), dim3(256), 0, s2, a, b, n);
        hipStreamSynchronize(s2);
    } else if (!strcmp(which, "ordered")) {
        hipLaunchKernelGGL(read_all, dim3(n / 256), dim3(256), 0, s2, a, b, n);
        hipEventRecord(e, s2);
        hipStreamWaitEvent(s1, e, 0);
        hipFreeAsync(a, s1);
        hipStreamSynchronize(s1);
    } else if (!strcmp(which, "barrier")) {
        const int g = 3;
        float* out = nullptr;
        hipMalloc(&out, g * 256 * sizeof(float));
        hipLaunchKernelGGL(rotate, dim3(g), dim3(256), 0, s1, out);
        hipStreamSynchronize(s1);
        for (int b = 0; b < g; ++b)
            for (int t = 0; t < 256; ++t) {
                float want = 0.0f;
                for (int r = 0; r < 4; ++r) want += (float)(1000 * r + (t + 1 + r) % 256 + 256 * b);
                if (out[b * 256 + t] != want) {
                    fprintf(stderr, "barrier: workgroup %d thread %d has %g, want %g\n", b, t, out[b * 256 + t], want);
                    return 3;
                }
            }
    } else {
        fprintf(stderr, "usage: selftest oob|after-free|ordered|barrier\n");
        return 2;
    }
    printf("no fault\n");
    return 0;
}
