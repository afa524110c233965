// Streaming ceiling for kernels that read K arrays and write M arrays point by point.
//
// Profiling aid for the FDTD passes: a pass of the 3-D elastic scheme reads 10-11 arrays
// and writes 3-6, the 2-D passes read 6-7 and write 2-3, a plain copy reads 1 and writes 1.
// This program times out_j[i] = (j + 1) * sum_k in_k[i] over separately allocated arrays of
// the same size, for several (K, M), with 4-, 8- and 16-byte accesses per lane (the
// stencil kernels load 4 bytes per lane; 4-byte lanes also with one element per thread,
// as the stencil kernels have), and with a z-stencil's extra load: each lane also
// reads the element just past its own (in_k[i + 1] for 4-byte lanes, the first float of
// the next vector for wider ones), so that a pass's achieved bandwidth can be set beside
// what a plain stream of the same number of arrays and the same access width reaches.
//   usage: stream_ceiling [MiB per array=178] [reps=5]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

constexpr int kMaxIn = 12;
constexpr int kMaxOut = 6;

struct Arrays {
    const float* in[kMaxIn];
    float* out[kMaxOut];
};

typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v2f __attribute__((ext_vector_type(2)));

// U elements per thread, 256 apart; every load of the thread is issued before its first
// store.  T is float, v2f or v4f; n counts elements of T.  With HALO, every lane also
// loads the float just past its element (in bounds: the last element of an array skips it).
template <int K, int M, int U, typename T, bool HALO>
__global__ __launch_bounds__(256) void streams(Arrays a, size_t n)
{
    constexpr int W = sizeof(T) / sizeof(float);
    const size_t base = (size_t)blockIdx.x * 256 * U + threadIdx.x;
    T s[U];
#pragma unroll
    for (int u = 0; u < U; ++u) s[u] = T(0);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const T* p = reinterpret_cast<const T*>(a.in[k]);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t i = base + (size_t)u * 256;
            if (i < n) {
                s[u] += p[i];
                if (HALO && i + 1 < n) s[u] += a.in[k][(i + 1) * W];
            }
        }
    }
#pragma unroll
    for (int j = 0; j < M; ++j) {
        T* q = reinterpret_cast<T*>(a.out[j]);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t i = base + (size_t)u * 256;
            if (i < n) q[i] = s[u] * (float)(j + 1);
        }
    }
}

template <int K, int M, int U, typename T, bool HALO>
static float run(const Arrays& a, size_t nfloat, int reps, hipEvent_t e0, hipEvent_t e1)
{
    const size_t n = nfloat / (sizeof(T) / sizeof(float));
    const unsigned blocks = (unsigned)((n + 256 * U - 1) / (256 * U));
    float best = 1e30f;
    for (int r = 0; r < reps; ++r) {
        (void)hipEventRecord(e0, 0);
        hipLaunchKernelGGL((streams<K, M, U, T, HALO>), dim3(blocks), dim3(256), 0, 0, a, n);
        (void)hipEventRecord(e1, 0);
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) best = ms;
    }
    return best;
}

template <int K, int M, bool HALO>
static int widths(const Arrays& a, size_t nfloat, int reps, hipEvent_t e0, hipEvent_t e1)
{
    const float t41 = run<K, M, 1, float, HALO>(a, nfloat, reps, e0, e1);
    const float t4 = run<K, M, 4, float, HALO>(a, nfloat, reps, e0, e1);
    const float t8 = run<K, M, 4, v2f, HALO>(a, nfloat, reps, e0, e1);
    const float t16 = run<K, M, 4, v4f, HALO>(a, nfloat, reps, e0, e1);
    if (hipGetLastError() != hipSuccess) return 1;
    const double bytes = (double)(K + M) * nfloat * sizeof(float);
    printf("stream_ceiling read %2d write %d%s: %.3e bytes; GB/s with 4-byte lanes, 1 element per thread %.1f, "
           "4 per thread %.1f; 8-byte lanes %.1f, 16-byte %.1f (4 per thread)\n",
           K, M, HALO ? " + z-neighbour" : "", bytes, bytes / (t41 * 1e-3) / 1e9, bytes / (t4 * 1e-3) / 1e9,
           bytes / (t8 * 1e-3) / 1e9, bytes / (t16 * 1e-3) / 1e9);
    fflush(stdout);
    return 0;
}

template <int K, int M>
static int both(const Arrays& a, size_t nfloat, int reps, hipEvent_t e0, hipEvent_t e1)
{
    return widths<K, M, false>(a, nfloat, reps, e0, e1) | widths<K, M, true>(a, nfloat, reps, e0, e1);
}

int main(int argc, char** argv)
{
    const size_t mib = argc > 1 ? strtoul(argv[1], nullptr, 10) : 178;
    const int reps = argc > 2 ? atoi(argv[2]) : 5;
    const size_t bytes = mib << 20;
    const size_t nfloat = bytes / sizeof(float);
    Arrays a{};
    for (int k = 0; k < kMaxIn; ++k) {
        float* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess || hipMemset(p, 0, bytes) != hipSuccess) {
            fprintf(stderr, "hipMalloc failed\n");
            return 1;
        }
        a.in[k] = p;
    }
    for (int j = 0; j < kMaxOut; ++j)
        if (hipMalloc(&a.out[j], bytes) != hipSuccess) {
            fprintf(stderr, "hipMalloc failed\n");
            return 1;
        }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    int rc = 0;
    rc |= both<1, 1>(a, nfloat, reps, e0, e1);
    rc |= both<4, 2>(a, nfloat, reps, e0, e1);
    rc |= both<7, 3>(a, nfloat, reps, e0, e1);
    rc |= both<11, 6>(a, nfloat, reps, e0, e1);
    for (int k = 0; k < kMaxIn; ++k) (void)hipFree(const_cast<float*>(a.in[k]));
    for (int j = 0; j < kMaxOut; ++j) (void)hipFree(a.out[j]);
    return rc;
}
