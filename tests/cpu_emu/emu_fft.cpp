// hipFFT stand-in of the CPU build: batched 1-D real transforms by a recursive mixed-radix
// decimation-in-time FFT (any length; prime factors beyond 7 fall back to direct sums),
// queued on the plan's stream like a kernel.
#include <hipfft/hipfft.h>

#include <cmath>
#include <complex>
#include <functional>
#include <vector>

namespace emu {
void enqueue(hipStream_t hs, std::function<void()> fn);
}

struct EmuFftPlan {
    int n = 0, batch = 0, idist = 0, odist = 0;
    hipfftType type = HIPFFT_D2Z;
    hipStream_t stream = nullptr;
};

namespace {

using cd = std::complex<double>;

// out[k] = sum_j in[j stride] exp(sign 2 pi i j k / n), k < n
void fft(const cd* in, int stride, cd* out, int n, int sign)
{
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    int p = n;
    for (int f : {2, 3, 5, 7})
        if (n % f == 0) {
            p = f;
            break;
        }
    const double w0 = sign * 2.0 * M_PI / n;
    if (p == n) {   // a prime beyond 7: direct sums
        for (int k = 0; k < n; ++k) {
            cd acc = 0.0;
            for (int j = 0; j < n; ++j) acc += in[(size_t)j * stride] * std::polar(1.0, w0 * (double)(((long)j * k) % n));
            out[k] = acc;
        }
        return;
    }
    const int m = n / p;
    std::vector<cd> sub((size_t)n);
    for (int r = 0; r < p; ++r) fft(in + (size_t)r * stride, stride * p, sub.data() + (size_t)r * m, m, sign);
    for (int k = 0; k < m; ++k)
        for (int q = 0; q < p; ++q) {
            const int kk = k + q * m;
            cd acc = 0.0;
            for (int r = 0; r < p; ++r) acc += sub[(size_t)r * m + k] * std::polar(1.0, w0 * (double)(((long)r * kk) % n));
            out[kk] = acc;
        }
}

}  // namespace

hipfftResult hipfftPlanMany(hipfftHandle* plan, int rank, int* n, int*, int istride, int idist, int*, int ostride,
                            int odist, hipfftType type, int batch)
{
    if (!plan || rank != 1 || !n || n[0] < 1 || istride != 1 || ostride != 1 || batch < 1)
        return HIPFFT_INVALID_VALUE;
    *plan = new EmuFftPlan{n[0], batch, idist, odist, type, nullptr};
    return HIPFFT_SUCCESS;
}

hipfftResult hipfftSetStream(hipfftHandle plan, hipStream_t stream)
{
    if (!plan) return HIPFFT_INVALID_PLAN;
    plan->stream = stream;
    return HIPFFT_SUCCESS;
}

hipfftResult hipfftExecD2Z(hipfftHandle plan, double* in, hipfftDoubleComplex* out)
{
    if (!plan || plan->type != HIPFFT_D2Z) return HIPFFT_INVALID_PLAN;
    const EmuFftPlan P = *plan;
    emu::enqueue(P.stream, [P, in, out] {
        std::vector<cd> x((size_t)P.n), y((size_t)P.n);
        for (int b = 0; b < P.batch; ++b) {
            for (int j = 0; j < P.n; ++j) x[j] = in[(size_t)b * P.idist + j];
            fft(x.data(), 1, y.data(), P.n, -1);
            for (int k = 0; k <= P.n / 2; ++k) out[(size_t)b * P.odist + k] = double2{y[k].real(), y[k].imag()};
        }
    });
    return HIPFFT_SUCCESS;
}

hipfftResult hipfftExecZ2D(hipfftHandle plan, hipfftDoubleComplex* in, double* out)
{
    if (!plan || plan->type != HIPFFT_Z2D) return HIPFFT_INVALID_PLAN;
    const EmuFftPlan P = *plan;
    emu::enqueue(P.stream, [P, in, out] {
        std::vector<cd> x((size_t)P.n), y((size_t)P.n);
        for (int b = 0; b < P.batch; ++b) {
            const double2* X = in + (size_t)b * P.idist;
            for (int k = 0; k < P.n; ++k) {
                if (k <= P.n / 2) x[k] = cd(X[k].x, X[k].y);
                else x[k] = std::conj(cd(X[P.n - k].x, X[P.n - k].y));
            }
            x[0] = cd(x[0].real(), 0.0);
            if (P.n % 2 == 0) x[P.n / 2] = cd(x[P.n / 2].real(), 0.0);
            fft(x.data(), 1, y.data(), P.n, +1);
            for (int j = 0; j < P.n; ++j) out[(size_t)b * P.odist + j] = y[j].real();
        }
    });
    return HIPFFT_SUCCESS;
}

hipfftResult hipfftDestroy(hipfftHandle plan)
{
    delete plan;
    return HIPFFT_SUCCESS;
}
