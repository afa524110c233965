// Source-signature estimation on the device: the convolution operator of src/Conv.jl
// (Pconv with d = g * s, d and g of nt samples and ntr traces, s of nt samples shared by
// all traces) by batched real FFTs (hipFFT), and the LSMR solve of update!(paw, pac)
// (src/fwi/fwi.jl:37-64: IterativeSolvers.lsmr! on Conv.operator(pa, G()), from zero).
//
// Signals and lags (Conv.jl:83-101, 154-224): d and g hold lags 0 .. nt-1; s holds
// nnl = nt - 1 - npl negative lags (most negative first), lag 0 at index nnl, then npl =
// (nt - 1) / 2 positive lags.  The FFT length np2 is the smallest 2-3-5-7-smooth number at
// or above len(s) + len(g) - 1 = 2 nt - 1 (DSP.nextfastfft), so the circular convolution
// equals the linear one.  Forward: S = rfft(pad(s)), D_tr = G_tr S, d = trunc(irfft(D)).
// Adjoint (mod!(pa, S()) with one filter shared by the traces): s = trunc(irfft(sum_tr
// D_tr conj(G_tr))).  Everything is FP64; the sum over traces and every norm are
// fixed-order reductions, so a solve is reproducible run to run.
//
// LSMR follows Fong and Saunders, "LSMR: an iterative algorithm for sparse least-squares
// problems", SIAM J. Sci. Comput. 33 (2011), with damping 0: Golub-Kahan
// bidiagonalisation, the two plane-rotation sequences of the paper's Section 2, the
// estimates of ||r||, ||A'r||, ||A|| and cond(A) of Section 5, and its stopping rules with
// tolerances atol, btol and conlim.  The long vector u (nt x ntr) stays in HBM; the
// filter-sized vectors (v, h, hbar, x: nt each) and the scalar recurrences stay on the host.
#include <hipfft/hipfft.h>

#include <cmath>
#include <vector>

#include "gpi_util.h"

using namespace gpi;

namespace {

constexpr int kThreads = 256;
constexpr int kRedBlocks = 1024;
constexpr int kChunks = 64;          // trace chunks of the adjoint's sum over traces

int grid_of(int64_t n) { return (int)std::min<int64_t>(std::max<int64_t>((n + kThreads - 1) / kThreads, 1), 16384); }

// out[t, tr] = scale x[t, tr] for t < nt, 0 up to np2 (lags 0 .. nt-1 padded, Conv.jl:170-197)
__global__ __launch_bounds__(256) void pad_traces(const double* __restrict__ x, double* __restrict__ out, int64_t nt,
                                                  int64_t np2, int64_t ntr, double scale)
{
    const int64_t n = np2 * ntr;
    for (int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x; e < n; e += (int64_t)gridDim.x * kThreads) {
        const int64_t t = e % np2, tr = e / np2;
        out[e] = t < nt ? scale * x[t + nt * tr] : 0.0;
    }
}

// The filter padded: lag 0 at 0, positive lags after it, negative lags wrapped to the end.
__global__ __launch_bounds__(256) void pad_filter(const double* __restrict__ s, double* __restrict__ out, int64_t nnl,
                                                  int64_t npl, int64_t np2)
{
    for (int64_t q = (int64_t)blockIdx.x * kThreads + threadIdx.x; q < np2; q += (int64_t)gridDim.x * kThreads) {
        double v = 0.0;
        if (q <= npl) v = s[nnl + q];
        else if (q >= np2 - nnl) v = s[q - (np2 - nnl)];
        out[q] = v;
    }
}

// out[f, tr] = G[f, tr] S[f]
__global__ __launch_bounds__(256) void spec_mul(const double2* __restrict__ G, const double2* __restrict__ S,
                                                double2* __restrict__ out, int64_t nf, int64_t ntr)
{
    const int64_t n = nf * ntr;
    for (int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x; e < n; e += (int64_t)gridDim.x * kThreads) {
        const double2 g = G[e], s = S[e % nf];
        out[e] = double2{g.x * s.x - g.y * s.y, g.x * s.y + g.y * s.x};
    }
}

// part[f, k] = sum over the traces of chunk k (ascending) of D[f, tr] conj(G[f, tr])
__global__ __launch_bounds__(256) void spec_corr_part(const double2* __restrict__ D, const double2* __restrict__ G,
                                                      double2* __restrict__ part, int64_t nf, int64_t ntr, int64_t per)
{
    const int64_t f = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const int64_t k = blockIdx.y;
    if (f >= nf) return;
    const int64_t t0 = k * per, t1 = std::min<int64_t>(t0 + per, ntr);
    double re = 0.0, im = 0.0;
    for (int64_t tr = t0; tr < t1; ++tr) {
        const double2 d = D[f + nf * tr], g = G[f + nf * tr];
        re += d.x * g.x + d.y * g.y;
        im += d.y * g.x - d.x * g.y;
    }
    part[f + nf * k] = double2{re, im};
}

// out[f] = sum over the chunks (ascending) of part[f, k]
__global__ __launch_bounds__(256) void spec_corr_sum(const double2* __restrict__ part, double2* __restrict__ out,
                                                     int64_t nf, int64_t nchunks)
{
    for (int64_t f = (int64_t)blockIdx.x * kThreads + threadIdx.x; f < nf; f += (int64_t)gridDim.x * kThreads) {
        double re = 0.0, im = 0.0;
        for (int64_t k = 0; k < nchunks; ++k) {
            re += part[f + nf * k].x;
            im += part[f + nf * k].y;
        }
        out[f] = double2{re, im};
    }
}

// u[t, tr] = r[t, tr] / np2 - alpha u[t, tr] (the inverse FFT is unnormalised); alpha = 0
// writes without reading u
__global__ __launch_bounds__(256) void trunc_traces(const double* __restrict__ r, double* __restrict__ u, int64_t nt,
                                                    int64_t np2, int64_t ntr, double inv_n, double alpha)
{
    const int64_t n = nt * ntr;
    for (int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x; e < n; e += (int64_t)gridDim.x * kThreads) {
        const int64_t t = e % nt, tr = e / nt;
        const double a = r[t + np2 * tr] * inv_n;
        u[e] = alpha == 0.0 ? a : a - alpha * u[e];
    }
}

// the filter from its padded form (Conv.jl:199-224), scaled by 1/np2
__global__ __launch_bounds__(256) void trunc_filter(const double* __restrict__ r, double* __restrict__ s, int64_t nnl,
                                                    int64_t npl, int64_t np2, double inv_n)
{
    const int64_t n = nnl + npl + 1;
    for (int64_t q = (int64_t)blockIdx.x * kThreads + threadIdx.x; q < n; q += (int64_t)gridDim.x * kThreads)
        s[q] = (q < nnl ? r[np2 - nnl + q] : r[q - nnl]) * inv_n;
}

__global__ __launch_bounds__(256) void sumsq_stage1(const double* __restrict__ a, double* __restrict__ part, int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    double acc = 0.0;
    for (int64_t i = t; i < n; i += (int64_t)kRedBlocks * kThreads) acc += a[i] * a[i];
    part[t] = acc;
}

// 256 threads, each summing the stage-1 partials strided by 256, into part[kRedBlocks * 256 + t]
__global__ __launch_bounds__(256) void sumsq_stage2(double* __restrict__ part)
{
    const int t = threadIdx.x;
    double acc = 0.0;
    for (int i = t; i < kRedBlocks * kThreads; i += kThreads) acc += part[i];
    part[kRedBlocks * kThreads + t] = acc;
}

// one thread: the sum of the 256 stage-2 sums, into part[kRedBlocks * 256 + 256]
__global__ void sumsq_stage3(double* __restrict__ part)
{
    if (threadIdx.x != 0) return;
    double s = 0.0;
    for (int i = 0; i < kThreads; ++i) s += part[kRedBlocks * kThreads + i];
    part[kRedBlocks * kThreads + kThreads] = s;
}

__global__ __launch_bounds__(256) void scale_vec(double* __restrict__ u, double c, int64_t n)
{
    for (int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x; e < n; e += (int64_t)gridDim.x * kThreads) u[e] *= c;
}

int64_t next_fast(int64_t n)
{
    for (int64_t m = std::max<int64_t>(n, 1);; ++m) {
        int64_t k = m;
        for (int p : {2, 3, 5, 7})
            while (k % p == 0) k /= p;
        if (k == 1) return m;
    }
}

// The operator of one filter shared by ntr traces of g, with its FFT plans and buffers.
struct ConvOp {
    gpi_ctx* c = nullptr;
    int64_t nt = 0, ntr = 0, np2 = 0, nf = 0, nnl = 0, npl = 0, nchunks = 1;
    hipfftHandle many_r2c = 0, many_c2r = 0, one_r2c = 0, one_c2r = 0;
    double* G = nullptr;    // spectra of the padded g traces, nf x ntr complex
    double* R = nullptr;    // real work, np2 x ntr
    double* C = nullptr;    // complex work, nf x ntr
    double* P = nullptr;    // per-chunk partial spectra, nf x nchunks complex
    double* r1 = nullptr;   // one padded filter, np2
    double* c1 = nullptr;   // one spectrum, nf complex
    double* s1 = nullptr;   // one filter, nt
    double* red = nullptr;  // reduction partials
    std::vector<void*> allocs;

    ~ConvOp()
    {
        for (hipfftHandle* h : {&many_r2c, &many_c2r, &one_r2c, &one_c2r})
            if (*h) hipfftDestroy(*h);
        for (void* p : allocs) (void)hipFree(p);
    }

    int alloc(double** p, size_t n)
    {
        hipError_t e = hipMalloc((void**)p, n * sizeof(double));
        if (e != hipSuccess) return fail(c, GPI_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        allocs.push_back(*p);
        return GPI_OK;
    }

    int plan(hipfftHandle* h, hipfftType type, int64_t batch)
    {
        int n[1] = {(int)np2};
        const bool r2c = type == HIPFFT_D2Z;
        const int idist = (int)(r2c ? np2 : nf), odist = (int)(r2c ? nf : np2);
        hipfftResult r = hipfftPlanMany(h, 1, n, nullptr, 1, idist, nullptr, 1, odist, type, (int)batch);
        if (r != HIPFFT_SUCCESS) return fail(c, GPI_EHIP, "hipfftPlanMany failed (" + std::to_string((int)r) + ")");
        r = hipfftSetStream(*h, c->stream);
        if (r != HIPFFT_SUCCESS) return fail(c, GPI_EHIP, "hipfftSetStream failed");
        return GPI_OK;
    }

    int fft(hipfftHandle h, bool r2c, double* in, double* out)
    {
        const hipfftResult r = r2c ? hipfftExecD2Z(h, in, (hipfftDoubleComplex*)out)
                                   : hipfftExecZ2D(h, (hipfftDoubleComplex*)in, out);
        if (r != HIPFFT_SUCCESS) return fail(c, GPI_EHIP, "hipfft execution failed (" + std::to_string((int)r) + ")");
        return GPI_OK;
    }

    // g: nt x ntr on the host
    int setup(gpi_ctx* ctx, int64_t nt_, int64_t ntr_, const double* g)
    {
        c = ctx;
        nt = nt_;
        ntr = ntr_;
        np2 = next_fast(2 * nt - 1);
        nf = np2 / 2 + 1;
        npl = (nt - 1) / 2;
        nnl = nt - 1 - npl;
        nchunks = std::min<int64_t>(kChunks, ntr);
        if (np2 > (1LL << 30) || ntr > (1LL << 30)) return fail(c, GPI_EINVAL, "convolution too large");
        int rc;
        if ((rc = alloc(&G, 2 * nf * ntr)) || (rc = alloc(&R, np2 * ntr)) || (rc = alloc(&C, 2 * nf * ntr)) ||
            (rc = alloc(&P, 2 * nf * nchunks)) || (rc = alloc(&r1, np2)) || (rc = alloc(&c1, 2 * nf)) ||
            (rc = alloc(&s1, nt)) || (rc = alloc(&red, (size_t)kRedBlocks * kThreads + kThreads + 1)))
            return rc;
        if ((rc = plan(&many_r2c, HIPFFT_D2Z, ntr)) || (rc = plan(&many_c2r, HIPFFT_Z2D, ntr)) ||
            (rc = plan(&one_r2c, HIPFFT_D2Z, 1)) || (rc = plan(&one_c2r, HIPFFT_Z2D, 1)))
            return rc;
        // G = rfft(pad(g)); g goes through the d-sized work buffer of the truncated layout
        double* gd = nullptr;
        if ((rc = alloc(&gd, nt * ntr))) return rc;
        HIPTRY(c, hipMemcpyAsync(gd, g, sizeof(double) * nt * ntr, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(pad_traces, dim3(grid_of(np2 * ntr)), dim3(kThreads), 0, c->stream, gd, R, nt, np2, ntr, 1.0);
        return fft(many_r2c, true, R, G);
    }

    // u = (g * s) - alpha u, s on the device (nt)
    int forward(const double* s_dev, double* u, double alpha)
    {
        hipLaunchKernelGGL(pad_filter, dim3(grid_of(np2)), dim3(kThreads), 0, c->stream, s_dev, r1, nnl, npl, np2);
        int rc = fft(one_r2c, true, r1, c1);
        if (rc) return rc;
        hipLaunchKernelGGL(spec_mul, dim3(grid_of(nf * ntr)), dim3(kThreads), 0, c->stream, (const double2*)G,
                           (const double2*)c1, (double2*)C, nf, ntr);
        if ((rc = fft(many_c2r, false, C, R))) return rc;
        hipLaunchKernelGGL(trunc_traces, dim3(grid_of(nt * ntr)), dim3(kThreads), 0, c->stream, R, u, nt, np2, ntr,
                           1.0 / (double)np2, alpha);
        return GPI_OK;
    }

    // s_dev = sum over traces of (scale d) correlated with g, d on the device (nt x ntr)
    int adjoint(const double* d, double scale, double* s_dev)
    {
        hipLaunchKernelGGL(pad_traces, dim3(grid_of(np2 * ntr)), dim3(kThreads), 0, c->stream, d, R, nt, np2, ntr,
                           scale);
        int rc = fft(many_r2c, true, R, C);
        if (rc) return rc;
        const int64_t per = (ntr + nchunks - 1) / nchunks;
        hipLaunchKernelGGL(spec_corr_part, dim3((unsigned)((nf + kThreads - 1) / kThreads), (unsigned)nchunks),
                           dim3(kThreads), 0, c->stream, (const double2*)C, (const double2*)G, (double2*)P, nf, ntr,
                           per);
        hipLaunchKernelGGL(spec_corr_sum, dim3(grid_of(nf)), dim3(kThreads), 0, c->stream, (const double2*)P,
                           (double2*)c1, nf, nchunks);
        if ((rc = fft(one_c2r, false, c1, r1))) return rc;
        hipLaunchKernelGGL(trunc_filter, dim3(grid_of(nt)), dim3(kThreads), 0, c->stream, r1, s_dev, nnl, npl, np2,
                           1.0 / (double)np2);
        return GPI_OK;
    }

    // ||u||^2 (fixed order), waited for
    int sumsq(const double* u, int64_t n, double* out)
    {
        hipLaunchKernelGGL(sumsq_stage1, dim3(kRedBlocks), dim3(kThreads), 0, c->stream, u, red, n);
        hipLaunchKernelGGL(sumsq_stage2, dim3(1), dim3(kThreads), 0, c->stream, red);
        hipLaunchKernelGGL(sumsq_stage3, dim3(1), dim3(64), 0, c->stream, red);
        HIPTRY(c, hipMemcpyAsync(out, red + (size_t)kRedBlocks * kThreads + kThreads, sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream));
        HIPTRY(c, hipStreamSynchronize(c->stream));
        return GPI_OK;
    }
};

double host_norm(const std::vector<double>& x)
{
    double s = 0.0;
    for (double v : x) s += v * v;
    return std::sqrt(s);
}

// Stable plane rotation [c s; -s c] [a; b] = [r; 0] (the SymOrtho of the paper).
void sym_ortho(double a, double b, double& c, double& s, double& r)
{
    auto sign = [](double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); };
    if (b == 0.0) {
        c = sign(a);
        s = 0.0;
        r = std::fabs(a);
    } else if (a == 0.0) {
        c = 0.0;
        s = sign(b);
        r = std::fabs(b);
    } else if (std::fabs(b) > std::fabs(a)) {
        const double tau = a / b;
        s = sign(b) / std::sqrt(1.0 + tau * tau);
        c = s * tau;
        r = b / s;
    } else {
        const double tau = b / a;
        c = sign(a) / std::sqrt(1.0 + tau * tau);
        s = c * tau;
        r = a / c;
    }
}

}  // namespace

extern "C" int gpi_conv_apply(gpi_ctx* c, int64_t nt, int64_t ntr, const double* g, int32_t mode, const double* in,
                              double* out)
{
    if (!c) return GPI_EINVAL;
    if (nt < 1 || ntr < 1 || !g || !in || !out || (mode != 0 && mode != 1))
        return fail(c, GPI_EINVAL, "gpi_conv_apply: bad arguments");
    HIPTRY(c, hipSetDevice(c->device));
    ConvOp op;
    int rc = op.setup(c, nt, ntr, g);
    if (rc) return rc;
    double *x = nullptr, *y = nullptr;
    if ((rc = op.alloc(&x, mode == 0 ? nt : nt * ntr)) || (rc = op.alloc(&y, mode == 0 ? nt * ntr : nt))) return rc;
    const size_t nin = mode == 0 ? nt : nt * ntr, nout = mode == 0 ? nt * ntr : nt;
    HIPTRY(c, hipMemcpyAsync(x, in, sizeof(double) * nin, hipMemcpyHostToDevice, c->stream));
    rc = mode == 0 ? op.forward(x, y, 0.0) : op.adjoint(x, 1.0, y);
    if (rc) return rc;
    HIPTRY(c, hipMemcpyAsync(out, y, sizeof(double) * nout, hipMemcpyDeviceToHost, c->stream));
    HIPTRY(c, hipStreamSynchronize(c->stream));
    return GPI_OK;
}

extern "C" int gpi_conv_lsmr(gpi_ctx* c, int64_t nt, int64_t ntr, const double* g, const double* d, int32_t maxiter,
                             double atol, double btol, double conlim, double* s_out, double* gs_out, int32_t* iters,
                             int32_t* istop_out)
{
    if (!c) return GPI_EINVAL;
    if (nt < 1 || ntr < 1 || !g || !d || !s_out || maxiter < 0 || !(conlim > 0.0))
        return fail(c, GPI_EINVAL, "gpi_conv_lsmr: bad arguments");
    HIPTRY(c, hipSetDevice(c->device));
    ConvOp op;
    int rc = op.setup(c, nt, ntr, g);
    if (rc) return rc;
    const int64_t m = nt * ntr;
    double *u = nullptr, *vd = nullptr;
    if ((rc = op.alloc(&u, m)) || (rc = op.alloc(&vd, nt))) return rc;
    std::vector<double> v(nt), h(nt), hbar(nt, 0.0), x(nt, 0.0), w(nt);

    // A v on the device into u (u = A v - alpha u); A' u back to the host into w
    auto apply_a = [&](const std::vector<double>& vec, double alpha) -> int {
        HIPTRY(c, hipMemcpyAsync(vd, vec.data(), sizeof(double) * nt, hipMemcpyHostToDevice, c->stream));
        return op.forward(vd, u, alpha);
    };
    auto apply_at = [&](double scale) -> int {
        int r = op.adjoint(u, scale, op.s1);
        if (r) return r;
        HIPTRY(c, hipMemcpyAsync(w.data(), op.s1, sizeof(double) * nt, hipMemcpyDeviceToHost, c->stream));
        HIPTRY(c, hipStreamSynchronize(c->stream));
        return GPI_OK;
    };

    // u = b, beta = ||b||
    HIPTRY(c, hipMemcpyAsync(u, d, sizeof(double) * m, hipMemcpyHostToDevice, c->stream));
    double bb = 0.0;
    if ((rc = op.sumsq(u, m, &bb))) return rc;
    double beta = std::sqrt(bb);
    const double normb = beta;
    int istop = 0, itn = 0;
    if (beta > 0.0) {
        // v = A' u / beta (u kept unnormalised until the first iteration scales it)
        if ((rc = apply_at(1.0 / beta))) return rc;
        hipLaunchKernelGGL(scale_vec, dim3(grid_of(m)), dim3(kThreads), 0, c->stream, u, 1.0 / beta, m);
        v = w;
        double alpha = host_norm(v);
        if (alpha > 0.0)
            for (double& e : v) e /= alpha;
        if (alpha > 0.0) {
            double zetabar = alpha * beta, alphabar = alpha, rho = 1.0, rhobar = 1.0, cbar = 1.0, sbar = 0.0;
            h = v;
            // ||r|| estimate
            double betadd = beta, betad = 0.0, rhodold = 1.0, tautildeold = 0.0, thetatilde = 0.0, zeta = 0.0,
                   dsum = 0.0;
            // ||A|| and cond(A) estimates
            double normA2 = alpha * alpha, maxrbar = 0.0, minrbar = 1e100;
            const double ctol = 1.0 / conlim;
            while (itn < maxiter) {
                ++itn;
                // bidiagonalisation: beta u = A v - alpha u, alpha v = A' u - beta v
                if ((rc = apply_a(v, alpha))) return rc;
                if ((rc = op.sumsq(u, m, &bb))) return rc;
                beta = std::sqrt(bb);
                if (beta > 0.0) {
                    hipLaunchKernelGGL(scale_vec, dim3(grid_of(m)), dim3(kThreads), 0, c->stream, u, 1.0 / beta, m);
                    if ((rc = apply_at(1.0))) return rc;
                    for (int64_t i = 0; i < nt; ++i) v[i] = w[i] - beta * v[i];
                    alpha = host_norm(v);
                    if (alpha > 0.0)
                        for (double& e : v) e /= alpha;
                }
                // rotation Qhat with the damping row (damping 0: it only takes |alphabar|)
                double chat, shat, alphahat;
                sym_ortho(alphabar, 0.0, chat, shat, alphahat);
                // rotation P_k
                const double rhoold = rho;
                double ck, sk;
                sym_ortho(alphahat, beta, ck, sk, rho);
                const double thetanew = sk * alpha;
                alphabar = ck * alpha;
                // rotation Pbar_k
                const double rhobarold = rhobar, zetaold = zeta;
                const double thetabar = sbar * rho;
                const double rhotemp = cbar * rho;
                sym_ortho(cbar * rho, thetanew, cbar, sbar, rhobar);
                zeta = cbar * zetabar;
                zetabar = -sbar * zetabar;
                // h, hbar, x
                const double fh = thetabar * rho / (rhoold * rhobarold);
                const double fx = zeta / (rho * rhobar);
                const double fv = thetanew / rho;
                for (int64_t i = 0; i < nt; ++i) {
                    hbar[i] = h[i] - fh * hbar[i];
                    x[i] += fx * hbar[i];
                    h[i] = v[i] - fv * h[i];
                }
                // ||r||: the rotations applied to the right-hand side (Section 5.2)
                const double betaacute = chat * betadd;
                const double betacheck = -shat * betadd;
                const double betahat = ck * betaacute;
                betadd = -sk * betaacute;
                const double thetatildeold = thetatilde;
                double ctildeold, stildeold, rhotildeold;
                sym_ortho(rhodold, thetabar, ctildeold, stildeold, rhotildeold);
                thetatilde = stildeold * rhobar;
                rhodold = ctildeold * rhobar;
                betad = -stildeold * betad + ctildeold * betahat;
                tautildeold = (zetaold - thetatildeold * tautildeold) / rhotildeold;
                const double taud = (zeta - thetatilde * tautildeold) / rhodold;
                dsum += betacheck * betacheck;
                const double normr = std::sqrt(dsum + (betad - taud) * (betad - taud) + betadd * betadd);
                // ||A|| (Frobenius estimate) and cond(A)
                normA2 += beta * beta;
                const double normA = std::sqrt(normA2);
                normA2 += alpha * alpha;
                maxrbar = std::max(maxrbar, rhobarold);
                if (itn > 1) minrbar = std::min(minrbar, rhobarold);
                const double condA = std::max(maxrbar, rhotemp) / std::min(minrbar, rhotemp);
                // stopping rules
                const double normar = std::fabs(zetabar);
                const double normx = host_norm(x);
                const double test1 = normr / normb;
                const double test2 = (normA * normr) != 0.0 ? normar / (normA * normr) : INFINITY;
                const double test3 = 1.0 / condA;
                const double t1 = test1 / (1.0 + normA * normx / normb);
                const double rtol = btol + atol * normA * normx / normb;
                if (itn >= maxiter) istop = 7;
                if (1.0 + test3 <= 1.0) istop = 6;
                if (1.0 + test2 <= 1.0) istop = 5;
                if (1.0 + t1 <= 1.0) istop = 4;
                if (test3 <= ctol) istop = 3;
                if (test2 <= atol) istop = 2;
                if (test1 <= rtol) istop = 1;
                if (istop > 0) break;
            }
        }
    }
    for (int64_t i = 0; i < nt; ++i) s_out[i] = x[i];
    if (gs_out) {
        if ((rc = apply_a(x, 0.0))) return rc;
        HIPTRY(c, hipMemcpyAsync(gs_out, u, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
        HIPTRY(c, hipStreamSynchronize(c->stream));
    }
    if (iters) *iters = itn;
    if (istop_out) *istop_out = istop;
    return GPI_OK;
}
