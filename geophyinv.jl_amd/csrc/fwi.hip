// The FWI outer loop next to the gradient: the inversion-grid projection and the L-BFGS
// direction, both in HBM.
//
// gpi_interp_apply: the separable linear map of apply_proj_matrix! (src/proj_mat.jl:208-226):
// y = x contracted with W_0 along axis 0, then W_1 along axis 1 (then W_2), each W_a a dense
// FP32 (nin x nout) matrix of 1-D interpolation weights (get_proj_matrix, :176-200), FP64
// sums.  Each W_a is banded (two nonzeros per column for the interpolation, a short run of
// rows for its transpose), so every output element sums the nonzero rows of its column,
// ascending.
//
// gpi_lbfgs_*: the two-loop recursion of L-BFGS (Nocedal & Wright, Algorithm 7.4) with the
// initial inverse Hessian gamma I, gamma = s'y / y'y of the newest pair: the history of m
// pairs (s, y) lives on the device in FP64, and every dot product is a fixed-order
// reduction (per-thread strided partial sums, then 256 partial sums, then one), so a
// direction is reproducible run to run.  The reference drives Optim.jl's LBFGS (memory 10)
// from src/fwi/fwi.jl:66-75.
#include <algorithm>
#include <deque>
#include <vector>

#include "gpi_util.h"

using namespace gpi;

namespace {

constexpr int kThreads = 256;
constexpr int kBlocks = 1024;                 // stage-1 threads: kThreads * kBlocks partials
constexpr int kPartials = kThreads * kBlocks;

__global__ __launch_bounds__(256) void dot_stage1(const double* __restrict__ a, const double* __restrict__ b,
                                                  double* __restrict__ part, int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    double acc = 0.0;
    for (int64_t i = t; i < n; i += kPartials) acc += a[i] * b[i];
    part[t] = acc;
}

__global__ __launch_bounds__(256) void dot_stage2(double* __restrict__ part)
{
    // 256 threads, each summing 1024 stage-1 partials (strided), in place at part[kPartials + t]
    const int t = threadIdx.x;
    double acc = 0.0;
    for (int i = t; i < kPartials; i += kThreads) acc += part[i];
    part[kPartials + t] = acc;
}

// One thread: the dot product from the 256 partials, then out = scale * (dot) + shift
// (alpha_i = rho_i s_i'q in the first loop, alpha_i - rho_i y_i'r in the second).
__global__ void dot_stage3(const double* __restrict__ part, double* __restrict__ out, double scale,
                           const double* __restrict__ shift, double* __restrict__ raw)
{
    if (threadIdx.x != 0) return;
    double acc = 0.0;
    for (int i = 0; i < kThreads; ++i) acc += part[kPartials + i];
    if (raw) *raw = acc;
    if (out) *out = (shift ? *shift : 0.0) - scale * acc;
}

// y += c * x with c read from device memory (c = coef * cscale)
__global__ __launch_bounds__(256) void axpy_dev(double* __restrict__ y, const double* __restrict__ x,
                                                const double* __restrict__ coef, double cscale, int64_t n)
{
    const double c = *coef * cscale;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
        y[i] += c * x[i];
}

// out = sa * a + sb * b
__global__ __launch_bounds__(256) void lincomb(double* __restrict__ out, const double* __restrict__ a, double sa,
                                               const double* __restrict__ b, double sb, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
        out[i] = sa * a[i] + sb * b[i];
}

// One pass of the separable map along `axis` of a (d0, d1, d2) array (z fastest): the
// output has nout in place of that axis.  Column j of W (nin x nout, column-major) has its
// nonzero rows in [lo[j], hi[j]).
__global__ __launch_bounds__(256) void interp_pass(const double* __restrict__ x, double* __restrict__ y,
                                                   const float* __restrict__ W, const int32_t* __restrict__ lo,
                                                   const int32_t* __restrict__ hi, int64_t d0, int64_t d1,
                                                   int64_t d2, int axis, int64_t nin, int64_t nout)
{
    int64_t o[3] = {d0, d1, d2};
    o[axis] = nout;
    const int64_t total = o[0] * o[1] * o[2];
    const int64_t stride_in = axis == 0 ? 1 : axis == 1 ? d0 : d0 * d1;
    for (int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x; e < total; e += (int64_t)gridDim.x * kThreads) {
        const int64_t i0 = e % o[0], r = e / o[0], i1 = r % o[1], i2 = r / o[1];
        const int64_t idx[3] = {i0, i1, i2};
        const int64_t j = idx[axis];
        int64_t base = i0 + d0 * (i1 + d1 * i2);   // input offset with the axis index at 0
        base -= idx[axis] * stride_in;
        double acc = 0.0;
        for (int64_t k = lo[j]; k < hi[j]; ++k) acc += (double)W[k + nin * j] * x[base + k * stride_in];
        y[e] = acc;
    }
}

int grid_for(int64_t n) { return (int)std::min<int64_t>(std::max<int64_t>((n + kThreads - 1) / kThreads, 1), 8192); }

}  // namespace

struct gpi_lbfgs {
    gpi_ctx* c = nullptr;
    int64_t n = 0;
    int mem = 0;
    // mem + 1 slots of (s, y): `order` lists the held pairs oldest first, `spare` receives
    // the next candidate, which joins the history only when its curvature is positive
    std::deque<int> order;
    std::vector<int> free_slots;
    int spare = 0;
    bool have_prev = false;
    std::vector<double> rho;           // 1 / s'y per slot
    double gamma = 1.0;
    double *S = nullptr, *Y = nullptr, *mprev = nullptr, *gprev = nullptr, *q = nullptr, *m = nullptr;
    double *part = nullptr, *alpha = nullptr, *scal = nullptr;   // partials; alpha per slot; sy, yy
    void clear()
    {
        order.clear();
        free_slots.clear();
        for (int k = mem; k >= 1; --k) free_slots.push_back(k);
        spare = 0;
        have_prev = false;
        gamma = 1.0;
    }
    std::vector<void*> allocs;
    uint64_t bytes = 0;
};

namespace {

void dot(const gpi_lbfgs& L, const double* a, const double* b, double* out, double scale, const double* shift,
         double* raw)
{
    hipStream_t s = L.c->stream;
    hipLaunchKernelGGL(dot_stage1, dim3(kBlocks), dim3(kThreads), 0, s, a, b, L.part, L.n);
    hipLaunchKernelGGL(dot_stage2, dim3(1), dim3(kThreads), 0, s, L.part);
    hipLaunchKernelGGL(dot_stage3, dim3(1), dim3(64), 0, s, (const double*)L.part, out, scale, shift, raw);
}

}  // namespace

extern "C" {

int gpi_lbfgs_create(gpi_ctx* c, int64_t n, int32_t memory, gpi_lbfgs** out)
{
    if (!c || !out || n < 1 || memory < 1 || memory > 64) return fail(c, GPI_EINVAL, "bad L-BFGS size");
    *out = nullptr;
    HIPTRY(c, hipSetDevice(c->device));
    gpi_lbfgs* L = new gpi_lbfgs();
    L->c = c;
    L->n = n;
    L->mem = memory;
    L->rho.assign(memory + 1, 0.0);
    L->clear();
    auto get = [&](double** p, size_t count) {
        return dev_alloc(c, &L->allocs, &L->bytes, (void**)p, sizeof(double) * count);
    };
    int rc = get(&L->S, (size_t)n * (memory + 1));
    if (!rc) rc = get(&L->Y, (size_t)n * (memory + 1));
    if (!rc) rc = get(&L->mprev, (size_t)n);
    if (!rc) rc = get(&L->gprev, (size_t)n);
    if (!rc) rc = get(&L->q, (size_t)n);
    if (!rc) rc = get(&L->m, (size_t)n);
    if (!rc) rc = get(&L->part, (size_t)kPartials + kThreads);
    if (!rc) rc = get(&L->alpha, (size_t)memory + 1);
    if (!rc) rc = get(&L->scal, 2);
    if (rc) {
        free_allocs(c, L->allocs, L->bytes);
        delete L;
        return rc;
    }
    *out = L;
    return GPI_OK;
}

void gpi_lbfgs_destroy(gpi_lbfgs* L)
{
    if (!L) return;
    (void)hipSetDevice(L->c->device);
    (void)hipStreamSynchronize(L->c->stream);
    free_allocs(L->c, L->allocs, L->bytes);
    delete L;
}

int gpi_lbfgs_reset(gpi_lbfgs* L)
{
    if (!L) return GPI_EINVAL;
    L->clear();
    return GPI_OK;
}

int gpi_lbfgs_direction(gpi_lbfgs* L, const double* m, const double* g, double* d, int32_t* pairs)
{
    if (!L || !m || !g || !d) return GPI_EINVAL;
    gpi_ctx* c = L->c;
    HIPTRY(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int64_t n = L->n;
    const size_t bytes = sizeof(double) * (size_t)n;
    const int nb = grid_for(n);
    HIPTRY(c, hipMemcpyAsync(L->m, m, bytes, hipMemcpyHostToDevice, s));
    HIPTRY(c, hipMemcpyAsync(L->q, g, bytes, hipMemcpyHostToDevice, s));
    if (L->have_prev) {   // the pair of the last step: s = m - m_prev, y = g - g_prev
        double* sk = L->S + (size_t)L->spare * n;
        double* yk = L->Y + (size_t)L->spare * n;
        hipLaunchKernelGGL(lincomb, dim3(nb), dim3(kThreads), 0, s, sk, (const double*)L->m, 1.0,
                           (const double*)L->mprev, -1.0, n);
        hipLaunchKernelGGL(lincomb, dim3(nb), dim3(kThreads), 0, s, yk, (const double*)L->q, 1.0,
                           (const double*)L->gprev, -1.0, n);
        dot(*L, sk, yk, nullptr, 0.0, nullptr, L->scal);
        dot(*L, yk, yk, nullptr, 0.0, nullptr, L->scal + 1);
        double sy_yy[2] = {0.0, 0.0};
        HIPTRY(c, hipMemcpyAsync(sy_yy, L->scal, sizeof sy_yy, hipMemcpyDeviceToHost, s));
        HIPTRY(c, hipStreamSynchronize(s));
        // keep the pair only with positive curvature (the update keeps H positive definite)
        if (sy_yy[0] > 2.220446049250313e-16 * sy_yy[1] && sy_yy[1] > 0.0) {
            L->rho[L->spare] = 1.0 / sy_yy[0];
            L->gamma = sy_yy[0] / sy_yy[1];
            L->order.push_back(L->spare);
            if ((int)L->order.size() > L->mem) {   // the oldest pair leaves; its slot is the spare
                L->spare = L->order.front();
                L->order.pop_front();
            } else {
                L->spare = L->free_slots.back();
                L->free_slots.pop_back();
            }
        }
    }
    HIPTRY(c, hipMemcpyAsync(L->mprev, L->m, bytes, hipMemcpyDeviceToDevice, s));
    HIPTRY(c, hipMemcpyAsync(L->gprev, L->q, bytes, hipMemcpyDeviceToDevice, s));
    // first loop, newest to oldest: alpha_i = rho_i s_i'q; q -= alpha_i y_i
    const int count = (int)L->order.size();
    for (int k = count - 1; k >= 0; --k) {
        const int i = L->order[(size_t)k];
        dot(*L, L->S + (size_t)i * n, L->q, L->alpha + i, -L->rho[i], nullptr, nullptr);
        hipLaunchKernelGGL(axpy_dev, dim3(nb), dim3(kThreads), 0, s, L->q, (const double*)(L->Y + (size_t)i * n),
                           (const double*)(L->alpha + i), -1.0, n);
    }
    // r = gamma q (in place), second loop, oldest to newest: r += s_i (alpha_i - rho_i y_i'r)
    hipLaunchKernelGGL(lincomb, dim3(nb), dim3(kThreads), 0, s, L->q, (const double*)L->q,
                       count ? L->gamma : 1.0, (const double*)L->q, 0.0, n);
    for (int k = 0; k < count; ++k) {
        const int i = L->order[(size_t)k];
        double* coef = L->scal;   // alpha_i - beta
        dot(*L, L->Y + (size_t)i * n, L->q, coef, L->rho[i], L->alpha + i, nullptr);
        hipLaunchKernelGGL(axpy_dev, dim3(nb), dim3(kThreads), 0, s, L->q, (const double*)(L->S + (size_t)i * n),
                           (const double*)coef, 1.0, n);
    }
    hipLaunchKernelGGL(lincomb, dim3(nb), dim3(kThreads), 0, s, L->q, (const double*)L->q, -1.0,
                       (const double*)L->q, 0.0, n);
    HIPTRY(c, hipGetLastError());
    HIPTRY(c, hipMemcpyAsync(d, L->q, bytes, hipMemcpyDeviceToHost, s));
    HIPTRY(c, hipStreamSynchronize(s));
    L->have_prev = true;
    if (pairs) *pairs = count;
    return GPI_OK;
}

int gpi_interp_apply(gpi_ctx* c, int32_t ndim, const int64_t nin[3], const int64_t nout[3], const float* const w[3],
                     const double* x, double* y)
{
    if (!c || !nin || !nout || !w || !x || !y || ndim < 1 || ndim > 3) return fail(c, GPI_EINVAL, "bad interpolation");
    for (int a = 0; a < ndim; ++a)
        if (nin[a] < 1 || nout[a] < 1 || !w[a]) return fail(c, GPI_EINVAL, "bad interpolation axis");
    HIPTRY(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    int64_t dims[3] = {1, 1, 1}, big = 1;
    for (int a = 0; a < ndim; ++a) dims[a] = nin[a];
    int64_t cur = dims[0] * dims[1] * dims[2];
    big = cur;
    for (int a = 0; a < ndim; ++a) {   // size after each pass
        int64_t t[3] = {dims[0], dims[1], dims[2]};
        for (int b = 0; b <= a; ++b) t[b] = nout[b];
        big = std::max(big, t[0] * t[1] * t[2]);
    }
    Scratch S(c);
    double *bufA = nullptr, *bufB = nullptr;
    int rc = S.alloc(&bufA, (size_t)big);
    if (!rc) rc = S.alloc(&bufB, (size_t)big);
    if (rc) return rc;
    HIPTRY(c, hipMemcpyAsync(bufA, x, sizeof(double) * (size_t)cur, hipMemcpyHostToDevice, s));
    for (int a = 0; a < ndim; ++a) {
        // the band of each column of W_a, and W_a itself
        const int64_t ni = nin[a], no = nout[a];
        std::vector<int32_t> lo((size_t)no), hi((size_t)no);
        for (int64_t j = 0; j < no; ++j) {
            int64_t l = 0, h = 0;
            for (int64_t k = 0; k < ni; ++k)
                if (w[a][k + ni * j] != 0.0f) {
                    if (h == 0) l = k;
                    h = k + 1;
                }
            lo[(size_t)j] = (int32_t)l;
            hi[(size_t)j] = (int32_t)h;
        }
        float* Wd = nullptr;
        int32_t *lod = nullptr, *hid = nullptr;
        std::vector<float> wv(w[a], w[a] + ni * no);
        if ((rc = S.put(&Wd, wv)) || (rc = S.put(&lod, lo)) || (rc = S.put(&hid, hi))) return rc;
        int64_t o[3] = {dims[0], dims[1], dims[2]};
        o[a] = no;
        const int64_t total = o[0] * o[1] * o[2];
        hipLaunchKernelGGL(interp_pass, dim3(grid_for(total)), dim3(kThreads), 0, s, (const double*)bufA, bufB,
                           (const float*)Wd, (const int32_t*)lod, (const int32_t*)hid, dims[0], dims[1], dims[2], a,
                           ni, no);
        HIPTRY(c, hipGetLastError());
        std::swap(bufA, bufB);
        dims[a] = no;
        cur = total;
    }
    HIPTRY(c, hipMemcpyAsync(y, bufA, sizeof(double) * (size_t)cur, hipMemcpyDeviceToHost, s));
    HIPTRY(c, hipStreamSynchronize(s));
    return GPI_OK;
}

}  // extern "C"
