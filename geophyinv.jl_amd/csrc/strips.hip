// Time-reversal kernels: boundary strips, snapshot negation and the imaging condition.
//
// strip_copy restates boundary_half{x,y,z}! (src/fdtd/boundary.jl:17-20) as used by
// boundary_save! (store = -field on the nbound = 3 planes just inside the PML at both ends
// of an axis, :217-264) and boundary_force! (field = store, :115-169).  image_acoustic_2d
// restates compute_gmodKI!, compute_gmodrho! and combine_gmodrho! (src/fdtd/gradient.jl:
// 27-56) with the velocity products formed on the fly instead of in full-grid buffers, and
// fuses save_tp! of the following step (src/fdtd/save_tp.jl:5-11) into the same pass.
#include "gpi_internal.h"

namespace gpi {

namespace {

constexpr int NBOUND = 3;   // _fd_nbound (src/GeoPhyInv.jl:92)

struct StripParams {
    float* field;
    float* store;
    int64_t sd0, sd1, sd2;   // store dims (the field's with the axis = 2*nbound)
    int64_t fn_axis;         // field length along the axis
    int64_t np;              // PML width on the axis' min face (used for both ends)
    int64_t p0, p1;          // pitches of the field
    int axis;
    int save;
};

__global__ __launch_bounds__(256) void strip_kernel(StripParams P)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = P.sd0 * P.sd1 * P.sd2;
    if (t >= n) return;
    int64_t i = t % P.sd0;
    const int64_t r = t / P.sd0;
    int64_t j = r % P.sd1;
    int64_t k = r / P.sd1;
    int64_t* ax = P.axis == 0 ? &i : P.axis == 1 ? &j : &k;
    const int64_t a = *ax;
    *ax = a < NBOUND ? P.np + a : P.fn_axis - P.np - 2 * NBOUND + a;
    const int64_t fo = i + P.p0 * (j + P.p1 * k);
    if (P.save) P.store[t] = P.field[fo] * -1.0f;   // copyto! then rmul!(b, -1)
    else P.field[fo] = P.store[t];
}

__global__ __launch_bounds__(256) void negate_kernel(float* dst, const float* src, int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = src[t] * -1.0f;
}

// Fields of the imaging condition, in the order p1, vx1, vz1, p2, vx2, vz2 (first and
// second wavefield).
struct ImageParams {
    const float* cur[6];   // after the step
    const float* prev[6];  // before the step (save_tp!, src/fdtd/save_tp.jl:5-11)
    float* next[6];        // receives `cur`: the previous values of the following step
    float* gK;
    float* gR;
    int nz, nx, ld, ncol, h;
    float dtI;
};

// One thread per cell of the pitched slab: the copy into `next` covers every field's
// extent, the gradients their own ranges.  `prev` is never written during the step, so the
// neighbour reads of the density term cannot race with the copy.
__global__ __launch_bounds__(256) void image2d_kernel(ImageParams P)
{
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    if (i >= P.ld || j >= P.ncol) return;
    const int o = i + P.ld * j;
    if (i < P.nz && j < P.nx) {
        // compute_gmodKI!: g + p2_prev * (p1_prev - p1) * dtI
        P.gK[o] = P.gK[o] + (P.prev[3][o] * (P.prev[0][o] - P.cur[0][o])) * P.dtI;
    }
    const int h = P.h;
    if (i < P.nz - 2 * h && j < P.nx - 2 * h) {
        // buffers of compute_gmodrho! at the two samples of @av_xi and of @av_zi
        const int a0 = (i + h) + P.ld * j, a1 = a0 + P.ld;
        const int b0 = i + P.ld * (j + h), b1 = b0 + 1;
        const float* vx1 = P.cur[1];
        const float* vx1p = P.prev[1];
        const float* vx2p = P.prev[4];
        const float* vz1 = P.cur[2];
        const float* vz1p = P.prev[2];
        const float* vz2p = P.prev[5];
        const float bx0 = (vx2p[a0] * (vx1[a0] - vx1p[a0])) * P.dtI;
        const float bx1 = (vx2p[a1] * (vx1[a1] - vx1p[a1])) * P.dtI;
        const float bz0 = (vz2p[b0] * (vz1[b0] - vz1p[b0])) * P.dtI;
        const float bz1 = (vz2p[b1] * (vz1[b1] - vz1p[b1])) * P.dtI;
        const int g = (i + h) + P.ld * (j + h);
        // combine_gmodrho!: @inn(g) = @inn(g) - @av_xi(vxbuffer) - @av_zi(vzbuffer)
        P.gR[g] = (P.gR[g] - (bx0 + bx1) * 0.5f) - (bz0 + bz1) * 0.5f;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) P.next[k][o] = P.cur[k][o];
}

// Imaging with the exact forward wavefield, one thread per cell of the pitched slab
// (z rows to the pitch, y columns to the y pitch, x to nx + h): the copy of the adjoint
// fields into the next set covers every field's extent, the gradients their own ranges.
// Each step's term is formed in FP32 with the reference's expressions and added to the
// FP64 accumulators.
struct ImageExactParams {
    ImageSets S;
    double* gK;
    double* gR;
    int nz, ny, nx;   // stress grid (ny = 1 in 2-D)
    int ld;           // z pitch
    int py;           // y pitch (1 in 2-D)
    int ncol;         // x extent of the slab (nx + h)
    int h;
    int ndim;
    float dtI;
};

__global__ __launch_bounds__(256) void image_exact_kernel(ImageExactParams P)
{
    const int nbz = (P.ld + 63) / 64;
    const int i = (int)(blockIdx.x % nbz) * 64 + threadIdx.x;
    const int64_t jk = (int64_t)(blockIdx.x / nbz) * 4 + threadIdx.y;
    const int j = (int)(jk % P.py), k = (int)(jk / P.py);
    if (i >= P.ld || k >= P.ncol) return;
    const int64_t sy = P.ld, sx = (int64_t)P.ld * P.py;
    const int64_t o = i + sy * j + sx * k;
    const ImageSets& S = P.S;
    if (P.gK && i < P.nz && j < P.ny && k < P.nx) {
        // compute_gmodKI!: p2_prev * (p1_prev - p1) * dtI, p1 = -p_forward
        const float t = (S.a2[0][o] * (S.pa[o] - S.pb[o])) * P.dtI;
        P.gK[o] = P.gK[o] + (double)t;
    }
    const int h = P.h;
    if (P.gR && i < P.nz - 2 * h && k < P.nx - 2 * h && (P.ndim == 2 || j < P.ny - 2 * h)) {
        // compute_gmodrho! buffers at the two samples of each @av_*i, then combine_gmodrho!
        const int jh = P.ndim == 2 ? 0 : j + h;
        auto buf = [&](int a, int64_t q) { return (S.a2[1 + a][q] * (S.va[a][q] - S.vb[a][q])) * P.dtI; };
        const int64_t ax = (i + h) + sy * jh + sx * k;     // @av_xi: x pair
        const int64_t az = i + sy * jh + sx * (k + h);     // @av_zi: z pair
        const float avx = (buf(0, ax) + buf(0, ax + sx)) * 0.5f;
        const float avz = (buf(2, az) + buf(2, az + 1)) * 0.5f;
        const int64_t g = (i + h) + sy * jh + sx * (k + h);
        double r = P.gR[g] - (double)avx;
        if (P.ndim == 3) {
            const int64_t ay = (i + h) + sy * j + sx * (k + h);   // @av_yi: y pair
            r = r - (double)((buf(1, ay) + buf(1, ay + sy)) * 0.5f);
        }
        P.gR[g] = r - (double)avz;
    }
    if (S.n2[0]) S.n2[0][o] = S.c2[0][o];
    if (P.gR)
        for (int a = 0; a < 3; ++a)
            if (S.n2[1 + a]) S.n2[1 + a][o] = S.c2[1 + a][o];
}

// ---- elastic imaging ----------------------------------------------------------------
// Derivation (interior, no CPML): with tau -= C E v and v -= B D tau, the Lagrange
// multipliers (Lambda of the stress updates, M of the velocity updates) satisfy the same
// scheme run backwards in time with tau~ = -C Lambda and v~ = B M, because E^T = -D on the
// staggered grids.  Then dL/dC = -Lambda^{n+1} . E v^{n+1/2} with E v = -C^{-1} dtau and
// dL/dB = -M . D tau with D tau = -B^{-1} dv, which the chain rule through
// dtM = dt(1/invlambda + 2/invmu), dtlambda = dt/invlambda, dtav*mu = dt/avg4(invmu) and
// dtinvav*irho = dt/avg2(rho) (src/fdtd/medium.jl:143-186) turns into the sums of
// ElasticImage: with d normal stresses, a = M + (d-1)L and b = M - L the eigenvalues of
// the normal block, Lambda . E v = T/(d a^2) + (P - T/d)/b^2.  The adjoint sources are
// the reference's 2 (dobs - dcal) = -dJ/dv (database.jl:371-380), so the adjoint run holds
// -(tau~, v~) and every sum enters negated.

// compute_illum!: illum += abs2(p) with the square in FP32 and the sum in FP64, over the
// whole pitched slab (the padding lanes are summed too and never read back)
__global__ __launch_bounds__(256) void illum_kernel(const float* __restrict__ p, double* __restrict__ acc, int64_t n)
{
    const int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (o >= n) return;
    const float v = p[o];
    acc[o] += (double)(v * v);
}

struct ElasticParams {
    ElasticImage E;
    int nz, ny, nx, ld, py, h;
    int64_t slab;
};

__device__ __forceinline__ void slab_coords(int64_t o, int ld, int py, int c[3])
{
    c[0] = (int)(o % ld);
    const int64_t r = o / ld;
    c[1] = (int)(r % py);
    c[2] = (int)(r / py);
}

__global__ __launch_bounds__(256) void image_elastic_kernel(ElasticParams P)
{
    const int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (o >= P.slab) return;
    int c[3];
    slab_coords(o, P.ld, P.py, c);
    const ElasticImage& E = P.E;
    if (c[0] < P.nz && c[1] < P.ny && c[2] < P.nx) {
        double ta = 0.0, td = 0.0, dot = 0.0;
        for (int q = 0; q < E.nn; ++q) {
            const double t = (double)E.an[q][o];
            const double d = (double)E.f1n[q][o] - (double)E.f0n[q][o];
            ta += t;
            td += d;
            dot += t * d;
        }
        E.T[o] += ta * td;
        E.P[o] += dot;
    }
    for (int q = 0; q < E.ns; ++q)
        if (c[0] < E.shi[q][0] && c[1] < E.shi[q][1] && c[2] < E.shi[q][2])
            E.S[q][o] += (double)E.as[q][o] * ((double)E.f1s[q][o] - (double)E.f0s[q][o]);
    for (int q = 0; q < E.nv; ++q)
        if (c[0] >= E.vlo[q][0] && c[0] < E.vhi[q][0] && c[1] >= E.vlo[q][1] && c[1] < E.vhi[q][1] &&
            c[2] >= E.vlo[q][2] && c[2] < E.vhi[q][2])
            E.R[q][o] += (double)E.av[q][o] * ((double)E.f1v[q][o] - (double)E.f0v[q][o]);
}

__global__ __launch_bounds__(256) void elastic_gradient_kernel(ElasticParams P, const float* __restrict__ il,
                                                               const float* __restrict__ im, double dt,
                                                               double* __restrict__ gl, double* __restrict__ gm,
                                                               double* __restrict__ gr)
{
    const int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (o >= P.slab) return;
    int c[3];
    slab_coords(o, P.ld, P.py, c);
    if (!(c[0] < P.nz && c[1] < P.ny && c[2] < P.nx)) return;
    const ElasticImage& E = P.E;
    const int64_t st[3] = {1, P.ld, (int64_t)P.ld * P.py};
    const double d = (double)E.nn;
    const double L = (double)il[o], M = (double)im[o];   // invlambda, invmu
    const double T = E.T[o], Pn = E.P[o];
    // a = dt (d/invlambda + 2/invmu), b = 2 dt/invmu
    const double ra = d + 2.0 * L / M;       // a invlambda / dt
    const double rb = d * M / L + 2.0;       // a invmu / dt
    gl[o] = -T / (dt * ra * ra);
    double g = 2.0 * T / (d * dt * rb * rb) + (Pn - T / d) / (2.0 * dt);
    // shear: each shear point spreads its term over the four cells of its invmu average
    double sh = 0.0;
    for (int q = 0; q < E.ns; ++q)
        for (int k = 0; k < 4; ++k) {
            int p[3] = {c[0] - E.sbase[q][0], c[1] - E.sbase[q][1], c[2] - E.sbase[q][2]};
            if (k & 1) --p[E.se1[q]];
            if (k & 2) --p[E.se2[q]];
            if (p[0] < 0 || p[1] < 0 || p[2] < 0 || p[0] >= E.shi[q][0] || p[1] >= E.shi[q][1] ||
                p[2] >= E.shi[q][2])
                continue;
            sh += E.S[q][p[0] + st[1] * p[1] + st[2] * p[2]];
        }
    gm[o] = -(g + sh / (4.0 * dt));
    // density: each velocity point spreads over the two cells of its rho average
    double r = 0.0;
    for (int q = 0; q < E.nv; ++q)
        for (int off = P.h; off >= P.h - 1; --off) {
            int v[3] = {c[0], c[1], c[2]};
            v[E.vaxis[q]] += off;
            if (v[0] < E.vlo[q][0] || v[0] >= E.vhi[q][0] || v[1] < E.vlo[q][1] || v[1] >= E.vhi[q][1] ||
                v[2] < E.vlo[q][2] || v[2] >= E.vhi[q][2])
                continue;
            r += E.R[q][v[0] + st[1] * v[1] + st[2] * v[2]];
        }
    gr[o] = r / (2.0 * dt);
}

// The normal-stress source of forward step n at each target: a_n = sum of weight x wavelet
// (the injection sum of the stress pass), times the trace of the adjoint normal stresses.
__global__ __launch_bounds__(256) void source_trace_kernel(const int64_t* __restrict__ toff,
                                                           const int32_t* __restrict__ tptr,
                                                           const int32_t* __restrict__ tsrc,
                                                           const float* __restrict__ tw, const float* __restrict__ wav,
                                                           int npts, int ntarget, int64_t n, const float* t0,
                                                           const float* t1, const float* t2, double* acc)
{
    const int t = (int)(blockIdx.x * 256 + threadIdx.x);
    if (t >= ntarget) return;
    float a = 0.0f;
    for (int k = tptr[t]; k < tptr[t + 1]; ++k) a = a + tw[k] * wav[n * npts + tsrc[k]];
    const int64_t c = toff[t];
    double tr = (double)t0[c] + (double)t1[c];
    if (t2) tr += (double)t2[c];
    acc[t] += tr * (double)a;
}

// With a = M + (d-1)L the trace eigenvalue of the normal block and Q = acc (adjoint run
// fields): the exact gradient adds (dt/il^2)(d-1)(M-L)/a^2 Q to invlambda and
// -(2dt/im^2)(d-1)L/a^2 Q to invmu at the target (the source term's own dependence on dtM
// and the Delta-tau form's source share).
__global__ __launch_bounds__(256) void source_correction_kernel(const int64_t* __restrict__ toff, int ntarget,
                                                                const double* __restrict__ acc,
                                                                const float* __restrict__ il,
                                                                const float* __restrict__ im, double dt, double d,
                                                                double* gl, double* gm)
{
    const int t = (int)(blockIdx.x * 256 + threadIdx.x);
    if (t >= ntarget) return;
    const int64_t c = toff[t];
    const double L0 = (double)il[c], M0 = (double)im[c];
    const double L = dt / L0, M = dt * (1.0 / L0 + 2.0 / M0), a = M + (d - 1.0) * L;
    gl[c] += dt / (L0 * L0) * (d - 1.0) * (M - L) / (a * a) * acc[t];
    gm[c] -= 2.0 * dt / (M0 * M0) * (d - 1.0) * L / (a * a) * acc[t];
}

ElasticParams elastic_params(const Model& m, const ElasticImage& E)
{
    ElasticParams P{};
    P.E = E;
    P.nz = (int)m.g.n[0];
    P.ny = m.g.ndim == 3 ? (int)m.g.n[1] : 1;
    P.nx = (int)m.g.n[2];
    P.ld = (int)m.pitch0;
    P.py = m.g.ndim == 3 ? (int)m.pitch1 : 1;
    P.h = m.h;
    P.slab = (int64_t)m.slab;
    return P;
}

}  // namespace

void image_elastic(const Model& m, const ElasticImage& E, hipStream_t s)
{
    const ElasticParams P = elastic_params(m, E);
    hipLaunchKernelGGL(image_elastic_kernel, dim3((unsigned)((P.slab + 255) / 256)), dim3(256), 0, s, P);
}

void elastic_gradient(const Model& m, const ElasticImage& E, double* g_il, double* g_im, double* g_rho,
                      hipStream_t s)
{
    const ElasticParams P = elastic_params(m, E);
    hipLaunchKernelGGL(elastic_gradient_kernel, dim3((unsigned)((P.slab + 255) / 256)), dim3(256), 0, s, P,
                       (const float*)m.coef[GPI_C_INVLAMBDA], (const float*)m.coef[GPI_C_INVMU], (double)m.g.dt,
                       g_il, g_im, g_rho);
}

void source_trace(const DevPoints& pts, int64_t n, const float* const an[3], int nn, double* acc, hipStream_t s)
{
    if (pts.ntarget <= 0) return;
    hipLaunchKernelGGL(source_trace_kernel, dim3((unsigned)((pts.ntarget + 255) / 256)), dim3(256), 0, s,
                       (const int64_t*)pts.toff, (const int32_t*)pts.tptr, (const int32_t*)pts.tsrc,
                       (const float*)pts.tw, (const float*)pts.wav, pts.npts, pts.ntarget, n, an[0], an[1],
                       nn > 2 ? an[2] : (const float*)nullptr, acc);
}

void source_correction(const Model& m, const DevPoints& pts, const double* acc, int nn, double* g_il, double* g_im,
                       hipStream_t s)
{
    if (pts.ntarget <= 0) return;
    hipLaunchKernelGGL(source_correction_kernel, dim3((unsigned)((pts.ntarget + 255) / 256)), dim3(256), 0, s,
                       (const int64_t*)pts.toff, pts.ntarget, acc, (const float*)m.coef[GPI_C_INVLAMBDA],
                       (const float*)m.coef[GPI_C_INVMU], (double)m.g.dt, (double)nn, g_il, g_im);
}

void illum_accumulate(const float* p, double* acc, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(illum_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, acc, n);
}

void image_acoustic_exact(const Model& m, const ImageSets& S, double* g_invK, double* g_rho, float dtI, hipStream_t s)
{
    ImageExactParams P{};
    P.S = S;
    P.gK = g_invK;
    P.gR = g_rho;
    P.nz = (int)m.g.n[0];
    P.ny = m.g.ndim == 3 ? (int)m.g.n[1] : 1;
    P.nx = (int)m.g.n[2];
    P.ld = (int)m.pitch0;
    P.py = m.g.ndim == 3 ? (int)m.pitch1 : 1;
    P.ncol = (int)(m.g.n[2] + m.h);
    P.h = m.h;
    P.ndim = m.g.ndim;
    P.dtI = dtI;
    const int64_t cols = (int64_t)P.py * P.ncol;
    const int64_t nb = (int64_t)((P.ld + 63) / 64) * ((cols + 3) / 4);
    hipLaunchKernelGGL(image_exact_kernel, dim3((unsigned)nb), dim3(64, 4), 0, s, P);
}

void strip_copy(const Model& m, float* field, const int64_t fn[3], int axis, int64_t np, float* store, bool save,
                hipStream_t s)
{
    StripParams P{};
    P.field = field;
    P.store = store;
    int64_t sd[3] = {fn[0], fn[1], fn[2]};
    sd[axis] = 2 * NBOUND;
    P.sd0 = sd[0];
    P.sd1 = sd[1];
    P.sd2 = sd[2];
    P.fn_axis = fn[axis];
    P.np = np;
    P.p0 = m.pitch0;
    P.p1 = m.pitch1;
    P.axis = axis;
    P.save = save ? 1 : 0;
    const int64_t n = sd[0] * sd[1] * sd[2];
    hipLaunchKernelGGL(strip_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, P);
}

void negate_copy(float* dst, const float* src, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(negate_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dst, src, n);
}

void image_acoustic_2d(const Model& m, const float* const cur[6], const float* const prev[6], float* const next[6],
                       float* g_invK, float* g_rho, float dtI, hipStream_t s)
{
    ImageParams P{};
    for (int k = 0; k < 6; ++k) {
        P.cur[k] = cur[k];
        P.prev[k] = prev[k];
        P.next[k] = next[k];
    }
    P.gK = g_invK;
    P.gR = g_rho;
    P.nz = (int)m.g.n[0];
    P.nx = (int)m.g.n[2];
    P.ld = (int)m.pitch0;
    P.ncol = (int)(m.g.n[2] + m.h);   // slab = pitch0 * (nx + h) in 2-D
    P.h = m.h;
    P.dtI = dtI;
    dim3 grd((P.ld + 63) / 64, (P.ncol + 3) / 4), blk(64, 4);
    hipLaunchKernelGGL(image2d_kernel, grd, blk, 0, s, P);
}

}  // namespace gpi
