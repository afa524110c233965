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
    if (i < P.nz && j < P.ny && k < P.nx) {
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
    S.n2[0][o] = S.c2[0][o];
    if (P.gR)
        for (int a = 0; a < 3; ++a)
            if (S.n2[1 + a]) S.n2[1 + a][o] = S.c2[1 + a][o];
}

}  // namespace

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
