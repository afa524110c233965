// Coefficient preparation on the device: the medium side of update!(pac, medium)
// (src/fdtd/medium.jl:128-186) for a medium given on its unpadded grid.
//
// One thread per cell of the extended grid reads the replicate-padded medium
// (padarray, src/media/media.jl:260-304: index clamped into the unpadded grid), forms
// the independent parameters (invK, or invlambda and invmu: media.jl:111-130) and writes
// every coefficient array whose grid contains the cell, with the same FP32 expressions
// (and the one FP64 promotion of dtM) as the host preparation, so the arrays are equal
// bit for bit to the host's.
#include "gpi_util.h"

namespace gpi {

namespace {

struct PrepParams {
    const float* vp;
    const float* vs;
    const float* rho;
    int mz, mx;          // unpadded medium size
    int lz, lx;          // pad widths at zmin and xmin
    int nz, nx, ld, h;   // extended grid, pitch, order - 1
    float dt;
    int physics;
    float* rho_out;
    float* cx;           // dtinvavxirho, (nz-2h) x (nx-h)
    float* cz;           // dtinvavzirho, (nz-h) x (nx-2h)
    float* dtk;          // acoustic dtK | elastic dtM
    float* invk;         // acoustic invK | elastic invlambda
    float* invmu;        // elastic invmu
    float* dtl;          // elastic dtlambda
    float* avmu;         // elastic dtavmu, (nz-h) x (nx-h)
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

struct Cell {
    float vp, vs, rho;
};

__device__ __forceinline__ Cell cell(const PrepParams& P, int i, int j)
{
    const int a = clampi(i - P.lz, 0, P.mz - 1) + P.mz * clampi(j - P.lx, 0, P.mx - 1);
    Cell c;
    c.vp = P.vp[a];
    c.vs = P.vs ? P.vs[a] : 0.0f;
    c.rho = P.rho[a];
    return c;
}

__device__ __forceinline__ float inv_mu(const PrepParams& P, int i, int j)
{
    const Cell c = cell(P, i, j);
    return 1.0f / ((c.vs * c.vs) * c.rho);
}

__global__ __launch_bounds__(256) void prep2d_kernel(PrepParams P)
{
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    if (i >= P.nz || j >= P.nx) return;
    const int h = P.h;
    const int o = i + P.ld * j;
    const Cell c = cell(P, i, j);
    P.rho_out[o] = c.rho;
    if (P.physics == 0) {
        const float ik = 1.0f / ((c.vp * c.vp) * c.rho);          // invK(vp, rho)
        P.invk[o] = ik;
        P.dtk[o] = (1.0f / ik) * P.dt;                             // dtK = dt * inv(invK)
    } else {
        const float il = 1.0f / (((c.vp * c.vp) - 2.0f * (c.vs * c.vs)) * c.rho);
        const float im = 1.0f / ((c.vs * c.vs) * c.rho);
        P.invk[o] = il;
        P.invmu[o] = im;
        P.dtl[o] = (1.0f / il) * P.dt;
        // inv(invλ) + 2.0 * inv(invμ): a host broadcast with a Float64 literal
        const double m = (double)(1.0f / il) + 2.0 * (double)(1.0f / im);
        P.dtk[o] = ((float)m) * P.dt;
        if (i < P.nz - h && j < P.nx - h) {                       // store_invav!: dt / @av(invmu)
            const float a = inv_mu(P, i, j), b = inv_mu(P, i + 1, j);
            const float cc = inv_mu(P, i, j + 1), d = inv_mu(P, i + 1, j + 1);
            P.avmu[o] = P.dt / ((((a + b) + cc) + d) * 0.25f);
        }
    }
    // dt / @av_xi(rho) on (nz-2h) x (nx-h) and dt / @av_zi(rho) on (nz-h) x (nx-2h)
    if (i < P.nz - 2 * h && j < P.nx - h) {
        const float r0 = cell(P, i + h, j).rho, r1 = cell(P, i + h, j + 1).rho;
        P.cx[o] = P.dt / ((r0 + r1) * 0.5f);
    }
    if (i < P.nz - h && j < P.nx - 2 * h) {
        const float r0 = cell(P, i, j + h).rho, r1 = cell(P, i + 1, j + h).rho;
        P.cz[o] = P.dt / ((r0 + r1) * 0.5f);
    }
}

}  // namespace

int prep_coefficients_2d(gpi_ctx* c, Model& m, const float* vp, const float* vs, const float* rho, const int64_t nm[3],
                         hipStream_t s)
{
    const gpi_grid& g = m.g;
    const bool el = g.physics == 1;
    const int ids_ac[] = {GPI_C_RHO, GPI_C_DTINVAVXIRHO, GPI_C_DTINVAVZIRHO, GPI_C_DTK, GPI_C_INVK};
    const int ids_el[] = {GPI_C_RHO,  GPI_C_DTINVAVXIRHO, GPI_C_DTINVAVZIRHO, GPI_C_DTM,
                          GPI_C_INVLAMBDA, GPI_C_INVMU, GPI_C_DTLAMBDA, GPI_C_DTAVMU};
    const int* ids = el ? ids_el : ids_ac;
    const int nid = el ? 8 : 5;
    for (int k = 0; k < nid; ++k)
        if (!m.coef[ids[k]]) {
            int rc = alloc_coef(c, m, ids[k]);
            if (rc) return rc;
        }
    PrepParams P{};
    P.vp = vp;
    P.vs = vs;
    P.rho = rho;
    P.mz = (int)nm[0];
    P.mx = (int)nm[2];
    P.lz = (g.pml_mask & GPI_FACE(0, 0)) ? g.npml : 0;
    P.lx = (g.pml_mask & GPI_FACE(2, 0)) ? g.npml : 0;
    P.nz = (int)g.n[0];
    P.nx = (int)g.n[2];
    P.ld = (int)m.pitch0;
    P.h = m.h;
    P.dt = g.dt;
    P.physics = g.physics;
    P.rho_out = m.coef[GPI_C_RHO];
    P.cx = m.coef[GPI_C_DTINVAVXIRHO];
    P.cz = m.coef[GPI_C_DTINVAVZIRHO];
    if (el) {
        P.dtk = m.coef[GPI_C_DTM];
        P.invk = m.coef[GPI_C_INVLAMBDA];
        P.invmu = m.coef[GPI_C_INVMU];
        P.dtl = m.coef[GPI_C_DTLAMBDA];
        P.avmu = m.coef[GPI_C_DTAVMU];
    } else {
        P.dtk = m.coef[GPI_C_DTK];
        P.invk = m.coef[GPI_C_INVK];
    }
    dim3 grd((P.nz + 63) / 64, (P.nx + 3) / 4), blk(64, 4);
    hipLaunchKernelGGL(prep2d_kernel, grd, blk, 0, s, P);
    HIPTRY(c, hipGetLastError());
    return GPI_OK;
}

}  // namespace gpi
