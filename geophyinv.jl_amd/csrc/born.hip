// Secondary sources of Born modelling (FdtdAcoustic{Born}, src/fdtd/born.jl): after both
// wavefields' velocity passes the scattered velocities receive d(dtinvav*irho) times the
// background pressure gradient, after both stress passes the scattered pressure receives
// d(dtK) times the background velocity divergence.  These are the linearisation of the
// two updates (advance_acou.jl compute_v!, compute_p!) with respect to their coefficients;
// the reference's own secondary-source calls pass arguments that do not match those
// kernels, so the terms are re-derived here.  Derivatives are the plain stencils (no CPML
// memory): the perturbations vanish on the CPML strips.  One thread per cell of the pitched
// slab; each scattered cell is written by its own thread only.
#include "gpi_util.h"
#include "stencil.h"

namespace gpi {

namespace {

struct BornParams {
    const float* p1;
    const float* v1[3];   // background vx, vy, vz (vy null in 2-D)
    float* p2;
    float* v2[3];         // scattered
    const float* dc[3];   // d(dtinvavxirho), d(dtinvavyirho), d(dtinvavzirho)
    const float* dk;      // d(dtK)
    int nz, ny, nx;       // stress grid (ny = 1 in 2-D)
    int ld, py, ncol;     // z pitch, y pitch (1 in 2-D), x extent of the slab
    int ndim;
    float dxI, dyI, dzI;
};

template <int O>
__device__ __forceinline__ float dsamp(const float* a, int64_t off, int64_t stride)
{
    float v[O];
#pragma unroll
    for (int m = 0; m < O; ++m) v[m] = a[off + m * stride];
    return stencil<O>(v);
}

__device__ __forceinline__ bool cell(const BornParams& P, int& i, int& j, int& k)
{
    const int nbz = (P.ld + 63) / 64;
    i = (int)(blockIdx.x % nbz) * 64 + threadIdx.x;
    const int64_t jk = (int64_t)(blockIdx.x / nbz) * 4 + threadIdx.y;
    j = (int)(jk % P.py);
    k = (int)(jk / P.py);
    return i < P.ld && k < P.ncol;
}

// v2[inn] += d(dtinvav*irho) * D_*i(p1) (compute_v!, advance_acou.jl:273-284)
template <int O>
__global__ __launch_bounds__(256) void born_velocity_kernel(BornParams P)
{
    constexpr int H = O - 1;
    int i, j, k;
    if (!cell(P, i, j, k)) return;
    const bool d3 = P.ndim == 3;
    const int64_t sy = P.ld, sx = (int64_t)P.ld * P.py;
    const int64_t jy = d3 ? sy * (j + H) : 0;            // y offset of @inn
    const int64_t c = i + sy * j + sx * k;                 // derivative-grid point
    const int64_t o = (i + H) + jy + sx * (k + H);         // @inn of every velocity
    const bool ymid = !d3 || j < P.ny - 2 * H;
    if (i < P.nz - 2 * H && ymid && k < P.nx - H) {       // vx: D_xi(p)
        const float d = dsamp<O>(P.p1, (i + H) + jy + sx * k, sx) * P.dxI;
        P.v2[0][o] = P.v2[0][o] + P.dc[0][c] * d;
    }
    if (d3 && i < P.nz - 2 * H && j < P.ny - H && k < P.nx - 2 * H) {   // vy: D_yi(p)
        const float d = dsamp<O>(P.p1, (i + H) + sy * j + sx * (k + H), sy) * P.dyI;
        P.v2[1][o] = P.v2[1][o] + P.dc[1][c] * d;
    }
    if (i < P.nz - H && ymid && k < P.nx - 2 * H) {       // vz: D_zi(p)
        const float d = dsamp<O>(P.p1, i + jy + sx * (k + H), 1) * P.dzI;
        P.v2[2][o] = P.v2[2][o] + P.dc[2][c] * d;
    }
}

// p2 += (D_xa(vx1) + D_za(vz1) [+ D_ya(vy1)]) * d(dtK) (compute_p!, advance_acou.jl:303-313)
template <int O>
__global__ __launch_bounds__(256) void born_stress_kernel(BornParams P)
{
    int i, j, k;
    if (!cell(P, i, j, k)) return;
    if (i >= P.nz || j >= P.ny || k >= P.nx) return;
    const int64_t sy = P.ld, sx = (int64_t)P.ld * P.py;
    const int64_t o = i + sy * j + sx * k;
    const float dvx = dsamp<O>(P.v1[0], o, sx) * P.dxI;
    const float dvz = dsamp<O>(P.v1[2], o, 1) * P.dzI;
    float s = dvx + dvz;
    if (P.ndim == 3) s = s + dsamp<O>(P.v1[1], o, sy) * P.dyI;
    P.p2[o] = P.p2[o] + s * P.dk[o];
}

BornParams born_params(const Model& m, const State& w1, const State& w2, const float* const dc[3], const float* dk)
{
    BornParams P{};
    P.p1 = w1.field[GPI_P];
    P.p2 = w2.field[GPI_P];
    const int vf[3] = {GPI_VX, GPI_VY, GPI_VZ};
    for (int a = 0; a < 3; ++a) {
        P.v1[a] = w1.field[vf[a]];
        P.v2[a] = w2.field[vf[a]];
        P.dc[a] = dc[a];
    }
    P.dk = dk;
    P.nz = (int)m.g.n[0];
    P.ny = m.g.ndim == 3 ? (int)m.g.n[1] : 1;
    P.nx = (int)m.g.n[2];
    P.ld = (int)m.pitch0;
    P.py = m.g.ndim == 3 ? (int)m.pitch1 : 1;
    P.ncol = (int)(m.g.n[2] + m.h);
    P.ndim = m.g.ndim;
    P.dzI = m.g.dsI[0];
    P.dyI = m.g.dsI[1];
    P.dxI = m.g.dsI[2];
    return P;
}

}  // namespace

void born_sources(const Model& m, const State& w1, const State& w2, const float* const dc[3], const float* dk,
                  bool velocity, hipStream_t s)
{
    const BornParams P = born_params(m, w1, w2, dc, dk);
    const int64_t cols = (int64_t)P.py * P.ncol;
    const dim3 grd((unsigned)((int64_t)((P.ld + 63) / 64) * ((cols + 3) / 4))), blk(64, 4);
    switch (m.g.order) {
    case 2:
        if (velocity) hipLaunchKernelGGL(born_velocity_kernel<2>, grd, blk, 0, s, P);
        else hipLaunchKernelGGL(born_stress_kernel<2>, grd, blk, 0, s, P);
        break;
    case 4:
        if (velocity) hipLaunchKernelGGL(born_velocity_kernel<4>, grd, blk, 0, s, P);
        else hipLaunchKernelGGL(born_stress_kernel<4>, grd, blk, 0, s, P);
        break;
    case 6:
        if (velocity) hipLaunchKernelGGL(born_velocity_kernel<6>, grd, blk, 0, s, P);
        else hipLaunchKernelGGL(born_stress_kernel<6>, grd, blk, 0, s, P);
        break;
    default:
        if (velocity) hipLaunchKernelGGL(born_velocity_kernel<8>, grd, blk, 0, s, P);
        else hipLaunchKernelGGL(born_stress_kernel<8>, grd, blk, 0, s, P);
        break;
    }
}

}  // namespace gpi
