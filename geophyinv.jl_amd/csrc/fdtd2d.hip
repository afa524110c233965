// 2-D staggered-grid velocity-stress kernels for gfx950.
//
// One launch per half step.  The velocity pass fuses what the reference runs as
// compute_dp!/compute_dstress!, the CPML memory strips, compute_v!, the rigid-face kernels
// and add_velocity_source! (src/fdtd/advance_acou.jl:11-58, src/fdtd/advance_elastic.jl:
// 534-608, src/fdtd/cpml.jl:158-215, src/fdtd/dirichlet.jl:3-78, src/fdtd/source.jl:124-157);
// its extra workgroups record p (src/fdtd/propagate.jl:177).  The stress pass fuses
// compute_dv!, the CPML strips, compute_p!/compute_stressii!/compute_stressij!, the free
// surface and add_stress_source! (src/fdtd/advance_acou.jl:59-96, src/fdtd/advance_elastic.jl:
// 609-682, src/fdtd/source.jl:61-119); its extra workgroups record v (propagate.jl:208).
// Derivatives never leave registers.  Arithmetic keeps the reference's FP32 evaluation
// order (the library is built with -ffp-contract=off), so results match the unfused CPU
// restatement bit for bit.
//
// Layout: every wavefield and coefficient array of a model uses the same column pitch
// `ld` (floats), z contiguous.  A thread owns one derivative-grid point (i, j); lanes of a
// wavefront run along z so every load and store is a coalesced 256-byte row segment.
#include <cstdlib>

#include "fdtd2d.h"
#include "pass_extras.h"
#include "stencil.h"

namespace gpi {

namespace {

// One CPML memory update (cpml.jl:178-180): psi = b psi + a d, d = d kI + psi, with every
// load before the store (the compiler cannot tell that psi aliases none of the coefficient
// arrays, so a coefficient read after the store would wait for it).
__device__ __forceinline__ float cpml_step(float d, const Pml& q, float* m, int c)
{
    const float b = q.b[c], a = q.a[c], ki = q.kI[c], mo = *m;
    const float mv = b * mo + a * d;
    *m = mv;
    return d * ki + mv;
}

// CPML of a derivative that varies along x (column j of a derivative grid with q.nd
// columns).  Min face first, then max face, as the reference launches memorynpx! for xmin
// before xmax.
__device__ __forceinline__ float cpml_x(float d, const Pml& q, int i, int j, int npml, uint32_t mask)
{
    if ((mask & GPI_FACE(2, 0)) && j < npml) d = cpml_step(d, q, q.psi + i + q.ps2 * j, j);
    if ((mask & GPI_FACE(2, 1)) && j >= q.nd - npml) {
        const int c = j - (q.nd - npml) + npml;
        d = cpml_step(d, q, q.psi + i + q.ps2 * c, c);
    }
    return d;
}

// Same along z (row i of a derivative grid with q.nd rows); psi rows are the 2*npml strip.
__device__ __forceinline__ float cpml_z(float d, const Pml& q, int i, int j, int npml, uint32_t mask)
{
    if ((mask & GPI_FACE(0, 0)) && i < npml) d = cpml_step(d, q, q.psi + i + q.ps2 * j, i);
    if ((mask & GPI_FACE(0, 1)) && i >= q.nd - npml) {
        const int r = i - (q.nd - npml) + npml;
        d = cpml_step(d, q, q.psi + r + q.ps2 * j, r);
    }
    return d;
}

template <int O>
__device__ __forceinline__ float dz_at(const float* a, int off)
{
    float v[O];
#pragma unroll
    for (int k = 0; k < O; ++k) v[k] = a[off + k];
    return stencil<O>(v);
}

template <int O>
__device__ __forceinline__ float dx_at(const float* a, int off, int ld)
{
    float v[O];
#pragma unroll
    for (int k = 0; k < O; ++k) v[k] = a[off + k * ld];
    return stencil<O>(v);
}

// ---------------------------------------------------------------------------------------
// Points.  A thread owns derivative-grid point (i, j); `bk` is its injection bucket (the
// 64-row segment of column j), `it` the shot's time step and `ninj` 0 when the pass takes
// no sources (the trailing half step).
// ---------------------------------------------------------------------------------------

// Samples of a stencil along z (contiguous) or x (stride ld), loaded without evaluating.
template <int O>
__device__ __forceinline__ void load_z(float (&v)[O], const float* a, int off)
{
#pragma unroll
    for (int k = 0; k < O; ++k) v[k] = a[off + k];
}

template <int O>
__device__ __forceinline__ void load_x(float (&v)[O], const float* a, int off, int ld)
{
#pragma unroll
    for (int k = 0; k < O; ++k) v[k] = a[off + k * ld];
}

// Each point function first issues every load its updates need (no store precedes them,
// so they go out together), then evaluates and stores in the reference's order.

template <int O>
__device__ __forceinline__ void vel_ac_point(const VelParams2D& P, const PassExtras& X, int64_t it, int ninj, int i,
                                             int j)
{
    constexpr int H = O - 1;
    const int ld = P.ld;
    const int bk = (i >> 6) + P.gx * j;
    const int o = (i + H) + ld * (j + H);
    const bool ux = i < P.nz - 2 * H && j < P.nx - H;   // vx on (nz-2h) x (nx-h)
    const bool uz = i < P.nz - H && j < P.nx - 2 * H;   // vz on (nz-h) x (nx-2h)
    float px[O], pz[O], vxo = 0.0f, vzo = 0.0f, rx0 = 0.0f, rx1 = 0.0f, rz0 = 0.0f, rz1 = 0.0f;
    if (ux) {
        load_x<O>(px, P.s0, (i + H) + ld * j, ld);
        vxo = P.vx[o];
        rx0 = P.rho[(i + H) + ld * j];
        rx1 = P.rho[(i + H) + ld * (j + 1)];
    }
    if (uz) {
        load_z<O>(pz, P.s0, i + ld * (j + H));
        vzo = P.vz[o];
        rz0 = P.rho[i + ld * (j + H)];
        rz1 = P.rho[(i + 1) + ld * (j + H)];
    }
    const float cxo = dt_over_av2(P.dt, rx0, rx1);   // dtinvavxirho
    const float czo = dt_over_av2(P.dt, rz0, rz1);   // dtinvavzirho
    // vx[inn] += dtinvavxirho * dpdx, dpdx = D_xi(p) * dxI
    if (ux) {
        float d = stencil<O>(px) * P.dxI;
        d = cpml_x(d, P.qx0, i, j, P.npml, P.pml_mask);
        float nv = vxo + cxo * d;
        // rigid x faces: the only nonzero ghost mirrors the first/last interior column
        // (dirichlet.jl ighostmin/ighostmax with ifd = 1), taken before the source term
        if ((P.rigid_mask & GPI_FACE(2, 0)) && j == 0) P.vx[(i + H)] = -nv;
        if ((P.rigid_mask & GPI_FACE(2, 1)) && j == P.nx - H - 1) P.vx[(i + H) + ld * (P.nx + H - 1)] = -nv;
        if (ninj) nv = inject_velocity(X, it, INJ_VX, bk, o, nv, P.rho, P.dt);
        P.vx[o] = nv;
    }
    // vz[inn] += dtinvavzirho * dpdz, dpdz = D_zi(p) * dzI
    if (uz) {
        float d = stencil<O>(pz) * P.dzI;
        d = cpml_z(d, P.qz0, i, j, P.npml, P.pml_mask);
        float nv = vzo + czo * d;
        if ((P.rigid_mask & GPI_FACE(0, 0)) && i == 0) P.vz[ld * (j + H)] = -nv;
        if ((P.rigid_mask & GPI_FACE(0, 1)) && i == P.nz - H - 1) P.vz[(P.nz + H - 1) + ld * (j + H)] = -nv;
        if (ninj) nv = inject_velocity(X, it, INJ_VZ, bk, o, nv, P.rho, P.dt);
        P.vz[o] = nv;
    }
}

template <int O>
__device__ __forceinline__ void str_ac_point(const StrParams2D& P, const PassExtras& X, int64_t it, int ninj, int i,
                                             int j)
{
    if (i >= P.nz || j >= P.nx) return;
    const int ld = P.ld;
    const int o = i + ld * j;
    float vxs[O], vzs[O];
    load_x<O>(vxs, P.vx, o, ld);
    load_z<O>(vzs, P.vz, o);
    const float c = P.c0[o];
    const float po = P.s0[o];
    float dvx = stencil<O>(vxs) * P.dxI;             // D_xa(vx)
    dvx = cpml_x(dvx, P.qx0, i, j, P.npml, P.pml_mask);
    float dvz = stencil<O>(vzs) * P.dzI;             // D_za(vz)
    dvz = cpml_z(dvz, P.qz0, i, j, P.npml, P.pml_mask);
    float p = po + (dvx + dvz) * c;
    // p = p + pv * dtK per source field (muladd_tauii!, source.jl:160-163)
    const int bk = (i >> 6) + P.gx * j;
    for (int q = 0; q < ninj; ++q) {
        float acc;
        int t;
        if (X.inj[q].kind == INJ_P && injection_sum(X.inj[q], bk, o, it, acc, t)) p = p + (acc * c);
    }
    P.s0[o] = p;
}

// Elastic updates from their operands (the point functions below load them).
template <int O>
__device__ __forceinline__ float vel_el_vx(const VelParams2D& P, int i, int j, const float (&a1)[O],
                                           const float (&a2)[O], float vxo, float cxo)
{
    // vx[inn] -= dtinvavxirho*(dtauxxdx + dtauxzdz)
    float d1 = stencil<O>(a1) * P.dxI;               // D_xi(tauxx)
    d1 = cpml_x(d1, P.qx0, i, j, P.npml, P.pml_mask);
    float d2 = stencil<O>(a2) * P.dzI;               // D_za(tauxz)
    d2 = cpml_z(d2, P.qz1, i, j, P.npml, P.pml_mask);
    return vxo - cxo * (d1 + d2);
}

template <int O>
__device__ __forceinline__ float vel_el_vz(const VelParams2D& P, int i, int j, const float (&b1)[O],
                                           const float (&b2)[O], float vzo, float czo)
{
    // vz[inn] -= dtinvavzirho*(dtauxzdx + dtauzzdz)
    float d1 = stencil<O>(b1) * P.dxI;               // D_xa(tauxz)
    d1 = cpml_x(d1, P.qx1, i, j, P.npml, P.pml_mask);
    float d2 = stencil<O>(b2) * P.dzI;               // D_zi(tauzz)
    d2 = cpml_z(d2, P.qz0, i, j, P.npml, P.pml_mask);
    return vzo - czo * (d1 + d2);
}

template <int O>
__device__ __forceinline__ void str_el_normal(const StrParams2D& P, int i, int j, const float (&ax)[O],
                                              const float (&az)[O], float s0o, float s1o, float m, float l,
                                              float& txx, float& tzz)
{
    float dvx = stencil<O>(ax) * P.dxI;              // D_xa(vx)
    dvx = cpml_x(dvx, P.qx0, i, j, P.npml, P.pml_mask);
    float dvz = stencil<O>(az) * P.dzI;              // D_za(vz)
    dvz = cpml_z(dvz, P.qz0, i, j, P.npml, P.pml_mask);
    txx = (s0o - (m * dvx)) - (l * dvz);
    tzz = (s1o - (m * dvz)) - (l * dvx);
}

// tauxx += pv*dtM and tauzz += pv*dtM per source field (source.jl:79-97)
__device__ __forceinline__ void inject_normal2(const PassExtras& X, int64_t it, int ninj, int bk, int64_t o, float m,
                                               float& txx, float& tzz)
{
    for (int q = 0; q < ninj; ++q) {
        float acc;
        int t;
        if (X.inj[q].kind == INJ_NORMAL2 && injection_sum(X.inj[q], bk, o, it, acc, t)) {
            txx = txx + (acc * m);
            tzz = tzz + (acc * m);
        }
    }
}

template <int O>
__device__ __forceinline__ float str_el_shear(const StrParams2D& P, int i, int j, const float (&bz)[O],
                                              const float (&bx)[O], float s2o, float c2o)
{
    // tauxz -= dtavmu*(dvxdz + dvzdx)
    float a = stencil<O>(bz) * P.dzI;                // D_zi(vx)
    a = cpml_z(a, P.qz1, i, j, P.npml, P.pml_mask);
    float c = stencil<O>(bx) * P.dxI;                // D_xi(vz)
    c = cpml_x(c, P.qx1, i, j, P.npml, P.pml_mask);
    return s2o - c2o * (a + c);
}

template <int O>
__device__ __forceinline__ void vel_el_point(const VelParams2D& P, const PassExtras& X, int64_t it, int ninj, int i,
                                             int j)
{
    constexpr int H = O - 1;
    const int ld = P.ld;
    const int bk = (i >> 6) + P.gx * j;
    const int o = (i + H) + ld * (j + H);
    const bool ux = i < P.nz - 2 * H && j < P.nx - H;   // vx on (nz-2h) x (nx-h)
    const bool uz = i < P.nz - H && j < P.nx - 2 * H;   // vz on (nz-h) x (nx-2h)
    float a1[O], a2[O], b1[O], b2[O], vxo = 0.0f, vzo = 0.0f, rx0 = 0.0f, rx1 = 0.0f, rz0 = 0.0f, rz1 = 0.0f;
    if (ux) {
        load_x<O>(a1, P.s0, (i + H) + ld * j, ld);   // tauxx along x
        load_z<O>(a2, P.s2, i + ld * j);             // tauxz along z
        vxo = P.vx[o];
        rx0 = P.rho[(i + H) + ld * j];
        rx1 = P.rho[(i + H) + ld * (j + 1)];
    }
    if (uz) {
        load_x<O>(b1, P.s2, i + ld * j, ld);         // tauxz along x
        load_z<O>(b2, P.s1, i + ld * (j + H));       // tauzz along z
        vzo = P.vz[o];
        rz0 = P.rho[i + ld * (j + H)];
        rz1 = P.rho[(i + 1) + ld * (j + H)];
    }
    const float cxo = dt_over_av2(P.dt, rx0, rx1);   // dtinvavxirho
    const float czo = dt_over_av2(P.dt, rz0, rz1);   // dtinvavzirho
    if (ux) {
        float nv = vel_el_vx<O>(P, i, j, a1, a2, vxo, cxo);
        if ((P.rigid_mask & GPI_FACE(2, 0)) && j == 0) P.vx[(i + H)] = -nv;
        if ((P.rigid_mask & GPI_FACE(2, 1)) && j == P.nx - H - 1) P.vx[(i + H) + ld * (P.nx + H - 1)] = -nv;
        if (ninj) nv = inject_velocity(X, it, INJ_VX, bk, o, nv, P.rho, P.dt);
        P.vx[o] = nv;
    }
    if (uz) {
        float nv = vel_el_vz<O>(P, i, j, b1, b2, vzo, czo);
        if ((P.rigid_mask & GPI_FACE(0, 0)) && i == 0) P.vz[ld * (j + H)] = -nv;
        if ((P.rigid_mask & GPI_FACE(0, 1)) && i == P.nz - H - 1) P.vz[(P.nz + H - 1) + ld * (j + H)] = -nv;
        if (ninj) nv = inject_velocity(X, it, INJ_VZ, bk, o, nv, P.rho, P.dt);
        P.vz[o] = nv;
    }
}

template <int O>
__device__ __forceinline__ void str_el_point(const StrParams2D& P, const PassExtras& X, int64_t it, int ninj, int i,
                                             int j)
{
    constexpr int H = O - 1;
    const int ld = P.ld;
    const bool fs = (P.free_mask & GPI_FACE(0, 0)) != 0;
    const int o = i + ld * j;
    const bool un = i < P.nz && j < P.nx;             // tauxx, tauzz
    const bool us = i < P.nz - H && j < P.nx - H;     // tauxz on (nz-h) x (nx-h)
    float ax[O], az[O], bz[O], bx[O];
    float il = 1.0f, im = 1.0f, s0o = 0.0f, s1o = 0.0f, s2o = 0.0f;
    float m01 = 1.0f, m10 = 1.0f, m11 = 1.0f;
    if (un) {
        load_x<O>(ax, P.vx, o, ld);
        load_z<O>(az, P.vz, o);
        il = P.c0[o];
        im = P.c1[o];
        s0o = P.s0[o];
        s1o = P.s1[o];
    }
    if (us) {
        load_z<O>(bz, P.vx, i + ld * (j + H));
        load_x<O>(bx, P.vz, (i + H) + ld * j, ld);
        s2o = P.s2[o];
        m10 = P.c1[o + 1];
        m01 = P.c1[o + ld];
        m11 = P.c1[o + 1 + ld];
    }
    const float m = dtm_of(il, im, P.dt);                     // dtM
    const float l = (1.0f / il) * P.dt;                       // dtlambda
    const float c2o = dt_over_av4(P.dt, im, m10, m01, m11);   // dtavmu
    if (un) {
        float txx, tzz;
        str_el_normal<O>(P, i, j, ax, az, s0o, s1o, m, l, txx, tzz);
        // free surface at zmin, before the sources: tauzz[1,:] = -tauzz[2,:]
        if (fs && i == 1) P.s1[ld * j] = -tzz;
        if (ninj) inject_normal2(X, it, ninj, (i >> 6) + P.gx * j, o, m, txx, tzz);
        P.s0[o] = txx;
        if (!(fs && i == 0)) P.s1[o] = tzz;
    }
    if (us) {
        const float t = str_el_shear<O>(P, i, j, bz, bx, s2o, c2o);
        P.s2[o] = (fs && i == 0) ? 0.0f : t;         // tauxz[1,:] = 0
    }
}

// ---------------------------------------------------------------------------------------
// Tiles: a workgroup of 64 x TY threads covers 64 rows and TY*RX columns; each thread
// updates RX consecutive columns.  Workgroup b covers rows 64*(b % gx).. and columns
// TY*RX*(b / gx)..
// ---------------------------------------------------------------------------------------

template <int O, int PHYS, int TY, int RX>
__device__ __forceinline__ void vel_tile(const VelParams2D& P, const PassExtras& X, int64_t it, int ninj, int b)
{
    const int i = (b % P.gx) * 64 + (int)threadIdx.x;
    const int j0 = (b / P.gx) * (TY * RX) + (int)threadIdx.y * RX;
#pragma unroll
    for (int r = 0; r < RX; ++r) {
        if (PHYS == 0) vel_ac_point<O>(P, X, it, ninj, i, j0 + r);
        else vel_el_point<O>(P, X, it, ninj, i, j0 + r);
    }
}

template <int O, int PHYS, int TY, int RX>
__device__ __forceinline__ void str_tile(const StrParams2D& P, const PassExtras& X, int64_t it, int ninj, int b)
{
    const int i = (b % P.gx) * 64 + (int)threadIdx.x;
    const int j0 = (b / P.gx) * (TY * RX) + (int)threadIdx.y * RX;
#pragma unroll
    for (int r = 0; r < RX; ++r) {
        if (PHYS == 0) str_ac_point<O>(P, X, it, ninj, i, j0 + r);
        else str_el_point<O>(P, X, it, ninj, i, j0 + r);
    }
}

// ---------------------------------------------------------------------------------------
// One shot per launch (the drop-in path): 64 x 4 workgroups, one column per thread.
// ---------------------------------------------------------------------------------------

// Workgroup shape of the single-shot kernels: index into TILE_SHAPES, from GPI_TILE2D
// (default 0: 64 x 4 threads, one column each).
int plain_shape()
{
    static int t = -1;
    if (t < 0) {
        const char* e = getenv("GPI_TILE2D");
        t = e ? atoi(e) : 0;
        if (t < 0 || t > 3) t = 0;
    }
    return t;
}

template <int O, int PHYS, int TY, int RX>
__global__ __launch_bounds__(64 * TY) void vel2d_kernel(VelParams2D P, PassExtras X)
{
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int nbm = P.gx * P.gy;
    if (b >= nbm) {
        gather_block<4>(X, X.it, b - nbm);
        return;
    }
    vel_tile<O, PHYS, TY, RX>(P, X, X.it, X.ninj, b);
}

template <int O, int PHYS, int TY, int RX>
__global__ __launch_bounds__(64 * TY) void str2d_kernel(StrParams2D P, PassExtras X)
{
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int nbm = P.gx * P.gy;
    if (b >= nbm) {
        gather_block<4>(X, X.it, b - nbm);
        return;
    }
    str_tile<O, PHYS, TY, RX>(P, X, X.it, X.ninj, b);
}

template <int O, int PHYS, int TY, int RX>
void launch_plain(bool velocity, hipStream_t s, VelParams2D V, StrParams2D S, const PassExtras& X)
{
    constexpr int H = O - 1;
    if (velocity) {
        V.gx = (V.nz - H + 63) / 64;
        V.gy = (V.nx - H + TY * RX - 1) / (TY * RX);
        hipLaunchKernelGGL((vel2d_kernel<O, PHYS, TY, RX>), dim3(V.gx * V.gy + X.gather_blocks), dim3(64, TY), 0, s,
                           V, X);
    } else {
        S.gx = (S.nz + 63) / 64;
        S.gy = (S.nx + TY * RX - 1) / (TY * RX);
        hipLaunchKernelGGL((str2d_kernel<O, PHYS, TY, RX>), dim3(S.gx * S.gy + X.gather_blocks), dim3(64, TY), 0, s,
                           S, X);
    }
}

template <int O, int PHYS>
void launch_plain_o(bool velocity, hipStream_t s, const VelParams2D& V, const StrParams2D& S, const PassExtras& X)
{
    switch (plain_shape()) {
    case 1: launch_plain<O, PHYS, 8, 1>(velocity, s, V, S, X); break;
    case 2: launch_plain<O, PHYS, 2, 2>(velocity, s, V, S, X); break;
    case 3: launch_plain<O, PHYS, 4, 2>(velocity, s, V, S, X); break;
    default: launch_plain<O, PHYS, 4, 1>(velocity, s, V, S, X); break;
    }
}

template <int O>
void launch_vel(const Model& m, hipStream_t s, VelParams2D P, const PassExtras& X)
{
    if (m.g.physics == 0) launch_plain_o<O, 0>(true, s, P, StrParams2D{}, X);
    else launch_plain_o<O, 1>(true, s, P, StrParams2D{}, X);
}

template <int O>
void launch_str(const Model& m, hipStream_t s, StrParams2D P, const PassExtras& X)
{
    if (m.g.physics == 0) launch_plain_o<O, 0>(false, s, VelParams2D{}, P, X);
    else launch_plain_o<O, 1>(false, s, VelParams2D{}, P, X);
}

// ---------------------------------------------------------------------------------------
// Many shots per launch (batched serving).  The launch covers every resident shot's
// workgroups back to back; first[s] is shot s's first workgroup (first[n] the total).
// Each shot advances at its own time step it = step - start: shots past their last step
// idle, and at it == nt a shot takes the trailing velocity half step of
// src/fdtd/propagate.jl:254-255 (no sources, no receivers).  The batch's step is read from
// device memory (*step0, set by batch_set_step before a chunk) plus the launch's offset
// koff in the chunk, and so is the number of shots (*nshots); a launch may cover more
// workgroups than first[n] (the extra ones return at once).  One captured graph of a
// chunk's launches therefore replays at any step and for any batch whose workgroups fit.
// ---------------------------------------------------------------------------------------

__device__ __forceinline__ int shot_of_block(const int* __restrict__ first, int n, int b)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

// Workgroup b of a batched pass at batch step `step` (the shot's own step it = step - start).
template <int O, int PHYS, int TY, int RX>
__device__ __forceinline__ void vel_batch_block(const BatchShot2D* __restrict__ D, const int* __restrict__ first, int n,
                                                int64_t step, int b)
{
    if (b >= first[n]) return;
    const int s = shot_of_block(first, n, b);
    const BatchShot2D& S = D[s];
    const int64_t it = step - S.start;
    if (it < 0 || it > S.nt) return;
    const int lb = b - first[s];
    const int nbm = S.vp.gx * S.vp.gy;
    if (lb >= nbm) {
        if (it < S.nt) gather_block<4>(S.xv, it, lb - nbm);
        return;
    }
    vel_tile<O, PHYS, TY, RX>(S.vp, S.xv, it, it < S.nt ? S.xv.ninj : 0, lb);
}

template <int O, int PHYS, int TY, int RX>
__device__ __forceinline__ void str_batch_block(const BatchShot2D* __restrict__ D, const int* __restrict__ first, int n,
                                                int64_t step, int b)
{
    if (b >= first[n]) return;
    const int s = shot_of_block(first, n, b);
    const BatchShot2D& S = D[s];
    const int64_t it = step - S.start;
    if (it < 0 || it >= S.nt) return;
    const int lb = b - first[s];
    const int nbm = S.sp.gx * S.sp.gy;
    if (lb >= nbm) {
        gather_block<4>(S.xs, it, lb - nbm);
        return;
    }
    str_tile<O, PHYS, TY, RX>(S.sp, S.xs, it, S.xs.ninj, lb);
}

template <int O, int PHYS, int TY, int RX>
__global__ __launch_bounds__(64 * TY) void vel2d_batch(const BatchShot2D* __restrict__ D, const int* __restrict__ first,
                                                       const int* __restrict__ nshots,
                                                       const int64_t* __restrict__ step0, int koff)
{
    vel_batch_block<O, PHYS, TY, RX>(D, first, *nshots, *step0 + koff, xcd_remap(blockIdx.x, gridDim.x));
}

template <int O, int PHYS, int TY, int RX>
__global__ __launch_bounds__(64 * TY) void str2d_batch(const BatchShot2D* __restrict__ D, const int* __restrict__ first,
                                                       const int* __restrict__ nshots,
                                                       const int64_t* __restrict__ step0, int koff)
{
    str_batch_block<O, PHYS, TY, RX>(D, first, *nshots, *step0 + koff, xcd_remap(blockIdx.x, gridDim.x));
}

__global__ void batch_set_step_kernel(int64_t* step0, int64_t step)
{
    if (threadIdx.x == 0) *step0 = step;
}

template <int O, int PHYS, int TY, int RX>
void launch_batch_t(bool velocity, const BatchShot2D* D, const int* first, const int* n, int nblocks,
                    const int64_t* step0, int koff, hipStream_t s)
{
    dim3 grd(nblocks), blk(64, TY);
    if (velocity) hipLaunchKernelGGL((vel2d_batch<O, PHYS, TY, RX>), grd, blk, 0, s, D, first, n, step0, koff);
    else hipLaunchKernelGGL((str2d_batch<O, PHYS, TY, RX>), grd, blk, 0, s, D, first, n, step0, koff);
}

template <int O, int PHYS>
void launch_batch_o(const TileShape& t, bool velocity, const BatchShot2D* D, const int* first, const int* n,
                    int nblocks, const int64_t* step0, int koff, hipStream_t s)
{
    if (t.ty == 4 && t.rx == 1) launch_batch_t<O, PHYS, 4, 1>(velocity, D, first, n, nblocks, step0, koff, s);
    else if (t.ty == 8 && t.rx == 1) launch_batch_t<O, PHYS, 8, 1>(velocity, D, first, n, nblocks, step0, koff, s);
    else if (t.ty == 2 && t.rx == 2) launch_batch_t<O, PHYS, 2, 2>(velocity, D, first, n, nblocks, step0, koff, s);
    else launch_batch_t<O, PHYS, 4, 2>(velocity, D, first, n, nblocks, step0, koff, s);
}

}  // namespace

const TileShape TILE_SHAPES[NTILE_SHAPES] = {{4, 1}, {8, 1}, {2, 2}, {4, 2}};

int velocity_buckets_2d(const Model& m, int* gz)
{
    *gz = (int)((m.g.n[0] - m.h + 63) / 64);
    return *gz * (int)(m.g.n[2] - m.h);
}

int stress_buckets_2d(const Model& m, int* gz)
{
    *gz = (int)((m.g.n[0] + 63) / 64);
    return *gz * (int)m.g.n[2];
}

void velocity_grid_2d(const Model& m, const TileShape& t, int* gx, int* gy)
{
    *gx = (int)((m.g.n[0] - m.h + 63) / 64);
    *gy = (int)((m.g.n[2] - m.h + t.ty * t.rx - 1) / (t.ty * t.rx));
}

void stress_grid_2d(const Model& m, const TileShape& t, int* gx, int* gy)
{
    *gx = (int)((m.g.n[0] + 63) / 64);
    *gy = (int)((m.g.n[2] + t.ty * t.rx - 1) / (t.ty * t.rx));
}

void batch_pass_2d(int physics, int order, const TileShape& t, bool velocity, const BatchShot2D* D, const int* first,
                   const int* n, int nblocks, const int64_t* step0, int koff, hipStream_t s)
{
    if (nblocks <= 0) return;
    if (physics == 0) {
        switch (order) {
        case 2: launch_batch_o<2, 0>(t, velocity, D, first, n, nblocks, step0, koff, s); break;
        case 4: launch_batch_o<4, 0>(t, velocity, D, first, n, nblocks, step0, koff, s); break;
        case 6: launch_batch_o<6, 0>(t, velocity, D, first, n, nblocks, step0, koff, s); break;
        default: launch_batch_o<8, 0>(t, velocity, D, first, n, nblocks, step0, koff, s); break;
        }
    } else {
        switch (order) {
        case 2: launch_batch_o<2, 1>(t, velocity, D, first, n, nblocks, step0, koff, s); break;
        case 4: launch_batch_o<4, 1>(t, velocity, D, first, n, nblocks, step0, koff, s); break;
        case 6: launch_batch_o<6, 1>(t, velocity, D, first, n, nblocks, step0, koff, s); break;
        default: launch_batch_o<8, 1>(t, velocity, D, first, n, nblocks, step0, koff, s); break;
        }
    }
}

void batch_set_step(int64_t* step0, int64_t step, hipStream_t s)
{
    hipLaunchKernelGGL(batch_set_step_kernel, dim3(1), dim3(64), 0, s, step0, step);
}

bool vel_params_2d(const Model& m, const State& st, VelParams2D* out)
{
    VelParams2D P{};
    const bool ac = m.g.physics == 0;
    P.s0 = st.field[ac ? GPI_P : GPI_TAUXX];
    P.s1 = ac ? nullptr : st.field[GPI_TAUZZ];
    P.s2 = ac ? nullptr : st.field[GPI_TAUXZ];
    P.vx = st.field[GPI_VX];
    P.vz = st.field[GPI_VZ];
    P.rho = m.coef[GPI_C_RHO];
    P.qx0 = pml_of(m, st, ac ? GPI_DPDX : GPI_DTAUXXDX);
    P.qx1 = ac ? Pml{} : pml_of(m, st, GPI_DTAUXZDX);
    P.qz0 = pml_of(m, st, ac ? GPI_DPDZ : GPI_DTAUZZDZ);
    P.qz1 = ac ? Pml{} : pml_of(m, st, GPI_DTAUXZDZ);
    P.nz = (int)m.g.n[0];
    P.nx = (int)m.g.n[2];
    P.ld = (int)m.pitch0;
    P.npml = m.g.npml;
    P.pml_mask = m.g.pml_mask;
    P.rigid_mask = m.g.rigid_mask;
    P.dxI = m.g.dsI[2];
    P.dzI = m.g.dsI[0];
    P.dt = m.g.dt;
    const Pml* q[4] = {&P.qx0, &P.qx1, &P.qz0, &P.qz1};
    if (!state_complete(m, q, 4) || !P.s0 || !P.vx || !P.vz || !P.rho || (!ac && (!P.s1 || !P.s2)))
        return false;
    *out = P;
    return true;
}

bool str_params_2d(const Model& m, const State& st, StrParams2D* out)
{
    StrParams2D P{};
    const bool ac = m.g.physics == 0;
    P.vx = st.field[GPI_VX];
    P.vz = st.field[GPI_VZ];
    P.s0 = st.field[ac ? GPI_P : GPI_TAUXX];
    P.s1 = ac ? nullptr : st.field[GPI_TAUZZ];
    P.s2 = ac ? nullptr : st.field[GPI_TAUXZ];
    P.c0 = m.coef[ac ? GPI_C_DTK : GPI_C_INVLAMBDA];
    P.c1 = ac ? nullptr : m.coef[GPI_C_INVMU];
    P.dt = m.g.dt;
    P.qx0 = pml_of(m, st, GPI_DVXDX);
    P.qx1 = ac ? Pml{} : pml_of(m, st, GPI_DVZDX);
    P.qz0 = pml_of(m, st, GPI_DVZDZ);
    P.qz1 = ac ? Pml{} : pml_of(m, st, GPI_DVXDZ);
    P.nz = (int)m.g.n[0];
    P.nx = (int)m.g.n[2];
    P.ld = (int)m.pitch0;
    P.npml = m.g.npml;
    P.pml_mask = m.g.pml_mask;
    P.free_mask = m.g.free_mask;
    P.dxI = m.g.dsI[2];
    P.dzI = m.g.dsI[0];
    const Pml* q[4] = {&P.qx0, &P.qx1, &P.qz0, &P.qz1};
    if (!state_complete(m, q, 4) || !P.s0 || !P.vx || !P.vz || !P.c0 || (!ac && (!P.s1 || !P.s2 || !P.c1)))
        return false;
    *out = P;
    return true;
}

int velocity_pass_2d(const Model& m, const State& st, hipStream_t s, const PassExtras& X)
{
    VelParams2D P;
    if (!vel_params_2d(m, st, &P)) return -1;
    switch (m.g.order) {
    case 2: launch_vel<2>(m, s, P, X); break;
    case 4: launch_vel<4>(m, s, P, X); break;
    case 6: launch_vel<6>(m, s, P, X); break;
    default: launch_vel<8>(m, s, P, X); break;
    }
    return 0;
}

int stress_pass_2d(const Model& m, const State& st, hipStream_t s, const PassExtras& X)
{
    StrParams2D P;
    if (!str_params_2d(m, st, &P)) return -1;
    switch (m.g.order) {
    case 2: launch_str<2>(m, s, P, X); break;
    case 4: launch_str<4>(m, s, P, X); break;
    case 6: launch_str<6>(m, s, P, X); break;
    default: launch_str<8>(m, s, P, X); break;
    }
    return 0;
}

}  // namespace gpi
