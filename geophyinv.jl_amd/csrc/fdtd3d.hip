// 3-D staggered-grid velocity-stress kernels for gfx950.
//
// Same structure as the 2-D passes: the velocity pass fuses the stress derivatives, their
// CPML strips, compute_v!, the rigid faces and the velocity sources
// (src/fdtd/advance_acou.jl:98-194, src/fdtd/advance_elastic.jl:233-394); the stress pass
// fuses the velocity derivatives, CPML, compute_p!/compute_stressii!/compute_stressij!, the
// free surface and the stress sources (src/fdtd/advance_acou.jl:196-250,
// src/fdtd/advance_elastic.jl:395-531).  Workgroups are 64 (z) x 4 (y) threads on one x
// plane; workgroup b covers z-rows 64*(b % gx).., y-rows 4*((b / gx) % gy).., plane
// b / (gx*gy).  Extra workgroups after the stencil range sample receivers.
#include <algorithm>
#include <cstdlib>

#include "fdtd3d.h"
#include "pass_extras.h"
#include "stencil.h"

namespace gpi {

namespace {

__device__ __forceinline__ float cpml_update(float d, const Pml& q, int idx, int c)
{
    float* m = q.psi + idx;
    const float mv = q.b[c] * *m + q.a[c] * d;
    *m = mv;
    return d * q.kI[c] + mv;
}

// CPML of a derivative along x at derivative-grid point (i, j, k); q.nd = its x length
__device__ __forceinline__ float cpml3_x(float d, const Pml& q, int i, int j, int k, int npml, uint32_t mask)
{
    if ((mask & GPI_FACE(2, 0)) && k < npml) d = cpml_update(d, q, i + (int)q.ps1 * j + (int)q.ps2 * k, k);
    if ((mask & GPI_FACE(2, 1)) && k >= q.nd - npml) {
        const int c = k - (q.nd - npml) + npml;
        d = cpml_update(d, q, i + (int)q.ps1 * j + (int)q.ps2 * c, c);
    }
    return d;
}

__device__ __forceinline__ float cpml3_y(float d, const Pml& q, int i, int j, int k, int npml, uint32_t mask)
{
    if ((mask & GPI_FACE(1, 0)) && j < npml) d = cpml_update(d, q, i + (int)q.ps1 * j + (int)q.ps2 * k, j);
    if ((mask & GPI_FACE(1, 1)) && j >= q.nd - npml) {
        const int r = j - (q.nd - npml) + npml;
        d = cpml_update(d, q, i + (int)q.ps1 * r + (int)q.ps2 * k, r);
    }
    return d;
}

__device__ __forceinline__ float cpml3_z(float d, const Pml& q, int i, int j, int k, int npml, uint32_t mask)
{
    if ((mask & GPI_FACE(0, 0)) && i < npml) d = cpml_update(d, q, i + (int)q.ps1 * j + (int)q.ps2 * k, i);
    if ((mask & GPI_FACE(0, 1)) && i >= q.nd - npml) {
        const int r = i - (q.nd - npml) + npml;
        d = cpml_update(d, q, r + (int)q.ps1 * j + (int)q.ps2 * k, r);
    }
    return d;
}

template <int O>
__device__ __forceinline__ float dstride(const float* a, int off, int stride)
{
    float v[O];
#pragma unroll
    for (int m = 0; m < O; ++m) v[m] = a[off + m * stride];
    return stencil<O>(v);
}

// Store of a pass output: streaming (non-temporal) when NT, for the 3-D acoustic passes
// whose outputs are not read again before they leave the caches.
template <bool NT>
__device__ __forceinline__ void store_out(float* p, float v)
{
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

template <int O>
__device__ __forceinline__ void load_s(float (&v)[O], const float* a, int off, int stride)
{
#pragma unroll
    for (int m = 0; m < O; ++m) v[m] = a[off + m * stride];
}

struct Idx3 {
    int i, j, k;
};

// Workgroup b -> point (i, j, k) of a 64 x TY tile of plane k (workgroup chunk times
// blockDim.z plus threadIdx.z); returns the per-plane tile id bx + gx*(by + gy*k) (the
// injection bucket).  With xfast the chunks of one tile are consecutive workgroups (nk
// chunks), so that with the XCD-aware order the x-neighbour planes a point's x-stencil
// reads are handled on the same XCD.
__device__ __forceinline__ int decode(int b, int gx, int gy, int nk, bool xfast, Idx3& p)
{
    int bx, by, kc;
    if (xfast) {
        kc = b % nk;
        const int t = b / nk;
        bx = t % gx;
        by = t / gx;
    } else {
        bx = b % gx;
        const int r = b / gx;
        by = r % gy;
        kc = r / gy;
    }
    p.k = kc * (int)blockDim.z + (int)threadIdx.z;
    p.i = bx * 64 + (int)threadIdx.x;
    p.j = by * (int)blockDim.y + (int)threadIdx.y;
    return bx + gx * (by + gy * p.k);
}

// ---------------------------------------------------------------------------------------
// acoustic
// ---------------------------------------------------------------------------------------

// Acoustic passes march along x: workgroup (bx, by, chunk) updates planes
// k0 = chunk*kc .. k0+kc-1 of its (z, y) tile in order and keeps the x-stencil samples of
// its own column in registers, so each value on the column is loaded once per chunk
// instead of once per plane that uses it.  The injection bucket of a point is the
// per-plane tile id bx + gx*(by + gy*k), as planned on the host.

template <int O, bool NT>
__global__ __launch_bounds__(512) void vel3d_acoustic(VelParams3D P, PassExtras X)
{
    constexpr int H = O - 1;
    const int b = P.remap ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int nxw = P.nx - H;
    const int nch = (nxw + P.kc - 1) / P.kc;
    const int nbm = P.gx * P.gy * nch;
    if (b >= nbm) {
        gather_block(X, b - nbm);
        return;
    }
    const int bx = b % P.gx, r = b / P.gx, by = r % P.gy, ch = r / P.gy;
    const int i = bx * 64 + (int)threadIdx.x - H;   // rows shifted by h
    const int j = by * (int)blockDim.y + (int)threadIdx.y;
    const int k0 = ch * P.kc, k1 = min(k0 + P.kc, nxw);
    const int s1 = (int)P.s1, s2 = (int)P.s2;
    const float* p = P.s[0];
    const bool row = i >= 0;
    const bool colx = row && i < P.nz - 2 * H && j < P.ny - 2 * H;   // this thread updates vx
    const int cx0 = (i + H) + s1 * (j + H);                     // its p column
    float w[O];
#pragma unroll
    for (int m = 0; m < O - 1; ++m) w[m + 1] = colx ? p[cx0 + s2 * (k0 + m)] : 0.0f;
    // operands of one plane; the next plane's are loaded before this plane's stores so
    // that two planes of loads are in flight per wave
    const bool ycol = row && i < P.nz - 2 * H && j < P.ny - H;   // vy column (plane k < nx-2h)
    const bool zcol = row && i < P.nz - H && j < P.ny - 2 * H;   // vz column (plane k < nx-2h)
    // (the precomputed density coefficients are read here: forming them from rho in this
    // pass measured 8 % slower on 440^3, unlike the 2-D and elastic passes)
    struct Ops {
        float wn, vxo, cxo, vyo, cyo, vzo, czo;
        float py[O], pz[O];
    };
    auto fetch = [&](int k, Ops& L) {
        const int o = (i + H) + s1 * (j + H) + s2 * (k + H);
        const int c = i + s1 * j + s2 * k;
        if (colx) {
            L.wn = p[cx0 + s2 * (k + O - 1)];
            L.vxo = P.vx[o];
            L.cxo = P.cx[c];
        }
        if (ycol && k < P.nx - 2 * H) {
            load_s<O>(L.py, p, (i + H) + s1 * j + s2 * (k + H), s1);
            L.vyo = P.vy[o];
            L.cyo = P.cy[c];
        }
        if (zcol && k < P.nx - 2 * H) {
            load_s<O>(L.pz, p, i + s1 * (j + H) + s2 * (k + H), 1);
            L.vzo = P.vz[o];
            L.czo = P.cz[c];
        }
    };
    Ops cur{}, nxt{};
    if (k0 < k1) fetch(k0, cur);
    for (int k = k0; k < k1; ++k) {
        if (k + 1 < k1) fetch(k + 1, nxt);
        const int bt = bx + P.gx * (by + P.gy * k);
        const int o = (i + H) + s1 * (j + H) + s2 * (k + H);   // @inn of every velocity
        // vx on (nz-2h, ny-2h, nx-h): dpdx = D_xi(p)
        if (colx) {
#pragma unroll
            for (int m = 0; m < O - 1; ++m) w[m] = w[m + 1];
            w[O - 1] = cur.wn;
            float d = stencil<O>(w) * P.dxI;
            d = cpml3_x(d, P.qx[0], i, j, k, P.npml, P.pml_mask);
            float nv = cur.vxo + cur.cxo * d;
            if ((P.rigid_mask & GPI_FACE(2, 0)) && k == 0) P.vx[(i + H) + s1 * (j + H)] = -nv;
            if ((P.rigid_mask & GPI_FACE(2, 1)) && k == P.nx - H - 1)
                P.vx[(i + H) + s1 * (j + H) + s2 * (P.nx + H - 1)] = -nv;
            if (X.ninj) nv = inject_velocity(X, INJ_VX, bt, o, nv, P.rho, P.dt);
            store_out<NT>(&P.vx[o], nv);
        }
        // vy on (nz-2h, ny-h, nx-2h): dpdy = D_yi(p)
        if (ycol && k < P.nx - 2 * H) {
            float d = stencil<O>(cur.py) * P.dyI;
            d = cpml3_y(d, P.qy[0], i, j, k, P.npml, P.pml_mask);
            float nv = cur.vyo + cur.cyo * d;
            if ((P.rigid_mask & GPI_FACE(1, 0)) && j == 0) P.vy[(i + H) + s2 * (k + H)] = -nv;
            if ((P.rigid_mask & GPI_FACE(1, 1)) && j == P.ny - H - 1)
                P.vy[(i + H) + s1 * (P.ny + H - 1) + s2 * (k + H)] = -nv;
            if (X.ninj) nv = inject_velocity(X, INJ_VY, bt, o, nv, P.rho, P.dt);
            store_out<NT>(&P.vy[o], nv);
        }
        // vz on (nz-h, ny-2h, nx-2h): dpdz = D_zi(p)
        if (zcol && k < P.nx - 2 * H) {
            float d = stencil<O>(cur.pz) * P.dzI;
            d = cpml3_z(d, P.qz[0], i, j, k, P.npml, P.pml_mask);
            float nv = cur.vzo + cur.czo * d;
            if ((P.rigid_mask & GPI_FACE(0, 0)) && i == 0) P.vz[s1 * (j + H) + s2 * (k + H)] = -nv;
            if ((P.rigid_mask & GPI_FACE(0, 1)) && i == P.nz - H - 1)
                P.vz[(P.nz + H - 1) + s1 * (j + H) + s2 * (k + H)] = -nv;
            if (X.ninj) nv = inject_velocity(X, INJ_VZ, bt, o, nv, P.rho, P.dt);
            store_out<NT>(&P.vz[o], nv);
        }
        cur = nxt;
    }
}

template <int O, bool NT>
__global__ __launch_bounds__(512) void str3d_acoustic(StrParams3D P, PassExtras X)
{
    const int b = P.remap ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int nch = (P.nx + P.kc - 1) / P.kc;
    const int nbm = P.gx * P.gy * nch;
    if (b >= nbm) {
        gather_block(X, b - nbm);
        return;
    }
    const int bx = b % P.gx, r = b / P.gx, by = r % P.gy, ch = r / P.gy;
    const int i = bx * 64 + (int)threadIdx.x, j = by * (int)blockDim.y + (int)threadIdx.y;
    if (i >= P.nz || j >= P.ny) return;
    const int k0 = ch * P.kc, k1 = min(k0 + P.kc, P.nx);
    const int s1 = (int)P.s1, s2 = (int)P.s2;
    const int col = i + s1 * j;
    float w[O];
#pragma unroll
    for (int m = 0; m < O - 1; ++m) w[m + 1] = P.vx[col + s2 * (k0 + m)];
    // operands of one plane; the next plane's are loaded before this plane's store
    struct Ops {
        float wn, c, p;
        float vy[O], vz[O];
    };
    auto fetch = [&](int k, Ops& L) {
        const int o = col + s2 * k;
        L.wn = P.vx[o + s2 * (O - 1)];
        load_s<O>(L.vy, P.vy, o, s1);
        load_s<O>(L.vz, P.vz, o, 1);
        L.c = P.c[0][o];
        L.p = P.s[0][o];
    };
    Ops cur{}, nxt{};
    if (k0 < k1) fetch(k0, cur);
    for (int k = k0; k < k1; ++k) {
        if (k + 1 < k1) fetch(k + 1, nxt);
        const int bt = bx + P.gx * (by + P.gy * k);
        const int o = col + s2 * k;
#pragma unroll
        for (int m = 0; m < O - 1; ++m) w[m] = w[m + 1];
        w[O - 1] = cur.wn;
        float dvx = stencil<O>(w) * P.dxI;                // D_xa(vx)
        dvx = cpml3_x(dvx, P.qx[0], i, j, k, P.npml, P.pml_mask);
        float dvy = stencil<O>(cur.vy) * P.dyI;          // D_ya(vy)
        dvy = cpml3_y(dvy, P.qy[0], i, j, k, P.npml, P.pml_mask);
        float dvz = stencil<O>(cur.vz) * P.dzI;          // D_za(vz)
        dvz = cpml3_z(dvz, P.qz[0], i, j, k, P.npml, P.pml_mask);
        const float c = cur.c;
        float pn = cur.p + ((dvx + dvz) + dvy) * c;       // compute_p! 3-D, advance_acou.jl:310-313
        for (int q = 0; q < X.ninj; ++q) {
            float acc;
            int tt;
            if (X.inj[q].kind == INJ_P && injection_sum(X.inj[q], bt, o, X.it, acc, tt)) pn = pn + (acc * c);
        }
        store_out<NT>(&P.s[0][o], pn);
        cur = nxt;
    }
}

// ---------------------------------------------------------------------------------------
// elastic
// ---------------------------------------------------------------------------------------

// One (z, y) point of plane k of the elastic velocity pass; bt is the injection bucket.
template <int O>
__device__ __forceinline__ void vel3d_elastic_point(const VelParams3D& P, const PassExtras& X, int i, int j, int k,
                                                    int bt)
{
    constexpr int H = O - 1;
    const int s1 = (int)P.s1, s2 = (int)P.s2;
    const float *txx = P.s[0], *tyy = P.s[1], *tzz = P.s[2], *txy = P.s[3], *txz = P.s[4], *tyz = P.s[5];
    const int o = (i + H) + s1 * (j + H) + s2 * (k + H);
    const int c = i + s1 * j + s2 * k;
    const bool row = i >= 0;
    const bool ux = row && i < P.nz - 2 * H && j < P.ny - 2 * H && k < P.nx - H;
    const bool uy = row && i < P.nz - 2 * H && j < P.ny - H && k < P.nx - 2 * H;
    const bool uz = row && i < P.nz - H && j < P.ny - 2 * H && k < P.nx - 2 * H;
    // loads of the three updates first, then evaluation and stores
    float a1[O], a2[O], a3[O], b1[O], b2[O], b3[O], e1[O], e2[O], e3[O];
    float vxo = 0.0f, vyo = 0.0f, vzo = 0.0f;
    float rx0 = 0.0f, rx1 = 0.0f, ry0 = 0.0f, ry1 = 0.0f, rz0 = 0.0f, rz1 = 0.0f;
    const float* rho = P.rho;
    if (ux) {
        const int a = (i + H) + s1 * (j + H) + s2 * k;
        load_s<O>(a1, txx, a, s2);
        load_s<O>(a2, txy, c, s1);
        load_s<O>(a3, txz, c, 1);
        vxo = P.vx[o];
        rx0 = rho[a];
        rx1 = rho[a + s2];
    }
    if (uy) {
        const int a = (i + H) + s1 * j + s2 * (k + H);
        load_s<O>(b1, txy, c, s2);
        load_s<O>(b2, tyy, a, s1);
        load_s<O>(b3, tyz, c, 1);
        vyo = P.vy[o];
        ry0 = rho[a];
        ry1 = rho[a + s1];
    }
    if (uz) {
        const int a = i + s1 * (j + H) + s2 * (k + H);
        load_s<O>(e1, txz, c, s2);
        load_s<O>(e2, tyz, c, s1);
        load_s<O>(e3, tzz, a, 1);
        vzo = P.vz[o];
        rz0 = rho[a];
        rz1 = rho[a + 1];
    }
    const float cxo = dt_over_av2(P.dt, rx0, rx1);   // dtinvavxirho
    const float cyo = dt_over_av2(P.dt, ry0, ry1);   // dtinvavyirho
    const float czo = dt_over_av2(P.dt, rz0, rz1);   // dtinvavzirho
    // vx -= dtinvavxirho*((dtauxxdx + dtauxydy) + dtauxzdz) on (nz-2h, ny-2h, nx-h)
    if (ux) {
        float d1 = stencil<O>(a1) * P.dxI;   // D_xi(tauxx)
        d1 = cpml3_x(d1, P.qx[0], i, j, k, P.npml, P.pml_mask);
        float d2 = stencil<O>(a2) * P.dyI;   // D_ya(tauxy)
        d2 = cpml3_y(d2, P.qy[0], i, j, k, P.npml, P.pml_mask);
        float d3 = stencil<O>(a3) * P.dzI;   // D_za(tauxz)
        d3 = cpml3_z(d3, P.qz[0], i, j, k, P.npml, P.pml_mask);
        float nv = vxo - cxo * ((d1 + d2) + d3);
        if ((P.rigid_mask & GPI_FACE(2, 0)) && k == 0) P.vx[(i + H) + s1 * (j + H)] = -nv;
        if ((P.rigid_mask & GPI_FACE(2, 1)) && k == P.nx - H - 1)
            P.vx[(i + H) + s1 * (j + H) + s2 * (P.nx + H - 1)] = -nv;
        if (X.ninj) nv = inject_velocity(X, INJ_VX, bt, o, nv, P.rho, P.dt);
        P.vx[o] = nv;
    }
    // vy -= dtinvavyirho*((dtauxydx + dtauyydy) + dtauyzdz) on (nz-2h, ny-h, nx-2h)
    if (uy) {
        float d1 = stencil<O>(b1) * P.dxI;   // D_xa(tauxy)
        d1 = cpml3_x(d1, P.qx[1], i, j, k, P.npml, P.pml_mask);
        float d2 = stencil<O>(b2) * P.dyI;   // D_yi(tauyy)
        d2 = cpml3_y(d2, P.qy[1], i, j, k, P.npml, P.pml_mask);
        float d3 = stencil<O>(b3) * P.dzI;   // D_za(tauyz)
        d3 = cpml3_z(d3, P.qz[1], i, j, k, P.npml, P.pml_mask);
        float nv = vyo - cyo * ((d1 + d2) + d3);
        if ((P.rigid_mask & GPI_FACE(1, 0)) && j == 0) P.vy[(i + H) + s2 * (k + H)] = -nv;
        if ((P.rigid_mask & GPI_FACE(1, 1)) && j == P.ny - H - 1)
            P.vy[(i + H) + s1 * (P.ny + H - 1) + s2 * (k + H)] = -nv;
        if (X.ninj) nv = inject_velocity(X, INJ_VY, bt, o, nv, P.rho, P.dt);
        P.vy[o] = nv;
    }
    // vz -= dtinvavzirho*((dtauxzdx + dtauyzdy) + dtauzzdz) on (nz-h, ny-2h, nx-2h)
    if (uz) {
        float d1 = stencil<O>(e1) * P.dxI;   // D_xa(tauxz)
        d1 = cpml3_x(d1, P.qx[2], i, j, k, P.npml, P.pml_mask);
        float d2 = stencil<O>(e2) * P.dyI;   // D_ya(tauyz)
        d2 = cpml3_y(d2, P.qy[2], i, j, k, P.npml, P.pml_mask);
        float d3 = stencil<O>(e3) * P.dzI;   // D_zi(tauzz)
        d3 = cpml3_z(d3, P.qz[2], i, j, k, P.npml, P.pml_mask);
        float nv = vzo - czo * ((d1 + d2) + d3);
        if ((P.rigid_mask & GPI_FACE(0, 0)) && i == 0) P.vz[s1 * (j + H) + s2 * (k + H)] = -nv;
        if ((P.rigid_mask & GPI_FACE(0, 1)) && i == P.nz - H - 1)
            P.vz[(P.nz + H - 1) + s1 * (j + H) + s2 * (k + H)] = -nv;
        if (X.ninj) nv = inject_velocity(X, INJ_VZ, bt, o, nv, P.rho, P.dt);
        P.vz[o] = nv;
    }
}

// Elastic workgroups are 64 (z) x TY (y) x KZ (x) threads: the layer threadIdx.z of
// workgroup chunk c updates plane c*KZ + threadIdx.z, so the planes whose x-stencils
// share samples are updated by one workgroup at the same time and the shared samples are
// served by its CU's caches rather than fetched again by another workgroup later (as a
// loop over planes would do, but without the registers the compiler keeps live across
// such a loop).  With KZ = 1 the xfast order is available as before.
template <int O>
__global__ __launch_bounds__(512) void vel3d_elastic(VelParams3D P, PassExtras X)
{
    constexpr int H = O - 1;
    const int b = P.remap ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int nxw = P.nx - H;
    const int kz = (int)blockDim.z;
    const int nch = (nxw + kz - 1) / kz;
    if (b >= P.gx * P.gy * nch) {
        if (threadIdx.z == 0) gather_block(X, b - P.gx * P.gy * nch);
        return;
    }
    Idx3 t;
    const int bt = decode(b, P.gx, P.gy, nch, P.xfast != 0, t);
    if (t.k >= nxw) return;
    vel3d_elastic_point<O>(P, X, t.i - H, t.j, t.k, bt);   // rows shifted by h
}

// One (z, y) point of plane k of the elastic stress pass; bt is the injection bucket.
template <int O>
__device__ __forceinline__ void str3d_elastic_point(const StrParams3D& P, const PassExtras& X, int i, int j, int k,
                                                    int bt)
{
    constexpr int H = O - 1;
    const int s1 = (int)P.s1, s2 = (int)P.s2;
    const bool fs = (P.free_mask & GPI_FACE(0, 0)) != 0;
    const int o = i + s1 * j + s2 * k;
    const bool un = i < P.nz && j < P.ny;
    const bool uxz = i < P.nz - H && j < P.ny - 2 * H && k < P.nx - H;
    const bool uxy = i < P.nz - 2 * H && j < P.ny - H && k < P.nx - H;
    const bool uyz = i < P.nz - H && j < P.ny - H && k < P.nx - 2 * H;
    // loads of the four updates first, then evaluation and stores
    float nx_[O], ny_[O], nz_[O], xz1[O], xz2[O], xy1[O], xy2[O], yz1[O], yz2[O];
    float il = 1.0f, im = 1.0f, s0o = 0.0f, s1o = 0.0f, s2o = 0.0f;
    float s4o = 0.0f, s3o = 0.0f, s5o = 0.0f;
    float q3[4] = {1.0f, 1.0f, 1.0f, 1.0f}, q2[4] = {1.0f, 1.0f, 1.0f, 1.0f}, q4[4] = {1.0f, 1.0f, 1.0f, 1.0f};
    const float* imu = P.c[1];
    if (un) {
        load_s<O>(nx_, P.vx, o, s2);
        load_s<O>(ny_, P.vy, o, s1);
        load_s<O>(nz_, P.vz, o, 1);
        il = P.c[0][o];
        im = imu[o];
        s0o = P.s[0][o];
        s1o = P.s[1][o];
        s2o = P.s[2][o];
    }
    if (uxz) {   // @av_xzi(invmu) on the dvxdz grid: (i, j+h, k) (i+1, ..) (.., k+1) (i+1, .., k+1)
        load_s<O>(xz1, P.vx, i + s1 * (j + H) + s2 * (k + H), 1);
        load_s<O>(xz2, P.vz, (i + H) + s1 * (j + H) + s2 * k, s2);
        s4o = P.s[4][o];
        const int a = i + s1 * (j + H) + s2 * k;
        q3[0] = imu[a];
        q3[1] = imu[a + 1];
        q3[2] = imu[a + s2];
        q3[3] = imu[a + 1 + s2];
    }
    if (uxy) {   // @av_xyi(invmu) on the dvxdy grid: (i+h, j, k) (.., j+1, ..) (.., k+1) (.., j+1, k+1)
        load_s<O>(xy1, P.vx, (i + H) + s1 * j + s2 * (k + H), s1);
        load_s<O>(xy2, P.vy, (i + H) + s1 * (j + H) + s2 * k, s2);
        s3o = P.s[3][o];
        const int a = (i + H) + s1 * j + s2 * k;
        q2[0] = imu[a];
        q2[1] = imu[a + s1];
        q2[2] = imu[a + s2];
        q2[3] = imu[a + s1 + s2];
    }
    if (uyz) {   // @av_yzi(invmu) on the dvydz grid: (i, j, k+h) (i+1, ..) (.., j+1, ..) (i+1, j+1, ..)
        load_s<O>(yz1, P.vy, i + s1 * (j + H) + s2 * (k + H), 1);
        load_s<O>(yz2, P.vz, (i + H) + s1 * j + s2 * (k + H), s1);
        s5o = P.s[5][o];
        const int a = i + s1 * j + s2 * (k + H);
        q4[0] = imu[a];
        q4[1] = imu[a + 1];
        q4[2] = imu[a + s1];
        q4[3] = imu[a + 1 + s1];
    }
    const float m = dtm_of(il, im, P.dt);                                  // dtM
    const float l = (1.0f / il) * P.dt;                                    // dtlambda
    const float c3o = dt_over_av4(P.dt, q3[0], q3[1], q3[2], q3[3]);      // dtavxzimu
    const float c2o = dt_over_av4(P.dt, q2[0], q2[1], q2[2], q2[3]);      // dtavxyimu
    const float c4o = dt_over_av4(P.dt, q4[0], q4[1], q4[2], q4[3]);      // dtavyzimu
    // normal stresses on the stress grid (compute_stressii!, advance_elastic.jl:155-177)
    if (un) {
        float dvx = stencil<O>(nx_) * P.dxI;    // D_xa(vx)
        dvx = cpml3_x(dvx, P.qx[0], i, j, k, P.npml, P.pml_mask);
        float dvy = stencil<O>(ny_) * P.dyI;    // D_ya(vy)
        dvy = cpml3_y(dvy, P.qy[0], i, j, k, P.npml, P.pml_mask);
        float dvz = stencil<O>(nz_) * P.dzI;    // D_za(vz)
        dvz = cpml3_z(dvz, P.qz[0], i, j, k, P.npml, P.pml_mask);
        float txx = (s0o - (m * dvx)) - (l * (dvy + dvz));
        float tyy = (s1o - (m * dvy)) - (l * (dvx + dvz));
        float tzz = (s2o - (m * dvz)) - (l * (dvy + dvx));
        if (fs && i == 1) P.s[2][s1 * j + s2 * k] = -tzz;   // tauzz[1,:,:] = -tauzz[2,:,:]
        for (int q = 0; q < X.ninj; ++q) {
            float acc;
            int tt;
            if (X.inj[q].kind == INJ_NORMAL3 && injection_sum(X.inj[q], bt, o, X.it, acc, tt)) {
                txx = txx + (acc * m);
                tyy = tyy + (acc * m);
                tzz = tzz + (acc * m);
            }
        }
        P.s[0][o] = txx;
        P.s[1][o] = tyy;
        if (!(fs && i == 0)) P.s[2][o] = tzz;
    }
    // tauxz -= dtavxzimu*(dvxdz + dvzdx) on (nz-h, ny-2h, nx-h)
    if (uxz) {
        float a = stencil<O>(xz1) * P.dzI;      // D_zi(vx)
        a = cpml3_z(a, P.qz[1], i, j, k, P.npml, P.pml_mask);
        float c = stencil<O>(xz2) * P.dxI;      // D_xi(vz)
        c = cpml3_x(c, P.qx[2], i, j, k, P.npml, P.pml_mask);
        const float v = s4o - c3o * (a + c);
        P.s[4][o] = (fs && i == 0) ? 0.0f : v;
    }
    // tauxy -= dtavxyimu*(dvxdy + dvydx) on (nz-2h, ny-h, nx-h)
    if (uxy) {
        float a = stencil<O>(xy1) * P.dyI;      // D_yi(vx)
        a = cpml3_y(a, P.qy[1], i, j, k, P.npml, P.pml_mask);
        float c = stencil<O>(xy2) * P.dxI;      // D_xi(vy)
        c = cpml3_x(c, P.qx[1], i, j, k, P.npml, P.pml_mask);
        P.s[3][o] = s3o - c2o * (a + c);
    }
    // tauyz -= dtavyzimu*(dvydz + dvzdy) on (nz-h, ny-h, nx-2h)
    if (uyz) {
        float a = stencil<O>(yz1) * P.dzI;      // D_zi(vy)
        a = cpml3_z(a, P.qz[2], i, j, k, P.npml, P.pml_mask);
        float c = stencil<O>(yz2) * P.dyI;      // D_yi(vz)
        c = cpml3_y(c, P.qy[2], i, j, k, P.npml, P.pml_mask);
        const float v = s5o - c4o * (a + c);
        P.s[5][o] = (fs && i == 0) ? 0.0f : v;
    }
}

template <int O>
__global__ __launch_bounds__(512) void str3d_elastic(StrParams3D P, PassExtras X)
{
    const int b = P.remap ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int kz = (int)blockDim.z;
    const int nch = (P.nx + kz - 1) / kz;
    if (b >= P.gx * P.gy * nch) {
        if (threadIdx.z == 0) gather_block(X, b - P.gx * P.gy * nch);
        return;
    }
    Idx3 t;
    const int bt = decode(b, P.gx, P.gy, nch, P.xfast != 0, t);
    if (t.k >= P.nx) return;
    str3d_elastic_point<O>(P, X, t.i, t.j, t.k, bt);
}

// XCD-aware workgroup order of the 3-D passes: GPI_XCD_REMAP3 (default 0: measured on
// 440^3 acoustic and 360^3 elastic, the plain order is 3-5 % faster in 3-D)
int remap3()
{
    static int r = -1;
    if (r < 0) {
        const char* e = getenv("GPI_XCD_REMAP3");
        r = e ? atoi(e) != 0 : 0;
    }
    return r;
}

// x-planes per workgroup of the acoustic passes: GPI_3D_MARCH planes (default 4, measured
// on 440^3 against 2, 8 and 16; 1 gives one plane per workgroup), at most the axis length
int march_planes(int n)
{
    static int kc = -1;
    if (kc < 0) {
        const char* e = getenv("GPI_3D_MARCH");
        kc = e ? atoi(e) : 4;
        if (kc < 1) kc = 1;
    }
    return std::min(kc, std::max(n, 1));
}

// 3-D acoustic outputs stored non-temporally when GPI_3D_NT=1.
bool nt_stores()
{
    static int e = -1;
    if (e < 0) {
        const char* v = getenv("GPI_3D_NT");
        e = v ? atoi(v) != 0 : 0;
    }
    return e == 1;
}

// Elastic workgroup order: GPI_3D_EORDER=1 walks each tile's x planes consecutively.
// Measured on 360^3 (order 2): plain order 24.9 Gpt/s, x-fastest 23.3, x-fastest with
// the XCD-aware order 21.2; marching along x with the x-stencils in registers (as the
// acoustic passes do) 16-18, its 166-256 VGPRs leaving 1-3 waves per SIMD against 7.
// x-planes per elastic workgroup (its z dimension): GPI_3D_EKZ = 1, 2 or 4 (default 1),
// reduced so that a workgroup has at most 512 threads
int elastic_planes(int ty)
{
    static int kz = -1;
    if (kz < 0) {
        const char* e = getenv("GPI_3D_EKZ");
        kz = e ? atoi(e) : 1;
        if (kz != 2 && kz != 4) kz = 1;
    }
    int k = kz;
    while (k > 1 && 64 * ty * k > 512) k /= 2;
    return k;
}

bool elastic_xfast()
{
    static int e = -1;
    if (e < 0) {
        const char* v = getenv("GPI_3D_EORDER");
        e = v ? atoi(v) != 0 : 0;
    }
    return e == 1;
}

template <int O>
void launch_vel3(const Model& m, hipStream_t s, VelParams3D P, const PassExtras& X)
{
    constexpr int H = O - 1;
    P.gx = (P.nz - H + vel3d_row_shift(m) + 63) / 64;
    const int ty = tile_rows_3d(m);
    P.gy = (P.ny - H + ty - 1) / ty;
    const int kz = m.g.physics == 0 ? 1 : elastic_planes(ty);
    P.kc = m.g.physics == 0 ? march_planes(P.nx - H) : kz;
    P.remap = remap3();
    P.xfast = m.g.physics == 1 && elastic_xfast();
    dim3 grd(P.gx * P.gy * ((P.nx - H + P.kc - 1) / P.kc) + X.gather_blocks), blk(64, ty, kz);
    if (m.g.physics == 0 && nt_stores()) hipLaunchKernelGGL((vel3d_acoustic<O, true>), grd, blk, 0, s, P, X);
    else if (m.g.physics == 0) hipLaunchKernelGGL((vel3d_acoustic<O, false>), grd, blk, 0, s, P, X);
    else hipLaunchKernelGGL(vel3d_elastic<O>, grd, blk, 0, s, P, X);
}

template <int O>
void launch_str3(const Model& m, hipStream_t s, StrParams3D P, const PassExtras& X)
{
    P.gx = (P.nz + 63) / 64;
    const int ty = tile_rows_3d(m);
    P.gy = (P.ny + ty - 1) / ty;
    const int kz = m.g.physics == 0 ? 1 : elastic_planes(ty);
    P.kc = m.g.physics == 0 ? march_planes(P.nx) : kz;
    P.remap = remap3();
    P.xfast = m.g.physics == 1 && elastic_xfast();
    dim3 grd(P.gx * P.gy * ((P.nx + P.kc - 1) / P.kc) + X.gather_blocks), blk(64, ty, kz);
    if (m.g.physics == 0 && nt_stores()) hipLaunchKernelGGL((str3d_acoustic<O, true>), grd, blk, 0, s, P, X);
    else if (m.g.physics == 0) hipLaunchKernelGGL((str3d_acoustic<O, false>), grd, blk, 0, s, P, X);
    else hipLaunchKernelGGL(str3d_elastic<O>, grd, blk, 0, s, P, X);
}

}  // namespace

int vel3d_row_shift(const Model& m) { return m.h; }

int tile_rows_3d(const Model& m)
{
    static int ty = -1;
    if (ty < 0) {
        const char* e = getenv("GPI_3D_TY");
        ty = e && atoi(e) == 8 ? 8 : 4;
    }
    static int tye = -1;
    if (tye < 0) {
        const char* e = getenv("GPI_3D_TYE");
        tye = e && (atoi(e) == 8 || atoi(e) == 2) ? atoi(e) : 4;
    }
    return m.g.physics == 0 ? ty : tye;
}

int velocity_blocks_3d(const Model& m, int* gx, int* gy)
{
    const int H = m.h;
    *gx = (int)((m.g.n[0] - H + vel3d_row_shift(m) + 63) / 64);
    *gy = (int)((m.g.n[1] - H + tile_rows_3d(m) - 1) / tile_rows_3d(m));
    return *gx * *gy * (int)(m.g.n[2] - H);
}

int stress_blocks_3d(const Model& m, int* gx, int* gy)
{
    *gx = (int)((m.g.n[0] + 63) / 64);
    *gy = (int)((m.g.n[1] + tile_rows_3d(m) - 1) / tile_rows_3d(m));
    return *gx * *gy * (int)m.g.n[2];
}

int velocity_pass_3d(const Model& m, const State& st, hipStream_t s, const PassExtras& X)
{
    VelParams3D P{};
    const bool ac = m.g.physics == 0;
    if (ac) {
        P.s[0] = st.field[GPI_P];
        P.qx[0] = pml_of(m, st, GPI_DPDX);
        P.qy[0] = pml_of(m, st, GPI_DPDY);
        P.qz[0] = pml_of(m, st, GPI_DPDZ);
    } else {
        const int sf[6] = {GPI_TAUXX, GPI_TAUYY, GPI_TAUZZ, GPI_TAUXY, GPI_TAUXZ, GPI_TAUYZ};
        for (int q = 0; q < 6; ++q) P.s[q] = st.field[sf[q]];
        P.qx[0] = pml_of(m, st, GPI_DTAUXXDX);
        P.qy[0] = pml_of(m, st, GPI_DTAUXYDY);
        P.qz[0] = pml_of(m, st, GPI_DTAUXZDZ);
        P.qx[1] = pml_of(m, st, GPI_DTAUXYDX);
        P.qy[1] = pml_of(m, st, GPI_DTAUYYDY);
        P.qz[1] = pml_of(m, st, GPI_DTAUYZDZ);
        P.qx[2] = pml_of(m, st, GPI_DTAUXZDX);
        P.qy[2] = pml_of(m, st, GPI_DTAUYZDY);
        P.qz[2] = pml_of(m, st, GPI_DTAUZZDZ);
    }
    P.vx = st.field[GPI_VX];
    P.vy = st.field[GPI_VY];
    P.vz = st.field[GPI_VZ];
    P.cx = m.coef[GPI_C_DTINVAVXIRHO];   // acoustic pass
    P.cy = m.coef[GPI_C_DTINVAVYIRHO];
    P.cz = m.coef[GPI_C_DTINVAVZIRHO];
    P.rho = m.coef[GPI_C_RHO];
    P.nz = (int)m.g.n[0];
    P.ny = (int)m.g.n[1];
    P.nx = (int)m.g.n[2];
    P.s1 = m.pitch0;
    P.s2 = m.pitch0 * m.pitch1;
    P.npml = m.g.npml;
    P.pml_mask = m.g.pml_mask;
    P.rigid_mask = m.g.rigid_mask;
    P.dzI = m.g.dsI[0];
    P.dyI = m.g.dsI[1];
    P.dxI = m.g.dsI[2];
    P.dt = m.g.dt;
    const Pml* q[9] = {&P.qx[0], &P.qx[1], &P.qx[2], &P.qy[0], &P.qy[1], &P.qy[2], &P.qz[0], &P.qz[1], &P.qz[2]};
    if (!state_complete(m, q, 9) || !P.vx || !P.vy || !P.vz || !P.s[0] || !P.rho || (ac && (!P.cx || !P.cy || !P.cz)))
        return -1;
    switch (m.g.order) {
    case 2: launch_vel3<2>(m, s, P, X); break;
    case 4: launch_vel3<4>(m, s, P, X); break;
    case 6: launch_vel3<6>(m, s, P, X); break;
    default: launch_vel3<8>(m, s, P, X); break;
    }
    return 0;
}

int stress_pass_3d(const Model& m, const State& st, hipStream_t s, const PassExtras& X)
{
    StrParams3D P{};
    const bool ac = m.g.physics == 0;
    P.vx = st.field[GPI_VX];
    P.vy = st.field[GPI_VY];
    P.vz = st.field[GPI_VZ];
    if (ac) {
        P.s[0] = st.field[GPI_P];
        P.c[0] = m.coef[GPI_C_DTK];
        P.qx[0] = pml_of(m, st, GPI_DVXDX);
        P.qy[0] = pml_of(m, st, GPI_DVYDY);
        P.qz[0] = pml_of(m, st, GPI_DVZDZ);
    } else {
        const int sf[6] = {GPI_TAUXX, GPI_TAUYY, GPI_TAUZZ, GPI_TAUXY, GPI_TAUXZ, GPI_TAUYZ};
        for (int q = 0; q < 6; ++q) P.s[q] = st.field[sf[q]];
        P.c[0] = m.coef[GPI_C_INVLAMBDA];
        P.c[1] = m.coef[GPI_C_INVMU];
        P.qx[0] = pml_of(m, st, GPI_DVXDX);
        P.qx[1] = pml_of(m, st, GPI_DVYDX);
        P.qx[2] = pml_of(m, st, GPI_DVZDX);
        P.qy[0] = pml_of(m, st, GPI_DVYDY);
        P.qy[1] = pml_of(m, st, GPI_DVXDY);
        P.qy[2] = pml_of(m, st, GPI_DVZDY);
        P.qz[0] = pml_of(m, st, GPI_DVZDZ);
        P.qz[1] = pml_of(m, st, GPI_DVXDZ);
        P.qz[2] = pml_of(m, st, GPI_DVYDZ);
    }
    P.nz = (int)m.g.n[0];
    P.ny = (int)m.g.n[1];
    P.nx = (int)m.g.n[2];
    P.s1 = m.pitch0;
    P.s2 = m.pitch0 * m.pitch1;
    P.npml = m.g.npml;
    P.pml_mask = m.g.pml_mask;
    P.free_mask = m.g.free_mask;
    P.dzI = m.g.dsI[0];
    P.dyI = m.g.dsI[1];
    P.dxI = m.g.dsI[2];
    P.dt = m.g.dt;
    const Pml* q[9] = {&P.qx[0], &P.qx[1], &P.qx[2], &P.qy[0], &P.qy[1], &P.qy[2], &P.qz[0], &P.qz[1], &P.qz[2]};
    if (!state_complete(m, q, 9) || !P.vx || !P.vy || !P.vz || !P.s[0] || !P.c[0] || (!ac && !P.c[1])) return -1;
    switch (m.g.order) {
    case 2: launch_str3<2>(m, s, P, X); break;
    case 4: launch_str3<4>(m, s, P, X); break;
    case 6: launch_str3<6>(m, s, P, X); break;
    default: launch_str3<8>(m, s, P, X); break;
    }
    return 0;
}

}  // namespace gpi
