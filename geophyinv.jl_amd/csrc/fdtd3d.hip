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

// EARLY: every load before the store (the compiler cannot tell that psi aliases none of
// the coefficient arrays, so a coefficient read after the store waits for it): 0.917 /
// 0.755 ms against 0.925 / 0.812 for the staged elastic passes on 360^3, stress pass of
// 440^3 acoustic 0.458 against 0.478; the acoustic velocity pass measured 0.753 against
// 0.735 that way and keeps kI after the store.
template <bool EARLY = true>
__device__ __forceinline__ float cpml_update(float d, const Pml& q, int idx, int c)
{
    float* m = q.psi + idx;
    if constexpr (EARLY) {
        const float b = q.b[c], a = q.a[c], ki = q.kI[c], mo = *m;
        const float mv = b * mo + a * d;
        *m = mv;
        return d * ki + mv;
    } else {
        const float mv = q.b[c] * *m + q.a[c] * d;
        *m = mv;
        return d * q.kI[c] + mv;
    }
}

// CPML of a derivative along x at derivative-grid point (i, j, k); q.nd = its x length
template <bool EARLY = true>
__device__ __forceinline__ float cpml3_x(float d, const Pml& q, int i, int j, int k, int npml, uint32_t mask)
{
    if ((mask & GPI_FACE(2, 0)) && k < npml) d = cpml_update<EARLY>(d, q, i + (int)q.ps1 * j + (int)q.ps2 * k, k);
    if ((mask & GPI_FACE(2, 1)) && k >= q.nd - npml) {
        const int c = k - (q.nd - npml) + npml;
        d = cpml_update<EARLY>(d, q, i + (int)q.ps1 * j + (int)q.ps2 * c, c);
    }
    return d;
}

template <bool EARLY = true>
__device__ __forceinline__ float cpml3_y(float d, const Pml& q, int i, int j, int k, int npml, uint32_t mask)
{
    if ((mask & GPI_FACE(1, 0)) && j < npml) d = cpml_update<EARLY>(d, q, i + (int)q.ps1 * j + (int)q.ps2 * k, j);
    if ((mask & GPI_FACE(1, 1)) && j >= q.nd - npml) {
        const int r = j - (q.nd - npml) + npml;
        d = cpml_update<EARLY>(d, q, i + (int)q.ps1 * r + (int)q.ps2 * k, r);
    }
    return d;
}

template <bool EARLY = true>
__device__ __forceinline__ float cpml3_z(float d, const Pml& q, int i, int j, int k, int npml, uint32_t mask)
{
    if ((mask & GPI_FACE(0, 0)) && i < npml) d = cpml_update<EARLY>(d, q, i + (int)q.ps1 * j + (int)q.ps2 * k, i);
    if ((mask & GPI_FACE(0, 1)) && i >= q.nd - npml) {
        const int r = i - (q.nd - npml) + npml;
        d = cpml_update<EARLY>(d, q, r + (int)q.ps1 * j + (int)q.ps2 * k, r);
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

// Loads through a buffer resource whose base is a wave-uniform plane pointer: the lane
// offsets (bytes) are shared by all arrays of one tile geometry, so a loop over planes
// keeps one offset register per geometry instead of a 64-bit address per array.  An
// offset at or past `bytes` (the rest of the array from the base) reads 0.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const float* a, int64_t s2, int p, int nplanes)
{
    const int64_t rest = 4 * s2 * (int64_t)(nplanes - p);
    const int bytes = (int)(rest < 0x7fffffff ? rest : 0x7fffffff);
    return __builtin_amdgcn_make_buffer_rsrc((void*)(a + s2 * p), (short)0, bytes, 0x00020000);
}

__device__ __forceinline__ float bload(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}

// v, opaque to the optimiser: the addresses an update forms from its point (CPML strips,
// rigid faces, injection) are then formed per plane of a marching loop instead of being
// hoisted out of it, where each would hold a 64-bit register pair for the whole loop.
__device__ __forceinline__ int opaque(int v)
{
    asm volatile("" : "+v"(v));
    return v;
}

// Workgroup order of the marching passes: mode 0 the launch order, 1 each XCD one
// contiguous range of the grid (xcd_remap), M > 1 windows of 8*M consecutive workgroups in
// which each XCD takes a contiguous run of M, so that a tile's z and y neighbours mostly
// share its XCD's L2 while all XCDs stay in one region of the arrays.
__device__ __forceinline__ int tile_order(int w, int n, int mode)
{
    if (mode == 0) return w;
    if (mode == 1) return xcd_remap(w, n);
    const int win = 8 * mode, base = w / win * win;
    return base + xcd_remap(w - base, min(win, n - base));
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
    const int b = tile_order((int)blockIdx.x, (int)gridDim.x, P.remap);
    const int nxw = P.nx - H;
    const int nch = (nxw + P.kc - 1) / P.kc;
    const int nbm = P.gx * P.gy * nch;
    if (b >= nbm) {
        gather_block<8>(X, b - nbm);
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
    const int np = P.nx + H;                                    // planes of every array
    // lane offsets (bytes from a plane's start) for buffer loads with wave-uniform plane
    // bases: the p column and the velocities at (i+h, j+h), the y-stencil of p from (i+h,
    // j), its z-stencil from (i, j+h), the density coefficients at (i, j) (negative only
    // for rows i < 0, which update nothing: such an offset reads 0)
    const uint32_t lo = 4u * (uint32_t)cx0;
    const uint32_t ly = lo - 4u * (uint32_t)(s1 * H);
    const uint32_t lz = lo - 4u * (uint32_t)H;
    const uint32_t lc = lz - 4u * (uint32_t)(s1 * H);
    float w[O];
#pragma unroll
    for (int m = 0; m < O - 1; ++m) w[m + 1] = bload(plane_rsrc(p, s2, k0 + m, np), lo);
    const bool ycol = row && i < P.nz - 2 * H && j < P.ny - H;   // vy column (plane k < nx-2h)
    const bool zcol = row && i < P.nz - H && j < P.ny - 2 * H;   // vz column (plane k < nx-2h)
    // (the precomputed density coefficients are read here: forming them from rho in this
    // pass measured 8 % slower on 440^3, unlike the 2-D and elastic passes)
    struct Ops {
        float wn, vxo, cxo, vyo, cyo, vzo, czo;
        float py[O], pz[O];
    };
    // Operands of plane k, loaded by every thread whether or not its updates run: a load
    // whose condition could skip it would make every wait count the loads after it as
    // possibly skipped, and such reads stay inside the arrays or read 0.
    auto fetch = [&](int k, Ops& L) {
        const __amdgpu_buffer_rsrc_t pk = plane_rsrc(p, s2, k + H, np);
        L.wn = bload(pk, lo);
        L.vxo = bload(plane_rsrc(P.vx, s2, k + H, np), lo);
        L.cxo = bload(plane_rsrc(P.cx, s2, k, np), lc);
#pragma unroll
        for (int m = 0; m < O; ++m) L.py[m] = bload(pk, ly + 4u * (uint32_t)(m * s1));
        L.vyo = bload(plane_rsrc(P.vy, s2, k + H, np), lo);
        L.cyo = bload(plane_rsrc(P.cy, s2, k, np), lc);
#pragma unroll
        for (int m = 0; m < O; ++m) L.pz[m] = bload(pk, lz + 4u * (uint32_t)m);
        L.vzo = bload(plane_rsrc(P.vz, s2, k + H, np), lo);
        L.czo = bload(plane_rsrc(P.cz, s2, k, np), lc);
    };
    auto update = [&](int k, const Ops& cur) {
        const int bt = bx + P.gx * (by + P.gy * k);
        const int iu = opaque(i), ju = opaque(j);
        const int o = (iu + H) + s1 * (ju + H) + s2 * (k + H);   // @inn of every velocity
        // vx on (nz-2h, ny-2h, nx-h): dpdx = D_xi(p)
        if (colx) {
#pragma unroll
            for (int m = 0; m < O - 1; ++m) w[m] = w[m + 1];
            w[O - 1] = cur.wn;
            float d = stencil<O>(w) * P.dxI;
            d = cpml3_x<false>(d, P.qx[0], iu, ju, k, P.npml, P.pml_mask);
            float nv = cur.vxo + cur.cxo * d;
            if ((P.rigid_mask & GPI_FACE(2, 0)) && k == 0) P.vx[(iu + H) + s1 * (ju + H)] = -nv;
            if ((P.rigid_mask & GPI_FACE(2, 1)) && k == P.nx - H - 1)
                P.vx[(iu + H) + s1 * (ju + H) + s2 * (P.nx + H - 1)] = -nv;
            if (X.ninj) nv = inject_velocity(X, INJ_VX, bt, o, nv, P.rho, P.dt);
            store_out<NT>(&P.vx[o], nv);
        }
        // vy on (nz-2h, ny-h, nx-2h): dpdy = D_yi(p)
        if (ycol && k < P.nx - 2 * H) {
            float d = stencil<O>(cur.py) * P.dyI;
            d = cpml3_y<false>(d, P.qy[0], iu, ju, k, P.npml, P.pml_mask);
            float nv = cur.vyo + cur.cyo * d;
            if ((P.rigid_mask & GPI_FACE(1, 0)) && ju == 0) P.vy[(iu + H) + s2 * (k + H)] = -nv;
            if ((P.rigid_mask & GPI_FACE(1, 1)) && ju == P.ny - H - 1)
                P.vy[(iu + H) + s1 * (P.ny + H - 1) + s2 * (k + H)] = -nv;
            if (X.ninj) nv = inject_velocity(X, INJ_VY, bt, o, nv, P.rho, P.dt);
            store_out<NT>(&P.vy[o], nv);
        }
        // vz on (nz-h, ny-2h, nx-2h): dpdz = D_zi(p)
        if (zcol && k < P.nx - 2 * H) {
            float d = stencil<O>(cur.pz) * P.dzI;
            d = cpml3_z<false>(d, P.qz[0], iu, ju, k, P.npml, P.pml_mask);
            float nv = cur.vzo + cur.czo * d;
            if ((P.rigid_mask & GPI_FACE(0, 0)) && iu == 0) P.vz[s1 * (ju + H) + s2 * (k + H)] = -nv;
            if ((P.rigid_mask & GPI_FACE(0, 1)) && iu == P.nz - H - 1)
                P.vz[(P.nz + H - 1) + s1 * (ju + H) + s2 * (k + H)] = -nv;
            if (X.ninj) nv = inject_velocity(X, INJ_VZ, bt, o, nv, P.rho, P.dt);
            store_out<NT>(&P.vz[o], nv);
        }
    };
    // two planes per trip with alternating operand sets: no register whose load may be
    // pending is copied, and on every path the next plane's loads follow the current one's
    // (vector memory loads complete in order, so an update waits only for its own plane)
    Ops A{}, B{};
    int k = k0;
    if (k < k1) fetch(k, A);
    for (; k + 2 < k1; k += 2) {
        fetch(k + 1, B);
        update(k, A);
        fetch(k + 2, A);
        update(k + 1, B);
    }
    if (k + 1 < k1) {
        fetch(k + 1, B);
        update(k, A);
        update(k + 1, B);
    } else if (k < k1) {
        update(k, A);
    }
}

template <int O, bool NT>
__global__ __launch_bounds__(512) void str3d_acoustic(StrParams3D P, PassExtras X)
{
    const int b = tile_order((int)blockIdx.x, (int)gridDim.x, P.remap);
    const int nch = (P.nx + P.kc - 1) / P.kc;
    const int nbm = P.gx * P.gy * nch;
    if (b >= nbm) {
        gather_block<8>(X, b - nbm);
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

// The elastic velocity update of point (i, j, k) from its derivative samples (each a
// stencil times its inverse spacing, before CPML) and the density samples of its three
// averages; shared by the per-plane and the staged passes.  o is the point's offset in
// the velocity arrays, bt its injection bucket.
struct VelElSamples {
    float vxo, vyo, vzo;                 // velocities before the update
    float rx0, rx1, ry0, ry1, rz0, rz1;  // rho pairs of dtinvav{x,y,z}irho
    float a1, a2, a3;                    // D_xi(tauxx), D_ya(tauxy), D_za(tauxz)
    float b1, b2, b3;                    // D_xa(tauxy), D_yi(tauyy), D_za(tauyz)
    float e1, e2, e3;                    // D_xa(tauxz), D_ya(tauyz), D_zi(tauzz)
};

template <int O>
__device__ __forceinline__ void vel3d_elastic_update(const VelParams3D& P, const PassExtras& X, int i, int j, int k,
                                                     int bt, int o, bool ux, bool uy, bool uz, const VelElSamples& v)
{
    constexpr int H = O - 1;
    const int s1 = (int)P.s1, s2 = (int)P.s2;
    const float cxo = dt_over_av2(P.dt, v.rx0, v.rx1);   // dtinvavxirho
    const float cyo = dt_over_av2(P.dt, v.ry0, v.ry1);   // dtinvavyirho
    const float czo = dt_over_av2(P.dt, v.rz0, v.rz1);   // dtinvavzirho
    // vx -= dtinvavxirho*((dtauxxdx + dtauxydy) + dtauxzdz) on (nz-2h, ny-2h, nx-h)
    if (ux) {
        const float d1 = cpml3_x(v.a1, P.qx[0], i, j, k, P.npml, P.pml_mask);
        const float d2 = cpml3_y(v.a2, P.qy[0], i, j, k, P.npml, P.pml_mask);
        const float d3 = cpml3_z(v.a3, P.qz[0], i, j, k, P.npml, P.pml_mask);
        float nv = v.vxo - cxo * ((d1 + d2) + d3);
        if ((P.rigid_mask & GPI_FACE(2, 0)) && k == 0) P.vx[(i + H) + s1 * (j + H)] = -nv;
        if ((P.rigid_mask & GPI_FACE(2, 1)) && k == P.nx - H - 1)
            P.vx[(i + H) + s1 * (j + H) + s2 * (P.nx + H - 1)] = -nv;
        if (X.ninj) nv = inject_velocity(X, INJ_VX, bt, o, nv, P.rho, P.dt);
        P.vx[o] = nv;
    }
    // vy -= dtinvavyirho*((dtauxydx + dtauyydy) + dtauyzdz) on (nz-2h, ny-h, nx-2h)
    if (uy) {
        const float d1 = cpml3_x(v.b1, P.qx[1], i, j, k, P.npml, P.pml_mask);
        const float d2 = cpml3_y(v.b2, P.qy[1], i, j, k, P.npml, P.pml_mask);
        const float d3 = cpml3_z(v.b3, P.qz[1], i, j, k, P.npml, P.pml_mask);
        float nv = v.vyo - cyo * ((d1 + d2) + d3);
        if ((P.rigid_mask & GPI_FACE(1, 0)) && j == 0) P.vy[(i + H) + s2 * (k + H)] = -nv;
        if ((P.rigid_mask & GPI_FACE(1, 1)) && j == P.ny - H - 1)
            P.vy[(i + H) + s1 * (P.ny + H - 1) + s2 * (k + H)] = -nv;
        if (X.ninj) nv = inject_velocity(X, INJ_VY, bt, o, nv, P.rho, P.dt);
        P.vy[o] = nv;
    }
    // vz -= dtinvavzirho*((dtauxzdx + dtauyzdy) + dtauzzdz) on (nz-h, ny-2h, nx-2h)
    if (uz) {
        const float d1 = cpml3_x(v.e1, P.qx[2], i, j, k, P.npml, P.pml_mask);
        const float d2 = cpml3_y(v.e2, P.qy[2], i, j, k, P.npml, P.pml_mask);
        const float d3 = cpml3_z(v.e3, P.qz[2], i, j, k, P.npml, P.pml_mask);
        float nv = v.vzo - czo * ((d1 + d2) + d3);
        if ((P.rigid_mask & GPI_FACE(0, 0)) && i == 0) P.vz[s1 * (j + H) + s2 * (k + H)] = -nv;
        if ((P.rigid_mask & GPI_FACE(0, 1)) && i == P.nz - H - 1)
            P.vz[(P.nz + H - 1) + s1 * (j + H) + s2 * (k + H)] = -nv;
        if (X.ninj) nv = inject_velocity(X, INJ_VZ, bt, o, nv, P.rho, P.dt);
        P.vz[o] = nv;
    }
}

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
    VelElSamples v{};
    const float* rho = P.rho;
    if (ux) {
        const int a = (i + H) + s1 * (j + H) + s2 * k;
        load_s<O>(a1, txx, a, s2);
        load_s<O>(a2, txy, c, s1);
        load_s<O>(a3, txz, c, 1);
        v.vxo = P.vx[o];
        v.rx0 = rho[a];
        v.rx1 = rho[a + s2];
    }
    if (uy) {
        const int a = (i + H) + s1 * j + s2 * (k + H);
        load_s<O>(b1, txy, c, s2);
        load_s<O>(b2, tyy, a, s1);
        load_s<O>(b3, tyz, c, 1);
        v.vyo = P.vy[o];
        v.ry0 = rho[a];
        v.ry1 = rho[a + s1];
    }
    if (uz) {
        const int a = i + s1 * (j + H) + s2 * (k + H);
        load_s<O>(e1, txz, c, s2);
        load_s<O>(e2, tyz, c, s1);
        load_s<O>(e3, tzz, a, 1);
        v.vzo = P.vz[o];
        v.rz0 = rho[a];
        v.rz1 = rho[a + 1];
    }
    if (ux) {
        v.a1 = stencil<O>(a1) * P.dxI;
        v.a2 = stencil<O>(a2) * P.dyI;
        v.a3 = stencil<O>(a3) * P.dzI;
    }
    if (uy) {
        v.b1 = stencil<O>(b1) * P.dxI;
        v.b2 = stencil<O>(b2) * P.dyI;
        v.b3 = stencil<O>(b3) * P.dzI;
    }
    if (uz) {
        v.e1 = stencil<O>(e1) * P.dxI;
        v.e2 = stencil<O>(e2) * P.dyI;
        v.e3 = stencil<O>(e3) * P.dzI;
    }
    vel3d_elastic_update<O>(P, X, i, j, k, bt, o, ux, uy, uz, v);
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
        if (threadIdx.z == 0) gather_block<8>(X, b - P.gx * P.gy * nch);
        return;
    }
    Idx3 t;
    const int bt = decode(b, P.gx, P.gy, nch, P.xfast != 0, t);
    if (t.k >= nxw) return;
    vel3d_elastic_point<O>(P, X, t.i - H, t.j, t.k, bt);   // rows shifted by h
}

// The elastic stress update of point (i, j, k) from its derivative samples (stencil times
// inverse spacing, before CPML), the stresses before the update and the invlambda /
// invmu samples of its coefficients; shared by the per-plane and the staged passes.
struct StrElSamples {
    float il, im;                        // invlambda, invmu at the point
    float s0, s1, s2, s3, s4, s5;        // tauxx, tauyy, tauzz, tauxy, tauxz, tauyz before
    float q3[4], q2[4], q4[4];           // invmu of @av_xzi, @av_xyi, @av_yzi
    float dvx, dvy, dvz;                 // D_xa(vx), D_ya(vy), D_za(vz)
    float xza, xzc;                      // D_zi(vx), D_xi(vz)
    float xya, xyc;                      // D_yi(vx), D_xi(vy)
    float yza, yzc;                      // D_zi(vy), D_yi(vz)
};

// ZP: the caller has applied the derivatives' z-CPML (the staged pass, from its own loads).
template <int O, bool ZP = false>
__device__ __forceinline__ void str3d_elastic_update(const StrParams3D& P, const PassExtras& X, int i, int j, int k,
                                                     int bt, int o, bool un, bool uxz, bool uxy, bool uyz,
                                                     const StrElSamples& v)
{
    const int s1 = (int)P.s1, s2 = (int)P.s2;
    const bool fs = (P.free_mask & GPI_FACE(0, 0)) != 0;
    const float m = dtm_of(v.il, v.im, P.dt);                                       // dtM
    const float l = (1.0f / v.il) * P.dt;                                           // dtlambda
    const float c3o = dt_over_av4(P.dt, v.q3[0], v.q3[1], v.q3[2], v.q3[3]);        // dtavxzimu
    const float c2o = dt_over_av4(P.dt, v.q2[0], v.q2[1], v.q2[2], v.q2[3]);        // dtavxyimu
    const float c4o = dt_over_av4(P.dt, v.q4[0], v.q4[1], v.q4[2], v.q4[3]);        // dtavyzimu
    // normal stresses on the stress grid (compute_stressii!, advance_elastic.jl:155-177)
    if (un) {
        const float dvx = cpml3_x(v.dvx, P.qx[0], i, j, k, P.npml, P.pml_mask);
        const float dvy = cpml3_y(v.dvy, P.qy[0], i, j, k, P.npml, P.pml_mask);
        const float dvz = ZP ? v.dvz : cpml3_z(v.dvz, P.qz[0], i, j, k, P.npml, P.pml_mask);
        float txx = (v.s0 - (m * dvx)) - (l * (dvy + dvz));
        float tyy = (v.s1 - (m * dvy)) - (l * (dvx + dvz));
        float tzz = (v.s2 - (m * dvz)) - (l * (dvy + dvx));
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
        const float a = ZP ? v.xza : cpml3_z(v.xza, P.qz[1], i, j, k, P.npml, P.pml_mask);
        const float c = cpml3_x(v.xzc, P.qx[2], i, j, k, P.npml, P.pml_mask);
        const float r = v.s4 - c3o * (a + c);
        P.s[4][o] = (fs && i == 0) ? 0.0f : r;
    }
    // tauxy -= dtavxyimu*(dvxdy + dvydx) on (nz-2h, ny-h, nx-h)
    if (uxy) {
        const float a = cpml3_y(v.xya, P.qy[1], i, j, k, P.npml, P.pml_mask);
        const float c = cpml3_x(v.xyc, P.qx[1], i, j, k, P.npml, P.pml_mask);
        P.s[3][o] = v.s3 - c2o * (a + c);
    }
    // tauyz -= dtavyzimu*(dvydz + dvzdy) on (nz-h, ny-h, nx-2h)
    if (uyz) {
        const float a = ZP ? v.yza : cpml3_z(v.yza, P.qz[2], i, j, k, P.npml, P.pml_mask);
        const float c = cpml3_y(v.yzc, P.qy[2], i, j, k, P.npml, P.pml_mask);
        const float r = v.s5 - c4o * (a + c);
        P.s[5][o] = (fs && i == 0) ? 0.0f : r;
    }
}

// One (z, y) point of plane k of the elastic stress pass; bt is the injection bucket.
template <int O>
__device__ __forceinline__ void str3d_elastic_point(const StrParams3D& P, const PassExtras& X, int i, int j, int k,
                                                    int bt)
{
    constexpr int H = O - 1;
    const int s1 = (int)P.s1, s2 = (int)P.s2;
    const int o = i + s1 * j + s2 * k;
    const bool un = i < P.nz && j < P.ny;
    const bool uxz = i < P.nz - H && j < P.ny - 2 * H && k < P.nx - H;
    const bool uxy = i < P.nz - 2 * H && j < P.ny - H && k < P.nx - H;
    const bool uyz = i < P.nz - H && j < P.ny - H && k < P.nx - 2 * H;
    // loads of the four updates first, then evaluation and stores
    float nx_[O], ny_[O], nz_[O], xz1[O], xz2[O], xy1[O], xy2[O], yz1[O], yz2[O];
    StrElSamples v{};
    v.il = v.im = 1.0f;
    for (int q = 0; q < 4; ++q) v.q3[q] = v.q2[q] = v.q4[q] = 1.0f;
    const float* imu = P.c[1];
    if (un) {
        load_s<O>(nx_, P.vx, o, s2);
        load_s<O>(ny_, P.vy, o, s1);
        load_s<O>(nz_, P.vz, o, 1);
        v.il = P.c[0][o];
        v.im = imu[o];
        v.s0 = P.s[0][o];
        v.s1 = P.s[1][o];
        v.s2 = P.s[2][o];
    }
    if (uxz) {   // @av_xzi(invmu) on the dvxdz grid: (i, j+h, k) (i+1, ..) (.., k+1) (i+1, .., k+1)
        load_s<O>(xz1, P.vx, i + s1 * (j + H) + s2 * (k + H), 1);
        load_s<O>(xz2, P.vz, (i + H) + s1 * (j + H) + s2 * k, s2);
        v.s4 = P.s[4][o];
        const int a = i + s1 * (j + H) + s2 * k;
        v.q3[0] = imu[a];
        v.q3[1] = imu[a + 1];
        v.q3[2] = imu[a + s2];
        v.q3[3] = imu[a + 1 + s2];
    }
    if (uxy) {   // @av_xyi(invmu) on the dvxdy grid: (i+h, j, k) (.., j+1, ..) (.., k+1) (.., j+1, k+1)
        load_s<O>(xy1, P.vx, (i + H) + s1 * j + s2 * (k + H), s1);
        load_s<O>(xy2, P.vy, (i + H) + s1 * (j + H) + s2 * k, s2);
        v.s3 = P.s[3][o];
        const int a = (i + H) + s1 * j + s2 * k;
        v.q2[0] = imu[a];
        v.q2[1] = imu[a + s1];
        v.q2[2] = imu[a + s2];
        v.q2[3] = imu[a + s1 + s2];
    }
    if (uyz) {   // @av_yzi(invmu) on the dvydz grid: (i, j, k+h) (i+1, ..) (.., j+1, ..) (i+1, j+1, ..)
        load_s<O>(yz1, P.vy, i + s1 * (j + H) + s2 * (k + H), 1);
        load_s<O>(yz2, P.vz, (i + H) + s1 * j + s2 * (k + H), s1);
        v.s5 = P.s[5][o];
        const int a = i + s1 * j + s2 * (k + H);
        v.q4[0] = imu[a];
        v.q4[1] = imu[a + 1];
        v.q4[2] = imu[a + s1];
        v.q4[3] = imu[a + 1 + s1];
    }
    if (un) {
        v.dvx = stencil<O>(nx_) * P.dxI;
        v.dvy = stencil<O>(ny_) * P.dyI;
        v.dvz = stencil<O>(nz_) * P.dzI;
    }
    if (uxz) {
        v.xza = stencil<O>(xz1) * P.dzI;
        v.xzc = stencil<O>(xz2) * P.dxI;
    }
    if (uxy) {
        v.xya = stencil<O>(xy1) * P.dyI;
        v.xyc = stencil<O>(xy2) * P.dxI;
    }
    if (uyz) {
        v.yza = stencil<O>(yz1) * P.dzI;
        v.yzc = stencil<O>(yz2) * P.dyI;
    }
    str3d_elastic_update<O>(P, X, i, j, k, bt, o, un, uxz, uxy, uyz, v);
}

template <int O>
__global__ __launch_bounds__(512) void str3d_elastic(StrParams3D P, PassExtras X)
{
    const int b = P.remap ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int kz = (int)blockDim.z;
    const int nch = (P.nx + kz - 1) / kz;
    if (b >= P.gx * P.gy * nch) {
        if (threadIdx.z == 0) gather_block<8>(X, b - P.gx * P.gy * nch);
        return;
    }
    Idx3 t;
    const int bt = decode(b, P.gx, P.gy, nch, P.xfast != 0, t);
    if (t.k >= P.nx) return;
    str3d_elastic_point<O>(P, X, t.i, t.j, t.k, bt);
}

// ---------------------------------------------------------------------------------------
// elastic, marching along x through workgroup memory
// ---------------------------------------------------------------------------------------
//
// Workgroup (bx, by, chunk) of 64 (z) x TY (y) threads updates planes k0 .. k1-1 of its
// tile in order.  Its iteration p stages plane p of the arrays whose samples it shares
// between threads in workgroup memory (tile rows y0 .. y0+TY+h-1, columns zb .. zb+63+h),
// loading each sample of the plane once, then updates plane k = p - h:
//   - samples of planes k+h (the y- and z-stencils at x offset h, and the invmu / rho
//     samples of the averages at offset h) are read from the staged plane;
//   - x-stencils at the thread's own column (and the averages' samples at x offsets 0
//     and 1) come from per-thread rings of the last h+1 planes' samples;
//   - y- and z-derivatives at x offset 0 are evaluated when their plane is staged and kept
//     for h planes in per-thread delay lines.
// Every load is issued one plane ahead of its use and in the order of use (vector memory
// loads complete in order, so a use waits for everything issued before it): after the
// update of plane k the thread writes plane p+1, loaded during that update, to the other
// slot of the two-slot staging buffer, then issues the update samples of plane k+1 and
// the loads of plane p+2.  One barrier per plane then orders both the writes of a slot
// and the reads of the plane before.  The updates themselves are those of the per-plane
// pass (shared helpers).

// One x-plane of an array over a workgroup's tile as staged in workgroup memory: R = TY+h
// rows of W = 64+h floats, column c holding z = zb + c.  The row loads of 64 lanes start
// at column CM (memory z = zb + CM + lane, a 64-float boundary; lane offset offm of row
// ty, rows ty + TY*q at offm + q*rstep); the h other columns of every row (from column
// CE) are loaded by the first R*h threads of the workgroup (lane offset offe).  Samples
// outside the array's padded extent are never read by a computed update.
template <int TY, int H, int CM, int CE>
struct PlaneStage {
    static constexpr int R = TY + H, W = 64 + H, NQ = (R + TY - 1) / TY;
    // (written only by the loads: a register that a pending load may still write is not
    // assigned otherwise, which would make the wave wait for every load in flight)
    float m[NQ] = {};
    float e = 0.0f;
    __device__ __forceinline__ void load(__amdgpu_buffer_rsrc_t a, uint32_t offm, uint32_t rstep, uint32_t offe)
    {
        const int ty = (int)threadIdx.y, t = (int)threadIdx.x + 64 * ty;
        m[0] = bload(a, offm);   // row ty < TY: every thread (so the loads after a point's
                                 // own samples are counted on every path)
#pragma unroll
        for (int q = 1; q < NQ; ++q)
            if (ty + TY * q < R) m[q] = bload(a, offm + q * rstep);
        if (t < R * H) e = bload(a, offe);
    }
    __device__ __forceinline__ void store(float* L) const
    {
        const int tx = (int)threadIdx.x, ty = (int)threadIdx.y, t = tx + 64 * ty;
        L[ty * W + CM + tx] = m[0];
#pragma unroll
        for (int q = 1; q < NQ; ++q) {
            const int r = ty + TY * q;
            if (r < R) L[r * W + CM + tx] = m[q];
        }
        if (t < R * H) L[(t / H) * W + CE + t % H] = e;
    }
    // lane offset (bytes) of the extra columns for tile origin (zb, y0)
    static __device__ __forceinline__ uint32_t extra_offset(int zb, int y0, int s1)
    {
        const int t = (int)threadIdx.x + 64 * (int)threadIdx.y;
        return 4u * (uint32_t)((zb + CE + t % H) + s1 * (y0 + t / H));
    }
};

// z-CPML of the staged stress pass.  A thread's row i fixes its strip row along z (its
// coefficients are the same on every plane: they are staged once per workgroup in
// workgroup memory), so its memory variables are loaded with the point's other samples a
// plane ahead, by buffer loads at an offset that reads 0 outside the strip (no branch:
// the waits then count every later load), and applied before the update, which skips
// its own z-CPML.  Used where the two z faces' strips do not overlap (q.nd >= 2 npml).
// On 360^3 + 20-cell CPML: stress pass 0.857 against 0.916 ms (80 against 72 VGPRs).  The
// velocity pass measured slower that way (0.836 ms at 81 VGPRs and 5 waves per SIMD, 0.884
// held to 80 and 6 waves, against 0.754) and keeps its z-CPML in the update.
__device__ __forceinline__ int zstrip_row(const Pml& q, int i, int npml, uint32_t mask)
{
    if (i < 0 || i >= q.nd) return -1;   // (a tile's rows past the derivative grid)
    if ((mask & GPI_FACE(0, 0)) && i < npml) return i;
    if ((mask & GPI_FACE(0, 1)) && i >= q.nd - npml) return i - (q.nd - npml) + npml;
    return -1;
}

// stage the coefficients b, a, kI of three z-derivatives at row i (64 rows per tile) in
// L[(3 n + coefficient) * 64 + lane], by the first wave of the workgroup
__device__ __forceinline__ void zstrip_stage(float* L, const Pml* const (&q)[3], int i, int npml, uint32_t mask)
{
    if (threadIdx.y != 0) return;
    const int lane = (int)threadIdx.x;
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        const int c = zstrip_row(*q[n], i, npml, mask);
        L[(3 * n + 0) * 64 + lane] = c >= 0 ? q[n]->b[c] : 0.0f;
        L[(3 * n + 1) * 64 + lane] = c >= 0 ? q[n]->a[c] : 0.0f;
        L[(3 * n + 2) * 64 + lane] = c >= 0 ? q[n]->kI[c] : 0.0f;
    }
}

// memory variable of z-derivative q at (i, j, plane k), 0 where act is false
__device__ __forceinline__ float zstrip_load(const Pml& q, int i, int j, int k, int npml, uint32_t mask, bool act)
{
    const int c = zstrip_row(q, i, npml, mask);
    const uint32_t off = act && c >= 0 ? 4u * (uint32_t)(c + (int)q.ps1 * j) : 0xffffffffu;
    const __amdgpu_buffer_rsrc_t r =
        __builtin_amdgcn_make_buffer_rsrc((void*)(q.psi + q.ps2 * k), (short)0, 0x7fffffff, 0x00020000);
    return bload(r, off);
}

// the z-CPML update of derivative d with memory variable m (loaded) and the staged
// coefficients of derivative n, where act is true
__device__ __forceinline__ float zstrip_apply(float d, float m, const float* L, int n, const Pml& q, int i, int j,
                                              int k, int npml, uint32_t mask, bool act)
{
    const int c = zstrip_row(q, i, npml, mask);
    if (!act || c < 0) return d;
    const int lane = (int)threadIdx.x;
    const float b = L[(3 * n + 0) * 64 + lane], a = L[(3 * n + 1) * 64 + lane], ki = L[(3 * n + 2) * 64 + lane];
    const float mv = b * m + a * d;
    q.psi[c + (int)q.ps1 * j + (int)q.ps2 * k] = mv;
    return d * ki + mv;
}

// ring of the last N samples of a column, oldest first
template <int N>
__device__ __forceinline__ void push(float (&r)[N], float v)
{
#pragma unroll
    for (int m = 0; m < N - 1; ++m) r[m] = r[m + 1];
    r[N - 1] = v;
}

// stencils over a staged plane: along z from column c, along y from row r
template <int O, int W>
__device__ __forceinline__ float lds_z(const float* L, int r, int c)
{
    float v[O];
#pragma unroll
    for (int m = 0; m < O; ++m) v[m] = L[r * W + c + m];
    return stencil<O>(v);
}

template <int O, int W>
__device__ __forceinline__ float lds_y(const float* L, int r, int c)
{
    float v[O];
#pragma unroll
    for (int m = 0; m < O; ++m) v[m] = L[(r + m) * W + c];
    return stencil<O>(v);
}

template <int O, int TY>
__global__ __launch_bounds__(64 * TY) void vel3d_elastic_staged(VelParams3D P, PassExtras X)
{
    constexpr int H = O - 1;
    using S = PlaneStage<TY, H, H, 0>;
    constexpr int W = S::W, A = S::R * S::W;
    // tauxy, tauxz, tauyz, tauyy, tauzz, rho of the staged plane; two slots
    __shared__ float L[2][6][A];
    const int b = tile_order((int)blockIdx.x, (int)gridDim.x, P.remap);
    const int nxw = P.nx - H;
    const int nch = (nxw + P.kc - 1) / P.kc;
    const int nbm = P.gx * P.gy * nch;
    if (b >= nbm) {
        gather_block<8>(X, b - nbm);
        return;
    }
    const int bx = b % P.gx, r0 = b / P.gx, by = r0 % P.gy, ch = r0 / P.gy;
    const int tx = (int)threadIdx.x, ty = (int)threadIdx.y;
    // (j is uniform in a wave, whose 64 lanes are one row: read from the first lane, the
    // y-strip CPML coefficients it indexes are scalar loads)
    const int i = bx * 64 + tx - H, j = __builtin_amdgcn_readfirstlane(by * TY + ty);   // rows shifted by h
    const int k0 = ch * P.kc, k1 = min(k0 + P.kc, nxw);
    const int s1 = (int)P.s1, s2 = (int)P.s2;
    const int zb = bx * 64 - H, y0 = by * TY;
    const bool row = i >= 0;
    const bool cx = row && i < P.nz - 2 * H && j < P.ny - 2 * H;
    const bool cy = row && i < P.nz - 2 * H && j < P.ny - H;
    const bool cz = row && i < P.nz - H && j < P.ny - 2 * H;
    const float* src[6] = {P.s[3], P.s[4], P.s[5], P.s[1], P.s[2], P.rho};
    const int np = P.nx + H;   // planes of every array
    // lane offsets (bytes from a plane's start): staged rows, extra columns, own column
    const uint32_t offm = 4u * (uint32_t)((zb + H + tx) + s1 * (y0 + ty));
    const uint32_t rstep = 4u * (uint32_t)(s1 * TY);
    const uint32_t offe = S::extra_offset(zb, y0, s1);
    const uint32_t offo = offm + 4u * (uint32_t)(s1 * H);   // (i+h, j+h)
    S st[6];
    auto stage = [&](int p) {
#pragma unroll
        for (int a = 0; a < 6; ++a) st[a].load(plane_rsrc(src[a], s2, p, np), offm, rstep, offe);
    };
    // x rings (planes k .. k+h): tauxx (i+h, j+h), tauxy (i, j), tauxz (i, j), rho (i+h, j+h);
    // delay lines of D_ya(tauxy), D_za(tauxz), D_za(tauyz), D_ya(tauyz) at (i, j)
    float rxx[O], rxy[O], rxz[O], rrh[O], da2[O], da3[O], db3[O], de2[O];
#pragma unroll
    for (int m = 0; m < O; ++m) rxx[m] = rxy[m] = rxz[m] = rrh[m] = da2[m] = da3[m] = db3[m] = de2[m] = 0.0f;
    // Loads are issued a plane ahead of their use, in the order of use: vector memory loads
    // complete in order, so a use waits for every load issued before it.  v's loaded
    // samples and st are written only by their loads (a register that a pending load may
    // still write is not assigned otherwise, which would wait for every load in flight).
    // (tauxx is loaded with the plane's own samples, one plane ahead like every operand: a
    // register loaded two planes ahead would be copied at the loop's back edge after the
    // next loads are issued, and that copy waits for all of them)
    VelElSamples v{};
    float xx = 0.0f;   // tauxx at (i+h, j+h) of plane q
    auto own = [&](int q) {   // tauxx of plane q; velocities of the update of plane q - h
        xx = bload(plane_rsrc(P.s[0], s2, q, np), offo);
        const int k = q - H;
        if (k < k0 || k >= k1) return;
        if (cx) v.vxo = bload(plane_rsrc(P.vx, s2, q, np), offo);
        if (cy && k < P.nx - 2 * H) v.vyo = bload(plane_rsrc(P.vy, s2, q, np), offo);
        if (cz && k < P.nx - 2 * H) v.vzo = bload(plane_rsrc(P.vz, s2, q, np), offo);
    };
    const int plast = k0 < k1 ? k1 - 1 + H : k0 - 1;
    if (k0 < k1) {
        stage(k0);
#pragma unroll
        for (int a = 0; a < 6; ++a) st[a].store(L[0][a]);
        own(k0);
        if (k0 < plast) stage(k0 + 1);
    }
    for (int p = k0; p <= plast; ++p) {
        const int slot = (p - k0) & 1;
        __syncthreads();   // plane p staged; plane p-1's slot read by every wave
        const float *Lxy = L[slot][0], *Lxz = L[slot][1], *Lyz = L[slot][2], *Lyy = L[slot][3], *Lzz = L[slot][4],
                    *Lrh = L[slot][5];
        push(rxx, xx);
        push(rxy, Lxy[ty * W + tx]);
        push(rxz, Lxz[ty * W + tx]);
        push(rrh, Lrh[(ty + H) * W + tx + H]);
        push(da2, lds_y<O, W>(Lxy, ty, tx) * P.dyI);
        push(da3, lds_z<O, W>(Lxz, ty, tx) * P.dzI);
        push(db3, lds_z<O, W>(Lyz, ty, tx) * P.dzI);
        push(de2, lds_y<O, W>(Lyz, ty, tx) * P.dyI);
        const int k = p - H;
        if (k >= k0) {
            const bool ux = cx, uy = cy && k < P.nx - 2 * H, uz = cz && k < P.nx - 2 * H;
            v.rx0 = rrh[0];                          // rho (i+h, j+h, k), (.., k+1)
            v.rx1 = rrh[1];
            v.ry0 = Lrh[ty * W + tx + H];            // rho (i+h, j, k+h), (.., j+1, ..)
            v.ry1 = Lrh[(ty + 1) * W + tx + H];
            v.rz0 = Lrh[(ty + H) * W + tx];          // rho (i, j+h, k+h), (i+1, ..)
            v.rz1 = Lrh[(ty + H) * W + tx + 1];
            v.a1 = stencil<O>(rxx) * P.dxI;          // D_xi(tauxx)
            v.a2 = da2[0];
            v.a3 = da3[0];
            v.b1 = stencil<O>(rxy) * P.dxI;          // D_xa(tauxy)
            v.b2 = lds_y<O, W>(Lyy, ty, tx + H) * P.dyI;   // D_yi(tauyy) at (i+h, j, k+h)
            v.b3 = db3[0];
            v.e1 = stencil<O>(rxz) * P.dxI;          // D_xa(tauxz)
            v.e2 = de2[0];
            v.e3 = lds_z<O, W>(Lzz, ty + H, tx) * P.dzI;   // D_zi(tauzz) at (i, j+h, k+h)
            const int bt = bx + P.gx * (by + P.gy * k);
            const int io = opaque(i), jo = __builtin_amdgcn_readfirstlane(opaque(j));
            vel3d_elastic_update<O>(P, X, io, jo, k, bt, (io + H) + s1 * (jo + H) + s2 * p, ux, uy, uz, v);
        }
        if (p < plast) {   // plane p+1 (loaded) into the other slot; loads for the next planes
#pragma unroll
            for (int a = 0; a < 6; ++a) st[a].store(L[slot ^ 1][a]);
            own(p + 1);
            stage(min(p + 2, plast));   // (also on the last trip: the waits then count these loads on every path)
        }
    }
}

template <int O, int TY, bool ZP>
__global__ __launch_bounds__(64 * TY) void str3d_elastic_staged(StrParams3D P, PassExtras X)
{
    constexpr int H = O - 1;
    using S = PlaneStage<TY, H, 0, 64>;
    constexpr int W = S::W, A = S::R * S::W;
    // vx, vy, vz, invmu of the staged plane; two slots
    __shared__ float L[2][4][A];
    __shared__ float Lz[ZP ? 9 * 64 : 1];   // z-strip coefficients of the tile's rows
    const int b = tile_order((int)blockIdx.x, (int)gridDim.x, P.remap);
    const int nch = (P.nx + P.kc - 1) / P.kc;
    const int nbm = P.gx * P.gy * nch;
    if (b >= nbm) {
        gather_block<8>(X, b - nbm);
        return;
    }
    const int bx = b % P.gx, r0 = b / P.gx, by = r0 % P.gy, ch = r0 / P.gy;
    const int tx = (int)threadIdx.x, ty = (int)threadIdx.y;
    const int i = bx * 64 + tx, j = __builtin_amdgcn_readfirstlane(by * TY + ty);   // (uniform in a wave)
    const int k0 = ch * P.kc, k1 = min(k0 + P.kc, P.nx);
    const int s1 = (int)P.s1, s2 = (int)P.s2;
    const int zb = bx * 64, y0 = by * TY;
    const bool cn = i < P.nz && j < P.ny;
    const bool cxz = i < P.nz - H && j < P.ny - 2 * H;
    const bool cxy = i < P.nz - 2 * H && j < P.ny - H;
    const bool cyz = i < P.nz - H && j < P.ny - H;
    const float* src[4] = {P.vx, P.vy, P.vz, P.c[1]};
    const int np = P.nx + H;   // planes of every array
    // lane offsets (bytes from a plane's start): staged rows (row ty: the thread's own
    // point), extra columns
    const uint32_t offm = 4u * (uint32_t)(i + s1 * j);
    const uint32_t rstep = 4u * (uint32_t)(s1 * TY);
    const uint32_t offe = S::extra_offset(zb, y0, s1);
    S st[4];
    auto stage = [&](int p) {
#pragma unroll
        for (int a = 0; a < 4; ++a) st[a].load(plane_rsrc(src[a], s2, p, np), offm, rstep, offe);
    };
    // x rings (planes k .. k+h): vx (i, j), vy (i+h, j+h), vz (i+h, j+h), invmu (i, j),
    // invmu (i, j+h) and (i+1, j+h), invmu (i+h, j) and (i+h, j+1);
    // delay lines of D_ya(vy), D_za(vz) at (i, j)
    float rvx[O], rvy[O], rvz[O], rim[O], q3a[O], q3b[O], q2a[O], q2b[O], dvy[O], dvz[O];
#pragma unroll
    for (int m = 0; m < O; ++m)
        rvx[m] = rvy[m] = rvz[m] = rim[m] = q3a[m] = q3b[m] = q2a[m] = q2b[m] = dvy[m] = dvz[m] = 0.0f;
    // Loads are issued a plane ahead of their use, in the order of use (vector memory loads
    // complete in order); v's loaded samples and st are written only by their loads.
    StrElSamples v{};
    v.il = 1.0f;
    const Pml* const qz3[3] = {&P.qz[0], &P.qz[1], &P.qz[2]};
    float zm[3] = {0.0f, 0.0f, 0.0f};   // z-strip memory variables of the next update
    if constexpr (ZP) zstrip_stage(Lz, qz3, i, P.npml, P.pml_mask);   // ordered by the loop's first barrier
    auto own = [&](int q) {   // stresses and invlambda of the update of plane k = q - h
        const int k = q - H;
        if (k < k0 || k >= k1) return;
        if (cn) {
            v.il = bload(plane_rsrc(P.c[0], s2, k, np), offm);
            v.s0 = bload(plane_rsrc(P.s[0], s2, k, np), offm);
            v.s1 = bload(plane_rsrc(P.s[1], s2, k, np), offm);
            v.s2 = bload(plane_rsrc(P.s[2], s2, k, np), offm);
        }
        if (cxz && k < P.nx - H) v.s4 = bload(plane_rsrc(P.s[4], s2, k, np), offm);
        if (cxy && k < P.nx - H) v.s3 = bload(plane_rsrc(P.s[3], s2, k, np), offm);
        if (cyz && k < P.nx - 2 * H) v.s5 = bload(plane_rsrc(P.s[5], s2, k, np), offm);
        if constexpr (ZP) {
            zm[0] = zstrip_load(P.qz[0], i, j, k, P.npml, P.pml_mask, cn);
            zm[1] = zstrip_load(P.qz[1], i, j, k, P.npml, P.pml_mask, cxz && k < P.nx - H);
            zm[2] = zstrip_load(P.qz[2], i, j, k, P.npml, P.pml_mask, cyz && k < P.nx - 2 * H);
        }
    };
    const int plast = k0 < k1 ? k1 - 1 + H : k0 - 1;
    if (k0 < k1) {
        stage(k0);
#pragma unroll
        for (int a = 0; a < 4; ++a) st[a].store(L[0][a]);
        own(k0);
        if (k0 < plast) stage(k0 + 1);
    }
    for (int p = k0; p <= plast; ++p) {
        const int slot = (p - k0) & 1;
        __syncthreads();   // plane p staged; plane p-1's slot read by every wave
        const float *Lvx = L[slot][0], *Lvy = L[slot][1], *Lvz = L[slot][2], *Lmu = L[slot][3];
        push(rvx, Lvx[ty * W + tx]);
        push(rvy, Lvy[(ty + H) * W + tx + H]);
        push(rvz, Lvz[(ty + H) * W + tx + H]);
        push(rim, Lmu[ty * W + tx]);
        push(q3a, Lmu[(ty + H) * W + tx]);
        push(q3b, Lmu[(ty + H) * W + tx + 1]);
        push(q2a, Lmu[ty * W + tx + H]);
        push(q2b, Lmu[(ty + 1) * W + tx + H]);
        push(dvy, lds_y<O, W>(Lvy, ty, tx) * P.dyI);
        push(dvz, lds_z<O, W>(Lvz, ty, tx) * P.dzI);
        const int k = p - H;
        if (k >= k0) {
            const bool un = cn, uxz = cxz && k < P.nx - H, uxy = cxy && k < P.nx - H, uyz = cyz && k < P.nx - 2 * H;
            v.im = rim[0];
            v.q3[0] = q3a[0];   // invmu (i, j+h, k), (i+1, ..), (.., k+1), (i+1, .., k+1)
            v.q3[1] = q3b[0];
            v.q3[2] = q3a[1];
            v.q3[3] = q3b[1];
            v.q2[0] = q2a[0];   // invmu (i+h, j, k), (.., j+1, ..), (.., k+1), (.., j+1, k+1)
            v.q2[1] = q2b[0];
            v.q2[2] = q2a[1];
            v.q2[3] = q2b[1];
            v.q4[0] = Lmu[ty * W + tx];   // invmu (i, j, k+h), (i+1, ..), (.., j+1, ..), (i+1, j+1, ..)
            v.q4[1] = Lmu[ty * W + tx + 1];
            v.q4[2] = Lmu[(ty + 1) * W + tx];
            v.q4[3] = Lmu[(ty + 1) * W + tx + 1];
            v.dvx = stencil<O>(rvx) * P.dxI;                 // D_xa(vx)
            v.dvy = dvy[0];
            v.dvz = dvz[0];
            v.xza = lds_z<O, W>(Lvx, ty + H, tx) * P.dzI;    // D_zi(vx) at (i, j+h, k+h)
            v.xzc = stencil<O>(rvz) * P.dxI;                 // D_xi(vz)
            v.xya = lds_y<O, W>(Lvx, ty, tx + H) * P.dyI;    // D_yi(vx) at (i+h, j, k+h)
            v.xyc = stencil<O>(rvy) * P.dxI;                 // D_xi(vy)
            v.yza = lds_z<O, W>(Lvy, ty + H, tx) * P.dzI;    // D_zi(vy) at (i, j+h, k+h)
            v.yzc = lds_y<O, W>(Lvz, ty, tx + H) * P.dyI;    // D_yi(vz) at (i+h, j, k+h)
            const int bt = bx + P.gx * (by + P.gy * k);
            const int io = opaque(i), jo = __builtin_amdgcn_readfirstlane(opaque(j));
            if constexpr (ZP) {
                v.dvz = zstrip_apply(v.dvz, zm[0], Lz, 0, P.qz[0], io, jo, k, P.npml, P.pml_mask, un);
                v.xza = zstrip_apply(v.xza, zm[1], Lz, 1, P.qz[1], io, jo, k, P.npml, P.pml_mask, uxz);
                v.yza = zstrip_apply(v.yza, zm[2], Lz, 2, P.qz[2], io, jo, k, P.npml, P.pml_mask, uyz);
            }
            str3d_elastic_update<O, ZP>(P, X, io, jo, k, bt, io + s1 * jo + s2 * k, un, uxz, uxy, uyz, v);
        }
        if (p < plast) {   // plane p+1 (loaded) into the other slot; loads for the next planes
#pragma unroll
            for (int a = 0; a < 4; ++a) st[a].store(L[slot ^ 1][a]);
            own(p + 1);
            stage(min(p + 2, plast));   // (also on the last trip: the waits then count these loads on every path)
        }
    }
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

// Elastic passes staged through workgroup memory while marching along x: GPI_3D_ESTAGE =
// 1 (default) at order 2, 2 at every order, 0 never.  Measured on 320^3 + 20-cell CPML
// (ms per launch, stress / velocity): order 2 staged 0.937 / 0.829 against per-plane
// 1.041 / 0.825; order 4 1.228 / 1.162 against 1.167 / 0.952, order 8 2.216 / 2.414
// against 1.679 / 1.343 (the staged passes' rings and delay lines need 107-193 VGPRs from
// order 4 on).  GPI_3D_EMARCH planes per workgroup (default 6: 0.849 / 0.761 ms against
// 0.856 / 0.763 with 8, 5 and 7 in between, 12 and 16 slower); workgroup order
// GPI_3D_XWIN = R > 0 (default 4) gives each XCD runs of R rows of tiles (tile_order),
// 0 the order of GPI_XCD_REMAP3 (windows: 0.929 / 0.815 ms against 0.937 / 0.829).
// Workgroup order of the acoustic passes: GPI_3D_AWIN = R > 0 (default 4) runs of R tile
// rows per XCD (tile_order), 0 the order of GPI_XCD_REMAP3.  On 440^3 the windows cut the
// measured traffic from 1.14x / 1.19x to 1.05x / 1.07x of the algorithmic bytes at the
// same speed (0.738 / 0.455 ms; R = 2, 4 alike, 8 0.744 / 0.462).
int acoustic_window()
{
    static int r = -1;
    if (r < 0) {
        const char* e = getenv("GPI_3D_AWIN");
        r = e ? std::max(0, atoi(e)) : 4;
    }
    return r;
}

int elastic_window()
{
    static int r = -1;
    if (r < 0) {
        const char* e = getenv("GPI_3D_XWIN");
        r = e ? std::max(0, atoi(e)) : 4;
    }
    return r;
}

bool elastic_staged(int order)
{
    static int e = -1;
    if (e < 0) {
        const char* v = getenv("GPI_3D_ESTAGE");
        e = v ? atoi(v) : 1;
    }
    return e == 2 || (e == 1 && order == 2);
}

// The staged stress pass takes the z-CPML from its own loads (zstrip_*) when some z face has
// a CPML and the two faces' strips do not overlap (GPI_3D_ZPRE=0 turns it off).
bool zstrip_prefetch(const Pml (&qz)[3], int npml, uint32_t mask)
{
    static int e = -1;
    if (e < 0) {
        const char* v = getenv("GPI_3D_ZPRE");
        e = v ? atoi(v) != 0 : 1;
    }
    if (!e || !(mask & (GPI_FACE(0, 0) | GPI_FACE(0, 1)))) return false;
    for (int n = 0; n < 3; ++n)
        if (qz[n].nd < 2 * npml) return false;
    return true;
}

int elastic_march(int n)
{
    static int kc = -1;
    if (kc < 0) {
        const char* e = getenv("GPI_3D_EMARCH");
        kc = e ? atoi(e) : 6;
        if (kc < 1) kc = 1;
    }
    return std::min(kc, std::max(n, 1));
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
    if (m.g.physics == 1 && elastic_staged(O)) {
        P.kc = elastic_march(P.nx - H);
        P.remap = elastic_window() ? std::max(2, P.gx * elastic_window()) : remap3();
        dim3 grd(P.gx * P.gy * ((P.nx - H + P.kc - 1) / P.kc) + X.gather_blocks), blk(64, ty);
        if (ty == 2) hipLaunchKernelGGL((vel3d_elastic_staged<O, 2>), grd, blk, 0, s, P, X);
        else if (ty == 8) hipLaunchKernelGGL((vel3d_elastic_staged<O, 8>), grd, blk, 0, s, P, X);
        else hipLaunchKernelGGL((vel3d_elastic_staged<O, 4>), grd, blk, 0, s, P, X);
        return;
    }
    const int kz = m.g.physics == 0 ? 1 : elastic_planes(ty);
    P.kc = m.g.physics == 0 ? march_planes(P.nx - H) : kz;
    P.remap = m.g.physics == 0 && acoustic_window() ? std::max(2, P.gx * acoustic_window()) : remap3();
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
    if (m.g.physics == 1 && elastic_staged(O)) {
        P.kc = elastic_march(P.nx);
        P.remap = elastic_window() ? std::max(2, P.gx * elastic_window()) : remap3();
        dim3 grd(P.gx * P.gy * ((P.nx + P.kc - 1) / P.kc) + X.gather_blocks), blk(64, ty);
        if (zstrip_prefetch(P.qz, P.npml, P.pml_mask)) {
            if (ty == 2) hipLaunchKernelGGL((str3d_elastic_staged<O, 2, true>), grd, blk, 0, s, P, X);
            else if (ty == 8) hipLaunchKernelGGL((str3d_elastic_staged<O, 8, true>), grd, blk, 0, s, P, X);
            else hipLaunchKernelGGL((str3d_elastic_staged<O, 4, true>), grd, blk, 0, s, P, X);
        } else {
            if (ty == 2) hipLaunchKernelGGL((str3d_elastic_staged<O, 2, false>), grd, blk, 0, s, P, X);
            else if (ty == 8) hipLaunchKernelGGL((str3d_elastic_staged<O, 8, false>), grd, blk, 0, s, P, X);
            else hipLaunchKernelGGL((str3d_elastic_staged<O, 4, false>), grd, blk, 0, s, P, X);
        }
        return;
    }
    const int kz = m.g.physics == 0 ? 1 : elastic_planes(ty);
    P.kc = m.g.physics == 0 ? march_planes(P.nx) : kz;
    P.remap = m.g.physics == 0 && acoustic_window() ? std::max(2, P.gx * acoustic_window()) : remap3();
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
