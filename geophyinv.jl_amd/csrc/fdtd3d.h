// Parameters and launchers of the 3-D stencil passes.
#pragma once
#include "gpi_internal.h"
#include "pass_extras.h"

namespace gpi {

struct VelParams3D {
    // acoustic: s[0] = p; elastic: s = tauxx, tauyy, tauzz, tauxy, tauxz, tauyz
    const float* s[6];
    float* vx;
    float* vy;
    float* vz;
    const float* cx;   // acoustic: dtinvavxirho (the elastic pass forms it from rho)
    const float* cy;   // acoustic: dtinvavyirho
    const float* cz;   // acoustic: dtinvavzirho
    const float* rho;  // density (velocity sources; elastic: the density coefficients)
    // CPML of the derivatives; acoustic: x[0] dpdx, y[0] dpdy, z[0] dpdz
    // elastic: vx uses x[0] dtauxxdx, y[0] dtauxydy, z[0] dtauxzdz
    //          vy uses x[1] dtauxydx, y[1] dtauyydy, z[1] dtauyzdz
    //          vz uses x[2] dtauxzdx, y[2] dtauyzdy, z[2] dtauzzdz
    Pml qx[3], qy[3], qz[3];
    int nz, ny, nx, npml;
    int64_t s1, s2;    // y and x strides of the pitched layout
    int gx, gy;        // workgroups along z and y; the grid's third factor runs over x
    int kc;            // x planes per workgroup (acoustic passes march along x)
    int remap;         // workgroup order: 0 launch order, 1 XCD-aware; staged elastic passes: window run
    int xfast;         // elastic: consecutive workgroups walk along x (then z, y tiles)
    uint32_t pml_mask, rigid_mask;
    float dxI, dyI, dzI, dt;
};

struct StrParams3D {
    const float* vx;
    const float* vy;
    const float* vz;
    // acoustic: s[0] = p; elastic: tauxx, tauyy, tauzz, tauxy, tauxz, tauyz
    float* s[6];
    // acoustic: c[0] = dtK; elastic: c[0] = invlambda, c[1] = invmu (dtM, dtlambda and the
    // averaged mu coefficients are formed from them in the pass)
    const float* c[5];
    float dt;
    // acoustic: x[0] dvxdx, y[0] dvydy, z[0] dvzdz
    // elastic:  x: dvxdx, dvydx, dvzdx; y: dvydy, dvxdy, dvzdy; z: dvzdz, dvxdz, dvydz
    Pml qx[3], qy[3], qz[3];
    int nz, ny, nx, npml;
    int64_t s1, s2;
    int gx, gy;
    int kc;            // x planes per workgroup (acoustic passes march along x)
    int remap;         // workgroup order: 0 launch order, 1 XCD-aware; staged elastic passes: window run
    int xfast;         // elastic: consecutive workgroups walk along x (then z, y tiles)
    uint32_t pml_mask, free_mask;
    float dxI, dyI, dzI;
};

// Rows of the 3-D velocity passes are shifted by h: workgroup row bx covers
// derivative-grid rows 64*bx - h .. 64*bx + 63 - h, so the velocity (and, acoustic,
// pressure) columns it reads and writes at row i + h start on a 128-byte boundary (no
// cache line is shared with the neighbouring workgroup, which may run on another XCD).
// The acoustic density coefficients it reads at row i are stored h floats past their
// allocation for the same reason.  Returns the shift.
int vel3d_row_shift(const Model& m);
// y rows per workgroup of the 3-D passes: acoustic GPI_3D_TY (4 or 8, default 4: on 440^3
// the 64 x 8 tile ran 9 % slower), elastic GPI_3D_TYE (2, 4 or 8, default 4: on 360^3 the
// 64 x 8 tile ran 12 % slower).  The elastic passes also take GPI_3D_EKZ x-planes per
// workgroup (1, 2 or 4, default 1, at most 512 threads): 64 x 2 x 2 fetches 20 % fewer
// bytes than 64 x 4 x 1 at the same speed, the re-fetched planes being cache hits.
int tile_rows_3d(const Model& m);
int velocity_blocks_3d(const Model& m, int* gx, int* gy);
int stress_blocks_3d(const Model& m, int* gx, int* gy);
int velocity_pass_3d(const Model& m, const State& st, hipStream_t s, const PassExtras& X);
int stress_pass_3d(const Model& m, const State& st, hipStream_t s, const PassExtras& X);

}  // namespace gpi
