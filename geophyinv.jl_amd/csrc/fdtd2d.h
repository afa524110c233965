// Parameters and launchers of the 2-D stencil passes.
#pragma once
#include "gpi_internal.h"
#include "pass_extras.h"

namespace gpi {

struct VelParams2D {
    const float* s0;   // acoustic p | elastic tauxx
    const float* s1;   // elastic tauzz
    const float* s2;   // elastic tauxz
    float* vx;
    float* vz;
    const float* cx;   // dtinvavxirho
    const float* cz;   // dtinvavzirho
    const float* rho;  // density, for velocity sources
    Pml qx0, qx1;      // x-derivative CPML: acoustic dpdx | elastic dtauxxdx, dtauxzdx
    Pml qz0, qz1;      // z-derivative CPML: acoustic dpdz | elastic dtauzzdz, dtauxzdz
    int nz, nx, ld, npml;
    int gx, gy;        // stencil workgroups along z and x
    uint32_t pml_mask, rigid_mask;
    float dxI, dzI, dt;
};

struct StrParams2D {
    const float* vx;
    const float* vz;
    float* s0;         // acoustic p | elastic tauxx
    float* s1;         // elastic tauzz
    float* s2;         // elastic tauxz
    const float* c0;   // acoustic dtK | elastic dtM
    const float* c1;   // elastic dtlambda
    const float* c2;   // elastic dtavmu
    Pml qx0, qx1;      // acoustic dvxdx | elastic dvxdx, dvzdx
    Pml qz0, qz1;      // acoustic dvzdz | elastic dvzdz, dvxdz
    int nz, nx, ld, npml;
    int gx, gy;
    uint32_t pml_mask, free_mask;
    float dxI, dzI;
};

// workgroup grid of each pass (64 x 4 threads per workgroup, workgroup b covers rows
// 64*(b % gx).. and columns 4*(b / gx)..); returns gx*gy
int velocity_blocks_2d(const Model& m, int* gx, int* gy);
int stress_blocks_2d(const Model& m, int* gx, int* gy);

// The pass launchers return -1 without launching when the state lacks an array the pass
// writes (a wavefield, or the CPML memory of an active derivative).
int velocity_pass_2d(const Model& m, const State& st, hipStream_t s, const PassExtras& X);
int stress_pass_2d(const Model& m, const State& st, hipStream_t s, const PassExtras& X);

}  // namespace gpi
