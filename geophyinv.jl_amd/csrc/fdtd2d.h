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
    const float* rho;  // density: dtinvav{x,z}irho are formed from it, and velocity sources
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
    const float* c0;   // acoustic dtK | elastic invlambda
    const float* c1;   // elastic invmu: dtM, dtlambda and dtavmu are formed from c0, c1
    float dt;
    Pml qx0, qx1;      // acoustic dvxdx | elastic dvxdx, dvzdx
    Pml qz0, qz1;      // acoustic dvzdz | elastic dvzdz, dvxdz
    int nz, nx, ld, npml;
    int gx, gy;
    uint32_t pml_mask, free_mask;
    float dxI, dzI;
};

// Injection buckets (64-row segment of one derivative-grid column): bucket of point (i, j)
// = i/64 + gz*j; returns the number of buckets.  Velocity pass: (i, j) = (iz-h, ix-h) of
// the interior velocity cell; stress pass: (iz, ix).
int velocity_buckets_2d(const Model& m, int* gz);
int stress_buckets_2d(const Model& m, int* gz);

// Workgroup shape of a batched launch: 64 x ty threads, rx columns per thread.
struct TileShape {
    int ty, rx;
};
constexpr int NTILE_SHAPES = 4;
extern const TileShape TILE_SHAPES[NTILE_SHAPES];
void velocity_grid_2d(const Model& m, const TileShape& t, int* gx, int* gy);
void stress_grid_2d(const Model& m, const TileShape& t, int* gx, int* gy);

// Pass parameters of a model and state (gx, gy left for the launcher); false when the state
// lacks an array the pass writes.
bool vel_params_2d(const Model& m, const State& st, VelParams2D* out);
bool str_params_2d(const Model& m, const State& st, StrParams2D* out);

// One resident shot of a batched launch (device table entry).
struct BatchShot2D {
    VelParams2D vp;
    StrParams2D sp;
    PassExtras xv, xs;
    int64_t start;      // batch step at which the shot's it = 0
    int64_t nt;
};

// One pass of every shot of a device table: n = *nshots shots, first[s] = first workgroup
// of shot s, first[n] = the workgroups they need; the launch has nblocks >= first[n]
// workgroups.  All shots share physics, order and tile shape.
void batch_pass_2d(int physics, int order, const TileShape& t, bool velocity, const BatchShot2D* D,
                   const int* first, const int* nshots, int nblocks, const int64_t* step0, int koff,
                   hipStream_t s);
// *step0 = step, stream-ordered (a one-thread kernel: no memset node in a captured chunk).
void batch_set_step(int64_t* step0, int64_t step, hipStream_t s);

// The pass launchers return -1 without launching when the state lacks an array the pass
// writes (a wavefield, or the CPML memory of an active derivative).
int velocity_pass_2d(const Model& m, const State& st, hipStream_t s, const PassExtras& X);
int stress_pass_2d(const Model& m, const State& st, hipStream_t s, const PassExtras& X);

}  // namespace gpi
