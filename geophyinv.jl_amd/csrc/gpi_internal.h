// Internal declarations shared by the libgpi translation units (HIP, gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "gpi.h"

namespace gpi {

// Device field view: element (iz, iy, ix) at p[iz + s1*iy + s2*ix].  All wavefields and
// coefficient arrays of a model share the same pitches, so one offset addresses every
// array; CPML memory arrays carry their own pitches.
struct View {
    float* p = nullptr;
    int32_t n[3] = {0, 0, 0};   // logical sizes z, y, x
    int64_t s1 = 0, s2 = 0;     // strides of y and x in floats
};

// CPML state of one derivative field: its memory array and the a, b, kI vectors.
struct Pml {
    float* psi = nullptr;       // memory variables, axis dimension = 2*npml
    int64_t ps1 = 0, ps2 = 0;   // strides of psi (z stride is 1)
    const float* a = nullptr;
    const float* b = nullptr;
    const float* kI = nullptr;
    int32_t nd = 0;             // length of the derivative field along its own axis
    int32_t active = 0;         // nonzero when the derivative exists
};

// Sparse point set prepared on the device.
//  - gather form (receivers): one column per point, nnz (index, weight) in CSC order.
//  - scatter form (sources): unique target cells, each with its contributions listed in
//    ascending source order (the accumulation order of SparseArrays' CSC mul!).
struct DevPoints {
    int32_t field = -1;
    int32_t npts = 0;
    int32_t nnz = 0;
    int64_t* idx = nullptr;     // gather: npts*nnz padded-layout offsets
    float* w = nullptr;         // gather: npts*nnz weights
    int32_t ntarget = 0;        // scatter: unique target cells (interior only for velocity)
    int32_t* blk = nullptr;     // scatter: per-workgroup target ranges of the owning pass
    bool fused = false;         // scatter: applied inside the owning pass
    int64_t* toff = nullptr;    // scatter: padded offsets of targets
    int32_t* tptr = nullptr;    // scatter: ntarget+1 CSR row pointers
    int32_t* tsrc = nullptr;    // scatter: source column per contribution
    float* tw = nullptr;        // scatter: weight per contribution
    float* wav = nullptr;       // scatter: nt x npts wavelets, point fastest
    float* rec = nullptr;       // gather: nt x npts records, point fastest
};

struct Model {
    gpi_grid g{};
    int h = 1;
    int64_t pitch0 = 0, pitch1 = 1;    // wavefield/coefficient pitches (z, y)
    int64_t slab = 0;                  // floats per array (pitch0*pitch1*(nx+h))
    std::vector<void*> allocs;
    float* coef[GPI_NCOEF] = {};
    float* field[GPI_NFIELDS] = {};    // wavefields only (no derivative arrays)
    Pml pml[GPI_NFIELDS];
    float* pmlvec = nullptr;           // all a, b, kI vectors
    uint64_t bytes = 0;
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {};
    std::vector<hipEvent_t> sample_ev;   // 4 events per sampled step
    uint64_t budget = 0, used = 0;
    std::string err;
    std::vector<Model*> models;
    gpi_timing timing{};
};

// stand-alone sparse kernels (sparse.hip): injection after a pass, for source sets that
// cannot ride along with the pass (targets on a free-surface row or dense target lists)
void inject(const Model& m, const DevPoints& pts, int64_t it, hipStream_t s);

}  // namespace gpi
