// Internal declarations shared by the libgpi translation units (HIP, gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <map>
#include <unordered_map>
#include <string>
#include <vector>

#include "gpi.h"

namespace gpi {

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
    int32_t* blk = nullptr;     // scatter: target ranges per injection bucket of the owning pass
    int32_t nbucket = 0;
    int64_t* toff = nullptr;    // scatter: padded offsets of targets (then 2 rho offsets each)
    int32_t* tptr = nullptr;    // scatter: ntarget+1 CSR row pointers
    int32_t* tsrc = nullptr;    // scatter: source column per contribution
    float* tw = nullptr;        // scatter: weight per contribution
    float* wav = nullptr;       // scatter: nt x npts wavelets, point fastest
    float* rec = nullptr;       // gather: nt x npts records, point fastest
};

// Wavefields and CPML memory of one propagating wavefield.  A model can serve several
// states (two in the adjoint run, many in batched serving).
struct State {
    float* field[GPI_NFIELDS] = {};   // wavefields only (no derivative arrays)
    float* psi[GPI_NFIELDS] = {};     // CPML memory of derivative fields
    std::vector<void*> allocs;
    uint64_t bytes = 0;
};

struct Model {
    gpi_grid g{};
    int h = 1;
    int64_t pitch0 = 0, pitch1 = 1;    // wavefield/coefficient pitches (z, y)
    int64_t slab = 0;                  // floats per array (pitch0*pitch1*(nx+h))
    std::vector<void*> allocs;
    uint64_t bytes = 0;
    float* coef[GPI_NCOEF] = {};
    Pml pml[GPI_NFIELDS];              // strides and coefficient vectors; psi lives in a State
    float* pmlvec = nullptr;           // all a, b, kI vectors
    State* st0 = nullptr;              // the state of gpi_forward / gpi_forward_save
};

struct Replay;

// Store of one :forward_save shot.  Time reversal (src/fdtd/boundary.jl, every == 0): nt
// strips of the stored stress fields per axis, plus the final snapshots.  Checkpoints
// (every > 0): the whole state after every `every`-th step, the velocities at nt+1/2 and
// the forward sources (Replay) for recomputing the forward wavefield segment by segment.
struct Store {
    Model* model = nullptr;
    int64_t nt = 0;
    int nfield = 0;
    int field[3] = {};                 // stored stress fields
    float* strip[3][3] = {};           // [field][axis]: nt x strip_size
    int64_t strip_size[3][3] = {};
    float* snap[GPI_NFIELDS] = {};     // pitched snapshots: -stress at nt, v at nt+1/2
    int64_t every = 0;                 // checkpoint interval (0: boundary strips)
    std::vector<float*> ckpt;          // state after step k*every - 1, k = 1 .. ceil(nt/every)-1
    Replay* replay = nullptr;
    std::vector<void*> allocs;
    uint64_t bytes = 0;
};

// Sub-allocator of device memory for contexts that allocate and free at a high rate (the
// serving scheduler): segments from hipMalloc, carved first-fit into 256-byte aligned
// blocks; a freed block coalesces with its free neighbours.  Not thread-safe: one thread
// (the scheduler) allocates and frees; reserved() may be read from any thread.
class DevicePool {
public:
    ~DevicePool();
    hipError_t alloc(void** p, size_t bytes);
    void free(void* p);
    void trim();                           // hipFree every segment with no live block
    uint64_t reserved() const { return reserved_.load(std::memory_order_relaxed); }

private:
    struct Seg {
        char* base;
        size_t size, live;
        std::map<size_t, size_t> free;     // offset -> size
    };
    std::vector<Seg*> segs_;
    std::unordered_map<void*, std::pair<Seg*, size_t>> blocks_;   // block -> (segment, size)
    std::atomic<uint64_t> reserved_{0};
    bool carve(Seg* g, size_t n, void** p);
};

// Pinned staging of host-to-device copies (copy_h2d): the source is copied into the ring
// and the upload is queued on the context's stream; the ring wraps after the stream drained.
struct PinnedStage {
    char* buf = nullptr;
    size_t cap = 0, off = 0;
};

// Fixed device region that a call carves its scratch buffers from instead of allocating
// (gpi_forward_static): 256-byte aligned bump allocation, nothing is ever freed.
struct Arena {
    char* base = nullptr;
    size_t cap = 0, off = 0;
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    DevicePool* pool = nullptr;          // when set, dev_alloc carves from it
    PinnedStage* stage = nullptr;        // when set, copy_h2d is asynchronous on `stream`
    // When set, frees are deferred: pointers and bytes go here and the owner releases them
    // once no queued work can reference them (the serving scheduler drains its streams first).
    std::vector<void*>* defer_free = nullptr;
    uint64_t* defer_bytes = nullptr;
    // When set, Scratch buffers are carved from this region (no device allocation).
    Arena* arena = nullptr;
    hipEvent_t ev[4] = {};
    std::vector<hipEvent_t> sample_ev;   // 4 events per sampled step
    uint64_t budget = 0, used = 0;
    std::string err;
    std::vector<Model*> models;
    std::vector<Store*> stores;
    gpi_timing timing{};
};

// stand-alone sparse kernels (sparse.hip): injection after a pass, for source sets that
// cannot ride along with the pass (targets on a free-surface row or dense target lists)
void inject(const Model& m, const State& st, const DevPoints& pts, int64_t it, hipStream_t s);
// illumination: acc[o] += p[o]^2 over n floats (FP32 square, FP64 sum)
void illum_accumulate(const float* p, double* acc, int64_t n, hipStream_t s);
// receiver records, nt x npts point fastest -> npts x nt time fastest (the host layout)
// nonfinite: device int set to 1 when a record sample is Inf or NaN (not cleared here)
void records_transpose(const float* rec, float* out, int64_t nt, int npts, int* nonfinite, hipStream_t s);

// time-reversal kernels (strips.hip)
void strip_copy(const Model& m, float* field, const int64_t fn[3], int axis, int64_t np, float* store,
                bool save, hipStream_t s);
void negate_copy(float* dst, const float* src, int64_t n, hipStream_t s);
// cur/prev/next: p1, vx1, vz1, p2, vx2, vz2 (pitched slabs); next <- cur after imaging
void image_acoustic_2d(const Model& m, const float* const cur[6], const float* const prev[6], float* const next[6],
                       float* g_invK, float* g_rho, float dtI, hipStream_t s);
// Imaging of one adjoint step with the exact forward wavefield (acoustic, 2-D and 3-D).
// a2/c2/n2: the adjoint p and velocities (x, y, z; y unused in 2-D) before the step, after
// it, and their copy destination; pa, pb: forward p_{n-1}, p_n; va, vb: forward v_n, v_{n+1}.
// g_rho (and the velocity arguments) may be null.  Accumulators are FP64.
struct ImageSets {
    const float* a2[4];
    const float* c2[4];
    float* n2[4];
    const float* pa;
    const float* pb;
    const float* va[3];
    const float* vb[3];
};
// Elastic imaging of the checkpointed gradient (strips.hip).  Per forward step n: the
// adjoint stresses before the adjoint stress pass (tau~^{n+1}), the adjoint velocities after
// the adjoint velocity pass (v~^{n+1/2}) and the forward states around the step (tau^n,
// tau^{n+1}, v^{n-1/2}, v^{n+1/2}).  Accumulated in FP64 on the slab layout:
//   T = sum (tr tau~)(tr dtau) and P = sum tau~ . dtau over the normal stresses,
//   S[s] = sum tau~_s dtau_s per shear stress, R[a] = sum v~_a dv_a per velocity.
struct ElasticImage {
    int nn = 0, ns = 0, nv = 0;                     // normal, shear, velocity components
    const float* an[3] = {}; const float* f0n[3] = {}; const float* f1n[3] = {};
    const float* as[3] = {}; const float* f0s[3] = {}; const float* f1s[3] = {};
    const float* av[3] = {}; const float* f0v[3] = {}; const float* f1v[3] = {};
    int shi[3][3] = {};                             // shear s occupies [0, shi) in (z, y, x)
    int sbase[3][3] = {};                           // its invmu average: cells p + sbase + {0, e1, e2, e1 + e2}
    int se1[3] = {}, se2[3] = {};                   // axes (0 z, 1 y, 2 x) of e1, e2
    int vlo[3][3] = {}, vhi[3][3] = {};             // velocity a occupies [vlo, vhi)
    int vaxis[3] = {};                              // its density pair: cells v - h e_a and v - h e_a + e_a
    double* T = nullptr;
    double* P = nullptr;
    double* S[3] = {};
    double* R[3] = {};
};
void image_elastic(const Model& m, const ElasticImage& E, hipStream_t s);
// gradients w.r.t. invlambda, invmu and rho on the stress grid from the accumulators
void elastic_gradient(const Model& m, const ElasticImage& E, double* g_il, double* g_im, double* g_rho,
                      hipStream_t s);
// Normal-stress sources inject dtM * a_n into every normal stress (src/fdtd/source.jl:79-119):
// per target cell, acc += (sum of the adjoint normal stresses) * a_n at forward step n, and at
// the end the part of the invlambda/invmu gradient that the Delta-tau form leaves out there.
void source_trace(const DevPoints& pts, int64_t n, const float* const an[3], int nn, double* acc, hipStream_t s);
void source_correction(const Model& m, const DevPoints& pts, const double* acc, int nn, double* g_il, double* g_im,
                       hipStream_t s);
void image_acoustic_exact(const Model& m, const ImageSets& S, double* g_invK, double* g_rho, float dtI, hipStream_t s);

// Born secondary sources (born.hip): scattered state w2 from background state w1 after the
// velocity passes (velocity) or after the stress passes; dc: d(dtinvav{x,y,z}irho), dk: d(dtK)
void born_sources(const Model& m, const State& w1, const State& w2, const float* const dc[3], const float* dk,
                  bool velocity, hipStream_t s);

// CPML strides and vectors of derivative f with the memory array of state st
inline Pml pml_of(const Model& m, const State& st, int f)
{
    Pml q = m.pml[f];
    q.psi = st.psi[f];
    return q;
}

// every active derivative of a CPML model has its memory array
inline bool state_complete(const Model& m, const Pml* const* q, int n)
{
    for (int k = 0; k < n; ++k)
        if (q[k]->active && m.g.pml_mask && !q[k]->psi) return false;
    return true;
}

}  // namespace gpi
