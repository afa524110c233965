// Host-side helpers shared by the C-ABI translation units: errors, field tables,
// allocation accounting, pitched copies and per-call scratch buffers.
#pragma once
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "gpi_internal.h"
#include "pass_extras.h"

struct gpi_ctx : public gpi::Ctx {};

namespace gpi {

inline int fail(gpi_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIPTRY(ctx, expr)                                                                  \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return ::gpi::fail(ctx, GPI_EHIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

// Per-axis grid code of each field (src/fields.jl:92-671): N = n points, A = n+h,
// H = n-h, I = n-2h.  '.' marks the absent y axis of 2-D grids.
const char* field_codes(int ndim, int physics, int f);
int shape_of(int ndim, int order, int physics, const int64_t n[3], int f, int64_t out[3]);
// axis along which a derivative field differentiates (0 z, 1 y, 2 x); -1 for wavefields
int deriv_axis(int f);
// grid (field id) of each coefficient array (src/fdtd/medium.jl:81-95); -1 if unused
int coef_grid(int ndim, int physics, int c);
bool required_coef(int ndim, int physics, int c);

inline bool is_wavefield(int f) { return f >= GPI_P && f <= GPI_TAUYZ; }
inline bool is_velocity(int f) { return f == GPI_VX || f == GPI_VY || f == GPI_VZ; }
inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

Model* find_model(gpi_ctx* c, uint64_t h);
int dev_alloc(gpi_ctx* c, std::vector<void*>* owner, uint64_t* owner_bytes, void** p, size_t bytes);
void free_allocs(gpi_ctx* c, std::vector<void*>& allocs, uint64_t& bytes);
void free_model(gpi_ctx* c, Model* m);

// device array of coefficient `id` (one pitched slab; the 3-D acoustic density
// coefficients start vel3d_row_shift floats into their allocation, fdtd3d.h)
int alloc_coef(gpi_ctx* c, Model& m, int id);

// wavefields + CPML memory of one propagating wavefield of model m
int alloc_state(gpi_ctx* c, const Model& m, State* st);
void free_state(gpi_ctx* c, State* st);
int zero_state(gpi_ctx* c, const Model& m, const State& st, hipStream_t s);
size_t psi_floats(const Model& m, int f);

// padded offset of compact index `lin` of an array of logical size n[3]
inline int64_t padded(const Model& m, const int64_t n[3], int64_t lin)
{
    const int64_t iz = lin % n[0];
    const int64_t r = lin / n[0];
    const int64_t iy = r % n[1];
    const int64_t ix = r / n[1];
    return iz + m.pitch0 * (iy + m.pitch1 * ix);
}

struct Scratch;

// Source sets of one wavefield: fused into the passes' PassExtras when every set of a pass
// can ride along, else applied by the stand-alone kernel after the pass (all-or-nothing per
// pass keeps the reference's order of updates).  gpi_shot.hip.
struct Sources {
    std::vector<DevPoints> pts;
    std::vector<int> vfall, sfall;   // sets injected after the velocity / stress pass
};
// Forward sources kept by a checkpoint store: replayed, without receivers, when the
// adjoint run recomputes a segment of the forward wavefield.
struct Replay {
    Scratch* S = nullptr;            // owns the device tables
    PassExtras xv{}, xs{};
    Sources src;
};
// free a store and everything it owns (strips, checkpoints, replayed sources)
void free_store(gpi_ctx* c, Store* s);
// floats of one checkpoint: every wavefield and CPML memory array of a state
size_t state_floats(const Model& m);
// copy a whole state into (to_ckpt) or out of a checkpoint buffer, on stream s
int copy_state(gpi_ctx* c, const Model& m, const State& st, float* ckpt, bool to_ckpt, hipStream_t s);

int plan_sources(gpi_ctx* c, const Model& m, Scratch& S, int nsrc, const gpi_points* src, const float* const* wav,
                 int64_t nt, bool velocity_only, PassExtras& XV, PassExtras& XS, Sources& out);
int plan_receivers(gpi_ctx* c, const Model& m, const State& st, Scratch& S, const gpi_shot* shot, PassExtras& XV,
                   PassExtras& XS, std::vector<DevPoints>& rec);
// the body of gpi_forward: one supersource on a resident model; every scratch buffer is
// taken before the first step (from the context's arena when it has one)
int forward_shot(gpi_ctx* c, uint64_t model, const gpi_shot* shot);

// model construction pieces (gpi_api.hip) and the device coefficient preparation
// (medium.hip), shared by the C ABI and the serving scheduler
int model_begin(gpi_ctx* c, const gpi_grid* g, Model** out);
int model_pml(gpi_ctx* c, Model* m, const gpi_pml* pml, int32_t npml_arrays);
int model_state(gpi_ctx* c, Model* m);
int create_model_medium(gpi_ctx* c, const gpi_grid* g, const gpi_medium* med, const gpi_pml* pml,
                        int32_t npml_arrays, Model** out);
int prep_coefficients_2d(gpi_ctx* c, Model& m, const float* vp, const float* vs, const float* rho,
                         const int64_t nm[3], hipStream_t s);

// host -> device copy: blocking, or queued on c->stream through the context's pinned
// staging ring when it has one (the source may be reused as soon as this returns)
int copy_h2d(gpi_ctx* c, void* dst, const void* src, size_t bytes);

int upload_pitched(gpi_ctx* c, const Model& m, float* dst, const float* src, const int64_t n[3]);
// pitched device slab -> compact host array of logical size n[3] (blocking); packed through
// `stage` (device, n[0]*n[1]*n[2] elements) when given, else through a scratch buffer
int download_pitched(gpi_ctx* c, const Model& m, float* dst, const float* src, const int64_t n[3],
                     void* stage = nullptr);
int download_pitched(gpi_ctx* c, const Model& m, double* dst, const double* src, const int64_t n[3],
                     void* stage = nullptr);

// carve `bytes` (256-byte aligned) out of the context's arena; GPI_ENOMEM when it is full
int arena_alloc(gpi_ctx* c, void** p, size_t bytes);

// Scratch device buffers of one call, freed on every exit path (after the stream drained).
// On a context with an arena the buffers are carved from it and nothing is freed.
struct Scratch {
    gpi_ctx* c;
    std::vector<void*> p;
    uint64_t bytes = 0;
    explicit Scratch(gpi_ctx* cc) : c(cc) {}
    ~Scratch()
    {
        if (p.empty()) return;
        if (!c->defer_free) (void)hipStreamSynchronize(c->stream);
        free_allocs(c, p, bytes);
    }
    template <typename T>
    int alloc(T** out, size_t count)
    {
        void* q = nullptr;
        const size_t b = std::max<size_t>(sizeof(T) * count, 4);
        int rc = c->arena ? arena_alloc(c, &q, b) : dev_alloc(c, &p, &bytes, &q, b);
        if (rc) return rc;
        *out = (T*)q;
        return GPI_OK;
    }
    template <typename T>
    int put(T** out, const std::vector<T>& v)
    {
        int rc = alloc(out, v.size());
        if (rc) return rc;
        if (!v.empty()) return copy_h2d(c, *out, v.data(), sizeof(T) * v.size());
        return GPI_OK;
    }
};

}  // namespace gpi
