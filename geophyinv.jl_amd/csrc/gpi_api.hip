// C-ABI of libgpi: context, resident models and the forward time loop.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>

#include "fdtd2d.h"
#include "fdtd3d.h"
#include "gpi_internal.h"
#include "pass_extras.h"

using namespace gpi;

struct gpi_ctx : public Ctx {};

namespace {

int fail(gpi_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIPTRY(ctx, expr)                                                              \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(ctx, GPI_EHIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

// Per-axis grid code of each field (src/fields.jl:92-671): N = n points, A = n+h,
// H = n-h, I = n-2h.  '.' marks the absent y axis of 2-D grids.
const char* field_codes(int ndim, int physics, int f)
{
    static const char* ac2[GPI_NFIELDS] = {};
    static const char* ac3[GPI_NFIELDS] = {};
    static const char* el2[GPI_NFIELDS] = {};
    static const char* el3[GPI_NFIELDS] = {};
    static bool init = false;
    if (!init) {
        ac2[GPI_P] = "N.N"; ac2[GPI_VX] = "N.A"; ac2[GPI_VZ] = "A.N";
        ac2[GPI_DPDX] = "I.H"; ac2[GPI_DPDZ] = "H.I"; ac2[GPI_DVXDX] = "N.N"; ac2[GPI_DVZDZ] = "N.N";
        ac3[GPI_P] = "NNN"; ac3[GPI_VX] = "NNA"; ac3[GPI_VY] = "NAN"; ac3[GPI_VZ] = "ANN";
        ac3[GPI_DPDX] = "IIH"; ac3[GPI_DPDY] = "IHI"; ac3[GPI_DPDZ] = "HII";
        ac3[GPI_DVXDX] = "NNN"; ac3[GPI_DVYDY] = "NNN"; ac3[GPI_DVZDZ] = "NNN";
        el2[GPI_TAUXX] = "N.N"; el2[GPI_TAUZZ] = "N.N"; el2[GPI_TAUXZ] = "H.H";
        el2[GPI_VX] = "N.A"; el2[GPI_VZ] = "A.N";
        el2[GPI_DTAUXXDX] = "I.H"; el2[GPI_DTAUZZDZ] = "H.I"; el2[GPI_DTAUXZDX] = "H.I";
        el2[GPI_DTAUXZDZ] = "I.H"; el2[GPI_DVXDX] = "N.N"; el2[GPI_DVZDZ] = "N.N";
        el2[GPI_DVXDZ] = "H.H"; el2[GPI_DVZDX] = "H.H";
        el3[GPI_TAUXX] = "NNN"; el3[GPI_TAUYY] = "NNN"; el3[GPI_TAUZZ] = "NNN";
        el3[GPI_TAUXY] = "IHH"; el3[GPI_TAUXZ] = "HIH"; el3[GPI_TAUYZ] = "HHI";
        el3[GPI_VX] = "NNA"; el3[GPI_VY] = "NAN"; el3[GPI_VZ] = "ANN";
        el3[GPI_DTAUXXDX] = "IIH"; el3[GPI_DTAUYYDY] = "IHI"; el3[GPI_DTAUZZDZ] = "HII";
        el3[GPI_DTAUXYDX] = "IHI"; el3[GPI_DTAUXYDY] = "IIH"; el3[GPI_DTAUXZDX] = "HII";
        el3[GPI_DTAUXZDZ] = "IIH"; el3[GPI_DTAUYZDY] = "HII"; el3[GPI_DTAUYZDZ] = "IHI";
        el3[GPI_DVXDX] = "NNN"; el3[GPI_DVYDY] = "NNN"; el3[GPI_DVZDZ] = "NNN";
        el3[GPI_DVXDY] = "IHH"; el3[GPI_DVXDZ] = "HIH"; el3[GPI_DVYDX] = "IHH";
        el3[GPI_DVYDZ] = "HHI"; el3[GPI_DVZDX] = "HIH"; el3[GPI_DVZDY] = "HHI";
        init = true;
    }
    if (f < 0 || f >= GPI_NFIELDS) return nullptr;
    if (physics == 0) return ndim == 2 ? ac2[f] : ac3[f];
    return ndim == 2 ? el2[f] : el3[f];
}

int shape_of(int ndim, int order, int physics, const int64_t n[3], int f, int64_t out[3])
{
    const char* c = field_codes(ndim, physics, f);
    if (!c) return GPI_EINVAL;
    const int h = order - 1;
    for (int a = 0; a < 3; ++a) {
        switch (c[a]) {
        case 'N': out[a] = n[a]; break;
        case 'A': out[a] = n[a] + h; break;
        case 'H': out[a] = n[a] - h; break;
        case 'I': out[a] = n[a] - 2 * h; break;
        default: out[a] = 1; break;
        }
    }
    return GPI_OK;
}

// axis along which a derivative field differentiates (0 z, 1 y, 2 x); -1 for wavefields
int deriv_axis(int f)
{
    switch (f) {
    case GPI_DPDZ: case GPI_DVZDZ: case GPI_DVXDZ: case GPI_DVYDZ: case GPI_DTAUZZDZ:
    case GPI_DTAUXZDZ: case GPI_DTAUYZDZ:
        return 0;
    case GPI_DPDY: case GPI_DVYDY: case GPI_DVXDY: case GPI_DVZDY: case GPI_DTAUYYDY:
    case GPI_DTAUXYDY: case GPI_DTAUYZDY:
        return 1;
    case GPI_DPDX: case GPI_DVXDX: case GPI_DVYDX: case GPI_DVZDX: case GPI_DTAUXXDX:
    case GPI_DTAUXYDX: case GPI_DTAUXZDX:
        return 2;
    }
    return -1;
}

// grid (field id) of each coefficient array (src/fdtd/medium.jl:81-95)
int coef_grid(int ndim, int physics, int c)
{
    const bool ac = physics == 0;
    switch (c) {
    case GPI_C_DTINVAVXIRHO: return ac ? GPI_DPDX : GPI_DTAUXXDX;
    case GPI_C_DTINVAVYIRHO: return ndim == 3 ? (ac ? GPI_DPDY : GPI_DTAUYYDY) : -1;
    case GPI_C_DTINVAVZIRHO: return ac ? GPI_DPDZ : GPI_DTAUZZDZ;
    case GPI_C_DTK: return ac ? GPI_P : -1;
    case GPI_C_DTM: case GPI_C_DTLAMBDA: return ac ? -1 : GPI_TAUXX;
    case GPI_C_DTAVMU: return (!ac && ndim == 2) ? GPI_DVXDZ : -1;
    case GPI_C_DTAVXZIMU: return (!ac && ndim == 3) ? GPI_DVXDZ : -1;
    case GPI_C_DTAVXYIMU: return (!ac && ndim == 3) ? GPI_DVXDY : -1;
    case GPI_C_DTAVYZIMU: return (!ac && ndim == 3) ? GPI_DVYDZ : -1;
    case GPI_C_RHO: return ac ? GPI_P : GPI_TAUXX;
    case GPI_C_INVK: return ac ? GPI_P : -1;
    case GPI_C_INVLAMBDA: case GPI_C_INVMU: return ac ? -1 : GPI_TAUXX;
    }
    return -1;
}

bool required_coef(int ndim, int physics, int c)
{
    if (physics == 0)
        return c == GPI_C_DTINVAVXIRHO || c == GPI_C_DTINVAVZIRHO || c == GPI_C_DTK ||
               c == GPI_C_RHO || (ndim == 3 && c == GPI_C_DTINVAVYIRHO);
    if (c == GPI_C_DTINVAVXIRHO || c == GPI_C_DTINVAVZIRHO || c == GPI_C_DTM ||
        c == GPI_C_DTLAMBDA || c == GPI_C_RHO)
        return true;
    if (ndim == 2) return c == GPI_C_DTAVMU;
    return c == GPI_C_DTINVAVYIRHO || c == GPI_C_DTAVXZIMU || c == GPI_C_DTAVXYIMU ||
           c == GPI_C_DTAVYZIMU;
}

bool is_wavefield(int f) { return f >= GPI_P && f <= GPI_TAUYZ; }

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

Model* find_model(gpi_ctx* c, uint64_t h)
{
    for (Model* m : c->models)
        if ((uint64_t)(uintptr_t)m == h) return m;
    return nullptr;
}

int dev_alloc(gpi_ctx* c, Model* m, void** p, size_t bytes)
{
    if (c->budget && c->used + bytes > c->budget)
        return fail(c, GPI_ENOMEM, "allocation of " + std::to_string(bytes) +
                                       " bytes exceeds the context's HBM budget");
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(c, GPI_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->used += bytes;
    if (m) {
        m->allocs.push_back(*p);
        m->bytes += bytes;
    }
    return GPI_OK;
}

void free_model(gpi_ctx* c, Model* m)
{
    for (void* p : m->allocs) (void)hipFree(p);
    c->used -= std::min<uint64_t>(c->used, m->bytes);
    delete m;
}

// padded offset of compact index `lin` of an array of logical size n[3]
inline int64_t padded(const Model& m, const int64_t n[3], int64_t lin)
{
    const int64_t iz = lin % n[0];
    const int64_t r = lin / n[0];
    const int64_t iy = r % n[1];
    const int64_t ix = r / n[1];
    return iz + m.pitch0 * (iy + m.pitch1 * ix);
}

// copy a compact host array into the model's pitched layout (zero elsewhere)
int upload_pitched(gpi_ctx* c, const Model& m, float* dst, const float* src, const int64_t n[3])
{
    std::vector<float> stage((size_t)m.slab, 0.0f);
    for (int64_t ix = 0; ix < n[2]; ++ix)
        for (int64_t iy = 0; iy < n[1]; ++iy)
            std::memcpy(&stage[(size_t)(m.pitch0 * (iy + m.pitch1 * ix))],
                        src + n[0] * (iy + n[1] * ix), sizeof(float) * (size_t)n[0]);
    HIPTRY(c, hipMemcpy(dst, stage.data(), sizeof(float) * (size_t)m.slab, hipMemcpyHostToDevice));
    return GPI_OK;
}

int download_pitched(gpi_ctx* c, const Model& m, float* dst, const float* src, const int64_t n[3])
{
    std::vector<float> stage((size_t)m.slab);
    HIPTRY(c, hipMemcpy(stage.data(), src, sizeof(float) * (size_t)m.slab, hipMemcpyDeviceToHost));
    for (int64_t ix = 0; ix < n[2]; ++ix)
        for (int64_t iy = 0; iy < n[1]; ++iy)
            std::memcpy(dst + n[0] * (iy + n[1] * ix),
                        &stage[(size_t)(m.pitch0 * (iy + m.pitch1 * ix))], sizeof(float) * (size_t)n[0]);
    return GPI_OK;
}

// Scratch device buffers of one forward call, freed on every exit path.
struct Scratch {
    gpi_ctx* c;
    std::vector<void*> p;
    uint64_t bytes = 0;
    explicit Scratch(gpi_ctx* cc) : c(cc) {}
    ~Scratch()
    {
        for (void* q : p) (void)hipFree(q);
        c->used -= std::min<uint64_t>(c->used, bytes);
    }
    template <typename T>
    int alloc(T** out, size_t count)
    {
        void* q = nullptr;
        size_t b = std::max<size_t>(sizeof(T) * count, 4);
        int rc = dev_alloc(c, nullptr, &q, b);
        if (rc) return rc;
        p.push_back(q);
        bytes += b;
        *out = (T*)q;
        return GPI_OK;
    }
    template <typename T>
    int put(T** out, const std::vector<T>& v)
    {
        int rc = alloc(out, v.size());
        if (rc) return rc;
        if (!v.empty()) {
            hipError_t e = hipMemcpy(*out, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
            if (e != hipSuccess) return fail(c, GPI_EHIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
        }
        return GPI_OK;
    }
};

bool recordable(int physics, int f)
{
    // propagate.jl:177 records [:p] and :208 records [:vx,:vy,:vz]; other fields of
    // rfields keep zero records, as in the reference.
    (void)physics;
    return f == GPI_P || f == GPI_VX || f == GPI_VY || f == GPI_VZ;
}

}  // namespace

extern "C" {

int gpi_abi_version(void) { return GPI_ABI_VERSION; }

int gpi_field_shape(int32_t ndim, int32_t order, int32_t physics, const int64_t n[3],
                    int32_t field, int64_t out[3])
{
    if ((ndim != 2 && ndim != 3) || (order != 2 && order != 4 && order != 6 && order != 8) ||
        (physics != 0 && physics != 1) || !n || !out)
        return GPI_EINVAL;
    int64_t nn[3] = {n[0], ndim == 2 ? 1 : n[1], n[2]};
    return shape_of(ndim, order, physics, nn, field, out);
}

int gpi_create(gpi_ctx** out, int device, uint64_t hbm_budget_bytes)
{
    if (!out) return GPI_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return GPI_EHIP;
    if (device < 0 || device >= ndev) return GPI_EINVAL;
    gpi_ctx* c = new (std::nothrow) gpi_ctx();
    if (!c) return GPI_ENOMEM;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return GPI_EHIP;
    }
    for (auto& e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete c;
            return GPI_EHIP;
        }
    size_t fr = 0, tot = 0;
    (void)hipMemGetInfo(&fr, &tot);
    c->budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t)fr;
    *out = c;
    return GPI_OK;
}

void gpi_destroy(gpi_ctx* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (Model* m : c->models) free_model(c, m);
    for (auto& e : c->ev) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->sample_ev) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* gpi_last_error(const gpi_ctx* c) { return c ? c->err.c_str() : "null context"; }

int gpi_model_create(gpi_ctx* c, const gpi_grid* g, const gpi_array* coefs, int32_t ncoefs,
                     const gpi_pml* pml, int32_t npml_arrays, uint64_t* model_out)
{
    if (!c || !g || !model_out) return fail(c, GPI_EINVAL, "null argument");
    if (g->ndim != 2 && g->ndim != 3) return fail(c, GPI_EINVAL, "ndim must be 2 or 3");
    if (g->order != 2 && g->order != 4 && g->order != 6 && g->order != 8)
        return fail(c, GPI_EINVAL, "order must be 2, 4, 6 or 8");
    if (g->physics != 0 && g->physics != 1) return fail(c, GPI_EINVAL, "physics must be 0 or 1");
    const int h = g->order - 1;
    const int64_t ny = g->ndim == 2 ? 1 : g->n[1];
    for (int a = 0; a < 3; ++a) {
        if (g->ndim == 2 && a == 1) continue;
        if (g->n[a] < 2 * h + 2) return fail(c, GPI_EINVAL, "grid too small for the stencil order");
        for (int side = 0; side < 2; ++side)
            if ((g->pml_mask & GPI_FACE(a, side)) && g->n[a] - 2 * h < g->npml)
                return fail(c, GPI_EINVAL, "grid shorter than the CPML strip");
    }
    if (g->pml_mask && g->npml < 1) return fail(c, GPI_EINVAL, "npml must be positive");
    if (!(g->dt > 0.0f) || !std::isfinite(g->dt)) return fail(c, GPI_EINVAL, "dt must be positive");
    HIPTRY(c, hipSetDevice(c->device));

    Model* m = new Model();
    m->g = *g;
    m->g.n[1] = ny;
    m->h = h;
    m->pitch0 = round_up(g->n[0] + h, 32);
    m->pitch1 = g->ndim == 2 ? 1 : ny + h;
    m->slab = m->pitch0 * m->pitch1 * (g->n[2] + h);
    if (m->slab >= (int64_t)1 << 31) {
        delete m;
        return fail(c, GPI_EINVAL, "grid exceeds 2^31 points per array");
    }
    auto bail = [&](int rc) {
        free_model(c, m);
        return rc;
    };
    // coefficient arrays
    std::vector<bool> have(GPI_NCOEF, false);
    for (int q = 0; q < ncoefs; ++q) {
        const gpi_array& A = coefs[q];
        if (A.id < 0 || A.id >= GPI_NCOEF) return bail(fail(c, GPI_EINVAL, "bad coefficient id"));
        int grid = coef_grid(g->ndim, g->physics, A.id);
        if (grid < 0) continue;   // not used by this physics
        int64_t n3[3];
        shape_of(g->ndim, g->order, g->physics, m->g.n, grid, n3);
        if (A.count != n3[0] * n3[1] * n3[2] || !A.data)
            return bail(fail(c, GPI_EINVAL, "coefficient " + std::to_string(A.id) + " has the wrong size"));
        if (!m->coef[A.id]) {
            int rc = dev_alloc(c, m, (void**)&m->coef[A.id], sizeof(float) * m->slab);
            if (rc) return bail(rc);
        }
        int rc = upload_pitched(c, *m, m->coef[A.id], A.data, n3);
        if (rc) return bail(rc);
        have[A.id] = true;
    }
    for (int q = 0; q < GPI_NCOEF; ++q)
        if (required_coef(g->ndim, g->physics, q) && !have[q])
            return bail(fail(c, GPI_EINVAL, "missing coefficient " + std::to_string(q)));
    // wavefields
    for (int f = GPI_P; f <= GPI_TAUYZ; ++f) {
        if (!field_codes(g->ndim, g->physics, f)) continue;
        int rc = dev_alloc(c, m, (void**)&m->field[f], sizeof(float) * m->slab);
        if (rc) return bail(rc);
    }
    // CPML memory and coefficient vectors of every derivative field
    const int L = 2 * g->npml;
    std::vector<float> vec;
    std::vector<int> order_f;
    for (int f = GPI_DPDX; f < GPI_NFIELDS; ++f) {
        if (!field_codes(g->ndim, g->physics, f)) continue;
        const gpi_pml* src = nullptr;
        for (int q = 0; q < npml_arrays; ++q)
            if (pml[q].field == f) src = &pml[q];
        if (g->pml_mask && !src)
            return bail(fail(c, GPI_EINVAL, "missing CPML vectors for field " + std::to_string(f)));
        if (src && (src->len != L || !src->a || !src->b || !src->kI))
            return bail(fail(c, GPI_EINVAL, "CPML vectors must have length 2*npml"));
        int64_t n3[3];
        shape_of(g->ndim, g->order, g->physics, m->g.n, f, n3);
        const int ax = deriv_axis(f);
        Pml& P = m->pml[f];
        P.active = 1;
        P.nd = (int32_t)n3[ax];
        int64_t rows, cols;
        if (ax == 2) {         // (nz_d, [ny_d,] 2npml)
            P.ps1 = m->pitch0;
            P.ps2 = m->pitch0 * m->pitch1;
            rows = P.ps2;
            cols = L;
        } else if (ax == 0) {  // (2npml, [ny_d,] nx_d)
            const int64_t pz = round_up(std::max(L, 1), 32);
            P.ps1 = pz;
            P.ps2 = pz * n3[1];
            rows = P.ps2;
            cols = n3[2];
        } else {               // (nz_d, 2npml, nx_d)
            P.ps1 = m->pitch0;
            P.ps2 = m->pitch0 * L;
            rows = P.ps2;
            cols = n3[2];
        }
        if (g->pml_mask) {
            int rc = dev_alloc(c, m, (void**)&P.psi, sizeof(float) * (size_t)std::max<int64_t>(rows * cols, 1));
            if (rc) return bail(rc);
            order_f.push_back(f);
            vec.insert(vec.end(), src->a, src->a + L);
            vec.insert(vec.end(), src->b, src->b + L);
            vec.insert(vec.end(), src->kI, src->kI + L);
        }
    }
    if (!vec.empty()) {
        int rc = dev_alloc(c, m, (void**)&m->pmlvec, sizeof(float) * vec.size());
        if (rc) return bail(rc);
        hipError_t e = hipMemcpy(m->pmlvec, vec.data(), sizeof(float) * vec.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) return bail(fail(c, GPI_EHIP, hipGetErrorString(e)));
        for (size_t q = 0; q < order_f.size(); ++q) {
            Pml& P = m->pml[order_f[q]];
            P.a = m->pmlvec + 3 * L * q;
            P.b = P.a + L;
            P.kI = P.b + L;
        }
    }
    c->models.push_back(m);
    *model_out = (uint64_t)(uintptr_t)m;
    return GPI_OK;
}

int gpi_model_update(gpi_ctx* c, uint64_t model, const gpi_array* coefs, int32_t ncoefs)
{
    if (!c) return GPI_EINVAL;
    Model* m = find_model(c, model);
    if (!m) return fail(c, GPI_ENOTFOUND, "unknown model handle");
    HIPTRY(c, hipSetDevice(c->device));
    for (int q = 0; q < ncoefs; ++q) {
        const gpi_array& A = coefs[q];
        if (A.id < 0 || A.id >= GPI_NCOEF) return fail(c, GPI_EINVAL, "bad coefficient id");
        int grid = coef_grid(m->g.ndim, m->g.physics, A.id);
        if (grid < 0) continue;
        int64_t n3[3];
        shape_of(m->g.ndim, m->g.order, m->g.physics, m->g.n, grid, n3);
        if (A.count != n3[0] * n3[1] * n3[2] || !A.data)
            return fail(c, GPI_EINVAL, "coefficient has the wrong size");
        if (!m->coef[A.id]) {
            int rc = dev_alloc(c, m, (void**)&m->coef[A.id], sizeof(float) * m->slab);
            if (rc) return rc;
        }
        int rc = upload_pitched(c, *m, m->coef[A.id], A.data, n3);
        if (rc) return rc;
    }
    return GPI_OK;
}

int gpi_model_free(gpi_ctx* c, uint64_t model)
{
    if (!c) return GPI_EINVAL;
    Model* m = find_model(c, model);
    if (!m) return fail(c, GPI_ENOTFOUND, "unknown model handle");
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->models.erase(std::find(c->models.begin(), c->models.end(), m));
    free_model(c, m);
    return GPI_OK;
}

int gpi_forward(gpi_ctx* c, uint64_t model, const gpi_shot* shot)
{
    if (!c || !shot) return fail(c, GPI_EINVAL, "null argument");
    Model* mp = find_model(c, model);
    if (!mp) return fail(c, GPI_ENOTFOUND, "unknown model handle");
    Model& m = *mp;
    const gpi_grid& g = m.g;
    const int64_t nt = shot->nt;
    if (nt < 1) return fail(c, GPI_EINVAL, "nt must be positive");
    if (shot->nsrc < 0 || shot->nrec < 0 || shot->nsrc > 8 || shot->nrec > 16)
        return fail(c, GPI_EINVAL, "bad number of source/receiver fields");
    HIPTRY(c, hipSetDevice(c->device));
    Scratch S(c);
    const int nnz_expect = g.ndim == 2 ? 4 : 8;

    // ---- sources: unique targets, grouped by the workgroup of the pass that owns them ----
    int vgx, vgy, sgx, sgy;
    const bool d3 = g.ndim == 3;
    const int vnb = d3 ? velocity_blocks_3d(m, &vgx, &vgy) : velocity_blocks_2d(m, &vgx, &vgy);
    const int snb = d3 ? stress_blocks_3d(m, &sgx, &sgy) : stress_blocks_2d(m, &sgx, &sgy);
    std::vector<DevPoints> src(shot->nsrc);
    std::vector<int> kinds(shot->nsrc);
    bool vel_fused = true, str_fused = true;
    for (int q = 0; q < shot->nsrc; ++q) {
        const gpi_points& P = shot->src[q];
        const int f = P.field;
        if (!field_codes(g.ndim, g.physics, f) || !is_wavefield(f))
            return fail(c, GPI_EINVAL, "source field " + std::to_string(f) + " is not a wavefield of this physics");
        const bool vel = f == GPI_VX || f == GPI_VY || f == GPI_VZ;
        const bool str = f == GPI_P || f == GPI_TAUXX || f == GPI_TAUYY || f == GPI_TAUZZ;
        if (!vel && !str) return fail(c, GPI_EINVAL, "shear stresses cannot carry sources");
        if (P.nnz != nnz_expect || P.npts < 0 || (P.npts > 0 && (!P.idx || !P.w || !shot->wavelets || !shot->wavelets[q])))
            return fail(c, GPI_EINVAL, "bad source point set");
        kinds[q] = f == GPI_VX ? INJ_VX : f == GPI_VY ? INJ_VY : f == GPI_VZ ? INJ_VZ
                 : g.physics == 0 ? INJ_P : (g.ndim == 3 ? INJ_NORMAL3 : INJ_NORMAL2);
        int64_t n3[3];
        shape_of(g.ndim, g.order, g.physics, g.n, f, n3);
        const int64_t N = n3[0] * n3[1] * n3[2];
        const int h = m.h;
        // target -> contributions in ascending source order
        std::map<int64_t, std::vector<std::pair<int, float>>> tg;
        for (int s = 0; s < P.npts; ++s)
            for (int k = 0; k < P.nnz; ++k) {
                int64_t lin = P.idx[(size_t)s * P.nnz + k];
                if (lin < 0 || lin >= N) return fail(c, GPI_EINVAL, "source index out of range");
                tg[lin].push_back({s, P.w[(size_t)s * P.nnz + k]});
            }
        struct Tgt {
            int64_t lin;
            int blk;
            const std::vector<std::pair<int, float>>* con;
        };
        std::vector<Tgt> tl;
        for (auto& kv : tg) {
            const int64_t lin = kv.first;
            const int64_t iz = lin % n3[0], r = lin / n3[0], iy = r % n3[1], ix = r / n3[1];
            int blk;
            if (vel) {
                // muladd_with_density_v*! touches @inn(v) only (source.jl:166-177)
                if (!(iz >= h && iz < n3[0] - h && ix >= h && ix < n3[2] - h)) continue;
                if (d3 && !(iy >= h && iy < n3[1] - h)) continue;
                blk = d3 ? (int)((iz - h) / 64 + vgx * ((iy - h) / 4 + vgy * (ix - h)))
                         : (int)((iz - h) / 64 + vgx * ((ix - h) / 4));
            } else {
                blk = d3 ? (int)(iz / 64 + sgx * (iy / 4 + sgy * ix)) : (int)(iz / 64 + sgx * (ix / 4));
                // a source on the free-surface row is applied after the surface condition
                // by a separate kernel (source.jl:61-119 runs after update_stress!)
                if ((g.free_mask & GPI_FACE(0, 0)) && iz == 0) str_fused = false;
            }
            tl.push_back({lin, blk, &kv.second});
        }
        std::stable_sort(tl.begin(), tl.end(), [](const Tgt& x, const Tgt& y) { return x.blk < y.blk; });
        const int nb = vel ? vnb : snb;
        std::vector<int32_t> blk(nb + 1, 0);
        for (const Tgt& t : tl) blk[t.blk + 1]++;
        int maxper = 0;
        for (int k = 0; k < nb; ++k) {
            maxper = std::max(maxper, blk[k + 1]);
            blk[k + 1] += blk[k];
        }
        if (maxper > 32) (vel ? vel_fused : str_fused) = false;   // keep owner scans short
        DevPoints& D = src[q];
        D.field = f;
        D.npts = P.npts;
        D.nnz = P.nnz;
        std::vector<int64_t> toff, roff;
        std::vector<int32_t> tptr{0}, tsrc;
        std::vector<float> tw;
        for (const Tgt& t : tl) {
            if (vel) {
                // @av_{x,y,z}i(rho) at the interior point: rho indices of the two samples
                const int64_t iz = t.lin % n3[0], r = t.lin / n3[0], iy = r % n3[1], ix = r / n3[1];
                int64_t b0[3] = {iz, iy, ix};
                const int ax = f == GPI_VZ ? 0 : f == GPI_VY ? 1 : 2;
                b0[ax] -= h;
                int64_t b1[3] = {b0[0], b0[1], b0[2]};
                b1[ax] += 1;
                roff.push_back(b0[0] + m.pitch0 * (b0[1] + m.pitch1 * b0[2]));
                roff.push_back(b1[0] + m.pitch0 * (b1[1] + m.pitch1 * b1[2]));
            }
            toff.push_back(padded(m, n3, t.lin));
            for (auto& pr : *t.con) {
                tsrc.push_back(pr.first);
                tw.push_back(pr.second);
            }
            tptr.push_back((int32_t)tsrc.size());
        }
        D.ntarget = (int32_t)toff.size();
        toff.insert(toff.end(), roff.begin(), roff.end());
        int rc;
        if ((rc = S.put(&D.toff, toff)) || (rc = S.put(&D.tptr, tptr)) || (rc = S.put(&D.tsrc, tsrc)) ||
            (rc = S.put(&D.tw, tw)) || (rc = S.put(&D.blk, blk)))
            return rc;
        // wavelets: host nt x npts time fastest -> device point fastest per step
        std::vector<float> wv((size_t)nt * std::max(P.npts, 1), 0.0f);
        for (int s = 0; s < P.npts; ++s)
            for (int64_t it = 0; it < nt; ++it) wv[(size_t)it * P.npts + s] = shot->wavelets[q][(size_t)s * nt + it];
        if ((rc = S.put(&D.wav, wv))) return rc;
    }

    // ---- receivers: gather form ---------------------------------------------------------
    std::vector<DevPoints> rec(shot->nrec);
    PassExtras XV{}, XS{}, X0{};
    int vmax = 0, smax = 0;
    for (int q = 0; q < shot->nrec; ++q) {
        const gpi_points& P = shot->rec[q];
        const int f = P.field;
        if (!field_codes(g.ndim, g.physics, f) || !is_wavefield(f))
            return fail(c, GPI_EINVAL, "receiver field " + std::to_string(f) + " is not a wavefield of this physics");
        if (P.nnz != nnz_expect || P.npts < 0 || (P.npts > 0 && (!P.idx || !P.w)) || !shot->records || !shot->records[q])
            return fail(c, GPI_EINVAL, "bad receiver point set");
        int64_t n3[3];
        shape_of(g.ndim, g.order, g.physics, g.n, f, n3);
        const int64_t N = n3[0] * n3[1] * n3[2];
        DevPoints& D = rec[q];
        D.field = f;
        D.npts = P.npts;
        D.nnz = P.nnz;
        std::vector<int64_t> off((size_t)P.npts * P.nnz);
        for (size_t k = 0; k < off.size(); ++k) {
            if (P.idx[k] < 0 || P.idx[k] >= N) return fail(c, GPI_EINVAL, "receiver index out of range");
            off[k] = padded(m, n3, P.idx[k]);
        }
        std::vector<float> w(P.w, P.w + (size_t)P.npts * P.nnz);
        int rc;
        if ((rc = S.put(&D.idx, off)) || (rc = S.put(&D.w, w))) return rc;
        if ((rc = S.alloc(&D.rec, (size_t)nt * std::max(P.npts, 1)))) return rc;
        HIPTRY(c, hipMemsetAsync(D.rec, 0, sizeof(float) * (size_t)nt * std::max(P.npts, 1), c->stream));
        if (!recordable(g.physics, f) || P.npts == 0) continue;
        // p is sampled by the velocity pass (before it), v by the stress pass (after v's sources)
        const bool vfield = f == GPI_VX || f == GPI_VY || f == GPI_VZ;
        GatherSet& G = vfield ? XS.g : XV.g;
        if (G.nset == 4) return fail(c, GPI_EINVAL, "too many receiver fields");
        const int k = G.nset++;
        G.nnz = P.nnz;
        G.npts[k] = P.npts;
        G.field[k] = m.field[f];
        G.idx[k] = D.idx;
        G.w[k] = D.w;
        G.rec[k] = D.rec;
        int& mx = vfield ? smax : vmax;
        mx = std::max(mx, P.npts);
    }
    XV.gather_blocks = XV.g.nset * ((vmax + 255) / 256);
    XS.gather_blocks = XS.g.nset * ((smax + 255) / 256);

    // ---- injection sets riding along with the passes ------------------------------------
    std::vector<int> vfall, sfall;
    for (int q = 0; q < shot->nsrc; ++q) {
        const DevPoints& D = src[q];
        if (D.ntarget == 0) continue;
        const bool vel = kinds[q] >= INJ_VX;
        PassExtras& X = vel ? XV : XS;
        bool& fused = vel ? vel_fused : str_fused;
        if (fused && X.ninj == MAX_INJ) fused = false;
        if (fused) {
            InjSet& I = X.inj[X.ninj++];
            I.kind = kinds[q];
            I.npts = D.npts;
            I.blk = D.blk;
            I.toff = D.toff;
            I.tptr = D.tptr;
            I.tsrc = D.tsrc;
            I.tw = D.tw;
            I.wav = D.wav;
            I.roff = D.toff + D.ntarget;
        }
    }
    // a pass either applies all of its source sets or none (keeps the reference's order)
    if (!vel_fused) XV.ninj = 0;
    if (!str_fused) XS.ninj = 0;
    for (int q = 0; q < shot->nsrc; ++q) {
        if (src[q].ntarget == 0) continue;
        const bool vel = kinds[q] >= INJ_VX;
        if (vel && !vel_fused) vfall.push_back(q);
        if (!vel && !str_fused) sfall.push_back(q);
    }

    // ---- reset wavefields and CPML memory (reset_w2!, types.jl:101-113) ----------------
    for (int f = GPI_P; f <= GPI_TAUYZ; ++f)
        if (m.field[f]) HIPTRY(c, hipMemsetAsync(m.field[f], 0, sizeof(float) * m.slab, c->stream));
    for (int f = GPI_DPDX; f < GPI_NFIELDS; ++f) {
        const Pml& P = m.pml[f];
        if (!P.psi) continue;
        int64_t n3[3];
        shape_of(g.ndim, g.order, g.physics, g.n, f, n3);
        const int ax = deriv_axis(f);
        size_t count = ax == 2 ? (size_t)(P.ps2 * 2 * g.npml) : (size_t)(P.ps2 * n3[2]);
        HIPTRY(c, hipMemsetAsync(P.psi, 0, sizeof(float) * count, c->stream));
    }

    // ---- time loop (propagate.jl:170-258) ------------------------------------------------
    // per step: velocity pass (records p at :177, sources of :202 ride along), then stress
    // pass (records v at :208, sources of :223 ride along)
    const int64_t nsamp = std::min<int64_t>((int64_t)c->sample_ev.size() / 4, nt);
    int64_t next_sample = nsamp ? 0 : -1, isamp = 0;
    int64_t launches = 0;
    HIPTRY(c, hipEventRecord(c->ev[0], c->stream));
    for (int64_t it = 0; it < nt; ++it) {
        const bool samp = it == next_sample;
        hipEvent_t* E = samp ? &c->sample_ev[4 * isamp] : nullptr;
        XV.it = it;
        XS.it = it;
        if (samp) HIPTRY(c, hipEventRecord(E[0], c->stream));
        if (d3) velocity_pass_3d(m, c->stream, XV);                                  // :177-202
        else velocity_pass_2d(m, c->stream, XV);
        if (samp) HIPTRY(c, hipEventRecord(E[1], c->stream));
        for (int q : vfall) inject(m, src[q], it, c->stream);
        if (samp) HIPTRY(c, hipEventRecord(E[2], c->stream));
        if (d3) stress_pass_3d(m, c->stream, XS);                                    // :208-223
        else stress_pass_2d(m, c->stream, XS);
        if (samp) HIPTRY(c, hipEventRecord(E[3], c->stream));
        for (int q : sfall) inject(m, src[q], it, c->stream);
        launches += 2 + (int64_t)vfall.size() + (int64_t)sfall.size();
        if (samp) {
            ++isamp;
            next_sample = isamp < nsamp ? (isamp * nt) / nsamp : -1;
        }
    }
    if (d3) velocity_pass_3d(m, c->stream, X0);                                      // :254-255
    else velocity_pass_2d(m, c->stream, X0);
    ++launches;
    HIPTRY(c, hipEventRecord(c->ev[1], c->stream));
    HIPTRY(c, hipGetLastError());
    HIPTRY(c, hipEventSynchronize(c->ev[1]));
    float ms = 0.0f;
    HIPTRY(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->timing = gpi_timing{};
    c->timing.total_ms = ms;
    c->timing.steps = nt;
    c->timing.launches = launches;
    c->timing.sampled = isamp;
    double vs = 0.0, ss = 0.0;
    for (int64_t k = 0; k < isamp; ++k) {
        float a = 0.0f, b = 0.0f;
        HIPTRY(c, hipEventElapsedTime(&a, c->sample_ev[4 * k], c->sample_ev[4 * k + 1]));
        HIPTRY(c, hipEventElapsedTime(&b, c->sample_ev[4 * k + 2], c->sample_ev[4 * k + 3]));
        vs += a;
        ss += b;
    }
    if (isamp) {
        c->timing.velocity_avg_ms = vs / (double)isamp;
        c->timing.stress_avg_ms = ss / (double)isamp;
    }

    // ---- copy back ------------------------------------------------------------------------
    for (int q = 0; q < shot->nrec; ++q) {
        const DevPoints& D = rec[q];
        std::vector<float> buf((size_t)nt * std::max(D.npts, 1));
        HIPTRY(c, hipMemcpy(buf.data(), D.rec, sizeof(float) * buf.size(), hipMemcpyDeviceToHost));
        float* out = shot->records[q];
        for (int r = 0; r < D.npts; ++r)
            for (int64_t it = 0; it < nt; ++it) out[(size_t)r * nt + it] = buf[(size_t)it * D.npts + r];
    }
    for (int q = 0; q < shot->nfinal; ++q) {
        const int f = shot->final_fields[q];
        if (f < 0 || f >= GPI_NFIELDS || !is_wavefield(f) || !m.field[f] || !shot->final_out[q])
            return fail(c, GPI_EINVAL, "bad final field request");
        int64_t n3[3];
        shape_of(g.ndim, g.order, g.physics, g.n, f, n3);
        int rc = download_pitched(c, m, shot->final_out[q], m.field[f], n3);
        if (rc) return rc;
    }
    return GPI_OK;
}

int gpi_last_timing(const gpi_ctx* c, gpi_timing* out)
{
    if (!c || !out) return GPI_EINVAL;
    *out = c->timing;
    return GPI_OK;
}

int gpi_set_pass_sampling(gpi_ctx* c, int32_t nsamples)
{
    if (!c || nsamples < 0 || nsamples > 4096) return fail(c, GPI_EINVAL, "nsamples must be in [0, 4096]");
    HIPTRY(c, hipSetDevice(c->device));
    HIPTRY(c, hipStreamSynchronize(c->stream));
    for (hipEvent_t e : c->sample_ev) (void)hipEventDestroy(e);
    c->sample_ev.clear();
    for (int k = 0; k < 4 * nsamples; ++k) {
        hipEvent_t e;
        HIPTRY(c, hipEventCreate(&e));
        c->sample_ev.push_back(e);
    }
    return GPI_OK;
}

}  // extern "C"
