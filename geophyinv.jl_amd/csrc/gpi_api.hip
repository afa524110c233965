// C-ABI of libgpi: context, resident models, timing.
#include <cmath>
#include <new>

#include "gpi_util.h"

using namespace gpi;

namespace gpi {

// validation and pitches of a new model (no device memory yet)
int model_begin(gpi_ctx* c, const gpi_grid* g, Model** out)
{
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
    *out = m;
    return GPI_OK;
}

// CPML strides and coefficient vectors of every derivative field
int model_pml(gpi_ctx* c, Model* m, const gpi_pml* pml, int32_t npml_arrays)
{
    const gpi_grid& g = m->g;
    if (npml_arrays > 0 && !pml) return fail(c, GPI_EINVAL, "null array list");
    const int L = 2 * g.npml;
    std::vector<float> vec;
    std::vector<int> order_f;
    for (int f = GPI_DPDX; f < GPI_NFIELDS; ++f) {
        if (!field_codes(g.ndim, g.physics, f)) continue;
        const gpi_pml* src = nullptr;
        for (int q = 0; q < npml_arrays; ++q)
            if (pml[q].field == f) src = &pml[q];
        if (g.pml_mask && !src) return fail(c, GPI_EINVAL, "missing CPML vectors for field " + std::to_string(f));
        if (src && (src->len != L || !src->a || !src->b || !src->kI))
            return fail(c, GPI_EINVAL, "CPML vectors must have length 2*npml");
        int64_t n3[3];
        shape_of(g.ndim, g.order, g.physics, g.n, f, n3);
        const int ax = deriv_axis(f);
        Pml& P = m->pml[f];
        P.active = 1;
        P.nd = (int32_t)n3[ax];
        if (ax == 2) {          // (nz_d, [ny_d,] 2npml)
            P.ps1 = m->pitch0;
            P.ps2 = m->pitch0 * m->pitch1;
        } else if (ax == 0) {   // (2npml, [ny_d,] nx_d)
            const int64_t pz = round_up(std::max(L, 1), 32);
            P.ps1 = pz;
            P.ps2 = pz * n3[1];
        } else {                // (nz_d, 2npml, nx_d)
            P.ps1 = m->pitch0;
            P.ps2 = m->pitch0 * L;
        }
        if (g.pml_mask) {
            order_f.push_back(f);
            vec.insert(vec.end(), src->a, src->a + L);
            vec.insert(vec.end(), src->b, src->b + L);
            vec.insert(vec.end(), src->kI, src->kI + L);
        }
    }
    if (!vec.empty()) {
        int rc = dev_alloc(c, &m->allocs, &m->bytes, (void**)&m->pmlvec, sizeof(float) * vec.size());
        if (rc) return rc;
        if ((rc = copy_h2d(c, m->pmlvec, vec.data(), sizeof(float) * vec.size()))) return rc;
        for (size_t q = 0; q < order_f.size(); ++q) {
            Pml& P = m->pml[order_f[q]];
            P.a = m->pmlvec + 3 * L * q;
            P.b = P.a + L;
            P.kI = P.b + L;
        }
    }
    return GPI_OK;
}

// the state of gpi_forward / gpi_forward_save
int model_state(gpi_ctx* c, Model* m)
{
    m->st0 = new State();
    return alloc_state(c, *m, m->st0);
}

int create_model_medium(gpi_ctx* c, const gpi_grid* g, const gpi_medium* med, const gpi_pml* pml,
                        int32_t npml_arrays, Model** out)
{
    if (g->ndim != 2) return fail(c, GPI_EINVAL, "device coefficient preparation is 2-D only");
    if (!med->vp || !med->rho || (g->physics == 1 && !med->vs)) return fail(c, GPI_EINVAL, "missing medium arrays");
    const int64_t nm[3] = {med->n[0], 1, med->n[2]};
    for (int a = 0; a < 3; a += 2) {
        const int64_t pad = ((g->pml_mask & GPI_FACE(a, 0)) ? g->npml : 0) + ((g->pml_mask & GPI_FACE(a, 1)) ? g->npml : 0);
        if (nm[a] < 1 || nm[a] + pad != g->n[a])
            return fail(c, GPI_EINVAL, "extended grid must be the medium padded by npml on the PML faces");
    }
    Model* m = nullptr;
    int rc = model_begin(c, g, &m);
    if (rc) return rc;
    auto bail = [&](int code) {
        free_model(c, m);
        return code;
    };
    {
        Scratch S(c);
        const size_t cnt = (size_t)(nm[0] * nm[2]);
        std::vector<const float*> host = {med->vp, med->rho, med->vs};
        float* dev[3] = {nullptr, nullptr, nullptr};
        for (int k = 0; k < 3; ++k) {
            if (!host[k]) continue;
            if ((rc = S.alloc(&dev[k], cnt)) || (rc = copy_h2d(c, dev[k], host[k], sizeof(float) * cnt)))
                return bail(rc);
        }
        if ((rc = prep_coefficients_2d(c, *m, dev[0], dev[2], dev[1], nm, c->stream))) return bail(rc);
    }
    if ((rc = model_pml(c, m, pml, npml_arrays)) || (rc = model_state(c, m))) return bail(rc);
    *out = m;
    return GPI_OK;
}

}  // namespace gpi

extern "C" {

int gpi_abi_version(void) { return GPI_ABI_VERSION; }

int gpi_field_shape(int32_t ndim, int32_t order, int32_t physics, const int64_t n[3], int32_t field, int64_t out[3])
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
    for (Store* s : c->stores) free_store(c, s);
    for (Model* m : c->models) free_model(c, m);
    for (auto& e : c->ev) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->sample_ev) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* gpi_last_error(const gpi_ctx* c) { return c ? c->err.c_str() : "null context"; }

int gpi_model_create(gpi_ctx* c, const gpi_grid* g, const gpi_array* coefs, int32_t ncoefs, const gpi_pml* pml,
                     int32_t npml_arrays, uint64_t* model_out)
{
    if (!c || !g || !model_out) return fail(c, GPI_EINVAL, "null argument");
    if ((ncoefs > 0 && !coefs) || (npml_arrays > 0 && !pml)) return fail(c, GPI_EINVAL, "null array list");
    Model* m = nullptr;
    int rc = model_begin(c, g, &m);
    if (rc) return rc;
    auto bail = [&](int code) {
        free_model(c, m);
        return code;
    };
    // coefficient arrays
    std::vector<bool> have(GPI_NCOEF, false);
    for (int q = 0; q < ncoefs; ++q) {
        const gpi_array& A = coefs[q];
        if (A.id < 0 || A.id >= GPI_NCOEF) return bail(fail(c, GPI_EINVAL, "bad coefficient id"));
        const int grid = coef_grid(g->ndim, g->physics, A.id);
        if (grid < 0) continue;   // not used by this physics
        int64_t n3[3];
        shape_of(g->ndim, g->order, g->physics, m->g.n, grid, n3);
        if (A.count != n3[0] * n3[1] * n3[2] || !A.data)
            return bail(fail(c, GPI_EINVAL, "coefficient " + std::to_string(A.id) + " has the wrong size"));
        if (!m->coef[A.id]) {
            rc = alloc_coef(c, *m, A.id);
            if (rc) return bail(rc);
        }
        rc = upload_pitched(c, *m, m->coef[A.id], A.data, n3);
        if (rc) return bail(rc);
        have[A.id] = true;
    }
    for (int q = 0; q < GPI_NCOEF; ++q)
        if (required_coef(g->ndim, g->physics, q) && !have[q])
            return bail(fail(c, GPI_EINVAL, "missing coefficient " + std::to_string(q)));
    if ((rc = model_pml(c, m, pml, npml_arrays)) || (rc = model_state(c, m))) return bail(rc);
    c->models.push_back(m);
    *model_out = (uint64_t)(uintptr_t)m;
    return GPI_OK;
}

int gpi_model_create_medium(gpi_ctx* c, const gpi_grid* g, const gpi_medium* med, const gpi_pml* pml,
                            int32_t npml_arrays, uint64_t* model_out)
{
    if (!c || !g || !med || !model_out) return fail(c, GPI_EINVAL, "null argument");
    Model* m = nullptr;
    int rc = create_model_medium(c, g, med, pml, npml_arrays, &m);
    if (rc) return rc;
    c->models.push_back(m);
    *model_out = (uint64_t)(uintptr_t)m;
    return GPI_OK;
}

int gpi_model_get_coef(gpi_ctx* c, uint64_t model, int32_t coef, float* out, int64_t count)
{
    if (!c || !out) return fail(c, GPI_EINVAL, "null argument");
    Model* m = find_model(c, model);
    if (!m) return fail(c, GPI_ENOTFOUND, "unknown model handle");
    if (coef < 0 || coef >= GPI_NCOEF || !m->coef[coef]) return fail(c, GPI_EINVAL, "coefficient not held by the model");
    const int grid = coef == GPI_C_INVK || coef == GPI_C_INVLAMBDA || coef == GPI_C_INVMU
                         ? (m->g.physics == 0 ? GPI_P : GPI_TAUXX)
                         : coef_grid(m->g.ndim, m->g.physics, coef);
    int64_t n3[3];
    shape_of(m->g.ndim, m->g.order, m->g.physics, m->g.n, grid, n3);
    if (count != n3[0] * n3[1] * n3[2]) return fail(c, GPI_EINVAL, "wrong output size");
    HIPTRY(c, hipSetDevice(c->device));
    return download_pitched(c, *m, out, m->coef[coef], n3);
}

int gpi_model_update(gpi_ctx* c, uint64_t model, const gpi_array* coefs, int32_t ncoefs)
{
    if (!c) return GPI_EINVAL;
    Model* m = find_model(c, model);
    if (!m) return fail(c, GPI_ENOTFOUND, "unknown model handle");
    if (ncoefs > 0 && !coefs) return fail(c, GPI_EINVAL, "null array list");
    HIPTRY(c, hipSetDevice(c->device));
    HIPTRY(c, hipStreamSynchronize(c->stream));
    for (int q = 0; q < ncoefs; ++q) {
        const gpi_array& A = coefs[q];
        if (A.id < 0 || A.id >= GPI_NCOEF) return fail(c, GPI_EINVAL, "bad coefficient id");
        const int grid = coef_grid(m->g.ndim, m->g.physics, A.id);
        if (grid < 0) continue;
        int64_t n3[3];
        shape_of(m->g.ndim, m->g.order, m->g.physics, m->g.n, grid, n3);
        if (A.count != n3[0] * n3[1] * n3[2] || !A.data) return fail(c, GPI_EINVAL, "coefficient has the wrong size");
        if (!m->coef[A.id]) {
            int rc = alloc_coef(c, *m, A.id);
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
    for (size_t k = 0; k < c->stores.size();) {   // stores of this model go with it
        Store* s = c->stores[k];
        if (s->model == m) {
            free_store(c, s);
            c->stores.erase(c->stores.begin() + k);
        } else {
            ++k;
        }
    }
    c->models.erase(std::find(c->models.begin(), c->models.end(), m));
    free_model(c, m);
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
