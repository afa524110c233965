// Host-side helpers of the C ABI (field tables, allocation accounting, pitched copies).
#include "fdtd3d.h"
#include "gpi_util.h"

namespace gpi {

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
        return c == GPI_C_DTINVAVXIRHO || c == GPI_C_DTINVAVZIRHO || c == GPI_C_DTK || c == GPI_C_RHO ||
               (ndim == 3 && c == GPI_C_DTINVAVYIRHO);
    // the elastic passes form dtM, dtlambda and the averaged mu coefficients from invlambda
    // and invmu in the kernels; the derived arrays stay for gpi_model_get_coef
    if (c == GPI_C_DTINVAVXIRHO || c == GPI_C_DTINVAVZIRHO || c == GPI_C_DTM || c == GPI_C_DTLAMBDA ||
        c == GPI_C_RHO || c == GPI_C_INVLAMBDA || c == GPI_C_INVMU)
        return true;
    if (ndim == 2) return c == GPI_C_DTAVMU;
    return c == GPI_C_DTINVAVYIRHO || c == GPI_C_DTAVXZIMU || c == GPI_C_DTAVXYIMU || c == GPI_C_DTAVYZIMU;
}

Model* find_model(gpi_ctx* c, uint64_t h)
{
    for (Model* m : c->models)
        if ((uint64_t)(uintptr_t)m == h) return m;
    return nullptr;
}

int dev_alloc(gpi_ctx* c, std::vector<void*>* owner, uint64_t* owner_bytes, void** p, size_t bytes)
{
    if (c->budget && c->used + bytes > c->budget)
        return fail(c, GPI_ENOMEM, "allocation of " + std::to_string(bytes) + " bytes exceeds the context's HBM budget");
    hipError_t e = c->pool ? c->pool->alloc(p, bytes) : hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(c, GPI_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->used += bytes;
    owner->push_back(*p);
    *owner_bytes += bytes;
    return GPI_OK;
}

int arena_alloc(gpi_ctx* c, void** p, size_t bytes)
{
    Arena& a = *c->arena;
    const size_t n = (size_t)round_up((int64_t)std::max<size_t>(bytes, 1), 256);
    if (n > a.cap - a.off)
        return fail(c, GPI_ENOMEM, "the call's device buffers exceed the static scratch of " + std::to_string(a.cap) +
                                       " bytes (" + std::to_string(a.off) + " carved, " + std::to_string(n) +
                                       " more asked)");
    *p = a.base + a.off;
    a.off += n;
    return GPI_OK;
}

int alloc_coef(gpi_ctx* c, Model& m, int id)
{
    const bool shifted = m.g.ndim == 3 && (id == GPI_C_DTINVAVXIRHO || id == GPI_C_DTINVAVYIRHO ||
                                           id == GPI_C_DTINVAVZIRHO);
    const int64_t off = shifted ? vel3d_row_shift(m) : 0;
    void* base = nullptr;
    const int rc = dev_alloc(c, &m.allocs, &m.bytes, &base, sizeof(float) * (size_t)(m.slab + off));
    if (rc) return rc;
    m.coef[id] = (float*)base + off;
    return GPI_OK;
}

void free_allocs(gpi_ctx* c, std::vector<void*>& allocs, uint64_t& bytes)
{
    if (c->defer_free) {
        c->defer_free->insert(c->defer_free->end(), allocs.begin(), allocs.end());
        if (c->defer_bytes) *c->defer_bytes += bytes;
    } else if (c->pool) {
        for (void* p : allocs) c->pool->free(p);
    } else {
        for (void* p : allocs) (void)hipFree(p);
    }
    allocs.clear();
    c->used -= std::min<uint64_t>(c->used, bytes);
    bytes = 0;
}

void free_state(gpi_ctx* c, State* st)
{
    if (!st) return;
    free_allocs(c, st->allocs, st->bytes);
    delete st;
}

void free_model(gpi_ctx* c, Model* m)
{
    free_state(c, m->st0);
    free_allocs(c, m->allocs, m->bytes);
    delete m;
}

void free_store(gpi_ctx* c, Store* s)
{
    if (!s) return;
    if (s->replay) {
        delete s->replay->S;   // frees the replayed source tables (after the stream drained)
        delete s->replay;
    }
    free_allocs(c, s->allocs, s->bytes);
    delete s;
}

size_t state_floats(const Model& m)
{
    size_t n = 0;
    for (int f = GPI_P; f <= GPI_TAUYZ; ++f)
        if (field_codes(m.g.ndim, m.g.physics, f)) n += (size_t)m.slab;
    for (int f = GPI_DPDX; f < GPI_NFIELDS; ++f) n += psi_floats(m, f);
    return n;
}

int copy_state(gpi_ctx* c, const Model& m, const State& st, float* ckpt, bool to_ckpt, hipStream_t s)
{
    size_t off = 0;
    auto one = [&](float* a, size_t n) -> int {
        if (!a) return fail(c, GPI_EINVAL, "internal error: incomplete wavefield state");
        float* dst = to_ckpt ? ckpt + off : a;
        const float* src = to_ckpt ? a : ckpt + off;
        HIPTRY(c, hipMemcpyAsync(dst, src, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
        off += n;
        return GPI_OK;
    };
    for (int f = GPI_P; f <= GPI_TAUYZ; ++f)
        if (field_codes(m.g.ndim, m.g.physics, f)) {
            const int rc = one(st.field[f], (size_t)m.slab);
            if (rc) return rc;
        }
    for (int f = GPI_DPDX; f < GPI_NFIELDS; ++f) {
        const size_t n = psi_floats(m, f);
        if (!n) continue;
        const int rc = one(st.psi[f], n);
        if (rc) return rc;
    }
    return GPI_OK;
}

size_t psi_floats(const Model& m, int f)
{
    const Pml& P = m.pml[f];
    if (!P.active || !m.g.pml_mask) return 0;
    int64_t n3[3];
    shape_of(m.g.ndim, m.g.order, m.g.physics, m.g.n, f, n3);
    const int ax = deriv_axis(f);
    return ax == 2 ? (size_t)(P.ps2 * 2 * m.g.npml) : (size_t)(P.ps2 * n3[2]);
}

int alloc_state(gpi_ctx* c, const Model& m, State* st)
{
    for (int f = GPI_P; f <= GPI_TAUYZ; ++f) {
        if (!field_codes(m.g.ndim, m.g.physics, f)) continue;
        int rc = dev_alloc(c, &st->allocs, &st->bytes, (void**)&st->field[f], sizeof(float) * (size_t)m.slab);
        if (rc) return rc;
    }
    for (int f = GPI_DPDX; f < GPI_NFIELDS; ++f) {
        const size_t n = psi_floats(m, f);
        if (!n) continue;
        int rc = dev_alloc(c, &st->allocs, &st->bytes, (void**)&st->psi[f], sizeof(float) * n);
        if (rc) return rc;
    }
    return GPI_OK;
}

int zero_state(gpi_ctx* c, const Model& m, const State& st, hipStream_t s)
{
    for (int f = GPI_P; f <= GPI_TAUYZ; ++f)
        if (st.field[f]) HIPTRY(c, hipMemsetAsync(st.field[f], 0, sizeof(float) * (size_t)m.slab, s));
    for (int f = GPI_DPDX; f < GPI_NFIELDS; ++f)
        if (st.psi[f]) HIPTRY(c, hipMemsetAsync(st.psi[f], 0, sizeof(float) * psi_floats(m, f), s));
    return GPI_OK;
}

int copy_h2d(gpi_ctx* c, void* dst, const void* src, size_t bytes)
{
    PinnedStage* st = c->stage;
    if (!st) {
        HIPTRY(c, hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
        return GPI_OK;
    }
    if (!bytes) return GPI_OK;
    const size_t need = (size_t)round_up((int64_t)bytes, 256);
    if (need > st->cap) {   // a larger ring: the queued copies out of the old one finish first
        HIPTRY(c, hipStreamSynchronize(c->stream));
        if (st->buf) (void)hipHostFree(st->buf);
        st->buf = nullptr;
        st->cap = st->off = 0;
        const size_t cap = std::max<size_t>(need, (size_t)64 << 20);
        HIPTRY(c, hipHostMalloc((void**)&st->buf, cap, hipHostMallocDefault));
        st->cap = cap;
    } else if (st->off + need > st->cap) {   // wrap once every queued copy has finished
        HIPTRY(c, hipStreamSynchronize(c->stream));
        st->off = 0;
    }
    std::memcpy(st->buf + st->off, src, bytes);
    HIPTRY(c, hipMemcpyAsync(dst, st->buf + st->off, bytes, hipMemcpyHostToDevice, c->stream));
    st->off += need;
    return GPI_OK;
}

int upload_pitched(gpi_ctx* c, const Model& m, float* dst, const float* src, const int64_t n[3])
{
    std::vector<float> stage((size_t)m.slab, 0.0f);
    for (int64_t ix = 0; ix < n[2]; ++ix)
        for (int64_t iy = 0; iy < n[1]; ++iy)
            std::memcpy(&stage[(size_t)(m.pitch0 * (iy + m.pitch1 * ix))], src + n[0] * (iy + n[1] * ix),
                        sizeof(float) * (size_t)n[0]);
    return copy_h2d(c, dst, stage.data(), sizeof(float) * (size_t)m.slab);
}

// pitched slab -> compact array of n0 x n1 x n2 (grid: x over n0 in 256s, y over n1, z over n2)
template <typename T>
__global__ __launch_bounds__(256) void pack_pitched_kernel(const T* __restrict__ src, T* __restrict__ dst,
                                                           int n0, int n1, int64_t pitch0, int64_t pitch1)
{
    const int i = (int)blockIdx.x * 256 + (int)threadIdx.x;
    if (i >= n0) return;
    const int64_t y = blockIdx.y, z = blockIdx.z;
    dst[i + (int64_t)n0 * (y + (int64_t)n1 * z)] = src[i + pitch0 * (y + pitch1 * z)];
}

// Packed on the device, then one copy into the caller's array (no host-side staging slab).
template <typename T>
static int download_packed(gpi_ctx* c, const Model& m, T* dst, const T* src, const int64_t n[3], void* stage)
{
    if (n[0] < 1 || n[1] < 1 || n[2] < 1 || n[0] > m.pitch0 || n[1] > 65535 || n[2] > 65535)
        return fail(c, GPI_EINVAL, "download: bad field size");
    const size_t cnt = (size_t)(n[0] * n[1] * n[2]);
    Scratch S(c);
    T* tmp = (T*)stage;
    int rc = tmp ? GPI_OK : S.alloc(&tmp, cnt);
    if (rc) return rc;
    hipLaunchKernelGGL(pack_pitched_kernel<T>, dim3((unsigned)((n[0] + 255) / 256), (unsigned)n[1], (unsigned)n[2]),
                       dim3(256), 0, c->stream, src, tmp, (int)n[0], (int)n[1], m.pitch0, m.pitch1);
    HIPTRY(c, hipGetLastError());
    HIPTRY(c, hipStreamSynchronize(c->stream));
    HIPTRY(c, hipMemcpy(dst, tmp, sizeof(T) * cnt, hipMemcpyDeviceToHost));
    return GPI_OK;
}

int download_pitched(gpi_ctx* c, const Model& m, float* dst, const float* src, const int64_t n[3], void* stage)
{
    return download_packed(c, m, dst, src, n, stage);
}

int download_pitched(gpi_ctx* c, const Model& m, double* dst, const double* src, const int64_t n[3], void* stage)
{
    return download_packed(c, m, dst, src, n, stage);
}

}  // namespace gpi
