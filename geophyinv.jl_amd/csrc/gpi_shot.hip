// C-ABI of libgpi: the time loops of one supersource (forward, forward with boundary saving,
// and the adjoint run with the reference imaging condition).
#include <algorithm>
#include <map>

#include "fdtd2d.h"
#include "fdtd3d.h"
#include "gpi_util.h"
#include "pass_extras.h"

using namespace gpi;

namespace {

constexpr int NBOUND = 3;   // _fd_nbound (src/GeoPhyInv.jl:92)

bool recordable(int f)
{
    // propagate.jl:177 records [:p] and :208 records [:vx,:vy,:vz]; other fields of
    // rfields keep zero records, as in the reference.
    return f == GPI_P || f == GPI_VX || f == GPI_VY || f == GPI_VZ;
}

int velocity_pass(gpi_ctx* c, const Model& m, const State& st, const PassExtras& X)
{
    const int rc = m.g.ndim == 3 ? velocity_pass_3d(m, st, c->stream, X) : velocity_pass_2d(m, st, c->stream, X);
    return rc ? fail(c, GPI_EINVAL, "internal error: incomplete wavefield state") : GPI_OK;
}

int stress_pass(gpi_ctx* c, const Model& m, const State& st, const PassExtras& X)
{
    const int rc = m.g.ndim == 3 ? stress_pass_3d(m, st, c->stream, X) : stress_pass_2d(m, st, c->stream, X);
    return rc ? fail(c, GPI_EINVAL, "internal error: incomplete wavefield state") : GPI_OK;
}

}  // namespace

namespace gpi {

int plan_sources(gpi_ctx* c, const Model& m, Scratch& S, int nsrc, const gpi_points* src, const float* const* wav,
                 int64_t nt, bool velocity_only, PassExtras& XV, PassExtras& XS, Sources& out)
{
    const gpi_grid& g = m.g;
    const int nnz_expect = g.ndim == 2 ? 4 : 8;
    int vgx = 0, vgy = 0, sgx = 0, sgy = 0;
    const bool d3 = g.ndim == 3;
    const int vnb = d3 ? velocity_blocks_3d(m, &vgx, &vgy) : velocity_buckets_2d(m, &vgx);
    const int snb = d3 ? stress_blocks_3d(m, &sgx, &sgy) : stress_buckets_2d(m, &sgx);
    const int ty3 = d3 ? tile_rows_3d(m) : 1;
    out.pts.assign(nsrc, DevPoints{});
    std::vector<int> kinds(nsrc);
    bool vel_fused = true, str_fused = true;
    for (int q = 0; q < nsrc; ++q) {
        const gpi_points& P = src[q];
        const int f = P.field;
        if (!field_codes(g.ndim, g.physics, f) || !is_wavefield(f))
            return fail(c, GPI_EINVAL, "source field " + std::to_string(f) + " is not a wavefield of this physics");
        const bool vel = is_velocity(f);
        const bool str = f == GPI_P || f == GPI_TAUXX || f == GPI_TAUYY || f == GPI_TAUZZ;
        if (!vel && !str) return fail(c, GPI_EINVAL, "shear stresses cannot carry sources");
        if (velocity_only && !vel) return fail(c, GPI_EINVAL, "adjoint sources must be velocity fields");
        if (P.nnz != nnz_expect || P.npts < 0 || (P.npts > 0 && (!P.idx || !P.w || !wav || !wav[q])))
            return fail(c, GPI_EINVAL, "bad source point set");
        kinds[q] = f == GPI_VX ? INJ_VX : f == GPI_VY ? INJ_VY : f == GPI_VZ ? INJ_VZ
                 : g.physics == 0 ? INJ_P : (d3 ? INJ_NORMAL3 : INJ_NORMAL2);
        int64_t n3[3];
        shape_of(g.ndim, g.order, g.physics, g.n, f, n3);
        const int64_t N = n3[0] * n3[1] * n3[2];
        const int h = m.h;
        // target -> contributions in ascending source order
        std::map<int64_t, std::vector<std::pair<int, float>>> tg;
        for (int s = 0; s < P.npts; ++s)
            for (int k = 0; k < P.nnz; ++k) {
                const int64_t lin = P.idx[(size_t)s * P.nnz + k];
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
                blk = d3 ? (int)((iz - h + vel3d_row_shift(m)) / 64 + vgx * ((iy - h) / ty3 + vgy * (ix - h)))
                         : (int)((iz - h) / 64 + vgx * (ix - h));
            } else {
                blk = d3 ? (int)(iz / 64 + sgx * (iy / ty3 + sgy * ix)) : (int)(iz / 64 + sgx * ix);
                // a source on the free-surface row is applied after the surface condition
                // by the stand-alone kernel (source.jl:61-119 runs after update_stress!)
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
        DevPoints& D = out.pts[q];
        D.nbucket = nb;
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
            for (int64_t it = 0; it < nt; ++it) wv[(size_t)it * P.npts + s] = wav[q][(size_t)s * nt + it];
        if ((rc = S.put(&D.wav, wv))) return rc;
    }
    for (int q = 0; q < nsrc; ++q) {
        const DevPoints& D = out.pts[q];
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
    if (!vel_fused) XV.ninj = 0;
    if (!str_fused) XS.ninj = 0;
    for (int q = 0; q < nsrc; ++q) {
        if (out.pts[q].ntarget == 0) continue;
        const bool vel = kinds[q] >= INJ_VX;
        if (vel && !vel_fused) out.vfall.push_back(q);
        if (!vel && !str_fused) out.sfall.push_back(q);
    }
    return GPI_OK;
}

// Receivers in gather form; p is sampled by the velocity pass (before it updates v), v by
// the stress pass (after v's sources).
int plan_receivers(gpi_ctx* c, const Model& m, const State& st, Scratch& S, const gpi_shot* shot, PassExtras& XV,
                   PassExtras& XS, std::vector<DevPoints>& rec)
{
    const gpi_grid& g = m.g;
    const int64_t nt = shot->nt;
    const int nnz_expect = g.ndim == 2 ? 4 : 8;
    rec.assign(shot->nrec, DevPoints{});
    int vmax = 0, smax = 0;
    for (int q = 0; q < shot->nrec; ++q) {
        const gpi_points& P = shot->rec[q];
        const int f = P.field;
        if (!field_codes(g.ndim, g.physics, f) || !is_wavefield(f))
            return fail(c, GPI_EINVAL, "receiver field " + std::to_string(f) + " is not a wavefield of this physics");
        if (P.nnz != nnz_expect || P.npts < 0 || (P.npts > 0 && (!P.idx || !P.w)) || !shot->records ||
            !shot->records[q])
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
        if (!recordable(f) || P.npts == 0) continue;
        const bool vfield = is_velocity(f);
        GatherSet& G = vfield ? XS.g : XV.g;
        if (G.nset == 4) return fail(c, GPI_EINVAL, "too many receiver fields");
        const int k = G.nset++;
        G.nnz = P.nnz;
        G.npts[k] = P.npts;
        G.field[k] = st.field[f];
        G.idx[k] = D.idx;
        G.w[k] = D.w;
        G.rec[k] = D.rec;
        int& mx = vfield ? smax : vmax;
        mx = std::max(mx, P.npts);
    }
    const int vper = gather_per_block(XV.g.nnz), sper = gather_per_block(XS.g.nnz);
    XV.gather_blocks = XV.g.nset * ((vmax + vper - 1) / vper);
    XS.gather_blocks = XS.g.nset * ((smax + sper - 1) / sper);
    return GPI_OK;
}

}  // namespace gpi

namespace {

// stress fields kept by boundary_save! (boundary.jl:217-264): p, or tauxx, tauxz, tauzz
int stored_fields(const Model& m, int out[3])
{
    if (m.g.physics == 0) {
        out[0] = GPI_P;
        return 1;
    }
    out[0] = GPI_TAUXX;
    out[1] = GPI_TAUXZ;
    out[2] = GPI_TAUZZ;
    return 3;
}

// boundary_{save,force}{x,y,z}! use the min-face flag of the axis for both strips
// (boundary.jl:24,40)
int64_t strip_np(const Model& m, int axis) { return (m.g.pml_mask & GPI_FACE(axis, 0)) ? m.g.npml : 0; }

int alloc_store(gpi_ctx* c, Model& m, int64_t nt, Store** out)
{
    const gpi_grid& g = m.g;
    if (g.physics == 1 && g.ndim == 3) return fail(c, GPI_EINVAL, "boundary saving of 3-D elastic is not in the reference");
    Store* s = new Store();
    s->model = &m;
    s->nt = nt;
    s->nfield = stored_fields(m, s->field);
    auto bail = [&](int rc) {
        free_store(c, s);
        return rc;
    };
    for (int q = 0; q < s->nfield; ++q) {
        int64_t fn[3];
        shape_of(g.ndim, g.order, g.physics, g.n, s->field[q], fn);
        for (int ax = 0; ax < 3; ++ax) {
            if (g.ndim == 2 && ax == 1) continue;
            if (fn[ax] < strip_np(m, ax) + 2 * NBOUND)
                return bail(fail(c, GPI_EINVAL, "grid too short for the boundary strips"));
            int64_t sd[3] = {fn[0], fn[1], fn[2]};
            sd[ax] = 2 * NBOUND;
            s->strip_size[q][ax] = sd[0] * sd[1] * sd[2];
            int rc = dev_alloc(c, &s->allocs, &s->bytes, (void**)&s->strip[q][ax],
                               sizeof(float) * (size_t)(s->strip_size[q][ax] * nt));
            if (rc) return bail(rc);
        }
        int rc = dev_alloc(c, &s->allocs, &s->bytes, (void**)&s->snap[s->field[q]], sizeof(float) * (size_t)m.slab);
        if (rc) return bail(rc);
    }
    for (int f : {GPI_VX, GPI_VY, GPI_VZ}) {
        if (!field_codes(g.ndim, g.physics, f)) continue;
        int rc = dev_alloc(c, &s->allocs, &s->bytes, (void**)&s->snap[f], sizeof(float) * (size_t)m.slab);
        if (rc) return bail(rc);
    }
    *out = s;
    return GPI_OK;
}

// Checkpoint store: ceil(nt/every) - 1 states, the velocities at nt+1/2 and a Replay.
int alloc_checkpoint_store(gpi_ctx* c, Model& m, int64_t nt, int64_t every, Store** out)
{
    const gpi_grid& g = m.g;
    if (every < 1) return fail(c, GPI_EINVAL, "the checkpoint interval must be positive");
    Store* s = new Store();
    s->model = &m;
    s->nt = nt;
    s->every = every;
    s->replay = new Replay();
    s->replay->S = new Scratch(c);
    auto bail = [&](int rc) {
        free_store(c, s);
        return rc;
    };
    const int64_t K = (nt + every - 1) / every;
    const size_t per = state_floats(m);
    s->ckpt.assign((size_t)std::max<int64_t>(K - 1, 0), nullptr);
    for (float*& p : s->ckpt) {
        const int rc = dev_alloc(c, &s->allocs, &s->bytes, (void**)&p, sizeof(float) * per);
        if (rc) return bail(rc);
    }
    for (int f : {GPI_VX, GPI_VY, GPI_VZ}) {
        if (!field_codes(g.ndim, g.physics, f)) continue;
        const int rc = dev_alloc(c, &s->allocs, &s->bytes, (void**)&s->snap[f], sizeof(float) * (size_t)m.slab);
        if (rc) return bail(rc);
    }
    *out = s;
    return GPI_OK;
}

// boundary_save! (save) or boundary_force! (!save) of step `it`: axes x, (y,) z
void boundary_strips(const Model& m, const State& st, const Store& s, int64_t it, bool save, hipStream_t stream,
                     int64_t* launches)
{
    for (int q = 0; q < s.nfield; ++q) {
        int64_t fn[3];
        shape_of(m.g.ndim, m.g.order, m.g.physics, m.g.n, s.field[q], fn);
        for (int ax = 2; ax >= 0; --ax) {
            if (!s.strip[q][ax]) continue;
            strip_copy(m, st.field[s.field[q]], fn, ax, strip_np(m, ax), s.strip[q][ax] + s.strip_size[q][ax] * it,
                       save, stream);
            ++*launches;
        }
    }
}

Store* find_store(gpi_ctx* c, uint64_t h)
{
    for (Store* s : c->stores)
        if ((uint64_t)(uintptr_t)s == h) return s;
    return nullptr;
}

// Pass timing of the sampled steps.
struct Sampler {
    gpi_ctx* c;
    int64_t nt, nsamp, next, k = 0;
    explicit Sampler(gpi_ctx* cc, int64_t n) : c(cc), nt(n)
    {
        nsamp = std::min<int64_t>((int64_t)c->sample_ev.size() / 4, nt);
        next = nsamp ? 0 : -1;
    }
    hipEvent_t* at(int64_t it) { return it == next ? &c->sample_ev[4 * k] : nullptr; }
    void done(int64_t it)
    {
        if (it != next) return;
        ++k;
        next = k < nsamp ? (k * nt) / nsamp : -1;
    }
};

int finish_timing(gpi_ctx* c, const Sampler& smp, int64_t nt, int64_t launches)
{
    HIPTRY(c, hipEventRecord(c->ev[1], c->stream));
    HIPTRY(c, hipGetLastError());
    HIPTRY(c, hipEventSynchronize(c->ev[1]));
    float ms = 0.0f;
    HIPTRY(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->timing = gpi_timing{};
    c->timing.total_ms = ms;
    c->timing.steps = nt;
    c->timing.launches = launches;
    c->timing.sampled = smp.k;
    double vs = 0.0, ss = 0.0;
    for (int64_t k = 0; k < smp.k; ++k) {
        float a = 0.0f, b = 0.0f;
        HIPTRY(c, hipEventElapsedTime(&a, c->sample_ev[4 * k], c->sample_ev[4 * k + 1]));
        HIPTRY(c, hipEventElapsedTime(&b, c->sample_ev[4 * k + 2], c->sample_ev[4 * k + 3]));
        vs += a;
        ss += b;
    }
    if (smp.k) {
        c->timing.velocity_avg_ms = vs / (double)smp.k;
        c->timing.stress_avg_ms = ss / (double)smp.k;
    }
    return GPI_OK;
}

// Floats of the staging buffer copy_records transposes one receiver set into.
size_t records_stage_floats(const std::vector<DevPoints>& rec, int64_t nt)
{
    size_t most = 0;
    for (const DevPoints& D : rec) most = std::max(most, (size_t)nt * (size_t)std::max(D.npts, 0));
    return most;
}

// Records of every receiver set into the caller's buffers (time fastest): transposed on
// the device into one staging buffer (`tmp`, records_stage_floats, and the int `flag`, or
// scratch buffers when null), then copied; the stream has drained on return.
// Returns GPI_ENONFINITE (after the copies) when a record holds Inf or NaN, as a shot
// that violates the stability condition does (src/fdtd/stability.jl:8-60 only warns).
int copy_records(gpi_ctx* c, const std::vector<DevPoints>& rec, int64_t nt, float* const* out,
                 float* tmp = nullptr, int* flag = nullptr)
{
    const size_t most = records_stage_floats(rec, nt);
    if (!most) return GPI_OK;
    Scratch S(c);
    int rc = tmp ? GPI_OK : S.alloc(&tmp, most);
    if (!rc && !flag) rc = S.alloc(&flag, 1);
    if (rc) return rc;
    int bad = -1;
    for (size_t q = 0; q < rec.size(); ++q) {
        const DevPoints& D = rec[q];
        if (D.npts <= 0) continue;
        int nonfinite = 0;
        HIPTRY(c, hipMemsetAsync(flag, 0, sizeof(int), c->stream));
        records_transpose(D.rec, tmp, nt, D.npts, flag, c->stream);
        HIPTRY(c, hipGetLastError());
        HIPTRY(c, hipMemcpyAsync(out[q], tmp, sizeof(float) * (size_t)nt * (size_t)D.npts, hipMemcpyDeviceToHost,
                                 c->stream));
        HIPTRY(c, hipMemcpyAsync(&nonfinite, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPTRY(c, hipStreamSynchronize(c->stream));   // tmp is reused by the next set
        if (nonfinite && bad < 0) bad = (int)q;
    }
    if (bad >= 0)
        return fail(c, GPI_ENONFINITE, "receiver set " + std::to_string(bad) +
                                           " recorded non-finite values (is dt within the stability limit?)");
    return GPI_OK;
}

// mod_x_proc! for one supersource in :forward (save == nullptr) or :forward_save mode
// (src/fdtd/propagate.jl:138-261).
int run_shot(gpi_ctx* c, Model& m, const gpi_shot* shot, Store* save)
{
    const gpi_grid& g = m.g;
    const int64_t nt = shot->nt;
    const State& st = *m.st0;
    Scratch S(c);
    PassExtras XV{}, XS{}, X0{};
    Sources src;
    const bool ckpt = save && save->every > 0;
    // a checkpoint store keeps the source tables for its replays
    int rc = plan_sources(c, m, ckpt ? *save->replay->S : S, shot->nsrc, shot->src, shot->wavelets, nt, false, XV,
                          XS, src);
    if (rc) return rc;
    if (ckpt) {   // no receivers yet in XV, XS
        save->replay->xv = XV;
        save->replay->xs = XS;
        save->replay->src = src;
    }
    std::vector<DevPoints> rec;
    if ((rc = plan_receivers(c, m, st, S, shot, XV, XS, rec))) return rc;
    for (int q = 0; q < shot->nfinal; ++q) {
        const int f = shot->final_fields[q];
        if (f < 0 || f >= GPI_NFIELDS || !is_wavefield(f) || !st.field[f] || !shot->final_out[q])
            return fail(c, GPI_EINVAL, "bad final field request");
    }
    // snapshots: device copies of the field after the requested steps (snaps_save!,
    // src/fdtd/fdtd.jl:587-591, called at propagate.jl:237-244)
    std::vector<float*> snap;
    if (shot->nsnap > 0) {
        const int f = shot->snap_field;
        if (f < 0 || f >= GPI_NFIELDS || !is_wavefield(f) || !st.field[f] || !shot->snap_steps || !shot->snap_out)
            return fail(c, GPI_EINVAL, "bad snapshot request");
        snap.resize(shot->nsnap);
        for (int q = 0; q < shot->nsnap; ++q) {
            if (!shot->snap_out[q]) return fail(c, GPI_EINVAL, "null snapshot output");
            if ((rc = S.alloc(&snap[q], (size_t)m.slab))) return rc;
        }
    }

    // illumination (compute_illum!, src/fdtd/fdtd.jl:570-581): acoustic p only
    double* illum = nullptr;
    if (shot->illum) {
        if (g.physics != 0) return fail(c, GPI_EINVAL, "illumination is of the acoustic pressure only");
        if ((rc = S.alloc(&illum, (size_t)m.slab))) return rc;
        HIPTRY(c, hipMemsetAsync(illum, 0, sizeof(double) * (size_t)m.slab, c->stream));
    }

    // staging of the copy-back, taken before the first step so that a call whose buffers
    // do not fit fails before any work is queued: transposed records + non-finite flag,
    // and one packed field (final fields, snapshots, FP64 illumination)
    float* rec_stage = nullptr;
    int* rec_flag = nullptr;
    if (const size_t most = records_stage_floats(rec, nt))
        if ((rc = S.alloc(&rec_stage, most)) || (rc = S.alloc(&rec_flag, 1))) return rc;
    size_t pack_bytes = 0;
    auto field_bytes = [&](int f, size_t elem) {
        int64_t n3[3];
        shape_of(g.ndim, g.order, g.physics, g.n, f, n3);
        pack_bytes = std::max(pack_bytes, elem * (size_t)(n3[0] * n3[1] * n3[2]));
    };
    for (int q = 0; q < shot->nfinal; ++q) field_bytes(shot->final_fields[q], sizeof(float));
    if (shot->nsnap > 0) field_bytes(shot->snap_field, sizeof(float));
    if (illum) field_bytes(GPI_P, sizeof(double));
    char* pack = nullptr;
    if (pack_bytes && (rc = S.alloc(&pack, pack_bytes))) return rc;

    // reset wavefields and CPML memory (reset_w2!, types.jl:101-113)
    if ((rc = zero_state(c, m, st, c->stream))) return rc;

    // per step: velocity pass (records p at :177, sources of :202 ride along), then stress
    // pass (records v at :208, sources of :223 ride along), then boundary_save! (:229)
    Sampler smp(c, nt);
    int64_t launches = 0;
    HIPTRY(c, hipEventRecord(c->ev[0], c->stream));
    for (int64_t it = 0; it < nt; ++it) {
        hipEvent_t* E = smp.at(it);
        XV.it = it;
        XS.it = it;
        if (E) HIPTRY(c, hipEventRecord(E[0], c->stream));
        if ((rc = velocity_pass(c, m, st, XV))) return rc;
        if (E) HIPTRY(c, hipEventRecord(E[1], c->stream));
        for (int q : src.vfall) inject(m, st, src.pts[q], it, c->stream);
        if (E) HIPTRY(c, hipEventRecord(E[2], c->stream));
        if ((rc = stress_pass(c, m, st, XS))) return rc;
        if (E) HIPTRY(c, hipEventRecord(E[3], c->stream));
        for (int q : src.sfall) inject(m, st, src.pts[q], it, c->stream);
        launches += 2 + (int64_t)src.vfall.size() + (int64_t)src.sfall.size();
        if (ckpt) {
            const int64_t k = (it + 1) / save->every;
            if ((it + 1) % save->every == 0 && k >= 1 && k <= (int64_t)save->ckpt.size())
                if ((rc = copy_state(c, m, st, save->ckpt[k - 1], true, c->stream))) return rc;
        } else if (save) {
            boundary_strips(m, st, *save, it, true, c->stream, &launches);
        }
        if (illum) {   // after the step's boundary saving and imaging, as propagate.jl:236
            illum_accumulate(st.field[GPI_P], illum, m.slab, c->stream);
            ++launches;
        }
        for (int q = 0; q < shot->nsnap; ++q)
            if (shot->snap_steps[q] == it)
                HIPTRY(c, hipMemcpyAsync(snap[q], st.field[shot->snap_field], sizeof(float) * (size_t)m.slab,
                                         hipMemcpyDeviceToDevice, c->stream));
        smp.done(it);
    }
    if (save && !ckpt)   // boundary_save_snap_tau!: -stress at nt (boundary.jl:268-287)
        for (int q = 0; q < save->nfield; ++q) {
            negate_copy(save->snap[save->field[q]], st.field[save->field[q]], m.slab, c->stream);
            ++launches;
        }
    if ((rc = velocity_pass(c, m, st, X0))) return rc;   // :254-255
    ++launches;
    if (save)   // boundary_save_snap_v!: v at nt+1/2 (boundary.jl:289-306)
        for (int f : {GPI_VX, GPI_VY, GPI_VZ})
            if (save->snap[f])
                HIPTRY(c, hipMemcpyAsync(save->snap[f], st.field[f], sizeof(float) * (size_t)m.slab,
                                         hipMemcpyDeviceToDevice, c->stream));
    if ((rc = finish_timing(c, smp, nt, launches))) return rc;

    // copy back (everything, also when the records hold non-finite values: that code is
    // returned last, with the message copy_records left)
    const int rec_rc = copy_records(c, rec, nt, shot->records, rec_stage, rec_flag);
    if (rec_rc && rec_rc != GPI_ENONFINITE) return rec_rc;
    const std::string rec_msg = c->err;
    for (int q = 0; q < shot->nfinal; ++q) {
        const int f = shot->final_fields[q];
        int64_t n3[3];
        shape_of(g.ndim, g.order, g.physics, g.n, f, n3);
        if ((rc = download_pitched(c, m, shot->final_out[q], st.field[f], n3, pack))) return rc;
    }
    for (int q = 0; q < shot->nsnap; ++q) {
        int64_t n3[3];
        shape_of(g.ndim, g.order, g.physics, g.n, shot->snap_field, n3);
        if ((rc = download_pitched(c, m, shot->snap_out[q], snap[q], n3, pack))) return rc;
    }
    if (illum) {
        int64_t n3[3];
        shape_of(g.ndim, g.order, g.physics, g.n, GPI_P, n3);
        if ((rc = download_pitched(c, m, shot->illum, illum, n3, pack))) return rc;
    }
    if (rec_rc) return fail(c, rec_rc, rec_msg);
    return GPI_OK;
}

int check_shot(gpi_ctx* c, uint64_t model, const gpi_shot* shot, Model** out)
{
    if (!c || !shot) return fail(c, GPI_EINVAL, "null argument");
    Model* m = find_model(c, model);
    if (!m) return fail(c, GPI_ENOTFOUND, "unknown model handle");
    if (shot->nt < 1) return fail(c, GPI_EINVAL, "nt must be positive");
    if (shot->nsrc < 0 || shot->nrec < 0 || shot->nsrc > 8 || shot->nrec > 16 || shot->nfinal < 0 ||
        (shot->nsrc > 0 && !shot->src) || (shot->nrec > 0 && !shot->rec) ||
        (shot->nfinal > 0 && (!shot->final_fields || !shot->final_out)) || shot->nsnap < 0)
        return fail(c, GPI_EINVAL, "bad number of source/receiver fields");
    HIPTRY(c, hipSetDevice(c->device));
    *out = m;
    return GPI_OK;
}

}  // namespace

namespace gpi {

int forward_shot(gpi_ctx* c, uint64_t model, const gpi_shot* shot)
{
    Model* m = nullptr;
    int rc = check_shot(c, model, shot, &m);
    if (rc) return rc;
    return run_shot(c, *m, shot, nullptr);
}

}  // namespace gpi

extern "C" {

int gpi_forward(gpi_ctx* c, uint64_t model, const gpi_shot* shot) { return forward_shot(c, model, shot); }

int gpi_forward_save(gpi_ctx* c, uint64_t model, const gpi_shot* shot, uint64_t* store_out)
{
    if (!store_out) return fail(c, GPI_EINVAL, "null argument");
    *store_out = 0;
    Model* m = nullptr;
    int rc = check_shot(c, model, shot, &m);
    if (rc) return rc;
    Store* s = nullptr;
    if ((rc = alloc_store(c, *m, shot->nt, &s))) return rc;
    if ((rc = run_shot(c, *m, shot, s))) {
        (void)hipStreamSynchronize(c->stream);
        free_store(c, s);
        return rc;
    }
    c->stores.push_back(s);
    *store_out = (uint64_t)(uintptr_t)s;
    return GPI_OK;
}

int gpi_store_free(gpi_ctx* c, uint64_t store)
{
    if (!c) return GPI_EINVAL;
    Store* s = find_store(c, store);
    if (!s) return fail(c, GPI_ENOTFOUND, "unknown store handle");
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->stores.erase(std::find(c->stores.begin(), c->stores.end(), s));
    free_store(c, s);
    return GPI_OK;
}

int gpi_adjoint(gpi_ctx* c, uint64_t model, uint64_t store, const gpi_adjoint_shot* shot, float dtI, float* g_invK,
                float* g_rho)
{
    if (!c || !shot || !g_invK || !g_rho) return fail(c, GPI_EINVAL, "null argument");
    Model* mp = find_model(c, model);
    if (!mp) return fail(c, GPI_ENOTFOUND, "unknown model handle");
    Store* sp = find_store(c, store);
    if (!sp || sp->model != mp) return fail(c, GPI_ENOTFOUND, "unknown store handle for this model");
    Model& m = *mp;
    const Store& sv = *sp;
    const gpi_grid& g = m.g;
    if (g.physics != 0 || g.ndim != 2) return fail(c, GPI_EINVAL, "the reference gradient exists for 2-D acoustic only");
    const int64_t nt = shot->nt;
    if (nt != sv.nt) return fail(c, GPI_EINVAL, "nt differs from the stored forward run");
    if (shot->nsrc1 < 0 || shot->nsrc2 < 0 || shot->nsrc1 > 8 || shot->nsrc2 > 8 ||
        (shot->nsrc1 > 0 && !shot->src1) || (shot->nsrc2 > 0 && !shot->src2))
        return fail(c, GPI_EINVAL, "bad number of source fields");
    if (!(dtI > 0.0f) || !std::isfinite(dtI)) return fail(c, GPI_EINVAL, "dtI must be positive");
    HIPTRY(c, hipSetDevice(c->device));
    const State& s1 = *m.st0;
    Scratch S(c);
    // second wavefield: its own wavefields and CPML memory
    struct Held {
        gpi_ctx* c;
        State st;
        ~Held()
        {
            (void)hipStreamSynchronize(c->stream);
            free_allocs(c, st.allocs, st.bytes);
        }
    } h2{c, State{}};
    int rc = alloc_state(c, m, &h2.st);
    if (rc) return rc;
    const State& s2 = h2.st;
    PassExtras XV1{}, XS1{}, XV2{}, XS2{};
    Sources src1, src2;
    if ((rc = plan_sources(c, m, S, shot->nsrc1, shot->src1, shot->wav1, nt, false, XV1, XS1, src1))) return rc;
    if ((rc = plan_sources(c, m, S, shot->nsrc2, shot->src2, shot->wav2, nt, true, XV2, XS2, src2))) return rc;
    // previous-step copies (save_tp!) in two alternating sets, and the gradients
    float* tp[2][6];
    for (auto& set : tp)
        for (float*& p : set)
            if ((rc = S.alloc(&p, (size_t)m.slab))) return rc;
    float *gK, *gR;
    if ((rc = S.alloc(&gK, (size_t)m.slab)) || (rc = S.alloc(&gR, (size_t)m.slab))) return rc;
    const size_t bytes = sizeof(float) * (size_t)m.slab;
    HIPTRY(c, hipMemsetAsync(gK, 0, bytes, c->stream));
    HIPTRY(c, hipMemsetAsync(gR, 0, bytes, c->stream));

    // initial conditions: the first wavefield from the snapshots (boundary_force_snap_tau!,
    // boundary_force_snap_v!, boundary.jl:173-212), the second at rest
    if ((rc = zero_state(c, m, s1, c->stream)) || (rc = zero_state(c, m, s2, c->stream))) return rc;
    for (int f : {GPI_P, GPI_VX, GPI_VZ})
        HIPTRY(c, hipMemcpyAsync(s1.field[f], sv.snap[f], bytes, hipMemcpyDeviceToDevice, c->stream));
    const float* cur[6] = {s1.field[GPI_P], s1.field[GPI_VX], s1.field[GPI_VZ],
                           s2.field[GPI_P], s2.field[GPI_VX], s2.field[GPI_VZ]};
    for (int k = 0; k < 6; ++k) {
        if (k < 3) HIPTRY(c, hipMemcpyAsync(tp[0][k], cur[k], bytes, hipMemcpyDeviceToDevice, c->stream));
        else HIPTRY(c, hipMemsetAsync(tp[0][k], 0, bytes, c->stream));
    }

    // time loop of :adjoint with npw = 2 (propagate.jl:170-247): save_tp!, boundary_force!
    // on the first wavefield, both wavefields' passes and sources, then compute_gradient!
    Sampler smp(c, nt);
    int64_t launches = 0;
    HIPTRY(c, hipEventRecord(c->ev[0], c->stream));
    for (int64_t it = 0; it < nt; ++it) {
        hipEvent_t* E = smp.at(it);
        XV1.it = XS1.it = XV2.it = XS2.it = it;
        boundary_strips(m, s1, sv, nt - 1 - it, false, c->stream, &launches);
        if (E) HIPTRY(c, hipEventRecord(E[0], c->stream));
        if ((rc = velocity_pass(c, m, s1, XV1)) || (rc = velocity_pass(c, m, s2, XV2))) return rc;
        if (E) HIPTRY(c, hipEventRecord(E[1], c->stream));
        for (int q : src1.vfall) inject(m, s1, src1.pts[q], it, c->stream);
        for (int q : src2.vfall) inject(m, s2, src2.pts[q], it, c->stream);
        if (E) HIPTRY(c, hipEventRecord(E[2], c->stream));
        if ((rc = stress_pass(c, m, s1, XS1)) || (rc = stress_pass(c, m, s2, XS2))) return rc;
        if (E) HIPTRY(c, hipEventRecord(E[3], c->stream));
        for (int q : src1.sfall) inject(m, s1, src1.pts[q], it, c->stream);
        // gradlame! and gradrho! (gradient.jl:17-64), fused with the next step's save_tp!
        image_acoustic_2d(m, cur, tp[it & 1], tp[(it + 1) & 1], gK, gR, dtI, c->stream);
        launches += 5 + (int64_t)(src1.vfall.size() + src2.vfall.size() + src1.sfall.size());
        smp.done(it);
    }
    if ((rc = finish_timing(c, smp, nt, launches))) return rc;
    // the trailing velocity half step of :254-255 changes no output of this mode: omitted
    int64_t n3[3];
    shape_of(g.ndim, g.order, g.physics, g.n, GPI_P, n3);
    if ((rc = download_pitched(c, m, g_invK, gK, n3)) || (rc = download_pitched(c, m, g_rho, gR, n3))) return rc;
    return GPI_OK;
}

int gpi_forward_checkpoint(gpi_ctx* c, uint64_t model, const gpi_shot* shot, int32_t every, uint64_t* store_out)
{
    if (!store_out) return fail(c, GPI_EINVAL, "null argument");
    *store_out = 0;
    Model* m = nullptr;
    int rc = check_shot(c, model, shot, &m);
    if (rc) return rc;
    Store* s = nullptr;
    if ((rc = alloc_checkpoint_store(c, *m, shot->nt, every, &s))) return rc;
    if ((rc = run_shot(c, *m, shot, s))) {
        (void)hipStreamSynchronize(c->stream);
        free_store(c, s);
        return rc;
    }
    c->stores.push_back(s);
    *store_out = (uint64_t)(uintptr_t)s;
    return GPI_OK;
}

int gpi_adjoint_checkpoint(gpi_ctx* c, uint64_t model, uint64_t store, const gpi_adjoint_shot* shot, float dtI,
                           double* g_invK, double* g_rho)
{
    if (!c || !shot || !g_invK) return fail(c, GPI_EINVAL, "null argument");
    Model* mp = find_model(c, model);
    if (!mp) return fail(c, GPI_ENOTFOUND, "unknown model handle");
    Store* sp = find_store(c, store);
    if (!sp || sp->model != mp) return fail(c, GPI_ENOTFOUND, "unknown store handle for this model");
    if (sp->every < 1 || !sp->replay) return fail(c, GPI_EINVAL, "the store holds no checkpoints");
    Model& m = *mp;
    const Store& sv = *sp;
    const Replay& R = *sv.replay;
    const gpi_grid& g = m.g;
    if (g.physics != 0) return fail(c, GPI_EINVAL, "checkpointed gradients are acoustic only");
    const int64_t nt = shot->nt;
    if (nt != sv.nt) return fail(c, GPI_EINVAL, "nt differs from the stored forward run");
    if (shot->nsrc2 < 0 || shot->nsrc2 > 8 || (shot->nsrc2 > 0 && !shot->src2))
        return fail(c, GPI_EINVAL, "bad number of adjoint source fields");
    if (!(dtI > 0.0f) || !std::isfinite(dtI)) return fail(c, GPI_EINVAL, "dtI must be positive");
    HIPTRY(c, hipSetDevice(c->device));
    const bool rho = g_rho != nullptr;
    const State& s1 = *m.st0;
    Scratch S(c);
    struct Held {
        gpi_ctx* c;
        State st;
        ~Held()
        {
            (void)hipStreamSynchronize(c->stream);
            free_allocs(c, st.allocs, st.bytes);
        }
    } h2{c, State{}};
    int rc = alloc_state(c, m, &h2.st);
    if (rc) return rc;
    const State& s2 = h2.st;
    PassExtras XV2{}, XS2{};
    Sources src2;
    if ((rc = plan_sources(c, m, S, shot->nsrc2, shot->src2, shot->wav2, nt, true, XV2, XS2, src2))) return rc;
    PassExtras XV1 = R.xv, XS1 = R.xs;

    // velocity components present: x, (y,) z
    std::vector<int> vf;
    for (int f : {GPI_VX, GPI_VY, GPI_VZ})
        if (field_codes(g.ndim, g.physics, f)) vf.push_back(f);
    const int nv = (int)vf.size();
    const int64_t C = sv.every;
    const int64_t K = (nt + C - 1) / C;
    const int64_t L = std::min(C, nt);
    const size_t slab = (size_t)m.slab;
    const size_t bytes = sizeof(float) * slab;
    // segment buffers: p_{n0-1} .. p_{n1-1} and (for rho) v_{n0} .. v_{n1}
    float* segP = nullptr;
    float* segV[3] = {};
    float* vnext[3] = {};
    if ((rc = S.alloc(&segP, (size_t)(L + 1) * slab))) return rc;
    if (rho)
        for (int d = 0; d < nv; ++d)
            if ((rc = S.alloc(&segV[d], (size_t)(L + 1) * slab)) || (rc = S.alloc(&vnext[d], slab))) return rc;
    // adjoint velocities before the step (for rho), two alternating sets; the adjoint p
    // before the step is the one the stress pass starts from (the velocity pass leaves p),
    // so the imaging reads it there and keeps no copy
    float* tp[2][4] = {};
    if (rho)
        for (auto& set : tp)
            for (int d = 0; d < nv; ++d)
                if ((rc = S.alloc(&set[1 + d], slab))) return rc;
    double *gK = nullptr, *gR = nullptr;
    if ((rc = S.alloc(&gK, slab))) return rc;
    if (rho && (rc = S.alloc(&gR, slab))) return rc;
    HIPTRY(c, hipMemsetAsync(gK, 0, sizeof(double) * slab, c->stream));
    if (rho) HIPTRY(c, hipMemsetAsync(gR, 0, sizeof(double) * slab, c->stream));
    if ((rc = zero_state(c, m, s2, c->stream))) return rc;
    for (auto& set : tp)
        for (float* p : set)
            if (p) HIPTRY(c, hipMemsetAsync(p, 0, bytes, c->stream));
    if (rho)   // v_nt: the trailing half step of the stored run
        for (int d = 0; d < nv; ++d)
            HIPTRY(c, hipMemcpyAsync(vnext[d], sv.snap[vf[d]], bytes, hipMemcpyDeviceToDevice, c->stream));

    Sampler smp(c, nt);
    int64_t launches = 0, it = 0;
    HIPTRY(c, hipEventRecord(c->ev[0], c->stream));
    for (int64_t k = K - 1; k >= 0; --k) {
        const int64_t n0 = k * C, n1 = std::min(n0 + C, nt);
        // the forward state before step n0, then the segment recomputed with its sources
        if (k == 0) rc = zero_state(c, m, s1, c->stream);
        else rc = copy_state(c, m, s1, sv.ckpt[k - 1], false, c->stream);
        if (rc) return rc;
        HIPTRY(c, hipMemcpyAsync(segP, s1.field[GPI_P], bytes, hipMemcpyDeviceToDevice, c->stream));
        for (int64_t n = n0; n < n1; ++n) {
            XV1.it = XS1.it = n;
            if ((rc = velocity_pass(c, m, s1, XV1))) return rc;
            for (int q : R.src.vfall) inject(m, s1, R.src.pts[q], n, c->stream);
            if ((rc = stress_pass(c, m, s1, XS1))) return rc;
            for (int q : R.src.sfall) inject(m, s1, R.src.pts[q], n, c->stream);
            HIPTRY(c, hipMemcpyAsync(segP + (size_t)(n - n0 + 1) * slab, s1.field[GPI_P], bytes,
                                     hipMemcpyDeviceToDevice, c->stream));
            if (rho)
                for (int d = 0; d < nv; ++d)
                    HIPTRY(c, hipMemcpyAsync(segV[d] + (size_t)(n - n0) * slab, s1.field[vf[d]], bytes,
                                             hipMemcpyDeviceToDevice, c->stream));
            launches += 2 + (int64_t)(R.src.vfall.size() + R.src.sfall.size());
        }
        if (rho)
            for (int d = 0; d < nv; ++d)
                HIPTRY(c, hipMemcpyAsync(segV[d] + (size_t)(n1 - n0) * slab, vnext[d], bytes,
                                         hipMemcpyDeviceToDevice, c->stream));
        // the adjoint steps of this segment, forward step n = nt - 1 - it
        for (int64_t n = n1 - 1; n >= n0; --n, ++it) {
            hipEvent_t* E = smp.at(it);
            XV2.it = XS2.it = it;
            if (E) HIPTRY(c, hipEventRecord(E[0], c->stream));
            if ((rc = velocity_pass(c, m, s2, XV2))) return rc;
            if (E) HIPTRY(c, hipEventRecord(E[1], c->stream));
            for (int q : src2.vfall) inject(m, s2, src2.pts[q], it, c->stream);
            ImageSets I{};
            float* const* prev = tp[it & 1];
            float* const* next = tp[(it + 1) & 1];
            I.a2[0] = s2.field[GPI_P];                   // p2 before the step
            I.pa = segP + (size_t)(n - n0) * slab;       // p_{n-1}
            I.pb = segP + (size_t)(n - n0 + 1) * slab;   // p_n
            image_acoustic_exact(m, I, gK, nullptr, dtI, c->stream);   // invK term
            if (E) HIPTRY(c, hipEventRecord(E[2], c->stream));
            if ((rc = stress_pass(c, m, s2, XS2))) return rc;
            if (E) HIPTRY(c, hipEventRecord(E[3], c->stream));
            if (rho) {   // density term: v2 before the step (a copy), then keep this step's v2
                ImageSets Iv{};
                for (int d = 0; d < nv; ++d) {
                    const int a = vf[d] - GPI_VX;   // 0 x, 1 y, 2 z
                    Iv.a2[1 + a] = prev[1 + d];
                    Iv.c2[1 + a] = s2.field[vf[d]];
                    Iv.n2[1 + a] = next[1 + d];
                    Iv.va[a] = segV[d] + (size_t)(n - n0) * slab;       // v_n
                    Iv.vb[a] = segV[d] + (size_t)(n - n0 + 1) * slab;   // v_{n+1}
                }
                image_acoustic_exact(m, Iv, nullptr, gR, dtI, c->stream);
            }
            launches += (rho ? 4 : 3) + (int64_t)src2.vfall.size();
            smp.done(it);
        }
        if (rho)
            for (int d = 0; d < nv; ++d)
                HIPTRY(c, hipMemcpyAsync(vnext[d], segV[d], bytes, hipMemcpyDeviceToDevice, c->stream));
    }
    if ((rc = finish_timing(c, smp, nt, launches))) return rc;
    int64_t n3[3];
    shape_of(g.ndim, g.order, g.physics, g.n, GPI_P, n3);
    if ((rc = download_pitched(c, m, g_invK, gK, n3))) return rc;
    if (rho && (rc = download_pitched(c, m, g_rho, gR, n3))) return rc;
    return GPI_OK;
}

int gpi_adjoint_checkpoint_elastic(gpi_ctx* c, uint64_t model, uint64_t store, const gpi_adjoint_shot* shot,
                                   double* g_invlambda, double* g_invmu, double* g_rho)
{
    if (!c || !shot || !g_invlambda || !g_invmu || !g_rho) return fail(c, GPI_EINVAL, "null argument");
    Model* mp = find_model(c, model);
    if (!mp) return fail(c, GPI_ENOTFOUND, "unknown model handle");
    Store* sp = find_store(c, store);
    if (!sp || sp->model != mp) return fail(c, GPI_ENOTFOUND, "unknown store handle for this model");
    if (sp->every < 1 || !sp->replay) return fail(c, GPI_EINVAL, "the store holds no checkpoints");
    Model& m = *mp;
    const Store& sv = *sp;
    const Replay& R = *sv.replay;
    const gpi_grid& g = m.g;
    if (g.physics != 1) return fail(c, GPI_EINVAL, "the elastic gradient needs an elastic model");
    if (!m.coef[GPI_C_INVLAMBDA] || !m.coef[GPI_C_INVMU]) return fail(c, GPI_EINVAL, "the model lacks invlambda/invmu");
    const int64_t nt = shot->nt;
    if (nt != sv.nt) return fail(c, GPI_EINVAL, "nt differs from the stored forward run");
    if (shot->nsrc2 < 0 || shot->nsrc2 > 8 || (shot->nsrc2 > 0 && !shot->src2))
        return fail(c, GPI_EINVAL, "bad number of adjoint source fields");
    for (int q = 0; q < shot->nsrc2; ++q)   // the adjoint of velocity records only
        if (!is_velocity(shot->src2[q].field))
            return fail(c, GPI_EINVAL, "elastic gradients take velocity records only (adjoint sources on vx, vy, vz)");
    HIPTRY(c, hipSetDevice(c->device));
    const bool d3 = g.ndim == 3;
    const int H = m.h;
    const State& s1 = *m.st0;
    Scratch S(c);
    struct Held {
        gpi_ctx* c;
        State st;
        ~Held()
        {
            (void)hipStreamSynchronize(c->stream);
            free_allocs(c, st.allocs, st.bytes);
        }
    } h2{c, State{}};
    int rc = alloc_state(c, m, &h2.st);
    if (rc) return rc;
    const State& s2 = h2.st;
    PassExtras XV2{}, XS2{};
    Sources src2;
    if ((rc = plan_sources(c, m, S, shot->nsrc2, shot->src2, shot->wav2, nt, true, XV2, XS2, src2))) return rc;
    PassExtras XV1 = R.xv, XS1 = R.xs;

    // components, their regions on the slab (z, y, x) and their coefficient stencils
    const int nz = (int)g.n[0], ny = d3 ? (int)g.n[1] : 1, nx = (int)g.n[2];
    const std::vector<int> nf = d3 ? std::vector<int>{GPI_TAUXX, GPI_TAUYY, GPI_TAUZZ}
                                   : std::vector<int>{GPI_TAUXX, GPI_TAUZZ};
    const std::vector<int> sf = d3 ? std::vector<int>{GPI_TAUXY, GPI_TAUXZ, GPI_TAUYZ} : std::vector<int>{GPI_TAUXZ};
    const std::vector<int> vf = d3 ? std::vector<int>{GPI_VX, GPI_VY, GPI_VZ} : std::vector<int>{GPI_VX, GPI_VZ};
    ElasticImage E{};
    E.nn = (int)nf.size();
    E.ns = (int)sf.size();
    E.nv = (int)vf.size();
    for (int q = 0; q < E.ns; ++q) {
        // stress grids of str3d_elastic / str_el_point and the @av_* cells of dtav*imu
        int* hi = E.shi[q];
        int* b = E.sbase[q];
        if (!d3) {                        // tauxz: (nz-h) x (nx-h), cells (i..i+1, k..k+1)
            hi[0] = nz - H, hi[1] = 1, hi[2] = nx - H;
            E.se1[q] = 0, E.se2[q] = 2;
        } else if (sf[q] == GPI_TAUXY) {  // dvxdy grid, cells (i+h, j..j+1, k..k+1)
            hi[0] = nz - 2 * H, hi[1] = ny - H, hi[2] = nx - H;
            b[0] = H;
            E.se1[q] = 1, E.se2[q] = 2;
        } else if (sf[q] == GPI_TAUXZ) {  // dvxdz grid, cells (i..i+1, j+h, k..k+1)
            hi[0] = nz - H, hi[1] = ny - 2 * H, hi[2] = nx - H;
            b[1] = H;
            E.se1[q] = 0, E.se2[q] = 2;
        } else {                          // dvydz grid, cells (i..i+1, j..j+1, k+h)
            hi[0] = nz - H, hi[1] = ny - H, hi[2] = nx - 2 * H;
            b[2] = H;
            E.se1[q] = 0, E.se2[q] = 1;
        }
    }
    for (int q = 0; q < E.nv; ++q) {
        // vel*_point updates v[inn]: [h, n-h) across, [h, n) along its own axis
        const int ax = vf[q] == GPI_VX ? 2 : vf[q] == GPI_VY ? 1 : 0;
        const int n3[3] = {nz, ny, nx};
        for (int a = 0; a < 3; ++a) {
            E.vlo[q][a] = (a == 1 && !d3) ? 0 : H;
            E.vhi[q][a] = (a == 1 && !d3) ? 1 : (a == ax ? n3[a] : n3[a] - H);
        }
        E.vaxis[q] = ax;
    }
    const int64_t C = sv.every;
    const int64_t L = std::min(C, nt);
    const size_t slab = (size_t)m.slab;
    const size_t bytes = sizeof(float) * slab;
    // segment buffers: the forward stresses tau^{n0} .. tau^{n1} and velocities v^{n0-1/2} .. v^{n1-1/2}
    std::vector<int> all = nf;
    all.insert(all.end(), sf.begin(), sf.end());
    all.insert(all.end(), vf.begin(), vf.end());
    std::vector<float*> seg(all.size(), nullptr);
    for (float*& p : seg)
        if ((rc = S.alloc(&p, (size_t)(L + 1) * slab))) return rc;
    double* acc[8] = {};
    const int nacc = 2 + E.ns + E.nv;
    for (int q = 0; q < nacc; ++q) {
        if ((rc = S.alloc(&acc[q], slab))) return rc;
        HIPTRY(c, hipMemsetAsync(acc[q], 0, sizeof(double) * slab, c->stream));
    }
    E.T = acc[0];
    E.P = acc[1];
    for (int q = 0; q < E.ns; ++q) E.S[q] = acc[2 + q];
    for (int q = 0; q < E.nv; ++q) E.R[q] = acc[2 + E.ns + q];
    if ((rc = zero_state(c, m, s2, c->stream))) return rc;
    // normal-stress sources of the forward run: their per-target traces (source_trace)
    std::vector<int> nsrc;
    std::vector<double*> sacc;
    for (size_t q = 0; q < R.src.pts.size(); ++q) {
        const int f = R.src.pts[q].field;
        if ((f != GPI_TAUXX && f != GPI_TAUYY && f != GPI_TAUZZ) || R.src.pts[q].ntarget <= 0) continue;
        double* a = nullptr;
        if ((rc = S.alloc(&a, (size_t)R.src.pts[q].ntarget))) return rc;
        HIPTRY(c, hipMemsetAsync(a, 0, sizeof(double) * (size_t)R.src.pts[q].ntarget, c->stream));
        nsrc.push_back((int)q);
        sacc.push_back(a);
    }
    auto save = [&](int64_t slot) -> int {
        for (size_t q = 0; q < all.size(); ++q)
            HIPTRY(c, hipMemcpyAsync(seg[q] + (size_t)slot * slab, s1.field[all[q]], bytes, hipMemcpyDeviceToDevice,
                                     c->stream));
        return GPI_OK;
    };

    Sampler smp(c, nt);
    int64_t launches = 0, it = 0;
    const int64_t K = (nt + C - 1) / C;
    HIPTRY(c, hipEventRecord(c->ev[0], c->stream));
    for (int64_t k = K - 1; k >= 0; --k) {
        const int64_t n0 = k * C, n1 = std::min(n0 + C, nt);
        if (k == 0) rc = zero_state(c, m, s1, c->stream);
        else rc = copy_state(c, m, s1, sv.ckpt[k - 1], false, c->stream);
        if (rc || (rc = save(0))) return rc;
        for (int64_t n = n0; n < n1; ++n) {
            XV1.it = XS1.it = n;
            if ((rc = velocity_pass(c, m, s1, XV1))) return rc;
            for (int q : R.src.vfall) inject(m, s1, R.src.pts[q], n, c->stream);
            if ((rc = stress_pass(c, m, s1, XS1))) return rc;
            for (int q : R.src.sfall) inject(m, s1, R.src.pts[q], n, c->stream);
            if ((rc = save(n - n0 + 1))) return rc;
            launches += 2 + (int64_t)(R.src.vfall.size() + R.src.sfall.size());
        }
        // adjoint steps of this segment (forward step n = nt - 1 - it), imaging between the passes
        for (int64_t n = n1 - 1; n >= n0; --n, ++it) {
            hipEvent_t* Ev = smp.at(it);
            XV2.it = XS2.it = it;
            if (Ev) HIPTRY(c, hipEventRecord(Ev[0], c->stream));
            if ((rc = velocity_pass(c, m, s2, XV2))) return rc;
            if (Ev) HIPTRY(c, hipEventRecord(Ev[1], c->stream));
            for (int q : src2.vfall) inject(m, s2, src2.pts[q], it, c->stream);
            const size_t a = (size_t)(n - n0) * slab, b = a + slab;
            int w = 0;
            for (int q = 0; q < E.nn; ++q, ++w) {
                E.an[q] = s2.field[nf[q]];
                E.f0n[q] = seg[w] + a;
                E.f1n[q] = seg[w] + b;
            }
            for (int q = 0; q < E.ns; ++q, ++w) {
                E.as[q] = s2.field[sf[q]];
                E.f0s[q] = seg[w] + a;
                E.f1s[q] = seg[w] + b;
            }
            for (int q = 0; q < E.nv; ++q, ++w) {
                E.av[q] = s2.field[vf[q]];
                E.f0v[q] = seg[w] + a;
                E.f1v[q] = seg[w] + b;
            }
            image_elastic(m, E, c->stream);
            for (size_t q = 0; q < nsrc.size(); ++q) source_trace(R.src.pts[nsrc[q]], n, E.an, E.nn, sacc[q], c->stream);
            if (Ev) HIPTRY(c, hipEventRecord(Ev[2], c->stream));
            if ((rc = stress_pass(c, m, s2, XS2))) return rc;
            for (int q : src2.sfall) inject(m, s2, src2.pts[q], it, c->stream);
            if (Ev) HIPTRY(c, hipEventRecord(Ev[3], c->stream));
            launches += 3 + (int64_t)(src2.vfall.size() + src2.sfall.size());
            smp.done(it);
        }
    }
    if ((rc = finish_timing(c, smp, nt, launches))) return rc;
    double* gout[3] = {};
    for (auto& p : gout)
        if ((rc = S.alloc(&p, slab))) return rc;
    elastic_gradient(m, E, gout[0], gout[1], gout[2], c->stream);
    for (size_t q = 0; q < nsrc.size(); ++q) source_correction(m, R.src.pts[nsrc[q]], sacc[q], E.nn, gout[0], gout[1], c->stream);
    HIPTRY(c, hipGetLastError());
    int64_t n3[3];
    shape_of(g.ndim, g.order, g.physics, g.n, GPI_TAUXX, n3);
    double* host[3] = {g_invlambda, g_invmu, g_rho};
    for (int q = 0; q < 3; ++q)
        if ((rc = download_pitched(c, m, host[q], gout[q], n3))) return rc;
    return GPI_OK;
}

int gpi_born(gpi_ctx* c, uint64_t model, const gpi_shot* shot, const gpi_array* dcoefs, int32_t ndcoefs)
{
    Model* mp = nullptr;
    int rc = check_shot(c, model, shot, &mp);
    if (rc) return rc;
    Model& m = *mp;
    const gpi_grid& g = m.g;
    if (g.physics != 0) return fail(c, GPI_EINVAL, "Born modelling is acoustic only");
    if (shot->nfinal != 0 || shot->nsnap != 0) return fail(c, GPI_EINVAL, "Born runs return records only");
    if (ndcoefs < 0 || (ndcoefs > 0 && !dcoefs)) return fail(c, GPI_EINVAL, "bad perturbation list");
    const int64_t nt = shot->nt;
    const State& w1 = *m.st0;
    Scratch S(c);
    struct Held {
        gpi_ctx* c;
        State st;
        ~Held()
        {
            (void)hipStreamSynchronize(c->stream);
            free_allocs(c, st.allocs, st.bytes);
        }
    } h2{c, State{}};
    if ((rc = alloc_state(c, m, &h2.st))) return rc;
    const State& w2 = h2.st;
    // coefficient perturbations on the device (absent ones are zero)
    const int ids[4] = {GPI_C_DTINVAVXIRHO, GPI_C_DTINVAVYIRHO, GPI_C_DTINVAVZIRHO, GPI_C_DTK};
    float* dev[4] = {};
    for (int q = 0; q < 4; ++q) {
        if (g.ndim == 2 && ids[q] == GPI_C_DTINVAVYIRHO) continue;
        if ((rc = S.alloc(&dev[q], (size_t)m.slab))) return rc;
        HIPTRY(c, hipMemsetAsync(dev[q], 0, sizeof(float) * (size_t)m.slab, c->stream));
    }
    for (int a = 0; a < ndcoefs; ++a) {
        const gpi_array& A = dcoefs[a];
        int q = -1;
        for (int t = 0; t < 4; ++t)
            if (ids[t] == A.id && dev[t]) q = t;
        if (q < 0) return fail(c, GPI_EINVAL, "perturbation " + std::to_string(A.id) + " is not a Born coefficient");
        int64_t n3[3];
        shape_of(g.ndim, g.order, g.physics, g.n, coef_grid(g.ndim, g.physics, A.id), n3);
        if (A.count != n3[0] * n3[1] * n3[2] || !A.data)
            return fail(c, GPI_EINVAL, "perturbation " + std::to_string(A.id) + " has the wrong size");
        HIPTRY(c, hipStreamSynchronize(c->stream));
        if ((rc = upload_pitched(c, m, dev[q], A.data, n3))) return rc;
    }
    const float* dc[3] = {dev[0], dev[1], dev[2]};
    PassExtras XV1{}, XS1{}, XV2{}, XS2{};
    Sources src;
    if ((rc = plan_sources(c, m, S, shot->nsrc, shot->src, shot->wavelets, nt, false, XV1, XS1, src))) return rc;
    std::vector<DevPoints> rec;
    if ((rc = plan_receivers(c, m, w2, S, shot, XV2, XS2, rec))) return rc;
    if ((rc = zero_state(c, m, w1, c->stream)) || (rc = zero_state(c, m, w2, c->stream))) return rc;

    // per step (propagate.jl:170-226 with activepw = [1, 2], src_flags = [true, false],
    // rec_flags = [false, true]): velocity passes, velocity sources of the background,
    // velocity Born sources, stress passes, stress sources, stress Born sources
    Sampler smp(c, nt);
    int64_t launches = 0;
    HIPTRY(c, hipEventRecord(c->ev[0], c->stream));
    for (int64_t it = 0; it < nt; ++it) {
        hipEvent_t* E = smp.at(it);
        XV1.it = XS1.it = XV2.it = XS2.it = it;
        if (E) HIPTRY(c, hipEventRecord(E[0], c->stream));
        if ((rc = velocity_pass(c, m, w1, XV1)) || (rc = velocity_pass(c, m, w2, XV2))) return rc;
        if (E) HIPTRY(c, hipEventRecord(E[1], c->stream));
        for (int q : src.vfall) inject(m, w1, src.pts[q], it, c->stream);
        born_sources(m, w1, w2, dc, dev[3], true, c->stream);
        if (E) HIPTRY(c, hipEventRecord(E[2], c->stream));
        if ((rc = stress_pass(c, m, w1, XS1)) || (rc = stress_pass(c, m, w2, XS2))) return rc;
        if (E) HIPTRY(c, hipEventRecord(E[3], c->stream));
        for (int q : src.sfall) inject(m, w1, src.pts[q], it, c->stream);
        born_sources(m, w1, w2, dc, dev[3], false, c->stream);
        launches += 6 + (int64_t)(src.vfall.size() + src.sfall.size());
        smp.done(it);
    }
    if ((rc = finish_timing(c, smp, nt, launches))) return rc;
    if ((rc = copy_records(c, rec, nt, shot->records))) return rc;
    return GPI_OK;
}

}  // extern "C"
