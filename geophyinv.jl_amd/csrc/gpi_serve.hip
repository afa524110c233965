// Serving: continuous batching of 2-D shots on one GPU.
//
// A scheduler thread owns two streams: the admission stream builds a shot's resident
// state (medium -> coefficients on the device, CPML vectors, wavefields, sources,
// receivers) while the compute stream runs the batch.  Every time step of the batch is
// one velocity and one stress launch over all resident shots (fdtd2d.hip, batched
// kernels), each shot at its own step.  Between chunks of steps, finished shots leave
// (records copied to pinned memory) and waiting shots enter as slots and the HBM budget
// allow.  Device memory comes from hipMalloc; the memory of departed shots is released
// only after both streams have drained, so no queued launch can reference freed memory.
// At most two chunks are in flight, so admission and copy-back overlap the device work.
// The workgroup shape of the batched kernels is chosen per batch size by timing each
// candidate on live chunks; a candidate's time is read when its chunk is accounted anyway
// (the older of the two chunks in flight), so testing never drains the stream.  A full
// chunk whose layout (shape and workgroup buckets) has run before is one hipGraphLaunch of
// a captured graph of its 2 x chunk_steps launches: the kernels read the batch step and
// shot count from device memory and the launches cover a bucketed workgroup count (extra
// workgroups return), so one graph per shape and bucket serves every batch composition
// (least recently used graphs evicted); a layout's first chunk is launched pass by pass,
// so layouts that never recur are not captured.
//
// No reference counterpart: the reference propagates supersources one after another on
// each worker (src/fdtd/propagate.jl:100-106, :138-261).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>

#include "fdtd2d.h"
#include "gpi_util.h"
#include "pass_extras.h"

using namespace gpi;

namespace {

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// A request with every input copied (the caller's buffers are free after gpi_submit),
// except the record outputs.
struct Req {
    uint64_t ticket = 0;
    gpi_grid g{};
    gpi_medium med{};
    std::vector<float> vp, vs, rho;
    std::vector<gpi_pml> pml;
    std::vector<std::vector<float>> pmlv;
    int64_t nt = 0;
    std::vector<gpi_points> src, rec;
    std::vector<std::vector<int64_t>> idx;
    std::vector<std::vector<float>> w, wav;
    std::vector<const float*> wavp;
    std::vector<float*> out;
    uint64_t est = 0;
    double t_submit = 0.0;
};

struct Slot {
    std::unique_ptr<Req> r;
    Model* m = nullptr;
    Scratch* S = nullptr;
    PassExtras XV{}, XS{};
    std::vector<DevPoints> rec;
    int64_t start = 0;
    float* pin = nullptr;
    size_t pin_cap = 0;
    float* trec = nullptr;   // device staging of the records in the caller's layout
    hipEvent_t ev = nullptr;
};

struct PinnedPool {
    std::mutex mu;   // get() on the scheduler thread, put() on the completion thread
    std::vector<std::pair<float*, size_t>> free_list;
    std::vector<float*> all;
    float* get(size_t count, size_t* cap)
    {
        std::lock_guard<std::mutex> lk(mu);
        for (size_t k = 0; k < free_list.size(); ++k)
            if (free_list[k].second >= count) {
                float* p = free_list[k].first;
                *cap = free_list[k].second;
                free_list.erase(free_list.begin() + k);
                return p;
            }
        void* p = nullptr;
        const size_t n = std::max<size_t>(count, 1);
        if (hipHostMalloc(&p, sizeof(float) * n, hipHostMallocDefault) != hipSuccess) return nullptr;
        all.push_back((float*)p);
        *cap = n;
        return (float*)p;
    }
    void put(float* p, size_t cap)
    {
        std::lock_guard<std::mutex> lk(mu);
        free_list.push_back({p, cap});
    }
    ~PinnedPool()
    {
        for (float* p : all) (void)hipHostFree(p);
    }
};

uint64_t estimate_bytes(const Req& r)
{
    const gpi_grid& g = r.g;
    const int h = g.order - 1;
    const uint64_t ld = (uint64_t)round_up(g.n[0] + h, 32);
    const uint64_t slab = ld * (uint64_t)(g.n[2] + h) * sizeof(float);
    const uint64_t arrays = g.physics == 0 ? 5 + 3 : 8 + 5;
    const uint64_t nder = g.physics == 0 ? 4 : 8;
    const uint64_t psi = nder * (uint64_t)(2 * g.npml) * (uint64_t)(std::max(g.n[0], g.n[2]) + ld) * sizeof(float);
    uint64_t pts = 0;
    for (const auto& p : r.src) pts += (uint64_t)p.npts * (uint64_t)(r.nt + 64) * sizeof(float);
    for (const auto& p : r.rec) pts += 2 * (uint64_t)p.npts * (uint64_t)(r.nt + 16) * sizeof(float);   // + staging
    const uint64_t buckets = 2 * ((uint64_t)(g.n[0] / 64 + 1) * (uint64_t)g.n[2] + 1) * sizeof(int32_t) * r.src.size();
    return (arrays * slab + psi + pts + buckets) * 21 / 20 + (1 << 20);
}

}  // namespace

struct gpi_server {
    gpi_server_opts o{};
    gpi_ctx* cc = nullptr;   // compute stream
    gpi_ctx* ac = nullptr;   // admission stream
    std::vector<void*> grave;     // device memory of departed shots, freed once drained
    uint64_t grave_bytes = 0;
    std::thread th;
    // completion thread: hands finished shots' records to the caller off the scheduler
    std::thread fin_th;
    std::mutex fin_mu;
    std::condition_variable fin_cv;
    std::deque<std::pair<Slot*, int>> fin_q;
    bool fin_stop = false;
    void complete(Slot* s, int code);   // queue a finished (or failed) shot for finalize()
    void run_completions();
    std::mutex mu;
    std::condition_variable cv, cvd;
    bool stop = false;
    bool fatal = false;
    std::deque<std::unique_ptr<Req>> queue;
    std::map<uint64_t, int> status;
    uint64_t next_ticket = 1;
    std::string err;
    // statistics (guarded by mu)
    int64_t submitted = 0, completed = 0, failed = 0, steps = 0, launches = 0;
    int64_t graph_chunks = 0, graph_captures = 0;
    double t_admit = 0.0, t_launch = 0.0, t_retire = 0.0, t_finalize = 0.0, t_collect = 0.0;   // host ms
    // the scheduler's own accumulators, published into t_* under mu once per chunk
    double h_admit = 0.0, h_launch = 0.0, h_retire = 0.0, h_collect = 0.0;
    void publish_times()
    {
        std::lock_guard<std::mutex> lk(mu);
        t_admit = h_admit;
        t_launch = h_launch;
        t_retire = h_retire;
        t_collect = h_collect;
    }
    double busy_ms = 0.0, batch_steps = 0.0;
    std::vector<double> latency;
    std::map<uint64_t, double> lat_of;   // per finished ticket
    int last_ty = 4, last_rx = 1;
    uint64_t used = 0, budget = 0;
    int64_t nactive = 0;
    // scheduler-private state
    std::vector<Slot*> active, finishing;
    PinnedPool pool;
    BatchShot2D* dtab = nullptr;
    int* dfirst = nullptr;
    char* htab[2] = {nullptr, nullptr};
    hipEvent_t tab_ev[2] = {nullptr, nullptr};
    int tab_cur = 0;
    hipEvent_t adm_ev = nullptr;
    hipEvent_t ck_ev[2][2] = {};   // [chunk parity][start, end]
    std::map<int, int> best_shape; // batch-size class -> shape index
    // captured chunks: one per batch layout (shape, shots, workgroups of both passes)
    struct ChunkGraph {
        int shape, nbv, nbs;   // workgroup buckets of the two passes
        hipGraphExec_t exec;
        uint64_t used;
    };
    static constexpr size_t kMaxGraphs = 16;
    static constexpr int kAdmitPerPass = 8;    // shots admitted between two chunk launches
    std::vector<ChunkGraph> graphs;
    std::map<std::tuple<int, int, int>, int> layout_seen;   // chunks run per layout
    std::vector<hipGraphExec_t> graph_grave;   // evicted; destroyed once the streams drain
    uint64_t graph_clock = 0;
    hipStream_t cap = nullptr;                 // capture stream (never runs work)
    int64_t* dstep = nullptr;                  // the batch step the kernels read
    // device memory of both contexts comes from `pool`; freed blocks return to it once an
    // event recorded after their last use on the stream that used them has completed
    DevicePool pool_dev;
    PinnedStage stage;
    struct Pending {
        hipEvent_t ev;
        std::vector<void*> ptrs;
        uint64_t bytes;
    };
    std::deque<Pending> pending;
    uint64_t pending_bytes = 0;
    void defer_grave(hipStream_t st);
    void release_done();

    void run();
    void collect(bool idle);
    int admit(std::unique_ptr<Req> r, Slot** out);
    void retire(Slot* s);
    void finalize(Slot* s, int code);
    void fail_req(uint64_t ticket, int code, const std::string& msg);
    int upload_table(const TileShape& t, int* nbv, int* nbs);
    int chunk_graph(int shape, int bv, int bs, hipGraphExec_t* out);
    void launch_chunk(const TileShape& t, int nbv, int nbs, int64_t K, hipStream_t st);
};

// Blocks freed since the last call were last used by work queued on `st`: they return to
// the pool once an event recorded now has completed.
void gpi_server::defer_grave(hipStream_t st)
{
    if (grave.empty()) return;
    Pending e{nullptr, {}, grave_bytes - pending_bytes};
    if (hipEventCreateWithFlags(&e.ev, hipEventDisableTiming) != hipSuccess || hipEventRecord(e.ev, st) != hipSuccess) {
        if (e.ev) (void)hipEventDestroy(e.ev);
        return;   // left in the graveyard for collect()
    }
    e.ptrs.swap(grave);
    pending_bytes += e.bytes;
    pending.push_back(std::move(e));
}

void gpi_server::release_done()
{
    for (auto it = pending.begin(); it != pending.end();) {
        if (hipEventQuery(it->ev) != hipSuccess) {
            ++it;
            continue;
        }
        for (void* p : it->ptrs) pool_dev.free(p);
        (void)hipEventDestroy(it->ev);
        grave_bytes -= std::min(grave_bytes, it->bytes);
        pending_bytes -= std::min(pending_bytes, it->bytes);
        it = pending.erase(it);
    }
}

// Drain both streams and return every freed block to the pool; when the server is idle
// also give empty pool segments back to the device.
void gpi_server::collect(bool idle)
{
    if (grave.empty() && pending.empty() && graph_grave.empty() && !(idle && pool_dev.reserved())) return;
    const double t0 = now_ms();
    (void)hipStreamSynchronize(ac->stream);
    (void)hipStreamSynchronize(cc->stream);
    for (Pending& e : pending) {
        for (void* p : e.ptrs) pool_dev.free(p);
        (void)hipEventDestroy(e.ev);
    }
    pending.clear();
    pending_bytes = 0;
    for (void* p : grave) pool_dev.free(p);
    grave.clear();
    grave_bytes = 0;
    for (hipGraphExec_t g : graph_grave) (void)hipGraphExecDestroy(g);
    graph_grave.clear();
    if (idle) pool_dev.trim();
    h_collect += now_ms() - t0;
}

// The 2 x K launches of a chunk (kernel offsets 0 .. K - 1) over nbv and nbs workgroups.
void gpi_server::launch_chunk(const TileShape& t, int nbv, int nbs, int64_t K, hipStream_t st)
{
    const int* nsh = dfirst + 2 * (o.max_shots + 1);
    for (int64_t k = 0; k < K; ++k) {
        batch_pass_2d(o.physics, o.order, t, true, dtab, dfirst, nsh, nbv, dstep, (int)k, st);
        batch_pass_2d(o.physics, o.order, t, false, dtab, dfirst + o.max_shots + 1, nsh, nbs, dstep, (int)k, st);
    }
}

// Workgroup count of a captured launch: the smallest of 256, 320, 384, ... (steps of about
// 1/8, multiples of 64) that holds nb, so batches of similar size share one graph.
int launch_bucket(int nb)
{
    int v = 256;
    while (v < nb) v = (v + v / 8 + 63) / 64 * 64;
    return v;
}

// The executable graph of a full chunk for this shape and workgroup buckets, captured when
// the layout runs for the second time (*out stays null the first time: launch it pass by
// pass).  Everything its launches read (table, shot count, step) lives at addresses fixed
// for the server's life and is written before the chunk that needs it.
int gpi_server::chunk_graph(int shape, int bv, int bs, hipGraphExec_t* out)
{
    *out = nullptr;
    for (ChunkGraph& g : graphs)
        if (g.shape == shape && g.nbv == bv && g.nbs == bs) {
            g.used = ++graph_clock;
            *out = g.exec;
            return GPI_OK;
        }
    if (layout_seen[std::make_tuple(shape, bv, bs)]++ == 0) return GPI_OK;
    hipGraph_t gr = nullptr;
    HIPTRY(cc, hipStreamBeginCapture(cap, hipStreamCaptureModeRelaxed));
    launch_chunk(TILE_SHAPES[shape], bv, bs, o.chunk_steps, cap);
    const hipError_t le = hipGetLastError();
    const hipError_t ee = hipStreamEndCapture(cap, &gr);
    if (le != hipSuccess || ee != hipSuccess) {
        if (gr) (void)hipGraphDestroy(gr);
        return fail(cc, GPI_EHIP, std::string("chunk capture: ") + hipGetErrorString(le != hipSuccess ? le : ee));
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (ie != hipSuccess) return fail(cc, GPI_EHIP, std::string("chunk graph instantiate: ") + hipGetErrorString(ie));
    if (graphs.size() >= kMaxGraphs) {
        auto lru = std::min_element(graphs.begin(), graphs.end(),
                                    [](const ChunkGraph& a, const ChunkGraph& b) { return a.used < b.used; });
        graph_grave.push_back(lru->exec);   // may still be queued on the compute stream
        graphs.erase(lru);
    }
    graphs.push_back(ChunkGraph{shape, bv, bs, ex, ++graph_clock});
    {
        std::lock_guard<std::mutex> lk(mu);
        ++graph_captures;
    }
    *out = ex;
    return GPI_OK;
}

void gpi_server::fail_req(uint64_t ticket, int code, const std::string& msg)
{
    std::lock_guard<std::mutex> lk(mu);
    status[ticket] = code;
    err = msg;
    ++failed;
    cvd.notify_all();
}

int gpi_server::admit(std::unique_ptr<Req> r, Slot** out)
{
    Slot* s = new Slot();
    Req& q = *r;
    s->r = std::move(r);
    auto bail = [&](int rc) {
        if (s->S) {
            free_allocs(ac, s->S->p, s->S->bytes);
            delete s->S;
        }
        if (s->m) free_model(ac, s->m);
        s->r.reset();
        delete s;
        return rc;
    };
    int rc = create_model_medium(ac, &q.g, &q.med, q.pml.data(), (int32_t)q.pml.size(), &s->m);
    if (rc) {
        s->m = nullptr;
        return bail(rc);
    }
    s->S = new Scratch(ac);
    Sources src;
    rc = plan_sources(ac, *s->m, *s->S, (int)q.src.size(), q.src.data(), q.wavp.data(), q.nt, false, s->XV, s->XS, src);
    if (rc) return bail(rc);
    if (!src.vfall.empty() || !src.sfall.empty())
        return bail(fail(ac, GPI_EINVAL, "sources on the free-surface row or dense target lists are not batchable"));
    gpi_shot sh{};
    sh.nt = q.nt;
    sh.nrec = (int32_t)q.rec.size();
    sh.rec = q.rec.data();
    sh.records = q.out.data();
    if ((rc = plan_receivers(ac, *s->m, *s->m->st0, *s->S, &sh, s->XV, s->XS, s->rec))) return bail(rc);
    if ((rc = zero_state(ac, *s->m, *s->m->st0, ac->stream))) return bail(rc);
    size_t nrec = 0;
    for (const DevPoints& D : s->rec) nrec += (size_t)q.nt * (size_t)std::max(D.npts, 1);
    // staging of the transposed records plus one int: the non-finite flag of the transpose
    if ((rc = s->S->alloc(&s->trec, nrec + 1))) return bail(rc);
    s->pin = pool.get(nrec + 1, &s->pin_cap);
    if (!s->pin) return bail(fail(ac, GPI_ENOMEM, "pinned host memory for the records"));
    if (hipEventCreateWithFlags(&s->ev, hipEventDisableTiming) != hipSuccess) {
        pool.put(s->pin, s->pin_cap);
        return bail(fail(ac, GPI_EHIP, "hipEventCreate"));
    }
    *out = s;
    return GPI_OK;
}

void gpi_server::retire(Slot* s)
{
    // records -> the caller's time-fastest layout on the device -> pinned staging, with the
    // non-finite flag after them; the device memory goes to the graveyard
    size_t total = 0;
    for (const DevPoints& D : s->rec) total += (size_t)s->r->nt * (size_t)std::max(D.npts, 1);
    int* flag = reinterpret_cast<int*>(s->trec + total);
    (void)hipMemsetAsync(flag, 0, sizeof(int), cc->stream);
    size_t off = 0;
    for (const DevPoints& D : s->rec) {
        const size_t n = (size_t)s->r->nt * (size_t)std::max(D.npts, 1);
        records_transpose(D.rec, s->trec + off, s->r->nt, D.npts, flag, cc->stream);
        off += n;
    }
    (void)hipMemcpyAsync(s->pin, s->trec, sizeof(float) * (total + 1), hipMemcpyDeviceToHost, cc->stream);
    (void)hipEventRecord(s->ev, cc->stream);
    free_allocs(cc, s->S->p, s->S->bytes);
    delete s->S;
    s->S = nullptr;
    free_model(cc, s->m);
    s->m = nullptr;
    defer_grave(cc->stream);
}

void gpi_server::complete(Slot* s, int code)
{
    {
        std::lock_guard<std::mutex> lk(fin_mu);
        fin_q.emplace_back(s, code);
    }
    fin_cv.notify_one();
}

void gpi_server::run_completions()
{
    for (;;) {
        std::pair<Slot*, int> job;
        {
            std::unique_lock<std::mutex> lk(fin_mu);
            fin_cv.wait(lk, [&] { return fin_stop || !fin_q.empty(); });
            if (fin_q.empty()) return;   // stopping and drained
            job = fin_q.front();
            fin_q.pop_front();
        }
        finalize(job.first, job.second);
    }
}

void gpi_server::finalize(Slot* s, int code)
{
    const Req& q = *s->r;
    if (code == GPI_OK) {   // the staging already holds each field in the caller's layout
        const double t0 = now_ms();
        size_t off = 0;
        for (size_t k = 0; k < s->rec.size(); ++k) {
            const int np = s->rec[k].npts;
            if (np > 0) std::memcpy(q.out[k], s->pin + off, sizeof(float) * (size_t)q.nt * (size_t)np);
            off += (size_t)q.nt * (size_t)std::max(np, 1);
        }
        int nonfinite = 0;
        std::memcpy(&nonfinite, s->pin + off, sizeof(int));
        std::lock_guard<std::mutex> lk(mu);
        t_finalize += now_ms() - t0;
        if (nonfinite) {   // records are delivered, the shot reports the error
            code = GPI_ENONFINITE;
            err = "shot " + std::to_string(q.ticket) +
                  " recorded non-finite values (is dt within the stability limit?)";
        }
    }
    pool.put(s->pin, s->pin_cap);
    (void)hipEventDestroy(s->ev);
    {
        std::lock_guard<std::mutex> lk(mu);
        used -= std::min(used, q.est);
        if (code == GPI_OK) {
            status[q.ticket] = 0;
            ++completed;
            latency.push_back(now_ms() - q.t_submit);
            lat_of[q.ticket] = latency.back();
        } else {
            status[q.ticket] = code;
            ++failed;
        }
        cvd.notify_all();
    }
    delete s;
}

int gpi_server::upload_table(const TileShape& t, int* nbv, int* nbs)
{
    const int n = (int)active.size();
    char* buf = htab[tab_cur];
    HIPTRY(cc, hipEventSynchronize(tab_ev[tab_cur]));   // the buffer's previous copy is done
    BatchShot2D* T = (BatchShot2D*)buf;
    int* fv = (int*)(buf + sizeof(BatchShot2D) * o.max_shots);
    int* fs = fv + (o.max_shots + 1);
    int nv = 0, ns = 0;
    for (int k = 0; k < n; ++k) {
        Slot* s = active[k];
        BatchShot2D e{};
        if (!vel_params_2d(*s->m, *s->m->st0, &e.vp) || !str_params_2d(*s->m, *s->m->st0, &e.sp))
            return fail(cc, GPI_EINVAL, "internal error: incomplete wavefield state");
        if (e.vp.nz != (int)s->m->g.n[0] || e.vp.ld != (int)s->m->pitch0 || !s->S)
            return fail(cc, GPI_EINVAL, "internal error: inconsistent shot state");
        velocity_grid_2d(*s->m, t, &e.vp.gx, &e.vp.gy);
        stress_grid_2d(*s->m, t, &e.sp.gx, &e.sp.gy);
        e.xv = s->XV;
        e.xs = s->XS;
        e.start = s->start;
        e.nt = s->r->nt;
        T[k] = e;
        fv[k] = nv;
        fs[k] = ns;
        nv += e.vp.gx * e.vp.gy + e.xv.gather_blocks;
        ns += e.sp.gx * e.sp.gy + e.xs.gather_blocks;
    }
    fv[n] = nv;
    fs[n] = ns;
    fv[2 * (o.max_shots + 1)] = n;   // the shot count the kernels read
    HIPTRY(cc, hipMemcpyAsync(dtab, T, sizeof(BatchShot2D) * n, hipMemcpyHostToDevice, cc->stream));
    HIPTRY(cc, hipMemcpyAsync(dfirst, fv, sizeof(int) * (2 * (o.max_shots + 1) + 1), hipMemcpyHostToDevice,
                              cc->stream));
    HIPTRY(cc, hipEventRecord(tab_ev[tab_cur], cc->stream));
    tab_cur ^= 1;
    *nbv = nv;
    *nbs = ns;
    return GPI_OK;
}

void gpi_server::run()
{
    (void)hipSetDevice(cc->device);
    int64_t step = 0;
    int64_t chunk_no = 0;
    bool dirty = false;
    int nbv = 0, nbs = 0;
    int shape = o.autotune ? 0 : o.tile;
    int table_shape = -1;       // the shape the uploaded table's workgroup grids are for
    // autotuning state: class of the current batch, next candidate to launch (-1: none),
    // per-step times measured for the class, and the candidate each in-flight chunk tests
    int cls = -1, testing = -1;
    double tms[NTILE_SHAPES];
    int ptest_shape[2] = {-1, -1}, ptest_cls[2] = {-1, -1};
    int64_t pend[2] = {0, 0};   // steps of an in-flight chunk not yet accounted
    // device time of a finished chunk (also the autotuner's measurement)
    auto account = [&](hipEvent_t a, hipEvent_t b, int64_t kk) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, a, b) == hipSuccess) {
            std::lock_guard<std::mutex> lk(mu);
            busy_ms += ms;
        }
        return (double)ms / (double)std::max<int64_t>(kk, 1);
    };
    // a finished chunk of parity slot k: its device time, and the autotuner's measurement
    auto account_slot = [&](int k) {
        const double per = account(ck_ev[k][0], ck_ev[k][1], pend[k]);
        pend[k] = 0;
        const int ts = ptest_shape[k];
        ptest_shape[k] = -1;
        if (ts < 0 || ptest_cls[k] != cls || best_shape.count(cls)) return;
        tms[ts] = per;
        int best = 0;
        for (int q = 0; q < NTILE_SHAPES; ++q) {
            if (tms[q] < 0.0) return;   // a candidate still in flight or not launched yet
            if (tms[q] < tms[best]) best = q;
        }
        best_shape[cls] = best;
        if (testing < 0) shape = best;
    };
    auto shot_fail_all = [&](int code, const std::string& msg) {
        for (Slot* s : active) {
            retire(s);
            (void)hipStreamSynchronize(cc->stream);
            complete(s, code);
        }
        active.clear();
        std::lock_guard<std::mutex> lk(mu);
        err = msg;
        fatal = true;
        nactive = 0;
    };
    for (;;) {
        // ---- admission ----------------------------------------------------------------
        std::vector<std::unique_ptr<Req>> take;
        bool blocked = false;
        release_done();
        if (active.empty() && finishing.empty()) collect(true);   // idle: release departed shots
        {
            std::unique_lock<std::mutex> lk(mu);
            if (active.empty() && finishing.empty() && queue.empty()) {
                if (stop) break;
                cv.wait(lk, [&] { return stop || !queue.empty(); });
                if (stop && queue.empty()) break;
            }
            // a few shots per pass, so chunks start while the rest of a backlog is admitted
            while (!queue.empty() && (int)(active.size() + take.size()) < o.max_shots &&
                   (int)take.size() < kAdmitPerPass) {
                const uint64_t need = queue.front()->est;
                if (used + grave_bytes + need > budget && !(active.empty() && take.empty() && used == 0)) {
                    blocked = true;
                    break;
                }
                used += need;
                take.push_back(std::move(queue.front()));
                queue.pop_front();
            }
        }
        // memory of departed shots holds back admission: drain and release it
        if ((blocked && grave_bytes > 0) || grave_bytes > budget / 8 || graph_grave.size() >= kMaxGraphs)
            collect(false);
        bool admitted = false;
        for (auto& r : take) {
            const uint64_t ticket = r->ticket;
            const uint64_t est = r->est;
            Slot* s = nullptr;
            const double t0 = now_ms();
            const int rc = admit(std::move(r), &s);
            defer_grave(ac->stream);
            h_admit += now_ms() - t0;
            if (rc) {
                {
                    std::lock_guard<std::mutex> lk(mu);
                    used -= std::min(used, est);
                }
                fail_req(ticket, rc, ac->err);
                continue;
            }
            s->start = step;
            active.push_back(s);
            admitted = true;
        }
        if (admitted) {
            (void)hipEventRecord(adm_ev, ac->stream);
            (void)hipStreamWaitEvent(cc->stream, adm_ev, 0);
            dirty = true;
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            nactive = (int64_t)active.size();
        }
        // ---- finished shots whose records have arrived --------------------------------
        for (size_t k = 0; k < finishing.size();) {
            Slot* s = finishing[k];
            if (hipEventQuery(s->ev) == hipSuccess) {
                complete(s, GPI_OK);
                finishing.erase(finishing.begin() + k);
            } else {
                ++k;
            }
        }
        if (active.empty()) {
            if (!finishing.empty()) (void)hipEventSynchronize(finishing.front()->ev);
            continue;
        }
        // ---- workgroup shape: measured per batch-size class ----------------------------
        if (dirty && o.autotune) {
            double cells = 0.0;
            for (Slot* s : active) cells += (double)s->m->g.n[0] * (double)s->m->g.n[2];
            const int c = (int)std::floor(std::log2(std::max(cells, 1.0)));   // classes a factor 2 apart
            if (c != cls) {
                cls = c;
                auto it = best_shape.find(c);
                if (it != best_shape.end()) {
                    shape = it->second;
                    testing = -1;
                } else {
                    testing = 0;
                    std::fill(tms, tms + NTILE_SHAPES, -1.0);
                }
            }
        }
        int this_test = -1;
        if (testing >= 0) {   // the next candidate; its time is read when its chunk is accounted
            shape = this_test = testing;
            testing = testing + 1 < NTILE_SHAPES ? testing + 1 : -1;
        }
        // the chunk runs with this shape: accounting an older chunk below may pick the
        // class's best shape, which the next chunk uses
        const int cshape = shape;
        const TileShape t = TILE_SHAPES[cshape];
        const double tl0 = now_ms();
        if (dirty || cshape != table_shape) {   // the workgroup grids depend on the shape
            if (upload_table(t, &nbv, &nbs)) {
                shot_fail_all(GPI_EHIP, cc->err);
                continue;
            }
            dirty = false;
            table_shape = cshape;
        }
        // ---- one chunk of steps ---------------------------------------------------------
        int64_t K = o.chunk_steps;
        for (Slot* s : active) K = std::min<int64_t>(K, s->start + s->r->nt + 1 - step);
        K = std::max<int64_t>(K, 1);
        const int slot = (int)(chunk_no & 1);
        hipEvent_t* E = ck_ev[slot];
        if (pend[slot] > 0) {   // at most two chunks in flight; account the older one
            (void)hipEventSynchronize(E[1]);
            account_slot(slot);
        }
        (void)hipEventRecord(E[0], cc->stream);
        const int ns = (int)active.size();
        batch_set_step(dstep, step, cc->stream);
        bool graphed = false;
        hipGraphExec_t ex = nullptr;
        if (!o.direct_launch && K == o.chunk_steps &&
            chunk_graph(cshape, launch_bucket(nbv), launch_bucket(nbs), &ex)) {
            shot_fail_all(GPI_EHIP, cc->err);
            continue;
        }
        if (ex) {
            const hipError_t ge = hipGraphLaunch(ex, cc->stream);
            if (ge != hipSuccess) {
                shot_fail_all(GPI_EHIP, std::string("chunk graph launch: ") + hipGetErrorString(ge));
                continue;
            }
            graphed = true;
        } else {
            launch_chunk(t, nbv, nbs, K, cc->stream);
        }
        (void)hipEventRecord(E[1], cc->stream);
        const hipError_t le = hipGetLastError();
        if (le != hipSuccess) {
            shot_fail_all(GPI_EHIP, std::string("batched launch: ") + hipGetErrorString(le));
            continue;
        }
        step += K;
        ++chunk_no;
        h_launch += now_ms() - tl0;
        {
            std::lock_guard<std::mutex> lk(mu);
            steps += K;
            launches += 2 * K;
            graph_chunks += graphed ? 1 : 0;
            t_admit = h_admit;
            t_launch = h_launch;
            t_retire = h_retire;
            t_collect = h_collect;
            batch_steps += (double)ns * (double)K;
            last_ty = t.ty;
            last_rx = t.rx;
        }
        pend[slot] = K;
        ptest_shape[slot] = this_test;
        ptest_cls[slot] = cls;
        // ---- shots past their trailing half step leave -----------------------------------
        for (size_t k = 0; k < active.size();) {
            Slot* s = active[k];
            if (step - s->start > s->r->nt) {
                const double t0 = now_ms();
                retire(s);
                h_retire += now_ms() - t0;
                finishing.push_back(s);
                active.erase(active.begin() + k);
                dirty = true;
            } else {
                ++k;
            }
        }
    }
    (void)hipStreamSynchronize(cc->stream);
    for (int k = 0; k < 2; ++k)
        if (pend[k] > 0) account_slot(k);
    collect(true);
    publish_times();
}

extern "C" {

int gpi_server_create(int device, const gpi_server_opts* opts, gpi_server** out)
{
    if (!out || !opts) return GPI_EINVAL;
    *out = nullptr;
    const gpi_server_opts& o = *opts;
    if ((o.physics != 0 && o.physics != 1) || (o.order != 2 && o.order != 4 && o.order != 6 && o.order != 8) ||
        o.max_shots < 0 || o.max_shots > 4096 || o.chunk_steps < 0 || o.chunk_steps > 4096 || o.tile < 0 ||
        o.tile >= NTILE_SHAPES || o.direct_launch < 0 || o.direct_launch > 1 || o.reserved != 0)
        return GPI_EINVAL;
    gpi_server* v = new gpi_server();
    v->o = o;
    if (v->o.max_shots == 0) v->o.max_shots = 64;
    if (v->o.chunk_steps == 0) v->o.chunk_steps = 32;
    int rc = gpi_create(&v->cc, device, 0);
    if (!rc) rc = gpi_create(&v->ac, device, 0);
    if (rc) {
        if (v->cc) gpi_destroy(v->cc);
        delete v;
        return rc;
    }
    v->cc->pool = v->ac->pool = &v->pool_dev;
    v->ac->stage = &v->stage;
    v->cc->defer_free = &v->grave;
    v->cc->defer_bytes = &v->grave_bytes;
    v->ac->defer_free = &v->grave;
    v->ac->defer_bytes = &v->grave_bytes;
    v->cc->budget = 0;
    v->ac->budget = 0;
    size_t fr = 0, tot = 0;
    (void)hipMemGetInfo(&fr, &tot);
    v->budget = o.hbm_budget_bytes ? o.hbm_budget_bytes : (uint64_t)((double)fr * 0.9);
    const size_t tb = sizeof(BatchShot2D) * v->o.max_shots + sizeof(int) * (2 * (v->o.max_shots + 1) + 1);
    bool ok = hipMalloc((void**)&v->dtab, sizeof(BatchShot2D) * v->o.max_shots) == hipSuccess &&
              hipMalloc((void**)&v->dfirst, sizeof(int) * (2 * (v->o.max_shots + 1) + 1)) == hipSuccess &&
              hipMalloc((void**)&v->dstep, sizeof(int64_t)) == hipSuccess &&
              hipStreamCreateWithFlags(&v->cap, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&v->adm_ev, hipEventDisableTiming) == hipSuccess;
    for (int k = 0; k < 2 && ok; ++k) {
        ok = hipHostMalloc((void**)&v->htab[k], tb, hipHostMallocDefault) == hipSuccess &&
             hipEventCreateWithFlags(&v->tab_ev[k], hipEventDisableTiming) == hipSuccess &&
             hipEventCreate(&v->ck_ev[k][0]) == hipSuccess && hipEventCreate(&v->ck_ev[k][1]) == hipSuccess;
        if (ok) ok = hipEventRecord(v->tab_ev[k], v->cc->stream) == hipSuccess;
    }
    if (!ok) {
        gpi_server_destroy(v);
        return GPI_EHIP;
    }
    v->fin_th = std::thread([v] { v->run_completions(); });
    v->th = std::thread([v] { v->run(); });
    *out = v;
    return GPI_OK;
}

void gpi_server_destroy(gpi_server* v)
{
    if (!v) return;
    if (v->th.joinable()) {
        {
            std::lock_guard<std::mutex> lk(v->mu);
            v->stop = true;
        }
        v->cv.notify_all();
        v->th.join();
    }
    if (v->fin_th.joinable()) {   // the scheduler has handed over every shot: drain, then stop
        {
            std::lock_guard<std::mutex> lk(v->fin_mu);
            v->fin_stop = true;
        }
        v->fin_cv.notify_all();
        v->fin_th.join();
    }
    if (v->cc) {
        (void)hipSetDevice(v->cc->device);
        (void)hipStreamSynchronize(v->cc->stream);
    }
    for (auto& g : v->graphs) (void)hipGraphExecDestroy(g.exec);
    for (hipGraphExec_t g : v->graph_grave) (void)hipGraphExecDestroy(g);
    if (v->cap) (void)hipStreamDestroy(v->cap);
    if (v->dtab) (void)hipFree(v->dtab);
    if (v->dfirst) (void)hipFree(v->dfirst);
    if (v->dstep) (void)hipFree(v->dstep);
    for (int k = 0; k < 2; ++k) {
        if (v->htab[k]) (void)hipHostFree(v->htab[k]);
        if (v->tab_ev[k]) (void)hipEventDestroy(v->tab_ev[k]);
        for (int e = 0; e < 2; ++e)
            if (v->ck_ev[k][e]) (void)hipEventDestroy(v->ck_ev[k][e]);
    }
    if (v->adm_ev) (void)hipEventDestroy(v->adm_ev);
    if (v->ac) gpi_destroy(v->ac);
    if (v->cc) gpi_destroy(v->cc);
    if (v->stage.buf) (void)hipHostFree(v->stage.buf);   // its copies ran on the admission stream
    delete v;                                             // the pool's segments go with it
}

const char* gpi_server_error(const gpi_server* v) { return v ? v->err.c_str() : "null server"; }

int gpi_submit(gpi_server* v, const gpi_request* rq, uint64_t* ticket)
{
    if (!v || !rq || !ticket) return GPI_EINVAL;
    const gpi_grid& g = rq->grid;
    const gpi_shot& sh = rq->shot;
    auto bad = [&](const char* msg) {
        std::lock_guard<std::mutex> lk(v->mu);
        v->err = msg;
        return GPI_EINVAL;
    };
    if (g.ndim != 2 || g.physics != v->o.physics || g.order != v->o.order)
        return bad("the request's dimension, physics or order differs from the server's");
    if (sh.nt < 1 || sh.nsrc < 0 || sh.nrec < 0 || sh.nsrc > 4 || sh.nrec > 4 || sh.nfinal != 0 || sh.nsnap != 0 || sh.illum ||
        (sh.nsrc && (!sh.src || !sh.wavelets)) || (sh.nrec && (!sh.rec || !sh.records)))
        return bad("bad shot description");
    const gpi_medium& md = rq->medium;
    if (md.n[0] < 1 || md.n[2] < 1 || !md.vp || !md.rho || (g.physics == 1 && !md.vs))
        return bad("bad medium");
    if (rq->npml_arrays < 0 || (rq->npml_arrays && !rq->pml)) return bad("bad CPML vectors");
    auto r = std::make_unique<Req>();
    r->g = g;
    r->nt = sh.nt;
    const size_t nm = (size_t)(md.n[0] * md.n[2]);
    r->vp.assign(md.vp, md.vp + nm);
    r->rho.assign(md.rho, md.rho + nm);
    if (md.vs) r->vs.assign(md.vs, md.vs + nm);
    r->med = md;
    r->med.n[1] = 1;
    r->med.vp = r->vp.data();
    r->med.rho = r->rho.data();
    r->med.vs = md.vs ? r->vs.data() : nullptr;
    for (int q = 0; q < rq->npml_arrays; ++q) {
        const gpi_pml& p = rq->pml[q];
        if (p.len < 0 || (p.len && (!p.a || !p.b || !p.kI))) return bad("bad CPML vectors");
        r->pmlv.emplace_back(p.a, p.a + p.len);
        r->pmlv.emplace_back(p.b, p.b + p.len);
        r->pmlv.emplace_back(p.kI, p.kI + p.len);
    }
    for (int q = 0; q < rq->npml_arrays; ++q) {
        gpi_pml p = rq->pml[q];
        p.a = r->pmlv[3 * q].data();
        p.b = r->pmlv[3 * q + 1].data();
        p.kI = r->pmlv[3 * q + 2].data();
        r->pml.push_back(p);
    }
    auto copy_points = [&](const gpi_points& p, std::vector<gpi_points>& dst) {
        const size_t n = (size_t)std::max(p.npts, 0) * (size_t)std::max(p.nnz, 0);
        if (n && (!p.idx || !p.w)) return false;
        r->idx.emplace_back(p.idx, p.idx + n);
        r->w.emplace_back(p.w, p.w + n);
        dst.push_back(p);
        return true;
    };
    for (int q = 0; q < sh.nsrc; ++q) {
        if (!copy_points(sh.src[q], r->src) || (sh.src[q].npts > 0 && !sh.wavelets[q])) return bad("bad source");
        r->wav.emplace_back(sh.wavelets[q], sh.wavelets[q] + (size_t)sh.src[q].npts * (size_t)sh.nt);
    }
    for (int q = 0; q < sh.nrec; ++q) {
        if (!copy_points(sh.rec[q], r->rec) || !sh.records[q]) return bad("bad receiver");
        r->out.push_back(sh.records[q]);
    }
    size_t k = 0;
    for (auto& p : r->src) {
        p.idx = r->idx[k].data();
        p.w = r->w[k].data();
        ++k;
    }
    for (auto& p : r->rec) {
        p.idx = r->idx[k].data();
        p.w = r->w[k].data();
        ++k;
    }
    for (auto& wv : r->wav) r->wavp.push_back(wv.data());
    r->est = estimate_bytes(*r);
    r->t_submit = now_ms();
    {
        std::lock_guard<std::mutex> lk(v->mu);
        if (v->stop || v->fatal) {
            v->err = "the server is stopped";
            return GPI_EINVAL;
        }
        if (r->est > v->budget) {
            v->err = "the shot does not fit the server's HBM budget";
            return GPI_ENOMEM;
        }
        r->ticket = v->next_ticket++;
        *ticket = r->ticket;
        v->status[r->ticket] = 1;
        ++v->submitted;
        v->queue.push_back(std::move(r));
    }
    v->cv.notify_all();
    return GPI_OK;
}

int gpi_poll(gpi_server* v, uint64_t ticket, int32_t* st)
{
    if (!v || !st) return GPI_EINVAL;
    std::lock_guard<std::mutex> lk(v->mu);
    auto it = v->status.find(ticket);
    if (it == v->status.end()) return GPI_ENOTFOUND;
    *st = it->second;
    return GPI_OK;
}

int gpi_wait(gpi_server* v, uint64_t ticket, double timeout_s, int32_t* st)
{
    if (!v || !st) return GPI_EINVAL;
    std::unique_lock<std::mutex> lk(v->mu);
    auto it = v->status.find(ticket);
    if (it == v->status.end()) return GPI_ENOTFOUND;
    auto pred = [&] { return v->status[ticket] != 1; };
    if (timeout_s > 0) v->cvd.wait_for(lk, std::chrono::duration<double>(timeout_s), pred);
    else v->cvd.wait(lk, pred);
    *st = v->status[ticket];
    return GPI_OK;
}

int gpi_latency(gpi_server* v, uint64_t ticket, double* ms)
{
    if (!v || !ms) return GPI_EINVAL;
    std::lock_guard<std::mutex> lk(v->mu);
    auto it = v->lat_of.find(ticket);
    if (it == v->lat_of.end()) return GPI_ENOTFOUND;
    *ms = it->second;
    return GPI_OK;
}

int gpi_server_get_stats(gpi_server* v, gpi_server_stats* out)
{
    if (!v || !out) return GPI_EINVAL;
    std::lock_guard<std::mutex> lk(v->mu);
    gpi_server_stats s{};
    s.submitted = v->submitted;
    s.completed = v->completed;
    s.failed = v->failed;
    s.active = v->nactive;
    s.queued = (int64_t)v->queue.size();
    s.steps = v->steps;
    s.launches = v->launches;
    s.graph_chunks = v->graph_chunks;
    s.graph_captures = v->graph_captures;
    s.host_admit_ms = v->t_admit;
    s.host_launch_ms = v->t_launch;
    s.host_retire_ms = v->t_retire;
    s.host_finalize_ms = v->t_finalize;
    s.host_collect_ms = v->t_collect;
    s.pool_bytes = v->pool_dev.reserved();
    s.busy_ms = v->busy_ms;
    s.mean_batch = v->steps ? v->batch_steps / (double)v->steps : 0.0;
    s.tile_ty = v->last_ty;
    s.tile_rx = v->last_rx;
    s.hbm_used = v->used;
    s.hbm_budget = v->budget;
    if (!v->latency.empty()) {
        std::vector<double> l = v->latency;
        std::sort(l.begin(), l.end());
        auto pct = [&](double p) {
            const double x = p * (double)(l.size() - 1);
            const size_t i = (size_t)x;
            const double f = x - (double)i;
            return i + 1 < l.size() ? l[i] * (1.0 - f) + l[i + 1] * f : l[i];
        };
        s.p50_ms = pct(0.50);
        s.p99_ms = pct(0.99);
    }
    *out = s;
    return GPI_OK;
}

}  // extern "C"
