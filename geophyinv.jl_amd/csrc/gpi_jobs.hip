// Job queue: forward and gradient jobs of any dimension and physics on one GPU.
//
// Worker threads, each with its own context (stream), take jobs in arrival order when
// their estimated device memory fits the budget next to the running jobs; a worker runs
// its job with the library's own entry points (gpi_model_create, gpi_forward or
// gpi_forward_checkpoint + residuals + the checkpointed adjoint), so the device work,
// host work and copies of concurrent jobs overlap while each result stays that of the
// plain calls.
//
// No reference counterpart: the reference runs supersources one after another per worker
// (src/fdtd/propagate.jl:100-106); its misfit and adjoint source are database.jl:357-380.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

#include "gpi_util.h"

using namespace gpi;

namespace {

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Job {
    uint64_t ticket = 0;
    gpi_job j{};
    uint64_t est = 0;
    double t_submit = 0.0;
};

double pct(std::vector<double> v, double q)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double r = q / 100.0 * (double)(v.size() - 1);
    const size_t a = (size_t)std::floor(r), b = std::min(a + 1, v.size() - 1);
    return v[a] + (r - (double)a) * (v[b] - v[a]);
}

// Device bytes of a job: model coefficients and CPML vectors, one wavefield state, the
// record and source tables, and for a gradient job the checkpoints, the segment buffers,
// the adjoint state and the FP64 accumulators.
uint64_t job_bytes(const gpi_job& j)
{
    const gpi_grid& g = j.grid;
    const int h = g.order - 1;
    const bool d3 = g.ndim == 3;
    const uint64_t ld = (uint64_t)round_up(g.n[0] + h, 32);
    const uint64_t slab = 4 * ld * (uint64_t)(d3 ? g.n[1] + h : 1) * (uint64_t)(g.n[2] + h);
    const uint64_t nw = g.physics == 0 ? (d3 ? 4 : 3) : (d3 ? 9 : 5);   // wavefields
    const uint64_t nder = g.physics == 0 ? (d3 ? 6 : 4) : (d3 ? 18 : 8);
    uint64_t face = 0;
    for (int a = 0; a < 3; ++a) {
        if (!d3 && a == 1) continue;
        uint64_t f = 1;
        for (int b = 0; b < 3; ++b)
            if (b != a && (d3 || b != 1)) f *= (uint64_t)(g.n[b] + h);
        face = std::max(face, f);
    }
    const uint64_t state = nw * slab + nder * 2 * (uint64_t)std::max(g.npml, 0) * face * 4;
    uint64_t pts = 0;
    for (int q = 0; q < j.shot.nsrc; ++q) pts += (uint64_t)j.shot.src[q].npts * (uint64_t)(j.shot.nt + 64) * 4;
    for (int q = 0; q < j.shot.nrec; ++q) pts += (uint64_t)j.shot.rec[q].npts * (uint64_t)(j.shot.nt + 16) * 4;
    uint64_t b = (uint64_t)std::max(j.ncoefs, 0) * slab + state + 2 * pts;
    if (j.checkpoint > 0) {
        const uint64_t nt = (uint64_t)j.shot.nt, C = (uint64_t)j.checkpoint;
        const uint64_t K = (nt + C - 1) / C;
        const uint64_t L = std::min(C, nt);
        const uint64_t segf = g.physics == 0 ? 1 + (d3 ? 3 : 2) : nw;
        const uint64_t acc = g.physics == 0 ? 2 : (d3 ? 8 : 5);
        b += (K > 0 ? K - 1 : 0) * state + (L + 1) * segf * slab + state + (acc + 3) * 2 * slab;
    }
    return b * 21 / 20 + (1 << 20);
}

int run_job(gpi_ctx* c, const gpi_job& j)
{
    uint64_t model = 0;
    int rc = gpi_model_create(c, &j.grid, j.coefs, j.ncoefs, j.pml, j.npml_arrays, &model);
    if (rc) return rc;
    if (j.checkpoint <= 0) {
        rc = gpi_forward(c, model, &j.shot);
        (void)gpi_model_free(c, model);
        return rc;
    }
    uint64_t store = 0;
    rc = gpi_forward_checkpoint(c, model, &j.shot, j.checkpoint, &store);
    if (rc) {
        (void)gpi_model_free(c, model);
        return rc;
    }
    // misfit and adjoint sources from the records (database.jl:357-380)
    const int64_t nt = j.shot.nt;
    double J = 0.0;
    std::vector<std::vector<float>> adj((size_t)j.shot.nrec);
    std::vector<const float*> wav2((size_t)j.shot.nrec);
    for (int k = 0; k < j.shot.nrec; ++k) {
        const size_t np = (size_t)j.shot.rec[k].npts;
        const float* dc = j.shot.records[k];
        const float* ob = j.dobs[k];
        std::vector<float>& a = adj[(size_t)k];
        a.resize(np * (size_t)nt);
        for (size_t r = 0; r < np; ++r)
            for (int64_t it = 0; it < nt; ++it) {
                const size_t o = r * (size_t)nt + (size_t)it;
                const double d = (double)ob[o] - (double)dc[o];
                J += d * d;
                // 2 (dobs - dcal) in FP32, reversed in time
                a[r * (size_t)nt + (size_t)(nt - 1 - it)] = (ob[o] - dc[o]) * 2.0f;
            }
        wav2[(size_t)k] = a.data();
    }
    gpi_adjoint_shot sh{};
    sh.nt = nt;
    sh.nsrc1 = 0;
    sh.nsrc2 = j.shot.nrec;
    sh.src2 = j.shot.rec;
    sh.wav2 = wav2.data();
    if (j.grid.physics == 1) rc = gpi_adjoint_checkpoint_elastic(c, model, store, &sh, j.grad[0], j.grad[1], j.grad[2]);
    else rc = gpi_adjoint_checkpoint(c, model, store, &sh, j.dtI, j.grad[0], j.grad[1]);
    if (!rc && j.misfit) *j.misfit = J;
    (void)gpi_store_free(c, store);
    (void)gpi_model_free(c, model);
    return rc;
}

}  // namespace

struct gpi_jobs {
    int device = 0;
    gpi_jobs_opts o{};
    uint64_t budget = 0, used = 0;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv, cvd;
    std::deque<Job> queue;
    std::map<uint64_t, int> status;
    uint64_t next = 1;
    bool stop = false;
    std::string err;
    int64_t submitted = 0, completed = 0, failed = 0, running = 0, fwd_done = 0, grad_done = 0;
    std::vector<double> lat, lat_f, lat_g;

    void work()
    {
        gpi_ctx* c = nullptr;
        const int crc = gpi_create(&c, device, 0);
        for (;;) {
            Job job;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] {
                    if (queue.empty()) return stop;
                    return used + queue.front().est <= budget || running == 0;   // FIFO, by memory
                });
                if (queue.empty()) break;
                job = queue.front();
                queue.pop_front();
                used += job.est;
                ++running;
            }
            const int rc = crc ? crc : run_job(c, job.j);
            std::lock_guard<std::mutex> lk(mu);
            used -= std::min(used, job.est);
            --running;
            status[job.ticket] = rc ? rc : 0;
            if (rc) {
                ++failed;
                err = c ? c->err : "no HIP device";
            } else {
                ++completed;
                const double ms = now_ms() - job.t_submit;
                lat.push_back(ms);
                (job.j.checkpoint > 0 ? lat_g : lat_f).push_back(ms);
                ++(job.j.checkpoint > 0 ? grad_done : fwd_done);
            }
            cvd.notify_all();
            cv.notify_all();
        }
        if (c) gpi_destroy(c);
    }
};

extern "C" {

int gpi_jobs_create(int device, const gpi_jobs_opts* opts, gpi_jobs** out)
{
    if (!out || !opts) return GPI_EINVAL;
    *out = nullptr;
    if (opts->workers < 0 || opts->workers > 16 || opts->reserved != 0) return GPI_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return GPI_EHIP;
    if (device < 0 || device >= ndev) return GPI_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return GPI_EHIP;
    gpi_jobs* q = new gpi_jobs();
    q->device = device;
    q->o = *opts;
    if (q->o.workers == 0) q->o.workers = 2;
    size_t fr = 0, tot = 0;
    (void)hipMemGetInfo(&fr, &tot);
    q->budget = opts->hbm_budget_bytes ? opts->hbm_budget_bytes : (uint64_t)((double)fr * 0.9);
    for (int k = 0; k < q->o.workers; ++k) q->workers.emplace_back([q] { q->work(); });
    *out = q;
    return GPI_OK;
}

void gpi_jobs_destroy(gpi_jobs* q)
{
    if (!q) return;
    {
        std::lock_guard<std::mutex> lk(q->mu);
        q->stop = true;
    }
    q->cv.notify_all();
    for (auto& t : q->workers) t.join();
    delete q;
}

const char* gpi_jobs_error(const gpi_jobs* q) { return q ? q->err.c_str() : "null job queue"; }

int gpi_jobs_submit(gpi_jobs* q, const gpi_job* job, uint64_t* ticket)
{
    if (!q || !job || !ticket) return GPI_EINVAL;
    const gpi_job& j = *job;
    auto bad = [&](const char* msg) {
        std::lock_guard<std::mutex> lk(q->mu);
        q->err = msg;
        return GPI_EINVAL;
    };
    if (j.shot.nt < 1 || j.shot.nfinal != 0 || j.shot.nsnap != 0 || j.shot.illum || j.shot.nrec < 0 || j.shot.nsrc < 0 ||
        (j.shot.nrec > 0 && (!j.shot.rec || !j.shot.records)) || (j.shot.nsrc > 0 && (!j.shot.src || !j.shot.wavelets)))
        return bad("bad shot (nt, points, records; snapshots, final fields and illumination are not served)");
    if (j.checkpoint < 0) return bad("checkpoint must be >= 0");
    if (j.checkpoint > 0) {
        if (j.shot.nrec < 1 || !j.dobs || !j.grad[0] || (j.grid.physics == 1 && (!j.grad[1] || !j.grad[2])))
            return bad("a gradient job needs receivers, observed records and gradient outputs");
        for (int k = 0; k < j.shot.nrec; ++k)
            if (!j.dobs[k]) return bad("missing observed records");
        if (j.grid.physics == 0 && !(j.dtI > 0.0f)) return bad("an acoustic gradient job needs dtI > 0");
    }
    const uint64_t est = job_bytes(j);
    std::lock_guard<std::mutex> lk(q->mu);
    if (est > q->budget) {
        q->err = "the job needs more device memory than the whole budget";
        return GPI_ENOMEM;
    }
    Job x;
    x.ticket = q->next++;
    x.j = j;
    x.est = est;
    x.t_submit = now_ms();
    q->status[x.ticket] = 1;
    q->queue.push_back(x);
    ++q->submitted;
    *ticket = x.ticket;
    q->cv.notify_all();
    return GPI_OK;
}

int gpi_jobs_wait(gpi_jobs* q, uint64_t ticket, double timeout_s, int32_t* st)
{
    if (!q || !st) return GPI_EINVAL;
    std::unique_lock<std::mutex> lk(q->mu);
    auto it = q->status.find(ticket);
    if (it == q->status.end()) return GPI_ENOTFOUND;
    auto pred = [&] { return q->status[ticket] != 1; };
    if (timeout_s > 0) q->cvd.wait_for(lk, std::chrono::duration<double>(timeout_s), pred);
    else q->cvd.wait(lk, pred);
    *st = q->status[ticket];
    return GPI_OK;
}

int gpi_jobs_get_stats(gpi_jobs* q, gpi_jobs_stats* out)
{
    if (!q || !out) return GPI_EINVAL;
    std::lock_guard<std::mutex> lk(q->mu);
    gpi_jobs_stats s{};
    s.submitted = q->submitted;
    s.completed = q->completed;
    s.failed = q->failed;
    s.queued = (int64_t)q->queue.size();
    s.running = q->running;
    s.forward_done = q->fwd_done;
    s.gradient_done = q->grad_done;
    s.p50_ms = pct(q->lat, 50.0);
    s.p99_ms = pct(q->lat, 99.0);
    s.forward_p50_ms = pct(q->lat_f, 50.0);
    s.gradient_p50_ms = pct(q->lat_g, 50.0);
    s.hbm_used = q->used;
    s.hbm_budget = q->budget;
    *out = s;
    return GPI_OK;
}

}  // extern "C"
