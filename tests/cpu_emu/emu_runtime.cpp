// The CPU HIP runtime of tests/cpu_emu (see include/hip/hip_runtime.h).
//
// Streams: one worker thread each, running its queue in order; work queued on stream 0
// goes to a stream of its own.  Events carry a record count and a completed count, so
// hipEventQuery/Synchronize and hipStreamWaitEvent wait for the record queued last before
// the call, as HIP's do.  Kernels: the stream's worker splits the grid into workgroups
// that it and a shared pool of host threads run; within a workgroup the threads run one
// after another, each on a fiber of its own so that a barrier can suspend it.  Memory: every allocation is its own mapping, placed against a guard page
// after it (and one before); hipFree/hipHostFree wait for all streams (HIP's frees
// synchronise the device) and hipFreeAsync waits for its stream position, then the pages
// are made inaccessible and never mapped again, so a late access faults.  A fault prints
// the address, the kernel and workgroup (or host call) that made it and the allocation it
// falls in, then ends the process with status 86.  Blocking hipMemcpy runs at once on the
// calling thread: it is not ordered with the non-blocking streams the library creates.
#include <hip/hip_runtime.h>

#include <execinfo.h>
#include <signal.h>
#include <ucontext.h>
#include <sys/mman.h>
#include <unistd.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

thread_local dim3 threadIdx __attribute__((tls_model("initial-exec")));
thread_local dim3 blockIdx __attribute__((tls_model("initial-exec")));
thread_local dim3 blockDim __attribute__((tls_model("initial-exec")));
thread_local dim3 gridDim __attribute__((tls_model("initial-exec")));

namespace emu {

using Op = std::shared_ptr<std::function<void()>>;

struct Graph {
    std::vector<Op> ops;
};

struct Stream {
    std::mutex mu;
    std::condition_variable cv, idle;
    std::deque<Op> q;
    bool busy = false, stop = false;
    Graph* cap = nullptr;
    int id = 0;
    std::thread th;
};

struct Event {
    std::mutex mu;
    std::condition_variable cv;
    uint64_t rec = 0, done = 0;
    double t_ms = 0.0;
};

// ---- diagnostics -------------------------------------------------------------------------

thread_local const char* tl_kernel = nullptr;
thread_local uint64_t tl_block = 0;
thread_local int tl_stream = -1;

struct Alloc {
    char* map = nullptr;     // mapping start (leading guard page)
    size_t map_len = 0;
    char* user = nullptr;    // pointer handed out
    size_t size = 0;         // bytes asked
    const char* kind = "";
    bool freed = false;
};

std::mutex g_mem_mu;
std::map<uintptr_t, Alloc> g_allocs;   // by mapping start
size_t g_live_device = 0;
const size_t PAGE = 4096;

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

void fault_handler(int sig, siginfo_t* si, void*)
{
    const uintptr_t a = (uintptr_t)si->si_addr;
    char buf[512];
    int n = snprintf(buf, sizeof buf, "EMU FAULT: signal %d at address %p in %s (workgroup %llu, stream %d)\n", sig,
                     si->si_addr, tl_kernel ? tl_kernel : "host code", (unsigned long long)tl_block, tl_stream);
    (void)!write(2, buf, (size_t)n);
    auto it = g_allocs.upper_bound(a);   // best effort: the process is ending
    if (it != g_allocs.begin()) {
        --it;
        const Alloc& A = it->second;
        if (a < (uintptr_t)A.map + A.map_len) {
            const long off = (long)(a - (uintptr_t)A.user);
            n = snprintf(buf, sizeof buf, "EMU FAULT: %s allocation %p of %zu bytes, offset %ld, %s\n", A.kind,
                         (void*)A.user, A.size, off,
                         A.freed ? "ALREADY FREED" : (off < 0 ? "before its start" : "past its end"));
            (void)!write(2, buf, (size_t)n);
        }
    }
    void* frames[32];
    const int nf = backtrace(frames, 32);
    backtrace_symbols_fd(frames, nf, 2);
    _exit(86);
}

struct Init {
    Init()
    {
        static std::vector<char> alt(1 << 16);
        stack_t ss{};
        ss.ss_sp = alt.data();
        ss.ss_size = alt.size();
        sigaltstack(&ss, nullptr);
        struct sigaction sa {};
        sa.sa_sigaction = fault_handler;
        sa.sa_flags = SA_SIGINFO | SA_ONSTACK;
        sigaction(SIGSEGV, &sa, nullptr);
        sigaction(SIGBUS, &sa, nullptr);
    }
};
Init g_init;

// ---- memory ------------------------------------------------------------------------------

void* map_guarded(size_t bytes, const char* kind)
{
    const size_t sz = (std::max<size_t>(bytes, 1) + 255) / 256 * 256;
    const size_t data = (sz + PAGE - 1) / PAGE * PAGE;
    const size_t len = data + 2 * PAGE;
    char* m = (char*)mmap(nullptr, len, PROT_NONE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
    if (m == MAP_FAILED) return nullptr;
    if (mprotect(m + PAGE, data, PROT_READ | PROT_WRITE) != 0) {
        munmap(m, len);
        return nullptr;
    }
    Alloc A;
    A.map = m;
    A.map_len = len;
    A.user = m + PAGE + data - sz;   // the rounded end abuts the trailing guard page
    A.size = bytes;
    A.kind = kind;
    std::lock_guard<std::mutex> lk(g_mem_mu);
    g_allocs[(uintptr_t)m] = A;
    if (kind[0] == 'd') g_live_device += sz;
    return A.user;
}

Alloc* find_alloc(const void* p)
{
    auto it = g_allocs.upper_bound((uintptr_t)p);
    if (it == g_allocs.begin()) return nullptr;
    --it;
    return it->second.user == p ? &it->second : nullptr;
}

// Make an allocation inaccessible for good (its address range is never mapped again).
bool unmap_guarded(void* p)
{
    std::lock_guard<std::mutex> lk(g_mem_mu);
    Alloc* A = find_alloc(p);
    if (!A || A->freed) return false;
    A->freed = true;
    if (A->kind[0] == 'd') g_live_device -= (std::max<size_t>(A->size, 1) + 255) / 256 * 256;
    mprotect(A->map + PAGE, A->map_len - 2 * PAGE, PROT_NONE);
    madvise(A->map + PAGE, A->map_len - 2 * PAGE, MADV_DONTNEED);
    return true;
}

// ---- streams -----------------------------------------------------------------------------

std::mutex g_streams_mu;
std::vector<Stream*> g_streams;
Stream* g_null = nullptr;
std::atomic<int> g_next_id{0};

void worker(Stream* s)
{
    tl_stream = s->id;
    for (;;) {
        Op op;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] { return s->stop || !s->q.empty(); });
            if (s->q.empty()) return;
            op = s->q.front();
            s->q.pop_front();
            s->busy = true;
        }
        (*op)();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->busy = false;
            if (s->q.empty()) s->idle.notify_all();
        }
    }
}

Stream* new_stream()
{
    Stream* s = new Stream();
    s->id = g_next_id++;
    s->th = std::thread(worker, s);
    std::lock_guard<std::mutex> lk(g_streams_mu);
    g_streams.push_back(s);
    return s;
}

Stream* resolve(hipStream_t s)
{
    if (s) return s;
    std::lock_guard<std::mutex> lk(g_streams_mu);
    if (!g_null) {
        g_null = new Stream();
        g_null->id = g_next_id++;
        g_null->th = std::thread(worker, g_null);
        g_streams.push_back(g_null);
    }
    return g_null;
}

void enqueue(hipStream_t hs, std::function<void()> fn)
{
    Stream* s = resolve(hs);
    Op op = std::make_shared<std::function<void()>>(std::move(fn));
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->cap) {
        s->cap->ops.push_back(op);
        return;
    }
    s->q.push_back(op);
    s->cv.notify_one();
}

void sync_stream(Stream* s)
{
    std::unique_lock<std::mutex> lk(s->mu);
    s->idle.wait(lk, [&] { return s->q.empty() && !s->busy; });
}

void sync_all()
{
    std::vector<Stream*> ss;
    {
        std::lock_guard<std::mutex> lk(g_streams_mu);
        ss = g_streams;
    }
    for (Stream* s : ss) sync_stream(s);
}

// ---- kernels -----------------------------------------------------------------------------

struct KernelOp {
    void (*run)(void*);
    void* ctx;   // the launch's arguments, owned (dropped with the last reference)
    void (*drop)(void*);
    dim3 grid, block;
    const char* name;
    KernelOp(void (*r)(void*), void* c, void (*d)(void*), dim3 g, dim3 b, const char* n)
        : run(r), ctx(c), drop(d), grid(g), block(b), name(n)
    {
    }
    KernelOp(const KernelOp&) = delete;
    KernelOp& operator=(const KernelOp&) = delete;
    ~KernelOp() { drop(ctx); }
};

struct GridJob {
    const KernelOp* k;
    uint64_t total;
    std::atomic<uint64_t> next{0};
    std::atomic<uint64_t> done{0};
    int users = 0;   // pool threads working on it (under g_pool_mu)
    int stream = -1;
};

// Never destroyed: the pool threads still wait on them when the process exits (destroying
// a condition variable with waiters blocks).
std::mutex& g_pool_mu = *new std::mutex;
std::condition_variable& g_pool_cv = *new std::condition_variable;
std::condition_variable& g_job_cv = *new std::condition_variable;
std::deque<GridJob*>& g_jobs = *new std::deque<GridJob*>;
std::vector<std::thread> g_pool;
std::once_flag g_pool_once;

void run_blocks(GridJob* j, uint64_t b0, uint64_t b1)
{
    const KernelOp& k = *j->k;
    tl_kernel = k.name;
    blockDim = k.block;
    gridDim = k.grid;
    const uint64_t gx = k.grid.x, gy = k.grid.y;
    for (uint64_t b = b0; b < b1; ++b) {
        tl_block = b;
        blockIdx = dim3((uint32_t)(b % gx), (uint32_t)((b / gx) % gy), (uint32_t)(b / (gx * gy)));
        k.run(k.ctx);
    }
    tl_kernel = nullptr;
}

// Take chunks of workgroups until none is left; returns when this caller's share is done.
void drain(GridJob* j)
{
    const uint64_t chunk = std::max<uint64_t>(1, std::min<uint64_t>(64, j->total / 64 + 1));
    for (;;) {
        const uint64_t b0 = j->next.fetch_add(chunk);
        if (b0 >= j->total) return;
        const uint64_t b1 = std::min(j->total, b0 + chunk);
        run_blocks(j, b0, b1);
        if (j->done.fetch_add(b1 - b0) + (b1 - b0) == j->total) {
            std::lock_guard<std::mutex> lk(g_pool_mu);
            g_job_cv.notify_all();
        }
    }
}

void pool_thread()
{
    for (;;) {
        GridJob* j = nullptr;
        {
            std::unique_lock<std::mutex> lk(g_pool_mu);
            g_pool_cv.wait(lk, [&] { return !g_jobs.empty(); });
            j = g_jobs.front();
            if (j->next.load() >= j->total) {   // nothing left to take: retire it from the list
                g_jobs.pop_front();
                continue;
            }
            ++j->users;
        }
        tl_stream = j->stream;
        drain(j);
        {
            std::lock_guard<std::mutex> lk(g_pool_mu);
            --j->users;
            g_job_cv.notify_all();
        }
    }
}

void start_pool()
{
    const char* e = getenv("EMU_THREADS");
    int n = e ? atoi(e) : (int)std::thread::hardware_concurrency() - 1;
    n = std::max(0, std::min(n, 256));
    for (int k = 0; k < n; ++k) {
        g_pool.emplace_back(pool_thread);
        g_pool.back().detach();
    }
}

void run_grid(const KernelOp& k)
{
    std::call_once(g_pool_once, start_pool);
    GridJob j;
    j.k = &k;
    j.total = (uint64_t)k.grid.x * k.grid.y * k.grid.z;
    j.stream = tl_stream;
    if (j.total == 0) return;
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        g_jobs.push_back(&j);
    }
    g_pool_cv.notify_all();
    drain(&j);
    std::unique_lock<std::mutex> lk(g_pool_mu);
    for (auto it = g_jobs.begin(); it != g_jobs.end(); ++it)
        if (*it == &j) {
            g_jobs.erase(it);
            break;
        }
    g_job_cv.wait(lk, [&] { return j.done.load() == j.total && j.users == 0; });
}

void launch(hipStream_t s, dim3 grid, dim3 block, const char* name, void (*run)(void*), void* ctx,
            void (*drop)(void*))
{
    auto k = std::make_shared<KernelOp>(run, ctx, drop, grid, block, name);
    static const bool trace = getenv("EMU_TRACE") != nullptr;
    enqueue(s, [k] {
        if (trace) fprintf(stderr, "emu: stream %d runs %s on %u x %u x %u workgroups of %u x %u\n", tl_stream, k->name,
                           k->grid.x, k->grid.y, k->grid.z, k->block.x, k->block.y);
        run_grid(*k);
    });
}

// ---- workgroup fibers ---------------------------------------------------------------------
// Every thread of a workgroup runs on a fiber (a stack of its own on the worker's OS
// thread).  The worker resumes the unfinished fibers in thread order; each runs until it
// returns or reaches a barrier.  When a sweep over the fibers ends with fibers at the
// barrier, every thread that has not returned is there, so the next sweep releases them.
// Workgroup memory (__shared__, a thread_local of the OS thread) is therefore shared by
// the fibers.

extern "C" void emu_fiber_switch(void** save_sp, void* next_sp);
// Saves the callee-saved registers of the System V x86-64 ABI on the current stack, stores
// the stack pointer in *save_sp and resumes the fiber whose saved stack pointer is next_sp.
asm(R"(
    .text
    .p2align 4
    .globl emu_fiber_switch
    .type emu_fiber_switch, @function
emu_fiber_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
    .size emu_fiber_switch, .-emu_fiber_switch
)");

struct Fiber {
    void* sp = nullptr;
    dim3 tid;
    bool done = false;
};

struct FiberSet {
    std::vector<Fiber> f;
    std::vector<char*> stacks;   // FIBER_STACK bytes each, above a guard page
    std::vector<char> alt;       // this OS thread's signal stack (a fault on a fiber stack)
    void* sched = nullptr;       // the worker's own stack while a fiber runs
    int cur = -1;                // running fiber, -1 outside a workgroup
    void (*one)(const void*) = nullptr;
    const void* ctx = nullptr;
};

thread_local FiberSet tl_fibers;
constexpr size_t FIBER_STACK = 256 * 1024;

[[noreturn]] void fiber_main()
{
    FiberSet& w = tl_fibers;
    w.one(w.ctx);
    Fiber& me = w.f[(size_t)w.cur];
    me.done = true;
    emu_fiber_switch(&me.sp, w.sched);
    __builtin_trap();   // a finished fiber is never resumed
}

void run_workgroup(void (*one)(const void*), const void* ctx)
{
    FiberSet& w = tl_fibers;
    if (w.alt.empty()) {
        w.alt.resize(1 << 16);
        stack_t ss{};
        ss.ss_sp = w.alt.data();
        ss.ss_size = w.alt.size();
        sigaltstack(&ss, nullptr);
    }
    const size_t n = (size_t)blockDim.x * blockDim.y * blockDim.z;
    while (w.stacks.size() < n) {
        char* m = (char*)mmap(nullptr, FIBER_STACK + PAGE, PROT_READ | PROT_WRITE,
                              MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
        if (m == MAP_FAILED || mprotect(m, PAGE, PROT_NONE) != 0) {
            fprintf(stderr, "EMU: no memory for the fibers of %s\n", tl_kernel);
            _exit(88);
        }
        w.stacks.push_back(m + PAGE);
    }
    w.f.assign(n, Fiber{});
    size_t t = 0;
    for (uint32_t tz = 0; tz < blockDim.z; ++tz)
        for (uint32_t ty = 0; ty < blockDim.y; ++ty)
            for (uint32_t tx = 0; tx < blockDim.x; ++tx, ++t) {
                void** sp = (void**)(w.stacks[t] + FIBER_STACK);
                *--sp = nullptr;                 // return address of fiber_main (never used)
                *--sp = (void*)&fiber_main;      // where the first switch returns to
                for (int r = 0; r < 6; ++r) *--sp = nullptr;   // rbp rbx r12 r13 r14 r15
                w.f[t].sp = sp;
                w.f[t].tid = dim3(tx, ty, tz);
            }
    w.one = one;
    w.ctx = ctx;
    // EMU_REVERSE=1 resumes the fibers last thread first, so a write that a barrier does not
    // order before another thread's read of the same workgroup memory shows up as well
    static const bool reverse = getenv("EMU_REVERSE") && atoi(getenv("EMU_REVERSE")) != 0;
    for (;;) {
        bool waiting = false;
        for (size_t u = 0; u < n; ++u) {
            t = reverse ? n - 1 - u : u;
            Fiber& f = w.f[t];
            if (f.done) continue;
            w.cur = (int)t;
            threadIdx = f.tid;
            emu_fiber_switch(&w.sched, f.sp);
            waiting |= !f.done;
        }
        if (!waiting) break;
    }
    w.cur = -1;
}

}  // namespace emu

using namespace emu;

void __syncthreads()
{
    FiberSet& w = tl_fibers;
    if (w.cur < 0) {
        fprintf(stderr, "EMU: __syncthreads outside a workgroup (%s)\n", tl_kernel ? tl_kernel : "host code");
        _exit(87);
    }
    Fiber& me = w.f[(size_t)w.cur];
    const dim3 tid = me.tid;
    emu_fiber_switch(&me.sp, w.sched);
    threadIdx = tid;
}

// ---- runtime API -------------------------------------------------------------------------

hipError_t hipGetDeviceCount(int* n)
{
    *n = 1;
    return hipSuccess;
}
hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidValue; }
hipError_t hipGetLastError() { return hipSuccess; }
hipError_t hipPeekAtLastError() { return hipSuccess; }
const char* hipGetErrorString(hipError_t e)
{
    switch (e) {
    case hipSuccess: return "hipSuccess";
    case hipErrorInvalidValue: return "hipErrorInvalidValue";
    case hipErrorOutOfMemory: return "hipErrorOutOfMemory";
    case hipErrorNotReady: return "hipErrorNotReady";
    default: return "hipErrorUnknown";
    }
}
hipError_t hipDeviceSynchronize()
{
    sync_all();
    return hipSuccess;
}
hipError_t hipMemGetInfo(size_t* fr, size_t* tot)
{
    const char* e = getenv("EMU_HBM_GB");
    const size_t total = (size_t)(e ? atof(e) : 32.0) * ((size_t)1 << 30);
    std::lock_guard<std::mutex> lk(g_mem_mu);
    *tot = total;
    *fr = total > g_live_device ? total - g_live_device : 0;
    return hipSuccess;
}

hipError_t hipMalloc(void** p, size_t bytes)
{
    size_t fr = 0, tot = 0;
    hipMemGetInfo(&fr, &tot);
    if (bytes > fr) return hipErrorOutOfMemory;
    *p = map_guarded(bytes, "device");
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p)
{
    if (!p) return hipSuccess;
    sync_all();
    return unmap_guarded(p) ? hipSuccess : hipErrorInvalidValue;
}
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned)
{
    *p = map_guarded(bytes, "host");
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostFree(void* p)
{
    if (!p) return hipSuccess;
    sync_all();
    return unmap_guarded(p) ? hipSuccess : hipErrorInvalidValue;
}
hipError_t hipMallocAsync(void** p, size_t bytes, hipStream_t)
{
    *p = map_guarded(bytes, "device (stream-ordered)");
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFreeAsync(void* p, hipStream_t s)
{
    if (!p) return hipSuccess;
    enqueue(s, [p] { unmap_guarded(p); });
    return hipSuccess;
}

hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind)
{
    std::memcpy(dst, src, bytes);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind, hipStream_t s)
{
    enqueue(s, [=] {
        tl_kernel = "hipMemcpyAsync";
        std::memcpy(dst, src, bytes);
        tl_kernel = nullptr;
    });
    return hipSuccess;
}
hipError_t hipMemset(void* p, int v, size_t bytes)
{
    std::memset(p, v, bytes);
    return hipSuccess;
}
hipError_t hipMemsetAsync(void* p, int v, size_t bytes, hipStream_t s)
{
    enqueue(s, [=] {
        tl_kernel = "hipMemsetAsync";
        std::memset(p, v, bytes);
        tl_kernel = nullptr;
    });
    return hipSuccess;
}

hipError_t hipStreamCreate(hipStream_t* s)
{
    *s = new_stream();
    return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { return hipStreamCreate(s); }
hipError_t hipStreamDestroy(hipStream_t s)
{
    if (!s) return hipErrorInvalidValue;
    sync_stream(s);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv.notify_all();
    }
    s->th.join();
    std::lock_guard<std::mutex> lk(g_streams_mu);
    for (auto it = g_streams.begin(); it != g_streams.end(); ++it)
        if (*it == s) {
            g_streams.erase(it);
            break;
        }
    delete s;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s)
{
    sync_stream(resolve(s));
    return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned)
{
    uint64_t seq;
    {
        std::lock_guard<std::mutex> lk(e->mu);
        seq = e->rec;
    }
    enqueue(s, [e, seq] {
        std::unique_lock<std::mutex> lk(e->mu);
        e->cv.wait(lk, [&] { return e->done >= seq; });
    });
    return hipSuccess;
}

// Events are never deleted: queued work may still refer to a destroyed one.
hipError_t hipEventCreate(hipEvent_t* e)
{
    *e = new Event();
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t hs)
{
    Stream* s = resolve(hs);
    bool capturing;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        capturing = s->cap != nullptr;
    }
    if (capturing) {   // a record node: each replay records again
        enqueue(hs, [e] {
            std::lock_guard<std::mutex> lk(e->mu);
            e->done = ++e->rec;
            e->t_ms = now_ms();
            e->cv.notify_all();
        });
        return hipSuccess;
    }
    uint64_t seq;
    {
        std::lock_guard<std::mutex> lk(e->mu);
        seq = ++e->rec;
    }
    enqueue(hs, [e, seq] {
        std::lock_guard<std::mutex> lk(e->mu);
        e->done = std::max(e->done, seq);
        e->t_ms = now_ms();
        e->cv.notify_all();
    });
    return hipSuccess;
}
hipError_t hipEventQuery(hipEvent_t e)
{
    std::lock_guard<std::mutex> lk(e->mu);
    return e->done >= e->rec ? hipSuccess : hipErrorNotReady;
}
hipError_t hipEventSynchronize(hipEvent_t e)
{
    std::unique_lock<std::mutex> lk(e->mu);
    const uint64_t seq = e->rec;
    e->cv.wait(lk, [&] { return e->done >= seq; });
    return hipSuccess;
}
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b)
{
    std::lock(a->mu, b->mu);
    std::lock_guard<std::mutex> la(a->mu, std::adopt_lock), lb(b->mu, std::adopt_lock);
    if (a->done < a->rec || b->done < b->rec || !a->done || !b->done) return hipErrorNotReady;
    *ms = (float)(b->t_ms - a->t_ms);
    return hipSuccess;
}

hipError_t hipStreamBeginCapture(hipStream_t hs, hipStreamCaptureMode)
{
    Stream* s = resolve(hs);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->cap) return hipErrorInvalidValue;
    s->cap = new Graph();
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t hs, hipGraph_t* g)
{
    Stream* s = resolve(hs);
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->cap) return hipErrorInvalidValue;
    *g = s->cap;
    s->cap = nullptr;
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* out, hipGraph_t g, hipGraphNode_t*, char*, size_t)
{
    *out = new Graph(*g);
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t g, hipStream_t hs)
{
    Stream* s = resolve(hs);
    std::lock_guard<std::mutex> lk(s->mu);
    for (const Op& op : g->ops) s->q.push_back(op);
    s->cv.notify_one();
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g)
{
    delete g;
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t g)
{
    delete g;
    return hipSuccess;
}

hipError_t hipGetSymbolAddress(void** p, const void* symbol)
{
    *p = const_cast<void*>(symbol);
    return hipSuccess;
}

// Used by the pool stand-in (emu_pool.cpp).
namespace emu {
void* pool_block_alloc(size_t bytes) { return map_guarded(bytes, "device (pool block)"); }
void pool_block_free(void* p) { unmap_guarded(p); }
}  // namespace emu
