// CPU stand-in for the HIP runtime header, used only to build libgpi for the host
// (tests/cpu_emu): kernels run on host threads workgroup by workgroup, each stream has a
// worker thread that executes its queue in order, events and stream waits order work
// across streams, and every device or pinned allocation sits against a guard page and is
// made inaccessible when it is freed (hipFree, hipHostFree, or hipFreeAsync when the free
// is reached in stream order), so an out-of-bounds or stale access stops at the faulting
// instruction.  Test infrastructure: nothing of the library's product build uses it.
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <tuple>
#include <utility>

// ---- qualifiers ---------------------------------------------------------------------
#define __global__
#define __device__
#define __host__
#define __constant__
#define __forceinline__ inline __attribute__((always_inline))
#define __launch_bounds__(...)
#define __align__(n) __attribute__((aligned(n)))
#define __shared__ static thread_local

// ---- types --------------------------------------------------------------------------
enum hipError_t {
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorOutOfMemory = 2,
    hipErrorNotReady = 600,
    hipErrorLaunchFailure = 719,
};

enum hipMemcpyKind {
    hipMemcpyHostToHost = 0,
    hipMemcpyHostToDevice = 1,
    hipMemcpyDeviceToHost = 2,
    hipMemcpyDeviceToDevice = 3,
    hipMemcpyDefault = 4,
};

enum hipStreamCaptureMode {
    hipStreamCaptureModeGlobal = 0,
    hipStreamCaptureModeThreadLocal = 1,
    hipStreamCaptureModeRelaxed = 2,
};

#define hipStreamDefault 0u
#define hipStreamNonBlocking 1u
#define hipEventDefault 0u
#define hipEventDisableTiming 2u
#define hipHostMallocDefault 0u

struct dim3 {
    uint32_t x, y, z;
    constexpr dim3(uint32_t a = 1, uint32_t b = 1, uint32_t c = 1) : x(a), y(b), z(c) {}
};

// HIP's vector of two doubles (complex samples of the FFT work)
struct alignas(16) double2 {
    double x, y;
};

namespace emu {
struct Stream;
struct Event;
struct Graph;
}  // namespace emu
typedef emu::Stream* hipStream_t;
typedef emu::Event* hipEvent_t;
typedef emu::Graph* hipGraph_t;
typedef emu::Graph* hipGraphExec_t;
typedef void* hipGraphNode_t;

// ---- built-in variables and device functions -------------------------------------------
extern thread_local dim3 threadIdx __attribute__((tls_model("initial-exec")));
extern thread_local dim3 blockIdx __attribute__((tls_model("initial-exec")));
extern thread_local dim3 blockDim __attribute__((tls_model("initial-exec")));
extern thread_local dim3 gridDim __attribute__((tls_model("initial-exec")));

#define __builtin_amdgcn_readfirstlane(x) (x)

// Raw buffer resources: base address and record count in bytes; a load at an offset at or
// past the count reads 0 without touching memory, as the hardware's range check does
// (soffset is added to the address but, as on gfx9, not to the checked offset).
struct __amdgpu_buffer_rsrc_t {
    const char* base;
    uint32_t bytes;
};
namespace emu {
inline __amdgpu_buffer_rsrc_t make_buffer_rsrc(const void* p, short, int bytes, int)
{
    return __amdgpu_buffer_rsrc_t{(const char*)p, (uint32_t)bytes};
}
inline uint32_t raw_buffer_load_b32(__amdgpu_buffer_rsrc_t r, int off, int soff, int)
{
    if ((uint32_t)off >= r.bytes) return 0u;
    uint32_t v;
    std::memcpy(&v, r.base + (uint32_t)off + (uint32_t)soff, 4);
    return v;
}
}  // namespace emu
#define __builtin_amdgcn_make_buffer_rsrc(p, stride, bytes, flags) ::emu::make_buffer_rsrc((p), (stride), (bytes), (flags))
#define __builtin_amdgcn_raw_buffer_load_b32(r, off, soff, aux) ::emu::raw_buffer_load_b32((r), (off), (soff), (aux))
#define __builtin_nontemporal_store(v, p) (*(p) = (v))

inline unsigned int __float_as_uint(float f)
{
    unsigned int u;
    std::memcpy(&u, &f, 4);
    return u;
}
inline float __uint_as_float(unsigned int u)
{
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
using std::max;
using std::min;

// Each thread of a workgroup runs as a fiber of the worker that runs the workgroup; a
// barrier suspends the calling thread until every thread of its workgroup that has not
// returned reaches it (threads that returned no longer take part, as ended waves do).
// Between barriers the threads run in thread order, or last first with EMU_REVERSE=1.
void __syncthreads();

// Wave shuffle: the value of the lane d above this one within its segment of `width` lanes
// (its own value past the segment's end), lanes being the workgroup's linear thread ids
// (64 to a wave).  Exchanged through the workgroup between two barriers, so every thread of
// the workgroup must call it the same number of times (as the kernels' uniform uses do).
template <typename T>
inline T __shfl_down(T v, unsigned int d, int width = 64)
{
    static thread_local T x[1024];
    const int t = (int)(threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z));
    x[t] = v;
    __syncthreads();
    const T r = (t % width) + (int)d < width ? x[t + (int)d] : v;
    __syncthreads();
    return r;
}

// ---- runtime API --------------------------------------------------------------------
hipError_t hipGetDeviceCount(int* n);
hipError_t hipSetDevice(int d);
hipError_t hipGetLastError();
hipError_t hipPeekAtLastError();
const char* hipGetErrorString(hipError_t e);
hipError_t hipDeviceSynchronize();
hipError_t hipMemGetInfo(size_t* free_bytes, size_t* total_bytes);

hipError_t hipMalloc(void** p, size_t bytes);
hipError_t hipFree(void* p);
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned flags = 0);
hipError_t hipHostFree(void* p);
hipError_t hipMallocAsync(void** p, size_t bytes, hipStream_t s);
hipError_t hipFreeAsync(void* p, hipStream_t s);
template <typename T>
hipError_t hipMalloc(T** p, size_t bytes) { return hipMalloc((void**)p, bytes); }
template <typename T>
hipError_t hipHostMalloc(T** p, size_t bytes, unsigned flags = 0) { return hipHostMalloc((void**)p, bytes, flags); }
template <typename T>
hipError_t hipMallocAsync(T** p, size_t bytes, hipStream_t s) { return hipMallocAsync((void**)p, bytes, s); }

hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind);
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t s = nullptr);
hipError_t hipMemset(void* p, int v, size_t bytes);
hipError_t hipMemsetAsync(void* p, int v, size_t bytes, hipStream_t s = nullptr);

hipError_t hipStreamCreate(hipStream_t* s);
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);

hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
hipError_t hipEventQuery(hipEvent_t e);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);

hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode mode);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g);
hipError_t hipGraphInstantiate(hipGraphExec_t* out, hipGraph_t g, hipGraphNode_t* err_node, char* log, size_t n);
hipError_t hipGraphLaunch(hipGraphExec_t g, hipStream_t s);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t g);

#define HIP_SYMBOL(x) ((const void*)&(x))
hipError_t hipGetSymbolAddress(void** p, const void* symbol);

// ---- kernel launch ------------------------------------------------------------------
namespace emu {
// Queue a grid x block launch on s; `run` runs every thread of the workgroup whose
// blockIdx/gridDim/blockDim are set (threadIdx is set per thread by `run`).
void launch(hipStream_t s, dim3 grid, dim3 block, const char* name, void (*run)(void*), void* ctx,
            void (*drop)(void*));
// Run `one` (one thread's body for the current threadIdx) for every thread of the current
// workgroup, each on a fiber of its own so that __syncthreads can suspend it.
void run_workgroup(void (*one)(const void*), const void* ctx);

template <typename F>
struct KernelSig;
template <typename... P>
struct KernelSig<void (*)(P...)> {
    using Args = std::tuple<typename std::decay<P>::type...>;
};

// The kernel is a template argument, so each workgroup's thread loop calls it directly.
template <auto K>
struct KernelCall {
    typename KernelSig<decltype(K)>::Args args;
    static void one(const void* self) { std::apply(K, ((const KernelCall*)self)->args); }
    static void run(void* self) { run_workgroup(&one, self); }
    static void drop(void* self) { delete (KernelCall*)self; }
};

// Arguments are converted to the kernel's parameter types at the launch, as HIP does.
template <auto K, typename... A>
void launch_kernel(hipStream_t s, dim3 grid, dim3 block, const char* name, A&&... a)
{
    using Args = typename KernelSig<decltype(K)>::Args;
    static_assert(std::tuple_size<Args>::value == sizeof...(A), "kernel argument count");
    auto* c = new KernelCall<K>{Args(std::forward<A>(a)...)};
    launch(s, grid, block, name, &KernelCall<K>::run, c, &KernelCall<K>::drop);
}
}  // namespace emu

#define hipLaunchKernelGGL(K, G, B, SHM, S, ...) ::emu::launch_kernel<K>((S), dim3(G), dim3(B), #K, ##__VA_ARGS__)
