// gpi_forward_static: gpi_forward without device allocation.
//
// gpi_forward takes its per-shot device buffers (source and receiver tables, the records
// on the device, the transposed records and the non-finite flag, snapshot, illumination
// and packing slabs) from hipMalloc.  This entry point carves the same buffers out of one
// file-scope device array instead, for callers that cannot call hipMalloc.  The array is
// looked up once per device with hipGetSymbolAddress (the array's name in host code is a
// host-side stand-in, never a device address); every call on a device shares it, so calls
// that overlap on one device race, and a call's buffers keep the previous call's contents
// until it clears or overwrites them (records are cleared before the first step).
#include <mutex>

#include "gpi_util.h"

using namespace gpi;

namespace {

static __device__ __align__(256) char scratch[GPI_STATIC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
char* g_ptr[kMaxDevices] = {};

// The device address of `scratch` on the context's device (current on entry).
int scratch_ptr(gpi_ctx* c, char** out)
{
    if (c->device < 0 || c->device >= kMaxDevices) return fail(c, GPI_EINVAL, "device id out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[c->device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(c, GPI_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_ptr[c->device] = (char*)p;
    }
    *out = g_ptr[c->device];
    return GPI_OK;
}

}  // namespace

extern "C" int gpi_forward_static(gpi_ctx* c, uint64_t model, const gpi_shot* shot)
{
    if (!c) return GPI_EINVAL;
    HIPTRY(c, hipSetDevice(c->device));
    char* base = nullptr;
    int rc = scratch_ptr(c, &base);
    if (rc) return rc;
    Arena a;
    a.base = base;
    a.cap = GPI_STATIC_SCRATCH_BYTES;
    c->arena = &a;
    rc = forward_shot(c, model, shot);
    c->arena = nullptr;
    // work queued before an early error may still use the array: drain before returning
    const hipError_t e = hipStreamSynchronize(c->stream);
    if (!rc && e != hipSuccess) rc = fail(c, GPI_EHIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
    return rc;
}
