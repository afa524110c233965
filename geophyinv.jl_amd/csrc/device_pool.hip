// DevicePool: the serving scheduler's sub-allocator of device memory (gpi_internal.h).
// Segments come from hipMalloc (256 MB or the request rounded up to 2 MB) and are carved
// first-fit into 256-byte aligned blocks; a freed block coalesces with its free
// neighbours, and trim() gives segments without live blocks back to the device.  The
// caller decides when a block may be freed (after the last work that uses it completed).
#include "gpi_util.h"

namespace gpi {

DevicePool::~DevicePool()
{
    for (Seg* g : segs_) {
        (void)hipFree(g->base);
        delete g;
    }
}

bool DevicePool::carve(Seg* g, size_t n, void** p)
{
    for (auto it = g->free.begin(); it != g->free.end(); ++it) {
        if (it->second < n) continue;
        const size_t off = it->first, len = it->second;
        g->free.erase(it);
        if (len > n) g->free.emplace(off + n, len - n);
        *p = g->base + off;
        g->live += n;
        blocks_.emplace(*p, std::make_pair(g, n));
        return true;
    }
    return false;
}

hipError_t DevicePool::alloc(void** p, size_t bytes)
{
    const size_t n = (size_t)round_up((int64_t)std::max<size_t>(bytes, 1), 256);
    for (Seg* g : segs_)
        if (carve(g, n, p)) return hipSuccess;
    const size_t size = std::max<size_t>((size_t)256 << 20, (size_t)round_up((int64_t)n, (int64_t)2 << 20));
    char* base = nullptr;
    hipError_t e = hipMalloc((void**)&base, size);
    if (e != hipSuccess) {   // give back empty segments and try once more
        (void)hipGetLastError();
        trim();
        e = hipMalloc((void**)&base, size);
        if (e != hipSuccess) return e;
    }
    Seg* g = new Seg{base, size, 0, {}};
    g->free.emplace(0, size);
    segs_.push_back(g);
    reserved_ += size;
    carve(g, n, p);
    return hipSuccess;
}

void DevicePool::free(void* p)
{
    auto b = blocks_.find(p);
    if (b == blocks_.end()) return;
    Seg* g = b->second.first;
    size_t off = (size_t)((char*)p - g->base), len = b->second.second;
    blocks_.erase(b);
    g->live -= len;
    auto next = g->free.lower_bound(off);
    if (next != g->free.end() && off + len == next->first) {   // merge with the following block
        len += next->second;
        next = g->free.erase(next);
    }
    if (next != g->free.begin()) {   // and with the preceding one
        auto prev = std::prev(next);
        if (prev->first + prev->second == off) {
            prev->second += len;
            return;
        }
    }
    g->free.emplace(off, len);
}

void DevicePool::trim()
{
    for (size_t k = 0; k < segs_.size();) {
        Seg* g = segs_[k];
        if (g->live == 0) {
            (void)hipFree(g->base);
            reserved_ -= g->size;
            delete g;
            segs_.erase(segs_.begin() + k);
        } else {
            ++k;
        }
    }
}

}  // namespace gpi
