// DevicePool for the CPU build (replaces geophyinv.jl_amd/csrc/device_pool.hip there):
// every block is its own guarded mapping, and free() makes it inaccessible at once.  The
// serving scheduler frees a block when it may hand the memory to another shot, so any
// access that a kernel, copy or host thread makes to a block after that point faults
// instead of reading or writing another shot's data.
#include "gpi_util.h"

namespace emu {
void* pool_block_alloc(size_t bytes);
void pool_block_free(void* p);
}  // namespace emu

namespace gpi {

DevicePool::~DevicePool()
{
    for (auto& b : blocks_) emu::pool_block_free(b.first);
}

bool DevicePool::carve(Seg*, size_t, void**) { return false; }

hipError_t DevicePool::alloc(void** p, size_t bytes)
{
    const size_t n = (size_t)round_up((int64_t)std::max<size_t>(bytes, 1), 256);
    *p = emu::pool_block_alloc(n);
    if (!*p) return hipErrorOutOfMemory;
    blocks_.emplace(*p, std::make_pair((Seg*)nullptr, n));
    reserved_ += n;
    return hipSuccess;
}

void DevicePool::free(void* p)
{
    auto b = blocks_.find(p);
    if (b == blocks_.end()) return;
    reserved_ -= b->second.second;
    blocks_.erase(b);
    emu::pool_block_free(p);
}

void DevicePool::trim() {}

}  // namespace gpi
