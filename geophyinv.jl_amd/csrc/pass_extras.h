// Work that rides along with a stencil pass: source injection by the thread that owns the
// updated cell, and receiver sampling by extra workgroups appended to the launch.
//
// Injection: targets of one source field are sorted by the bucket that owns them: in 2-D
// a 64-row segment of one derivative-grid column (the cells one wavefront updates, so the
// bucket is wavefront-uniform whatever the workgroup shape), in 3-D the workgroup.
// blk[b]..blk[b+1] lists the targets of bucket b, so a bucket without targets pays two
// uniform loads.  The owner sums the contributions in ascending source order (the CSC
// spmv order of src/fdtd/source.jl:71,89,109,136,152) and applies the reference's update
// expression to the value it has just computed, so the result equals the separate
// injection pass of the reference bit for bit.
//
// Sampling: a pass reads but never writes the fields it samples (p before the velocity
// update, v before the stress update), so extra workgroups can gather receivers
// concurrently with the stencil workgroups (src/fdtd/receiver.jl:3-14).
#pragma once
#include "gpi_internal.h"

namespace gpi {

enum InjKind { INJ_P = 0, INJ_NORMAL2 = 1, INJ_NORMAL3 = 2, INJ_VX = 3, INJ_VY = 4, INJ_VZ = 5 };

// Workgroup order for the 8 XCDs of MI355X: workgroups are dispatched to XCDs round
// robin (workgroup w to XCD w % 8), so tile t = remap(w) gives each XCD one contiguous
// range of tiles.  Neighbouring tiles, which read each other's halo columns or planes,
// then share the XCD's L2.  A bijection of [0, n).
__device__ __forceinline__ int xcd_remap(int w, int n)
{
    const int q = n >> 3, r = n & 7;
    const int k = w & 7, i = w >> 3;
    return k * q + min(k, r) + i;
}

struct InjSet {
    int kind;
    int npts;              // sources of the field (wavelet row length)
    const int32_t* blk;    // nblocks+1 target ranges per workgroup
    const int64_t* toff;   // padded offsets of the targets, grouped by workgroup
    const int32_t* tptr;   // contribution ranges per target
    const int32_t* tsrc;
    const float* tw;
    const float* wav;      // nt x npts, point fastest
    const int64_t* roff;   // velocity sources: two density offsets per target
};

struct GatherSet {
    int nset;
    int nnz;
    int npts[4];
    const float* field[4];
    const int64_t* idx[4];
    const float* w[4];
    float* rec[4];
};

constexpr int MAX_INJ = 4;

struct PassExtras {
    int64_t it;
    int ninj;
    InjSet inj[MAX_INJ];
    int gather_blocks;     // extra workgroups appended after the stencil workgroups
    GatherSet g;
};

// Sum of the contributions to target `off` of workgroup `b`; returns false if none.
__device__ __forceinline__ bool injection_sum(const InjSet& S, int b, int64_t off, int64_t it, float& acc,
                                              int& tslot)
{
    const int s = S.blk[b], e = S.blk[b + 1];
    for (int t = s; t < e; ++t) {
        if (S.toff[t] == off) {
            const float* w = S.wav + it * (int64_t)S.npts;
            float a = 0.0f;
            for (int c = S.tptr[t]; c < S.tptr[t + 1]; ++c) a += S.tw[c] * w[S.tsrc[c]];
            acc = a;
            tslot = t;
            return true;
        }
    }
    return false;
}

// v[inn] = v[inn] + (pv / @av(rho)) * dt  (muladd_with_density_v*!, source.jl:166-177)
__device__ __forceinline__ float inject_velocity(const PassExtras& X, int64_t it, int kind, int b, int64_t off,
                                                 float v, const float* rho, float dt)
{
    for (int q = 0; q < X.ninj; ++q) {
        const InjSet& S = X.inj[q];
        if (S.kind != kind) continue;
        float acc;
        int t;
        if (injection_sum(S, b, off, it, acc, t)) {
            const float av = (rho[S.roff[2 * t]] + rho[S.roff[2 * t + 1]]) * 0.5f;
            v = v + ((acc / av) * dt);
        }
    }
    return v;
}

__device__ __forceinline__ float inject_velocity(const PassExtras& X, int kind, int b, int64_t off, float v,
                                                 const float* rho, float dt)
{
    return inject_velocity(X, X.it, kind, b, off, v, rho, dt);
}

// Receiver sampling by wavefront shuffles: the nnz taps of a receiver (its interpolation
// offsets and weights, 4 in 2-D and 8 in 3-D) go to gather_group(nnz) consecutive lanes of
// one wave, so that a wave reads the offsets, weights and field samples of its receivers
// with one load instruction each (the table reads coalesced); the lane of tap 0 then sums
// the products in tap order from its neighbours' registers, the same arithmetic as a loop
// over the taps (src/fdtd/receiver.jl:3-14).  A gather workgroup samples
// gather_per_block(nnz) receivers, one round for 256 threads: gather_blocks =
// nset * ceil(max npts / gather_per_block(nnz)).  The host admits nnz <= 64 only.
constexpr int GATHER_LANES = 256;

__host__ __device__ constexpr int gather_group(int nnz)
{
    int g = 1;
    while (g < nnz) g <<= 1;
    return g;
}

__host__ __device__ constexpr int gather_per_block(int nnz) { return GATHER_LANES / gather_group(nnz); }

// receiver sampling by workgroup gb of the extra range; every thread of the workgroup takes
// part (the shuffles read the lanes of whole waves).  NNZ: the taps per receiver when the
// caller's dimension fixes them (the host admits 4 in 2-D and 8 in 3-D only), 0 to read them.
template <int NNZ = 0>
__device__ __forceinline__ void gather_block(const PassExtras& X, int64_t it, int gb)
{
    const int per = (X.gather_blocks + X.g.nset - 1) / X.g.nset;
    const int s = gb / per;
    if (s >= X.g.nset) return;
    const int nnz = NNZ ? NNZ : X.g.nnz, G = gather_group(nnz);
    const int nth = (int)(blockDim.x * blockDim.y);
    const int t = (int)(threadIdx.y * blockDim.x + threadIdx.x);
    const int q = t & (G - 1);   // this lane's tap
    const int base = (gb - s * per) * gather_per_block(nnz);
    const int stop = min(base + gather_per_block(nnz), X.g.npts[s]);
    const float* F = X.g.field[s];
    for (int r0 = base; r0 < stop; r0 += nth / G) {   // (bounds uniform over the workgroup)
        const int r = r0 + t / G;
        float w = 0.0f, f = 0.0f;
        if (r < stop && q < nnz) {
            const int64_t e = (int64_t)r * nnz + q;
            w = X.g.w[s][e];
            f = F[X.g.idx[s][e]];
        }
        float tmp = 0.0f;
        for (int m = 0; m < nnz; ++m) tmp += __shfl_down(w, m, G) * __shfl_down(f, m, G);
        if (q == 0 && r < stop) X.g.rec[s][it * (int64_t)X.g.npts[s] + r] = tmp;
    }
}

template <int NNZ = 0>
__device__ __forceinline__ void gather_block(const PassExtras& X, int gb) { gather_block<NNZ>(X, X.it, gb); }

}  // namespace gpi
