// Stand-alone source injection kernel.
//
// Injection replaces the reference's two full-grid passes per source field and step
// (CSC spmv into a zeroed grid buffer, then a whole-grid muladd: src/fdtd/source.jl:61-177)
// with one thread per touched cell.  Each thread sums its contributions in ascending
// source order, which is the accumulation order of SparseArrays' CSC mul!, then applies the
// reference's update expression, so values are identical.
//
// The stencil passes normally apply injection themselves (pass_extras.h); this stand-alone
// kernel serves the source sets that cannot ride along with a pass.
#include "gpi_internal.h"

namespace gpi {

namespace {

enum InjectKind { INJ_P = 0, INJ_NORMAL2 = 1, INJ_NORMAL3 = 2, INJ_VEL = 3 };

struct InjectParams {
    int kind;
    int ntarget;
    int npts;
    int64_t it;
    const int64_t* toff;
    const int32_t* tptr;
    const int32_t* tsrc;
    const float* tw;
    const float* wav;
    float* f0;
    float* f1;
    float* f2;
    const float* coef;     // dtK or dtM on the stress grid
    const float* rho;      // velocity sources: density (mod[:rho])
    const int64_t* roff;   // velocity sources: two rho offsets per target
    float dt;
};

__global__ __launch_bounds__(256) void inject_kernel(InjectParams P)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P.ntarget) return;
    const float* w = P.wav + P.it * (int64_t)P.npts;
    float acc = 0.0f;
    for (int c = P.tptr[t]; c < P.tptr[t + 1]; ++c) acc += P.tw[c] * w[P.tsrc[c]];
    const int64_t o = P.toff[t];
    switch (P.kind) {
    case INJ_P:
        P.f0[o] = P.f0[o] + (acc * P.coef[o]);
        break;
    case INJ_NORMAL3:
        P.f2[o] = P.f2[o] + (acc * P.coef[o]);
        [[fallthrough]];
    case INJ_NORMAL2:
        P.f0[o] = P.f0[o] + (acc * P.coef[o]);
        P.f1[o] = P.f1[o] + (acc * P.coef[o]);
        break;
    default: {
        const float av = (P.rho[P.roff[2 * t]] + P.rho[P.roff[2 * t + 1]]) * 0.5f;
        P.f0[o] = P.f0[o] + ((acc / av) * P.dt);
    }
    }
}

// Records of one receiver set, nt x npts point fastest on the device, into the caller's
// layout (time fastest): lanes run along time, so the stores are coalesced (a few tens of
// microseconds for the largest record sets, against milliseconds for a host transpose).
// A sample whose exponent bits are all set (Inf or NaN) stores 1 into *nonfinite; every
// such store writes the same value, so no atomic is needed.
__global__ __launch_bounds__(256) void records_transpose_kernel(const float* __restrict__ src, float* __restrict__ dst,
                                                                int64_t nt, int np, int* __restrict__ nonfinite)
{
    const int64_t it = (int64_t)blockIdx.y * 256 + threadIdx.x;
    const int r = (int)blockIdx.x;
    if (it < nt) {
        const float v = src[it * np + r];
        dst[(int64_t)r * nt + it] = v;
        if ((__float_as_uint(v) & 0x7f800000u) == 0x7f800000u) *nonfinite = 1;
    }
}

}  // namespace

void records_transpose(const float* rec, float* out, int64_t nt, int npts, int* nonfinite, hipStream_t s)
{
    if (nt <= 0 || npts <= 0) return;
    dim3 grd((unsigned)npts, (unsigned)((nt + 255) / 256)), blk(256);
    hipLaunchKernelGGL(records_transpose_kernel, grd, blk, 0, s, rec, out, nt, npts, nonfinite);
}

void inject(const Model& m, const State& st, const DevPoints& pts, int64_t it, hipStream_t s)
{
    if (pts.ntarget == 0) return;
    InjectParams P{};
    P.ntarget = pts.ntarget;
    P.npts = pts.npts;
    P.it = it;
    P.toff = pts.toff;
    P.tptr = pts.tptr;
    P.tsrc = pts.tsrc;
    P.tw = pts.tw;
    P.wav = pts.wav;
    P.dt = m.g.dt;
    const int f = pts.field;
    if (f == GPI_VX || f == GPI_VY || f == GPI_VZ) {
        P.kind = INJ_VEL;
        P.f0 = st.field[f];
        P.rho = m.coef[GPI_C_RHO];
        P.roff = pts.toff + pts.ntarget;   // rho offsets stored after the target offsets
    } else if (m.g.physics == 0) {
        P.kind = INJ_P;
        P.f0 = st.field[GPI_P];
        P.coef = m.coef[GPI_C_DTK];
    } else {
        P.kind = m.g.ndim == 3 ? INJ_NORMAL3 : INJ_NORMAL2;
        P.f0 = st.field[GPI_TAUXX];
        P.f1 = st.field[GPI_TAUZZ];
        P.f2 = m.g.ndim == 3 ? st.field[GPI_TAUYY] : nullptr;
        P.coef = m.coef[GPI_C_DTM];
    }
    hipLaunchKernelGGL(inject_kernel, dim3((pts.ntarget + 255) / 256), dim3(256), 0, s, P);
}

}  // namespace gpi
