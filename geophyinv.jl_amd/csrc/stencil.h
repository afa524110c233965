// Staggered finite-difference stencils of orders 2, 4, 6, 8 with the literal order of
// src/fdtd/diff2D.jl:15-212 (and diff3D.jl).  v[k] is the sample at offset k along the
// derivative axis; the divisor (1, 24, 1920, 107520) is folded into dsI on the host
// (src/fdtd/fdtd.jl:313-332).  Evaluated strictly left to right in FP32.
#pragma once
#include <hip/hip_runtime.h>

namespace gpi {

template <int O>
__device__ __forceinline__ float stencil(const float* v);

template <>
__device__ __forceinline__ float stencil<2>(const float* v)
{
    return v[1] - v[0];
}

template <>
__device__ __forceinline__ float stencil<4>(const float* v)
{
    return ((v[2] * 27.0f - v[1] * 27.0f) + v[0]) - v[3];
}

template <>
__device__ __forceinline__ float stencil<6>(const float* v)
{
    return ((((v[3] * 2250.0f - v[2] * 2250.0f) + v[1] * 125.0f) - v[4] * 125.0f) + v[5] * 9.0f) -
           v[0] * 9.0f;
}

template <>
__device__ __forceinline__ float stencil<8>(const float* v)
{
    return ((((((v[4] * 128625.0f - v[3] * 128625.0f) + v[2] * 8575.0f) - v[5] * 8575.0f) +
              v[6] * 1029.0f) - v[1] * 1029.0f) + v[0] * 75.0f) - v[7] * 75.0f;
}

// Coefficients formed in the pass from the independent parameters, with the FP32
// expressions of update_dmod! (src/fdtd/medium.jl:143-186, one FP64 promotion in dtM), so
// the passes read rho (resp. invlambda and invmu) instead of the derived arrays:
//   dtinvav*irho = dt / ((rho_a + rho_b) * 0.5)          (@av_xi / @av_zi of rho)
//   dtlambda = inv(invlambda) * dt,  dtM = dt * Float32(inv(invlambda) + 2.0 * inv(invmu))
//   dtavmu = dt / ((((im_00 + im_10) + im_01) + im_11) * 0.25)   (@av of invmu)
__device__ __forceinline__ float dt_over_av2(float dt, float a, float b) { return dt / ((a + b) * 0.5f); }

__device__ __forceinline__ float dtm_of(float il, float im, float dt)
{
    const double m = (double)(1.0f / il) + 2.0 * (double)(1.0f / im);
    return ((float)m) * dt;
}

__device__ __forceinline__ float dt_over_av4(float dt, float a, float b, float c, float d)
{
    return dt / ((((a + b) + c) + d) * 0.25f);
}

}  // namespace gpi
