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

}  // namespace gpi
