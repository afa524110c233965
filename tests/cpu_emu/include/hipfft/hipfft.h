// Stand-in for hipFFT in the CPU build of libgpi: one-dimensional batched real transforms
// (D2Z forward, Z2D inverse, unnormalised, hipFFT's layout), run on the host in the
// stream's order by a mixed-radix FFT (emu_fft.cpp).
#pragma once
#include <hip/hip_runtime.h>

typedef struct EmuFftPlan* hipfftHandle;
typedef double2 hipfftDoubleComplex;
typedef enum { HIPFFT_SUCCESS = 0, HIPFFT_INVALID_PLAN = 1, HIPFFT_INVALID_VALUE = 4 } hipfftResult;
typedef enum { HIPFFT_D2Z = 0x6a, HIPFFT_Z2D = 0x6c } hipfftType;

hipfftResult hipfftPlanMany(hipfftHandle* plan, int rank, int* n, int* inembed, int istride, int idist, int* onembed,
                            int ostride, int odist, hipfftType type, int batch);
hipfftResult hipfftSetStream(hipfftHandle plan, hipStream_t stream);
hipfftResult hipfftExecD2Z(hipfftHandle plan, double* in, hipfftDoubleComplex* out);
hipfftResult hipfftExecZ2D(hipfftHandle plan, hipfftDoubleComplex* in, double* out);
hipfftResult hipfftDestroy(hipfftHandle plan);
