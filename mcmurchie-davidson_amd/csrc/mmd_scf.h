// Elementwise and reduction kernels of the GPU-resident SCF loop (host-side
// declarations; the kernels are in mmd_util.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace mmd {

hipError_t launch_scf_axpby(size_t n, double a, const double* x, double b, const double* y, double* out,
                            hipStream_t st);
hipError_t launch_scf_antisym(int N, const double* M, double* out, hipStream_t st);
hipError_t launch_scf_energy(int N, const double* H, const double* F, const double* P, const double* Pold,
                             double* out, hipStream_t st);
// out[k] = <E[k], e> for k < nset <= 8 (device vectors of length n)
hipError_t launch_scf_dots(int nset, size_t n, const double* const* E, const double* e, double* out,
                           hipStream_t st);
// out = sum_k w[k] F[k], k < nset <= 8
hipError_t launch_scf_combine(size_t n, int nset, const double* const* F, const double* w, double* out,
                              hipStream_t st);

}  // namespace mmd
