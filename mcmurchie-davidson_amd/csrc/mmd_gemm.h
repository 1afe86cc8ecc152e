// FP64 matrix-core products and the overlap powers built on them
// (host-side declarations).
#pragma once
#include <hip/hip_runtime.h>

namespace mmd {

// C = alpha op(A) op(B) + beta C, row-major, on the matrix cores (mmd_gemm.hip).
hipError_t launch_gemm_f64(bool tA, bool tB, int M, int N, int K, double alpha, const double* A, int lda,
                           const double* B, int ldb, double beta, double* C, int ldc, hipStream_t st);

// Newton-Schulz iteration for the principal square root and inverse square
// root of a symmetric positive definite n x n matrix S (mmd_linalg.hip).
// work: 5 n^2 doubles plus 2 for the norms; host_norm: pinned or pageable
// host double.  Returns the iteration count (> 0), 0 if it did not converge
// within max_iter, or -1 on a launch error.
int overlap_powers_device(int n, const double* S, double* X, double* U, double* work, int max_iter,
                          hipStream_t st);

}  // namespace mmd
