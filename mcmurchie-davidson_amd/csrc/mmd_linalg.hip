// Principal powers S^-1/2 and S^1/2 of the overlap matrix on the GPU
// (Molecule.one_electron_integrals: X = fractional_matrix_power(S, -0.5),
// U = fractional_matrix_power(S, 0.5), mmd/molecule.py:284-286).
//
// The reference diagonalises S; here the coupled Newton-Schulz iteration
// (Higham, "Functions of Matrices", 2008, eq. 6.35) computes the same two
// matrices from products only, on the matrix cores:
//   A = S / c,  Y_0 = A,  Z_0 = I,
//   T_k = (3 I - Z_k Y_k) / 2,  Y_{k+1} = Y_k T_k,  Z_{k+1} = T_k Z_k,
// with Y_k -> A^{1/2}, Z_k -> A^{-1/2} quadratically once ||I - Z_k Y_k|| is
// small; c = ||S||_F bounds the largest eigenvalue, so the spectrum of A lies
// in (0, 1] and the iteration converges for any positive definite S.  Then
// X = Z / sqrt(c), U = Y sqrt(c), symmetrised.
#include <hip/hip_runtime.h>
#include <math.h>
#include "mmd_gemm.h"

namespace mmd {

namespace {
// sum of squares of x (n elements) into out[0] (one workgroup)
__global__ void __launch_bounds__(1024) sumsq_kernel(size_t n, const double* __restrict__ x, double* out)
{
    __shared__ double part[16];
    double s = 0.0;
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) s = fma(x[i], x[i], s);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) part[wv] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += part[w];
        out[0] = t;
    }
}

// Y = S * inv_c, Z = I
__global__ void ns_init_kernel(int n, const double* __restrict__ S, const double* __restrict__ sumsq,
                               double* __restrict__ Y, double* __restrict__ Z)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * n) return;
    const double inv_c = 1.0 / sqrt(sumsq[0]);
    Y[i] = S[i] * inv_c;
    Z[i] = (i / n == i % n) ? 1.0 : 0.0;
}

// T = (3 I - P) / 2 and R = I - P (the residual whose norm decides convergence)
__global__ void ns_step_kernel(int n, const double* __restrict__ P, double* __restrict__ T, double* __restrict__ R)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * n) return;
    const double d = (i / n == i % n) ? 1.0 : 0.0;
    T[i] = 0.5 * (3.0 * d - P[i]);
    R[i] = d - P[i];
}

// X = (Z + Z^T) / (2 sqrt(c)),  U = (Y + Y^T) sqrt(c) / 2
__global__ void ns_finish_kernel(int n, const double* __restrict__ Z, const double* __restrict__ Y,
                                 const double* __restrict__ sumsq, double* __restrict__ X, double* __restrict__ U)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * n) return;
    const size_t r = i / n, c = i % n, t = c * n + r;
    const double sc = sqrt(sqrt(sumsq[0]));   // sqrt(c), c = ||S||_F
    if (X) X[i] = 0.5 * (Z[i] + Z[t]) / sc;
    if (U) U[i] = 0.5 * (Y[i] + Y[t]) * sc;
}
}  // namespace

int overlap_powers_device(int n, const double* S, double* X, double* U, double* work, int max_iter,
                          hipStream_t st)
{
    const size_t nn = (size_t)n * n;
    double* Y = work;
    double* Z = work + nn;
    double* P = work + 2 * nn;
    double* T = work + 3 * nn;
    double* W = work + 4 * nn;
    double* norms = work + 5 * nn;   // [0] ||S||_F^2, [1] ||I - Z Y||_F^2
    const unsigned g = (unsigned)((nn + 255) / 256);
    hipLaunchKernelGGL(sumsq_kernel, dim3(1), dim3(1024), 0, st, nn, S, norms);
    hipLaunchKernelGGL(ns_init_kernel, dim3(g), dim3(256), 0, st, n, S, norms, Y, Z);
    if (hipGetLastError() != hipSuccess) return -1;
    int converged_at = 0;
    for (int it = 1; it <= max_iter; ++it) {
        if (launch_gemm_f64(false, false, n, n, n, 1.0, Z, n, Y, n, 0.0, P, n, st) != hipSuccess) return -1;
        hipLaunchKernelGGL(ns_step_kernel, dim3(g), dim3(256), 0, st, n, P, T, W);
        hipLaunchKernelGGL(sumsq_kernel, dim3(1), dim3(1024), 0, st, nn, W, norms + 1);
        // Y <- Y T (into P), Z <- T Z (into W); then swap
        if (launch_gemm_f64(false, false, n, n, n, 1.0, Y, n, T, n, 0.0, P, n, st) != hipSuccess) return -1;
        if (launch_gemm_f64(false, false, n, n, n, 1.0, T, n, Z, n, 0.0, W, n, st) != hipSuccess) return -1;
        if (hipMemcpyAsync(Y, P, sizeof(double) * nn, hipMemcpyDeviceToDevice, st) != hipSuccess) return -1;
        if (hipMemcpyAsync(Z, W, sizeof(double) * nn, hipMemcpyDeviceToDevice, st) != hipSuccess) return -1;
        double res2 = 0.0;
        if (hipMemcpyAsync(&res2, norms + 1, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
        if (hipStreamSynchronize(st) != hipSuccess) return -1;
        // a residual that is not finite or grows without bound: S is not
        // positive definite (the iteration diverges)
        if (!(res2 == res2) || res2 > 1e30) return 0;
        // the residual is that of the iterate before this step: once it is
        // below 1e-9 (relative to sqrt(n)), this step has taken it to rounding
        if (res2 <= 1e-18 * n) {
            converged_at = it;
            break;
        }
    }
    if (converged_at == 0) return 0;
    hipLaunchKernelGGL(ns_finish_kernel, dim3(g), dim3(256), 0, st, n, Z, Y, norms, X, U);
    if (hipGetLastError() != hipSuccess) return -1;
    return converged_at;
}

}  // namespace mmd
