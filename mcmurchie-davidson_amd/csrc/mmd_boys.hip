// Boys function: the host-side Taylor table used by the ERI kernels and the
// general-order device evaluation behind the _boys test hook.
//
//   F_m(T) = int_0^1 t^{2m} exp(-T t^2) dt
//          = 1F1(m+1/2; m+3/2; -T) / (2m+1)     (cython/util.pxi:53-55)
#include <hip/hip_runtime.h>
#include <math.h>
#include <vector>
#include "mmd_device.h"
#include "mmd_kernels.h"

namespace mmd {

// Host: F_m(T) = e^{-T} sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1)), extended
// precision; every term is positive, so the sum is free of cancellation.
static long double boys_series_host(int m, long double T)
{
    long double term = 1.0L / (2.0L * m + 1.0L);
    long double sum = term;
    for (int k = 1; k < 4000; ++k) {
        term *= 2.0L * T / (2.0L * m + 2.0L * k + 1.0L);
        sum += term;
        if (term < sum * 1e-22L) break;
    }
    return expl(-T) * sum;
}

void boys_build_table(std::vector<double>& tab)
{
    tab.assign((size_t)BOYS_ROWS * BOYS_NT, 0.0);
    for (int m = 0; m < BOYS_ROWS; ++m)
        for (int i = 0; i < BOYS_NT; ++i)
            tab[(size_t)m * BOYS_NT + i] = (double)boys_series_host(m, (long double)i * (long double)BOYS_DT);
}

// Device, general real order a = m + 1/2: lower incomplete gamma by its
// power series for T < a + 1, otherwise Gamma(a) minus the upper part from
// Legendre's continued fraction (modified Lentz).
__device__ double boys_general(double m, double T)
{
    const double a = m + 0.5;
    if (T <= 0.0) return 1.0 / (2.0 * a);
    if (T < a + 1.0) {
        double term = 1.0 / (2.0 * a);
        double sum = term;
        for (int k = 1; k < 20000; ++k) {
            term *= T / (a + (double)k);
            sum += term;
            if (term < sum * 1e-17) break;
        }
        return exp(-T) * sum;
    }
    const double tiny = 1e-300;
    double b = T + 1.0 - a;
    double c = 1.0 / tiny;
    double d = 1.0 / b;
    double h = d;
    for (int i = 1; i < 20000; ++i) {
        const double an = -(double)i * ((double)i - a);
        b += 2.0;
        d = an * d + b;
        if (fabs(d) < tiny) d = tiny;
        c = b + an / c;
        if (fabs(c) < tiny) c = tiny;
        d = 1.0 / d;
        const double del = d * c;
        h *= del;
        if (fabs(del - 1.0) < 1e-16) break;
    }
    const double lg = lgamma(a);
    const double lT = log(T);
    const double Q = exp(-T + a * lT - lg) * h;
    return exp(lg - a * lT) * 0.5 * (1.0 - Q);
}

__global__ void boys_eval_kernel(int n, const double* __restrict__ m, const double* __restrict__ T, int table,
                                 const double* __restrict__ tab, double* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (table) {
        double F[BOYS_MMAX + 1];
        const int L = (int)m[i];
        boys_runtime(L, T[i], tab, F);
        out[i] = F[L];
    } else {
        out[i] = boys_general(m[i], T[i]);
    }
}

hipError_t launch_boys_eval(int n, const double* m, const double* T, int table, const double* tab, double* out,
                            hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(boys_eval_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, m, T, table, tab, out);
    return hipGetLastError();
}

}  // namespace mmd
