// Device-side building blocks of the McMurchie-Davidson integral kernels for
// gfx950 (CDNA4): compile-time Hermite index maps, Cartesian component tables,
// the Boys function from a Taylor table staged in LDS, in-register Hermite
// expansion (E) and Hermite Coulomb (R) recursions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include <utility>

namespace mmd {

// ---------------------------------------------------------------------------
// Compile-time index helpers
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
// Number of Hermite functions with t+u+v <= L.
__host__ __device__ constexpr int nherm(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }
// Flat index of (t,u,v) in the tetrahedral set t+u+v <= L (independent of L).
__host__ __device__ constexpr int hidx(int t, int u, int v)
{
    return (t + u + v) * (t + u + v + 1) * (t + u + v + 2) / 6 + (u + v) * (u + v + 1) / 2 + v;
}
// Cartesian components of a shell in the reference's order
// (mmd/molecule.py:108-114): x exponent descending, then y descending.
__host__ __device__ constexpr int comp_lx(int l, int c)
{
    for (int lx = l; lx >= 0; --lx) {
        int n = l - lx + 1;
        if (c < n) return lx;
        c -= n;
    }
    return 0;
}
__host__ __device__ constexpr int comp_ly(int l, int c)
{
    for (int lx = l; lx >= 0; --lx) {
        int n = l - lx + 1;
        if (c < n) return l - lx - c;
        c -= n;
    }
    return 0;
}
__host__ __device__ constexpr int comp_lz(int l, int c) { return l - comp_lx(l, c) - comp_ly(l, c); }

// Angular-momentum pair class: (0,0)=0 (1,0)=1 (1,1)=2 (2,0)=3 (2,1)=4 (2,2)=5.
__host__ __device__ constexpr int pair_class(int la, int lb) { return la * (la + 1) / 2 + lb; }

constexpr int MAX_L = 2;            // highest shell angular momentum supported
constexpr int MAX_LTOT = 4 * MAX_L; // highest total L of a quartet
constexpr int NPAIRCLASS = 6;
constexpr int NQCLASS = 21;         // quartet classes with pair_class(bra) >= pair_class(ket)

// ---------------------------------------------------------------------------
// Compile-time loops
// ---------------------------------------------------------------------------
template <class F, int... I>
__device__ __forceinline__ void sfor_impl(F&& f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void sfor(F&& f)
{
    if constexpr (N > 0) sfor_impl(f, std::make_integer_sequence<int, N>{});
}

struct Herm3 {
    int t, u, v;
};
// Inverse of hidx: (t,u,v) of flat Hermite index k.
__host__ __device__ constexpr Herm3 hdecode(int k)
{
    int n = 0;
    while (nherm(n) <= k) ++n;
    int r = k - (n > 0 ? nherm(n - 1) : 0);
    int s = 0;
    while ((s + 1) * (s + 2) / 2 <= r) ++s;
    const int v = r - s * (s + 1) / 2;
    return Herm3{n - s, s - v, v};
}


// ---------------------------------------------------------------------------
// Boys function table
//   F_m(T_i), T_i = i * BOYS_DT, i = 0..BOYS_NT-1, m = 0..BOYS_MMAX
// Evaluation: 6th-order Taylor expansion of F_L about the nearest grid point
// (|T - T_i| <= BOYS_DT/2), then downward recursion
//   F_m = (2T F_{m+1} + e^{-T}) / (2m+1).
// For T >= BOYS_TMAX: F_0 = sqrt(pi/T)/2 (erfc(sqrt(T)) < 3e-17) and the
// upward recursion F_{m+1} = ((2m+1) F_m - e^{-T}) / (2T).
// ---------------------------------------------------------------------------
constexpr double BOYS_DT = 0.1;
constexpr double BOYS_INV_DT = 10.0;
constexpr int BOYS_NT = 361;          // T in [0, 36.05]
constexpr double BOYS_TMAX = 36.0;
constexpr int BOYS_TAYLOR = 7;        // terms k = 0..6
constexpr int BOYS_MMAX = MAX_LTOT + BOYS_TAYLOR;       // 15: derivative integrals reach L + 1
constexpr int BOYS_ROWS = BOYS_MMAX + 1;

__device__ __forceinline__ double dev_rsqrt(double x) { return rsqrt(x); }
__device__ __forceinline__ float dev_rsqrt(float x) { return rsqrtf(x); }
__device__ __forceinline__ double dev_exp(double x) { return exp(x); }
__device__ __forceinline__ float dev_exp(float x) { return expf(x); }

// Rows needed by a class of total angular momentum L: m = L .. L+6.  The
// table is double (FP64 kernels) or float (FP32-tier kernels).
template <int L, typename R = double>
__device__ inline void boys_stage_lds(R* lds, const R* __restrict__ tab)
{
    const R* src = tab + (size_t)L * BOYS_NT;
    for (int i = threadIdx.x; i < BOYS_TAYLOR * BOYS_NT; i += blockDim.x) lds[i] = src[i];
}

// F[0..L] scaled by pref * (-2 alpha)^m, i.e. the seeds of the R recursion.
template <int L, typename R = double>
__device__ inline void boys_seeds(R T, R alpha, R pref, const R* lds, R* F)
{
    R Fm[L + 1];
    if (T < R(BOYS_TMAX)) {
        const int i = (int)(T * R(BOYS_INV_DT) + R(0.5));
        const R dT = (R)i * R(BOYS_DT) - T;   // expansion in (-dT)
        const R* r = lds + i;
        // sum_k F_{L+k}(T_i) (T_i - T)^k / k!   (Horner)
        R f = r[6 * BOYS_NT];
        f = r[5 * BOYS_NT] + dT * f * R(1.0 / 6.0);
        f = r[4 * BOYS_NT] + dT * f * R(1.0 / 5.0);
        f = r[3 * BOYS_NT] + dT * f * R(1.0 / 4.0);
        f = r[2 * BOYS_NT] + dT * f * R(1.0 / 3.0);
        f = r[1 * BOYS_NT] + dT * f * R(1.0 / 2.0);
        f = r[0] + dT * f;
        Fm[L] = f;
        if (L > 0) {
            const R e = dev_exp(-T);
            const R t2 = R(2.0) * T;
#pragma unroll
            for (int m = L - 1; m >= 0; --m) Fm[m] = (t2 * Fm[m + 1] + e) * R(1.0 / (2.0 * m + 1.0));
        }
    } else {
        const R rT = dev_rsqrt(T);
        const R inv2T = R(0.5) * rT * rT;
        Fm[0] = R(0.886226925452758013649083741671) * rT;
        if (L > 0) {
            const R e = dev_exp(-T);
#pragma unroll
            for (int m = 0; m < L; ++m) Fm[m + 1] = (R(2.0 * m + 1.0) * Fm[m] - e) * inv2T;
        }
    }
    const R m2a = R(-2.0) * alpha;
    R s = pref;
#pragma unroll
    for (int m = 0; m <= L; ++m) {
        F[m] = s * Fm[m];
        s *= m2a;
    }
}

// ---------------------------------------------------------------------------
// Hermite Coulomb integrals R_{tuv} = R^0_{tuv}, t+u+v <= L, computed in place
// from the seeds F[n] = pref (-2a)^n F_n (the recursion of
// cython/util.pxi:33-50, made iterative: level n from L down to 0, degree from
// L-n down to 1, reusing the level n+1 values in the same registers).
// ---------------------------------------------------------------------------
template <int L, typename Real = double>
__device__ inline void hermite_R(Real X, Real Y, Real Z, const Real* F, Real* R)
{
    R[0] = F[L];
    sfor<L>([&](auto in) {
        constexpr int n = L - 1 - decltype(in)::value;
        sfor<L - n>([&](auto id) {
            constexpr int d = L - n - decltype(id)::value;       // degree L-n down to 1
            sfor<nherm(d) - nherm(d - 1)>([&](auto ik) {
                constexpr int k = nherm(d - 1) + decltype(ik)::value;
                constexpr Herm3 h = hdecode(k);
                constexpr int t = h.t, u = h.u, v = h.v;
                Real val;
                if constexpr (t > 0) {
                    val = X * R[hidx(t - 1, u, v)];
                    if constexpr (t > 1) val = fma((Real)(t - 1), R[hidx(t - 2, u, v)], val);
                } else if constexpr (u > 0) {
                    val = Y * R[hidx(t, u - 1, v)];
                    if constexpr (u > 1) val = fma((Real)(u - 1), R[hidx(t, u - 2, v)], val);
                } else {
                    val = Z * R[hidx(t, u, v - 1)];
                    if constexpr (v > 1) val = fma((Real)(v - 1), R[hidx(t, u, v - 2)], val);
                }
                R[k] = val;
            });
        });
        R[0] = F[n];
    });
}

// ---------------------------------------------------------------------------
// 1-D Hermite expansion coefficients E^{ij}_t / K_AB for i <= LI, j <= LJ
// (cython/util.pxi:13-28 in iterative form; the Gaussian-product factor
// exp(-ab/p X_AB^2) is carried separately in the pair prefactor).
//   E^{i+1,j}_t = E^{ij}_{t-1}/(2p) + X_PA E^{ij}_t + (t+1) E^{ij}_{t+1}
//   E^{i,j+1}_t = E^{ij}_{t-1}/(2p) + X_PB E^{ij}_t + (t+1) E^{ij}_{t+1}
// Storage E[(i*(LJ+1)+j)*(LI+LJ+1)+t].  SIGN = -1 multiplies odd t by -1
// (the (-1)^{tau+nu+phi} factor of the ket, cython/twoe.pyx:89).
// ---------------------------------------------------------------------------
template <int LI, int LJ, int SIGN, typename Real = double>
__device__ inline void hermite_E(Real inv2p, Real PA, Real PB, Real* E)
{
    constexpr int NT = LI + LJ + 1;
    constexpr auto EI = [](int i, int j, int t) { return (i * (LJ + 1) + j) * NT + t; };
    sfor<(LI + 1) * (LJ + 1) * NT>([&](auto k) { E[k] = Real(0); });
    E[0] = Real(1);
    sfor<LI>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        sfor<i + 2>([&](auto it) {
            constexpr int t = decltype(it)::value;
            Real v = Real(0);
            if constexpr (t > 0) v = fma(inv2p, E[EI(i, 0, t - 1)], v);
            if constexpr (t <= i) v = fma(PA, E[EI(i, 0, t)], v);
            if constexpr (t + 1 <= i) v = fma((Real)(t + 1), E[EI(i, 0, t + 1)], v);
            E[EI(i + 1, 0, t)] = v;
        });
    });
    sfor<LI + 1>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        sfor<LJ>([&](auto ij) {
            constexpr int j = decltype(ij)::value;
            sfor<i + j + 2>([&](auto it) {
                constexpr int t = decltype(it)::value;
                Real v = Real(0);
                if constexpr (t > 0) v = fma(inv2p, E[EI(i, j, t - 1)], v);
                if constexpr (t <= i + j) v = fma(PB, E[EI(i, j, t)], v);
                if constexpr (t + 1 <= i + j) v = fma((Real)(t + 1), E[EI(i, j, t + 1)], v);
                E[EI(i, j + 1, t)] = v;
            });
        });
    });
    if constexpr (SIGN < 0) {
        sfor<LI + 1>([&](auto ii) {
            sfor<LJ + 1>([&](auto ij) {
                constexpr int i = decltype(ii)::value, j = decltype(ij)::value;
                sfor<(i + j + 1) / 2>([&](auto it) {
                    constexpr int t = 2 * decltype(it)::value + 1;
                    E[EI(i, j, t)] = -E[EI(i, j, t)];
                });
            });
        });
    }
}

// ---------------------------------------------------------------------------
// Runtime-L helpers (generic kernels; the class kernels use the templates)
// ---------------------------------------------------------------------------
__device__ inline void boys_runtime(int L, double T, const double* __restrict__ tab, double* F)
{
    if (T < BOYS_TMAX) {
        const int i = (int)(T * BOYS_INV_DT + 0.5);
        const double dT = (double)i * BOYS_DT - T;
        const double* r = tab + (size_t)L * BOYS_NT + i;
        double f = r[6 * BOYS_NT];
        f = r[5 * BOYS_NT] + dT * f * (1.0 / 6.0);
        f = r[4 * BOYS_NT] + dT * f * (1.0 / 5.0);
        f = r[3 * BOYS_NT] + dT * f * (1.0 / 4.0);
        f = r[2 * BOYS_NT] + dT * f * (1.0 / 3.0);
        f = r[1 * BOYS_NT] + dT * f * (1.0 / 2.0);
        f = r[0] + dT * f;
        F[L] = f;
        const double e = exp(-T);
        for (int m = L - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + e) / (2.0 * m + 1.0);
    } else {
        F[0] = 0.886226925452758013649083741671 / sqrt(T);
        const double e = exp(-T);
        for (int m = 0; m < L; ++m) F[m + 1] = ((2.0 * m + 1.0) * F[m] - e) / (2.0 * T);
    }
}

// R_tuv, t+u+v <= L, from seeds F[n] (already scaled by (-2 alpha)^n).
__device__ inline void hermite_R_rt(int L, double X, double Y, double Z, const double* F, double* R)
{
    R[0] = F[L];
    for (int n = L - 1; n >= 0; --n) {
        for (int d = L - n; d >= 1; --d)
            for (int t = d; t >= 0; --t)
                for (int u = d - t; u >= 0; --u) {
                    const int v = d - t - u;
                    double val;
                    if (t > 0) {
                        val = X * R[hidx(t - 1, u, v)];
                        if (t > 1) val += (double)(t - 1) * R[hidx(t - 2, u, v)];
                    } else if (u > 0) {
                        val = Y * R[hidx(t, u - 1, v)];
                        if (u > 1) val += (double)(u - 1) * R[hidx(t, u - 2, v)];
                    } else {
                        val = Z * R[hidx(t, u, v - 1)];
                        if (v > 1) val += (double)(v - 1) * R[hidx(t, u, v - 2)];
                    }
                    R[hidx(t, u, v)] = val;
                }
        R[0] = F[n];
    }
}

// 1-D Hermite coefficients / K_AB for i <= LI, j <= LJ:
// E[((i*(LJ+1)+j)*(LI+LJ+1)+t)].
__device__ inline void etab_rt(int LI, int LJ, double inv2p, double PA, double PB, double* E)
{
    const int NT = LI + LJ + 1;
    for (int k = 0; k < (LI + 1) * (LJ + 1) * NT; ++k) E[k] = 0.0;
    E[0] = 1.0;
    for (int i = 0; i < LI; ++i)
        for (int t = 0; t <= i + 1; ++t) {
            double v = 0.0;
            if (t > 0) v += inv2p * E[((i) * (LJ + 1)) * NT + t - 1];
            if (t <= i) v += PA * E[((i) * (LJ + 1)) * NT + t];
            if (t + 1 <= i) v += (double)(t + 1) * E[((i) * (LJ + 1)) * NT + t + 1];
            E[((i + 1) * (LJ + 1)) * NT + t] = v;
        }
    for (int i = 0; i <= LI; ++i)
        for (int j = 0; j < LJ; ++j)
            for (int t = 0; t <= i + j + 1; ++t) {
                double v = 0.0;
                if (t > 0) v += inv2p * E[(i * (LJ + 1) + j) * NT + t - 1];
                if (t <= i + j) v += PB * E[(i * (LJ + 1) + j) * NT + t];
                if (t + 1 <= i + j) v += (double)(t + 1) * E[(i * (LJ + 1) + j) * NT + t + 1];
                E[(i * (LJ + 1) + j + 1) * NT + t] = v;
            }
}

}  // namespace mmd
