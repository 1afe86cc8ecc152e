// One-electron integrals by shell pair for gfx950: overlap, kinetic energy
// and nuclear attraction (and, in a second pass, dipole and angular-momentum
// integrals about an origin) of every Cartesian component of a shell pair,
// with the pair class's angular momenta known at compile time so that the
// Hermite tables, Boys values and R_tuv live in registers.
//
// One wavefront per shell pair; lane items are (primitive pair, group of the
// molecule's nuclei), as many as fill the wavefront for short contractions;
// a wavefront reduction sums the items.  Everything the kinetic, dipole and
// angular-momentum integrals need is derived from the pair's own table
// E^{ij}_t (i <= l_a, j <= l_b) through the Hermite recursions
//   E^{i,j+1}_0 = X_PB E_0 + E_1,   E^{i+1,j}_0 = X_PA E_0 + E_1,
//   E^{i,j+2}_0 = (X_PB^2 + 1/(2p)) E_0 + 2 X_PB E_1 + 2 E_2,
// with E_t = E^{ij}_t.
//
// The Schwarz diagonal (ab|ab) of every component pair of a shell pair is
// computed the same way (schwarz_pair_kernel), with lane items (bra
// primitive pair, ket primitive pair) of the pair with itself.
//
// Reference semantics: cython/onee.pyx:14-192 (the primitive integrals and
// their contraction), mmd/molecule.py:253-276 (assembly of S, T, V, M, L),
// mmd/molecule.py:93-99 and cython/twoe.pyx:34-94 (the diagonal ERIs of the
// Schwarz screen).
#include <hip/hip_runtime.h>

#include "mmd_device.h"
#include "mmd_kernels.h"

namespace mmd {
namespace {

constexpr double PI = 3.141592653589793238462643383279;

template <int LA, int LB, bool MULT>
__global__ void __launch_bounds__(256) onee_pair_kernel(OneeArgs A, const int2* __restrict__ list, int npairs)
{
    constexpr int NA = ncart(LA), NB = ncart(LB), NAB = NA * NB, L = LA + LB;
    constexpr int NT = L + 1;
    constexpr int NE = (LA + 1) * (LB + 1) * NT;
    constexpr int NACC = MULT ? 6 * NAB : 3 * NAB;
    const int wid = (int)(((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (wid >= npairs) return;   // uniform across the wavefront
    const int2 pr = list[wid];
    const OneeShell a = A.shells[pr.x], b = A.shells[pr.y];
    const double Ac[3] = {a.x, a.y, a.z}, Bc[3] = {b.x, b.y, b.z}, Cc[3] = {A.cx, A.cy, A.cz};
    const double AB2 = (a.x - b.x) * (a.x - b.x) + (a.y - b.y) * (a.y - b.y) + (a.z - b.z) * (a.z - b.z);
    int at0 = 0, nat = A.natom;
    if (A.mol_atoms) {
        const int2 r = A.mol_atoms[a.mol];
        at0 = r.x;
        nat = r.y;
    }
    if (MULT) nat = 0;
    const int npp = a.np * b.np;
    int G = npp >= 64 ? 1 : 64 / npp;
    if (G > nat) G = nat > 0 ? nat : 1;
    double acc[NACC];
    sfor<NACC>([&](auto k) { acc[decltype(k)::value] = 0.0; });
    for (int it = lane; it < npp * G; it += 64) {
        const int ipp = it % npp, g = it / npp;
        const int ia = ipp / b.np, ib = ipp - ia * b.np;
        const double ea = A.exps[a.p0 + ia], eb = A.exps[b.p0 + ib];
        const double p = ea + eb, ip = 1.0 / p;
        const double w = A.wts[a.p0 + ia] * A.wts[b.p0 + ib] * exp(-ea * eb * ip * AB2);
        const double P[3] = {(ea * a.x + eb * b.x) * ip, (ea * a.y + eb * b.y) * ip, (ea * a.z + eb * b.z) * ip};
        double E[3][NE];
        sfor<3>([&](auto xx) {
            constexpr int x = decltype(xx)::value;
            hermite_E<LA, LB, 1, double>(0.5 * ip, P[x] - Ac[x], P[x] - Bc[x], E[x]);
        });
        if (g == 0) {
            const double sp = PI * ip;
            const double pre = sp * sqrt(sp) * w;
            sfor<NAB>([&](auto kk) {
                constexpr int k = decltype(kk)::value;
                constexpr int ca = k / NB, cb = k % NB;
                double S0[3], X1[3], X2[3], X3[3];
                sfor<3>([&](auto xx) {
                    constexpr int x = decltype(xx)::value;
                    constexpr int i = x == 0 ? comp_lx(LA, ca) : (x == 1 ? comp_ly(LA, ca) : comp_lz(LA, ca));
                    constexpr int j = x == 0 ? comp_lx(LB, cb) : (x == 1 ? comp_ly(LB, cb) : comp_lz(LB, cb));
                    constexpr int e0 = (i * (LB + 1) + j) * NT;
                    const double E0 = E[x][e0];
                    double E1 = 0.0;
                    if constexpr (i + j >= 1) E1 = E[x][e0 + 1];
                    const double pb = P[x] - Bc[x];
                    S0[x] = E0;
                    if constexpr (!MULT) {
                        // kinetic piece (cython/onee.pyx:109-136):
                        // (2j+1) b E^{ij}_0 - 2 b^2 E^{i,j+2}_0 - j(j-1)/2 E^{i,j-2}_0
                        double e2 = (pb * pb + 0.5 * ip) * E0 + 2.0 * pb * E1;
                        if constexpr (i + j >= 2) e2 += 2.0 * E[x][e0 + 2];
                        double tk = (2.0 * j + 1.0) * eb * E0 - 2.0 * eb * eb * e2;
                        if constexpr (j >= 2) tk -= 0.5 * j * (j - 1.0) * E[x][(i * (LB + 1) + j - 2) * NT];
                        X1[x] = tk;
                    } else {
                        // dipole (cython/onee.pyx:81-107): E_1 + (P - C) E_0
                        X1[x] = E1 + (P[x] - Cc[x]) * E0;
                        // angular momentum (cython/onee.pyx:139-173):
                        // S1 = E^{i+1,j}_0 + (A - C) E_0,  D1 = j E^{i,j-1}_0 - 2 b E^{i,j+1}_0
                        X2[x] = (P[x] - Ac[x]) * E0 + E1 + (Ac[x] - Cc[x]) * E0;
                        double d1 = -2.0 * eb * (pb * E0 + E1);
                        if constexpr (j >= 1) d1 += (double)j * E[x][(i * (LB + 1) + j - 1) * NT];
                        X3[x] = d1;
                    }
                });
                if constexpr (!MULT) {
                    acc[k] += S0[0] * S0[1] * S0[2] * pre;
                    acc[NAB + k] += (X1[0] * S0[1] * S0[2] + S0[0] * X1[1] * S0[2] + S0[0] * S0[1] * X1[2]) * pre;
                } else {
                    acc[k] += X1[0] * S0[1] * S0[2] * pre;
                    acc[NAB + k] += S0[0] * X1[1] * S0[2] * pre;
                    acc[2 * NAB + k] += S0[0] * S0[1] * X1[2] * pre;
                    acc[3 * NAB + k] += -S0[0] * (X2[1] * X3[2] - X2[2] * X3[1]) * pre;
                    acc[4 * NAB + k] += -S0[1] * (X2[2] * X3[0] - X2[0] * X3[2]) * pre;
                    acc[5 * NAB + k] += -S0[2] * (X2[0] * X3[1] - X2[1] * X3[0]) * pre;
                }
            });
        }
        if constexpr (!MULT) {
            // nuclear attraction of the nuclei g, g+G, ... of the pair's
            // molecule (cython/onee.pyx:175-192)
            const double vpre = -2.0 * PI * ip * w;
            for (int at = g; at < nat; at += G) {
                const double* nuc = A.atoms + 4 * (size_t)(at0 + at);
                const double X = P[0] - nuc[1], Y = P[1] - nuc[2], W = P[2] - nuc[3];
                double F[L + 1];
                boys_seeds<L, double>(p * (X * X + Y * Y + W * W), p, nuc[0] * vpre, A.boys + (size_t)L * BOYS_NT, F);
                double R[nherm(L)];
                hermite_R<L, double>(X, Y, W, F, R);
                sfor<NAB>([&](auto kk) {
                    constexpr int k = decltype(kk)::value;
                    constexpr int ca = k / NB, cb = k % NB;
                    constexpr int ix = comp_lx(LA, ca), iy = comp_ly(LA, ca), iz = comp_lz(LA, ca);
                    constexpr int jx = comp_lx(LB, cb), jy = comp_ly(LB, cb), jz = comp_lz(LB, cb);
                    double val = 0.0;
                    sfor<ix + jx + 1>([&](auto tt) {
                        constexpr int t = decltype(tt)::value;
                        double vu = 0.0;
                        sfor<iy + jy + 1>([&](auto uu) {
                            constexpr int u = decltype(uu)::value;
                            double vv = 0.0;
                            sfor<iz + jz + 1>([&](auto ww) {
                                constexpr int v = decltype(ww)::value;
                                vv = fma(E[2][(iz * (LB + 1) + jz) * NT + v], R[hidx(t, u, v)], vv);
                            });
                            vu = fma(E[1][(iy * (LB + 1) + jy) * NT + u], vv, vu);
                        });
                        val = fma(E[0][(ix * (LB + 1) + jx) * NT + t], vu, val);
                    });
                    acc[2 * NAB + k] += val;
                });
            }
        }
    }
    sfor<NACC>([&](auto k) {
        double v = acc[decltype(k)::value];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        acc[decltype(k)::value] = v;
    });
    if (lane != 0) return;
    const size_t n = (size_t)A.nbf, nn = n * n;
    const bool same = pr.x == pr.y;
    sfor<NAB>([&](auto kk) {
        constexpr int k = decltype(kk)::value;
        constexpr int ca = k / NB, cb = k % NB;
        if (same && ca < cb) return;   // one shell: each element pair once, copied to its transpose
        const size_t i = (size_t)a.bf0 + ca, j = (size_t)b.bf0 + cb;
        const double gs = A.gcomp[i] * A.gcomp[j];
        if constexpr (!MULT) {
            A.S[i * n + j] = A.S[j * n + i] = gs * acc[k];
            A.T[i * n + j] = A.T[j * n + i] = gs * acc[NAB + k];
            A.V[i * n + j] = A.V[j * n + i] = gs * acc[2 * NAB + k];
        } else {
            sfor<3>([&](auto xx) {
                constexpr int x = decltype(xx)::value;
                A.M[x * nn + i * n + j] = A.M[x * nn + j * n + i] = gs * acc[x * NAB + k];
                A.Lm[x * nn + i * n + j] = gs * acc[(3 + x) * NAB + k];
                A.Lm[x * nn + j * n + i] = -gs * acc[(3 + x) * NAB + k];
            });
        }
    });
}


// (ab|ab) for every component pair of the shell pair list[w] (one
// wavefront); out[i][j] = out[j][i] = (ij|ij).
template <int LA, int LB>
__global__ void __launch_bounds__(256) schwarz_pair_kernel(OneeArgs A, const int2* __restrict__ list, int npairs,
                                                           double* __restrict__ out)
{
    constexpr int NA = ncart(LA), NB = ncart(LB), NAB = NA * NB, LAB = LA + LB, L = 2 * LAB;
    constexpr int NT = LAB + 1;
    constexpr int NE = (LA + 1) * (LB + 1) * NT;
    constexpr double TWO_PI_2_5 = 34.98683665524972497;   // 2 pi^{5/2}
    const int wid = (int)(((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (wid >= npairs) return;   // uniform across the wavefront
    const int2 pr = list[wid];
    const OneeShell a = A.shells[pr.x], b = A.shells[pr.y];
    const double Ac[3] = {a.x, a.y, a.z}, Bc[3] = {b.x, b.y, b.z};
    const double AB2 = (a.x - b.x) * (a.x - b.x) + (a.y - b.y) * (a.y - b.y) + (a.z - b.z) * (a.z - b.z);
    const int npp = a.np * b.np;
    double acc[NAB];
    sfor<NAB>([&](auto k) { acc[decltype(k)::value] = 0.0; });
    // primitive pair k of the shell pair: exponent sum, weight x K_AB, centre
    auto prim = [&](int k, double& p, double& w, double* P) {
        const int ia = k / b.np, ib = k - (k / b.np) * b.np;
        const double ea = A.exps[a.p0 + ia], eb = A.exps[b.p0 + ib];
        p = ea + eb;
        const double ip = 1.0 / p;
        w = A.wts[a.p0 + ia] * A.wts[b.p0 + ib] * exp(-ea * eb * ip * AB2);
        P[0] = (ea * a.x + eb * b.x) * ip;
        P[1] = (ea * a.y + eb * b.y) * ip;
        P[2] = (ea * a.z + eb * b.z) * ip;
    };
    for (int it = lane; it < npp * npp; it += 64) {
        const int i = it / npp, j = it - (it / npp) * npp;
        double p, wp, P[3], q, wq, Q[3];
        prim(i, p, wp, P);
        prim(j, q, wq, Q);
        double Ep[3][NE], Eq[3][NE];
        sfor<3>([&](auto xx) {
            constexpr int x = decltype(xx)::value;
            hermite_E<LA, LB, 1, double>(0.5 / p, P[x] - Ac[x], P[x] - Bc[x], Ep[x]);
            hermite_E<LA, LB, -1, double>(0.5 / q, Q[x] - Ac[x], Q[x] - Bc[x], Eq[x]);
        });
        const double rs = rsqrt(p + q);
        const double alpha = p * q * rs * rs;
        const double X = P[0] - Q[0], Y = P[1] - Q[1], Z = P[2] - Q[2];
        double F[L + 1];
        boys_seeds<L, double>(alpha * (X * X + Y * Y + Z * Z), alpha, TWO_PI_2_5 * wp * wq / (p * q) * rs,
                              A.boys + (size_t)L * BOYS_NT, F);
        double R[nherm(L)];
        hermite_R<L, double>(X, Y, Z, F, R);
        sfor<NAB>([&](auto kk) {
            constexpr int k = decltype(kk)::value;
            constexpr int ca = k / NB, cb = k % NB;
            constexpr int ex = (comp_lx(LA, ca) * (LB + 1) + comp_lx(LB, cb)) * NT;
            constexpr int ey = (comp_ly(LA, ca) * (LB + 1) + comp_ly(LB, cb)) * NT;
            constexpr int ez = (comp_lz(LA, ca) * (LB + 1) + comp_lz(LB, cb)) * NT;
            constexpr int nx = comp_lx(LA, ca) + comp_lx(LB, cb) + 1;
            constexpr int ny = comp_ly(LA, ca) + comp_ly(LB, cb) + 1;
            constexpr int nz = comp_lz(LA, ca) + comp_lz(LB, cb) + 1;
            double val = 0.0;
            sfor<nx>([&](auto t_) {
                constexpr int t = decltype(t_)::value;
                sfor<ny>([&](auto u_) {
                    constexpr int u = decltype(u_)::value;
                    sfor<nz>([&](auto v_) {
                        constexpr int v = decltype(v_)::value;
                        // sum over the ket's Hermite indices for bra index (t,u,v)
                        double s = 0.0;
                        sfor<nx>([&](auto ta_) {
                            constexpr int ta = decltype(ta_)::value;
                            sfor<ny>([&](auto ua_) {
                                constexpr int ua = decltype(ua_)::value;
                                const double exy = Eq[0][ex + ta] * Eq[1][ey + ua];
                                sfor<nz>([&](auto va_) {
                                    constexpr int va = decltype(va_)::value;
                                    s = fma(exy * Eq[2][ez + va], R[hidx(t + ta, u + ua, v + va)], s);
                                });
                            });
                        });
                        val = fma(Ep[0][ex + t] * Ep[1][ey + u] * Ep[2][ez + v], s, val);
                    });
                });
            });
            acc[k] += val;
        });
    }
    sfor<NAB>([&](auto k) {
        double v = acc[decltype(k)::value];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        acc[decltype(k)::value] = v;
    });
    if (lane != 0) return;
    const size_t n = (size_t)A.nbf;
    const bool same = pr.x == pr.y;
    sfor<NAB>([&](auto kk) {
        constexpr int k = decltype(kk)::value;
        constexpr int ca = k / NB, cb = k % NB;
        if (same && ca < cb) return;
        const size_t i = (size_t)a.bf0 + ca, j = (size_t)b.bf0 + cb;
        const double g = A.gcomp[i] * A.gcomp[j];
        out[i * n + j] = out[j * n + i] = g * g * acc[k];
    });
}

template <int LA, int LB, bool MULT>
hipError_t launch_pc(const OneeArgs& A, const int2* list, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL((onee_pair_kernel<LA, LB, MULT>), dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, A, list,
                       n);
    return hipGetLastError();
}

template <int LA, int LB>
hipError_t launch_schwarz_pc(const OneeArgs& A, const int2* list, int n, double* out, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL((schwarz_pair_kernel<LA, LB>), dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, A, list, n,
                       out);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_schwarz_shells(const OneeArgs& A, const int2* list, const int* off, double* out, hipStream_t st)
{
    hipError_t e = hipSuccess;
#define MMD_SCHWARZ_PC(pc, la, lb) \
    if (e == hipSuccess) e = launch_schwarz_pc<la, lb>(A, list + off[pc], off[pc + 1] - off[pc], out, st);
    MMD_SCHWARZ_PC(0, 0, 0)
    MMD_SCHWARZ_PC(1, 1, 0)
    MMD_SCHWARZ_PC(2, 1, 1)
    MMD_SCHWARZ_PC(3, 2, 0)
    MMD_SCHWARZ_PC(4, 2, 1)
    MMD_SCHWARZ_PC(5, 2, 2)
#undef MMD_SCHWARZ_PC
    return e;
}

hipError_t launch_one_electron_shells(const OneeArgs& A, const int2* list, const int* off, hipStream_t st)
{
    for (int pass = 0; pass < (A.M && A.Lm ? 2 : 1); ++pass) {
        const bool mult = pass == 1;
        hipError_t e = hipSuccess;
#define MMD_ONEE_PC(pc, la, lb)                                                                               \
    if (e == hipSuccess)                                                                                     \
        e = mult ? launch_pc<la, lb, true>(A, list + off[pc], off[pc + 1] - off[pc], st)                     \
                 : launch_pc<la, lb, false>(A, list + off[pc], off[pc + 1] - off[pc], st);
        MMD_ONEE_PC(0, 0, 0)
        MMD_ONEE_PC(1, 1, 0)
        MMD_ONEE_PC(2, 1, 1)
        MMD_ONEE_PC(3, 2, 0)
        MMD_ONEE_PC(4, 2, 1)
        MMD_ONEE_PC(5, 2, 2)
#undef MMD_ONEE_PC
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace mmd
