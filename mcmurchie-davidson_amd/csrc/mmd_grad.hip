// Nuclear gradient of the converged RHF energy (mmd/forces.py:8-97) with the
// derivative integrals of cython/grad.pyx, on the device:
//
//   dE/dX = sum_pq P_pq (2 dH_pq/dX) + sum_pqrs Gamma_pqrs d(pq|rs)/dX
//           - 2 sum_pq W_pq dS_pq/dX            (+ dV_NN/dX on the host)
//
// with P the half density (C_occ C_occ^T), W = P F P and, for the
// two-electron part, the density products of the positions a canonical
// quartet occupies in the reference's derivative tensor (forces.py:73-84):
// Gamma = sum over the distinct permutations (a,b,c,d) of the quartet of
// 2 P_ba P_dc - P_ca P_bd.  The derivative of a Gaussian with respect to its
// centre raises and lowers its Cartesian power: in Hermite form the 1-D
// coefficient E^{ij}_t becomes 2a E^{i+1,j}_t - i E^{i-1,j}_t (first centre)
// or 2b E^{i,j+1}_t - j E^{i,j-1}_t (second centre), cython/grad.pyx:97-103.
// The fourth centre of an ERI follows from translational invariance,
// d/dD = -(d/dA + d/dB + d/dC).
#include <hip/hip_runtime.h>
#include <algorithm>
#include "mmd_device.h"
#include "mmd_kernels.h"

namespace mmd {

namespace {

// 1-D Hermite table of a primitive pair (divided by K_AB), i <= LI, j <= LJ.
struct Tab1 {
    const double* E;
    int LI, LJ;
    __device__ double e(int i, int j, int t) const
    {
        if (i < 0 || j < 0 || t < 0 || t > i + j || i > LI || j > LJ) return 0.0;
        return E[(i * (LJ + 1) + j) * (LI + LJ + 1) + t];
    }
    // mode 0: E^{ij}_t; 1: derivative for the first centre; 2: for the second
    __device__ double c(int i, int j, int t, int mode, double a, double b) const
    {
        if (mode == 1) return 2.0 * a * e(i + 1, j, t) - (double)i * e(i - 1, j, t);
        if (mode == 2) return 2.0 * b * e(i, j + 1, t) - (double)j * e(i, j - 1, t);
        return e(i, j, t);
    }
};

constexpr int GERI_TAB = (MAX_L + 2) * (MAX_L + 2) * (2 * MAX_L + 3);   // LI, LJ <= MAX_L + 1
constexpr int GONE_TAB = (MAX_L + 2) * (MAX_L + 4) * (2 * MAX_L + 5);   // LI <= MAX_L + 1, LJ <= MAX_L + 3
constexpr int GMAX_LTOT = MAX_LTOT + 1;

// A primitive pair with tables one order above its functions on both indices.
struct GPair {
    double a, b, p, w, P[3];
    double E[3][GERI_TAB];
};

__device__ void gpair(const BFDesc& fa, const BFDesc& fb, int ia, int ib, const double* __restrict__ pexp,
                      const double* __restrict__ pw, GPair& o)
{
    const double A[3] = {fa.x, fa.y, fa.z}, B[3] = {fb.x, fb.y, fb.z};
    const int la[3] = {fa.l, fa.m, fa.n}, lb[3] = {fb.l, fb.m, fb.n};
    o.a = pexp[fa.p0 + ia];
    o.b = pexp[fb.p0 + ib];
    o.p = o.a + o.b;
    double k = 1.0;
    for (int x = 0; x < 3; ++x) {
        o.P[x] = (o.a * A[x] + o.b * B[x]) / o.p;
        k *= exp(-o.a * o.b / o.p * (A[x] - B[x]) * (A[x] - B[x]));
        etab_rt(la[x] + 1, lb[x] + 1, 0.5 / o.p, o.P[x] - A[x], o.P[x] - B[x], o.E[x]);
    }
    o.w = pw[fa.p0 + ia] * pw[fb.p0 + ib] * k;
}

// d[ab|cd]/d(centre, axis) of one primitive quartet for the nine
// directions A x,y,z, B x,y,z, C x,y,z (cython/grad.pyx:107-350).
__device__ void prim_quartet_grad(const int* la, const int* lb, const int* lc, const int* ld, const GPair& P,
                                  const GPair& Q, const double* __restrict__ boys, double* out)
{
    const int L = la[0] + la[1] + la[2] + lb[0] + lb[1] + lb[2] + lc[0] + lc[1] + lc[2] + ld[0] + ld[1] + ld[2] + 1;
    const double alpha = P.p * Q.p / (P.p + Q.p);
    const double X = P.P[0] - Q.P[0], Y = P.P[1] - Q.P[1], Z = P.P[2] - Q.P[2];
    double F[GMAX_LTOT + 1];
    boys_runtime(L, alpha * (X * X + Y * Y + Z * Z), boys, F);
    double s = 1.0;
    for (int m = 0; m <= L; ++m) { F[m] *= s; s *= -2.0 * alpha; }
    double R[nherm(GMAX_LTOT)];
    hermite_R_rt(L, X, Y, Z, F, R);
    const double pref = 34.98683665524972497 / (P.p * Q.p * sqrt(P.p + Q.p)) * P.w * Q.w;
    Tab1 tb[3], tk[3];
    for (int x = 0; x < 3; ++x) {
        tb[x] = Tab1{P.E[x], la[x] + 1, lb[x] + 1};
        tk[x] = Tab1{Q.E[x], lc[x] + 1, ld[x] + 1};
    }
    for (int dir = 0; dir < 9; ++dir) {
        const int ax = dir % 3;
        const int cb = dir < 3 ? 1 : (dir < 6 ? 2 : 0);
        const int ck = dir >= 6 ? 1 : 0;
        int nb[3], nk[3], mb[3], mk[3];
        for (int x = 0; x < 3; ++x) {
            mb[x] = x == ax ? cb : 0;
            mk[x] = x == ax ? ck : 0;
            nb[x] = la[x] + lb[x] + (mb[x] ? 1 : 0);
            nk[x] = lc[x] + ld[x] + (mk[x] ? 1 : 0);
        }
        double val = 0.0;
        for (int t = 0; t <= nb[0]; ++t) {
            const double e0 = tb[0].c(la[0], lb[0], t, mb[0], P.a, P.b);
            if (e0 == 0.0) continue;
            for (int u = 0; u <= nb[1]; ++u) {
                const double e1 = e0 * tb[1].c(la[1], lb[1], u, mb[1], P.a, P.b);
                if (e1 == 0.0) continue;
                for (int v = 0; v <= nb[2]; ++v) {
                    const double eb = e1 * tb[2].c(la[2], lb[2], v, mb[2], P.a, P.b);
                    if (eb == 0.0) continue;
                    double inner = 0.0;
                    for (int tau = 0; tau <= nk[0]; ++tau) {
                        const double k0 = tk[0].c(lc[0], ld[0], tau, mk[0], Q.a, Q.b);
                        for (int nu = 0; nu <= nk[1]; ++nu) {
                            const double k1 = k0 * tk[1].c(lc[1], ld[1], nu, mk[1], Q.a, Q.b);
                            for (int phi = 0; phi <= nk[2]; ++phi) {
                                const double kk = k1 * tk[2].c(lc[2], ld[2], phi, mk[2], Q.a, Q.b);
                                const double sg = ((tau + nu + phi) & 1) ? -kk : kk;
                                inner = fma(sg, R[hidx(t + tau, u + nu, v + phi)], inner);
                            }
                        }
                    }
                    val = fma(eb, inner, val);
                }
            }
        }
        out[dir] = fma(val, pref, out[dir]);
    }
}

__device__ inline void tri_decode_g(int64_t idx, int64_t* r, int64_t* c)
{
    int64_t x = (int64_t)((sqrt(8.0 * (double)idx + 1.0) - 1.0) * 0.5);
    while (x * (x + 1) / 2 > idx) --x;
    while ((x + 1) * (x + 2) / 2 <= idx) ++x;
    *r = x;
    *c = idx - x * (x + 1) / 2;
}

// Two-electron part: one wavefront per canonical quartet i>=j, k>=l, ij>=kl
// (forces.py:62-80), its primitive quartets spread over the lanes and summed
// by a wavefront reduction; the workgroup sums the quartets' contributions
// per atom in LDS and adds them to grad once.
__global__ void __launch_bounds__(256) grad_eri_kernel(int N, const BFDesc* __restrict__ bfs,
                                                       const double* __restrict__ pexp,
                                                       const double* __restrict__ pw,
                                                       const double* __restrict__ boys,
                                                       const int* __restrict__ bf_atom, int natom,
                                                       const double* __restrict__ P, double* grad)
{
    extern __shared__ double gl[];   // 3 * natom
    for (int i = threadIdx.x; i < 3 * natom; i += blockDim.x) gl[i] = 0.0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t npair = (int64_t)N * (N + 1) / 2;
    const int64_t total = npair * (npair + 1) / 2;
    const size_t n = (size_t)N;
    const int64_t wave0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwave = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t idx = wave0; idx < total; idx += nwave) {   // uniform across the wavefront
        int64_t ij, kl, i, j, k, l;
        tri_decode_g(idx, &ij, &kl);
        tri_decode_g(ij, &i, &j);
        tri_decode_g(kl, &k, &l);
        // density weight of the quartet's distinct positions
        const int64_t pm[8][4] = {{i, j, k, l}, {k, l, i, j}, {j, i, l, k}, {l, k, j, i},
                                  {j, i, k, l}, {l, k, i, j}, {i, j, l, k}, {k, l, j, i}};
        double wgt = 0.0;
        for (int s = 0; s < 8; ++s) {
            bool dup = false;
            for (int r = 0; r < s && !dup; ++r)
                dup = pm[r][0] == pm[s][0] && pm[r][1] == pm[s][1] && pm[r][2] == pm[s][2] && pm[r][3] == pm[s][3];
            if (dup) continue;
            const int64_t a = pm[s][0], b = pm[s][1], c = pm[s][2], d = pm[s][3];
            wgt += 2.0 * P[b * n + a] * P[d * n + c] - P[c * n + a] * P[b * n + d];
        }
        if (wgt == 0.0) continue;
        const BFDesc fa = bfs[i], fb = bfs[j], fc = bfs[k], fd = bfs[l];
        const int la[3] = {fa.l, fa.m, fa.n}, lb[3] = {fb.l, fb.m, fb.n};
        const int lc[3] = {fc.l, fc.m, fc.n}, ld[3] = {fd.l, fd.m, fd.n};
        double g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        const int nab = fa.np * fb.np, ncd = fc.np * fd.np;
        GPair Pp, Qp;
        for (int q = lane; q < nab * ncd; q += 64) {
            const int pab = q / ncd, pcd = q - pab * ncd;
            gpair(fa, fb, pab / fb.np, pab % fb.np, pexp, pw, Pp);
            gpair(fc, fd, pcd / fd.np, pcd % fd.np, pexp, pw, Qp);
            prim_quartet_grad(la, lb, lc, ld, Pp, Qp, boys, g);
        }
        for (int off = 32; off > 0; off >>= 1)
            for (int d = 0; d < 9; ++d) g[d] += __shfl_xor(g[d], off);
        if (lane == 0) {
            const int at[4] = {bf_atom[i], bf_atom[j], bf_atom[k], bf_atom[l]};
            for (int x = 0; x < 3; ++x) {
                const double dA = g[x], dB = g[3 + x], dC = g[6 + x];
                atomicAdd(&gl[3 * at[0] + x], wgt * dA);
                atomicAdd(&gl[3 * at[1] + x], wgt * dB);
                atomicAdd(&gl[3 * at[2] + x], wgt * dC);
                atomicAdd(&gl[3 * at[3] + x], -wgt * (dA + dB + dC));
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 3 * natom; i += blockDim.x)
        if (gl[i] != 0.0) atomicAdd(&grad[i], gl[i]);
}

// One-electron part: one wavefront per basis-function pair i >= j; a lane
// item is a primitive pair with a share of the nuclei (every G-th), as in
// the one-electron integrals.  For centre A (function i) and B (function j)
// along each axis: the overlap and kinetic derivatives (overlapX,
// kineticX) and the nuclear attraction of every nucleus with a
// differentiated function (nuclear_attractionXb); for each nucleus C the
// derivative of its operator (nuclear_attractionXa), added to C's gradient.
__global__ void __launch_bounds__(256) grad_one_electron_kernel(int nbf, const BFDesc* __restrict__ bfs,
                                                                const double* __restrict__ pexp,
                                                                const double* __restrict__ pw, int natom,
                                                                const double* __restrict__ atoms,
                                                                const double* __restrict__ boys,
                                                                const int* __restrict__ bf_atom,
                                                                const double* __restrict__ P,
                                                                const double* __restrict__ W, double* grad)
{
    const int64_t idx = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int64_t npair = (int64_t)nbf * (nbf + 1) / 2;
    if (idx >= npair) return;   // uniform across the wavefront
    int64_t ii, jj;
    tri_decode_g(idx, &ii, &jj);
    const int i = (int)ii, j = (int)jj;
    const BFDesc a = bfs[i], b = bfs[j];
    const double A[3] = {a.x, a.y, a.z}, B[3] = {b.x, b.y, b.z};
    const int l1[3] = {a.l, a.m, a.n}, l2[3] = {b.l, b.m, b.n};
    const double f = (i == j) ? 1.0 : 2.0;
    const double pij = 2.0 * f * P[(size_t)i * nbf + j];
    const double wij = 2.0 * f * W[(size_t)i * nbf + j];
    double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0};
    const int npp = a.np * b.np;
    const int G = npp >= 64 ? 1 : 64 / npp;
    const int nitem = npp * G;
    for (int it = lane; it < nitem; it += 64) {
        const int ipp = it % npp, g = it / npp;
        const int ia = ipp / b.np, ib = ipp % b.np;
        const double ea = pexp[a.p0 + ia], eb = pexp[b.p0 + ib];
        const double p = ea + eb;
        const double w = pw[a.p0 + ia] * pw[b.p0 + ib];
        double Pc[3], kab = 1.0;
        double E[3][GONE_TAB];
        Tab1 tb[3];
        for (int x = 0; x < 3; ++x) {
            Pc[x] = (ea * A[x] + eb * B[x]) / p;
            kab *= exp(-ea * eb / p * (A[x] - B[x]) * (A[x] - B[x]));
            etab_rt(l1[x] + 1, l2[x] + 3, 0.5 / p, Pc[x] - A[x], Pc[x] - B[x], E[x]);
            tb[x] = Tab1{E[x], l1[x] + 1, l2[x] + 3};
        }
        if (g == 0) {
            // overlap and kinetic derivatives (cython/grad.pyx:352-507)
            const double pre = pow(3.141592653589793238462643383279 / p, 1.5) * kab * w;
            for (int ctr = 1; ctr <= 2; ++ctr)
                for (int ax = 0; ax < 3; ++ax) {
                    double S0[3], Tk[3];
                    for (int x = 0; x < 3; ++x) {
                        const int m = x == ax ? ctr : 0;
                        const double jq = (double)l2[x];
                        S0[x] = tb[x].c(l1[x], l2[x], 0, m, ea, eb);
                        Tk[x] = (2.0 * jq + 1.0) * eb * tb[x].c(l1[x], l2[x], 0, m, ea, eb)
                              - 2.0 * eb * eb * tb[x].c(l1[x], l2[x] + 2, 0, m, ea, eb)
                              - 0.5 * jq * (jq - 1.0) * (l2[x] >= 2 ? tb[x].c(l1[x], l2[x] - 2, 0, m, ea, eb) : 0.0);
                    }
                    const double ds = S0[0] * S0[1] * S0[2] * pre;
                    const double dt = (Tk[0] * S0[1] * S0[2] + Tk[1] * S0[0] * S0[2] + Tk[2] * S0[0] * S0[1]) * pre;
                    double& gc = ctr == 1 ? gA[ax] : gB[ax];
                    gc += pij * dt - wij * ds;
                }
        }
        // nuclear attraction (cython/grad.pyx:509-)
        const int Lv = l1[0] + l2[0] + l1[1] + l2[1] + l1[2] + l2[2] + 1;
        const double vpre = 2.0 * 3.141592653589793238462643383279 / p * kab * w;
        for (int at = g; at < natom; at += G) {
            const double* nuc = atoms + 4 * at;
            const double Z = nuc[0];
            const double X = Pc[0] - nuc[1], Y = Pc[1] - nuc[2], Wd = Pc[2] - nuc[3];
            double F[2 * MAX_L + 2];
            boys_runtime(Lv, p * (X * X + Y * Y + Wd * Wd), boys, F);
            double sc = 1.0;
            for (int m = 0; m <= Lv; ++m) { F[m] *= sc; sc *= -2.0 * p; }
            double R[nherm(2 * MAX_L + 1)];
            hermite_R_rt(Lv, X, Y, Wd, F, R);
            double opd[3] = {0, 0, 0};
            for (int ctr = 1; ctr <= 2; ++ctr)
                for (int ax = 0; ax < 3; ++ax) {
                    double val = 0.0;
                    const int m0 = ax == 0 ? ctr : 0, m1 = ax == 1 ? ctr : 0, m2 = ax == 2 ? ctr : 0;
                    for (int t = 0; t <= l1[0] + l2[0] + (m0 ? 1 : 0); ++t)
                        for (int u = 0; u <= l1[1] + l2[1] + (m1 ? 1 : 0); ++u)
                            for (int v = 0; v <= l1[2] + l2[2] + (m2 ? 1 : 0); ++v)
                                val += tb[0].c(l1[0], l2[0], t, m0, ea, eb) * tb[1].c(l1[1], l2[1], u, m1, ea, eb)
                                     * tb[2].c(l1[2], l2[2], v, m2, ea, eb) * R[hidx(t, u, v)];
                    double& gc = ctr == 1 ? gA[ax] : gB[ax];
                    gc += pij * (-Z) * val * vpre;
                }
            for (int t = 0; t <= l1[0] + l2[0]; ++t)
                for (int u = 0; u <= l1[1] + l2[1]; ++u)
                    for (int v = 0; v <= l1[2] + l2[2]; ++v) {
                        const double e3 = tb[0].e(l1[0], l2[0], t) * tb[1].e(l1[1], l2[1], u)
                                        * tb[2].e(l1[2], l2[2], v);
                        opd[0] -= e3 * R[hidx(t + 1, u, v)];
                        opd[1] -= e3 * R[hidx(t, u + 1, v)];
                        opd[2] -= e3 * R[hidx(t, u, v + 1)];
                    }
            for (int x = 0; x < 3; ++x) atomicAdd(&grad[3 * at + x], pij * (-Z) * opd[x] * vpre);
        }
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int x = 0; x < 3; ++x) {
            gA[x] += __shfl_xor(gA[x], off);
            gB[x] += __shfl_xor(gB[x], off);
        }
    if (lane != 0) return;
    for (int x = 0; x < 3; ++x) {
        atomicAdd(&grad[3 * bf_atom[i] + x], gA[x]);
        atomicAdd(&grad[3 * bf_atom[j] + x], gB[x]);
    }
}

}  // namespace

hipError_t launch_rhf_gradient(int nbf, const BFDesc* bfs, const double* pexp, const double* pw, int natom,
                               const double* atoms, const double* boys, const int* bf_atom, const double* P,
                               const double* W, double* grad, hipStream_t st)
{
    if (nbf <= 0 || natom <= 0) return hipSuccess;
    const int64_t npair = (int64_t)nbf * (nbf + 1) / 2;
    hipLaunchKernelGGL(grad_one_electron_kernel, dim3((unsigned)((npair + 3) / 4)), dim3(256), 0, st, nbf, bfs, pexp,
                       pw, natom, atoms, boys, bf_atom, P, W, grad);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t total = npair * (npair + 1) / 2;   // quartets: one wavefront each, four per workgroup
    const int64_t grid = std::min<int64_t>(std::max<int64_t>(1, (total + 3) / 4), 16384);
    hipLaunchKernelGGL(grad_eri_kernel, dim3((unsigned)grid), dim3(256), sizeof(double) * 3 * natom, st, nbf, bfs,
                       pexp, pw, boys, bf_atom, natom, P, grad);
    return hipGetLastError();
}

}  // namespace mmd
