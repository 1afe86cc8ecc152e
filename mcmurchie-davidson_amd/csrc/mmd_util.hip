// Support kernels of libmmd for gfx950: screening/compaction of shell
// quartets, density preprocessing, symmetrisation, the generic per
// basis-function quartet ERI kernel (the ERI(a,b,c,d) entry point), the
// one-electron integrals and the stored-tensor J/K contraction.
#include <hip/hip_runtime.h>
#include "mmd_device.h"
#include "mmd_kernels.h"
#include "mmd_scf.h"

namespace mmd {

// ---------------------------------------------------------------------------
// Generic contracted ERI of one basis-function quartet (cython/twoe.pyx:34-94)
// with runtime angular momenta (l <= 2 per function).  Used by the ERI(a,b,c,d)
// entry point, the Schwarz diagonal, the stored tensor (doERIs) and formPT's
// unsymmetrised output; the Fock build itself uses the class kernels.
// ---------------------------------------------------------------------------
// Hermite expansion of one primitive pair: centre, Gaussian-product weight
// (times the contraction weights) and the three 1-D E tables.
struct PrimPairRT {
    double p, w, P[3], E[3][45];
};

__device__ inline void prim_pair_rt(const BFDesc& a, const BFDesc& b, int ia, int ib, const double* __restrict__ pexp,
                                    const double* __restrict__ pw, PrimPairRT& o)
{
    const double A[3] = {a.x, a.y, a.z}, B[3] = {b.x, b.y, b.z};
    const int la[3] = {a.l, a.m, a.n}, lb[3] = {b.l, b.m, b.n};
    const double ea = pexp[a.p0 + ia], eb = pexp[b.p0 + ib];
    o.p = ea + eb;
    double k = 1.0;
    for (int x = 0; x < 3; ++x) {
        o.P[x] = (ea * A[x] + eb * B[x]) / o.p;
        k *= exp(-ea * eb / o.p * (A[x] - B[x]) * (A[x] - B[x]));
        etab_rt(la[x], lb[x], 0.5 / o.p, o.P[x] - A[x], o.P[x] - B[x], o.E[x]);
    }
    o.w = pw[a.p0 + ia] * pw[b.p0 + ib] * k;
}

// (ab|cd) of one primitive quartet (cython/twoe.pyx:52-94).
__device__ inline double prim_quartet_rt(const BFDesc& a, const BFDesc& b, const BFDesc& c, const BFDesc& d,
                                         const PrimPairRT& P, const PrimPairRT& Q, const double* __restrict__ boys)
{
    const int L = a.l + a.m + a.n + b.l + b.m + b.n + c.l + c.m + c.n + d.l + d.m + d.n;
    const int la[3] = {a.l, a.m, a.n}, lb[3] = {b.l, b.m, b.n}, lc[3] = {c.l, c.m, c.n}, ld[3] = {d.l, d.m, d.n};
    const int nbx = la[0] + lb[0] + 1, nby = la[1] + lb[1] + 1, nbz = la[2] + lb[2] + 1;
    const int nkx = lc[0] + ld[0] + 1, nky = lc[1] + ld[1] + 1, nkz = lc[2] + ld[2] + 1;
    const double p = P.p, qq = Q.p;
    const double alpha = p * qq / (p + qq);
    const double X = P.P[0] - Q.P[0], Y = P.P[1] - Q.P[1], Z = P.P[2] - Q.P[2];
    const double T = alpha * (X * X + Y * Y + Z * Z);
    double F[MAX_LTOT + 1];
    boys_runtime(L, T, boys, F);
    double s = 1.0;
    for (int m = 0; m <= L; ++m) { F[m] *= s; s *= -2.0 * alpha; }
    double R[nherm(MAX_LTOT)];
    hermite_R_rt(L, X, Y, Z, F, R);
    double val = 0.0;
    for (int t = 0; t < nbx; ++t)
        for (int u = 0; u < nby; ++u)
            for (int v = 0; v < nbz; ++v) {
                const double eb3 = P.E[0][(la[0] * (lb[0] + 1) + lb[0]) * nbx + t]
                                 * P.E[1][(la[1] * (lb[1] + 1) + lb[1]) * nby + u]
                                 * P.E[2][(la[2] * (lb[2] + 1) + lb[2]) * nbz + v];
                for (int tau = 0; tau < nkx; ++tau)
                    for (int nu = 0; nu < nky; ++nu)
                        for (int phi = 0; phi < nkz; ++phi) {
                            const double sg = ((tau + nu + phi) & 1) ? -1.0 : 1.0;
                            val += eb3 * sg
                                 * Q.E[0][(lc[0] * (ld[0] + 1) + ld[0]) * nkx + tau]
                                 * Q.E[1][(lc[1] * (ld[1] + 1) + ld[1]) * nky + nu]
                                 * Q.E[2][(lc[2] * (ld[2] + 1) + ld[2]) * nkz + phi]
                                 * R[hidx(t + tau, u + nu, v + phi)];
                        }
            }
    return val * 34.98683665524972497 / (p * qq * sqrt(p + qq)) * P.w * Q.w;
}

__device__ double eri_bf_generic(const BFDesc& a, const BFDesc& b, const BFDesc& c, const BFDesc& d,
                                 const double* __restrict__ pexp, const double* __restrict__ pw,
                                 const double* __restrict__ boys)
{
    double total = 0.0;
    PrimPairRT P, Q;
    for (int ia = 0; ia < a.np; ++ia)
        for (int ib = 0; ib < b.np; ++ib) {
            prim_pair_rt(a, b, ia, ib, pexp, pw, P);
            for (int ic = 0; ic < c.np; ++ic)
                for (int id = 0; id < d.np; ++id) {
                    prim_pair_rt(c, d, ic, id, pexp, pw, Q);
                    total += prim_quartet_rt(a, b, c, d, P, Q, boys);
                }
        }
    return total;
}

// Schwarz diagonal (pq|pq), p >= q (mmd/molecule.py:93-99): one wavefront per
// basis-function pair, its primitive quartets spread over the lanes and
// summed by a wavefront reduction, so that heavily contracted core pairs
// (up to 9^4 primitive quartets) do not serialise on one lane.
__global__ void __launch_bounds__(256) eri_diag_wave_kernel(int N, const BFDesc* __restrict__ bfs,
                                                            const double* __restrict__ pexp,
                                                            const double* __restrict__ pw,
                                                            const double* __restrict__ boys, double* out,
                                                            const int2* __restrict__ plist, int64_t npl)
{
    const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int64_t npair = plist ? npl : (int64_t)N * (N + 1) / 2;
    if (w >= npair) return;   // uniform across the wavefront
    int p, q;
    if (plist) {   // a block-diagonal basis: the pairs of each molecule
        p = plist[w].x;
        q = plist[w].y;
    } else {
        int64_t r = (int64_t)((sqrt(8.0 * (double)w + 1.0) - 1.0) * 0.5);
        while (r * (r + 1) / 2 > w) --r;
        while ((r + 1) * (r + 2) / 2 <= w) ++r;
        p = (int)r;
        q = (int)(w - r * (r + 1) / 2);
    }
    const BFDesc a = bfs[p], b = bfs[q];
    const int nab = a.np * b.np;
    double total = 0.0;
    PrimPairRT P, Q;
    for (int k = lane; k < nab * nab; k += 64) {
        const int i = k / nab, j = k - i * nab;
        prim_pair_rt(a, b, i / b.np, i % b.np, pexp, pw, P);
        prim_pair_rt(a, b, j / b.np, j % b.np, pexp, pw, Q);
        total += prim_quartet_rt(a, b, a, b, P, Q, boys);
    }
    for (int off = 32; off > 0; off >>= 1) total += __shfl_xor(total, off);
    if (lane == 0) {
        out[(size_t)p * N + q] = total;
        out[(size_t)q * N + p] = total;
    }
}

hipError_t launch_eri_diag(int N, const BFDesc* bfs, const double* pexp, const double* pw, const double* boys,
                           double* out, hipStream_t st, const int2* plist, int64_t npl)
{
    const int64_t npair = plist ? npl : (int64_t)N * (N + 1) / 2;
    if (npair <= 0) return hipSuccess;
    if (plist) {   // only the molecules' own pairs are computed: clear the rest
        const hipError_t e = launch_fill((size_t)N * N, out, 0.0, st);
        if (e != hipSuccess) return e;
    }
    const int64_t grid = (npair + 3) / 4;   // four wavefronts per 256-thread block
    hipLaunchKernelGGL(eri_diag_wave_kernel, dim3((unsigned)grid), dim3(256), 0, st, N, bfs, pexp, pw, boys, out,
                       plist, npl);
    return hipGetLastError();
}

__global__ void __launch_bounds__(256) eri_bf_list_kernel(int64_t nq, const int64_t* __restrict__ ijkl,
                                                          const BFDesc* __restrict__ bfs,
                                                          const double* __restrict__ pexp,
                                                          const double* __restrict__ pw,
                                                          const double* __restrict__ boys, double* out)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    out[q] = eri_bf_generic(bfs[ijkl[4 * q + 0]], bfs[ijkl[4 * q + 1]], bfs[ijkl[4 * q + 2]], bfs[ijkl[4 * q + 3]],
                            pexp, pw, boys);
}

hipError_t launch_eri_bf_list(int64_t nq, const int64_t* ijkl, const BFDesc* bfs, const double* pexp,
                              const double* pw, const double* boys, double* out, hipStream_t st)
{
    if (nq <= 0) return hipSuccess;
    const int64_t grid = (nq + 255) / 256;
    hipLaunchKernelGGL(eri_bf_list_kernel, dim3((unsigned)grid), dim3(256), 0, st, nq, ijkl, bfs, pexp, pw, boys,
                       out);
    return hipGetLastError();
}

// Triangular index -> (r, c), r >= c.
__device__ inline void tri_decode(int64_t idx, int64_t* r, int64_t* c)
{
    int64_t x = (int64_t)((sqrt(8.0 * (double)idx + 1.0) - 1.0) * 0.5);
    while (x * (x + 1) / 2 > idx) --x;
    while ((x + 1) * (x + 2) / 2 <= idx) ++x;
    *r = x;
    *c = idx - x * (x + 1) / 2;
}

// Canonical basis-function quartets i>=j, k>=l, ij>=kl (cython/twoe.pyx:15-21,
// cython/fock.pyx:38-44), enumerated as a triangle of pair indices.
//   mode 0: store (ij|kl) at all 8 permutations (doERIs, cython/twoe.pyx:23-30)
//   mode 1: formPT's unsymmetrised update (cython/fock.pyx:46-85)
__global__ void __launch_bounds__(256) canonical_quartet_kernel(int mode, int N, const BFDesc* __restrict__ bfs,
                                                                const double* __restrict__ pexp,
                                                                const double* __restrict__ pw,
                                                                const double* __restrict__ boys,
                                                                const double* __restrict__ D,
                                                                const double* __restrict__ absD,
                                                                const double* __restrict__ sqS, double tol,
                                                                double* __restrict__ out)
{
    const int64_t npair = (int64_t)N * (N + 1) / 2;
    const int64_t total = npair * (npair + 1) / 2;
    const size_t n = (size_t)N;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        int64_t ij, kl, i, j, k, l;
        tri_decode(idx, &ij, &kl);
        tri_decode(ij, &i, &j);
        tri_decode(kl, &k, &l);
        if (mode == 0) {
            const double v = eri_bf_generic(bfs[i], bfs[j], bfs[k], bfs[l], pexp, pw, boys);
            out[((i * n + j) * n + k) * n + l] = v;
            out[((k * n + l) * n + i) * n + j] = v;
            out[((j * n + i) * n + l) * n + k] = v;
            out[((l * n + k) * n + j) * n + i] = v;
            out[((j * n + i) * n + k) * n + l] = v;
            out[((l * n + k) * n + i) * n + j] = v;
            out[((i * n + j) * n + l) * n + k] = v;
            out[((k * n + l) * n + j) * n + i] = v;
        } else {
            const double bound = sqS[i * n + j] * sqS[k * n + l]
                * fmax(fmax(fmax(4.0 * absD[i * n + j], 4.0 * absD[k * n + l]), fmax(absD[i * n + k], absD[i * n + l])),
                       fmax(absD[j * n + k], absD[j * n + l]));
            if (bound < tol) continue;
            const double s12 = (i == j) ? 1.0 : 2.0;
            const double s34 = (k == l) ? 1.0 : 2.0;
            const double s1234 = (i == k && j == l) ? 1.0 : 2.0;
            const double eri = s12 * s34 * s1234 * eri_bf_generic(bfs[i], bfs[j], bfs[k], bfs[l], pexp, pw, boys);
            atomicAdd(&out[i * n + j], D[k * n + l] * eri);
            atomicAdd(&out[k * n + l], D[i * n + j] * eri);
            atomicAdd(&out[i * n + k], -0.25 * D[j * n + l] * eri);
            atomicAdd(&out[j * n + l], -0.25 * D[i * n + k] * eri);
            atomicAdd(&out[i * n + l], -0.25 * D[j * n + k] * eri);
            atomicAdd(&out[k * n + j], -0.25 * D[i * n + l] * eri);
        }
    }
}

hipError_t launch_canonical_quartets(int mode, int N, const BFDesc* bfs, const double* pexp, const double* pw,
                                     const double* boys, const double* D, const double* absD, const double* sqS,
                                     double tol, double* out, hipStream_t st)
{
    const int64_t npair = (int64_t)N * (N + 1) / 2;
    const int64_t total = npair * (npair + 1) / 2;
    if (total <= 0) return hipSuccess;
    int64_t grid = (total + 255) / 256;
    if (grid > 65536) grid = 65536;
    hipLaunchKernelGGL(canonical_quartet_kernel, dim3((unsigned)grid), dim3(256), 0, st, mode, N, bfs, pexp, pw, boys,
                       D, absD, sqS, tol, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// One-electron integrals (cython/onee.pyx:14-192, assembled as
// mmd/molecule.py:253-276).  One wavefront per basis-function pair i >= j;
// each lane item is a primitive pair with a share of the nuclei, so that the
// Hermite tables of a primitive pair are built once per item and the nuclear
// attraction of short contractions is still spread over the lanes.  A
// wavefront reduction sums the lanes.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) one_electron_kernel(int nbf, const BFDesc* __restrict__ bfs,
                                                           const double* __restrict__ pexp,
                                                           const double* __restrict__ pw, int natom,
                                                           const double* __restrict__ atoms, double cx, double cy,
                                                           double cz, const double* __restrict__ boys, double* S,
                                                           double* T, double* V, double* M, double* Lm,
                                                           const int2* __restrict__ plist, int64_t npl,
                                                           const int2* __restrict__ mol_atoms)
{
    const int64_t idx = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int64_t npair = plist ? npl : (int64_t)nbf * (nbf + 1) / 2;
    if (idx >= npair) return;   // uniform across the wavefront
    int i, j;
    if (plist) {   // a block-diagonal basis: the pairs of each molecule
        i = plist[idx].x;
        j = plist[idx].y;
    } else {
        i = (int)((sqrt(8.0 * (double)idx + 1.0) - 1.0) * 0.5);
        while ((int64_t)i * (i + 1) / 2 > idx) --i;
        while ((int64_t)(i + 1) * (i + 2) / 2 <= idx) ++i;
        j = (int)(idx - (int64_t)i * (i + 1) / 2);
    }
    const BFDesc a = bfs[i], b = bfs[j];
    const double A[3] = {a.x, a.y, a.z}, B[3] = {b.x, b.y, b.z}, Cc[3] = {cx, cy, cz};
    const int l1[3] = {a.l, a.m, a.n}, l2[3] = {b.l, b.m, b.n};
    double s = 0, t = 0, v = 0, mu[3] = {0, 0, 0}, lam[3] = {0, 0, 0};
    // nuclei of the pair's own molecule: a contiguous range of the atom list
    int at0 = 0, nat_mol = natom;
    if (mol_atoms) {
        const int2 r = mol_atoms[a.pad];
        at0 = r.x;
        nat_mol = r.y;
    }
    // lane items: a primitive pair and a group g of the molecule's nuclei
    // (every G-th from g; G = 64 / npp for short contractions, else 1); the
    // item's Hermite tables serve its nuclei, and the items of group 0 add
    // the overlap, kinetic and multipole terms
    const int npp = a.np * b.np;
    const int G = npp >= 64 ? 1 : 64 / npp;
    const int nitem = npp * G;
    for (int it = lane; it < nitem; it += 64) {
        const int ipp = it % npp, g = it / npp;
        const int ia = ipp / b.np, ib = ipp % b.np;
        const double ea = pexp[a.p0 + ia], eb = pexp[b.p0 + ib];
        const double p = ea + eb;
        const double w = pw[a.p0 + ia] * pw[b.p0 + ib];
        double P[3], kab = 1.0;
        // E tables up to (l1+1, l2+2) per dimension: 4 x 5 x 8 entries.
        double E[3][160];
        int LI[3], LJ[3], NT[3];
        for (int x = 0; x < 3; ++x) {
            P[x] = (ea * A[x] + eb * B[x]) / p;
            kab *= exp(-ea * eb / p * (A[x] - B[x]) * (A[x] - B[x]));
            LI[x] = l1[x] + 1;
            LJ[x] = l2[x] + 2;
            NT[x] = LI[x] + LJ[x] + 1;
            etab_rt(LI[x], LJ[x], 0.5 / p, P[x] - A[x], P[x] - B[x], E[x]);
        }
#define E1(x, ii, jj, tt) \
    (((jj) < 0 || (tt) > (ii) + (jj)) ? 0.0 : E[x][((ii) * (LJ[x] + 1) + (jj)) * NT[x] + (tt)])
        if (g == 0) {
            const double pre = pow(3.141592653589793238462643383279 / p, 1.5) * kab * w;
            double S0[3], S1[3], D1[3], Tk[3], Dp[3];
            for (int x = 0; x < 3; ++x) {
                const int ii = l1[x], jj = l2[x];
                S0[x] = E1(x, ii, jj, 0);
                // kinetic pieces (cython/onee.pyx:109-136)
                Tk[x] = (2.0 * jj + 1.0) * eb * E1(x, ii, jj, 0) - 2.0 * eb * eb * E1(x, ii, jj + 2, 0)
                      - 0.5 * jj * (jj - 1.0) * E1(x, ii, jj - 2, 0);
                // dipole (cython/onee.pyx:81-107)
                Dp[x] = E1(x, ii, jj, 1) + (P[x] - Cc[x]) * E1(x, ii, jj, 0);
                // angular momentum (cython/onee.pyx:139-173)
                S1[x] = E1(x, ii + 1, jj, 0) + (A[x] - Cc[x]) * E1(x, ii, jj, 0);
                D1[x] = jj * E1(x, ii, jj - 1, 0) - 2.0 * eb * E1(x, ii, jj + 1, 0);
            }
            s += S0[0] * S0[1] * S0[2] * pre;
            t += (Tk[0] * S0[1] * S0[2] + Tk[1] * S0[0] * S0[2] + Tk[2] * S0[0] * S0[1]) * pre;
            mu[0] += Dp[0] * S0[1] * S0[2] * pre;
            mu[1] += S0[0] * Dp[1] * S0[2] * pre;
            mu[2] += S0[0] * S0[1] * Dp[2] * pre;
            lam[0] += -S0[0] * (S1[1] * D1[2] - S1[2] * D1[1]) * pre;
            lam[1] += -S0[1] * (S1[2] * D1[0] - S1[0] * D1[2]) * pre;
            lam[2] += -S0[2] * (S1[0] * D1[1] - S1[1] * D1[0]) * pre;
        }
        // nuclear attraction of every nucleus of the pair's molecule
        // (cython/onee.pyx:175-192)
        const int Lv = l1[0] + l2[0] + l1[1] + l2[1] + l1[2] + l2[2];
        const double vpre = -2.0 * 3.141592653589793238462643383279 / p * kab * w;
        for (int at = g; at < nat_mol; at += G) {
            const double* nuc = atoms + 4 * (at0 + at);
            const double Z = nuc[0];
            const double X = P[0] - nuc[1], Y = P[1] - nuc[2], W = P[2] - nuc[3];
            double F[2 * MAX_L + 1];
            boys_runtime(Lv, p * (X * X + Y * Y + W * W), boys, F);
            double sc = 1.0;
            for (int m = 0; m <= Lv; ++m) { F[m] *= sc; sc *= -2.0 * p; }
            double R[nherm(2 * MAX_L)];
            hermite_R_rt(Lv, X, Y, W, F, R);
            double val = 0.0;
            for (int tt = 0; tt <= l1[0] + l2[0]; ++tt)
                for (int uu = 0; uu <= l1[1] + l2[1]; ++uu)
                    for (int vv = 0; vv <= l1[2] + l2[2]; ++vv)
                        val += E1(0, l1[0], l2[0], tt) * E1(1, l1[1], l2[1], uu) * E1(2, l1[2], l2[2], vv)
                             * R[hidx(tt, uu, vv)];
            v += Z * val * vpre;
        }
#undef E1
    }
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_xor(s, off);
        t += __shfl_xor(t, off);
        v += __shfl_xor(v, off);
        for (int x = 0; x < 3; ++x) {
            mu[x] += __shfl_xor(mu[x], off);
            lam[x] += __shfl_xor(lam[x], off);
        }
    }
    if (lane != 0) return;
    const size_t n = (size_t)nbf;
    S[i * n + j] = S[j * n + i] = s;
    T[i * n + j] = T[j * n + i] = t;
    V[i * n + j] = V[j * n + i] = v;
    for (int x = 0; x < 3; ++x) {
        M[x * n * n + i * n + j] = M[x * n * n + j * n + i] = mu[x];
        Lm[x * n * n + i * n + j] = lam[x];
        Lm[x * n * n + j * n + i] = -lam[x];
    }
}

hipError_t launch_one_electron(int nbf, const BFDesc* bfs, const double* pexp, const double* pw, int natom,
                               const double* atoms, double cx, double cy, double cz, const double* boys, double* S,
                               double* T, double* V, double* M, double* Lm, hipStream_t st, const int2* plist,
                               int64_t npl, const int2* mol_atoms)
{
    const int64_t npair = plist ? npl : (int64_t)nbf * (nbf + 1) / 2;
    if (npair <= 0) return hipSuccess;
    if (plist) {   // only the molecules' own pairs are computed: clear the rest
        const size_t nn = (size_t)nbf * nbf;
        for (double* x : {S, T, V}) {
            const hipError_t e = launch_fill(nn, x, 0.0, st);
            if (e != hipSuccess) return e;
        }
        for (double* x : {M, Lm}) {
            const hipError_t e = launch_fill(3 * nn, x, 0.0, st);
            if (e != hipSuccess) return e;
        }
    }
    // four wavefronts (basis-function pairs) per 256-thread block
    hipLaunchKernelGGL(one_electron_kernel, dim3((unsigned)((npair + 3) / 4)), dim3(256), 0, st, nbf, bfs, pexp,
                       pw, natom, atoms, cx, cy, cz, boys, S, T, V, M, Lm, plist, npl, mol_atoms);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Shell-quartet screening and compaction for one quartet class.  Quartets are
// enumerated implicitly: bra pair P of class b (local index < nb), ket pair Q
// of class k (< nk); P >= Q when both classes are the same (tri).  A quartet
// survives if q_P q_Q max(4 D_ab, 4 D_cd, D_ac, D_ad, D_bc, D_bd) >= tol with
// D the shell-block maxima of |dP| -- an upper bound of the per
// basis-function test of cython/fock.pyx:46-57.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) screen_class_kernel(int boff, int nb, int koff, int nk, int tri,
                                                           const ShellPair* __restrict__ pairs,
                                                           const double* __restrict__ Dsh, int nshell, double tol,
                                                           int keep_all, int2* __restrict__ surv, int* count,
                                                           double split, int2* __restrict__ surv2, int* count2)
{
    const int64_t total = tri ? (int64_t)nb * (nb + 1) / 2 : (int64_t)nb * nk;
    const int lane = threadIdx.x & 63;
    const int64_t wave0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t base = wave0 * 64; base < total; base += nwaves * 64) {
        const int64_t idx = base + lane;
        bool keep = false, low = false;
        int P = 0, Q = 0;
        if (idx < total) {
            if (tri) {
                int r = (int)((sqrt(8.0 * (double)idx + 1.0) - 1.0) * 0.5);
                while ((int64_t)r * (r + 1) / 2 > idx) --r;
                while ((int64_t)(r + 1) * (r + 2) / 2 <= idx) ++r;
                P = r;
                Q = (int)(idx - (int64_t)r * (r + 1) / 2);
            } else {
                P = (int)(idx / nk);
                Q = (int)(idx % nk);
            }
            if (keep_all) {
                keep = pairs[boff + P].mol == pairs[koff + Q].mol;
            } else {
                const ShellPair bra = pairs[boff + P], ket = pairs[koff + Q];
                const size_t ns = (size_t)nshell;
                const double dm = fmax(fmax(fmax(4.0 * Dsh[bra.sa * ns + bra.sb], 4.0 * Dsh[ket.sa * ns + ket.sb]),
                                            fmax(Dsh[bra.sa * ns + ket.sa], Dsh[bra.sa * ns + ket.sb])),
                                       fmax(Dsh[bra.sb * ns + ket.sa], Dsh[bra.sb * ns + ket.sb]));
                const double bound = bra.q * ket.q * dm;
                keep = bound >= tol && bra.mol == ket.mol;   // quartets of one molecule only
                // with a second list: the quartets below `split` go there
                low = keep && surv2 && bound < split;
            }
        }
        const unsigned long long m = __ballot(keep && !low);
        const int cnt = __popcll(m);
        int pos = 0;
        if (lane == 0 && cnt) pos = atomicAdd(count, cnt);
        pos = __shfl(pos, 0);
        if (keep && !low) {
            const int off = __popcll(m & ((1ull << lane) - 1ull));
            surv[pos + off] = make_int2(boff + P, koff + Q);
        }
        if (surv2) {
            const unsigned long long m2 = __ballot(low);
            const int cnt2 = __popcll(m2);
            int pos2 = 0;
            if (lane == 0 && cnt2) pos2 = atomicAdd(count2, cnt2);
            pos2 = __shfl(pos2, 0);
            if (low) surv2[pos2 + __popcll(m2 & ((1ull << lane) - 1ull))] = make_int2(boff + P, koff + Q);
        }
    }
}

hipError_t launch_screen_class(int boff, int nb, int koff, int nk, int tri, const ShellPair* pairs, const double* Dsh,
                               int nshell, double tol, int keep_all, int2* surv, int* count, hipStream_t st,
                               double split, int2* surv2, int* count2)
{
    const int64_t total = tri ? (int64_t)nb * (nb + 1) / 2 : (int64_t)nb * nk;
    if (total <= 0) return hipSuccess;
    int64_t grid = (total + 255) / 256;
    if (grid > 8192) grid = 8192;
    hipLaunchKernelGGL(screen_class_kernel, dim3((unsigned)grid), dim3(256), 0, st, boff, nb, koff, nk, tri, pairs,
                       Dsh, nshell, tol, keep_all, surv, count, split, surv2, count2);
    return hipGetLastError();
}

// Primitive shell quartets of a class's survivor list: sum of npp(bra) npp(ket).
__global__ void count_prims_kernel(const int2* __restrict__ surv, const int* __restrict__ count,
                                   const ShellPair* __restrict__ pairs, unsigned long long* __restrict__ out)
{
    const int n = *count;
    unsigned long long s = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int2 q = surv[i];
        s += (unsigned long long)pairs[q.x].npp * (unsigned long long)pairs[q.y].npp;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(out, s);
}

hipError_t launch_count_prims(const int2* surv, const int* count, const ShellPair* pairs, unsigned long long* out,
                              hipStream_t st)
{
    hipLaunchKernelGGL(count_prims_kernel, dim3(256), dim3(256), 0, st, surv, count, pairs, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Fock build for a shell-quartet class with runtime angular momenta: one
// wavefront per shell quartet.  Per primitive quartet, lane 0 evaluates the
// Boys function and the R_tuv recursion into LDS and lanes 0-5 the 1-D
// Hermite tables; then every lane contracts its share of the component
// quartets (ab|cd) from LDS, accumulating over primitives in registers.
// Digestion as the class kernels (shell degeneracy, basis-function Schwarz
// x |dP| test, Coulomb and exchange updates).  Used for the high-L classes
// whose fully unrolled kernels are not worth their code size.
// ---------------------------------------------------------------------------
constexpr int GW_WAVES = 4;                         // waves per block
constexpr int GW_MAXC = (6 * 6 * 6 * 6 + 63) / 64;  // components per lane (d-d|d-d)

__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ void __launch_bounds__(64 * GW_WAVES) fock_wave_class_kernel(ClassArgs A, int la, int lb, int lc, int ld)
{
    __shared__ double Rs[GW_WAVES][nherm(MAX_LTOT)];
    __shared__ double Es[GW_WAVES][6][(MAX_L + 1) * (MAX_L + 1) * (2 * MAX_L + 1)];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nq = *A.count;
    const int na = ncart(la), nb = ncart(lb), nc = ncart(lc), nd = ncart(ld);
    const int nab = na * nb, ncd = nc * nd, ntot = nab * ncd;
    const int L = la + lb + lc + ld;
    const int N = A.N;
    double* R = Rs[wv];
    for (int64_t q = (int64_t)blockIdx.x * GW_WAVES + wv; q < nq; q += (int64_t)gridDim.x * GW_WAVES) {
        const int2 qd = A.quart[q];
        const ShellPair bra = A.pairs[qd.x];
        const ShellPair ket = A.pairs[qd.y];
        double acc[GW_MAXC];
        for (int i = 0; i < GW_MAXC; ++i) acc[i] = 0.0;
        for (int ip = 0; ip < bra.npp; ++ip) {
            const PrimPair* P = A.pp + bra.pp0 + ip;
            for (int jq = 0; jq < ket.npp; ++jq) {
                const PrimPair* Q = A.pp + ket.pp0 + jq;
                if (lane == 0) {
                    const double pq = P->p * Q->p, ppq = P->p + Q->p;
                    const double alpha = pq / ppq;
                    const double X = P->Px - Q->Px, Y = P->Py - Q->Py, Z = P->Pz - Q->Pz;
                    double F[MAX_LTOT + 1];
                    boys_runtime(L, alpha * (X * X + Y * Y + Z * Z), A.boys, F);
                    double s = 34.98683665524972497 / (pq * sqrt(ppq)) * P->K * Q->K;
                    for (int m = 0; m <= L; ++m) { F[m] *= s; s *= -2.0 * alpha; }
                    hermite_R_rt(L, X, Y, Z, F, R);
                } else if (lane <= 6) {
                    const int x = (lane - 1) % 3;
                    const bool isbra = lane <= 3;
                    const PrimPair* S = isbra ? P : Q;
                    const double PA = x == 0 ? S->PAx : (x == 1 ? S->PAy : S->PAz);
                    const double PB = x == 0 ? S->PBx : (x == 1 ? S->PBy : S->PBz);
                    etab_rt(isbra ? la : lc, isbra ? lb : ld, S->inv2p, PA, PB, Es[wv][lane - 1]);
                }
                wave_sync();
                const int ntb = la + lb + 1, ntk = lc + ld + 1;
                for (int i = 0; i < GW_MAXC; ++i) {
                    const int k = lane + 64 * i;
                    if (k >= ntot) break;
                    const int ab = k / ncd, cd = k % ncd;
                    const int a = ab / nb, b = ab % nb, c = cd / nd, d = cd % nd;
                    const int ax = comp_lx(la, a), ay = comp_ly(la, a), az = comp_lz(la, a);
                    const int bx = comp_lx(lb, b), by = comp_ly(lb, b), bz = comp_lz(lb, b);
                    const int cx = comp_lx(lc, c), cy = comp_ly(lc, c), cz = comp_lz(lc, c);
                    const int dx = comp_lx(ld, d), dy = comp_ly(ld, d), dz = comp_lz(ld, d);
                    const double* ebx = Es[wv][0] + (ax * (lb + 1) + bx) * ntb;
                    const double* eby = Es[wv][1] + (ay * (lb + 1) + by) * ntb;
                    const double* ebz = Es[wv][2] + (az * (lb + 1) + bz) * ntb;
                    const double* ekx = Es[wv][3] + (cx * (ld + 1) + dx) * ntk;
                    const double* eky = Es[wv][4] + (cy * (ld + 1) + dy) * ntk;
                    const double* ekz = Es[wv][5] + (cz * (ld + 1) + dz) * ntk;
                    double v = 0.0;
                    for (int t = 0; t <= ax + bx; ++t)
                        for (int u = 0; u <= ay + by; ++u)
                            for (int w = 0; w <= az + bz; ++w) {
                                double h = 0.0;
                                for (int tau = 0; tau <= cx + dx; ++tau)
                                    for (int nu = 0; nu <= cy + dy; ++nu) {
                                        const double e2 = ((tau + nu) & 1 ? -1.0 : 1.0) * ekx[tau] * eky[nu];
                                        for (int phi = 0; phi <= cz + dz; ++phi)
                                            h += (phi & 1 ? -e2 : e2) * ekz[phi] * R[hidx(t + tau, u + nu, w + phi)];
                                    }
                                v += ebx[t] * eby[u] * ebz[w] * h;
                            }
                    acc[i] += v;
                }
                wave_sync();
            }
        }
        const int bfa = bra.bfa, bfb = bra.bfb, bfc = ket.bfa, bfd = ket.bfb;
        const double deg = (bra.same ? 1.0 : 2.0) * (ket.same ? 1.0 : 2.0) * (qd.x == qd.y ? 1.0 : 2.0);
        for (int i = 0; i < GW_MAXC; ++i) {
            const int k = lane + 64 * i;
            if (k >= ntot) break;
            const int ab = k / ncd, cd = k % ncd;
            const int p = bfa + ab / nb, qq = bfb + ab % nb, r = bfc + cd / nd, s = bfd + cd % nd;
            const double* aD = A.absD;
            const double bound = A.sqS[(size_t)p * N + qq] * A.sqS[(size_t)r * N + s]
                * fmax(fmax(fmax(4.0 * aD[(size_t)p * N + qq], 4.0 * aD[(size_t)r * N + s]),
                            fmax(aD[(size_t)p * N + r], aD[(size_t)p * N + s])),
                       fmax(aD[(size_t)qq * N + r], aD[(size_t)qq * N + s]));
            if (bound < A.tol) continue;
            const double v = acc[i] * deg * A.gcomp[p] * A.gcomp[qq] * A.gcomp[r] * A.gcomp[s];
            const double* D = A.D;
            double* G = A.G;
            atomicAdd(&G[(size_t)p * N + qq], v * D[(size_t)r * N + s]);
            atomicAdd(&G[(size_t)r * N + s], v * D[(size_t)p * N + qq]);
            atomicAdd(&G[(size_t)qq * N + s], -0.25 * v * D[(size_t)p * N + r]);
            atomicAdd(&G[(size_t)qq * N + r], -0.25 * v * D[(size_t)p * N + s]);
            atomicAdd(&G[(size_t)p * N + r], -0.25 * v * D[(size_t)qq * N + s]);
            atomicAdd(&G[(size_t)p * N + s], -0.25 * v * D[(size_t)qq * N + r]);
        }
    }
}

// Compile-time variant for one class (used for (dd|dd)): lane m evaluates
// F_m, the R_tuv levels are computed in parallel across the wavefront
// through two LDS buffers, the E tables by six lanes; the ket contraction
// W_tuv(cd) is spread over the lanes into LDS, then every lane applies the
// bra expansion to its component quartets.  TENSOR: the integrals are
// stored at their eight index permutations in A.G (an N^4 tensor) instead
// of being digested into a Fock matrix.
template <int LA, int LB, int LC, int LD, bool TENSOR = false>
__global__ void __launch_bounds__(64 * GW_WAVES) fock_wave_class_kernel_t(ClassArgs A)
{
    constexpr int L = LA + LB + LC + LD;
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    constexpr int NAB = NA * NB, NCD = NC * ND, NTOT = NAB * NCD;
    constexpr int NCPL = (NTOT + 63) / 64;
    constexpr int HL = nherm(L);
    constexpr int NTB = LA + LB + 1, NTK = LC + LD + 1;
    constexpr int EB = (LA + 1) * (LB + 1) * NTB, EK = (LC + 1) * (LD + 1) * NTK;
    constexpr int HAB = nherm(LA + LB);
    __shared__ double Ws[GW_WAVES][HAB * NCD];
    __shared__ double Rb[GW_WAVES][2][HL];
    __shared__ double Es[GW_WAVES][3 * EB + 3 * EK];
    __shared__ double Fs[GW_WAVES][L + 1];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nq = *A.count;
    const int N = A.N;
    for (int64_t q = (int64_t)blockIdx.x * GW_WAVES + wv; q < nq; q += (int64_t)gridDim.x * GW_WAVES) {
        const int2 qd = A.quart[q];
        const ShellPair bra = A.pairs[qd.x];
        const ShellPair ket = A.pairs[qd.y];
        double acc[NCPL];
        for (int i = 0; i < NCPL; ++i) acc[i] = 0.0;
        for (int ip = 0; ip < bra.npp; ++ip) {
            const PrimPair* P = A.pp + bra.pp0 + ip;
            for (int jq = 0; jq < ket.npp; ++jq) {
                const PrimPair* Q = A.pp + ket.pp0 + jq;
                const double rs = rsqrt(P->p + Q->p);
                const double alpha = P->p * Q->p * rs * rs;
                const double X = P->Px - Q->Px, Y = P->Py - Q->Py, Z = P->Pz - Q->Pz;
                const double T = alpha * (X * X + Y * Y + Z * Z);
                if (lane <= L) {
                    double F[MAX_LTOT + 1];
                    boys_runtime(lane, T, A.boys, F);
                    double s = 34.98683665524972497 * P->K * P->invp * Q->K * Q->invp * rs;
                    for (int m = 0; m < lane; ++m) s *= -2.0 * alpha;
                    Fs[wv][lane] = s * F[lane];
                } else if (lane >= 32 && lane < 38) {
                    const int x = (lane - 32) % 3;
                    double E[EB > EK ? EB : EK];
                    if (lane < 35) {
                        hermite_E<LA, LB, 1>(P->inv2p, x == 0 ? P->PAx : (x == 1 ? P->PAy : P->PAz),
                                             x == 0 ? P->PBx : (x == 1 ? P->PBy : P->PBz), E);
                        for (int k = 0; k < EB; ++k) Es[wv][x * EB + k] = E[k];
                    } else {
                        hermite_E<LC, LD, -1>(Q->inv2p, x == 0 ? Q->PAx : (x == 1 ? Q->PAy : Q->PAz),
                                              x == 0 ? Q->PBx : (x == 1 ? Q->PBy : Q->PBz), E);
                        for (int k = 0; k < EK; ++k) Es[wv][3 * EB + x * EK + k] = E[k];
                    }
                }
                wave_sync();
                if (lane == 0) Rb[wv][L & 1][0] = Fs[wv][L];
                wave_sync();
                for (int n = L - 1; n >= 0; --n) {
                    const double* prev = Rb[wv][(n + 1) & 1];
                    double* cur = Rb[wv][n & 1];
                    for (int k = lane; k < nherm(L - n); k += 64) {
                        double val;
                        if (k == 0) {
                            val = Fs[wv][n];
                        } else {
                            const Herm3 h = hdecode(k);
                            if (h.t > 0) {
                                val = X * prev[hidx(h.t - 1, h.u, h.v)];
                                if (h.t > 1) val += (double)(h.t - 1) * prev[hidx(h.t - 2, h.u, h.v)];
                            } else if (h.u > 0) {
                                val = Y * prev[hidx(h.t, h.u - 1, h.v)];
                                if (h.u > 1) val += (double)(h.u - 1) * prev[hidx(h.t, h.u - 2, h.v)];
                            } else {
                                val = Z * prev[hidx(h.t, h.u, h.v - 1)];
                                if (h.v > 1) val += (double)(h.v - 1) * prev[hidx(h.t, h.u, h.v - 2)];
                            }
                        }
                        cur[k] = val;
                    }
                    wave_sync();
                }
                const double* R = Rb[wv][0];
                const double* Eb = Es[wv];
                const double* Ek = Es[wv] + 3 * EB;
                // ket contraction into the bra Hermite space: W[k][cd]
                double* W = Ws[wv];
                for (int e = lane; e < HAB * NCD; e += 64) {
                    const int k = e / NCD, cd = e % NCD;
                    const Herm3 h = hdecode(k);
                    const int c = cd / ND, d = cd % ND;
                    const int cx = comp_lx(LC, c), cy = comp_ly(LC, c), cz = comp_lz(LC, c);
                    const int dx = comp_lx(LD, d), dy = comp_ly(LD, d), dz = comp_lz(LD, d);
                    const double* ekx = Ek + (cx * (LD + 1) + dx) * NTK;
                    const double* eky = Ek + EK + (cy * (LD + 1) + dy) * NTK;
                    const double* ekz = Ek + 2 * EK + (cz * (LD + 1) + dz) * NTK;
                    double w = 0.0;
                    for (int tau = 0; tau <= cx + dx; ++tau)
                        for (int nu = 0; nu <= cy + dy; ++nu) {
                            const double e2 = ekx[tau] * eky[nu];
                            for (int phi = 0; phi <= cz + dz; ++phi)
                                w = fma(e2 * ekz[phi], R[hidx(h.t + tau, h.u + nu, h.v + phi)], w);
                        }
                    W[e] = w;
                }
                wave_sync();
                // bra expansion
#pragma unroll
                for (int i = 0; i < NCPL; ++i) {
                    const int k = lane + 64 * i;
                    if (k < NTOT) {
                        const int ab = k / NCD, cd = k % NCD;
                        const int a = ab / NB, b = ab % NB;
                        const int ax = comp_lx(LA, a), ay = comp_ly(LA, a), az = comp_lz(LA, a);
                        const int bx = comp_lx(LB, b), by = comp_ly(LB, b), bz = comp_lz(LB, b);
                        const double* ebx = Eb + (ax * (LB + 1) + bx) * NTB;
                        const double* eby = Eb + EB + (ay * (LB + 1) + by) * NTB;
                        const double* ebz = Eb + 2 * EB + (az * (LB + 1) + bz) * NTB;
                        double v = 0.0;
                        for (int t = 0; t <= ax + bx; ++t)
                            for (int u = 0; u <= ay + by; ++u) {
                                const double exy = ebx[t] * eby[u];
                                for (int w = 0; w <= az + bz; ++w)
                                    v = fma(exy * ebz[w], W[hidx(t, u, w) * NCD + cd], v);
                            }
                        acc[i] += v;
                    }
                }
                wave_sync();
            }
        }
        const int bfa = bra.bfa, bfb = bra.bfb, bfc = ket.bfa, bfd = ket.bfb;
        if constexpr (TENSOR) {
            const size_t n = (size_t)N;
#pragma unroll
            for (int i = 0; i < NCPL; ++i) {
                const int k = lane + 64 * i;
                if (k >= NTOT) continue;
                const int ab = k / NCD, cd = k % NCD;
                const size_t p = bfa + ab / NB, qq = bfb + ab % NB, r = bfc + cd / ND, s = bfd + cd % ND;
                const size_t pq = p * n + qq, qp = qq * n + p, rs = r * n + s, sr = s * n + r;
                // one producer per orbit of the eight permutations
                if ((bra.same && p < qq) || (ket.same && r < s) || (qd.x == qd.y && pq < rs)) continue;
                const double v = acc[i] * A.gcomp[p] * A.gcomp[qq] * A.gcomp[r] * A.gcomp[s];
                A.G[pq * n * n + rs] = v;
                A.G[qp * n * n + rs] = v;
                A.G[pq * n * n + sr] = v;
                A.G[qp * n * n + sr] = v;
                A.G[rs * n * n + pq] = v;
                A.G[sr * n * n + pq] = v;
                A.G[rs * n * n + qp] = v;
                A.G[sr * n * n + qp] = v;
            }
            continue;
        }
        const double deg = (bra.same ? 1.0 : 2.0) * (ket.same ? 1.0 : 2.0) * (qd.x == qd.y ? 1.0 : 2.0);
#pragma unroll
        for (int i = 0; i < NCPL; ++i) {
            const int k = lane + 64 * i;
            if (k >= NTOT) continue;
            const int ab = k / NCD, cd = k % NCD;
            const int p = bfa + ab / NB, qq = bfb + ab % NB, r = bfc + cd / ND, s = bfd + cd % ND;
            const double* aD = A.absD;
            const double bound = A.sqS[(size_t)p * N + qq] * A.sqS[(size_t)r * N + s]
                * fmax(fmax(fmax(4.0 * aD[(size_t)p * N + qq], 4.0 * aD[(size_t)r * N + s]),
                            fmax(aD[(size_t)p * N + r], aD[(size_t)p * N + s])),
                       fmax(aD[(size_t)qq * N + r], aD[(size_t)qq * N + s]));
            if (bound < A.tol) continue;
            const double v = acc[i] * deg * A.gcomp[p] * A.gcomp[qq] * A.gcomp[r] * A.gcomp[s];
            const double* D = A.D;
            double* G = A.G;
            atomicAdd(&G[(size_t)p * N + qq], v * D[(size_t)r * N + s]);
            atomicAdd(&G[(size_t)r * N + s], v * D[(size_t)p * N + qq]);
            atomicAdd(&G[(size_t)qq * N + s], -0.25 * v * D[(size_t)p * N + r]);
            atomicAdd(&G[(size_t)qq * N + r], -0.25 * v * D[(size_t)p * N + s]);
            atomicAdd(&G[(size_t)p * N + r], -0.25 * v * D[(size_t)qq * N + s]);
            atomicAdd(&G[(size_t)p * N + s], -0.25 * v * D[(size_t)qq * N + r]);
        }
    }
}

hipError_t launch_fock_wave_class(int la, int lb, int lc, int ld, int fp32, const ClassArgs& A, int64_t max_quartets,
                                  hipStream_t st)
{
    if (fock_class_is_mfma(la, lb, lc, ld)) return launch_fock_mfma_class(la, lb, lc, ld, fp32, A, max_quartets, st);
    int64_t grid = (max_quartets + GW_WAVES - 1) / GW_WAVES;
    if (grid < 1) grid = 1;
    if (grid > 4096) grid = 4096;
    if (la == 2 && lb == 2 && lc == 2 && ld == 2)
        hipLaunchKernelGGL((fock_wave_class_kernel_t<2, 2, 2, 2>), dim3((unsigned)grid), dim3(64 * GW_WAVES), 0, st, A);
    else
        hipLaunchKernelGGL(fock_wave_class_kernel, dim3((unsigned)grid), dim3(64 * GW_WAVES), 0, st, A, la, lb, lc, ld);
    return hipGetLastError();
}

// Stored-tensor mode of the (dd|dd) class: A.quart lists the canonical shell
// quartets (*A.count of them) and A.G is the N^4 output tensor.
hipError_t launch_tensor_wave_class(const ClassArgs& A, int64_t nquartets, hipStream_t st)
{
    if (nquartets <= 0) return hipSuccess;
    int64_t grid = (nquartets + GW_WAVES - 1) / GW_WAVES;
    if (grid > 4096) grid = 4096;
    hipLaunchKernelGGL((fock_wave_class_kernel_t<2, 2, 2, 2, true>), dim3((unsigned)grid), dim3(64 * GW_WAVES), 0, st,
                       A);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Small elementwise kernels
// ---------------------------------------------------------------------------
__global__ void abs_density_kernel(int n, const double* __restrict__ Dre, const double* __restrict__ Dim,
                                   double* __restrict__ absD)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    absD[i] = Dim ? hypot(Dre[i], Dim[i]) : fabs(Dre[i]);
}

hipError_t launch_abs_density(int n, const double* Dre, const double* Dim, double* absD, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(abs_density_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, Dre, Dim, absD);
    return hipGetLastError();
}

__global__ void shell_block_max_kernel(int nshell, const int* __restrict__ bf0, const int* __restrict__ nbf, int N,
                                       const double* __restrict__ absD, double* __restrict__ Dsh)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nshell * nshell) return;
    const int s1 = idx / nshell, s2 = idx % nshell;
    double m = 0.0;
    for (int i = 0; i < nbf[s1]; ++i)
        for (int j = 0; j < nbf[s2]; ++j) m = fmax(m, absD[(size_t)(bf0[s1] + i) * N + bf0[s2] + j]);
    Dsh[idx] = m;
}

hipError_t launch_shell_block_max(int nshell, const int* sh_bf0, const int* sh_nbf, int N, const double* absD,
                                  double* Dsh, hipStream_t st)
{
    const int n = nshell * nshell;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(shell_block_max_kernel, dim3((n + 255) / 256), dim3(256), 0, st, nshell, sh_bf0, sh_nbf, N,
                       absD, Dsh);
    return hipGetLastError();
}

__global__ void fill_kernel(size_t n, double* x, double v)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

hipError_t launch_fill(size_t n, double* x, double v, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, x, v);
    return hipGetLastError();
}

__global__ void fill_int_kernel(int n, int* x, int v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

hipError_t launch_fill_int(int n, int* x, int v, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_int_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, x, v);
    return hipGetLastError();
}

__global__ void symmetrize_kernel(int N, const double* __restrict__ G, double* __restrict__ out)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N * N) return;
    const int i = idx / N, j = idx % N;
    out[idx] = 0.5 * (G[(size_t)i * N + j] + G[(size_t)j * N + i]);
}

hipError_t launch_symmetrize(int N, const double* G, double* out, hipStream_t st)
{
    if (N <= 0) return hipSuccess;
    hipLaunchKernelGGL(symmetrize_kernel, dim3((N * N + 255) / 256), dim3(256), 0, st, N, G, out);
    return hipGetLastError();
}

// Block-diagonal batch: the per-build preparation restricted to each
// molecule's diagonal block (one workgroup per molecule): |dP| and a zero G
// in the block, then the shell-block maxima of |dP| over the molecule's
// shell pairs.  Entries outside the blocks are never read by the builds.
__global__ void __launch_bounds__(256) block_prep_kernel(const int4* __restrict__ blocks, int N,
                                                         const double* __restrict__ dP, double* __restrict__ absD,
                                                         double* __restrict__ G, const int* __restrict__ bf0,
                                                         const int* __restrict__ nbf, int nshell,
                                                         double* __restrict__ Dsh, double* __restrict__ mol_dmax)
{
    __shared__ double red[4];
    const int4 b = blocks[blockIdx.x];
    const int f0 = b.x, n = b.y, s0 = b.z, ns = b.w - b.z;
    double mx = 0.0;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const size_t o = (size_t)(f0 + e / n) * N + f0 + e % n;
        const double a = fabs(dP[o]);
        absD[o] = a;
        mx = fmax(mx, a);
        G[o] = 0.0;
    }
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) mol_dmax[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    for (int e = threadIdx.x; e < ns * ns; e += blockDim.x) {
        const int s1 = s0 + e / ns, s2 = s0 + e % ns;
        double m = 0.0;
        for (int i = 0; i < nbf[s1]; ++i)
            for (int j = 0; j < nbf[s2]; ++j) m = fmax(m, absD[(size_t)(bf0[s1] + i) * N + bf0[s2] + j]);
        Dsh[(size_t)s1 * nshell + s2] = m;
    }
}

hipError_t launch_block_prep(int nmol, const int4* blocks, int N, const double* dP, double* absD, double* G,
                             const int* sh_bf0, const int* sh_nbf, int nshell, double* Dsh, double* mol_dmax,
                             hipStream_t st)
{
    if (nmol <= 0) return hipSuccess;
    hipLaunchKernelGGL(block_prep_kernel, dim3(nmol), dim3(256), 0, st, blocks, N, dP, absD, G, sh_bf0, sh_nbf,
                       nshell, Dsh, mol_dmax);
    return hipGetLastError();
}

__global__ void __launch_bounds__(256) block_symmetrize_kernel(const int4* __restrict__ blocks, int N,
                                                               const double* __restrict__ G, double* __restrict__ out)
{
    const int4 b = blocks[blockIdx.x];
    const int f0 = b.x, n = b.y;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int i = f0 + e / n, j = f0 + e % n;
        out[(size_t)i * N + j] = 0.5 * (G[(size_t)i * N + j] + G[(size_t)j * N + i]);
    }
}

hipError_t launch_block_symmetrize(int nmol, const int4* blocks, int N, const double* G, double* out, hipStream_t st)
{
    if (nmol <= 0) return hipSuccess;
    hipLaunchKernelGGL(block_symmetrize_kernel, dim3(nmol), dim3(256), 0, st, blocks, N, G, out);
    return hipGetLastError();
}

__global__ void sqrt_kernel(size_t n, const double* __restrict__ x, double* __restrict__ y)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = sqrt(x[i]);
}

hipError_t launch_sqrt_clamped(size_t n, const double* x, double* y, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(sqrt_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, x, y);
    return hipGetLastError();
}

// Conventional J/K from the stored tensor (mmd/scf.py:97-98):
//   J_pq = sum_rs (pq|rs) P_sr ;  K_pq = sum_rs (ps|qr) P_sr.
// One wavefront per (p,q): the lanes run along the tensor's contiguous last
// index, so that both the row (pq|..) of J and the N segments (ps|q.) of K
// are read coalesced; the tensor is read twice in all (HBM-bound).
__global__ void __launch_bounds__(256) jk_tensor_kernel(int N, const double* __restrict__ T,
                                                        const double* __restrict__ Pre,
                                                        const double* __restrict__ Pim, double* Jre, double* Jim,
                                                        double* Kre, double* Kim)
{
    const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (w >= (int64_t)N * N) return;   // uniform across the wavefront
    const int p = (int)(w / N), q = (int)(w - (int64_t)p * N);
    const size_t n = (size_t)N, nn = n * n;
    double jr = 0, ji = 0, kr = 0, ki = 0;
    const double* row = T + (size_t)w * nn;                  // (pq|r s) at row[r*N + s]
    for (int r = 0; r < N; ++r)
        for (int s = lane; s < N; s += 64) {
            const double t = row[(size_t)r * n + s];
            jr = fma(t, Pre[(size_t)s * n + r], jr);
            if (Pim) ji = fma(t, Pim[(size_t)s * n + r], ji);
        }
    const double* kb = T + (size_t)p * n * nn + (size_t)q * n; // (p s|q r) at kb[s*N*N + r]
    for (int s2 = 0; s2 < N; ++s2)
        for (int r2 = lane; r2 < N; r2 += 64) {
            const double t = kb[(size_t)s2 * nn + r2];
            kr = fma(t, Pre[(size_t)s2 * n + r2], kr);
            if (Pim) ki = fma(t, Pim[(size_t)s2 * n + r2], ki);
        }
    for (int off = 32; off > 0; off >>= 1) {
        jr += __shfl_xor(jr, off);
        kr += __shfl_xor(kr, off);
        ji += __shfl_xor(ji, off);
        ki += __shfl_xor(ki, off);
    }
    if (lane != 0) return;
    Jre[w] = jr;
    Kre[w] = kr;
    if (Pim) {
        Jim[w] = ji;
        Kim[w] = ki;
    }
}

hipError_t launch_jk_tensor(int N, const double* T, const double* Pre, const double* Pim, double* Jre, double* Jim,
                            double* Kre, double* Kim, hipStream_t st)
{
    if (N <= 0) return hipSuccess;
    const int64_t waves = (int64_t)N * N;   // four wavefronts per 256-thread block
    hipLaunchKernelGGL(jk_tensor_kernel, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, st, N, T, Pre, Pim, Jre,
                       Jim, Kre, Kim);
    return hipGetLastError();
}

}  // namespace mmd

// ---------------------------------------------------------------------------
// Elementwise and reduction kernels of the GPU-resident SCF loop
// (mmd/scf.py:10-180 in the reference; the loop is driven from the host).
// ---------------------------------------------------------------------------
namespace mmd {

__global__ void scf_axpby_kernel(size_t n, double a, const double* __restrict__ x, double b,
                                 const double* __restrict__ y, double* __restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a * x[i] + b * y[i];
}

hipError_t launch_scf_axpby(size_t n, double a, const double* x, double b, const double* y, double* out,
                            hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(scf_axpby_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, a, x, b, y, out);
    return hipGetLastError();
}

// out = M - M^T (N x N)
__global__ void scf_antisym_kernel(int N, const double* __restrict__ M, double* __restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * N) return;
    const int r = (int)(i / N), c = (int)(i - (size_t)r * N);
    out[i] = M[i] - M[(size_t)c * N + r];
}

hipError_t launch_scf_antisym(int N, const double* M, double* out, hipStream_t st)
{
    const size_t n = (size_t)N * N;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(scf_antisym_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, N, M, out);
    return hipGetLastError();
}

// One workgroup: out[0] = sum_pq (H + F)_pq P_qp, out[1] = sum (P - P_old)^2
// (deterministic order).
__global__ void __launch_bounds__(1024) scf_energy_kernel(int N, const double* __restrict__ H,
                                                          const double* __restrict__ F,
                                                          const double* __restrict__ P,
                                                          const double* __restrict__ Pold, double* out)
{
    __shared__ double re[1024], rr[1024];
    double e = 0.0, r = 0.0;
    const size_t n = (size_t)N * N;
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) {
        const int p = (int)(i / N), q = (int)(i - (size_t)p * N);
        e = fma(H[i] + F[i], P[(size_t)q * N + p], e);
        const double d = P[i] - Pold[i];
        r = fma(d, d, r);
    }
    re[threadIdx.x] = e;
    rr[threadIdx.x] = r;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            re[threadIdx.x] += re[threadIdx.x + s];
            rr[threadIdx.x] += rr[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = re[0];
        out[1] = rr[0];
    }
}

hipError_t launch_scf_energy(int N, const double* H, const double* F, const double* P, const double* Pold,
                             double* out, hipStream_t st)
{
    hipLaunchKernelGGL(scf_energy_kernel, dim3(1), dim3(1024), 0, st, N, H, F, P, Pold, out);
    return hipGetLastError();
}

struct ScfVecs {
    const double* v[8];
};

// out[k] = sum_i E[k][i] e[i], k < nset (one workgroup per k; deterministic)
__global__ void __launch_bounds__(1024) scf_dots_kernel(size_t n, ScfVecs E, const double* __restrict__ e,
                                                        double* out)
{
    __shared__ double red[1024];
    const double* a = E.v[blockIdx.x];
    double s = 0.0;
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) s = fma(a[i], e[i], s);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

hipError_t launch_scf_dots(int nset, size_t n, const double* const* E, const double* e, double* out, hipStream_t st)
{
    if (nset <= 0) return hipSuccess;
    if (nset > 8) return hipErrorInvalidValue;
    ScfVecs V;
    for (int k = 0; k < nset; ++k) V.v[k] = E[k];
    hipLaunchKernelGGL(scf_dots_kernel, dim3(nset), dim3(1024), 0, st, n, V, e, out);
    return hipGetLastError();
}

// out = sum_k w[k] F[k], k < nset <= 8
struct ScfWeights {
    const double* F[8];
    double w[8];
    int n;
};

__global__ void scf_combine_kernel(size_t n, ScfWeights W, double* __restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int k = 0; k < W.n; ++k) s = fma(W.w[k], W.F[k][i], s);
    out[i] = s;
}

hipError_t launch_scf_combine(size_t n, int nset, const double* const* F, const double* w, double* out,
                              hipStream_t st)
{
    if (nset < 1 || nset > 8) return hipErrorInvalidValue;
    ScfWeights W;
    for (int k = 0; k < nset; ++k) {
        W.F[k] = F[k];
        W.w[k] = w[k];
    }
    W.n = nset;
    hipLaunchKernelGGL(scf_combine_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, W, out);
    return hipGetLastError();
}

}  // namespace mmd
