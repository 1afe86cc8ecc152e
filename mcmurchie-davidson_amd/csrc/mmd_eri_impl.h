#pragma once
// Shell-quartet Fock-build kernel for gfx950, one instantiation per
// angular-momentum class (la lb | lc ld), la >= lb, lc >= ld,
// pair_class(bra) >= pair_class(ket).
//
// Work decomposition: one lane per (contracted shell quartet, chunk of ket
// Cartesian components); the chunk index is uniform across a wavefront, so
// every lane of a wave runs the same compile-time specialised code.
//
// Per bra primitive pair the lane loops over the ket primitive pairs: Boys
// function from the LDS-staged Taylor table, R_tuv recursion in registers,
// and the ket contraction accumulated in the bra's Hermite space,
//   W_tuv(cd) += sum_{tau nu phi} (-1)^{tau+nu+phi} E^{cd}_{tau nu phi} R_{t+tau,u+nu,v+phi}.
// The bra expansion is applied once per bra primitive pair,
//   (ab|cd) += sum_tuv E^{ab}_tuv W_tuv(cd),
// so the dense bra contraction is not repeated per primitive quartet.  All
// loops are unrolled at compile time (sfor) so that every array lives in
// registers.  The epilogue digests the block into the Fock matrix: Coulomb
// and exchange with the incremental density after the Schwarz x |dP| test.
//
// Reference semantics: cython/twoe.pyx:52-94 (primitive ERI),
// cython/twoe.pyx:34-50 (contraction), cython/fock.pyx:38-85 (screening,
// degeneracy, Coulomb/exchange scatter).
//
// This header is instantiated once per class by mmd_eri_class.hip.
#include <hip/hip_runtime.h>
#include <type_traits>
#include <utility>
#include "mmd_device.h"
#include "mmd_types.h"
#include "mmd_kernels.h"

namespace mmd {

constexpr double TWO_PI_2_5 = 34.98683665524972497;   // 2 pi^{5/2}

// Largest divisor of ncd whose chunk keeps the per-lane accumulators
// (bra Hermite intermediate + contracted block) within `limit` doubles.
__host__ __device__ constexpr int chunk_size(int nab, int hab, int ncd, int limit)
{
    int best = 1;
    for (int c = 1; c <= ncd; ++c)
        if (ncd % c == 0 && (nab + hab) * c <= limit) best = c;
    return best;
}
constexpr int ACC_LIMIT = 40;

template <int LA, int LB, int LC, int LD>
struct ClassShape {
    static constexpr int L = LA + LB + LC + LD;
    static constexpr int LAB = LA + LB;
    static constexpr int HAB = nherm(LAB);
    static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    static constexpr int NAB = NA * NB, NCD = NC * ND;
    static constexpr int CH = chunk_size(NAB, HAB, NCD, ACC_LIMIT);
    static constexpr int NCHUNK = NCD / CH;
    static constexpr int NTB = LA + LB + 1, NTK = LC + LD + 1;
};

__device__ __forceinline__ double dmax6(double a, double b, double c, double d, double e, double f)
{
    return fmax(fmax(fmax(a, b), fmax(c, d)), fmax(e, f));
}

// ---------------------------------------------------------------------------
// One (quartet, ket-chunk) work item.
// ---------------------------------------------------------------------------
// rowA / rowB: rows of G for the bra shells' functions (row x of shell a at
// rowA + x*N), either in LDS (accumulated per bra pair, flushed later) or
// the rows of G itself.  The ket-side Coulomb block always goes to G.
template <int LA, int LB, int LC, int LD, int C0, int C1>
__device__ __forceinline__ void quartet_item(const ClassArgs& A, const double* lds, int2 qd, const ShellPair& bra,
                                             const ShellPair& ket, double* rowA, double* rowB)
{
    using S = ClassShape<LA, LB, LC, LD>;
    constexpr int NCH = C1 - C0;
    constexpr int L = S::L;
    constexpr int HAB = S::HAB;

    double acc[S::NAB * NCH];
    sfor<S::NAB * NCH>([&](auto k) { acc[k] = 0.0; });

    for (int ip = 0; ip < bra.npp; ++ip) {
        const PrimPair* P = A.pp + bra.pp0 + ip;
        const double Pp = P->p, PK = TWO_PI_2_5 * P->K * P->invp, Px = P->Px, Py = P->Py, Pz = P->Pz;
        double W[HAB * NCH];
        sfor<HAB * NCH>([&](auto k) { W[k] = 0.0; });
        for (int jq = 0; jq < ket.npp; ++jq) {
            const PrimPair* Q = A.pp + ket.pp0 + jq;
            double Ekx[(LC + 1) * (LD + 1) * S::NTK], Eky[(LC + 1) * (LD + 1) * S::NTK],
                Ekz[(LC + 1) * (LD + 1) * S::NTK];
            if constexpr (LC + LD > 0) {
                const double qinv2 = Q->inv2p;
                hermite_E<LC, LD, -1>(qinv2, Q->PAx, Q->PBx, Ekx);
                hermite_E<LC, LD, -1>(qinv2, Q->PAy, Q->PBy, Eky);
                hermite_E<LC, LD, -1>(qinv2, Q->PAz, Q->PBz, Ekz);
            } else {
                Ekx[0] = Eky[0] = Ekz[0] = 1.0;
            }
            const double Qp = Q->p;
            const double rs = rsqrt(Pp + Qp);
            const double alpha = Pp * Qp * rs * rs;
            const double X = Px - Q->Px, Y = Py - Q->Py, Z = Pz - Q->Pz;
            const double T = alpha * (X * X + Y * Y + Z * Z);
            // 2 pi^{5/2} / (p q sqrt(p+q)) K_ab K_cd
            const double pref = PK * (Q->K * Q->invp) * rs;
            double F[L + 1];
            boys_seeds<L>(T, alpha, pref, lds, F);
            double R[nherm(L)];
            hermite_R<L>(X, Y, Z, F, R);
            // ket contraction into the bra Hermite space
            sfor<NCH>([&](auto ic) {
                constexpr int cd = C0 + decltype(ic)::value;
                constexpr int c = cd / S::ND, d = cd % S::ND;
                constexpr int cx = comp_lx(LC, c), cy = comp_ly(LC, c), cz = comp_lz(LC, c);
                constexpr int dx = comp_lx(LD, d), dy = comp_ly(LD, d), dz = comp_lz(LD, d);
                constexpr int kx = (cx * (LD + 1) + dx) * S::NTK;
                constexpr int ky = (cy * (LD + 1) + dy) * S::NTK;
                constexpr int kz = (cz * (LD + 1) + dz) * S::NTK;
                sfor<HAB>([&](auto ik) {
                    constexpr Herm3 h = hdecode(decltype(ik)::value);
                    double w = W[decltype(ik)::value * NCH + decltype(ic)::value];
                    sfor<cx + dx + 1>([&](auto itau) {
                        constexpr int tau = decltype(itau)::value;
                        sfor<cy + dy + 1>([&](auto inu) {
                            constexpr int nu = decltype(inu)::value;
                            const double exy = Ekx[kx + tau] * Eky[ky + nu];
                            sfor<cz + dz + 1>([&](auto iphi) {
                                constexpr int phi = decltype(iphi)::value;
                                w = fma(exy * Ekz[kz + phi], R[hidx(h.t + tau, h.u + nu, h.v + phi)], w);
                            });
                        });
                    });
                    W[decltype(ik)::value * NCH + decltype(ic)::value] = w;
                });
            });
        }
        // bra expansion, once per bra primitive pair
        double Ebx[(LA + 1) * (LB + 1) * S::NTB], Eby[(LA + 1) * (LB + 1) * S::NTB],
            Ebz[(LA + 1) * (LB + 1) * S::NTB];
        if constexpr (LA + LB > 0) {
            const double pinv2 = P->inv2p;
            hermite_E<LA, LB, 1>(pinv2, P->PAx, P->PBx, Ebx);
            hermite_E<LA, LB, 1>(pinv2, P->PAy, P->PBy, Eby);
            hermite_E<LA, LB, 1>(pinv2, P->PAz, P->PBz, Ebz);
        } else {
            Ebx[0] = Eby[0] = Ebz[0] = 1.0;
        }
        sfor<S::NAB>([&](auto iab) {
            constexpr int ab = decltype(iab)::value;
            constexpr int a = ab / S::NB, b = ab % S::NB;
            constexpr int ax = comp_lx(LA, a), ay = comp_ly(LA, a), az = comp_lz(LA, a);
            constexpr int bx = comp_lx(LB, b), by = comp_ly(LB, b), bz = comp_lz(LB, b);
            constexpr int ix = (ax * (LB + 1) + bx) * S::NTB;
            constexpr int iy = (ay * (LB + 1) + by) * S::NTB;
            constexpr int iz = (az * (LB + 1) + bz) * S::NTB;
            sfor<NCH>([&](auto ic) {
                double s = acc[ab * NCH + decltype(ic)::value];
                sfor<ax + bx + 1>([&](auto it) {
                    constexpr int t = decltype(it)::value;
                    sfor<ay + by + 1>([&](auto iu) {
                        constexpr int u = decltype(iu)::value;
                        const double exy = Ebx[ix + t] * Eby[iy + u];
                        sfor<az + bz + 1>([&](auto iv) {
                            constexpr int v = decltype(iv)::value;
                            s = fma(exy * Ebz[iz + v], W[hidx(t, u, v) * NCH + decltype(ic)::value], s);
                        });
                    });
                });
                acc[ab * NCH + decltype(ic)::value] = s;
            });
        });
    }

    // -----------------------------------------------------------------------
    // Epilogue: component scales, shell-level degeneracy, basis-function-level
    // Schwarz x |dP| test (cython/fock.pyx:46-57), then Coulomb and exchange.
    // Every component combination of the unique shell quartet is digested
    // with the shell degeneracy; the symmetrised result equals formPT's
    // symmetrised result for a real symmetric density difference.
    // -----------------------------------------------------------------------
    const int N = A.N;
    const int bfa = bra.bfa, bfb = bra.bfb, bfc = ket.bfa, bfd = ket.bfb;
    double ga[S::NA], gb[S::NB], gc[S::NC], gd[S::ND];
    sfor<S::NA>([&](auto i) { ga[i] = A.gcomp[bfa + i]; });
    sfor<S::NB>([&](auto i) { gb[i] = A.gcomp[bfb + i]; });
    sfor<S::NC>([&](auto i) { gc[i] = A.gcomp[bfc + i]; });
    sfor<S::ND>([&](auto i) { gd[i] = A.gcomp[bfd + i]; });
    const double deg = (bra.same ? 1.0 : 2.0) * (ket.same ? 1.0 : 2.0) * (qd.x == qd.y ? 1.0 : 2.0);
    const double* aD = A.absD;
    const double* sq = A.sqS;
    const double tol = A.tol;
    sfor<S::NAB>([&](auto iab) {
        constexpr int ab = decltype(iab)::value;
        constexpr int a = ab / S::NB, b = ab % S::NB;
        const int p = bfa + a, q = bfb + b;
        const double sab = sq[(size_t)p * N + q];
        const double dab = 4.0 * aD[(size_t)p * N + q];
        sfor<NCH>([&](auto ic) {
            constexpr int cd = C0 + decltype(ic)::value;
            constexpr int c = cd / S::ND, d = cd % S::ND;
            const int r = bfc + c, s = bfd + d;
            const double bound = sab * sq[(size_t)r * N + s]
                * dmax6(dab, 4.0 * aD[(size_t)r * N + s], aD[(size_t)p * N + r], aD[(size_t)p * N + s],
                        aD[(size_t)q * N + r], aD[(size_t)q * N + s]);
            const double v = acc[ab * NCH + decltype(ic)::value] * deg * ga[a] * gb[b] * gc[c] * gd[d];
            acc[ab * NCH + decltype(ic)::value] = (bound < tol) ? 0.0 : v;
        });
    });
    const double* D = A.D;
    double* G = A.G;
    // Coulomb: G_ab += sum_cd (ab|cd) D_cd
    sfor<S::NAB>([&](auto iab) {
        constexpr int ab = decltype(iab)::value;
        constexpr int a = ab / S::NB, b = ab % S::NB;
        double s = 0.0;
        sfor<NCH>([&](auto ic) {
            constexpr int cd = C0 + decltype(ic)::value;
            constexpr int c = cd / S::ND, d = cd % S::ND;
            s = fma(acc[ab * NCH + decltype(ic)::value], D[(size_t)(bfc + c) * N + bfd + d], s);
        });
        atomicAdd(&rowA[(size_t)a * N + bfb + b], s);
    });
    // Coulomb: G_cd += sum_ab (ab|cd) D_ab
    sfor<NCH>([&](auto ic) {
        constexpr int cd = C0 + decltype(ic)::value;
        constexpr int c = cd / S::ND, d = cd % S::ND;
        double s = 0.0;
        sfor<S::NAB>([&](auto iab) {
            constexpr int ab = decltype(iab)::value;
            constexpr int a = ab / S::NB, b = ab % S::NB;
            s = fma(acc[ab * NCH + decltype(ic)::value], D[(size_t)(bfa + a) * N + bfb + b], s);
        });
        atomicAdd(&G[(size_t)(bfc + c) * N + bfd + d], s);
    });
    // Exchange (cython/fock.pyx:82-85): G_bd -= 1/4 (ab|cd) D_ac ;
    // G_bc -= 1/4 (ab|cd) D_ad ; G_ac -= 1/4 (ab|cd) D_bd ; G_ad -= 1/4 (ab|cd) D_bc.
    // Each target element sums the chunk's contributions in registers first.
    sfor<S::NB>([&](auto ix) {
        constexpr int x = decltype(ix)::value;
        sfor<S::ND>([&](auto iy) {        // (b,d) block
            constexpr int y = decltype(iy)::value;
            double s = 0.0;
            bool any = false;
            sfor<S::NA>([&](auto ia) {
                constexpr int a = decltype(ia)::value;
                sfor<S::NC>([&](auto icc) {
                    constexpr int c = decltype(icc)::value;
                    constexpr int cd = c * S::ND + y;
                    if constexpr (cd >= C0 && cd < C1) {
                        any = true;
                        s = fma(acc[(a * S::NB + x) * NCH + (cd - C0)], D[(size_t)(bfa + a) * N + bfc + c], s);
                    }
                });
            });
            if (any) atomicAdd(&rowB[(size_t)x * N + bfd + y], -0.25 * s);
        });
        sfor<S::NC>([&](auto iy) {        // (b,c) block
            constexpr int y = decltype(iy)::value;
            double s = 0.0;
            bool any = false;
            sfor<S::NA>([&](auto ia) {
                constexpr int a = decltype(ia)::value;
                sfor<S::ND>([&](auto idd) {
                    constexpr int d = decltype(idd)::value;
                    constexpr int cd = y * S::ND + d;
                    if constexpr (cd >= C0 && cd < C1) {
                        any = true;
                        s = fma(acc[(a * S::NB + x) * NCH + (cd - C0)], D[(size_t)(bfa + a) * N + bfd + d], s);
                    }
                });
            });
            if (any) atomicAdd(&rowB[(size_t)x * N + bfc + y], -0.25 * s);
        });
    });
    sfor<S::NA>([&](auto ix) {
        constexpr int x = decltype(ix)::value;
        sfor<S::NC>([&](auto iy) {        // (a,c) block
            constexpr int y = decltype(iy)::value;
            double s = 0.0;
            bool any = false;
            sfor<S::NB>([&](auto ib) {
                constexpr int b = decltype(ib)::value;
                sfor<S::ND>([&](auto idd) {
                    constexpr int d = decltype(idd)::value;
                    constexpr int cd = y * S::ND + d;
                    if constexpr (cd >= C0 && cd < C1) {
                        any = true;
                        s = fma(acc[(x * S::NB + b) * NCH + (cd - C0)], D[(size_t)(bfb + b) * N + bfd + d], s);
                    }
                });
            });
            if (any) atomicAdd(&rowA[(size_t)x * N + bfc + y], -0.25 * s);
        });
        sfor<S::ND>([&](auto iy) {        // (a,d) block
            constexpr int y = decltype(iy)::value;
            double s = 0.0;
            bool any = false;
            sfor<S::NB>([&](auto ib) {
                constexpr int b = decltype(ib)::value;
                sfor<S::NC>([&](auto icc) {
                    constexpr int c = decltype(icc)::value;
                    constexpr int cd = c * S::ND + y;
                    if constexpr (cd >= C0 && cd < C1) {
                        any = true;
                        s = fma(acc[(x * S::NB + b) * NCH + (cd - C0)], D[(size_t)(bfb + b) * N + bfc + c], s);
                    }
                });
            });
            if (any) atomicAdd(&rowA[(size_t)x * N + bfd + y], -0.25 * s);
        });
    });
}

template <int I, int N, class F>
__device__ __forceinline__ void static_switch(int i, F&& f)
{
    if constexpr (I < N) {
        if (i == I) f(std::integral_constant<int, I>{});
        else static_switch<I + 1, N>(i, f);
    }
}

// ---------------------------------------------------------------------------
// Class kernel.  A work unit is one bra shell pair ("row") and a segment of
// its ket pairs; workgroups take units from a device counter.  The unit's
// ket pairs are screened; survivors are compacted in LDS and processed as
// (64-quartet block, ket chunk) wave items.  Coulomb
// G_ab and the four exchange blocks, all in the rows of the bra shells, are
// accumulated in LDS (ds_add_f64) and their nonzero entries added to G once
// per unit; only the
// ket-side Coulomb block is added to G directly.  (The host sends a class
// whose bra rows do not fit in LDS to the wavefront-per-quartet kernel.)
// ---------------------------------------------------------------------------
constexpr int ROW_SEG = 1024;   // maximum ket pairs of one work unit

template <int LA, int LB, int LC, int LD>
__global__ void __launch_bounds__(256) eri_fock_class_kernel(ClassArgs A, RowArgs R)
{
    using S = ClassShape<LA, LB, LC, LD>;
    extern __shared__ double dyn[];
    double* lds_boys = dyn;
    double* rows = dyn + BOYS_TAYLOR * BOYS_NT;
    int* qlist = reinterpret_cast<int*>(rows + (S::NA + S::NB) * A.N);
    int* misc = qlist + ROW_SEG;
    boys_stage_lds<S::L>(lds_boys, A.boys);
    const int N = A.N;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int upr = (R.nk + R.seg - 1) / R.seg;          // units per row
    const int nunits = R.nb * upr;
    const size_t ns = (size_t)R.nshell;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) misc[0] = atomicAdd(R.row_counter, 1);
        __syncthreads();
        const int u = misc[0];
        if (u >= nunits) break;
        // triangular classes hand out their long rows first
        const int P = R.tri ? (R.nb - 1 - u / upr) : u / upr;
        const int k0 = (u % upr) * R.seg;
        const int nk = R.tri ? P + 1 : R.nk;
        if (k0 >= nk) continue;
        const int segn = min(R.seg, nk - k0);
        const ShellPair bra = A.pairs[R.boff + P];
        const int nrows = bra.same ? S::NA : S::NA + S::NB;
        for (int i = threadIdx.x; i < nrows * N; i += blockDim.x) rows[i] = 0.0;
        if (threadIdx.x == 0) misc[1] = 0;
        __syncthreads();
        for (int t0 = 0; t0 < segn; t0 += blockDim.x) {
            const int t = t0 + threadIdx.x;
            bool keep = false;
            if (t < segn) {
                const ShellPair ket = A.pairs[R.koff + k0 + t];
                const double* Dsh = R.Dsh;
                const double dm = fmax(fmax(fmax(4.0 * Dsh[bra.sa * ns + bra.sb], 4.0 * Dsh[ket.sa * ns + ket.sb]),
                                            fmax(Dsh[bra.sa * ns + ket.sa], Dsh[bra.sa * ns + ket.sb])),
                                       fmax(Dsh[bra.sb * ns + ket.sa], Dsh[bra.sb * ns + ket.sb]));
                keep = bra.q * ket.q * dm >= A.tol;
            }
            const unsigned long long m = __ballot(keep);
            int pos = 0;
            if (lane == 0 && m) pos = atomicAdd(&misc[1], __popcll(m));
            pos = __shfl(pos, 0);
            if (keep) qlist[pos + __popcll(m & ((1ull << lane) - 1ull))] = k0 + t;
        }
        __syncthreads();
        const int nsurv = misc[1];
        if (nsurv == 0) continue;
        if (R.stat_nquart && threadIdx.x == 0) atomicAdd(R.stat_nquart, (unsigned long long)nsurv);
        const int total = ((nsurv + 63) / 64) * S::NCHUNK;
        for (int w = wv; w < total; w += (int)(blockDim.x >> 6)) {
            const int chunk = w % S::NCHUNK;
            const int idx = (w / S::NCHUNK) * 64 + lane;
            if (idx >= nsurv) continue;
            const int Q = qlist[idx];
            const int2 qd = make_int2(R.boff + P, R.koff + Q);
            const ShellPair ket = A.pairs[qd.y];
            if (R.stat_nprim && chunk == 0)
                atomicAdd(R.stat_nprim, (unsigned long long)bra.npp * (unsigned long long)ket.npp);
            static_switch<0, S::NCHUNK>(chunk, [&](auto ic) {
                constexpr int C = decltype(ic)::value;
                quartet_item<LA, LB, LC, LD, C * S::CH, (C + 1) * S::CH>(
                    A, lds_boys, qd, bra, ket, rows, rows + (bra.same ? 0 : S::NA * N));
            });
        }
        __syncthreads();
        for (int i = threadIdx.x; i < nrows * N; i += blockDim.x) {
            const double v = rows[i];
            const int x = i / N, col = i - x * N;
            const int g = x < S::NA ? bra.bfa + x : bra.bfb + (x - S::NA);
            if (v != 0.0) atomicAdd(&A.G[(size_t)g * N + col], v);
        }
    }
}

template <int LA, int LB, int LC, int LD>
inline size_t fock_class_lds_bytes(int N)
{
    using S = ClassShape<LA, LB, LC, LD>;
    return sizeof(double) * (BOYS_TAYLOR * BOYS_NT + (size_t)(S::NA + S::NB) * N) + sizeof(int) * (ROW_SEG + 4);
}

}  // namespace mmd
