#pragma once
// Shell-quartet Fock-build kernel for gfx950, one instantiation per
// angular-momentum class (la lb | lc ld), la >= lb, lc >= ld,
// pair_class(bra) >= pair_class(ket).
//
// Work decomposition: one lane per (contracted shell quartet, chunk of bra
// Cartesian components); the chunk index is uniform across a wavefront, so
// every lane of a wave runs the same compile-time specialised code.
//
// All lanes of a work unit share the bra shell pair, whose primitive pairs
// (with their Hermite tables) are staged in LDS; per ket primitive pair the
// lane loops over them: Boys function from the LDS-staged Taylor table, R_tuv
// recursion in registers, and the bra contraction accumulated in the ket's
// Hermite space,
//   V_{tau nu phi}(ab) += sum_tuv E^{ab}_tuv R_{t+tau,u+nu,v+phi}.
// The ket expansion is applied once per ket primitive pair,
//   (ab|cd) += sum (-1)^{tau+nu+phi} E^{cd}_{tau nu phi} V_{tau nu phi}(ab),
// so the dense ket contraction is not repeated per primitive quartet.  All
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

// Largest divisor of nab whose chunk of bra components keeps the per-lane
// accumulators (ket Hermite intermediate + contracted block) within `limit`
// doubles.
__host__ __device__ constexpr int chunk_size(int nab, int hcd, int ncd, int limit)
{
    int best = 1;
    for (int c = 1; c <= nab; ++c)
        if (nab % c == 0 && (hcd + ncd) * c <= limit) best = c;
    return best;
}
constexpr int ACC_LIMIT = 40;

template <int LA, int LB, int LC, int LD>
struct ClassShape {
    static constexpr int L = LA + LB + LC + LD;
    static constexpr int LAB = LA + LB, LCD = LC + LD;
    static constexpr int HCD = nherm(LCD);
    static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    static constexpr int NAB = NA * NB, NCD = NC * ND;
    static constexpr int CH = chunk_size(NAB, HCD, NCD, ACC_LIMIT);   // bra components per item
    static constexpr int NCHUNK = NAB / CH;
    static constexpr int NTB = LA + LB + 1, NTK = LC + LD + 1;
    static constexpr int EB = (LA + 1) * (LB + 1) * NTB;   // 1-D bra E table
    static constexpr int BPS = 8 + 3 * EB;                 // staged bra primitive stride (doubles)
};

__device__ __forceinline__ double dmax6(double a, double b, double c, double d, double e, double f)
{
    return fmax(fmax(fmax(a, b), fmax(c, d)), fmax(e, f));
}

// Stage the bra pair's primitive pairs in LDS: p, 2 pi^{5/2} K / p, P and the
// three 1-D Hermite tables E^{ab}_t (cython/util.pxi:13-28).
template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void stage_bra(const ClassArgs& A, const ShellPair& bra, double* bp)
{
    using S = ClassShape<LA, LB, LC, LD>;
    for (int idx = threadIdx.x; idx < 3 * bra.npp; idx += blockDim.x) {
        const int ip = idx / 3, x = idx - 3 * ip;
        const PrimPair* P = A.pp + bra.pp0 + ip;
        double* b = bp + ip * S::BPS;
        if (x == 0) {
            b[0] = P->p;
            b[1] = TWO_PI_2_5 * P->K * P->invp;
            b[2] = P->Px;
            b[3] = P->Py;
            b[4] = P->Pz;
        }
        double E[S::EB];
        hermite_E<LA, LB, 1>(P->inv2p, x == 0 ? P->PAx : (x == 1 ? P->PAy : P->PAz),
                             x == 0 ? P->PBx : (x == 1 ? P->PBy : P->PBz), E);
        sfor<S::EB>([&](auto k) { b[8 + x * S::EB + k] = E[k]; });
    }
}

// ---------------------------------------------------------------------------
// One work item: a quartet, a chunk of its bra components, and a slice of its
// bra and ket primitive pairs (the contributions of the slices add up, and
// so do their digestions).  Outer loop: the lane's ket primitive pairs
// (global memory).  Inner loop: the bra primitive pairs staged in LDS (the
// same for every lane of the work unit).  Per inner step: Boys,
// R_tuv, and the bra contraction accumulated in the ket's Hermite space,
//   V_{tau nu phi}(ab) += sum_tuv E^{ab}_tuv R_{t+tau,u+nu,v+phi};
// per ket primitive: (ab|cd) += sum (-1)^{tau+nu+phi} E^{cd}_{tau nu phi} V(ab).
// rowA / rowB: rows of G for the bra shells' functions (row x of shell a at
// rowA + x*N) accumulated in LDS; the ket-side Coulomb block goes to G.
// ---------------------------------------------------------------------------
template <int LA, int LB, int LC, int LD, int A0, int A1>
__device__ __forceinline__ void quartet_item(const ClassArgs& A, const double* lds, const double* bp, int2 qd,
                                             const ShellPair& bra, const ShellPair& ket, int ip0, int ip1, int jq0,
                                             int jq1, double* rowA, double* rowB)
{
    using S = ClassShape<LA, LB, LC, LD>;
    constexpr int NCH = A1 - A0;
    constexpr int NCD = S::NCD;
    constexpr int L = S::L;
    constexpr int HCD = S::HCD;
    auto AI = [](int ab, int cd) { return (ab - A0) * NCD + cd; };

    double acc[NCH * NCD];
    sfor<NCH * NCD>([&](auto k) { acc[k] = 0.0; });

    for (int jq = jq0; jq < jq1; ++jq) {
        const PrimPair* Q = A.pp + ket.pp0 + jq;
        const double Qp = Q->p, QK = Q->K * Q->invp, Qx = Q->Px, Qy = Q->Py, Qz = Q->Pz;
        double V[HCD * NCH];
        sfor<HCD * NCH>([&](auto k) { V[k] = 0.0; });
        for (int ip = ip0; ip < ip1; ++ip) {
            const double* b = bp + ip * S::BPS;
            const double Pp = b[0];
            const double rs = rsqrt(Pp + Qp);
            const double alpha = Pp * Qp * rs * rs;
            const double X = b[2] - Qx, Y = b[3] - Qy, Z = b[4] - Qz;
            const double T = alpha * (X * X + Y * Y + Z * Z);
            const double pref = b[1] * QK * rs;   // 2 pi^{5/2} / (p q sqrt(p+q)) K_ab K_cd
            double F[L + 1];
            boys_seeds<L>(T, alpha, pref, lds, F);
            double R[nherm(L)];
            hermite_R<L>(X, Y, Z, F, R);
            const double* Ebx = b + 8;
            const double* Eby = b + 8 + S::EB;
            const double* Ebz = b + 8 + 2 * S::EB;
            sfor<NCH>([&](auto ii) {
                constexpr int ab = A0 + decltype(ii)::value;
                constexpr int a = ab / S::NB, bb = ab % S::NB;
                constexpr int ax = comp_lx(LA, a), ay = comp_ly(LA, a), az = comp_lz(LA, a);
                constexpr int bx = comp_lx(LB, bb), by = comp_ly(LB, bb), bz = comp_lz(LB, bb);
                constexpr int ix = (ax * (LB + 1) + bx) * S::NTB;
                constexpr int iy = (ay * (LB + 1) + by) * S::NTB;
                constexpr int iz = (az * (LB + 1) + bz) * S::NTB;
                sfor<HCD>([&](auto ik) {
                    constexpr Herm3 h = hdecode(decltype(ik)::value);
                    double v = V[decltype(ik)::value * NCH + decltype(ii)::value];
                    sfor<ax + bx + 1>([&](auto it) {
                        constexpr int t = decltype(it)::value;
                        sfor<ay + by + 1>([&](auto iu) {
                            constexpr int u = decltype(iu)::value;
                            const double exy = Ebx[ix + t] * Eby[iy + u];
                            sfor<az + bz + 1>([&](auto iw) {
                                constexpr int w = decltype(iw)::value;
                                v = fma(exy * Ebz[iz + w], R[hidx(t + h.t, u + h.u, w + h.v)], v);
                            });
                        });
                    });
                    V[decltype(ik)::value * NCH + decltype(ii)::value] = v;
                });
            });
        }
        // ket expansion, once per ket primitive pair
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
        sfor<NCD>([&](auto icd) {
            constexpr int cd = decltype(icd)::value;
            constexpr int c = cd / S::ND, d = cd % S::ND;
            constexpr int cx = comp_lx(LC, c), cy = comp_ly(LC, c), cz = comp_lz(LC, c);
            constexpr int dx = comp_lx(LD, d), dy = comp_ly(LD, d), dz = comp_lz(LD, d);
            constexpr int kx = (cx * (LD + 1) + dx) * S::NTK;
            constexpr int ky = (cy * (LD + 1) + dy) * S::NTK;
            constexpr int kz = (cz * (LD + 1) + dz) * S::NTK;
            sfor<NCH>([&](auto ii) {
                double sacc = acc[decltype(ii)::value * NCD + cd];
                sfor<cx + dx + 1>([&](auto itau) {
                    constexpr int tau = decltype(itau)::value;
                    sfor<cy + dy + 1>([&](auto inu) {
                        constexpr int nu = decltype(inu)::value;
                        const double exy = Ekx[kx + tau] * Eky[ky + nu];
                        sfor<cz + dz + 1>([&](auto iphi) {
                            constexpr int phi = decltype(iphi)::value;
                            sacc = fma(exy * Ekz[kz + phi], V[hidx(tau, nu, phi) * NCH + decltype(ii)::value], sacc);
                        });
                    });
                });
                acc[decltype(ii)::value * NCD + cd] = sacc;
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
    sfor<NCH>([&](auto ii) {
        constexpr int ab = A0 + decltype(ii)::value;
        constexpr int a = ab / S::NB, b = ab % S::NB;
        const int p = bfa + a, q = bfb + b;
        const double sab = sq[(size_t)p * N + q];
        const double dab = 4.0 * aD[(size_t)p * N + q];
        sfor<NCD>([&](auto icd) {
            constexpr int cd = decltype(icd)::value;
            constexpr int c = cd / S::ND, d = cd % S::ND;
            const int r = bfc + c, s = bfd + d;
            const double bound = sab * sq[(size_t)r * N + s]
                * dmax6(dab, 4.0 * aD[(size_t)r * N + s], aD[(size_t)p * N + r], aD[(size_t)p * N + s],
                        aD[(size_t)q * N + r], aD[(size_t)q * N + s]);
            const double v = acc[AI(ab, cd)] * deg * ga[a] * gb[b] * gc[c] * gd[d];
            acc[AI(ab, cd)] = (bound < tol) ? 0.0 : v;
        });
    });
    const double* D = A.D;
    double* G = A.G;
    // Coulomb: G_ab += sum_cd (ab|cd) D_cd   (bra rows)
    sfor<NCH>([&](auto ii) {
        constexpr int ab = A0 + decltype(ii)::value;
        constexpr int a = ab / S::NB, b = ab % S::NB;
        double s = 0.0;
        sfor<NCD>([&](auto icd) {
            constexpr int cd = decltype(icd)::value;
            constexpr int c = cd / S::ND, d = cd % S::ND;
            s = fma(acc[AI(ab, cd)], D[(size_t)(bfc + c) * N + bfd + d], s);
        });
        atomicAdd(&rowA[(size_t)a * N + bfb + b], s);
    });
    // Coulomb: G_cd += sum_ab (ab|cd) D_ab   (ket block, global)
    sfor<NCD>([&](auto icd) {
        constexpr int cd = decltype(icd)::value;
        constexpr int c = cd / S::ND, d = cd % S::ND;
        double s = 0.0;
        sfor<NCH>([&](auto ii) {
            constexpr int ab = A0 + decltype(ii)::value;
            constexpr int a = ab / S::NB, b = ab % S::NB;
            s = fma(acc[AI(ab, cd)], D[(size_t)(bfa + a) * N + bfb + b], s);
        });
        atomicAdd(&G[(size_t)(bfc + c) * N + bfd + d], s);
    });
    // Exchange (cython/fock.pyx:82-85): G_bd -= 1/4 (ab|cd) D_ac ;
    // G_bc -= 1/4 (ab|cd) D_ad ; G_ac -= 1/4 (ab|cd) D_bd ; G_ad -= 1/4 (ab|cd) D_bc.
    // Each target element sums the item's contributions in registers first.
    sfor<S::NB>([&](auto ix) {
        constexpr int x = decltype(ix)::value;
        sfor<S::ND>([&](auto iy) {        // (b,d) block
            constexpr int y = decltype(iy)::value;
            double s = 0.0;
            bool any = false;
            sfor<S::NA>([&](auto ia) {
                constexpr int a = decltype(ia)::value;
                constexpr int ab = a * S::NB + x;
                if constexpr (ab >= A0 && ab < A1) {
                    any = true;
                    sfor<S::NC>([&](auto icc) {
                        constexpr int c = decltype(icc)::value;
                        s = fma(acc[AI(ab, c * S::ND + y)], D[(size_t)(bfa + a) * N + bfc + c], s);
                    });
                }
            });
            if (any) atomicAdd(&rowB[(size_t)x * N + bfd + y], -0.25 * s);
        });
        sfor<S::NC>([&](auto iy) {        // (b,c) block
            constexpr int y = decltype(iy)::value;
            double s = 0.0;
            bool any = false;
            sfor<S::NA>([&](auto ia) {
                constexpr int a = decltype(ia)::value;
                constexpr int ab = a * S::NB + x;
                if constexpr (ab >= A0 && ab < A1) {
                    any = true;
                    sfor<S::ND>([&](auto idd) {
                        constexpr int d = decltype(idd)::value;
                        s = fma(acc[AI(ab, y * S::ND + d)], D[(size_t)(bfa + a) * N + bfd + d], s);
                    });
                }
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
                constexpr int ab = x * S::NB + b;
                if constexpr (ab >= A0 && ab < A1) {
                    any = true;
                    sfor<S::ND>([&](auto idd) {
                        constexpr int d = decltype(idd)::value;
                        s = fma(acc[AI(ab, y * S::ND + d)], D[(size_t)(bfb + b) * N + bfd + d], s);
                    });
                }
            });
            if (any) atomicAdd(&rowA[(size_t)x * N + bfc + y], -0.25 * s);
        });
        sfor<S::ND>([&](auto iy) {        // (a,d) block
            constexpr int y = decltype(iy)::value;
            double s = 0.0;
            bool any = false;
            sfor<S::NB>([&](auto ib) {
                constexpr int b = decltype(ib)::value;
                constexpr int ab = x * S::NB + b;
                if constexpr (ab >= A0 && ab < A1) {
                    any = true;
                    sfor<S::NC>([&](auto icc) {
                        constexpr int c = decltype(icc)::value;
                        s = fma(acc[AI(ab, c * S::ND + y)], D[(size_t)(bfb + b) * N + bfc + c], s);
                    });
                }
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
// ket pairs are screened 256 at a time; survivors are appended to an LDS list
// that is digested whenever it holds a full batch, as wave items of (64
// quartets, bra-component chunk, bra primitive slice, ket primitive slice),
// so that lanes stay busy and heavily contracted quartets are spread over
// several lanes.  Coulomb
// G_ab and the four exchange blocks, all in the rows of the bra shells, are
// accumulated in LDS (ds_add_f64) and their nonzero entries added to G once
// per unit; only the
// ket-side Coulomb block is added to G directly.  (The host sends a class
// whose bra rows do not fit in LDS to the wavefront-per-quartet kernel.)
// ---------------------------------------------------------------------------
constexpr int QBATCH = 256;                // survivors digested per batch
constexpr int QLIST = QBATCH + 256;        // batch + one screening step
constexpr int PSLICE = 8;                  // primitive pairs per slice (bra and ket)

template <int LA, int LB, int LC, int LD>
__global__ void __launch_bounds__(256) eri_fock_class_kernel(ClassArgs A, RowArgs R)
{
    using S = ClassShape<LA, LB, LC, LD>;
    extern __shared__ double dyn[];
    double* lds_boys = dyn;
    double* rows = dyn + BOYS_TAYLOR * BOYS_NT;
    double* bp = rows + (S::NA + S::NB) * A.N;
    int* qlist = reinterpret_cast<int*>(bp + (size_t)R.maxpp * S::BPS);
    int* misc = qlist + QLIST;
    boys_stage_lds<S::L>(lds_boys, A.boys);
    const int N = A.N;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nwv = (int)(blockDim.x >> 6);
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
        stage_bra<LA, LB, LC, LD>(A, bra, bp);
        if (threadIdx.x == 0) {
            misc[1] = 0;   // survivors in the list
            misc[2] = 0;   // largest ket primitive-pair count in the list
        }
        __syncthreads();
        const int nbs = (bra.npp + PSLICE - 1) / PSLICE;
        for (int t0 = 0;; t0 += (int)blockDim.x) {
            // screen the next kets and append the survivors
            if (t0 < segn) {
                const int t = t0 + threadIdx.x;
                bool keep = false;
                int npk = 0;
                if (t < segn) {
                    const ShellPair ket = A.pairs[R.koff + k0 + t];
                    const double* Dsh = R.Dsh;
                    const double dm = fmax(fmax(fmax(4.0 * Dsh[bra.sa * ns + bra.sb], 4.0 * Dsh[ket.sa * ns + ket.sb]),
                                                fmax(Dsh[bra.sa * ns + ket.sa], Dsh[bra.sa * ns + ket.sb])),
                                           fmax(Dsh[bra.sb * ns + ket.sa], Dsh[bra.sb * ns + ket.sb]));
                    keep = bra.q * ket.q * dm >= A.tol;
                    npk = ket.npp;
                }
                const unsigned long long m = __ballot(keep);
                int pos = 0;
                if (lane == 0 && m) pos = atomicAdd(&misc[1], __popcll(m));
                pos = __shfl(pos, 0);
                if (keep) {
                    qlist[pos + __popcll(m & ((1ull << lane) - 1ull))] = k0 + t;
                    atomicMax(&misc[2], npk);
                }
            }
            __syncthreads();
            const int cnt = misc[1];
            const bool last = t0 + (int)blockDim.x >= segn;
            if (cnt >= QBATCH || (last && cnt > 0)) {
                // digest the batch: (64-quartet block, component chunk, bra slice, ket slice)
                if (R.stat_nquart && threadIdx.x == 0) atomicAdd(R.stat_nquart, (unsigned long long)cnt);
                const int nks = (misc[2] + PSLICE - 1) / PSLICE;
                const int total = ((cnt + 63) / 64) * S::NCHUNK * nbs * nks;
                for (int w = wv; w < total; w += nwv) {
                    const int chunk = w % S::NCHUNK;
                    int r = w / S::NCHUNK;
                    const int bs = r % nbs;
                    r /= nbs;
                    const int ks = r % nks;
                    const int idx = (r / nks) * 64 + lane;
                    if (idx >= cnt) continue;
                    const int Q = qlist[idx];
                    const int2 qd = make_int2(R.boff + P, R.koff + Q);
                    const ShellPair ket = A.pairs[qd.y];
                    const int jq0 = ks * PSLICE;
                    if (jq0 >= ket.npp) continue;
                    const int jq1 = min(ket.npp, jq0 + PSLICE);
                    const int ip0 = bs * PSLICE, ip1 = min(bra.npp, ip0 + PSLICE);
                    if (R.stat_nprim && chunk == 0)
                        atomicAdd(R.stat_nprim, (unsigned long long)(ip1 - ip0) * (unsigned long long)(jq1 - jq0));
                    static_switch<0, S::NCHUNK>(chunk, [&](auto ic) {
                        constexpr int C = decltype(ic)::value;
                        quartet_item<LA, LB, LC, LD, C * S::CH, (C + 1) * S::CH>(
                            A, lds_boys, bp, qd, bra, ket, ip0, ip1, jq0, jq1, rows,
                            rows + (bra.same ? 0 : S::NA * N));
                    });
                }
                __syncthreads();
                if (threadIdx.x == 0) {
                    misc[1] = 0;
                    misc[2] = 0;
                }
                __syncthreads();
            }
            if (last) break;
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
inline size_t fock_class_lds_bytes(int N, int maxpp)
{
    using S = ClassShape<LA, LB, LC, LD>;
    return sizeof(double) * (BOYS_TAYLOR * BOYS_NT + (size_t)(S::NA + S::NB) * N + (size_t)maxpp * S::BPS)
         + sizeof(int) * (QLIST + 4);
}

}  // namespace mmd
