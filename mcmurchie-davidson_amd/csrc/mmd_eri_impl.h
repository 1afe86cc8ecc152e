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
#ifndef MMD_IBUF_CAP
#define MMD_IBUF_CAP 2048
#endif
constexpr int IBUF_CAP = MMD_IBUF_CAP;     // doubles of the LDS integral buffer that set its batch size
constexpr int ACC_LIMIT_HIGH = 64;

template <int LA, int LB, int LC, int LD>
struct ClassShape {
    static constexpr int L = LA + LB + LC + LD;
    static constexpr int LAB = LA + LB, LCD = LC + LD;
    static constexpr int HCD = nherm(LCD);
    static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    static constexpr int NAB = NA * NB, NCD = NC * ND;
    // bra components per item: accumulators (ket Hermite + ket Cartesian
    // per component) up to ACC_LIMIT doubles, or ACC_LIMIT_HIGH for L >= 4,
    // whose R_tuv dominates the registers anyway and whose chunks each
    // recompute Boys and R
    static constexpr int CH = chunk_size(NAB, HCD, NCD, L >= 4 ? ACC_LIMIT_HIGH : ACC_LIMIT);
    static constexpr int NCHUNK = NAB / CH;
    static constexpr int NTB = LA + LB + 1, NTK = LC + LD + 1;
    static constexpr int EB = (LA + 1) * (LB + 1) * NTB;   // 1-D bra E table
    static constexpr int BPS = 8 + 3 * EB;                 // staged bra primitive stride (doubles)
    // Classes with few integrals per quartet (the heavily contracted low-L
    // ones) add the primitive-slice partials of a batch into an LDS buffer
    // and digest each quartet once; QB is their batch size (buffer <= 16 KB).
    static constexpr int NABCD = NAB * NCD;
    static constexpr bool IBUF = NABCD <= 32;
    // Items are (quartet, ket primitive slice) pairs, not whole quartets,
    // for the buffered classes and for the classes whose ket pairs are the
    // heavily contracted (ss) and (ps) pairs (up to 81 primitive pairs per
    // ket in cc-pVDZ): there a whole-ket lane would carry up to 81 primitive
    // quartets alone and set the length of its batch.  Unbuffered classes
    // digest each slice's partial block directly (the digestion is linear).
    // (pp|pp) is sliced too: its units hold few surviving kets in batches of
    // small molecules, and slices multiply its lane items (a 6-31G batch
    // pass measured 285 against 298 ms; (H2O)8 cc-pVDZ builds unchanged).
    static constexpr bool KSL = IBUF || (LC <= 1 && LD == 0) || (LA <= 1 && LC <= 1);
    // The other classes digest each (quartet, chunk) item; when a quartet
    // has several chunks, their ket-side Coulomb blocks are summed in an LDS
    // buffer per batch (KJ) and added to G once per quartet.
    static constexpr bool KJ = !IBUF && NCHUNK > 1;
    static constexpr int QB = IBUF ? (IBUF_CAP / NABCD >= 256 ? 256 : (IBUF_CAP / NABCD >= 128 ? 128 : 64))
                                   : (KJ ? (2048 / NCD >= 256 ? 256 : (2048 / NCD >= 128 ? 128 : 64)) : 256);
    static constexpr int IBUF_DOUBLES = IBUF ? QB * NABCD : 0;
    static constexpr int KJ_DOUBLES = KJ ? QB * NCD : 0;   // used when the launch has LDS for it
};

__device__ __forceinline__ double dmax6(double a, double b, double c, double d, double e, double f)
{
    return fmax(fmax(fmax(a, b), fmax(c, d)), fmax(e, f));
}

// Stage primitive pairs ip0 .. ip0+np-1 of the bra pair in LDS: p,
// 2 pi^{5/2} K / p, P and the three 1-D Hermite tables E^{ab}_t
// (cython/util.pxi:13-28).
template <int LA, int LB, int LC, int LD, typename T>
__device__ __forceinline__ void stage_bra(const ClassArgs& A, const ShellPair& bra, int ip0, int np, T* bp)
{
    using S = ClassShape<LA, LB, LC, LD>;
    for (int idx = threadIdx.x; idx < 3 * np; idx += blockDim.x) {
        const int ip = idx / 3, x = idx - 3 * ip;
        const PrimPair* P = A.pp + bra.pp0 + ip0 + ip;
        T* b = bp + ip * S::BPS;
        if (x == 0) {
            b[0] = (T)P->p;
            b[1] = (T)(TWO_PI_2_5 * P->K * P->invp);
            b[2] = (T)P->Px;
            b[3] = (T)P->Py;
            b[4] = (T)P->Pz;
        }
        double E[S::EB];
        hermite_E<LA, LB, 1>(P->inv2p, x == 0 ? P->PAx : (x == 1 ? P->PAy : P->PAz),
                             x == 0 ? P->PBx : (x == 1 ? P->PBy : P->PBz), E);
        sfor<S::EB>([&](auto k) { b[8 + x * S::EB + k] = (T)E[k]; });
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
// The integrals of the item (its bra components x all ket components) are
// returned in acc; quartet_digest adds them to the Fock matrix.
// ---------------------------------------------------------------------------
template <int LA, int LB, int LC, int LD, int A0, int A1, typename T>
__device__ __forceinline__ void quartet_compute(const ClassArgs& A, const T* lds, const T* bp,
                                                const ShellPair& ket, int ip0, int ip1, int jq0, int jq1,
                                                T* acc)
{
    using S = ClassShape<LA, LB, LC, LD>;
    constexpr int NCH = A1 - A0;
    constexpr int NCD = S::NCD;
    constexpr int L = S::L;
    constexpr int HCD = S::HCD;
    sfor<NCH * NCD>([&](auto k) { acc[k] = T(0); });

    for (int jq = jq0; jq < jq1; ++jq) {
        const PrimPair* Q = A.pp + ket.pp0 + jq;
        const T Qp = (T)Q->p, QK = (T)(Q->K * Q->invp), Qx = (T)Q->Px, Qy = (T)Q->Py, Qz = (T)Q->Pz;
        T V[HCD * NCH];
        sfor<HCD * NCH>([&](auto k) { V[k] = T(0); });
        for (int ip = ip0; ip < ip1; ++ip) {
            const T* b = bp + ip * S::BPS;
            const T Pp = b[0];
            const T rs = dev_rsqrt(Pp + Qp);
            const T alpha = Pp * Qp * rs * rs;
            const T X = b[2] - Qx, Y = b[3] - Qy, Z = b[4] - Qz;
            const T Targ = alpha * (X * X + Y * Y + Z * Z);
            const T pref = b[1] * QK * rs;   // 2 pi^{5/2} / (p q sqrt(p+q)) K_ab K_cd
            T F[L + 1];
            boys_seeds<L, T>(Targ, alpha, pref, lds, F);
            T R[nherm(L)];
            hermite_R<L, T>(X, Y, Z, F, R);
            const T* Ebx = b + 8;
            const T* Eby = b + 8 + S::EB;
            const T* Ebz = b + 8 + 2 * S::EB;
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
                    T v = V[decltype(ik)::value * NCH + decltype(ii)::value];
                    sfor<ax + bx + 1>([&](auto it) {
                        constexpr int t = decltype(it)::value;
                        sfor<ay + by + 1>([&](auto iu) {
                            constexpr int u = decltype(iu)::value;
                            const T exy = Ebx[ix + t] * Eby[iy + u];
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
        T Ekx[(LC + 1) * (LD + 1) * S::NTK], Eky[(LC + 1) * (LD + 1) * S::NTK],
            Ekz[(LC + 1) * (LD + 1) * S::NTK];
        if constexpr (LC + LD > 0) {
            const T qinv2 = (T)Q->inv2p;
            hermite_E<LC, LD, -1, T>(qinv2, (T)Q->PAx, (T)Q->PBx, Ekx);
            hermite_E<LC, LD, -1, T>(qinv2, (T)Q->PAy, (T)Q->PBy, Eky);
            hermite_E<LC, LD, -1, T>(qinv2, (T)Q->PAz, (T)Q->PBz, Ekz);
        } else {
            Ekx[0] = Eky[0] = Ekz[0] = T(1);
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
                T sacc = acc[decltype(ii)::value * NCD + cd];
                sfor<cx + dx + 1>([&](auto itau) {
                    constexpr int tau = decltype(itau)::value;
                    sfor<cy + dy + 1>([&](auto inu) {
                        constexpr int nu = decltype(inu)::value;
                        const T exy = Ekx[kx + tau] * Eky[ky + nu];
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
}

// Digest the integrals acc of a quartet (bra components A0..A1-1) into G.
// rowA / rowB: rows of G for the bra shells' functions accumulated in LDS,
// row x holding column `col` of G at rowA[x * RW + col] (the caller offsets
// the pointers by the first column of the molecule); the ket-side Coulomb
// block goes to G.
template <int LA, int LB, int LC, int LD, int A0, int A1>
__device__ __forceinline__ void quartet_digest(const ClassArgs& A, int2 qd, const ShellPair& bra,
                                               const ShellPair& ket, double* acc, double* rowA, double* rowB,
                                               int RW, double* kj = nullptr)
{
    using S = ClassShape<LA, LB, LC, LD>;
    constexpr int NCH = A1 - A0;
    constexpr int NCD = S::NCD;
    auto AI = [](int ab, int cd) { return (ab - A0) * NCD + cd; };
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
        atomicAdd(&rowA[a * RW + bfb + b], s);
    });
    // Coulomb: G_cd += sum_ab (ab|cd) D_ab   (ket block: to G, or to the
    // quartet's LDS row kj when given, added to G after the batch)
    sfor<NCD>([&](auto icd) {
        constexpr int cd = decltype(icd)::value;
        constexpr int c = cd / S::ND, d = cd % S::ND;
        double s = 0.0;
        sfor<NCH>([&](auto ii) {
            constexpr int ab = A0 + decltype(ii)::value;
            constexpr int a = ab / S::NB, b = ab % S::NB;
            s = fma(acc[AI(ab, cd)], D[(size_t)(bfa + a) * N + bfb + b], s);
        });
        if (kj)
            atomicAdd(&kj[cd], s);
        else
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
            if (any) atomicAdd(&rowB[x * RW + bfd + y], -0.25 * s);
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
            if (any) atomicAdd(&rowB[x * RW + bfc + y], -0.25 * s);
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
            if (any) atomicAdd(&rowA[x * RW + bfc + y], -0.25 * s);
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
            if (any) atomicAdd(&rowA[x * RW + bfd + y], -0.25 * s);
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
// Class kernel.  A work unit is one bra shell pair ("row"), a slice of at
// most PSLICE of its primitive pairs and a segment of its ket pairs (a
// host-built table, heaviest first); workgroups take units from a device
// counter.  The unit's ket pairs are screened a workgroup at a time;
// survivors are appended to an LDS list that is digested whenever it holds
// a full batch, as wave items of (bra-component chunk, 64 consecutive
// (ket primitive slice, quartet) lane items).  Classes with an integral
// buffer (IBUF) slice the ket primitives too, add each item's partial
// integrals into LDS and digest every quartet once, after the batch; the
// others contract a quartet's whole ket side in one lane and digest it.  Coulomb G_ab and the four exchange
// blocks, all in the rows of the bra shells, are accumulated in LDS
// (ds_add_f64) and their nonzero entries added to G once per unit; only the
// ket-side Coulomb block is added to G directly.  (The host sends a class
// whose bra rows do not fit in LDS to the wavefront-per-quartet kernel.)
// ---------------------------------------------------------------------------
constexpr int CLASS_BLOCK = 256;           // threads per workgroup

#ifndef MMD_BOYS_LDS_MIN_L
#define MMD_BOYS_LDS_MIN_L 0
#endif
constexpr int BOYS_LDS_MIN_L = MMD_BOYS_LDS_MIN_L;   // classes with L below it read Boys rows from L1

template <int QB>
struct ClassList {
    static constexpr int QLIST = QB + CLASS_BLOCK;   // batch + one screening step
};

// T is the arithmetic type of the integrals (double, or float for the
// FP32 tier); screening, digestion and all accumulation into G are FP64.
template <int LA, int LB, int LC, int LD, typename T>
__global__ void __launch_bounds__(CLASS_BLOCK) eri_fock_class_kernel(ClassArgs A, RowArgs R)
{
    using S = ClassShape<LA, LB, LC, LD>;
    constexpr int QB = S::QB;
    static_assert(QB <= CLASS_BLOCK, "one thread per batch entry in the slice-count scan");
    extern __shared__ double dyn[];
    double* rows = dyn;
    const int RW = R.rw;
    double* ibuf = rows + (S::NA + S::NB) * RW;
    // ket-side Coulomb buffer (KJ classes, when the launch has room for it)
    const bool kj = S::KJ && R.kj;
    const int nbuf = S::IBUF_DOUBLES + (kj ? S::KJ_DOUBLES : 0);
    // the class's Boys rows in LDS, or read from the global table (L1) for
    // the classes below BOYS_LDS_MIN_L, which leaves their workgroups the LDS
    // for more of them per CU (the default keeps every class in LDS: reading
    // the rows through L1 made the (ps|ss) class 20% slower)
    constexpr bool boys_lds = S::L >= BOYS_LDS_MIN_L;
    T* lds_boys = reinterpret_cast<T*>(ibuf + nbuf);
    T* bp = boys_lds ? lds_boys + BOYS_TAYLOR * BOYS_NT : lds_boys;
    int* qlist = reinterpret_cast<int*>(bp + (size_t)(R.maxpp < PSLICE ? R.maxpp : PSLICE) * S::BPS);
    int* misc = qlist + ClassList<QB>::QLIST;
    int* pre = misc + 4;         // IBUF classes: prefix sums of the batch's ket slice counts
    int* wsum = pre + QB + 1;    // per-wavefront partial sums of that scan
    const T* boys_tab;
    if constexpr (std::is_same<T, float>::value) {
        if constexpr (boys_lds) boys_stage_lds<S::L, float>(lds_boys, A.boysf);
        boys_tab = boys_lds ? lds_boys : A.boysf + (size_t)S::L * BOYS_NT;
    } else {
        if constexpr (boys_lds) boys_stage_lds<S::L, double>(lds_boys, A.boys);
        boys_tab = boys_lds ? lds_boys : A.boys + (size_t)S::L * BOYS_NT;
    }
    for (int i = threadIdx.x; i < nbuf; i += blockDim.x) ibuf[i] = 0.0;
    const int N = A.N;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nwv = (int)(blockDim.x >> 6);
    const int nunits = R.nunits;
    const size_t ns = (size_t)R.nshell;
    unsigned long long my_nprim = 0;   // primitive quartets of this lane (statistics)
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) misc[0] = atomicAdd(R.row_counter, 1);
        __syncthreads();
        const int u = misc[0];
        if (u >= nunits) {
            if (R.stat_nprim) {   // one atomic per wavefront
                for (int off = 32; off > 0; off >>= 1) my_nprim += __shfl_xor(my_nprim, off);
                if (lane == 0 && my_nprim) atomicAdd(R.stat_nprim, my_nprim);
            }
            break;
        }
        // unit: bra pair P, its primitive pairs [ip0, ip1) and the ket
        // pairs [k0, k1) of its molecule
        // a unit of a batch whose every quartet fails the screen: pair
        // factors q_P q_Q <= the unit's bound, and the density weight of the
        // screen is at most 4 max |dP| of the molecule (0 once converged)
        if (R.mol_dmax) {
            const UnitBound ub = R.ubound[u];
            if ((double)ub.q * 4.0 * R.mol_dmax[ub.mol] < 0.999999 * A.tol) continue;
        }
        const int4 U = R.units[u];
        const int P = U.x, k0 = U.z;
        const int ip0u = U.y & 0xffff, ip1u = U.y >> 16;
        const int segn = U.w - U.z;
        const int bnp = ip1u - ip0u;   // bra primitive pairs of the unit (<= PSLICE)
        const ShellPair bra = A.pairs[R.boff + P];
        const int nrows = bra.same ? S::NA : S::NA + S::NB;
        const int c0 = bra.c0;       // the rows hold the molecule's columns c0 .. c0+RW-1
        double* rowA = rows - c0;
        double* rowB = rows + (bra.same ? 0 : S::NA * RW) - c0;
        for (int i = threadIdx.x; i < nrows * RW; i += blockDim.x) rows[i] = 0.0;
        stage_bra<LA, LB, LC, LD, T>(A, bra, ip0u, bnp, bp);
        if (threadIdx.x == 0) misc[1] = 0;   // survivors in the list
        __syncthreads();
        for (int t0 = 0;; t0 += (int)blockDim.x) {
            // screen the next kets and append the survivors
            if (t0 < segn) {
                const int t = t0 + threadIdx.x;
                bool keep = false;
                if (t < segn) {
                    const ShellPair ket = A.pairs[R.koff + k0 + t];
                    const double* Dsh = R.Dsh;
                    const double dm = fmax(fmax(fmax(4.0 * Dsh[bra.sa * ns + bra.sb], 4.0 * Dsh[ket.sa * ns + ket.sb]),
                                                fmax(Dsh[bra.sa * ns + ket.sa], Dsh[bra.sa * ns + ket.sb])),
                                           fmax(Dsh[bra.sb * ns + ket.sa], Dsh[bra.sb * ns + ket.sb]));
                    const double bound = bra.q * ket.q * dm;
                    keep = bound >= A.tol && bound >= R.tier_lo && bound < R.tier_hi;
                }
                const unsigned long long m = __ballot(keep);
                int pos = 0;
                if (lane == 0 && m) pos = atomicAdd(&misc[1], __popcll(m));
                pos = __shfl(pos, 0);
                if (keep) qlist[pos + __popcll(m & ((1ull << lane) - 1ull))] = k0 + t;
            }
            __syncthreads();
            const int cnt0 = misc[1];
            const bool last = t0 + (int)blockDim.x >= segn;
            for (int b0 = 0; b0 < cnt0 && (cnt0 - b0 >= QB || last); b0 += QB) {
                // digest the batch qlist[b0, b0+cnt): a wavefront takes one
                // chunk of bra components and 64 consecutive lane items.
                // Classes that add partial integrals into the LDS buffer
                // split each ket's primitives into slices of R.kslice; their
                // lane items are the (quartet, ket slice) pairs that exist,
                // numbered through the prefix sums pre[] of the quartets'
                // slice counts, so that lanes stay busy however the batch
                // mixes contraction lengths.  The other classes take a
                // quartet's whole ket contraction in one lane and digest it.
                const int cnt = min(QB, cnt0 - b0);
                if (R.stat_nquart && threadIdx.x == 0) atomicAdd(R.stat_nquart, (unsigned long long)cnt);
                int nflat = cnt;
                if constexpr (S::KSL) {
                    const int j = threadIdx.x;   // cnt <= QB <= CLASS_BLOCK
                    int x = 0;
                    if (j < cnt) x = (A.pairs[R.koff + qlist[b0 + j]].npp + R.kslice - 1) / R.kslice;
                    for (int off = 1; off < 64; off <<= 1) {
                        const int y = __shfl_up(x, off);
                        if (lane >= off) x += y;
                    }
                    if (lane == 63) wsum[wv] = x;
                    __syncthreads();
                    for (int k = 0; k < wv; ++k) x += wsum[k];
                    if (j < cnt) pre[j + 1] = x;
                    if (j == 0) pre[0] = 0;
                    __syncthreads();
                    nflat = pre[cnt];
                }
                const int total = ((nflat + 63) / 64) * S::NCHUNK;
                for (int w = wv; w < total; w += nwv) {
                    const int chunk = w % S::NCHUNK;
                    const int flat = (w / S::NCHUNK) * 64 + lane;
                    if (flat >= nflat) continue;
                    int idx = flat, ks = 0;
                    if constexpr (S::KSL) {
                        // the quartet whose slices cover item flat: pre[idx] <= flat < pre[idx + 1]
                        int lo = 0, hi = cnt - 1;
                        while (lo < hi) {
                            const int mid = (lo + hi + 1) >> 1;
                            if (pre[mid] <= flat) lo = mid;
                            else hi = mid - 1;
                        }
                        idx = lo;
                        ks = flat - pre[lo];
                    }
                    const int Q = qlist[b0 + idx];
                    const int2 qd = make_int2(R.boff + P, R.koff + Q);
                    const ShellPair ket = A.pairs[qd.y];
                    const int jq0 = ks * R.kslice;
                    if (jq0 >= ket.npp) continue;
                    const int jq1 = S::KSL ? min(ket.npp, jq0 + R.kslice) : ket.npp;
                    const int ip0 = 0, ip1 = bnp;
                    if (chunk == 0) my_nprim += (unsigned long long)(ip1 - ip0) * (unsigned long long)(jq1 - jq0);
                    static_switch<0, S::NCHUNK>(chunk, [&](auto ic) {
                        constexpr int C = decltype(ic)::value;
                        constexpr int A0 = C * S::CH, A1 = (C + 1) * S::CH;
                        T acc[(A1 - A0) * S::NCD];
                        quartet_compute<LA, LB, LC, LD, A0, A1, T>(A, boys_tab, bp, ket, ip0, ip1, jq0, jq1, acc);
                        if constexpr (S::IBUF) {
                            double* dst = ibuf + (size_t)idx * S::NABCD + A0 * S::NCD;
                            sfor<(A1 - A0) * S::NCD>(
                                [&](auto k) { atomicAdd(&dst[decltype(k)::value], (double)acc[k]); });
                        } else {
                            double accd[(A1 - A0) * S::NCD];
                            sfor<(A1 - A0) * S::NCD>([&](auto k) { accd[k] = (double)acc[k]; });
                            quartet_digest<LA, LB, LC, LD, A0, A1>(A, qd, bra, ket, accd, rowA, rowB, RW,
                                                                   kj ? ibuf + (size_t)idx * S::NCD : nullptr);
                        }
                    });
                }
                if (kj) {
                    // the summed ket-side Coulomb blocks of the batch into G
                    // (kj is uniform across the workgroup)
                    __syncthreads();
                    for (int i = threadIdx.x; i < cnt * S::NCD; i += blockDim.x) {
                        const double v = ibuf[i];
                        if (v == 0.0) continue;
                        ibuf[i] = 0.0;
                        const int idx = i / S::NCD, cd = i - idx * S::NCD;
                        const ShellPair ket = A.pairs[R.koff + qlist[b0 + idx]];
                        atomicAdd(&A.G[(size_t)(ket.bfa + cd / S::ND) * N + ket.bfb + cd % S::ND], v);
                    }
                }
                if constexpr (S::IBUF) {
                    // every quartet of the batch once: its summed integrals into G
                    __syncthreads();
                    const int total2 = ((cnt + 63) / 64) * S::NCHUNK;
                    for (int w = wv; w < total2; w += nwv) {
                        const int chunk = w % S::NCHUNK;
                        const int idx = (w / S::NCHUNK) * 64 + lane;
                        if (idx >= cnt) continue;
                        const int Q = qlist[b0 + idx];
                        const int2 qd = make_int2(R.boff + P, R.koff + Q);
                        const ShellPair ket = A.pairs[qd.y];
                        static_switch<0, S::NCHUNK>(chunk, [&](auto ic) {
                            constexpr int C = decltype(ic)::value;
                            constexpr int A0 = C * S::CH, A1 = (C + 1) * S::CH;
                            double acc[(A1 - A0) * S::NCD];
                            double* src = ibuf + (size_t)idx * S::NABCD + A0 * S::NCD;
                            sfor<(A1 - A0) * S::NCD>([&](auto k) {
                                acc[k] = src[decltype(k)::value];
                                src[decltype(k)::value] = 0.0;
                            });
                            quartet_digest<LA, LB, LC, LD, A0, A1>(A, qd, bra, ket, acc, rowA, rowB, RW);
                        });
                    }
                }
                __syncthreads();
            }
            if (last) break;
            // keep the survivors that did not fill a batch at the front of the list
            const int done = (cnt0 / QB) * QB;
            const int rest = cnt0 - done;
            if (done > 0) {
                int v = 0;
                if ((int)threadIdx.x < rest) v = qlist[done + threadIdx.x];
                __syncthreads();
                if ((int)threadIdx.x < rest) qlist[threadIdx.x] = v;
                if (threadIdx.x == 0) misc[1] = rest;
            }
            __syncthreads();
        }
        __syncthreads();
        for (int i = threadIdx.x; i < nrows * RW; i += blockDim.x) {
            const double v = rows[i];
            const int x = i / RW, col = c0 + i - x * RW;
            const int g = x < S::NA ? bra.bfa + x : bra.bfb + (x - S::NA);
            if (v != 0.0) atomicAdd(&A.G[(size_t)g * N + col], v);
        }
    }
}

// ---------------------------------------------------------------------------
// Stored ERI tensor (doERIs, cython/twoe.pyx:9-31) from shell quartets.
// bra_table_kernel writes the staged bra data of stage_bra for every
// primitive pair of the class's bra pairs to global memory (table + pp*BPS);
// eri_tensor_class_kernel gives one lane per (shell quartet, bra-component
// chunk), contracts all primitive pairs with quartet_compute and stores each
// canonical integral at its eight index permutations (cython/twoe.pyx:23-30),
// so that every element of the tensor is written once and the tensor has the
// permutational symmetry exactly.
// ---------------------------------------------------------------------------
template <int LA, int LB, int LC, int LD>
__global__ void __launch_bounds__(256) bra_table_kernel(ClassArgs A, int boff, double* table)
{
    using S = ClassShape<LA, LB, LC, LD>;
    const ShellPair bra = A.pairs[boff + blockIdx.x];
    for (int idx = threadIdx.x; idx < 3 * bra.npp; idx += blockDim.x) {
        const int ip = idx / 3, x = idx - 3 * ip;
        const PrimPair* P = A.pp + bra.pp0 + ip;
        double* b = table + (size_t)(bra.pp0 + ip) * S::BPS;
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

template <int LA, int LB, int LC, int LD>
__global__ void __launch_bounds__(256) eri_tensor_class_kernel(ClassArgs A, int boff, int nb, int koff, int nk,
                                                               int tri, const double* __restrict__ table,
                                                               double* __restrict__ out)
{
    using S = ClassShape<LA, LB, LC, LD>;
    __shared__ double lds_boys[BOYS_TAYLOR * BOYS_NT];
    boys_stage_lds<S::L, double>(lds_boys, A.boys);
    __syncthreads();
    const int64_t nq = tri ? (int64_t)nb * (nb + 1) / 2 : (int64_t)nb * nk;
    const int64_t items = nq * S::NCHUNK;
    const size_t N = (size_t)A.N;
    for (int64_t it = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; it < items;
         it += (int64_t)gridDim.x * blockDim.x) {
        const int chunk = (int)(it / nq);   // uniform across most wavefronts
        const int64_t q = it - (int64_t)chunk * nq;
        int P, Q;
        if (tri) {
            int64_t r = (int64_t)((sqrt(8.0 * (double)q + 1.0) - 1.0) * 0.5);
            while (r * (r + 1) / 2 > q) --r;
            while ((r + 1) * (r + 2) / 2 <= q) ++r;
            P = (int)r;
            Q = (int)(q - r * (r + 1) / 2);
        } else {
            P = (int)(q / nk);
            Q = (int)(q - (int64_t)P * nk);
        }
        const ShellPair bra = A.pairs[boff + P];
        const ShellPair ket = A.pairs[koff + Q];
        static_switch<0, S::NCHUNK>(chunk, [&](auto ic) {
            constexpr int C = decltype(ic)::value;
            constexpr int A0 = C * S::CH, A1 = (C + 1) * S::CH;
            double acc[(A1 - A0) * S::NCD];
            quartet_compute<LA, LB, LC, LD, A0, A1, double>(A, lds_boys, table + (size_t)bra.pp0 * S::BPS, ket, 0,
                                                            bra.npp, 0, ket.npp, acc);
            sfor<A1 - A0>([&](auto ii) {
                constexpr int ab = A0 + decltype(ii)::value;
                constexpr int a = ab / S::NB, b = ab % S::NB;
                const size_t i = bra.bfa + a, j = bra.bfb + b;
                const double gab = A.gcomp[i] * A.gcomp[j];
                sfor<S::NCD>([&](auto icd) {
                    constexpr int cd = decltype(icd)::value;
                    constexpr int c = cd / S::ND, d = cd % S::ND;
                    const size_t k = ket.bfa + c, l = ket.bfb + d;
                    const size_t ij = i * N + j, ji = j * N + i, kl = k * N + l, lk = l * N + k;
                    // one producer per orbit of the eight permutations: the
                    // others (same shell, or the same pair on both sides)
                    // hold the same integral with different rounding
                    if ((bra.same && i < j) || (ket.same && k < l) || (tri && P == Q && ij < kl)) return;
                    const double v = acc[decltype(ii)::value * S::NCD + cd] * gab * A.gcomp[k] * A.gcomp[l];
                    out[ij * N * N + kl] = v;
                    out[ji * N * N + kl] = v;
                    out[ij * N * N + lk] = v;
                    out[ji * N * N + lk] = v;
                    out[kl * N * N + ij] = v;
                    out[lk * N * N + ij] = v;
                    out[kl * N * N + ji] = v;
                    out[lk * N * N + ji] = v;
                });
            });
        });
    }
}

template <int LA, int LB, int LC, int LD, typename T>
inline size_t fock_class_lds_bytes(int RW, int maxpp, int kj)
{
    using S = ClassShape<LA, LB, LC, LD>;
    return sizeof(double) * ((size_t)(S::NA + S::NB) * RW + S::IBUF_DOUBLES + (kj ? S::KJ_DOUBLES : 0))
         + sizeof(T) * ((S::L >= BOYS_LDS_MIN_L ? BOYS_TAYLOR * BOYS_NT : 0)
                        + (size_t)(maxpp < PSLICE ? maxpp : PSLICE) * S::BPS)
         + sizeof(int) * (ClassList<S::QB>::QLIST + 4 + S::QB + 1 + CLASS_BLOCK / 64);
}

}  // namespace mmd
