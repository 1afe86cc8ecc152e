#pragma once
// Shell-quartet ERI kernels for gfx950, one instantiation per angular-momentum
// class (la lb | lc ld), la >= lb, lc >= ld, pair_class(bra) >= pair_class(ket).
//
// Work decomposition: one lane per (contracted shell quartet, chunk of ket
// Cartesian components); the chunk index is uniform across a wavefront, so
// every lane of a wave runs the same compile-time specialised code.  For each
// primitive quartet the lane evaluates the Boys function from the LDS-staged
// Taylor table, runs the R_tuv recursion in registers, contracts the ket
// Hermite expansion into H_tuv and the bra expansion into the contracted
// integrals (ab|cd), accumulated over all primitive pairs in registers.
// The epilogue digests the block into the Fock matrix: Coulomb and exchange
// with the incremental density, after the Schwarz x |dP| test.
//
// Reference semantics: cython/twoe.pyx:52-94 (primitive ERI),
// cython/twoe.pyx:34-50 (contraction), cython/fock.pyx:38-85 (screening,
// degeneracy, Coulomb/exchange scatter).
//
// This header is instantiated once per class by mmd_eri_class.hip.
#include <hip/hip_runtime.h>
#include <type_traits>
#include "mmd_device.h"
#include "mmd_types.h"
#include "mmd_kernels.h"

namespace mmd {

constexpr double TWO_PI_2_5 = 34.98683665524972497;   // 2 pi^{5/2}

// Largest divisor of ncd whose chunk keeps nab*chunk accumulators <= limit.
__host__ __device__ constexpr int chunk_size(int nab, int ncd, int limit)
{
    int best = 1;
    for (int c = 1; c <= ncd; ++c)
        if (ncd % c == 0 && nab * c <= limit) best = c;
    return best;
}
constexpr int ACC_LIMIT = 54;

template <int LA, int LB, int LC, int LD>
struct ClassShape {
    static constexpr int L = LA + LB + LC + LD;
    static constexpr int LAB = LA + LB;
    static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    static constexpr int NAB = NA * NB, NCD = NC * ND;
    static constexpr int CH = chunk_size(NAB, NCD, ACC_LIMIT);
    static constexpr int NCHUNK = NCD / CH;
    static constexpr int NTB = LA + LB + 1, NTK = LC + LD + 1;
};

// ---------------------------------------------------------------------------
// Hermite contraction of one primitive quartet into acc[ab][cd - C0].
// ---------------------------------------------------------------------------
template <int LA, int LB, int LC, int LD, int C0, int C1>
__device__ __forceinline__ void contract(const double* R, const double* Ebx, const double* Eby,
                                         const double* Ebz, const double* Ekx, const double* Eky,
                                         const double* Ekz, double* acc)
{
    using S = ClassShape<LA, LB, LC, LD>;
    constexpr int NCH = C1 - C0;
#pragma unroll
    for (int cd = C0; cd < C1; ++cd) {
        const int c = cd / S::ND, d = cd % S::ND;
        const int cx = comp_lx(LC, c), cy = comp_ly(LC, c), cz = comp_lz(LC, c);
        const int dx = comp_lx(LD, d), dy = comp_ly(LD, d), dz = comp_lz(LD, d);
        const int kx = (cx * (LD + 1) + dx) * S::NTK;
        const int ky = (cy * (LD + 1) + dy) * S::NTK;
        const int kz = (cz * (LD + 1) + dz) * S::NTK;
        double H[nherm(S::LAB)];
#pragma unroll
        for (int t = 0; t <= S::LAB; ++t) {
#pragma unroll
            for (int u = 0; u <= S::LAB - t; ++u) {
#pragma unroll
                for (int v = 0; v <= S::LAB - t - u; ++v) {
                    double h = 0.0;
#pragma unroll
                    for (int tau = 0; tau <= cx + dx; ++tau) {
#pragma unroll
                        for (int nu = 0; nu <= cy + dy; ++nu) {
                            const double exy = Ekx[kx + tau] * Eky[ky + nu];
#pragma unroll
                            for (int phi = 0; phi <= cz + dz; ++phi)
                                h = fma(exy * Ekz[kz + phi], R[hidx(t + tau, u + nu, v + phi)], h);
                        }
                    }
                    H[hidx(t, u, v)] = h;
                }
            }
        }
#pragma unroll
        for (int ab = 0; ab < S::NAB; ++ab) {
            const int a = ab / S::NB, b = ab % S::NB;
            const int ax = comp_lx(LA, a), ay = comp_ly(LA, a), az = comp_lz(LA, a);
            const int bx = comp_lx(LB, b), by = comp_ly(LB, b), bz = comp_lz(LB, b);
            const int ix = (ax * (LB + 1) + bx) * S::NTB;
            const int iy = (ay * (LB + 1) + by) * S::NTB;
            const int iz = (az * (LB + 1) + bz) * S::NTB;
            double s = 0.0;
#pragma unroll
            for (int t = 0; t <= ax + bx; ++t) {
#pragma unroll
                for (int u = 0; u <= ay + by; ++u) {
                    const double exy = Ebx[ix + t] * Eby[iy + u];
#pragma unroll
                    for (int v = 0; v <= az + bz; ++v)
                        s = fma(exy * Ebz[iz + v], H[hidx(t, u, v)], s);
                }
            }
            acc[ab * NCH + (cd - C0)] += s;
        }
    }
}

__device__ __forceinline__ double dmax6(double a, double b, double c, double d, double e, double f)
{
    return fmax(fmax(fmax(a, b), fmax(c, d)), fmax(e, f));
}

// ---------------------------------------------------------------------------
// One (quartet, ket-chunk) work item.
// ---------------------------------------------------------------------------
template <int LA, int LB, int LC, int LD, int C0, int C1>
__device__ __forceinline__ void quartet_item(const ClassArgs& A, const double* lds, int2 qd)
{
    using S = ClassShape<LA, LB, LC, LD>;
    constexpr int NCH = C1 - C0;
    constexpr int L = S::L;
    const ShellPair bra = A.pairs[qd.x];
    const ShellPair ket = A.pairs[qd.y];

    double acc[S::NAB * NCH];
#pragma unroll
    for (int k = 0; k < S::NAB * NCH; ++k) acc[k] = 0.0;

    for (int ip = 0; ip < bra.npp; ++ip) {
        const PrimPair P = A.pp[bra.pp0 + ip];
        double Ebx[(LA + 1) * (LB + 1) * S::NTB], Eby[(LA + 1) * (LB + 1) * S::NTB],
            Ebz[(LA + 1) * (LB + 1) * S::NTB];
        hermite_E<LA, LB, 1>(P.inv2p, P.PAx, P.PBx, Ebx);
        hermite_E<LA, LB, 1>(P.inv2p, P.PAy, P.PBy, Eby);
        hermite_E<LA, LB, 1>(P.inv2p, P.PAz, P.PBz, Ebz);
        for (int jq = 0; jq < ket.npp; ++jq) {
            const PrimPair Q = A.pp[ket.pp0 + jq];
            double Ekx[(LC + 1) * (LD + 1) * S::NTK], Eky[(LC + 1) * (LD + 1) * S::NTK],
                Ekz[(LC + 1) * (LD + 1) * S::NTK];
            hermite_E<LC, LD, -1>(Q.inv2p, Q.PAx, Q.PBx, Ekx);
            hermite_E<LC, LD, -1>(Q.inv2p, Q.PAy, Q.PBy, Eky);
            hermite_E<LC, LD, -1>(Q.inv2p, Q.PAz, Q.PBz, Ekz);
            const double pq = P.p * Q.p;
            const double ppq = P.p + Q.p;
            const double alpha = pq / ppq;
            const double X = P.Px - Q.Px, Y = P.Py - Q.Py, Z = P.Pz - Q.Pz;
            const double T = alpha * (X * X + Y * Y + Z * Z);
            const double pref = TWO_PI_2_5 / (pq * sqrt(ppq)) * P.K * Q.K;
            double F[L + 1];
            boys_seeds<L>(T, alpha, pref, lds, F);
            double R[nherm(L)];
            hermite_R<L>(X, Y, Z, F, R);
            contract<LA, LB, LC, LD, C0, C1>(R, Ebx, Eby, Ebz, Ekx, Eky, Ekz, acc);
        }
    }

    const int N = A.N;
    const int bfa = bra.bfa, bfb = bra.bfb, bfc = ket.bfa, bfd = ket.bfb;
    double ga[S::NA], gb[S::NB], gc[S::NC], gd[S::ND];
#pragma unroll
    for (int i = 0; i < S::NA; ++i) ga[i] = A.gcomp[bfa + i];
#pragma unroll
    for (int i = 0; i < S::NB; ++i) gb[i] = A.gcomp[bfb + i];
#pragma unroll
    for (int i = 0; i < S::NC; ++i) gc[i] = A.gcomp[bfc + i];
#pragma unroll
    for (int i = 0; i < S::ND; ++i) gd[i] = A.gcomp[bfd + i];

    {
        // Shell-level degeneracy with every component combination (the
        // symmetrised result equals formPT's symmetrised result for a real
        // symmetric density difference) and the basis-function-level
        // Schwarz x |dP| test of cython/fock.pyx:46-57.
        const double deg = (bra.same ? 1.0 : 2.0) * (ket.same ? 1.0 : 2.0) * (qd.x == qd.y ? 1.0 : 2.0);
        const double* aD = A.absD;
        const double* sq = A.sqS;
#pragma unroll
        for (int ab = 0; ab < S::NAB; ++ab) {
            const int a = ab / S::NB, b = ab % S::NB;
            const int p = bfa + a, q = bfb + b;
            const double sab = sq[(size_t)p * N + q];
            const double dab = 4.0 * aD[(size_t)p * N + q];
#pragma unroll
            for (int cd = C0; cd < C1; ++cd) {
                const int c = cd / S::ND, d = cd % S::ND;
                const int r = bfc + c, s = bfd + d;
                const double bound = sab * sq[(size_t)r * N + s]
                    * dmax6(dab, 4.0 * aD[(size_t)r * N + s], aD[(size_t)p * N + r], aD[(size_t)p * N + s],
                            aD[(size_t)q * N + r], aD[(size_t)q * N + s]);
                const double v = acc[ab * NCH + (cd - C0)] * deg * ga[a] * gb[b] * gc[c] * gd[d];
                acc[ab * NCH + (cd - C0)] = (bound < A.tol) ? 0.0 : v;
            }
        }
        const double* D = A.D;
        double* G = A.G;
        // Coulomb: G_ab += sum_cd (ab|cd) D_cd
#pragma unroll
        for (int ab = 0; ab < S::NAB; ++ab) {
            const int a = ab / S::NB, b = ab % S::NB;
            double s = 0.0;
#pragma unroll
            for (int cd = C0; cd < C1; ++cd) {
                const int c = cd / S::ND, d = cd % S::ND;
                s = fma(acc[ab * NCH + (cd - C0)], D[(size_t)(bfc + c) * N + bfd + d], s);
            }
            atomicAdd(&G[(size_t)(bfa + a) * N + bfb + b], s);
        }
        // Coulomb: G_cd += sum_ab (ab|cd) D_ab
#pragma unroll
        for (int cd = C0; cd < C1; ++cd) {
            const int c = cd / S::ND, d = cd % S::ND;
            double s = 0.0;
#pragma unroll
            for (int ab = 0; ab < S::NAB; ++ab) {
                const int a = ab / S::NB, b = ab % S::NB;
                s = fma(acc[ab * NCH + (cd - C0)], D[(size_t)(bfa + a) * N + bfb + b], s);
            }
            atomicAdd(&G[(size_t)(bfc + c) * N + bfd + d], s);
        }
        // Exchange: G_bd -= 1/4 (ab|cd) D_ac ; G_bc -= 1/4 (ab|cd) D_ad ;
        //           G_ac -= 1/4 (ab|cd) D_bd ; G_ad -= 1/4 (ab|cd) D_bc
#pragma unroll
        for (int x = 0; x < S::NB; ++x) {
#pragma unroll
            for (int y = 0; y < S::ND; ++y) {     // (b,d) block
                double s = 0.0;
                bool any = false;
#pragma unroll
                for (int a = 0; a < S::NA; ++a)
#pragma unroll
                    for (int c = 0; c < S::NC; ++c) {
                        const int cd = c * S::ND + y;
                        if (cd >= C0 && cd < C1) {
                            any = true;
                            s = fma(acc[(a * S::NB + x) * NCH + (cd - C0)], D[(size_t)(bfa + a) * N + bfc + c], s);
                        }
                    }
                if (any) atomicAdd(&G[(size_t)(bfb + x) * N + bfd + y], -0.25 * s);
            }
#pragma unroll
            for (int y = 0; y < S::NC; ++y) {     // (b,c) block
                double s = 0.0;
                bool any = false;
#pragma unroll
                for (int a = 0; a < S::NA; ++a)
#pragma unroll
                    for (int d = 0; d < S::ND; ++d) {
                        const int cd = y * S::ND + d;
                        if (cd >= C0 && cd < C1) {
                            any = true;
                            s = fma(acc[(a * S::NB + x) * NCH + (cd - C0)], D[(size_t)(bfa + a) * N + bfd + d], s);
                        }
                    }
                if (any) atomicAdd(&G[(size_t)(bfb + x) * N + bfc + y], -0.25 * s);
            }
        }
#pragma unroll
        for (int x = 0; x < S::NA; ++x) {
#pragma unroll
            for (int y = 0; y < S::NC; ++y) {     // (a,c) block
                double s = 0.0;
                bool any = false;
#pragma unroll
                for (int b = 0; b < S::NB; ++b)
#pragma unroll
                    for (int d = 0; d < S::ND; ++d) {
                        const int cd = y * S::ND + d;
                        if (cd >= C0 && cd < C1) {
                            any = true;
                            s = fma(acc[(x * S::NB + b) * NCH + (cd - C0)], D[(size_t)(bfb + b) * N + bfd + d], s);
                        }
                    }
                if (any) atomicAdd(&G[(size_t)(bfa + x) * N + bfc + y], -0.25 * s);
            }
#pragma unroll
            for (int y = 0; y < S::ND; ++y) {     // (a,d) block
                double s = 0.0;
                bool any = false;
#pragma unroll
                for (int b = 0; b < S::NB; ++b)
#pragma unroll
                    for (int c = 0; c < S::NC; ++c) {
                        const int cd = c * S::ND + y;
                        if (cd >= C0 && cd < C1) {
                            any = true;
                            s = fma(acc[(x * S::NB + b) * NCH + (cd - C0)], D[(size_t)(bfb + b) * N + bfc + c], s);
                        }
                    }
                if (any) atomicAdd(&G[(size_t)(bfa + x) * N + bfd + y], -0.25 * s);
            }
        }
    }
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
// Class kernel: persistent grid-stride loop over (64-quartet block, chunk).
// ---------------------------------------------------------------------------
template <int LA, int LB, int LC, int LD>
__global__ void __launch_bounds__(256) eri_fock_class_kernel(ClassArgs A)
{
    using S = ClassShape<LA, LB, LC, LD>;
    __shared__ double lds[BOYS_TAYLOR * BOYS_NT];
    boys_stage_lds<S::L>(lds, A.boys);
    __syncthreads();
    const int nq = *A.count;
    const int lane = threadIdx.x & 63;
    const long nblk = ((long)nq + 63) / 64;
    const long total = nblk * S::NCHUNK;
    const long wave0 = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long nwaves = ((long)gridDim.x * blockDim.x) >> 6;
    for (long w = wave0; w < total; w += nwaves) {
        const int chunk = (int)(w % S::NCHUNK);
        const long q = (w / S::NCHUNK) * 64 + lane;
        if (q >= nq) continue;
        const int2 qd = A.quart[q];
        static_switch<0, S::NCHUNK>(chunk, [&](auto ic) {
            constexpr int C = decltype(ic)::value;
            quartet_item<LA, LB, LC, LD, C * S::CH, (C + 1) * S::CH>(A, lds, qd);
        });
    }
}

}  // namespace mmd
