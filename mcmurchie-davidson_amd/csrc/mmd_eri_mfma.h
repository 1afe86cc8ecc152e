#pragma once
// Fock-build kernel for the high angular-momentum shell-quartet classes on
// the gfx950 matrix cores: one wavefront per surviving shell quartet, the
// two dense Hermite contractions of the McMurchie-Davidson integral
// (cython/twoe.pyx:77-93) as 16 x 16 x 4 MFMA products.
//
// Per primitive quartet (bra pair i, ket pair j), with R_ij the Hermite
// Coulomb integrals R_{t+tau, u+nu, v+phi} (cython/util.pxi:33-50):
//   step A   W_i[tuv][cd] += sum_{tau nu phi} R_ij[tuv + tau nu phi] Ek_j[tau nu phi][cd]
//            (A = the Hankel-structured R matrix, HAB x HCD, read from the
//             R values in LDS; B = the signed ket expansion
//             (-1)^{tau+nu+phi} E^{cd}_{tau nu phi}, HCD x NCD, built from the
//             1-D E tables)
// and once per bra primitive pair, after all its ket pairs,
//   step B   I[ab][cd] += sum_tuv Eb_i[ab][tuv] W_i[tuv][cd]
//            (A = the bra expansion E^{ab}_{tuv}, NAB x HAB; B = W).
// I is the contracted integral block (ab|cd) of the quartet; it is digested
// into the Fock matrix with the reference's screening test, degeneracy and
// Coulomb / exchange weights (cython/fock.pyx:46-85), summed per target
// element in LDS so that each element of the six target blocks takes one
// atomic addition.
//
// Lane layout of a 16 x 16 x 4 MFMA (wave64): A[i][k] in lane i + 16 k,
// B[k][j] in lane j + 16 k, D[i][j] in lane j + 16 g, element r, with
// i = g + 4 r for FP64 (measured on gfx950) and i = 4 g + r for FP32.  With
// the FP64 layout the accumulator of step A already holds step B's B operand
// in the right lanes (k-step ks of step B is element ks % 4 of W's row block
// ks / 4); the FP32 layout is gathered from four lanes with shuffles.
//
// Boys function values are FP64 (table + Taylor, as the other kernels) in
// both arithmetic types; T is the type of R, the E tables and the products.
#include <hip/hip_runtime.h>
#include "mmd_device.h"
#include "mmd_types.h"

namespace mmd {

template <typename T>
struct MfmaOp;

template <>
struct MfmaOp<double> {
    typedef double acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
    static constexpr bool interleaved = true;
};

template <>
struct MfmaOp<float> {
    typedef float acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
    static constexpr bool interleaved = false;
};

// Primitive pairs of the bra (ket) staged in LDS at a time: eight, or four
// for the pair classes with the largest E tables.
__host__ __device__ constexpr int mfma_cap(int la, int lb) { return 3 * (la + 1) * (lb + 1) * (la + lb + 1) <= 60 ? 8 : 4; }

// LDS bytes of one wavefront of a class: staged E tables and scalars of the
// bra and ket primitive pairs, the Boys seeds and PQ vectors of one bra
// primitive against the staged ket primitives, the R double buffer and the
// digested integral block.
template <int LA, int LB, int LC, int LD, typename T>
__host__ __device__ constexpr int mfma_wave_lds()
{
    constexpr int L = LA + LB + LC + LD;
    constexpr int EB = (LA + 1) * (LB + 1) * (LA + LB + 1), EK = (LC + 1) * (LD + 1) * (LC + LD + 1);
    constexpr int CB = mfma_cap(LA, LB), CK = mfma_cap(LC, LD);
    return (int)sizeof(T) * (CB * 3 * EB + CK * 3 * EK + CK * (L + 1) + 2 * nherm(L))
         + (int)sizeof(double) * (CB * 8 + CK * 8 + CK * 4 + ncart(LA) * ncart(LB) * ncart(LC) * ncart(LD));
}

// Wavefronts (independent quartets) per workgroup: the most of 4, 2, 1 for
// which the class's LDS (operand descriptors, R recursion table and the
// per-wave buffers) fits in 64 KiB.
template <int LA, int LB, int LC, int LD, typename T>
__host__ __device__ constexpr int mfma_waves()
{
    constexpr int LAB = LA + LB, LCD = LC + LD;
    constexpr int NAB = ncart(LA) * ncart(LB), NCD = ncart(LC) * ncart(LD);
    constexpr int HAB = nherm(LAB), HCD = nherm(LCD);
    constexpr int ndesc = ((HAB + 15) / 16) * ((HCD + 3) / 4) + ((HCD + 3) / 4) * ((NCD + 15) / 16)
                        + ((NAB + 15) / 16) * ((HAB + 3) / 4);
    constexpr int fixed = 4 * 64 * ndesc + 4 * nherm(LA + LB + LC + LD);
    constexpr int per = mfma_wave_lds<LA, LB, LC, LD, T>();
    return fixed + 4 * per <= 65536 ? 4 : (fixed + 2 * per <= 65536 ? 2 : 1);
}

__device__ __forceinline__ void mfma_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int LA, int LB, int LC, int LD, typename T>
__global__ void __launch_bounds__(256, 2) fock_mfma_class_kernel(ClassArgs A)
{
    constexpr int MFMA_WAVES = mfma_waves<LA, LB, LC, LD, T>();
    using Op = MfmaOp<T>;
    using acc_t = typename Op::acc_t;
    constexpr int L = LA + LB + LC + LD, LAB = LA + LB, LCD = LC + LD;
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    constexpr int NAB = NA * NB, NCD = NC * ND;
    constexpr int HAB = nherm(LAB), HCD = nherm(LCD), HL = nherm(L);
    constexpr int MTA = (HAB + 15) / 16, NT = (NCD + 15) / 16, KSA = (HCD + 3) / 4;
    constexpr int MB = (NAB + 15) / 16, KSB = (HAB + 3) / 4;
    constexpr int NTB = LAB + 1, NTK = LCD + 1;
    constexpr int EB = (LA + 1) * (LB + 1) * NTB, EK = (LC + 1) * (LD + 1) * NTK;
    static_assert(3 * EB < 256 && 3 * EK < 256, "E offsets are packed in 8 bits");
    constexpr int CB = mfma_cap(LA, LB), CK = mfma_cap(LC, LD);
    __shared__ T Rb[MFMA_WAVES][2][HL];
    __shared__ T Ebs[MFMA_WAVES][CB][3 * EB];       // staged bra E tables
    __shared__ T Eks[MFMA_WAVES][CK][3 * EK];       // staged signed ket E tables
    __shared__ T Fs[MFMA_WAVES][CK][L + 1];         // seeds (-2 alpha)^m F_m of one bra primitive vs the kets
    __shared__ double Pbs[MFMA_WAVES][CB][8];       // bra p, 2 pi^{5/2} K / p (x 1/p for the ket), P
    __shared__ double Pks[MFMA_WAVES][CK][8];
    __shared__ double PQs[MFMA_WAVES][CK][4];       // P - Q of the current bra primitive and each ket
    __shared__ double Ib[MFMA_WAVES][NAB * NCD];    // the digested integral block of each wave
    __shared__ int rtab[HL];                        // R recursion: source terms of each Hermite index
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int lr = lane & 15, lg = lane >> 4;

    // Operand descriptors, fixed for the class, per (block, k-step, lane) in
    // LDS (shared by the workgroup's wavefronts, off the register file): the
    // R index of step A's A operand, and the packed offsets of the three 1-D
    // E factors of step A's B operand (ket) and step B's A operand (bra);
    // -1 marks the zero padding outside the matrices and the structural
    // zeros of E.
    constexpr int NDA = MTA * KSA, NDK = KSA * NT, NDB = MB * KSB;
    __shared__ int sdA[NDA * 64], sdK[NDK * 64], sdB[NDB * 64];
    for (int e = threadIdx.x; e < (NDA + NDK + NDB) * 64; e += blockDim.x) {
        const int l = e & 63, slot = e >> 6;
        const int elr = l & 15, elg = l >> 4;
        int d = -1;
        if (slot < NDA) {
            const int mt = slot / KSA, ks = slot % KSA;
            const int tuv = 16 * mt + elr, k = 4 * ks + elg;
            if (tuv < HAB && k < HCD) {
                const Herm3 h1 = hdecode(tuv), h2 = hdecode(k);
                d = hidx(h1.t + h2.t, h1.u + h2.u, h1.v + h2.v);
            }
            sdA[slot * 64 + l] = d;
        } else if (slot < NDA + NDK) {
            const int sl = slot - NDA, ks = sl / NT, nt = sl % NT;
            const int k = 4 * ks + elg, cd = 16 * nt + elr;
            if (k < HCD && cd < NCD) {
                const Herm3 h = hdecode(k);
                const int c = cd / ND, dd = cd % ND;
                const int cx = comp_lx(LC, c), cy = comp_ly(LC, c), cz = comp_lz(LC, c);
                const int dx = comp_lx(LD, dd), dy = comp_ly(LD, dd), dz = comp_lz(LD, dd);
                if (h.t <= cx + dx && h.u <= cy + dy && h.v <= cz + dz)
                    d = ((cx * (LD + 1) + dx) * NTK + h.t) | ((EK + (cy * (LD + 1) + dy) * NTK + h.u) << 8)
                        | ((2 * EK + (cz * (LD + 1) + dz) * NTK + h.v) << 16);
            }
            sdK[sl * 64 + l] = d;
        } else {
            const int sl = slot - NDA - NDK, mb = sl / KSB, ks = sl % KSB;
            const int ab = 16 * mb + elr, k = 4 * ks + elg;
            if (ab < NAB && k < HAB) {
                const Herm3 h = hdecode(k);
                const int a = ab / NB, b = ab % NB;
                const int ax = comp_lx(LA, a), ay = comp_ly(LA, a), az = comp_lz(LA, a);
                const int bx = comp_lx(LB, b), by = comp_ly(LB, b), bz = comp_lz(LB, b);
                if (h.t <= ax + bx && h.u <= ay + by && h.v <= az + bz)
                    d = ((ax * (LB + 1) + bx) * NTB + h.t) | ((EB + (ay * (LB + 1) + by) * NTB + h.u) << 8)
                        | ((2 * EB + (az * (LB + 1) + bz) * NTB + h.v) << 16);
            }
            sdB[sl * 64 + l] = d;
        }
    }
    // R_tuv = X_dir R_{prev}(i1) + c R_{prev}(i2) for the first nonzero of
    // (t, u, v) (cython/util.pxi:38-49): dir | i1 << 2 | i2 << 10 | c << 18,
    // i2 = 255 when the second term is absent
    for (int k = threadIdx.x; k < HL; k += blockDim.x) {
        const Herm3 h = hdecode(k);
        int dir = 0, i1 = 0, i2 = 255, c = 0;
        if (h.t > 0) {
            dir = 0; i1 = hidx(h.t - 1, h.u, h.v);
            if (h.t > 1) { i2 = hidx(h.t - 2, h.u, h.v); c = h.t - 1; }
        } else if (h.u > 0) {
            dir = 1; i1 = hidx(h.t, h.u - 1, h.v);
            if (h.u > 1) { i2 = hidx(h.t, h.u - 2, h.v); c = h.u - 1; }
        } else if (h.v > 0) {
            dir = 2; i1 = hidx(h.t, h.u, h.v - 1);
            if (h.v > 1) { i2 = hidx(h.t, h.u, h.v - 2); c = h.v - 1; }
        }
        rtab[k] = dir | (i1 << 2) | (i2 << 10) | (c << 18);
    }
    __syncthreads();
    const int nq = *A.count;
    const int N = A.N;
    for (int64_t q = (int64_t)blockIdx.x * MFMA_WAVES + wv; q < nq; q += (int64_t)gridDim.x * MFMA_WAVES) {
        const int2 qd = A.quart[q];
        const ShellPair bra = A.pairs[qd.x];
        const ShellPair ket = A.pairs[qd.y];
        acc_t I[MB][NT];
        sfor<MB>([&](auto i) { sfor<NT>([&](auto j) { I[i][j] = acc_t{T(0), T(0), T(0), T(0)}; }); });
        for (int ib0 = 0; ib0 < bra.npp; ib0 += CB) {
            const int nb = min(CB, bra.npp - ib0);
            // stage the bra primitive pairs of this chunk: E tables
            // (cython/util.pxi:13-28), one lane per (pair, direction)
            if (lane < 3 * nb) {
                const int ip = lane / 3, x = lane - 3 * ip;
                const PrimPair* P = A.pp + bra.pp0 + ib0 + ip;
                T E[EB];
                hermite_E<LA, LB, 1, T>((T)P->inv2p, (T)(x == 0 ? P->PAx : (x == 1 ? P->PAy : P->PAz)),
                                        (T)(x == 0 ? P->PBx : (x == 1 ? P->PBy : P->PBz)), E);
                sfor<EB>([&](auto k) { Ebs[wv][ip][x * EB + decltype(k)::value] = E[decltype(k)::value]; });
                if (x == 0) {
                    Pbs[wv][ip][0] = P->p;
                    Pbs[wv][ip][1] = 34.98683665524972497 * P->K * P->invp;   // 2 pi^{5/2} K_ab / p
                    Pbs[wv][ip][2] = P->Px;
                    Pbs[wv][ip][3] = P->Py;
                    Pbs[wv][ip][4] = P->Pz;
                }
            }
            for (int jk0 = 0; jk0 < ket.npp; jk0 += CK) {
                const int nk = min(CK, ket.npp - jk0);
                if (lane >= 32 && lane < 32 + 3 * nk) {   // signed ket tables (-1)^tau E^{cd}_tau
                    const int jq = (lane - 32) / 3, x = lane - 32 - 3 * jq;
                    const PrimPair* Q = A.pp + ket.pp0 + jk0 + jq;
                    T E[EK];
                    if constexpr (LC + LD > 0) {
                        hermite_E<LC, LD, -1, T>((T)Q->inv2p, (T)(x == 0 ? Q->PAx : (x == 1 ? Q->PAy : Q->PAz)),
                                                 (T)(x == 0 ? Q->PBx : (x == 1 ? Q->PBy : Q->PBz)), E);
                    } else {
                        E[0] = T(1);
                    }
                    sfor<EK>([&](auto k) { Eks[wv][jq][x * EK + decltype(k)::value] = E[decltype(k)::value]; });
                    if (x == 0) {
                        Pks[wv][jq][0] = Q->p;
                        Pks[wv][jq][1] = Q->K * Q->invp;
                        Pks[wv][jq][2] = Q->Px;
                        Pks[wv][jq][3] = Q->Py;
                        Pks[wv][jq][4] = Q->Pz;
                    }
                }
                mfma_wave_sync();
                for (int ip = 0; ip < nb; ++ip) {
                    // Boys seeds of this bra primitive against every staged
                    // ket primitive, one lane per ket primitive
                    if (lane < nk) {
                        const double* pb = Pbs[wv][ip];
                        const double* pk = Pks[wv][lane];
                        const double rs = rsqrt(pb[0] + pk[0]);
                        const double alpha = pb[0] * pk[0] * rs * rs;
                        const double X = pb[2] - pk[2], Y = pb[3] - pk[3], Z = pb[4] - pk[4];
                        double F[MAX_LTOT + 1];
                        boys_runtime(L, alpha * (X * X + Y * Y + Z * Z), A.boys, F);
                        double sc = pb[1] * pk[1] * rs;
                        sfor<L + 1>([&](auto m) {
                            Fs[wv][lane][decltype(m)::value] = (T)(sc * F[decltype(m)::value]);
                            sc *= -2.0 * alpha;
                        });
                        PQs[wv][lane][0] = X;
                        PQs[wv][lane][1] = Y;
                        PQs[wv][lane][2] = Z;
                    }
                    mfma_wave_sync();
                    acc_t W[MTA][NT];
                    sfor<MTA>([&](auto i) { sfor<NT>([&](auto j) { W[i][j] = acc_t{T(0), T(0), T(0), T(0)}; }); });
                    for (int jq = 0; jq < nk; ++jq) {
                        // R_tuv (t+u+v <= L) level by level in LDS (cython/util.pxi:33-50)
                        const T xyz[3] = {(T)PQs[wv][jq][0], (T)PQs[wv][jq][1], (T)PQs[wv][jq][2]};
                        const T* fs = Fs[wv][jq];
                        if (lane == 0) Rb[wv][L & 1][0] = fs[L];
                        mfma_wave_sync();
                        for (int n = L - 1; n >= 0; --n) {
                            const T* prev = Rb[wv][(n + 1) & 1];
                            T* cur = Rb[wv][n & 1];
                            for (int k = lane; k < nherm(L - n); k += 64) {
                                T val;
                                if (k == 0) {
                                    val = fs[n];
                                } else {
                                    const int e = rtab[k];
                                    const int i2 = (e >> 10) & 255;
                                    val = xyz[e & 3] * prev[(e >> 2) & 255];
                                    if (i2 != 255) val += (T)(e >> 18) * prev[i2];
                                }
                                cur[k] = val;
                            }
                            mfma_wave_sync();
                        }
                        const T* R = Rb[wv][0];
                        const T* Ek = Eks[wv][jq];
                        // step A: W += R_ij Ek_j
                        sfor<KSA>([&](auto iks) {
                            constexpr int ks = decltype(iks)::value;
                            T b[NT];
                            sfor<NT>([&](auto int_) {
                                const int d = sdK[(ks * NT + decltype(int_)::value) * 64 + lane];
                                b[decltype(int_)::value] =
                                    d < 0 ? T(0) : Ek[d & 255] * Ek[(d >> 8) & 255] * Ek[d >> 16];
                            });
                            sfor<MTA>([&](auto imt) {
                                const int d = sdA[(decltype(imt)::value * KSA + ks) * 64 + lane];
                                const T a = d < 0 ? T(0) : R[d];
                                sfor<NT>([&](auto int_) {
                                    W[decltype(imt)::value][decltype(int_)::value] = Op::mma(
                                        a, b[decltype(int_)::value], W[decltype(imt)::value][decltype(int_)::value]);
                                });
                            });
                        });
                        mfma_wave_sync();   // the next ket primitive rewrites Rb
                    }
                    // step B: I += Eb_i W_i
                    const T* Eb = Ebs[wv][ip];
                    sfor<KSB>([&](auto iks) {
                        constexpr int ks = decltype(iks)::value;
                        T wop[NT];
                        sfor<NT>([&](auto int_) {
                            const acc_t& w = W[ks / 4][decltype(int_)::value];
                            if constexpr (Op::interleaved) {
                                wop[decltype(int_)::value] = w[ks % 4];
                            } else {
                                const int src = 16 * (ks % 4) + lr;
                                const T v0 = __shfl(w[0], src), v1 = __shfl(w[1], src);
                                const T v2 = __shfl(w[2], src), v3 = __shfl(w[3], src);
                                wop[decltype(int_)::value] = lg == 0 ? v0 : (lg == 1 ? v1 : (lg == 2 ? v2 : v3));
                            }
                        });
                        sfor<MB>([&](auto imb) {
                            const int d = sdB[(decltype(imb)::value * KSB + ks) * 64 + lane];
                            const T a = d < 0 ? T(0) : Eb[d & 255] * Eb[(d >> 8) & 255] * Eb[d >> 16];
                            sfor<NT>([&](auto int_) {
                                I[decltype(imb)::value][decltype(int_)::value] = Op::mma(
                                    a, wop[decltype(int_)::value], I[decltype(imb)::value][decltype(int_)::value]);
                            });
                        });
                    });
                }
                mfma_wave_sync();   // the next ket chunk rewrites the staged ket data
            }
            mfma_wave_sync();   // the next bra chunk rewrites the staged bra data
        }
        // digestion (cython/fock.pyx:46-85): each integral, scaled by its
        // component norms and the shell degeneracy and zeroed when the
        // basis-function-level Schwarz x |dP| test drops it, goes to the
        // wave's LDS block; the six target blocks (Coulomb J_ab, J_cd and
        // the four exchange blocks) are then summed there, so that each
        // target element takes one atomic addition into G
        const int bfa = bra.bfa, bfb = bra.bfb, bfc = ket.bfa, bfd = ket.bfb;
        const double deg = (bra.same ? 1.0 : 2.0) * (ket.same ? 1.0 : 2.0) * (qd.x == qd.y ? 1.0 : 2.0);
        const double* aD = A.absD;
        const double* D = A.D;
        double* G = A.G;
        double* Iw = Ib[wv];
        sfor<MB>([&](auto imb) {
            sfor<NT>([&](auto int_) {
                sfor<4>([&](auto ir) {
                    const int ab = 16 * decltype(imb)::value + Op::row(lane, decltype(ir)::value);
                    const int cd = 16 * decltype(int_)::value + lr;
                    if (ab < NAB && cd < NCD) {
                        const int p = bfa + ab / NB, qq = bfb + ab % NB, r = bfc + cd / ND, s = bfd + cd % ND;
                        const double bound = A.sqS[(size_t)p * N + qq] * A.sqS[(size_t)r * N + s]
                            * fmax(fmax(fmax(4.0 * aD[(size_t)p * N + qq], 4.0 * aD[(size_t)r * N + s]),
                                        fmax(aD[(size_t)p * N + r], aD[(size_t)p * N + s])),
                                   fmax(aD[(size_t)qq * N + r], aD[(size_t)qq * N + s]));
                        const double v = (double)I[decltype(imb)::value][decltype(int_)::value][decltype(ir)::value]
                            * deg * A.gcomp[p] * A.gcomp[qq] * A.gcomp[r] * A.gcomp[s];
                        Iw[ab * NCD + cd] = bound >= A.tol ? v : 0.0;
                    }
                });
            });
        });
        mfma_wave_sync();
        constexpr int T_JAB = NAB, T_JCD = NCD, T_BD = NB * ND, T_BC = NB * NC, T_AC = NA * NC, T_AD = NA * ND;
        constexpr int NTARGET = T_JAB + T_JCD + T_BD + T_BC + T_AC + T_AD;
        for (int t = lane; t < NTARGET; t += 64) {
            double sum = 0.0;
            size_t dst;
            if (t < T_JAB) {                              // G_ab += sum_cd (ab|cd) D_cd
                const int ab = t, a = ab / NB, b = ab % NB;
                for (int cd = 0; cd < NCD; ++cd)
                    sum = fma(Iw[ab * NCD + cd], D[(size_t)(bfc + cd / ND) * N + bfd + cd % ND], sum);
                dst = (size_t)(bfa + a) * N + bfb + b;
            } else if (t < T_JAB + T_JCD) {               // G_cd += sum_ab (ab|cd) D_ab
                const int cd = t - T_JAB, c = cd / ND, d = cd % ND;
                for (int ab = 0; ab < NAB; ++ab)
                    sum = fma(Iw[ab * NCD + cd], D[(size_t)(bfa + ab / NB) * N + bfb + ab % NB], sum);
                dst = (size_t)(bfc + c) * N + bfd + d;
            } else if (t < T_JAB + T_JCD + T_BD) {        // G_bd -= 1/4 sum_ac (ab|cd) D_ac
                const int i = t - T_JAB - T_JCD, b = i / ND, d = i % ND;
                for (int a = 0; a < NA; ++a)
                    for (int c = 0; c < NC; ++c)
                        sum = fma(Iw[(a * NB + b) * NCD + c * ND + d], D[(size_t)(bfa + a) * N + bfc + c], sum);
                sum *= -0.25;
                dst = (size_t)(bfb + b) * N + bfd + d;
            } else if (t < T_JAB + T_JCD + T_BD + T_BC) { // G_bc -= 1/4 sum_ad (ab|cd) D_ad
                const int i = t - T_JAB - T_JCD - T_BD, b = i / NC, c = i % NC;
                for (int a = 0; a < NA; ++a)
                    for (int d = 0; d < ND; ++d)
                        sum = fma(Iw[(a * NB + b) * NCD + c * ND + d], D[(size_t)(bfa + a) * N + bfd + d], sum);
                sum *= -0.25;
                dst = (size_t)(bfb + b) * N + bfc + c;
            } else if (t < T_JAB + T_JCD + T_BD + T_BC + T_AC) {   // G_ac -= 1/4 sum_bd (ab|cd) D_bd
                const int i = t - T_JAB - T_JCD - T_BD - T_BC, a = i / NC, c = i % NC;
                for (int b = 0; b < NB; ++b)
                    for (int d = 0; d < ND; ++d)
                        sum = fma(Iw[(a * NB + b) * NCD + c * ND + d], D[(size_t)(bfb + b) * N + bfd + d], sum);
                sum *= -0.25;
                dst = (size_t)(bfa + a) * N + bfc + c;
            } else {                                      // G_ad -= 1/4 sum_bc (ab|cd) D_bc
                const int i = t - T_JAB - T_JCD - T_BD - T_BC - T_AC, a = i / ND, d = i % ND;
                for (int b = 0; b < NB; ++b)
                    for (int c = 0; c < NC; ++c)
                        sum = fma(Iw[(a * NB + b) * NCD + c * ND + d], D[(size_t)(bfb + b) * N + bfc + c], sum);
                sum *= -0.25;
                dst = (size_t)(bfa + a) * N + bfd + d;
            }
            if (sum != 0.0) atomicAdd(&G[dst], sum);
        }
        mfma_wave_sync();   // Iw is rewritten by the wave's next quartet
    }
}

}  // namespace mmd
