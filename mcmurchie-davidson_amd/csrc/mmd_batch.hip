// Device side of the batched RHF driver (mmd_batch_rhf): many molecules
// stepped together, one workgroup per molecule.
//
// The molecules of a batch form one block-diagonal basis: molecule m owns
// basis functions [f0, f0 + n) and the n x n diagonal block of every N x N
// matrix (S, H, X, F, P, P_old, dP, G; N = all functions of the batch).  The
// Fock build (the class kernels) runs once per iteration for the whole
// batch; these kernels do the per-molecule rest of the reference's loop
// (mmd/scf.py:10-80) in the block: the F update, the DIIS extrapolation
// (mmd/scf.py:146-180), FO = X^T F X, the symmetric eigensolver
// (scipy.linalg.eigh at mmd/scf.py:47, here a cyclic Jacobi method), the
// density, the energy (mmd/scf.py:118-121) and the convergence test on
// ||P - P_old||_F.  A molecule that has converged keeps its state and
// contributes dP = 0 to later Fock builds, so it costs the builds no work.
#include <hip/hip_runtime.h>
#include <math.h>
#include "mmd_batch.h"

namespace mmd {

namespace {

constexpr int BT = 256;   // threads per molecule workgroup

__device__ __forceinline__ double* blk(double* M, int N, int f0) { return M + (size_t)f0 * N + f0; }
__device__ __forceinline__ const double* blk(const double* M, int N, int f0) { return M + (size_t)f0 * N + f0; }

// C (n x n, leading dimension ldc) = op(A) op(B), A and B n x n with leading
// dimensions lda, ldb; the workgroup's threads share the output elements.
__device__ void wg_gemm(int n, bool tA, const double* A, int lda, bool tB, const double* B, int ldb, double* C,
                        int ldc)
{
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int i = e / n, j = e - i * n;
        double s = 0.0;
        for (int k = 0; k < n; ++k) {
            const double a = tA ? A[(size_t)k * lda + i] : A[(size_t)i * lda + k];
            const double b = tB ? B[(size_t)j * ldb + k] : B[(size_t)k * ldb + j];
            s = fma(a, b, s);
        }
        C[(size_t)i * ldc + j] = s;
    }
    __syncthreads();
}

// Workgroup sum of v (all threads get the result); red: BT/64 doubles of LDS.
__device__ double wg_sum(double v, double* red)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wv] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
    __syncthreads();
    return s;
}

typedef double dbl4_t __attribute__((ext_vector_type(4)));

// C (n x n, leading dimension ldc, default n; LDS or global) = op(A) op(B)
// with A, B n x n in LDS (leading dimension n), on the matrix cores: each wavefront takes 16 x 16 output tiles (the matrix
// padded with zeros to a multiple of 16) and runs v_mfma_f64_16x16x4_f64
// over k in steps of 4.  Operand layout (wave64): A[i][k] in lane i + 16k,
// B[k][j] in lane j + 16k; the accumulator holds D[i][j] in lane j + 16g,
// element r, with i = g + 4r.
__device__ void lds_gemm(int n, bool tA, const double* A, bool tB, const double* B, double* C, int ldc = -1)
{
    if (ldc < 0) ldc = n;
    const int nt = (n + 15) >> 4;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nw = (int)(blockDim.x >> 6);
    const int l16 = lane & 15, l4 = lane >> 4;
    for (int tile = wave; tile < nt * nt; tile += nw) {
        const int ti = tile / nt, tj = tile - ti * nt;
        const int i = ti * 16 + l16, j = tj * 16 + l16;
        dbl4_t acc = {0.0, 0.0, 0.0, 0.0};
        for (int k0 = 0; k0 < n; k0 += 4) {
            const int k = k0 + l4;
            const double a = (i < n && k < n) ? (tA ? A[k * n + i] : A[i * n + k]) : 0.0;
            const double b = (k < n && j < n) ? (tB ? B[j * n + k] : B[k * n + j]) : 0.0;
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
        }
        const int jc = tj * 16 + l16;
        for (int r = 0; r < 4; ++r) {
            const int ir = ti * 16 + l4 + 4 * r;
            if (ir < n && jc < n) C[(size_t)ir * ldc + jc] = acc[r];
        }
    }
    __syncthreads();
}

// Symmetric eigendecomposition A = V diag(w) V^T of the n x n matrix in LDS
// (A, leading dimension lda; destroyed) by the cyclic Jacobi method with a
// round-robin ordering: each round applies n/2 disjoint rotations at once.
// V (n x n, leading dimension ldv, global memory) receives the eigenvectors
// as columns and w the eigenvalues, both in ascending order of w (the order
// of LAPACK's drivers).  perm, cs: LDS scratch of n + 1 ints and 2 (n + 1)
// doubles; red: BT/64 doubles.
__device__ void wg_jacobi(int n, double* A, int lda, double* V, int ldv, double* w, int* perm, double* cs,
                          double* red)
{
    const int m = (n + 1) & ~1;   // even: an odd n takes a dummy index n that stays decoupled
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int i = e / n, j = e - i * n;
        V[(size_t)i * ldv + j] = (i == j) ? 1.0 : 0.0;
    }
    double fro = 0.0;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) fro += A[(e / n) * lda + e % n] * A[(e / n) * lda + e % n];
    fro = wg_sum(fro, red);
    double* c = cs;
    double* s = cs + (m / 2);
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0;
        for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
            const int i = e / n, j = e - i * n;
            if (i != j) off += A[i * lda + j] * A[i * lda + j];
        }
        off = wg_sum(off, red);
        if (!(off > 1e-30 * fro) ) break;
        for (int round = 0; round < m - 1; ++round) {
            // round-robin pairs: index 0 fixed, the others rotate
            for (int k = threadIdx.x; k < m / 2; k += blockDim.x) {
                auto slot = [&](int pos) { return pos == 0 ? 0 : 1 + (pos - 1 + round) % (m - 1); };
                int p = slot(k), q = slot(m - 1 - k);
                if (p > q) { const int t = p; p = q; q = t; }
                perm[2 * k] = p;
                perm[2 * k + 1] = q;
                double cc = 1.0, ss = 0.0;
                if (q < n) {
                    const double apq = A[p * lda + q];
                    if (apq != 0.0) {
                        const double app = A[p * lda + p], aqq = A[q * lda + q];
                        const double theta = (aqq - app) / (2.0 * apq);
                        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                        cc = 1.0 / sqrt(t * t + 1.0);
                        ss = t * cc;
                    }
                }
                c[k] = cc;
                s[k] = ss;
            }
            __syncthreads();
            // rows: (A_p., A_q.) <- (c A_p. - s A_q., s A_p. + c A_q.)
            for (int e = threadIdx.x; e < (m / 2) * n; e += blockDim.x) {
                const int k = e / n, j = e - k * n;
                const int p = perm[2 * k], q = perm[2 * k + 1];
                if (q >= n || s[k] == 0.0) continue;
                const double ap = A[p * lda + j], aq = A[q * lda + j];
                A[p * lda + j] = c[k] * ap - s[k] * aq;
                A[q * lda + j] = s[k] * ap + c[k] * aq;
            }
            __syncthreads();
            // columns of A and of V
            for (int e = threadIdx.x; e < (m / 2) * n; e += blockDim.x) {
                const int k = e / n, i = e - k * n;
                const int p = perm[2 * k], q = perm[2 * k + 1];
                if (q >= n || s[k] == 0.0) continue;
                const double ap = A[i * lda + p], aq = A[i * lda + q];
                A[i * lda + p] = c[k] * ap - s[k] * aq;
                A[i * lda + q] = s[k] * ap + c[k] * aq;
                double* vr = V + (size_t)i * ldv;
                const double vp = vr[p], vq = vr[q];
                vr[p] = c[k] * vp - s[k] * vq;
                vr[q] = s[k] * vp + c[k] * vq;
            }
            __syncthreads();
        }
    }
    // ascending order: rank of each eigenvalue (ties by index)
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double di = A[i * lda + i];
        int r = 0;
        for (int j = 0; j < n; ++j) {
            const double dj = A[j * lda + j];
            if (dj < di || (dj == di && j < i)) ++r;
        }
        perm[i] = r;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) w[perm[i]] = A[i * lda + i];
    __syncthreads();
    // permute the columns of V row by row through LDS (A is free now)
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int i = e / n, j = e - i * n;
        A[i * lda + perm[j]] = V[(size_t)i * ldv + j];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int i = e / n, j = e - i * n;
        V[(size_t)i * ldv + j] = A[i * lda + j];
    }
    __syncthreads();
}

constexpr int BTL = 1024;   // threads of the all-LDS molecule workgroups

// LDS vectors of a molecule workgroup (the matrices are dynamic LDS).
struct WgSmall {
    double w[BATCH_MAX_NBF];
    double cs[BATCH_MAX_NBF + 2];
    int perm[BATCH_MAX_NBF + 2];
    double red[BTL / 64];
    double Bm[9 * 9];
    double rhs[9];
};

// Cyclic Jacobi eigensolver with A (destroyed) and V both in LDS (leading
// dimension n).  V holds the starting basis on entry: the identity, or the
// eigenvectors of a nearby matrix, in which case A must be V^T A0 V; on exit
// A0 = V diag(w) V^T with V's columns in ascending order of w.  A round
// applies the m/2 disjoint rotations of the round-robin ordering (m = n
// rounded up to even) in one pass: each thread updates whole 2 x 2 blocks
// (rows of one rotation, columns of another), so a round needs only two
// barriers.  Convergence: sum of squared off-diagonal elements at most
// 1e-30 of the squared Frobenius norm.
__device__ void lds_jacobi(int n, double* A, double* V, double* w, int* perm, double* cs, double* red)
{
    const int m = (n + 1) & ~1;
    const int h = m / 2;
    double* c = cs;
    double* s = cs + h;
    double fro = 0.0;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) fro = fma(A[e], A[e], fro);
    fro = wg_sum(fro, red);
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0;
        for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
            const int i = e / n, j = e - i * n;
            if (i != j) off = fma(A[e], A[e], off);
        }
        off = wg_sum(off, red);
        if (!(off > 1e-30 * fro)) break;
        for (int round = 0; round < m - 1; ++round) {
            for (int k = threadIdx.x; k < h; k += blockDim.x) {
                auto slot = [&](int pos) { return pos == 0 ? 0 : 1 + (pos - 1 + round) % (m - 1); };
                int p = slot(k), q = slot(m - 1 - k);
                if (p > q) { const int t = p; p = q; q = t; }
                perm[2 * k] = p;
                perm[2 * k + 1] = q;
                double cc = 1.0, ss = 0.0;
                if (q < n) {
                    const double apq = A[p * n + q];
                    if (apq != 0.0) {
                        const double app = A[p * n + p], aqq = A[q * n + q];
                        const double theta = (aqq - app) / (2.0 * apq);
                        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                        cc = 1.0 / sqrt(t * t + 1.0);
                        ss = t * cc;
                    }
                }
                c[k] = cc;
                s[k] = ss;
            }
            __syncthreads();
            // A <- J^T A J, block (k1, k2) = rows p1, q1 x columns p2, q2
            for (int b = threadIdx.x; b < h * h; b += blockDim.x) {
                const int k1 = b / h, k2 = b - k1 * h;
                const double s1 = s[k1], s2 = s[k2];
                if (s1 == 0.0 && s2 == 0.0) continue;
                const double c1 = c[k1], c2 = c[k2];
                const int p1 = perm[2 * k1], q1 = perm[2 * k1 + 1], p2 = perm[2 * k2], q2 = perm[2 * k2 + 1];
                const bool r1 = q1 < n, r2 = q2 < n;   // the dummy index n has no row or column
                const double a = A[p1 * n + p2];
                const double bb = r2 ? A[p1 * n + q2] : 0.0;
                const double d = r1 ? A[q1 * n + p2] : 0.0;
                const double e = (r1 && r2) ? A[q1 * n + q2] : 0.0;
                const double a1 = c1 * a - s1 * d, b1 = c1 * bb - s1 * e;
                const double d1 = s1 * a + c1 * d, e1 = s1 * bb + c1 * e;
                A[p1 * n + p2] = c2 * a1 - s2 * b1;
                if (r2) A[p1 * n + q2] = s2 * a1 + c2 * b1;
                if (r1) A[q1 * n + p2] = c2 * d1 - s2 * e1;
                if (r1 && r2) A[q1 * n + q2] = s2 * d1 + c2 * e1;
            }
            // V <- V J
            for (int e = threadIdx.x; e < h * n; e += blockDim.x) {
                const int k = e / n, i = e - k * n;
                const int p = perm[2 * k], q = perm[2 * k + 1];
                if (q >= n || s[k] == 0.0) continue;
                double* vr = V + i * n;
                const double vp = vr[p], vq = vr[q];
                vr[p] = c[k] * vp - s[k] * vq;
                vr[q] = s[k] * vp + c[k] * vq;
            }
            __syncthreads();
        }
    }
    // ascending order: rank of each eigenvalue (ties by index)
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double di = A[i * n + i];
        int r = 0;
        for (int j = 0; j < n; ++j) {
            const double dj = A[j * n + j];
            if (dj < di || (dj == di && j < i)) ++r;
        }
        perm[i] = r;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) w[perm[i]] = A[i * n + i];
    __syncthreads();
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int i = e / n, j = e - i * n;
        A[i * n + perm[j]] = V[e];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) V[e] = A[e];
    __syncthreads();
}

// LDS of a molecule workgroup: one n x n matrix (leading dimension n) plus
// small vectors; n <= BATCH_MAX_NBF.
struct WgLds {
    double A[BATCH_MAX_NBF * BATCH_MAX_NBF];
    double w[BATCH_MAX_NBF];
    double cs[BATCH_MAX_NBF + 2];
    int perm[BATCH_MAX_NBF + 2];
    double red[BT / 64];
    double Bm[9 * 9];
    double rhs[9];
};

// X = S^-1/2 and U = S^1/2 of each molecule's overlap block (mmd/molecule.py:284-286).
__global__ void __launch_bounds__(BT) batch_overlap_kernel(BatchArgs B)
{
    __shared__ WgLds L;
    const BatchMol M = B.mol[blockIdx.x];
    const int n = M.n, N = B.N;
    double* V = B.work + M.wo;   // n x n eigenvectors
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) L.A[e] = blk(B.S, N, M.f0)[(size_t)(e / n) * N + e % n];
    __syncthreads();
    wg_jacobi(n, L.A, n, V, n, L.w, L.perm, L.cs, L.red);
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int i = e / n, j = e - i * n;
        double x = 0.0, u = 0.0;
        for (int k = 0; k < n; ++k) {
            const double vv = V[(size_t)i * n + k] * V[(size_t)j * n + k];
            x = fma(vv, 1.0 / sqrt(L.w[k]), x);
            u = fma(vv, sqrt(L.w[k]), u);
        }
        blk(B.X, N, M.f0)[(size_t)i * N + j] = x;
        blk(B.U, N, M.f0)[(size_t)i * N + j] = u;
    }
    if (threadIdx.x == 0 && !(L.w[0] > 0.0)) B.flag[blockIdx.x] = -1;   // not positive definite
}

// batch_overlap_kernel with S and its eigenvectors in LDS (two n x n
// matrices of dynamic shared memory).
__global__ void __launch_bounds__(BTL) batch_overlap_lds_kernel(BatchArgs B)
{
    extern __shared__ double dyn[];
    __shared__ WgSmall L;
    const BatchMol M = B.mol[blockIdx.x];
    const int n = M.n, N = B.N;
    double* A = dyn;
    double* V = dyn + (size_t)n * n;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int i = e / n, j = e - i * n;
        A[e] = blk(B.S, N, M.f0)[(size_t)i * N + j];
        V[e] = i == j ? 1.0 : 0.0;
    }
    __syncthreads();
    lds_jacobi(n, A, V, L.w, L.perm, L.cs, L.red);
    // X = V w^-1/2 V^T and U = V w^1/2 V^T: the scaled eigenvectors in A, then
    // one product each straight into the molecule's blocks
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) A[e] = V[e] / sqrt(L.w[e % n]);
    __syncthreads();
    lds_gemm(n, false, A, true, V, blk(B.X, N, M.f0), N);
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) A[e] = V[e] * sqrt(L.w[e % n]);
    __syncthreads();
    lds_gemm(n, false, A, true, V, blk(B.U, N, M.f0), N);
    if (threadIdx.x == 0 && !(L.w[0] > 0.0)) B.flag[blockIdx.x] = -1;   // not positive definite
}

// One iteration of the reference's loop for each unconverged molecule.
// build = 0: step 0 (FO = X^T F X); build = 1: after a Fock build into G:
// F += G, P_old = P, DIIS (window 8), FO = X^T F_diis X.  Then eigh(FO),
// C = X C', P = C_occ C_occ^T, the energy sum (H + F)_pq P_qp, ||P - P_old||,
// dP = P - P_old for the next build, and the convergence flag.
__global__ void __launch_bounds__(BT) batch_step_kernel(BatchArgs B, int build, int diis, double conver)
{
    __shared__ WgLds L;
    const int mi = blockIdx.x;
    const BatchMol M = B.mol[mi];
    BatchState& Sx = B.state[mi];
    const int n = M.n, N = B.N, f0 = M.f0;
    if (Sx.converged) return;
    double* F = blk(B.F, N, f0);
    double* P = blk(B.P, N, f0);
    double* Po = blk(B.Pold, N, f0);
    double* dP = blk(B.dP, N, f0);
    const double* G = blk(B.G, N, f0);
    const double* H = blk(B.H, N, f0);
    const double* S = blk(B.S, N, f0);
    const double* X = blk(B.X, N, f0);
    double* T1 = B.work + M.wo;                 // n x n work matrices (packed)
    double* T2 = T1 + (size_t)n * n;
    double* Fd = T2 + (size_t)n * n;
    double* Eset = B.diis + M.d8;               // 8 error matrices, then 8 Fock matrices
    double* Fset = Eset + (size_t)8 * n * n;
    const double* Fuse = F;
    int ld_use = N;
    if (build) {
        for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
            const size_t o = (size_t)(e / n) * N + e % n;
            F[o] += G[o];
            Po[o] = P[o];
        }
        __syncthreads();
        if (diis) {
            // error X (F P S - S P F) X of the new F (mmd/scf.py:146-150)
            const int cnt = Sx.ndiis;
            const int slot = cnt == 8 ? Sx.order[0] : cnt;   // a full window drops its oldest entry
            __syncthreads();
            if (cnt == 8 && threadIdx.x == 0) {
                for (int i = 0; i < 7; ++i) Sx.order[i] = Sx.order[i + 1];
                for (int i = 0; i < 7; ++i)
                    for (int j = 0; j < 7; ++j) Sx.Bm[i * 8 + j] = Sx.Bm[(i + 1) * 8 + (j + 1)];
            }
            __syncthreads();
            const int nw = cnt == 8 ? 8 : cnt + 1;
            double* Es = Eset + (size_t)slot * n * n;
            wg_gemm(n, false, P, N, false, S, N, T1, n);          // P S
            wg_gemm(n, false, F, N, false, T1, n, T2, n);         // F P S
            for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
                const int i = e / n, j = e - i * n;
                T1[e] = T2[i * n + j] - T2[j * n + i];            // FPS - SPF
            }
            __syncthreads();
            wg_gemm(n, false, X, N, false, T1, n, T2, n);
            wg_gemm(n, false, T2, n, false, X, N, Es, n);
            for (int e = threadIdx.x; e < n * n; e += blockDim.x)
                Fset[(size_t)slot * n * n + e] = F[(size_t)(e / n) * N + e % n];
            if (threadIdx.x == 0) Sx.order[nw - 1] = slot;
            __syncthreads();
            // new row of B: tr(e_k^T e_new) (mmd/scf.py:163-167)
            for (int k = 0; k < nw; ++k) {
                const double* Ek = Eset + (size_t)Sx.order[k] * n * n;
                double d = 0.0;
                for (int e = threadIdx.x; e < n * n; e += blockDim.x) d = fma(Ek[e], Es[e], d);
                d = wg_sum(d, L.red);
                if (threadIdx.x == 0) Sx.Bm[(nw - 1) * 8 + k] = Sx.Bm[k * 8 + (nw - 1)] = d;
            }
            if (threadIdx.x == 0) {
                Sx.ndiis = nw;
                // [B -1; -1 0] c = [0 .. 0 -1] (mmd/scf.py:160-170), Gaussian
                // elimination with partial pivoting
                const int mm = nw + 1;
                double* a = L.Bm;
                double* r = L.rhs;
                for (int i = 0; i < mm; ++i) {
                    for (int j = 0; j < mm; ++j)
                        a[i * 9 + j] = (i < nw && j < nw) ? Sx.Bm[i * 8 + j] : ((i == nw && j == nw) ? 0.0 : -1.0);
                    r[i] = i == nw ? -1.0 : 0.0;
                }
                bool ok = true;
                for (int col = 0; col < mm; ++col) {
                    int piv = col;
                    for (int i = col + 1; i < mm; ++i)
                        if (fabs(a[i * 9 + col]) > fabs(a[piv * 9 + col])) piv = i;
                    if (a[piv * 9 + col] == 0.0) { ok = false; break; }
                    if (piv != col) {
                        for (int j = 0; j < mm; ++j) { const double t = a[col * 9 + j]; a[col * 9 + j] = a[piv * 9 + j]; a[piv * 9 + j] = t; }
                        const double t = r[col]; r[col] = r[piv]; r[piv] = t;
                    }
                    for (int i = col + 1; i < mm; ++i) {
                        const double f = a[i * 9 + col] / a[col * 9 + col];
                        for (int j = col; j < mm; ++j) a[i * 9 + j] -= f * a[col * 9 + j];
                        r[i] -= f * r[col];
                    }
                }
                for (int i = mm - 1; i >= 0 && ok; --i) {
                    double v = r[i];
                    for (int j = i + 1; j < mm; ++j) v -= a[i * 9 + j] * r[j];
                    r[i] = v / a[i * 9 + i];
                }
                if (!ok) B.flag[mi] = -2;
            }
            __syncthreads();
            // F_diis = sum_k c_k F_k (mmd/scf.py:175-178)
            for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
                double v = 0.0;
                for (int k = 0; k < nw; ++k) v = fma(L.rhs[k], Fset[(size_t)Sx.order[k] * n * n + e], v);
                Fd[e] = v;
            }
            __syncthreads();
            Fuse = Fd;
            ld_use = n;
        }
    }
    // FO = X^T F X into LDS
    wg_gemm(n, true, X, N, false, Fuse, ld_use, T1, n);
    wg_gemm(n, false, T1, n, false, X, N, T2, n);
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) L.A[e] = T2[e];
    __syncthreads();
    double* CO = T1;   // eigenvectors C' (n x n)
    wg_jacobi(n, L.A, n, CO, n, L.w, L.perm, L.cs, L.red);
    // C = X C'  (T2); P = C_occ C_occ^T
    wg_gemm(n, false, X, N, false, CO, n, T2, n);
    const int nocc = M.nocc;
    double e_el = 0.0, rms = 0.0;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int i = e / n, j = e - i * n;
        double v = 0.0;
        for (int k = 0; k < nocc; ++k) v = fma(T2[i * n + k], T2[j * n + k], v);
        const size_t o = (size_t)i * N + j;
        P[o] = v;
        const double d = v - Po[o];
        dP[o] = d;
        rms = fma(d, d, rms);
        // sum_pq (H + F)_pq P_qp, P symmetric (mmd/scf.py:118-121)
        e_el = fma(H[o] + F[o], v, e_el);
    }
    e_el = wg_sum(e_el, L.red);
    rms = wg_sum(rms, L.red);
    if (threadIdx.x == 0) {
        const double energy = e_el + M.enuc;
        const int step = Sx.iterations;
        Sx.delta_energy = step > 0 ? energy - Sx.energy : 1e20;
        Sx.energy = energy;
        Sx.prms = step > 0 ? sqrt(rms) : 1e20;
        if (step > 0 && Sx.prms < conver) Sx.converged = 1;
        else Sx.iterations = step + 1;
    }
    for (int k = threadIdx.x; k < n; k += blockDim.x) B.mo[M.f0 + k] = L.w[k];
}

// load the n x n block at M (leading dimension N) into LDS
__device__ void lds_load(int n, const double* M, int N, double* L)
{
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) L[e] = M[(size_t)(e / n) * N + e % n];
    __syncthreads();
}

// The step of batch_step_kernel with its matrices in LDS: three n x n
// matrices (dynamic shared memory) hold P S, F P S, the error, X, F_diis, FO
// and the eigenvectors in turn, so the products and the Jacobi rotations run
// from LDS.  Same operations in the same order as batch_step_kernel.
__global__ void __launch_bounds__(BTL) batch_step_lds_kernel(BatchArgs B, int build, int diis, double conver)
{
    extern __shared__ double dyn[];
    __shared__ WgSmall L;
    const int mi = blockIdx.x;
    const BatchMol M = B.mol[mi];
    BatchState& Sx = B.state[mi];
    const int n = M.n, N = B.N, f0 = M.f0;
    if (Sx.converged) return;
    const size_t nn = (size_t)n * n;
    double* M0 = dyn;
    double* M1 = dyn + nn;
    double* M2 = dyn + 2 * nn;
    double* F = blk(B.F, N, f0);
    double* P = blk(B.P, N, f0);
    double* Po = blk(B.Pold, N, f0);
    double* dP = blk(B.dP, N, f0);
    const double* G = blk(B.G, N, f0);
    const double* H = blk(B.H, N, f0);
    const double* S = blk(B.S, N, f0);
    const double* X = blk(B.X, N, f0);
    double* Eset = B.diis + M.d8;
    double* Fset = Eset + (size_t)8 * nn;
    bool fo_from_lds = false;   // F_diis in M1
    if (build) {
        for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
            const size_t o = (size_t)(e / n) * N + e % n;
            F[o] += G[o];
            Po[o] = P[o];
        }
        __syncthreads();
        if (diis) {
            const int cnt = Sx.ndiis;
            const int slot = cnt == 8 ? Sx.order[0] : cnt;
            __syncthreads();
            if (cnt == 8 && threadIdx.x == 0) {
                for (int i = 0; i < 7; ++i) Sx.order[i] = Sx.order[i + 1];
                for (int i = 0; i < 7; ++i)
                    for (int j = 0; j < 7; ++j) Sx.Bm[i * 8 + j] = Sx.Bm[(i + 1) * 8 + (j + 1)];
            }
            __syncthreads();
            const int nw = cnt == 8 ? 8 : cnt + 1;
            double* Es = Eset + (size_t)slot * nn;
            lds_load(n, P, N, M0);
            lds_load(n, S, N, M1);
            lds_gemm(n, false, M0, false, M1, M2);                 // P S
            lds_load(n, F, N, M0);
            lds_gemm(n, false, M0, false, M2, M1);                 // F P S
            for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
                const int i = e / n, j = e - i * n;
                M2[e] = M1[i * n + j] - M1[j * n + i];             // FPS - SPF
            }
            __syncthreads();
            lds_load(n, X, N, M0);
            lds_gemm(n, false, M0, false, M2, M1);                 // X (FPS - SPF)
            lds_gemm(n, false, M1, false, M0, M2);                 // ... X: the error
            for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
                Es[e] = M2[e];
                Fset[(size_t)slot * nn + e] = F[(size_t)(e / n) * N + e % n];
            }
            if (threadIdx.x == 0) Sx.order[nw - 1] = slot;
            __syncthreads();
            for (int k = 0; k < nw; ++k) {
                const int sk = Sx.order[k];
                double d = 0.0;
                if (sk == slot) {
                    for (int e = threadIdx.x; e < n * n; e += blockDim.x) d = fma(M2[e], M2[e], d);
                } else {
                    const double* Ek = Eset + (size_t)sk * nn;
                    for (int e = threadIdx.x; e < n * n; e += blockDim.x) d = fma(Ek[e], M2[e], d);
                }
                d = wg_sum(d, L.red);
                if (threadIdx.x == 0) Sx.Bm[(nw - 1) * 8 + k] = Sx.Bm[k * 8 + (nw - 1)] = d;
            }
            if (threadIdx.x == 0) {
                Sx.ndiis = nw;
                const int mm = nw + 1;
                double* a = L.Bm;
                double* r = L.rhs;
                for (int i = 0; i < mm; ++i) {
                    for (int j = 0; j < mm; ++j)
                        a[i * 9 + j] = (i < nw && j < nw) ? Sx.Bm[i * 8 + j] : ((i == nw && j == nw) ? 0.0 : -1.0);
                    r[i] = i == nw ? -1.0 : 0.0;
                }
                bool ok = true;
                for (int col = 0; col < mm; ++col) {
                    int piv = col;
                    for (int i = col + 1; i < mm; ++i)
                        if (fabs(a[i * 9 + col]) > fabs(a[piv * 9 + col])) piv = i;
                    if (a[piv * 9 + col] == 0.0) { ok = false; break; }
                    if (piv != col) {
                        for (int j = 0; j < mm; ++j) { const double t = a[col * 9 + j]; a[col * 9 + j] = a[piv * 9 + j]; a[piv * 9 + j] = t; }
                        const double t = r[col]; r[col] = r[piv]; r[piv] = t;
                    }
                    for (int i = col + 1; i < mm; ++i) {
                        const double f = a[i * 9 + col] / a[col * 9 + col];
                        for (int j = col; j < mm; ++j) a[i * 9 + j] -= f * a[col * 9 + j];
                        r[i] -= f * r[col];
                    }
                }
                for (int i = mm - 1; i >= 0 && ok; --i) {
                    double v = r[i];
                    for (int j = i + 1; j < mm; ++j) v -= a[i * 9 + j] * r[j];
                    r[i] = v / a[i * 9 + i];
                }
                if (!ok) B.flag[mi] = -2;
            }
            __syncthreads();
            for (int e = threadIdx.x; e < n * n; e += blockDim.x) {   // F_diis into M1
                double v = 0.0;
                for (int k = 0; k < nw; ++k) v = fma(L.rhs[k], Fset[(size_t)Sx.order[k] * nn + e], v);
                M1[e] = v;
            }
            __syncthreads();
            fo_from_lds = true;
        }
    }
    // FO = X^T F X  (X in M0, F in M1)
    lds_load(n, X, N, M0);
    if (!fo_from_lds) lds_load(n, F, N, M1);
    lds_gemm(n, true, M0, false, M1, M2);
    lds_gemm(n, false, M2, false, M0, M1);                         // FO in M1
    // eigenvectors of FO, started from the last step's (kept in the work
    // area): A = V^T FO V is nearly diagonal once the SCF settles
    double* Vkeep = B.work + M.wo;
    if (Sx.warm) {
        lds_load(n, Vkeep, n, M0);
        lds_gemm(n, false, M1, false, M0, M2);                     // FO V
        lds_gemm(n, true, M0, false, M2, M1);                      // V^T FO V
    } else {
        for (int e = threadIdx.x; e < n * n; e += blockDim.x) M0[e] = (e / n == e % n) ? 1.0 : 0.0;
        __syncthreads();
    }
    lds_jacobi(n, M1, M0, L.w, L.perm, L.cs, L.red);               // C' in M0
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) Vkeep[e] = M0[e];
    if (threadIdx.x == 0) Sx.warm = 1;
    lds_load(n, X, N, M2);
    lds_gemm(n, false, M2, false, M0, M1);                         // C = X C' in M1
    const int nocc = M.nocc;
    double e_el = 0.0, rms = 0.0;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int i = e / n, j = e - i * n;
        double v = 0.0;
        for (int k = 0; k < nocc; ++k) v = fma(M1[i * n + k], M1[j * n + k], v);
        const size_t o = (size_t)i * N + j;
        P[o] = v;
        const double d = v - Po[o];
        dP[o] = d;
        rms = fma(d, d, rms);
        e_el = fma(H[o] + F[o], v, e_el);
    }
    e_el = wg_sum(e_el, L.red);
    rms = wg_sum(rms, L.red);
    if (threadIdx.x == 0) {
        const double energy = e_el + M.enuc;
        const int step = Sx.iterations;
        Sx.delta_energy = step > 0 ? energy - Sx.energy : 1e20;
        Sx.energy = energy;
        Sx.prms = step > 0 ? sqrt(rms) : 1e20;
        if (step > 0 && Sx.prms < conver) Sx.converged = 1;
        else Sx.iterations = step + 1;
    }
    for (int k = threadIdx.x; k < n; k += blockDim.x) B.mo[M.f0 + k] = L.w[k];
}

// dP of converged molecules: zero, so the next Fock build skips them.
__global__ void batch_clear_converged_kernel(BatchArgs B)
{
    const BatchMol M = B.mol[blockIdx.x];
    if (!B.state[blockIdx.x].converged) return;
    const int n = M.n, N = B.N;
    double* dP = blk(B.dP, N, M.f0);
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) dP[(size_t)(e / n) * N + e % n] = 0.0;
}

// Number of unconverged molecules into *out.
__global__ void batch_count_kernel(int nmol, const BatchState* state, int* out)
{
    int c = 0;
    for (int i = threadIdx.x; i < nmol; i += blockDim.x) c += state[i].converged ? 0 : 1;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    __shared__ int part[16];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += part[w];
        *out = t;
    }
}

// Copies each molecule's n x n block between the N x N block-diagonal
// matrix and the packed array (unpack: packed -> full).
__global__ void batch_pack_kernel(BatchArgs B, double* full, double* packed, int unpack)
{
    const BatchMol M = B.mol[blockIdx.x];
    const int n = M.n;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const size_t o = (size_t)(M.f0 + e / n) * B.N + M.f0 + e % n;
        if (unpack)
            full[o] = packed[M.po + e];
        else
            packed[M.po + e] = full[o];
    }
}

}  // namespace

hipError_t launch_batch_pack(const BatchArgs& B, const double* full, double* packed, hipStream_t st)
{
    hipLaunchKernelGGL(batch_pack_kernel, dim3(B.nmol), dim3(256), 0, st, B, const_cast<double*>(full), packed, 0);
    return hipGetLastError();
}

hipError_t launch_batch_unpack(const BatchArgs& B, const double* packed, double* full, hipStream_t st)
{
    hipLaunchKernelGGL(batch_pack_kernel, dim3(B.nmol), dim3(256), 0, st, B, full, const_cast<double*>(packed), 1);
    return hipGetLastError();
}

hipError_t launch_batch_overlap(const BatchArgs& B, hipStream_t st)
{
    if (B.lds_bytes >= 2 * (size_t)B.maxn * B.maxn * sizeof(double) + sizeof(WgSmall)) {
        const size_t dyn = 2 * (size_t)B.maxn * B.maxn * sizeof(double);
        hipLaunchKernelGGL(batch_overlap_lds_kernel, dim3(B.nmol), dim3(BTL), dyn, st, B);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(batch_overlap_kernel, dim3(B.nmol), dim3(BT), 0, st, B);
    return hipGetLastError();
}

hipError_t launch_batch_step(const BatchArgs& B, int build, int diis, double conver, hipStream_t st)
{
    // the all-LDS step where three n x n matrices of the largest molecule fit
    // beside the small vectors; the global-memory step otherwise
    if (B.lds_bytes >= 3 * (size_t)B.maxn * B.maxn * sizeof(double) + sizeof(WgSmall)) {
        const size_t dyn = 3 * (size_t)B.maxn * B.maxn * sizeof(double);
        hipLaunchKernelGGL(batch_step_lds_kernel, dim3(B.nmol), dim3(BTL), dyn, st, B, build, diis, conver);
    } else {
        hipLaunchKernelGGL(batch_step_kernel, dim3(B.nmol), dim3(BT), 0, st, B, build, diis, conver);
    }
    if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
    hipLaunchKernelGGL(batch_clear_converged_kernel, dim3(B.nmol), dim3(BT), 0, st, B);
    return hipGetLastError();
}

hipError_t launch_batch_count(const BatchArgs& B, int* out, hipStream_t st)
{
    hipLaunchKernelGGL(batch_count_kernel, dim3(1), dim3(1024), 0, st, B.nmol, B.state, out);
    return hipGetLastError();
}

}  // namespace mmd
