// FP64 matrix products on the gfx950 matrix cores (v_mfma_f64_16x16x4_f64),
// row-major:  C = alpha * op(A) * op(B) + beta * C,  op(X) = X or X^T.
//
// Used by the four-index transformation (PostSCF.ao2mo, mmd/postscf.py:21-28:
// four quarter transforms, each an N x N^3 by N x N product), by the
// overlap powers S^-1/2 and S^1/2 (mmd/molecule.py:284-286, Newton-Schulz
// iterations made of products) and by the GPU-resident SCF loop.
//
// Workgroup tile 128 x 128 (wide products) or 64 x 64, K staged 16 at a
// time in LDS (k-major for both operands, so that a lane reads its MFMA
// operand with one ds_read_b64) with the next slice prefetched into
// registers; four wavefronts in a 2 x 2 grid, each owning a quarter of the
// tile as 16 x 16 MFMA blocks.
//
// Operand layout of v_mfma_f64_16x16x4_f64 (wave64):
//   A (16 x 4):  lane l holds A[l % 16][l / 16]
//   B (4 x 16):  lane l holds B[l / 16][l % 16]
//   D (16 x 16): lane l holds D[l / 16 + 4 r][l % 16], r = 0..3 (rows
//                interleaved across the four lane groups; measured)
#include <hip/hip_runtime.h>
#include "mmd_gemm.h"

namespace mmd {

namespace {
constexpr int GBK = 16, GTHREADS = 256;

typedef double double4_t __attribute__((ext_vector_type(4)));

// Tile BM x BN, four wavefronts in a 2 x 2 grid, each (BM/2) x (BN/2) as
// MT x NT MFMA blocks.  The next K slice is loaded from global memory into
// registers while the current one is multiplied from LDS.
template <int BM, int BN>
__global__ void __launch_bounds__(GTHREADS)
    gemm_f64_kernel(int tA, int tB, int M, int N, int K, double alpha, const double* __restrict__ A, int lda,
                    const double* __restrict__ B, int ldb, double beta, double* __restrict__ C, int ldc)
{
    constexpr int MT = BM / 32, NT = BN / 32;
    constexpr int EA = BM * GBK / GTHREADS, EB = BN * GBK / GTHREADS;   // staged elements per thread
    __shared__ double As[GBK][BM + 1];
    __shared__ double Bs[GBK][BN + 1];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = wv >> 1, wn = wv & 1;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    double4_t acc[MT][NT];
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};
    double ra[EA], rb[EB];
    // element e of the thread's share of a K slice: (m or n, k) in the tile
    auto coordA = [&](int e, int& m, int& k) {
        const int x = tid + e * GTHREADS;
        if (tA) { m = x % BM; k = x / BM; }      // A is K x M: threads along m
        else    { k = x % GBK; m = x / GBK; }    // A is M x K: threads along k
    };
    auto coordB = [&](int e, int& n, int& k) {
        const int x = tid + e * GTHREADS;
        if (tB) { k = x % GBK; n = x / GBK; }    // B is N x K: threads along k
        else    { n = x % BN; k = x / BN; }      // B is K x N: threads along n
    };
    auto load = [&](int k0) {
#pragma unroll
        for (int e = 0; e < EA; ++e) {
            int m, k;
            coordA(e, m, k);
            const int gm = m0 + m, gk = k0 + k;
            ra[e] = (gm < M && gk < K) ? (tA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk]) : 0.0;
        }
#pragma unroll
        for (int e = 0; e < EB; ++e) {
            int n, k;
            coordB(e, n, k);
            const int gn = n0 + n, gk = k0 + k;
            rb[e] = (gn < N && gk < K) ? (tB ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn]) : 0.0;
        }
    };
    load(0);
    for (int k0 = 0; k0 < K; k0 += GBK) {
#pragma unroll
        for (int e = 0; e < EA; ++e) {
            int m, k;
            coordA(e, m, k);
            As[k][m] = ra[e];
        }
#pragma unroll
        for (int e = 0; e < EB; ++e) {
            int n, k;
            coordB(e, n, k);
            Bs[k][n] = rb[e];
        }
        __syncthreads();
        if (k0 + GBK < K) load(k0 + GBK);   // in flight during the products below
#pragma unroll
        for (int kk = 0; kk < GBK; kk += 4) {
            const int k = kk + (lane >> 4);
            double a[MT], b[NT];
#pragma unroll
            for (int i = 0; i < MT; ++i) a[i] = As[k][wm * (BM / 2) + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < NT; ++j) b[j] = Bs[k][wn * (BN / 2) + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < MT; ++i)
#pragma unroll
                for (int j = 0; j < NT; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            const int col = n0 + wn * (BN / 2) + j * 16 + (lane & 15);
            if (col >= N) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm * (BM / 2) + i * 16 + (lane >> 4) + 4 * r;
                if (row >= M) continue;
                double* c = C + (size_t)row * ldc + col;
                const double v = alpha * acc[i][j][r];
                *c = beta == 0.0 ? v : v + beta * *c;
            }
        }
}
}  // namespace

hipError_t launch_gemm_f64(bool tA, bool tB, int M, int N, int K, double alpha, const double* A, int lda,
                           const double* B, int ldb, double beta, double* C, int ldc, hipStream_t st)
{
    if (M <= 0 || N <= 0) return hipSuccess;
    // wide products (the transform's N x N^3) take 128 x 128 tiles; square
    // products of SCF size take 64 x 64 tiles, for more workgroups
    const int tA_ = tA ? 1 : 0, tB_ = tB ? 1 : 0;
    if (M > 64 && N >= 4096) {
        const dim3 grid((unsigned)((N + 127) / 128), (unsigned)((M + 127) / 128));
        hipLaunchKernelGGL((gemm_f64_kernel<128, 128>), grid, dim3(GTHREADS), 0, st, tA_, tB_, M, N, K, alpha, A,
                           lda, B, ldb, beta, C, ldc);
    } else {
        const dim3 grid((unsigned)((N + 63) / 64), (unsigned)((M + 63) / 64));
        hipLaunchKernelGGL((gemm_f64_kernel<64, 64>), grid, dim3(GTHREADS), 0, st, tA_, tB_, M, N, K, alpha, A, lda,
                           B, ldb, beta, C, ldc);
    }
    return hipGetLastError();
}

}  // namespace mmd
