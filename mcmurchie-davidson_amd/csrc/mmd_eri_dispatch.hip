// Dispatch table over the shell-quartet classes of the Fock-build kernel
// (each compiled in its own objects from mmd_eri_class.hip, in FP64 and in
// the FP32 tier).
#include "mmd_kernels.h"

namespace mmd {

#define MMD_CLASSES(X)                                                                                   \
    X(0, 0, 0, 0) X(1, 0, 0, 0) X(1, 0, 1, 0) X(1, 1, 0, 0) X(1, 1, 1, 0) X(1, 1, 1, 1) X(2, 0, 0, 0)   \
    X(2, 0, 1, 0) X(2, 0, 1, 1) X(2, 0, 2, 0) X(2, 1, 0, 0) X(2, 1, 1, 0) X(2, 1, 1, 1) X(2, 1, 2, 0)   \
    X(2, 1, 2, 1) X(2, 2, 0, 0) X(2, 2, 1, 0) X(2, 2, 1, 1) X(2, 2, 2, 0) X(2, 2, 2, 1)

#define DECL(a, b, c, d)                                                                                         \
    hipError_t fock_class_launch_##a##b##c##d##_d(const ClassArgs& A, const RowArgs& R, int grid, hipStream_t st); \
    hipError_t fock_class_launch_##a##b##c##d##_f(const ClassArgs& A, const RowArgs& R, int grid, hipStream_t st); \
    size_t fock_class_lds_##a##b##c##d##_d(int N, int maxpp, int kj);                                            \
    size_t fock_class_lds_##a##b##c##d##_f(int N, int maxpp, int kj);
MMD_CLASSES(DECL)
#undef DECL

#define DECLT(a, b, c, d)                                                                                    \
    hipError_t tensor_class_launch_##a##b##c##d##_d(const ClassArgs& A, int boff, int nb, int koff, int nk,  \
                                                    int tri, double* table, double* out, hipStream_t st);    \
    int tensor_class_bps_##a##b##c##d##_d();
MMD_CLASSES(DECLT)
#undef DECLT

hipError_t launch_tensor_class(int qc, const ClassArgs& A, int boff, int nb, int koff, int nk, int tri,
                               double* table, double* out, hipStream_t st)
{
    int la, lb, lc, ld;
    qclass_decode(qc, &la, &lb, &lc, &ld);
#define CASE(a, b, c, d)                          \
    if (la == a && lb == b && lc == c && ld == d) \
        return tensor_class_launch_##a##b##c##d##_d(A, boff, nb, koff, nk, tri, table, out, st);
    MMD_CLASSES(CASE)
#undef CASE
    return hipErrorInvalidValue;
}

int tensor_class_bps(int qc)
{
    int la, lb, lc, ld;
    qclass_decode(qc, &la, &lb, &lc, &ld);
#define CASE(a, b, c, d) \
    if (la == a && lb == b && lc == c && ld == d) return tensor_class_bps_##a##b##c##d##_d();
    MMD_CLASSES(CASE)
#undef CASE
    return 0;
}

hipError_t launch_fock_class(int qc, int fp32, const ClassArgs& A, const RowArgs& R, int grid, hipStream_t st)
{
    int la, lb, lc, ld;
    qclass_decode(qc, &la, &lb, &lc, &ld);
#define CASE(a, b, c, d)                                                                               \
    if (la == a && lb == b && lc == c && ld == d)                                                      \
        return fp32 ? fock_class_launch_##a##b##c##d##_f(A, R, grid, st)                               \
                    : fock_class_launch_##a##b##c##d##_d(A, R, grid, st);
    MMD_CLASSES(CASE)
#undef CASE
    return hipErrorInvalidValue;  // wavefront-per-quartet classes are launched separately
}

#define MMD_MFMA_CLASSES(X) X(2, 2, 2, 2) X(2, 2, 2, 1) X(2, 2, 1, 1) X(2, 2, 2, 0) X(2, 1, 2, 1) X(2, 1, 1, 1) X(2, 1, 2, 0)

#define DECLM(a, b, c, d)                                                                                    \
    hipError_t fock_mfma_launch_##a##b##c##d##_d(const ClassArgs& A, int64_t max_quartets, hipStream_t st); \
    hipError_t fock_mfma_launch_##a##b##c##d##_f(const ClassArgs& A, int64_t max_quartets, hipStream_t st);
MMD_MFMA_CLASSES(DECLM)
#undef DECLM

hipError_t launch_fock_mfma_class(int la, int lb, int lc, int ld, int fp32, const ClassArgs& A, int64_t max_quartets,
                                  hipStream_t st)
{
#define CASE(a, b, c, d)                                                                    \
    if (la == a && lb == b && lc == c && ld == d)                                           \
        return fp32 ? fock_mfma_launch_##a##b##c##d##_f(A, max_quartets, st)                \
                    : fock_mfma_launch_##a##b##c##d##_d(A, max_quartets, st);
    MMD_MFMA_CLASSES(CASE)
#undef CASE
    return hipErrorInvalidValue;
}

size_t fock_class_lds(int qc, int fp32, int N, int maxpp, int kj)
{
    int la, lb, lc, ld;
    qclass_decode(qc, &la, &lb, &lc, &ld);
#define CASE(a, b, c, d)                          \
    if (la == a && lb == b && lc == c && ld == d) \
        return fp32 ? fock_class_lds_##a##b##c##d##_f(N, maxpp, kj) : fock_class_lds_##a##b##c##d##_d(N, maxpp, kj);
    MMD_CLASSES(CASE)
#undef CASE
    return 0;
}

}  // namespace mmd
