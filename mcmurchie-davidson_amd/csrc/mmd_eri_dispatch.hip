// Dispatch table over the 21 shell-quartet classes of the Fock-build kernel
// (each compiled in its own object from mmd_eri_class.hip).
#include "mmd_kernels.h"

namespace mmd {

#define MMD_CLASSES(X)                                                                                   \
    X(0, 0, 0, 0) X(1, 0, 0, 0) X(1, 0, 1, 0) X(1, 1, 0, 0) X(1, 1, 1, 0) X(1, 1, 1, 1) X(2, 0, 0, 0)   \
    X(2, 0, 1, 0) X(2, 0, 1, 1) X(2, 0, 2, 0) X(2, 1, 0, 0) X(2, 1, 1, 0) X(2, 1, 1, 1) X(2, 1, 2, 0)   \
    X(2, 1, 2, 1) X(2, 2, 0, 0) X(2, 2, 1, 0) X(2, 2, 1, 1) X(2, 2, 2, 0) X(2, 2, 2, 1)

#define DECL(a, b, c, d)                                                                                       \
    hipError_t fock_class_launch_##a##b##c##d(const ClassArgs& A, const RowArgs& R, int grid, hipStream_t st); \
    int fock_class_nchunk_##a##b##c##d();                                                                     \
    size_t fock_class_lds_##a##b##c##d(int N, int maxpp);
MMD_CLASSES(DECL)
#undef DECL

hipError_t launch_fock_class(int qc, const ClassArgs& A, const RowArgs& R, int grid, hipStream_t st)
{
    int la, lb, lc, ld;
    qclass_decode(qc, &la, &lb, &lc, &ld);
#define CASE(a, b, c, d) \
    if (la == a && lb == b && lc == c && ld == d) return fock_class_launch_##a##b##c##d(A, R, grid, st);
    MMD_CLASSES(CASE)
#undef CASE
    return hipErrorInvalidValue;  // wavefront-per-quartet classes are launched separately
}

size_t fock_class_lds(int qc, int N, int maxpp)
{
    int la, lb, lc, ld;
    qclass_decode(qc, &la, &lb, &lc, &ld);
#define CASE(a, b, c, d) \
    if (la == a && lb == b && lc == c && ld == d) return fock_class_lds_##a##b##c##d(N, maxpp);
    MMD_CLASSES(CASE)
#undef CASE
    return 0;
}

int fock_class_nchunk(int qc)
{
    int la, lb, lc, ld;
    qclass_decode(qc, &la, &lb, &lc, &ld);
#define CASE(a, b, c, d) \
    if (la == a && lb == b && lc == c && ld == d) return fock_class_nchunk_##a##b##c##d();
    MMD_CLASSES(CASE)
#undef CASE
    return 1;
}

}  // namespace mmd
