// One shell-quartet class of the Fock-build kernel per translation unit; the
// build sets QC_LA, QC_LB, QC_LC, QC_LD, and QC_F32 for the FP32-tier
// instantiation.  Splitting the classes over separate objects lets them
// compile in parallel.
#include <algorithm>
#include "mmd_eri_impl.h"

#define MMD_PASTE6(p, a, b, c, d, s) p##a##b##c##d##s
#define MMD_NAME(p, a, b, c, d, s) MMD_PASTE6(p, a, b, c, d, s)

#ifdef QC_F32
#define QC_T float
#define QC_SUFFIX _f
#else
#define QC_T double
#define QC_SUFFIX _d
#endif

namespace mmd {

hipError_t MMD_NAME(fock_class_launch_, QC_LA, QC_LB, QC_LC, QC_LD, QC_SUFFIX)(const ClassArgs& A, const RowArgs& R,
                                                                               int grid, hipStream_t st)
{
    const size_t lds = fock_class_lds_bytes<QC_LA, QC_LB, QC_LC, QC_LD, QC_T>(R.rw, R.maxpp, R.kj);
    hipLaunchKernelGGL((eri_fock_class_kernel<QC_LA, QC_LB, QC_LC, QC_LD, QC_T>), dim3(grid), dim3(CLASS_BLOCK), lds,
                       st, A, R);
    return hipGetLastError();
}

size_t MMD_NAME(fock_class_lds_, QC_LA, QC_LB, QC_LC, QC_LD, QC_SUFFIX)(int rw, int maxpp, int kj)
{
    return fock_class_lds_bytes<QC_LA, QC_LB, QC_LC, QC_LD, QC_T>(rw, maxpp, kj);
}

#ifndef QC_F32
// Stored-tensor mode of the class (FP64 only).
hipError_t MMD_NAME(tensor_class_launch_, QC_LA, QC_LB, QC_LC, QC_LD, QC_SUFFIX)(const ClassArgs& A, int boff, int nb,
                                                                                int koff, int nk, int tri,
                                                                                double* table, double* out,
                                                                                hipStream_t st)
{
    if (nb <= 0 || (!tri && nk <= 0)) return hipSuccess;
    hipLaunchKernelGGL((bra_table_kernel<QC_LA, QC_LB, QC_LC, QC_LD>), dim3(nb), dim3(256), 0, st, A, boff, table);
    const int64_t nq = tri ? (int64_t)nb * (nb + 1) / 2 : (int64_t)nb * nk;
    const int64_t items = nq * ClassShape<QC_LA, QC_LB, QC_LC, QC_LD>::NCHUNK;
    const int64_t grid = std::min<int64_t>((items + 255) / 256, 8192);
    hipLaunchKernelGGL((eri_tensor_class_kernel<QC_LA, QC_LB, QC_LC, QC_LD>), dim3((unsigned)grid), dim3(256), 0, st,
                       A, boff, nb, koff, nk, tri, table, out);
    return hipGetLastError();
}

int MMD_NAME(tensor_class_bps_, QC_LA, QC_LB, QC_LC, QC_LD, QC_SUFFIX)()
{
    return ClassShape<QC_LA, QC_LB, QC_LC, QC_LD>::BPS;
}
#endif

}  // namespace mmd
