// One shell-quartet class of the Fock-build kernel per translation unit; the
// build sets QC_LA, QC_LB, QC_LC, QC_LD.  Splitting the 21 fully unrolled
// classes over separate objects lets them compile in parallel.
#include "mmd_eri_impl.h"

#define MMD_PASTE5(p, a, b, c, d) p##a##b##c##d
#define MMD_NAME(p, a, b, c, d) MMD_PASTE5(p, a, b, c, d)

namespace mmd {

hipError_t MMD_NAME(fock_class_launch_, QC_LA, QC_LB, QC_LC, QC_LD)(const ClassArgs& A, const RowArgs& R, int grid,
                                                                    hipStream_t st)
{
    const size_t lds = fock_class_lds_bytes<QC_LA, QC_LB, QC_LC, QC_LD>(A.N, R.maxpp);
    hipLaunchKernelGGL((eri_fock_class_kernel<QC_LA, QC_LB, QC_LC, QC_LD>), dim3(grid), dim3(256), lds, st, A, R);
    return hipGetLastError();
}

size_t MMD_NAME(fock_class_lds_, QC_LA, QC_LB, QC_LC, QC_LD)(int N, int maxpp)
{
    return fock_class_lds_bytes<QC_LA, QC_LB, QC_LC, QC_LD>(N, maxpp);
}

int MMD_NAME(fock_class_nchunk_, QC_LA, QC_LB, QC_LC, QC_LD)()
{
    return ClassShape<QC_LA, QC_LB, QC_LC, QC_LD>::NCHUNK;
}

}  // namespace mmd
