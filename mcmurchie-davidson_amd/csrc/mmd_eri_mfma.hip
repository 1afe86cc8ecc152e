// One high angular-momentum class of the matrix-core Fock kernel per
// translation unit; the build sets QC_LA, QC_LB, QC_LC, QC_LD and QC_F32 for
// the FP32-tier instantiation (FP32 MFMA with FP32 accumulation).
#include <algorithm>
#include "mmd_eri_mfma.h"

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

hipError_t MMD_NAME(fock_mfma_launch_, QC_LA, QC_LB, QC_LC, QC_LD, QC_SUFFIX)(const ClassArgs& A,
                                                                              int64_t max_quartets, hipStream_t st)
{
    constexpr int waves = mfma_waves<QC_LA, QC_LB, QC_LC, QC_LD, QC_T>();
    int64_t grid = (max_quartets + waves - 1) / waves;
    // persistent wavefronts: each takes quartets until the list is done,
    // so the per-workgroup set-up (descriptor tables) is paid once per CU slot
    grid = std::max<int64_t>(1, std::min<int64_t>(grid, 2048));
    hipLaunchKernelGGL((fock_mfma_class_kernel<QC_LA, QC_LB, QC_LC, QC_LD, QC_T>), dim3((unsigned)grid),
                       dim3(64 * waves), 0, st, A);
    return hipGetLastError();
}

}  // namespace mmd
