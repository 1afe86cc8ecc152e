// Kernel entry points of libmmd (host-side declarations).
#pragma once
#include <hip/hip_runtime.h>
#include <mutex>
#include <vector>
#include "mmd_types.h"

namespace mmd {

// Held around every device allocation and free of the library and around the
// capture of a HIP graph: allocation calls of one thread must not interleave
// with a stream capture running in another (contexts run in parallel threads).
std::recursive_mutex& device_alloc_mutex();


// Quartet class index: qc = pcb (pcb+1)/2 + pck with pcb >= pck.
inline __host__ __device__ void qclass_decode(int qc, int* la, int* lb, int* lc, int* ld)
{
    int pcb = 0;
    while ((pcb + 1) * (pcb + 2) / 2 <= qc) ++pcb;
    int pck = qc - pcb * (pcb + 1) / 2;
    int l1 = 0;
    while ((l1 + 1) * (l1 + 2) / 2 <= pcb) ++l1;
    *la = l1;
    *lb = pcb - l1 * (l1 + 1) / 2;
    int l3 = 0;
    while ((l3 + 1) * (l3 + 2) / 2 <= pck) ++l3;
    *lc = l3;
    *ld = pck - l3 * (l3 + 1) / 2;
}

// fp32 = 1 selects the FP32-tier instantiation of the class kernel.
hipError_t launch_fock_class(int qc, int fp32, const ClassArgs& A, const RowArgs& R, int grid, hipStream_t st);
// kj = 1: including the ket-side Coulomb buffer of the classes that have one.
size_t fock_class_lds(int qc, int fp32, int N, int maxpp, int kj);
// Classes whose shell quartets run one per wavefront from survivor lists:
// the high angular-momentum classes on the matrix cores (fp32 = 1: FP32
// MFMA), any other class on the generic FP64 wavefront kernel.
hipError_t launch_fock_wave_class(int la, int lb, int lc, int ld, int fp32, const ClassArgs& A, int64_t max_quartets,
                                  hipStream_t st);
hipError_t launch_fock_mfma_class(int la, int lb, int lc, int ld, int fp32, const ClassArgs& A, int64_t max_quartets,
                                  hipStream_t st);
// Stored ERI tensor from shell quartets: class kernels (every class but
// (dd|dd)) and the wavefront kernel for (dd|dd).  table: staged bra data,
// tensor_class_bps(qc) doubles per primitive pair.
hipError_t launch_tensor_class(int qc, const ClassArgs& A, int boff, int nb, int koff, int nk, int tri,
                               double* table, double* out, hipStream_t st);
int tensor_class_bps(int qc);
hipError_t launch_tensor_wave_class(const ClassArgs& A, int64_t nquartets, hipStream_t st);
// High angular-momentum classes whose two Hermite contractions run on the
// matrix cores (fock_mfma_class_kernel), one wavefront per shell quartet.
inline bool fock_class_is_mfma(int la, int lb, int lc, int ld)
{
    return la == 2 && ((lb == 2 && lc + ld >= 2) || (lb == 1 && lc == 2) || (lb == 1 && lc == 1 && ld == 1));
}
// Classes served by a wavefront-per-quartet kernel (survivor lists) instead of
// the unrolled row kernel by default.
inline bool fock_class_is_wave(int la, int lb, int lc, int ld) { return fock_class_is_mfma(la, lb, lc, ld); }
// Matrix-core classes that also have a compiled row kernel (all but (dd|dd)).
inline bool fock_class_has_both(int la, int lb, int lc, int ld)
{
    return fock_class_is_mfma(la, lb, lc, ld) && !(la == 2 && lb == 2 && lc == 2 && ld == 2);
}

// A single Cartesian basis function for the generic (per basis-function
// quartet) kernel: centre, exponents (l,m,n), primitive range; the
// primitive weights are norm*coef products as cython/basis.pxi stores them.
struct BFDesc {
    double x, y, z;
    int32_t l, m, n;
    int32_t p0, np;
    int32_t pad;   // molecule of the function (0 for a single molecule)
};

// plist (optional): the (p, q) pairs to compute, one molecule's pairs after
// another, for a basis that holds several molecules block-diagonally; the
// other elements of the output are cleared.
hipError_t launch_eri_diag(int N, const BFDesc* bfs, const double* pexp, const double* pw, const double* boys,
                           double* out, hipStream_t st, const int2* plist = nullptr, int64_t npl = 0);
hipError_t launch_eri_bf_list(int64_t nq, const int64_t* ijkl, const BFDesc* bfs, const double* pexp,
                              const double* pw, const double* boys, double* out, hipStream_t st);
hipError_t launch_canonical_quartets(int mode, int N, const BFDesc* bfs, const double* pexp, const double* pw,
                                     const double* boys, const double* D, const double* absD, const double* sqS,
                                     double tol, double* out, hipStream_t st);
hipError_t launch_one_electron(int nbf, const BFDesc* bfs, const double* pexp, const double* pw, int natom,
                               const double* atoms /* natom x 4: Z, x, y, z */, double cx, double cy, double cz,
                               const double* boys, double* S, double* T, double* V, double* M, double* Lm,
                               hipStream_t st, const int2* plist = nullptr, int64_t npl = 0,
                               const int2* mol_atoms = nullptr /* per molecule: first atom, count */);
// One-electron integrals by shell pair (compile-time angular momenta per
// pair class): a shell with its primitives at exps/wts[p0 .. p0+np) (the
// weights of its first Cartesian component; gcomp scales the others).
struct OneeShell {
    double x, y, z;
    int32_t l, np, p0, bf0, mol, pad;
};
struct OneeArgs {
    const OneeShell* shells;
    const double* exps;
    const double* wts;
    const double* gcomp;
    const double* atoms;       // natom x 4: Z, x, y, z
    int natom;
    const int2* mol_atoms;     // per molecule: first atom, count; or null (every atom)
    const double* boys;
    double cx, cy, cz;         // origin of the multipoles
    int nbf;
    double *S, *T, *V;         // N x N
    double *M, *Lm;            // 3 x N x N each, or null (not computed)
};
// list: shell pairs (a, b), l_a >= l_b, of pair class pc at
// list[off[pc] .. off[pc+1]) (off on the host).  Writes the pairs' blocks
// and their transposes; the other elements are left as they are.
hipError_t launch_one_electron_shells(const OneeArgs& A, const int2* list, const int* off, hipStream_t st);
// The Schwarz diagonal out[i][j] = (ij|ij) of the same shell pairs (only
// A.shells, exps, wts, gcomp, boys and nbf are read).
hipError_t launch_schwarz_shells(const OneeArgs& A, const int2* list, const int* off, double* out, hipStream_t st);
// Survivor list of one class; with surv2, the quartets whose bound lies in
// [tol, split) go to the second list instead (the FP32 tier).
hipError_t launch_screen_class(int boff, int nb, int koff, int nk, int tri, const ShellPair* pairs,
                               const double* Dsh, int nshell, double tol, int keep_all, int2* surv, int* count,
                               hipStream_t st, double split = 0.0, int2* surv2 = nullptr, int* count2 = nullptr);
hipError_t launch_abs_density(int n, const double* Dre, const double* Dim, double* absD, hipStream_t st);
// Electronic part of the RHF nuclear gradient (natom x 3, added to grad):
// P, W = P F P (N x N), bf_atom (N); atoms natom x 4 (Z, x, y, z).
hipError_t launch_rhf_gradient(int nbf, const BFDesc* bfs, const double* pexp, const double* pw, int natom,
                               const double* atoms, const double* boys, const int* bf_atom, const double* P,
                               const double* W, double* grad, hipStream_t st);
// Block-diagonal batches (blocks: per molecule {first function, functions,
// first shell, end shell}): |dP|, zero G, shell maxima and each molecule's
// max |dP| (mol_dmax) in the blocks only; the symmetrisation of the blocks.
hipError_t launch_block_prep(int nmol, const int4* blocks, int N, const double* dP, double* absD, double* G,
                             const int* sh_bf0, const int* sh_nbf, int nshell, double* Dsh, double* mol_dmax,
                             hipStream_t st);
hipError_t launch_block_symmetrize(int nmol, const int4* blocks, int N, const double* G, double* out,
                                   hipStream_t st);
hipError_t launch_shell_block_max(int nshell, const int* sh_bf0, const int* sh_nbf, int N, const double* absD,
                                  double* Dsh, hipStream_t st);
hipError_t launch_count_prims(const int2* surv, const int* count, const ShellPair* pairs, unsigned long long* out,
                              hipStream_t st);
hipError_t launch_fill(size_t n, double* x, double v, hipStream_t st);
hipError_t launch_fill_int(int n, int* x, int v, hipStream_t st);
hipError_t launch_symmetrize(int N, const double* G, double* out, hipStream_t st);
hipError_t launch_sqrt_clamped(size_t n, const double* x, double* y, hipStream_t st);
hipError_t launch_jk_tensor(int N, const double* T, const double* Pre, const double* Pim, double* Jre,
                            double* Jim, double* Kre, double* Kim, hipStream_t st);

}  // namespace mmd

namespace mmd {
void boys_build_table(std::vector<double>& tab);
hipError_t launch_boys_eval(int n, const double* m, const double* T, int table, const double* tab, double* out,
                            hipStream_t st);
}  // namespace mmd
