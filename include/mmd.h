/*
 * mmd.h -- C ABI of libmmd, the MI355X (gfx950) McMurchie-Davidson
 * electron-repulsion-integral and Fock-build library.
 *
 * Plain pointers and sizes only.  Matrices are row-major float64; an N x N
 * matrix has element (i,j) at [i*N+j]; the ERI tensor is chemists' (ij|kl)
 * at [((i*N+j)*N+k)*N+l].  Basis functions are the reference's Cartesian
 * components, in the reference's order (atoms -> shells -> components
 * x, y, z / xx, xy, xz, yy, yz, zz).
 *
 * Every call returns 0 (MMD_OK) or a negative MMD_ERR_* code; the message of
 * the last failure on the calling thread is returned by mmd_last_error().
 * All calls are synchronous: outputs are complete when the call returns.
 * A context belongs to one host thread at a time; calls on different
 * contexts may run concurrently.
 */
#ifndef MMD_H
#define MMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MMD_OK 0
#define MMD_ERR_ARG (-1)         /* invalid argument (Python: ValueError)          */
#define MMD_ERR_NOMEM (-3)       /* device or host allocation failed (MemoryError) */
#define MMD_ERR_DEVICE (-4)      /* HIP runtime / no device (RuntimeError)         */
#define MMD_ERR_STATE (-5)       /* call order: e.g. no basis set yet              */
#define MMD_ERR_UNSUPPORTED (-6) /* angular momentum above d, incomplete shells    */

typedef struct mmd_ctx mmd_ctx;

/* Message of the last failed call on this thread ("" if none). */
const char *mmd_last_error(void);

/* Library ABI version (this header: 1). */
int mmd_version(void);

/* Number of visible HIP devices (0 without a GPU). */
int mmd_device_count(int *count);

/* Create / destroy a context on one device (one HIP stream per context). */
int mmd_ctx_create(mmd_ctx **out, int device_id);
void mmd_ctx_destroy(mmd_ctx *ctx);

/*
 * Basis functions, one per Cartesian component (replaces the list of
 * `Basis` objects, cython/basis.pxi:6-73, built by Molecule.formBasis,
 * mmd/molecule.py:57-73).
 *   origin  : 3*nbf, bohr
 *   lmn     : 3*nbf exponents (l, m, n), l+m+n <= 2
 *   nprim   : nbf primitive counts
 *   exps    : sum(nprim) exponents, basis function after basis function
 *   weights : sum(nprim) products norm[k]*coefs[k] of the normalised
 *             function (cython/basis.pxi:87-120)
 * Replaces any previous basis and drops derived data (Schwarz, tensor).
 */
int mmd_set_basis(mmd_ctx *ctx, int nbf, const double *origin, const int32_t *lmn, const int32_t *nprim,
                  const double *exps, const double *weights);

/* Nuclei for the nuclear-attraction integrals: natom charges Z and 3*natom
   coordinates (bohr). */
int mmd_set_atoms(mmd_ctx *ctx, int natom, const double *Z, const double *xyz);

/*
 * Several molecules in one basis, block-diagonally: the basis of
 * mmd_set_basis and the nuclei of mmd_set_atoms are those of nmol molecules
 * one after the other, molecule m owning nbf_per_mol[m] functions and
 * natom_per_mol[m] nuclei.  Integrals, pairs and quartets never mix two
 * molecules: every N x N result is block-diagonal, and one Fock build serves
 * all molecules.  No counterpart in the reference (one Molecule at a time);
 * used by mmd_batch_rhf.
 */
int mmd_set_molecules(mmd_ctx *ctx, int nmol, const int32_t *nbf_per_mol, const int32_t *natom_per_mol);

/*
 * One-electron integrals (Molecule.one_electron_integrals,
 * mmd/molecule.py:235-281, with S/T/V/Mu/RxDel of cython/onee.pyx:14-69):
 * S, T, V (sum over nuclei of -Z V_C): N*N each; M (dipole, origin `center`)
 * and L (angular momentum, antisymmetric): 3*N*N each.  Host pointers.
 */
int mmd_one_electron(mmd_ctx *ctx, const double center[3], double *S, double *T, double *V, double *M,
                     double *L);

/*
 * Schwarz diagonal (pq|pq) for every pair (the `screen` dictionary of
 * Molecule.build(direct=True), mmd/molecule.py:93-99), as a full symmetric
 * N*N host matrix (may be NULL).  Also builds the shell-pair tables and
 * quartet bounds used by mmd_fock_delta.
 */
int mmd_schwarz(mmd_ctx *ctx, double *diag);

/*
 * Install a caller-supplied Schwarz diagonal (N*N, (pq|pq)) instead of the
 * computed one -- the `screen` argument of formPT (cython/fock.pyx:13).
 */
int mmd_set_schwarz(mmd_ctx *ctx, const double *diag);

/*
 * Full ERI tensor with 8-fold symmetry (doERIs, cython/twoe.pyx:9-31).
 * out: N^4 doubles, host memory (out_on_device = 0) or device memory
 * (out_on_device = 1).
 */
int mmd_eri_tensor(mmd_ctx *ctx, double *out, int out_on_device);

/*
 * Contracted ERIs of arbitrary basis-function quartets of the current basis
 * (ERI(a,b,c,d), cython/twoe.pyx:34-50): ijkl holds 4*n indices, out n
 * values.  Host pointers.
 */
int mmd_eri_list(mmd_ctx *ctx, int64_t n, const int64_t *ijkl, double *out);

/*
 * Direct incremental two-electron Fock contribution G(dP), dP = P - P_old
 * (formPT, cython/fock.pyx:10-87).  Needs mmd_schwarz first.
 *   dP_re, dP_im : N*N real / imaginary parts (dP_im may be NULL)
 *   tol          : screening threshold (acc2e)
 *   raw = 1      : formPT's own unsymmetrised output (per canonical
 *                  basis-function quartet, any dP)
 *   raw = 0      : the symmetrised matrix 0.5*(G + G^T), as mmd/scf.py:93
 *                  forms it; requires real symmetric dP (dP_im == NULL)
 *   G_re, G_im   : N*N outputs (G_im ignored if dP_im == NULL)
 *   on_device    : all of dP_*, G_* are device pointers (1) or host (0)
 *   nsurv        : if not NULL, receives the number of shell quartets that
 *                  passed the shell-level screen
 */
int mmd_fock_delta(mmd_ctx *ctx, const double *dP_re, const double *dP_im, double tol, int raw, double *G_re,
                   double *G_im, int on_device, int64_t *nsurv);

/*
 * mmd_fock_delta's symmetrised build (raw = 0, real dP) that allocates
 * nothing: no counterpart in the reference; it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  dP and G are N*N device pointers; nsurv may be
 * NULL.  The kernels are mmd_fock_delta's; their work buffers (|dP|, shell
 * maxima, the unsymmetrised accumulator, counters, survivor lists) live in a
 * static 16 MiB array per device instead of the context, so the basis is
 * limited to what fits there (about 700 functions with the default LDS
 * budget; a larger basis, or survivor lists beyond the array, returns
 * MMD_ERR_ARG).  The context must have its Schwarz screen (mmd_schwarz or
 * mmd_set_schwarz).  Every call on one GPU shares that GPU's array: calls
 * that overlap in time on one GPU (from several threads or contexts) race
 * and must be serialised by the caller.  Synchronous, like the other calls.
 */
int mmd_fock_delta_noalloc(mmd_ctx *ctx, const double *dP, double tol, double *G, int64_t *nsurv);

/*
 * Conventional J and K from a stored tensor (mmd/scf.py:97-98):
 *   J_pq = sum_rs (pq|rs) P_sr,   K_pq = sum_rs (ps|qr) P_sr
 * T: N^4 (host or device per t_on_device); P, J, K: N*N host pointers;
 * P_im / J_im / K_im may be NULL for a real density.
 */
int mmd_jk_tensor(mmd_ctx *ctx, const double *T, int t_on_device, const double *P_re, const double *P_im,
                  double *J_re, double *J_im, double *K_re, double *K_im);

/*
 * Boys function F_m(T) (boys, cython/util.pxi:53-55; test hook _boys,
 * cython/onee.pyx:195-196), evaluated on the device for n arguments:
 *   table = 0 : general real order m >= 0 (series / continued fraction)
 *   table = 1 : the tabulated kernel path (integer m <= 8)
 * Host pointers.
 */
int mmd_boys(mmd_ctx *ctx, int n, const double *m, const double *T, int table, double *out);

/*
 * Electronic part of the nuclear gradient of a converged RHF energy,
 * replacing the derivative-integral loops of Forces.forces
 * (mmd/forces.py:8-97; integrals cython/grad.pyx:11-350):
 *   grad[3*A + x] = sum_pq P_pq 2 dH_pq/dX_Ax + sum_pqrs Gamma_pqrs d(pq|rs)/dX_Ax
 *                   - 2 sum_pq W_pq dS_pq/dX_Ax
 * for the context's basis and nuclei (mmd_set_atoms); P the half density
 * C_occ C_occ^T and W = P F P (N x N, host), bf_atom the atom of each basis
 * function (the reference's atom.mask, mmd/molecule.py:74-80).  The nuclear
 * repulsion term (forces.py:52-60) is the caller's.  grad: natom x 3, host.
 * One molecule per context; up to 4096 atoms.  mmd_last_gradient_ms: device
 * time of the last call (HIP events around its kernels).
 */
int mmd_rhf_gradient(mmd_ctx *ctx, const double *P, const double *W, const int32_t *bf_atom, double *grad);
int mmd_last_gradient_ms(mmd_ctx *ctx, double *ms);

/*
 * Per-class statistics of the symmetrised Fock build.  With profiling on,
 * every mmd_fock_delta(raw = 0) records, per shell-quartet class
 * (la lb | lc ld): surviving shell quartets, surviving primitive shell
 * quartets and the class kernel's device time (HIP events on the stream the
 * class runs on).  on = 1 keeps the classes concurrent (times include the
 * sharing of the GPU), on = 2 runs them one after the other (each timed
 * alone), on = 0 turns profiling off.  No counterpart in the reference
 * (instrumentation).
 * mmd_fock_stats fills up to 21 entries (lclass: 4 per class) and returns
 * the number of classes.
 */
/*
 * Kernel choice for the six shell-quartet classes that have both a row
 * kernel and a matrix-core kernel ((dd|dp), (dd|pp), (dd|ds), (dp|dp),
 * (dp|pp), (dp|ds)); no counterpart in the reference (tuning).  mode 0: row
 * kernels; 1: matrix cores; 2 (default): the first two measured symmetrised
 * builds after a pair table is built time one variant each, side by side
 * with the other classes as usual, and the faster is kept per class; -1
 * leaves the mode as it is.  choice (optional, 21 entries in class order):
 * per class 0 = row kernel, 1 = matrix cores, -1 = not decided yet.
 */
int mmd_set_class_impl(mmd_ctx *ctx, int mode, int32_t *choice);
/*
 * Ket primitive pairs per lane item of the row kernels' sliced classes, per
 * class (21 entries in class order, 1..16; 0 keeps a class's value; slices
 * NULL changes nothing); current (optional, 21 entries) receives the values
 * in force.  Default 2.  No counterpart in the reference (tuning; see
 * mmd.tune.tune_ket_slices).
 */
int mmd_set_ket_slices(mmd_ctx *ctx, const int32_t *slices, int32_t *current);
int mmd_set_profiling(mmd_ctx *ctx, int on);
int mmd_fock_stats(mmd_ctx *ctx, int32_t *lclass, int64_t *nquart, int64_t *nprim, double *ms);
/*
 * Sums over every profiled build since profiling was last turned on
 * (symmetrised builds of mmd_fock_delta and the builds of mmd_batch_rhf,
 * whose iterations are then launched one by one instead of from a graph),
 * per class and tier: nquart, nprim and ms get 2 entries per class (FP64
 * launches, FP32-tier launches), builds 1 (builds in which the class ran).
 * Returns the number of classes (21).
 */
int mmd_fock_profile_sums(mmd_ctx *ctx, int32_t *lclass, int64_t *nquart, int64_t *nprim, double *ms,
                          int64_t *builds);

/* Device time (ms) of the last mmd_fock_delta call, from HIP events. */
int mmd_last_fock_ms(mmd_ctx *ctx, double *ms);

/*
 * Precision tier of the symmetrised Fock build.  No counterpart in the
 * reference (formPT is FP64 throughout, cython/fock.pyx:72-85): shell
 * quartets whose screening bound q_ab q_cd max|dP| lies in [tol, fp32_below)
 * have their integrals computed in FP32 (digestion into G stays FP64); the
 * rest in FP64.  fp32_below = 0 (the default) keeps every quartet in FP64.
 * mmd_fock_tiers reports the surviving shell quartets of the last
 * symmetrised build and how many of them took the FP32 tier.
 */
int mmd_set_precision(mmd_ctx *ctx, double fp32_below);
int mmd_fock_tiers(mmd_ctx *ctx, int64_t *nquart, int64_t *nquart_fp32);

/*
 * Number of streams (1..8; default min(8, GPU_MAX_HW_QUEUES), 4 when that
 * is unset) the shell-quartet classes of a Fock
 * build are spread over.  No counterpart in the reference: one stream suits
 * callers that run many contexts side by side (batches of small molecules),
 * where the other contexts supply the concurrency.  Classes on different
 * streams run concurrently only as far as the process has hardware queues
 * (HIP's GPU_MAX_HW_QUEUES).
 */
int mmd_set_streams(mmd_ctx *ctx, int n);

/*
 * LDS budget (bytes) of one Fock class launch; 0 restores the device's
 * limit (64 KiB on gfx950).  A class whose bra rows and buffers do not fit
 * runs without the ket-side Coulomb buffer, and without room for the rows it
 * runs on the wavefront-per-quartet kernel with survivor lists; large bases
 * take these paths by themselves, a smaller budget lets a small basis take
 * them too (for testing them).  No counterpart in the reference.
 */
int mmd_set_lds_limit(mmd_ctx *ctx, int64_t bytes);

/*
 * GPU-resident state of the direct RHF loop (mmd/scf.py:10-80 of the
 * reference; the caller runs the loop and the eigensolver).  No counterpart
 * in the reference: H, S, X = S^-1/2, F and the density stay in device
 * memory between Fock builds, with the DIIS window (depth 8, error
 * X (FPS - SPF) X, mmd/scf.py:146-180).  Host pointers, N*N row-major.
 *   scf_begin   : H (core), S, X, the current F and density P (= P_old)
 *   scf_step    : build = 1: F += G(P - P_old) (the symmetrised
 *                 mmd_fock_delta build), P_old = P and, with diis, the DIIS
 *                 extrapolation; then FO = X^T F X (extrapolated F with
 *                 diis) is copied to the host
 *   scf_density : from the eigenvectors C' of FO: C = X C',
 *                 P = C_occ C_occ^T; stats = {sum (H + F)_pq P_qp,
 *                 ||P - P_old||_F}
 *   scf_get     : copies F, P, P_old, G (last build) to the host (NULL skips)
 */
int mmd_scf_begin(mmd_ctx *ctx, int nocc, const double *H, const double *S, const double *X, const double *F,
                  const double *P);
int mmd_scf_step(mmd_ctx *ctx, double tol, int build, int diis, double *FO);
int mmd_scf_density(mmd_ctx *ctx, const double *CO, double *stats);
int mmd_scf_get(mmd_ctx *ctx, double *F, double *P, double *P_old, double *G);

/*
 * Direct RHF of every molecule of the context's basis at once (the loop of
 * SCF.RHF, mmd/scf.py:10-80 and 146-180, with direct=True), all on the GPU:
 * one-electron integrals, Schwarz screen and X = S^-1/2 per molecule, then
 * per iteration one Fock build for the whole batch (the class kernels) and
 * one workgroup per molecule for the F update, DIIS (window 8),
 * FO = X^T F X, a Jacobi eigensolver, the density, the energy and the test
 * ||P - P_old||_F < conver.  A converged molecule keeps its state and adds no
 * work to later builds.  The iteration is captured once and replayed as a
 * HIP graph when use_graph = 1.  No counterpart in the reference.
 *   nocc, enuc        : per molecule (nmol of mmd_set_molecules)
 *   P0                : NULL (core guess) or per-molecule initial densities,
 *                       n_m x n_m each, packed one after the other
 *   conver, acc2e     : as SCF.RHF; maxiter: the reference's 100
 *   diis              : 1 as SCF.RHF(DIIS=True)
 *   energy, iterations, converged : per molecule outputs (iterations: the
 *                       step index of convergence, as SCF.iterations)
 *   P_out (nullable)  : converged densities, packed as P0
 *   mo_out (nullable) : orbital energies, N (molecule after molecule)
 * Molecules of up to 88 basis functions.  The FP32 tier of mmd_set_precision
 * applies to the Fock builds.
 */
int mmd_batch_rhf(mmd_ctx *ctx, const int32_t *nocc, const double *enuc, const double *P0, double conver,
                  double acc2e, int maxiter, int diis, int use_graph, double *energy, int32_t *iterations,
                  int32_t *converged, double *P_out, double *mo_out);

/*
 * Four-index transformation of the stored ERI tensor to the molecular-orbital
 * basis (PostSCF.ao2mo, mmd/postscf.py:21-28):
 *   out[P,Q,R,S] = sum_pqrs C[p,P] C[q,Q] C[r,R] C[s,S] T[p,q,r,s]
 * T, out: N^4 row-major (host or device per the flags); C: N*N host.  Four
 * quarter transforms, each one FP64 matrix-core product (mmd_dgemm's
 * routine) of N x N by N x N^3;
 * gemm_ms (may be NULL) receives their device time.
 */
int mmd_ao2mo(mmd_ctx *ctx, const double *T, int t_on_device, const double *C, double *out, int out_on_device,
              double *gemm_ms);

/*
 * Overlap powers X = S^-1/2 and U = S^1/2 (Molecule.one_electron_integrals,
 * X = fractional_matrix_power(S, -0.5), U = fractional_matrix_power(S, 0.5),
 * mmd/molecule.py:284-286), computed on the GPU by Newton-Schulz iterations
 * of FP64 matrix-core products.  S: N*N host matrix, or NULL for the overlap
 * of the last mmd_one_electron (still on the device); X, U: N*N host
 * outputs (either may be NULL); iterations (may be NULL) receives the number
 * of iterations.  MMD_ERR_ARG if S is not positive definite.
 */
int mmd_overlap_powers(mmd_ctx *ctx, const double *S, double *X, double *U, int *iterations);

/*
 * Row-major FP64 product C = alpha op(A) op(B) + beta C on the matrix cores
 * (v_mfma_f64_16x16x4_f64), op = transpose when trans* = 1: the routine of
 * mmd_ao2mo and mmd_overlap_powers.  No counterpart in the reference (its
 * products are numpy's).  on_device = 1: A, B, C are device pointers;
 * 0: host pointers (copied for the call).
 */
int mmd_dgemm(mmd_ctx *ctx, int transA, int transB, int M, int N, int K, double alpha, const double *A, int lda,
              const double *B, int ldb, double beta, double *C, int ldc, int on_device);

/*
 * Multilayer perceptron of the learned initial densities (no counterpart in
 * the reference): tanh (default) or SiLU hidden layers
 * (mmd_mlp_set_activation: 0 tanh, 1 SiLU), linear output layer, FP32 HIP
 * kernels.
 *   dims            : nlayer + 1 widths (input, hidden..., output)
 *   params          : flat float32, per layer W (fan_in x fan_out, row
 *                     major) then b (fan_out); mmd_mlp_nparam gives the size
 *   forward         : Y (n x dims[nlayer]) = f(X (n x dims[0]))
 *   loss_grad       : loss = mean over all n*dims[nlayer] outputs of
 *                     (Y - Yt)^2 and its gradient with respect to params
 *   train           : epochs of Adam (beta 0.9 / 0.999, eps 1e-8) on
 *                     shuffled full mini-batches; epoch_loss (may be NULL)
 *                     receives the mean batch loss of each epoch
 * Host pointers; calls are synchronous.
 */
typedef struct mmd_mlp mmd_mlp;
int mmd_mlp_create(mmd_mlp **out, int nlayer, const int32_t *dims, int device_id);
void mmd_mlp_destroy(mmd_mlp *mlp);
int64_t mmd_mlp_nparam(mmd_mlp *mlp);
int mmd_mlp_set_activation(mmd_mlp *mlp, int activation);
int mmd_mlp_set_params(mmd_mlp *mlp, const float *params);
int mmd_mlp_get_params(mmd_mlp *mlp, float *params);
int mmd_mlp_forward(mmd_mlp *mlp, int n, const float *X, float *Y);
int mmd_mlp_loss_grad(mmd_mlp *mlp, int n, const float *X, const float *Yt, float *loss, float *grad);
int mmd_mlp_train(mmd_mlp *mlp, int n, const float *X, const float *Yt, int epochs, int batch, float lr,
                  uint64_t seed, float *epoch_loss);

#ifdef __cplusplus
}
#endif

#endif /* MMD_H */
