// libmmd host runtime: the C ABI declared in include/mmd.h.
//
// A context owns one HIP stream, the Boys table, the basis (per function and
// per shell), shell-pair / primitive-pair tables, Schwarz data, the density
// and Fock scratch buffers and the per-class survivor lists.  Callers own
// every input and output pointer.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mmd.h"
#include <rocblas/rocblas.h>
#include "mmd_batch.h"
#include "mmd_device.h"
#include "mmd_gemm.h"
#include "mmd_kernels.h"
#include "mmd_scf.h"
#include "mmd_types.h"

using namespace mmd;

namespace mmd {
std::recursive_mutex& device_alloc_mutex()
{
    static std::recursive_mutex mu;
    return mu;
}
}  // namespace mmd

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIP_TRY(expr)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            if (e_ == hipErrorOutOfMemory) return fail(MMD_ERR_NOMEM, "%s: %s", #expr, hipGetErrorString(e_)); \
            return fail(MMD_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_));                \
        }                                                                                       \
    } while (0)

struct Shell {
    int l, np, p0, bf0;
    int mol;   // molecule of the shell's functions
    double x, y, z;
};

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) {
            std::lock_guard<std::recursive_mutex> lk(device_alloc_mutex());
            (void)hipFree(p);
        }
        p = nullptr;
        n = 0;
    }
    hipError_t reserve(size_t count)
    {
        if (count <= n && p) return hipSuccess;
        std::lock_guard<std::recursive_mutex> lk(device_alloc_mutex());
        release();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
        if (e == hipSuccess) n = count;
        return e;
    }
};

}  // namespace

// ---------------------------------------------------------------------------
// Static scratch of mmd_fock_delta_noalloc: one array per device (each GPU
// the library runs on holds its own copy until the process ends).  Host code
// never passes the array's name to a kernel (on the host it is a stand-in,
// not a device address): every use goes through the address that
// hipGetSymbolAddress returned for the device, looked up once per device when
// the first context on it is created (scratch_for_device).
// ---------------------------------------------------------------------------
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(16) << 20;   // 16 MiB
constexpr int NPRIM_SLOTS = 4 * NQCLASS;   // survivor counters of one Fock build

// Per-class sums over profiled builds; index 0 the FP64 launches, 1 the FP32 tier.
struct ProfileSum {
    int64_t nquart[2] = {0, 0}, nprim[2] = {0, 0};
    double ms[2] = {0.0, 0.0};
    int64_t builds = 0;
};
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[MAX_DEVICES] = {};

// Device address of `scratch` on the current device (device_id), looked up
// on the first call for that device.
int scratch_for_device(int device_id, void** out)
{
    if (device_id < 0 || device_id >= MAX_DEVICES) return fail(MMD_ERR_ARG, "device id out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device_id]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(MMD_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[device_id] = ptr;
    }
    *out = g_scratch_ptr[device_id];
    return MMD_OK;
}
}  // namespace

struct mmd_ctx {
    int device = 0;
    void* noalloc_scratch = nullptr;    // this device's `scratch` (device address)
    hipStream_t st = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // side streams for concurrent class kernels (default priority), with the
    // fork / join events that order them against st
    static constexpr int NSIDE = 7;
    hipStream_t side[NSIDE] = {};
    hipEvent_t evfork = nullptr, evjoin[NSIDE] = {};
    int nstreams = 1 + NSIDE;           // streams the Fock classes are spread over
    double last_fock_ms = 0.0;
    int64_t last_nsurv = 0, last_nsurv_fp32 = 0;   // shell quartets of the last symmetrised build
    double fp32_below = 0.0;                        // FP32 tier: quartets with bound below this (0: off)
    DevBuf<float> d_boysf;

    // basis, per function
    int nbf = 0;
    std::vector<BFDesc> bf;
    std::vector<double> bf_exp, bf_w;
    DevBuf<BFDesc> d_bf;
    DevBuf<double> d_bf_exp, d_bf_w;
    DevBuf<double> d_boys;

    // basis, per shell
    bool shells_ok = false;
    std::string shells_err;
    std::vector<Shell> shells;
    std::vector<double> sh_exp, sh_w, gcomp;
    DevBuf<double> d_gcomp;
    DevBuf<int> d_sh_bf0, d_sh_nbf;
    DevBuf<OneeShell> d_oshell;         // shells for the one-electron kernels
    DevBuf<double> d_sh_exp, d_sh_w;
    DevBuf<int2> d_onee_list;           // every shell pair of each molecule, by pair class
    int onee_off[NPAIRCLASS + 1] = {0};
    bool onee_list_ok = false;

    // shell pairs (grouped by pair class, sorted inside each class)
    bool pairs_ok = false;
    double prim_qmax = 0.0;
    std::vector<ShellPair> pairs;
    int pc_off[NPAIRCLASS + 1] = {0};
    DevBuf<ShellPair> d_pairs;
    DevBuf<PrimPair> d_pp;
    DevBuf<double> d_onee;   // S, T, V, M (3), L (3) of mmd_one_electron
    DevBuf<double> d_ns;     // overlap powers: S, X, U and the Newton-Schulz work (8 N^2 + 2)
    int64_t npp_total = 0;

    // Schwarz
    bool schwarz_ok = false;
    DevBuf<double> d_schwarz, d_sqS;

    // quartet classes
    int64_t qc_total[NQCLASS] = {0};
    int64_t qc_off[NQCLASS + 1] = {0};
    DevBuf<int2> d_surv;
    DevBuf<int> d_count;

    // atoms
    int natom = 0;
    DevBuf<double> d_atoms;
    DevBuf<double> g_PW;                // gradient: P, W and the natom x 3 result
    DevBuf<int> g_atom;                 // gradient: atom of each basis function
    double last_grad_ms = 0.0;

    // Molecules of a block-diagonal basis (mmd_set_molecules; one molecule
    // unless set): molecule m owns functions [mol_f0[m], mol_f0[m] + mol_n[m])
    // and shells [mol_s0[m], mol_s0[m + 1]); pairs and quartets never mix
    // molecules.  rw: the largest molecule (width of the class kernels' LDS
    // rows); plist: the same-molecule function pairs p >= q.
    int nmol = 1;
    std::vector<int> mol_f0{0}, mol_n{0}, mol_s0{0, 0};
    std::vector<int2> mol_atoms;        // per molecule: first atom, number of atoms
    DevBuf<int2> d_mol_atoms;
    DevBuf<int2> d_plist;
    DevBuf<int4> d_mol_blocks;          // per molecule: first function, functions, shells [s0, s1)
    DevBuf<double> d_mol_dmax;          // per molecule: max |dP| of the build
    int64_t npl = 0;
    int rw = 0;
    // per pair class, per molecule: the class-local range [first, end) of its pairs
    std::vector<int> kr_first[NPAIRCLASS], kr_end[NPAIRCLASS];

    // density / Fock scratch
    DevBuf<double> d_D, d_Dim, d_absD, d_Dsh, d_G, d_Gout;

    // per-class statistics of the last symmetrised Fock build
    int profiling = 0;                  // 0 off; 1 events, classes concurrent; 2 events, classes serial
    hipEvent_t evc[3 * NQCLASS] = {};   // per class: start, after the FP64 launch, end
    // Stream plan of the Fock build: class durations measured with the
    // per-class events in one build out of PLAN_EVERY (0: not measured yet).
    static constexpr int PLAN_EVERY = 16;
    double plan_ms[NQCLASS] = {0};
    int plan_countdown = 0;
    bool plan_measuring = false;
    // The six classes with both a row kernel and a matrix-core kernel:
    // impl_mode 0 takes the row kernels, 1 the matrix cores, 2 times both on
    // the first two measured builds of a pair table (the classes running side
    // by side as usual) and keeps the faster per class (impl_choice: -1
    // undecided, 0 row, 1 matrix cores).
    int impl_mode = 2;
    int impl_choice[NQCLASS];
    int ket_slice[NQCLASS];             // mmd_set_ket_slices (KSLICE_DEFAULT until set)
    int tune_phase = 0;
    double tune_ms[2][NQCLASS] = {};
    bool tune_measuring = false;
    DevBuf<unsigned long long> d_nprim;   // NPRIM_SLOTS counters (FockScratch::nprim)
    DevBuf<int> d_rowctr;
    // Work units of the row-kernel classes (built once per pair table):
    // class qc owns units [unit_off[qc], unit_off[qc+1]) with ket segment
    // length unit_seg[qc].
    DevBuf<int4> d_units;
    DevBuf<UnitBound> d_ubound;         // per unit, in the same order
    int64_t unit_off[NQCLASS + 1] = {0};
    int unit_seg[NQCLASS] = {0};
    bool units_ok = false;
    size_t max_lds = 65536;             // LDS a class launch may use (mmd_set_lds_limit)
    size_t max_lds_device = 65536;      // the device's limit per workgroup (class kernels use up to 64 KiB)
    size_t lds_per_block = 65536;       // the device's LDS limit per workgroup, as reported
    bool last_row_kernel[NQCLASS] = {};
    int64_t st_nquart[NQCLASS] = {0};

    // batched RHF (mmd_batch_rhf): N x N block-diagonal matrices, per-molecule
    // work, DIIS windows and state
    DevBuf<double> b_H, b_X, b_U, b_F, b_P, b_Pold, b_dP, b_G, b_work, b_diis, b_mo;
    DevBuf<BatchMol> b_mol;
    DevBuf<BatchState> b_state;
    DevBuf<int> b_flag, b_count;

    // GPU-resident SCF loop (mmd_scf_*): matrices, DIIS window, BLAS handle
    rocblas_handle blas = nullptr;
    int scf_n = 0, scf_nocc = 0;
    DevBuf<double> s_H, s_S, s_X, s_F, s_P, s_Pold, s_G, s_dP, s_T1, s_T2, s_C, s_stats, s_dots;
    DevBuf<double> s_Fset, s_Eset;   // DIIS window: 8 slots of N*N each
    std::vector<int> s_order;        // slots of the window, oldest first
    std::vector<double> s_B;         // 8 x 8 error overlaps, window order
    int64_t st_nprim[NQCLASS] = {0};
    double st_ms[NQCLASS] = {0};
    ProfileSum psum[NQCLASS];           // sums over the builds since profiling was turned on
};

namespace {

int check_ctx(mmd_ctx* ctx)
{
    if (!ctx) return fail(MMD_ERR_ARG, "null context");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return fail(MMD_ERR_DEVICE, "hipSetDevice: %s", hipGetErrorString(e));
    return MMD_OK;
}

// Group consecutive basis functions into complete Cartesian shells in the
// reference's component order with a common centre and exponents and
// weights that differ by a per-component factor only.
void build_shells(mmd_ctx* c)
{
    c->shells.clear();
    c->sh_exp.clear();
    c->sh_w.clear();
    c->gcomp.assign(c->nbf, 1.0);
    c->shells_ok = false;
    int i = 0;
    while (i < c->nbf) {
        const BFDesc& b0 = c->bf[i];
        const int L = b0.l + b0.m + b0.n;
        const int nc = ncart(L);
        if (b0.l != L || i + nc > c->nbf) {
            c->shells_err = "basis functions do not form complete Cartesian shells";
            return;
        }
        for (int k = 0; k < nc; ++k) {
            const BFDesc& bk = c->bf[i + k];
            bool ok = bk.l == comp_lx(L, k) && bk.m == comp_ly(L, k) && bk.n == comp_lz(L, k) && bk.np == b0.np
                   && bk.x == b0.x && bk.y == b0.y && bk.z == b0.z;
            if (ok) {
                const double g = c->bf_w[bk.p0] / c->bf_w[b0.p0];
                for (int p = 0; p < b0.np && ok; ++p) {
                    ok = c->bf_exp[bk.p0 + p] == c->bf_exp[b0.p0 + p];
                    const double gp = c->bf_w[bk.p0 + p] / c->bf_w[b0.p0 + p];
                    ok = ok && fabs(gp - g) <= 1e-12 * fabs(g);
                }
                c->gcomp[i + k] = g;
            }
            if (!ok) {
                c->shells_err = "basis functions do not form complete Cartesian shells";
                return;
            }
        }
        Shell s;
        s.l = L;
        s.mol = b0.pad;
        s.np = b0.np;
        s.p0 = (int)c->sh_exp.size();
        s.bf0 = i;
        s.x = b0.x;
        s.y = b0.y;
        s.z = b0.z;
        for (int p = 0; p < b0.np; ++p) {
            c->sh_exp.push_back(c->bf_exp[b0.p0 + p]);
            c->sh_w.push_back(c->bf_w[b0.p0 + p]);
        }
        c->shells.push_back(s);
        i += nc;
    }
    c->shells_ok = true;
}

int upload_shells(mmd_ctx* c)
{
    HIP_TRY(c->d_gcomp.reserve(c->nbf));
    HIP_TRY(hipMemcpyAsync(c->d_gcomp.p, c->gcomp.data(), sizeof(double) * c->nbf, hipMemcpyHostToDevice, c->st));
    const int ns = (int)c->shells.size();
    std::vector<int> b0(ns), nb(ns);
    for (int s = 0; s < ns; ++s) {
        b0[s] = c->shells[s].bf0;
        nb[s] = ncart(c->shells[s].l);
    }
    HIP_TRY(c->d_sh_bf0.reserve(ns));
    HIP_TRY(c->d_sh_nbf.reserve(ns));
    HIP_TRY(hipMemcpyAsync(c->d_sh_bf0.p, b0.data(), sizeof(int) * ns, hipMemcpyHostToDevice, c->st));
    HIP_TRY(hipMemcpyAsync(c->d_sh_nbf.p, nb.data(), sizeof(int) * ns, hipMemcpyHostToDevice, c->st));
    std::vector<OneeShell> os(std::max(ns, 1));
    for (int k = 0; k < ns; ++k) {
        const Shell& sh = c->shells[k];
        os[k] = OneeShell{sh.x, sh.y, sh.z, sh.l, sh.np, sh.p0, sh.bf0, sh.mol, 0};
    }
    HIP_TRY(c->d_oshell.reserve(os.size()));
    HIP_TRY(c->d_sh_exp.reserve(c->sh_exp.size()));
    HIP_TRY(c->d_sh_w.reserve(c->sh_w.size()));
    HIP_TRY(hipMemcpyAsync(c->d_oshell.p, os.data(), sizeof(OneeShell) * os.size(), hipMemcpyHostToDevice, c->st));
    HIP_TRY(hipMemcpyAsync(c->d_sh_exp.p, c->sh_exp.data(), sizeof(double) * c->sh_exp.size(), hipMemcpyHostToDevice,
                           c->st));
    HIP_TRY(hipMemcpyAsync(c->d_sh_w.p, c->sh_w.data(), sizeof(double) * c->sh_w.size(), hipMemcpyHostToDevice,
                           c->st));
    HIP_TRY(hipStreamSynchronize(c->st));
    c->onee_list_ok = false;
    return MMD_OK;
}

// Schwarz-type bound of a primitive pair's product function (any component):
// sqrt([ab|ab]) of an s-type pair, 2 pi^{5/2} / (p^2 sqrt(2p)) K^2, times
// (1 + |AB| + 1/sqrt(p))^{la+lb} 4^{la+lb} to cover the Hermite expansion of
// higher angular momenta.
double prim_bound(double p, double K, double AB2, int lab)
{
    const double s = sqrt(34.98683665524972497 / (p * p * sqrt(2.0 * p))) * fabs(K);
    const double f = 4.0 * (1.0 + sqrt(AB2) + 1.0 / sqrt(p));
    double m = 1.0;
    for (int i = 0; i < lab; ++i) m *= f;
    return s * m;
}
constexpr double PRIM_DROP = 1e-20;

// Shells [mol_s0[m], mol_s0[m+1]) of each molecule (shells never span two),
// and, for a batch, the molecules' blocks {f0, n, s0, s1} on the device.
int molecule_shell_ranges(mmd_ctx* c)
{
    const int ns = (int)c->shells.size();
    c->mol_s0.assign(c->nmol + 1, ns);
    for (int I = ns - 1; I >= 0; --I) c->mol_s0[c->shells[I].mol] = I;
    for (int m = c->nmol - 1; m >= 0; --m)
        if (c->mol_s0[m] > c->mol_s0[m + 1]) c->mol_s0[m] = c->mol_s0[m + 1];
    if (c->nmol > 1) {
        std::vector<int4> blk(c->nmol);
        for (int m = 0; m < c->nmol; ++m)
            blk[m] = make_int4(c->mol_f0[m], c->mol_n[m], c->mol_s0[m], c->mol_s0[m + 1]);
        HIP_TRY(c->d_mol_blocks.reserve(c->nmol));
        HIP_TRY(c->d_mol_dmax.reserve(c->nmol));
        HIP_TRY(hipMemcpyAsync(c->d_mol_blocks.p, blk.data(), sizeof(int4) * c->nmol, hipMemcpyHostToDevice, c->st));
    }
    return MMD_OK;
}

// Every shell pair of each molecule (l_a >= l_b), grouped by pair class,
// for the shell-pair one-electron and Schwarz kernels; built once per basis.
int onee_pair_list(mmd_ctx* c)
{
    if (c->onee_list_ok) return MMD_OK;
    if (int rc = molecule_shell_ranges(c)) return rc;
    std::vector<std::vector<int2>> byclass(NPAIRCLASS);
    for (int m = 0; m < c->nmol; ++m)
        for (int I = c->mol_s0[m]; I < c->mol_s0[m + 1]; ++I)
            for (int J = c->mol_s0[m]; J <= I; ++J) {
                int a = I, b = J;
                if (c->shells[I].l < c->shells[J].l) std::swap(a, b);
                byclass[pair_class(c->shells[a].l, c->shells[b].l)].push_back(make_int2(a, b));
            }
    std::vector<int2> all;
    for (int pc = 0; pc < NPAIRCLASS; ++pc) {
        c->onee_off[pc] = (int)all.size();
        all.insert(all.end(), byclass[pc].begin(), byclass[pc].end());
    }
    c->onee_off[NPAIRCLASS] = (int)all.size();
    HIP_TRY(c->d_onee_list.reserve(std::max<size_t>(all.size(), 1)));
    HIP_TRY(hipMemcpyAsync(c->d_onee_list.p, all.data(), sizeof(int2) * all.size(), hipMemcpyHostToDevice, c->st));
    HIP_TRY(hipStreamSynchronize(c->st));
    c->onee_list_ok = true;
    return MMD_OK;
}

OneeArgs onee_args(const mmd_ctx* c)
{
    OneeArgs A{};
    A.shells = c->d_oshell.p;
    A.exps = c->d_sh_exp.p;
    A.wts = c->d_sh_w.p;
    A.gcomp = c->d_gcomp.p;
    A.atoms = c->natom ? c->d_atoms.p : nullptr;
    A.natom = c->natom;
    A.mol_atoms = c->nmol > 1 ? c->d_mol_atoms.p : nullptr;
    A.boys = c->d_boys.p;
    A.nbf = c->nbf;
    return A;
}

// One-electron integrals of the basis on the shell-pair kernels: S, T, V
// (N x N) and, when M and Lm are given, the dipole and angular-momentum
// integrals about `center`.  A basis of several molecules gets only the
// molecules' own blocks; the rest of the outputs is cleared.
int one_electron_shells(mmd_ctx* c, const double center[3], double* S, double* T, double* V, double* M, double* Lm)
{
    if (int rc = onee_pair_list(c)) return rc;
    const size_t nn = (size_t)c->nbf * c->nbf;
    if (c->nmol > 1) {
        for (double* x : {S, T, V}) HIP_TRY(launch_fill(nn, x, 0.0, c->st));
        if (M && Lm)
            for (double* x : {M, Lm}) HIP_TRY(launch_fill(3 * nn, x, 0.0, c->st));
    }
    OneeArgs A = onee_args(c);
    A.cx = center[0];
    A.cy = center[1];
    A.cz = center[2];
    A.S = S;
    A.T = T;
    A.V = V;
    A.M = M;
    A.Lm = Lm;
    HIP_TRY(launch_one_electron_shells(A, c->d_onee_list.p, c->onee_off, c->st));
    return MMD_OK;
}

// The Schwarz diagonal (pq|pq) of the basis into out (N x N; for several
// molecules, zero outside their blocks): shell-pair kernels where the basis
// forms complete shells, else the per-function kernel.
int schwarz_diagonal(mmd_ctx* c, double* out)
{
    const int N = c->nbf;
    if (!c->shells_ok) {
        HIP_TRY(launch_eri_diag(N, c->d_bf.p, c->d_bf_exp.p, c->d_bf_w.p, c->d_boys.p, out, c->st,
                                c->nmol > 1 ? c->d_plist.p : nullptr, c->npl));
        return MMD_OK;
    }
    if (int rc = onee_pair_list(c)) return rc;
    if (c->nmol > 1) HIP_TRY(launch_fill((size_t)N * N, out, 0.0, c->st));
    HIP_TRY(launch_schwarz_shells(onee_args(c), c->d_onee_list.p, c->onee_off, out, c->st));
    return MMD_OK;
}

// Host view of the Schwarz diagonal (pq|pq): element (i, j) = p[row[i] + j].
// The full N x N matrix, or the molecules' diagonal blocks packed one after
// another (pairs never mix molecules).
struct DiagView {
    const double* p = nullptr;
    std::vector<int64_t> row;
    double at(int i, int j) const { return p[row[i] + j]; }
};

// Shell pairs of each molecule, grouped by pair class and sorted in each
// class by (molecule, primitive pairs descending, Schwarz factor
// descending), with their primitive-pair tables.  diag: the full N x N
// Schwarz diagonal (pq|pq), or null before it is known (factors 1).
// Shell pairs I >= J in canonical orientation, with primitive-pair data.
// `diag` (optional): the Schwarz diagonal for the factors q; pairs are then
// sorted inside each pair class by (npp desc, q desc).
int build_pairs(mmd_ctx* c, const DiagView* diag)
{
    const int N = c->nbf;
    if (int rc = molecule_shell_ranges(c)) return rc;
    // largest primitive-pair bound, for the primitive-pair screen
    c->prim_qmax = 0.0;
    for (int m = 0; m < c->nmol; ++m)
        for (int I = c->mol_s0[m]; I < c->mol_s0[m + 1]; ++I)
            for (int J = c->mol_s0[m]; J <= I; ++J) {
                const Shell& A = c->shells[I];
                const Shell& B = c->shells[J];
                const double AB2 = (A.x - B.x) * (A.x - B.x) + (A.y - B.y) * (A.y - B.y) + (A.z - B.z) * (A.z - B.z);
                for (int i = 0; i < A.np; ++i)
                    for (int j = 0; j < B.np; ++j) {
                        const double ea = c->sh_exp[A.p0 + i], eb = c->sh_exp[B.p0 + j];
                        const double p = ea + eb;
                        const double K = c->sh_w[A.p0 + i] * c->sh_w[B.p0 + j] * exp(-ea * eb / p * AB2);
                        c->prim_qmax = std::max(c->prim_qmax, prim_bound(p, K, AB2, A.l + B.l));
                    }
            }
    std::vector<std::vector<ShellPair>> byclass(NPAIRCLASS);
    std::vector<PrimPair> pp;
    for (int m = 0; m < c->nmol; ++m)
        for (int I = c->mol_s0[m]; I < c->mol_s0[m + 1]; ++I)
            for (int J = c->mol_s0[m]; J <= I; ++J) {
                int a = I, b = J;
                if (c->shells[I].l < c->shells[J].l) std::swap(a, b);
                const Shell& A = c->shells[a];
                const Shell& B = c->shells[b];
                ShellPair sp;
                sp.sa = a;
                sp.sb = b;
                sp.bfa = A.bf0;
                sp.bfb = B.bf0;
                sp.pp0 = (int)pp.size();
                sp.same = (a == b) ? 1 : 0;
                sp.mol = m;
                sp.c0 = c->mol_f0[m];
                sp.pad = 0;
                // pair factor q = max sqrt((ab|ab)) over the pair's functions
                double q = 1.0;
                if (diag) {
                    q = 0.0;
                    for (int x = 0; x < ncart(A.l); ++x)
                        for (int y = 0; y < ncart(B.l); ++y) {
                            const double v = diag->at(A.bf0 + x, B.bf0 + y);
                            const double r = v > 0.0 ? sqrt(v) : 0.0;
                            if (r > q) q = r;
                        }
                }
                sp.q = q;
                const double AB2 = (A.x - B.x) * (A.x - B.x) + (A.y - B.y) * (A.y - B.y) + (A.z - B.z) * (A.z - B.z);
                for (int i = 0; i < A.np; ++i)
                    for (int j = 0; j < B.np; ++j) {
                        const double ea = c->sh_exp[A.p0 + i], eb = c->sh_exp[B.p0 + j];
                        PrimPair pq;
                        pq.p = ea + eb;
                        pq.K = c->sh_w[A.p0 + i] * c->sh_w[B.p0 + j] * exp(-ea * eb / pq.p * AB2);
                        // Primitive-pair screening (Fock-build tables only): a pair
                        // whose Schwarz-type bound, times the largest bound of any
                        // pair, is below PRIM_DROP contributes less than that to
                        // every primitive quartet and is left out.
                        if (prim_bound(pq.p, pq.K, AB2, A.l + B.l) * c->prim_qmax < PRIM_DROP) continue;
                        pq.Px = (ea * A.x + eb * B.x) / pq.p;
                        pq.Py = (ea * A.y + eb * B.y) / pq.p;
                        pq.Pz = (ea * A.z + eb * B.z) / pq.p;
                        pq.PAx = pq.Px - A.x;
                        pq.PAy = pq.Py - A.y;
                        pq.PAz = pq.Pz - A.z;
                        pq.PBx = pq.Px - B.x;
                        pq.PBy = pq.Py - B.y;
                        pq.PBz = pq.Pz - B.z;
                        pq.inv2p = 0.5 / pq.p;
                        pq.invp = 1.0 / pq.p;
                        pp.push_back(pq);
                    }
                sp.npp = (int)pp.size() - sp.pp0;
                if (sp.npp == 0) sp.q = 0.0;   // no primitive pair survives: never screened in
                byclass[pair_class(A.l, B.l)].push_back(sp);
            }
    c->pairs.clear();
    for (int pc = 0; pc < NPAIRCLASS; ++pc) {
        auto& v = byclass[pc];
        if (diag)
            std::stable_sort(v.begin(), v.end(), [](const ShellPair& x, const ShellPair& y) {
                if (x.mol != y.mol) return x.mol < y.mol;
                if (x.npp != y.npp) return x.npp > y.npp;
                return x.q > y.q;
            });
        c->pc_off[pc] = (int)c->pairs.size();
        c->kr_first[pc].assign(c->nmol, (int)v.size());
        c->kr_end[pc].assign(c->nmol, 0);
        for (int k = 0; k < (int)v.size(); ++k) {
            const int m = v[k].mol;
            c->kr_first[pc][m] = std::min(c->kr_first[pc][m], k);
            c->kr_end[pc][m] = std::max(c->kr_end[pc][m], k + 1);
        }
        for (int m = 0; m < c->nmol; ++m)
            if (c->kr_first[pc][m] > c->kr_end[pc][m]) c->kr_first[pc][m] = c->kr_end[pc][m] = 0;
        c->pairs.insert(c->pairs.end(), v.begin(), v.end());
    }
    c->pc_off[NPAIRCLASS] = (int)c->pairs.size();
    c->npp_total = (int64_t)pp.size();
    HIP_TRY(c->d_pairs.reserve(c->pairs.size()));
    HIP_TRY(c->d_pp.reserve(pp.size()));
    HIP_TRY(hipMemcpyAsync(c->d_pairs.p, c->pairs.data(), sizeof(ShellPair) * c->pairs.size(), hipMemcpyHostToDevice,
                           c->st));
    HIP_TRY(hipMemcpyAsync(c->d_pp.p, pp.data(), sizeof(PrimPair) * pp.size(), hipMemcpyHostToDevice, c->st));
    HIP_TRY(hipStreamSynchronize(c->st));
    // quartet-class sizes (implicit enumeration; for several molecules an
    // upper bound, the screen keeps the quartets of one molecule)
    c->qc_off[0] = 0;
    for (int qc = 0; qc < NQCLASS; ++qc) {
        int pcb = 0;
        while ((pcb + 1) * (pcb + 2) / 2 <= qc) ++pcb;
        const int pck = qc - pcb * (pcb + 1) / 2;
        const int64_t nb = c->pc_off[pcb + 1] - c->pc_off[pcb];
        const int64_t nk = c->pc_off[pck + 1] - c->pc_off[pck];
        c->qc_total[qc] = (pcb == pck) ? nb * (nb + 1) / 2 : nb * nk;
        c->qc_off[qc + 1] = c->qc_off[qc] + c->qc_total[qc];
    }
    HIP_TRY(c->d_count.reserve(2 * NQCLASS));
    c->pairs_ok = true;
    c->units_ok = false;
    for (int qc = 0; qc < NQCLASS; ++qc) {
        c->plan_ms[qc] = 0.0;
        c->impl_choice[qc] = -1;
    }
    c->plan_countdown = 0;
    c->tune_phase = 0;
    return MMD_OK;
}

// Work units of every class: one per (bra pair, slice of at most PSLICE of
// its primitive pairs, segment of seg ket pairs), in decreasing order of
// estimated cost (bra primitive pairs x ket primitive pairs of the
// segment), so that the heaviest units start first.  Slicing the bra
// primitives spreads the heavily contracted pairs over several workgroups.
void class_of(int qc, int* pcb, int* pck);

int build_units(mmd_ctx* c)
{
    std::vector<int4> all;
    std::vector<UnitBound> allb;
    std::vector<std::pair<double, int4>> v;
    std::vector<double> kpre, kqmax;
    for (int qc = 0; qc < NQCLASS; ++qc) {
        c->unit_off[qc] = (int64_t)all.size();
        int pcb, pck;
        class_of(qc, &pcb, &pck);
        const int boff = c->pc_off[pcb], nb = c->pc_off[pcb + 1] - boff;
        const int koff = c->pc_off[pck], nk = c->pc_off[pck + 1] - koff;
        const bool tri = pcb == pck;
        const int64_t quartets = tri ? (int64_t)nb * (nb + 1) / 2 : (int64_t)nb * nk;
        // ket segment: long (survivors are batched inside a unit), but short
        // enough for a few thousand segments per class.  (Shorter segments
        // make a class that runs alone faster, but cost more than they gain
        // when classes or molecules share the GPU.)
        int seg = 4096;
        while (seg > 256 && quartets / std::max(1, c->nmol) / seg < 4096) seg /= 2;
        c->unit_seg[qc] = seg;
        if (nb == 0 || nk == 0) continue;
        kpre.assign(nk + 1, 0.0);
        for (int k = 0; k < nk; ++k) kpre[k + 1] = kpre[k] + c->pairs[koff + k].npp;
        // largest ket pair factor of a segment [k0, k1): kets are sorted by
        // molecule, so a running maximum restarted at each segment start
        auto seg_qmax = [&](int k0, int k1) {
            double m = 0.0;
            for (int k = k0; k < k1; ++k) m = std::max(m, c->pairs[koff + k].q);
            return m;
        };
        v.clear();
        for (int P = 0; P < nb; ++P) {
            const ShellPair& bp = c->pairs[boff + P];
            const int npp = bp.npp;
            // the kets of the bra pair's molecule (up to the bra pair itself
            // in the triangular classes)
            const int kfirst = c->kr_first[pck][bp.mol];
            const int kend = tri ? P + 1 : c->kr_end[pck][bp.mol];
            for (int k0 = kfirst; k0 < kend; k0 += seg) {
                const int k1 = std::min(kend, k0 + seg);
                const double kc = kpre[k1] - kpre[k0];
                for (int ip = 0; ip < npp; ip += PSLICE) {
                    const int ip1 = std::min(npp, ip + PSLICE);
                    v.push_back({(ip1 - ip) * kc, make_int4(P, ip | (ip1 << 16), k0, k1)});
                }
            }
        }
        std::stable_sort(v.begin(), v.end(), [](const std::pair<double, int4>& x, const std::pair<double, int4>& y) {
            return x.first > y.first;
        });
        for (auto& e : v) {
            all.push_back(e.second);
            const ShellPair& bp = c->pairs[boff + e.second.x];
            const double qb = bp.q * seg_qmax(e.second.z, e.second.w);
            float qf = (float)qb;
            if ((double)qf < qb) qf = nextafterf(qf, HUGE_VALF);
            allb.push_back(UnitBound{qf, bp.mol});
        }
    }
    c->unit_off[NQCLASS] = (int64_t)all.size();
    HIP_TRY(c->d_units.reserve(std::max<size_t>(1, all.size())));
    HIP_TRY(c->d_ubound.reserve(std::max<size_t>(1, allb.size())));
    if (!all.empty()) {
        HIP_TRY(hipMemcpyAsync(c->d_units.p, all.data(), sizeof(int4) * all.size(), hipMemcpyHostToDevice, c->st));
        HIP_TRY(hipMemcpyAsync(c->d_ubound.p, allb.data(), sizeof(UnitBound) * allb.size(), hipMemcpyHostToDevice,
                               c->st));
    }
    HIP_TRY(hipStreamSynchronize(c->st));
    c->units_ok = true;
    return MMD_OK;
}

void class_of(int qc, int* pcb, int* pck)
{
    int b = 0;
    while ((b + 1) * (b + 2) / 2 <= qc) ++b;
    *pcb = b;
    *pck = qc - b * (b + 1) / 2;
}

int grid_for(int64_t items)
{
    int64_t g = (items + 255) / 256;
    if (g < 1) g = 1;
    if (g > 2048) g = 2048;
    return (int)g;
}

ClassArgs base_args(mmd_ctx* c)
{
    ClassArgs A;
    memset(&A, 0, sizeof(A));
    A.pairs = c->d_pairs.p;
    A.pp = c->d_pp.p;
    A.boys = c->d_boys.p;
    A.boysf = c->d_boysf.p;
    A.gcomp = c->d_gcomp.p;
    A.N = c->nbf;
    return A;
}

// Device memory a Fock build works in besides its inputs and outputs: work-unit
// counters, survivor counts and statistics, the shell-block maxima of |dP|,
// and the survivor lists of the classes that run on the wavefront-per-quartet
// kernel (class qc at surv + surv_off[qc]).  The context's buffers, or a
// caller-free static array (mmd_fock_delta_noalloc).
struct FockScratch {
    int* rowctr = nullptr;                 // 2 * NQCLASS
    int* count = nullptr;                  // 2 * NQCLASS: survivor-list lengths (FP64 lists, then FP32)
    // NPRIM_SLOTS counters: primitive quartets (FP64 tier), shell quartets of
    // the row kernels (FP64, FP32 tier), primitive quartets (FP32 tier)
    unsigned long long* nprim = nullptr;
    double* Dsh = nullptr;                 // nshell^2
    const double* mol_dmax = nullptr;      // batches: per-molecule max |dP| (units of molecules at 0 skipped)
    int2* surv = nullptr;
    int64_t surv_off[NQCLASS] = {0};       // survivor lists, packed: wave classes only
    int64_t surv_cap = -1;                 // entries available at surv; < 0: grow the context's buffer
    bool allow_measure = true;             // may create events and re-measure the stream plan
    bool profile_events = false;           // with profiling on, time the classes even without allow_measure
};

// Launch the Fock kernels of every class.  Classes with a row kernel screen
// inline; the wavefront-per-quartet classes (and any class whose bra rows do
// not fit in LDS) are screened into survivor lists first.
int run_fock_classes(mmd_ctx* c, ClassArgs A, double tol, FockScratch& W0)
{
    const bool tier32 = c->fp32_below > tol;
    if (!c->units_ok) {
        if (!W0.allow_measure) return fail(MMD_ERR_STATE, "work units are built by mmd_schwarz / mmd_set_schwarz");
        const int rc = build_units(c);
        if (rc) return rc;
    }
    bool row_kernel[NQCLASS], kj_lds[NQCLASS];
    int maxpp[NQCLASS];
    // a build that times the row or the matrix-core variant of the classes
    // that have both, while some of them are undecided
    bool tuning = false;
    if (c->impl_mode == 2 && W0.allow_measure)
        for (int qc = 0; qc < NQCLASS; ++qc) {
            int la, lb, lc, ld;
            qclass_decode(qc, &la, &lb, &lc, &ld);
            if (c->qc_total[qc] && fock_class_has_both(la, lb, lc, ld) && c->impl_choice[qc] < 0) tuning = true;
        }
    for (int qc = 0; qc < NQCLASS; ++qc) {
        int la, lb, lc, ld, pcb, pck;
        qclass_decode(qc, &la, &lb, &lc, &ld);
        class_of(qc, &pcb, &pck);
        maxpp[qc] = 1;
        for (int P = c->pc_off[pcb]; P < c->pc_off[pcb + 1]; ++P) maxpp[qc] = std::max(maxpp[qc], c->pairs[P].npp);
        bool mfma = fock_class_is_wave(la, lb, lc, ld);
        if (mfma && fock_class_has_both(la, lb, lc, ld)) {
            if (c->impl_mode == 0) mfma = false;
            else if (c->impl_mode == 2 && c->impl_choice[qc] >= 0) mfma = c->impl_choice[qc] == 1;
            else if (c->impl_mode == 2 && tuning) mfma = c->tune_phase == 1;
        }
        // the ket-side Coulomb buffer where it fits; the row kernel without it
        // where only that fits; else the wavefront-per-quartet kernel
        kj_lds[qc] = fock_class_lds(qc, 0, c->rw, maxpp[qc], 1) <= c->max_lds;
        row_kernel[qc] = !mfma && (kj_lds[qc] || fock_class_lds(qc, 0, c->rw, maxpp[qc], 0) <= c->max_lds);
    }
    c->tune_measuring = tuning;
    {
        // survivor lists only for the classes that use them, packed; with the
        // FP32 tier each class has a second list right after the first
        int64_t off = 0;
        for (int qc = 0; qc < NQCLASS; ++qc) {
            W0.surv_off[qc] = off;
            if (c->qc_total[qc] && !row_kernel[qc]) off += (tier32 ? 2 : 1) * c->qc_total[qc];
        }
        if (W0.surv_cap < 0) {
            HIP_TRY(c->d_surv.reserve((size_t)std::max<int64_t>(off, 1)));
            W0.surv = c->d_surv.p;
        } else if (off > W0.surv_cap) {
            return fail(MMD_ERR_ARG, "survivor lists need %lld entries, the scratch holds %lld", (long long)off,
                        (long long)W0.surv_cap);
        }
    }
    HIP_TRY(launch_fill_int(2 * NQCLASS, W0.rowctr, 0, c->st));
    HIP_TRY(launch_fill_int(2 * NQCLASS, W0.count, 0, c->st));
    HIP_TRY(launch_fill_int(2 * NPRIM_SLOTS, reinterpret_cast<int*>(W0.nprim), 0, c->st));
    // Classes run concurrently on the main and side streams (longest first,
    // each to the least loaded stream), so that the small classes fill the
    // GPU beside the large ones.  Profiling mode 1 brackets every class with
    // events on its own stream (times include sharing the GPU with the other
    // classes); mode 2 runs the classes one after the other on the main
    // stream, so that each class's events time it alone.
    double pc_npp[NPAIRCLASS] = {0};
    for (int pc = 0; pc < NPAIRCLASS; ++pc)
        for (int P = c->pc_off[pc]; P < c->pc_off[pc + 1]; ++P) pc_npp[pc] += c->pairs[P].npp;
    int order[NQCLASS], nord = 0;
    double cost[NQCLASS] = {0};
    for (int qc = 0; qc < NQCLASS; ++qc) {
        if (c->qc_total[qc] == 0) continue;
        int pcb, pck, la, lb, lc, ld;
        class_of(qc, &pcb, &pck);
        qclass_decode(qc, &la, &lb, &lc, &ld);
        const double prim = pc_npp[pcb] * pc_npp[pck] * (pcb == pck ? 0.5 : 1.0);
        const int L = la + lb + lc + ld;
        cost[qc] = prim * (40.0 + 4.0 * nherm(L) + (double)ncart(la) * ncart(lb) * ncart(lc) * ncart(ld));
        order[nord++] = qc;
    }
    const int nstream = c->profiling == 2 ? 1 : c->nstreams;
    // Plan on the measured class durations once they exist: the model above
    // misses the latency floor of the small classes (150-300 us each at
    // (H2O)8/cc-pVDZ, whatever their work), which left the longest stream
    // twice as long as the others.  One build in PLAN_EVERY measures again.
    const bool measure = (W0.allow_measure && c->profiling == 0 && nstream > 1 && c->plan_countdown <= 0) || tuning;
    bool measured = nord > 0;
    for (int oi = 0; oi < nord; ++oi) measured = measured && c->plan_ms[order[oi]] > 0.0;
    if (measured)
        for (int oi = 0; oi < nord; ++oi) cost[order[oi]] = c->plan_ms[order[oi]];
    if ((measure || (c->profiling && W0.profile_events)) && !c->evc[0])
        for (int i = 0; i < 3 * NQCLASS; ++i) HIP_TRY(hipEventCreate(&c->evc[i]));
    c->plan_measuring = measure;
    std::sort(order, order + nord, [&](int x, int y) { return cost[x] > cost[y]; });
    hipStream_t streams[1 + mmd_ctx::NSIDE] = {c->st};
    for (int i = 0; i < mmd_ctx::NSIDE; ++i) streams[1 + i] = c->side[i];
    double load[1 + mmd_ctx::NSIDE] = {0};
    bool used[1 + mmd_ctx::NSIDE] = {true};
    if (nstream > 1) HIP_TRY(hipEventRecord(c->evfork, c->st));
    for (int oi = 0; oi < nord; ++oi) {
        const int qc = order[oi];
        int si = 0;
        for (int i = 1; i < nstream; ++i)
            if (load[i] < load[si]) si = i;
        load[si] += cost[qc];
        hipStream_t sq = streams[si];
        if (!used[si]) {
            HIP_TRY(hipStreamWaitEvent(sq, c->evfork, 0));
            used[si] = true;
        }
        int pcb, pck, la, lb, lc, ld;
        class_of(qc, &pcb, &pck);
        qclass_decode(qc, &la, &lb, &lc, &ld);
        const int nb = c->pc_off[pcb + 1] - c->pc_off[pcb];
        const int nk = c->pc_off[pck + 1] - c->pc_off[pck];
        const bool events = measure || (c->profiling && (W0.allow_measure || W0.profile_events));
        if (events) HIP_TRY(hipEventRecord(c->evc[3 * qc], sq));
        if (row_kernel[qc]) {
            RowArgs R;
            R.boff = c->pc_off[pcb];
            R.nb = nb;
            R.koff = c->pc_off[pck];
            R.nk = nk;
            R.tri = pcb == pck ? 1 : 0;
            R.nshell = (int)c->shells.size();
            R.Dsh = W0.Dsh;
            R.mol_dmax = W0.mol_dmax;
            R.row_counter = W0.rowctr + qc;
            R.maxpp = maxpp[qc];
            R.kj = kj_lds[qc] ? 1 : 0;
            R.kslice = c->ket_slice[qc];
            R.seg = c->unit_seg[qc];
            R.rw = c->rw;
            R.units = c->d_units.p + c->unit_off[qc];
            R.ubound = c->d_ubound.p + c->unit_off[qc];
            const int64_t units = c->unit_off[qc + 1] - c->unit_off[qc];
            R.nunits = (int)units;
            if (units == 0) {
                if (events) {
                    HIP_TRY(hipEventRecord(c->evc[3 * qc + 1], sq));
                    HIP_TRY(hipEventRecord(c->evc[3 * qc + 2], sq));
                }
                continue;
            }
            R.stat_nquart = W0.nprim + NQCLASS + qc;
            R.stat_nprim = c->profiling ? W0.nprim + qc : nullptr;
            R.tier_lo = tier32 ? c->fp32_below : 0.0;
            R.tier_hi = HUGE_VAL;
            const int grid = (int)std::min<int64_t>(units, 2048);
            HIP_TRY(launch_fock_class(qc, 0, A, R, grid, sq));
            if (events) HIP_TRY(hipEventRecord(c->evc[3 * qc + 1], sq));
            if (tier32) {
                // FP32 tier: the same units, quartets with tol <= bound < fp32_below
                R.tier_lo = 0.0;
                R.tier_hi = c->fp32_below;
                R.row_counter = W0.rowctr + NQCLASS + qc;
                R.stat_nquart = W0.nprim + 2 * NQCLASS + qc;
                R.stat_nprim = c->profiling ? W0.nprim + 3 * NQCLASS + qc : nullptr;
                // the second traversal holds the many small-bound quartets
                // of a unit: longer ket slices keep its items few ((H2O)16
                // cc-pVDZ tier build 9.3 ms with the FP64 pass's slices of 2,
                // 8.8 ms with 8)
                R.kslice = std::max(R.kslice, 8);
                HIP_TRY(launch_fock_class(qc, 1, A, R, grid, sq));
            }
        } else {
            // one screening pass; with the FP32 tier the quartets with bound
            // in [tol, fp32_below) go to a second list and run in FP32
            int2* surv = W0.surv + W0.surv_off[qc];
            int2* surv32 = tier32 ? surv + c->qc_total[qc] : nullptr;
            HIP_TRY(launch_screen_class(c->pc_off[pcb], nb, c->pc_off[pck], nk, pcb == pck ? 1 : 0, c->d_pairs.p,
                                        W0.Dsh, (int)c->shells.size(), tol, 0, surv, W0.count + qc, sq,
                                        c->fp32_below, surv32, W0.count + NQCLASS + qc));
            ClassArgs W = A;
            W.quart = surv;
            W.count = W0.count + qc;
            if (c->profiling)
                HIP_TRY(launch_count_prims(W.quart, W.count, c->d_pairs.p, W0.nprim + qc, sq));
            HIP_TRY(launch_fock_wave_class(la, lb, lc, ld, 0, W, c->qc_total[qc], sq));
            if (events) HIP_TRY(hipEventRecord(c->evc[3 * qc + 1], sq));
            if (tier32) {
                W.quart = surv32;
                W.count = W0.count + NQCLASS + qc;
                if (c->profiling)
                    HIP_TRY(launch_count_prims(W.quart, W.count, c->d_pairs.p, W0.nprim + 3 * NQCLASS + qc, sq));
                HIP_TRY(launch_fock_wave_class(la, lb, lc, ld, 1, W, c->qc_total[qc], sq));
            }
        }
        if (events) HIP_TRY(hipEventRecord(c->evc[3 * qc + 2], sq));
    }
    for (int i = 1; i < nstream; ++i) {
        if (!used[i]) continue;
        HIP_TRY(hipEventRecord(c->evjoin[i - 1], streams[i]));
        HIP_TRY(hipStreamWaitEvent(c->st, c->evjoin[i - 1], 0));
    }
    for (int qc = 0; qc < NQCLASS; ++qc) c->last_row_kernel[qc] = row_kernel[qc];
    return MMD_OK;
}

// Per-class statistics and event times of the last build (profiling on),
// also added to the sums since profiling was turned on.  Called after the
// build's end event, so the readback is not timed.
int collect_profile(mmd_ctx* c)
{
    int h[2 * NQCLASS];
    unsigned long long hp[NPRIM_SLOTS];
    HIP_TRY(hipMemcpyAsync(h, c->d_count.p, sizeof(int) * 2 * NQCLASS, hipMemcpyDeviceToHost, c->st));
    HIP_TRY(hipMemcpyAsync(hp, c->d_nprim.p, sizeof(hp), hipMemcpyDeviceToHost, c->st));
    HIP_TRY(hipStreamSynchronize(c->st));
    for (int qc = 0; qc < NQCLASS; ++qc) {
        if (!c->qc_total[qc]) {
            c->st_nquart[qc] = c->st_nprim[qc] = 0;
            c->st_ms[qc] = 0.0;
            continue;
        }
        const int64_t q64 = c->last_row_kernel[qc] ? (int64_t)hp[NQCLASS + qc] : (int64_t)h[qc];
        const int64_t q32 = c->last_row_kernel[qc] ? (int64_t)hp[2 * NQCLASS + qc] : (int64_t)h[NQCLASS + qc];
        c->st_nquart[qc] = q64 + q32;
        c->st_nprim[qc] = (int64_t)(hp[qc] + hp[3 * NQCLASS + qc]);
        float ms = 0.f, ms64 = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, c->evc[3 * qc], c->evc[3 * qc + 2]));
        HIP_TRY(hipEventElapsedTime(&ms64, c->evc[3 * qc], c->evc[3 * qc + 1]));
        c->st_ms[qc] = ms;
        ProfileSum& a = c->psum[qc];
        a.nquart[0] += q64;
        a.nquart[1] += q32;
        a.nprim[0] += (int64_t)hp[qc];
        a.nprim[1] += (int64_t)hp[3 * NQCLASS + qc];
        a.ms[0] += ms64;
        a.ms[1] += ms - ms64;
        a.builds += 1;
    }
    return MMD_OK;
}

}  // namespace

namespace {
// The Schwarz matrix (pq|pq) is in d_schwarz: its square roots for the
// screen, and the sorted shell-pair tables with their factors
// q = max sqrt((ab|ab)) from the host view diag.
int finish_schwarz(mmd_ctx* ctx, const DiagView& diag)
{
    const size_t nn = (size_t)ctx->nbf * ctx->nbf;
    HIP_TRY(launch_sqrt_clamped(nn, ctx->d_schwarz.p, ctx->d_sqS.p, ctx->st));
    HIP_TRY(hipStreamSynchronize(ctx->st));
    ctx->schwarz_ok = true;
    if (!ctx->shells_ok) return MMD_OK;   // only the per-function paths are available
    const int rc = build_pairs(ctx, &diag);   // pair factors q from the diagonal
    if (rc) return rc;
    // the Fock build's work units belong to the pair table: built here, at
    // set-up, so that a build (mmd_fock_delta_noalloc) allocates nothing
    return build_units(ctx);
}

// Install a Schwarz diagonal (N*N, (pq|pq), host) and rebuild the pair tables.
int set_schwarz(mmd_ctx* ctx, const std::vector<double>& h)
{
    const int N = ctx->nbf;
    const size_t nn = (size_t)N * N;
    HIP_TRY(hipMemcpyAsync(ctx->d_schwarz.p, h.data(), sizeof(double) * nn, hipMemcpyHostToDevice, ctx->st));
    DiagView v;
    v.p = h.data();
    v.row.resize(N);
    for (int i = 0; i < N; ++i) v.row[i] = (int64_t)i * N;
    return finish_schwarz(ctx, v);
}

}  // namespace

namespace mmd {
void set_last_error(const char* msg) { g_last_error = msg; }
}  // namespace mmd

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* mmd_last_error(void) { return g_last_error.c_str(); }

int mmd_version(void) { return 1; }

int mmd_device_count(int* count)
{
    if (!count) return fail(MMD_ERR_ARG, "null pointer");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return MMD_OK;
}

int mmd_ctx_create(mmd_ctx** out, int device_id)
{
    if (!out) return fail(MMD_ERR_ARG, "null output pointer");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(MMD_ERR_DEVICE, "no HIP device available");
    if (device_id < 0 || device_id >= n) return fail(MMD_ERR_ARG, "device %d out of range (%d devices)", device_id, n);
    HIP_TRY(hipSetDevice(device_id));
    mmd_ctx* c = new (std::nothrow) mmd_ctx();
    if (!c) return fail(MMD_ERR_NOMEM, "context allocation failed");
    c->device = device_id;
    for (int qc = 0; qc < NQCLASS; ++qc) {
        c->impl_choice[qc] = -1;
        c->ket_slice[qc] = KSLICE_DEFAULT;
    }
    // Classes on more streams than the process has hardware queues would
    // share queues in an order the longest-first plan does not see: spread
    // them over min(8, GPU_MAX_HW_QUEUES) streams (HIP's default is 4).  The
    // variable is read once, at the first context of the process, so that a
    // later change of the environment (after HIP read it) does not matter.
    {
        static const int hw_queues = [] {
            const char* hq = getenv("GPU_MAX_HW_QUEUES");
            const int q = hq ? atoi(hq) : 4;
            return q > 0 ? q : 4;
        }();
        c->nstreams = std::max(1, std::min(hw_queues, 1 + mmd_ctx::NSIDE));
    }
    int lds_max = 0;
    if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, device_id) == hipSuccess &&
        lds_max > 0)
        c->max_lds = std::min<size_t>((size_t)lds_max, 65536);
    c->max_lds_device = c->max_lds;
    if (lds_max > 0) c->lds_per_block = (size_t)lds_max;
    hipError_t e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->evfork, hipEventDisableTiming);
    for (int i = 0; i < mmd_ctx::NSIDE && e == hipSuccess; ++i) {
        e = hipStreamCreateWithFlags(&c->side[i], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&c->evjoin[i], hipEventDisableTiming);
    }
    std::vector<double> tab;
    boys_build_table(tab);
    if (e == hipSuccess) e = c->d_boys.reserve(tab.size());
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->d_boys.p, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice, c->st);
    std::vector<float> tabf(tab.begin(), tab.end());
    if (e == hipSuccess) e = c->d_boysf.reserve(tabf.size());
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->d_boysf.p, tabf.data(), sizeof(float) * tabf.size(), hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) {
        mmd_ctx_destroy(c);
        return fail(MMD_ERR_DEVICE, "context setup: %s", hipGetErrorString(e));
    }
    // this device's address of the static scratch (once per device)
    const int rc = scratch_for_device(device_id, &c->noalloc_scratch);
    if (rc) {
        mmd_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return MMD_OK;
}

void mmd_ctx_destroy(mmd_ctx* ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->st) (void)hipStreamSynchronize(ctx->st);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    for (int i = 0; i < 3 * NQCLASS; ++i)
        if (ctx->evc[i]) (void)hipEventDestroy(ctx->evc[i]);
    for (int i = 0; i < mmd_ctx::NSIDE; ++i) {
        if (ctx->side[i]) (void)hipStreamSynchronize(ctx->side[i]);
        if (ctx->evjoin[i]) (void)hipEventDestroy(ctx->evjoin[i]);
        if (ctx->side[i]) (void)hipStreamDestroy(ctx->side[i]);
    }
    if (ctx->evfork) (void)hipEventDestroy(ctx->evfork);
    std::lock_guard<std::recursive_mutex> lk(device_alloc_mutex());
    if (ctx->blas) (void)rocblas_destroy_handle(ctx->blas);
    hipStream_t st = ctx->st;
    delete ctx;  // frees device buffers
    if (st) (void)hipStreamDestroy(st);
}

int mmd_set_basis(mmd_ctx* ctx, int nbf, const double* origin, const int32_t* lmn, const int32_t* nprim,
                  const double* exps, const double* weights)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (nbf <= 0 || !origin || !lmn || !nprim || !exps || !weights) return fail(MMD_ERR_ARG, "invalid basis arguments");
    ctx->bf.assign(nbf, BFDesc{});
    ctx->bf_exp.clear();
    ctx->bf_w.clear();
    for (int i = 0; i < nbf; ++i) {
        BFDesc& b = ctx->bf[i];
        b.x = origin[3 * i];
        b.y = origin[3 * i + 1];
        b.z = origin[3 * i + 2];
        b.l = lmn[3 * i];
        b.m = lmn[3 * i + 1];
        b.n = lmn[3 * i + 2];
        if (b.l < 0 || b.m < 0 || b.n < 0) return fail(MMD_ERR_ARG, "negative angular exponent");
        if (b.l + b.m + b.n > MAX_L)
            return fail(MMD_ERR_UNSUPPORTED, "basis function %d has l=%d; this build supports l <= %d", i,
                        b.l + b.m + b.n, MAX_L);
        if (nprim[i] <= 0) return fail(MMD_ERR_ARG, "basis function %d has no primitives", i);
        b.np = nprim[i];
        b.p0 = (int)ctx->bf_exp.size();
        b.pad = 0;
        for (int p = 0; p < b.np; ++p) {
            const double e = exps[b.p0 + p];
            if (!(e > 0.0)) return fail(MMD_ERR_ARG, "non-positive exponent");
            ctx->bf_exp.push_back(e);
            ctx->bf_w.push_back(weights[b.p0 + p]);
        }
    }
    ctx->nbf = nbf;
    ctx->pairs_ok = false;
    ctx->schwarz_ok = false;
    ctx->nmol = 1;
    ctx->mol_f0.assign(1, 0);
    ctx->mol_n.assign(1, nbf);
    ctx->mol_atoms.clear();
    ctx->npl = 0;
    ctx->rw = nbf;
    HIP_TRY(ctx->d_bf.reserve(nbf));
    HIP_TRY(ctx->d_bf_exp.reserve(ctx->bf_exp.size()));
    HIP_TRY(ctx->d_bf_w.reserve(ctx->bf_w.size()));
    HIP_TRY(hipMemcpyAsync(ctx->d_bf.p, ctx->bf.data(), sizeof(BFDesc) * nbf, hipMemcpyHostToDevice, ctx->st));
    HIP_TRY(hipMemcpyAsync(ctx->d_bf_exp.p, ctx->bf_exp.data(), sizeof(double) * ctx->bf_exp.size(),
                           hipMemcpyHostToDevice, ctx->st));
    HIP_TRY(hipMemcpyAsync(ctx->d_bf_w.p, ctx->bf_w.data(), sizeof(double) * ctx->bf_w.size(), hipMemcpyHostToDevice,
                           ctx->st));
    HIP_TRY(hipStreamSynchronize(ctx->st));
    build_shells(ctx);
    if (ctx->shells_ok) {
        rc = upload_shells(ctx);
        if (rc) return rc;
    }
    const size_t nn = (size_t)nbf * nbf;
    HIP_TRY(ctx->d_D.reserve(nn));
    HIP_TRY(ctx->d_Dim.reserve(nn));
    HIP_TRY(ctx->d_absD.reserve(nn));
    HIP_TRY(ctx->d_G.reserve(nn));
    HIP_TRY(ctx->d_Gout.reserve(nn));
    HIP_TRY(ctx->d_schwarz.reserve(nn));
    HIP_TRY(ctx->d_sqS.reserve(nn));
    const size_t ns = ctx->shells.size();
    HIP_TRY(ctx->d_Dsh.reserve(ns * ns));
    return MMD_OK;
}

int mmd_set_atoms(mmd_ctx* ctx, int natom, const double* Z, const double* xyz)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (natom < 0 || (natom > 0 && (!Z || !xyz))) return fail(MMD_ERR_ARG, "invalid atoms");
    std::vector<double> a(4 * (size_t)std::max(natom, 1));
    for (int i = 0; i < natom; ++i) {
        a[4 * i] = Z[i];
        a[4 * i + 1] = xyz[3 * i];
        a[4 * i + 2] = xyz[3 * i + 1];
        a[4 * i + 3] = xyz[3 * i + 2];
    }
    HIP_TRY(ctx->d_atoms.reserve(a.size()));
    HIP_TRY(hipMemcpyAsync(ctx->d_atoms.p, a.data(), sizeof(double) * a.size(), hipMemcpyHostToDevice, ctx->st));
    HIP_TRY(hipStreamSynchronize(ctx->st));
    ctx->natom = natom;
    return MMD_OK;
}

int mmd_set_molecules(mmd_ctx* ctx, int nmol, const int32_t* nbf_per_mol, const int32_t* natom_per_mol)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (ctx->nbf == 0) return fail(MMD_ERR_STATE, "no basis set");
    if (nmol < 1 || !nbf_per_mol || !natom_per_mol) return fail(MMD_ERR_ARG, "invalid molecule arguments");
    int64_t sf = 0, sa = 0;
    for (int m = 0; m < nmol; ++m) {
        if (nbf_per_mol[m] < 1 || natom_per_mol[m] < 0) return fail(MMD_ERR_ARG, "molecule %d is empty", m);
        sf += nbf_per_mol[m];
        sa += natom_per_mol[m];
    }
    if (sf != ctx->nbf) return fail(MMD_ERR_ARG, "molecule sizes add up to %lld functions, the basis has %d",
                                    (long long)sf, ctx->nbf);
    if (sa != ctx->natom) return fail(MMD_ERR_ARG, "molecule atom counts add up to %lld, %d atoms are set",
                                      (long long)sa, ctx->natom);
    ctx->nmol = nmol;
    ctx->mol_f0.assign(nmol, 0);
    ctx->mol_n.assign(nmol, 0);
    ctx->rw = 0;
    std::vector<int2> pl;
    for (int m = 0, f = 0; m < nmol; ++m) {
        ctx->mol_f0[m] = f;
        ctx->mol_n[m] = nbf_per_mol[m];
        ctx->rw = std::max(ctx->rw, nbf_per_mol[m]);
        for (int i = f; i < f + nbf_per_mol[m]; ++i) {
            ctx->bf[i].pad = m;
            for (int j = f; j <= i; ++j) pl.push_back(make_int2(i, j));
        }
        f += nbf_per_mol[m];
    }
    ctx->mol_atoms.clear();
    for (int m = 0, a0 = 0; m < nmol; a0 += natom_per_mol[m], ++m) ctx->mol_atoms.push_back(make_int2(a0, natom_per_mol[m]));
    ctx->npl = (int64_t)pl.size();
    HIP_TRY(hipMemcpyAsync(ctx->d_bf.p, ctx->bf.data(), sizeof(BFDesc) * ctx->nbf, hipMemcpyHostToDevice, ctx->st));
    HIP_TRY(ctx->d_plist.reserve(pl.size()));
    HIP_TRY(hipMemcpyAsync(ctx->d_plist.p, pl.data(), sizeof(int2) * pl.size(), hipMemcpyHostToDevice, ctx->st));
    HIP_TRY(ctx->d_mol_atoms.reserve(ctx->mol_atoms.size()));
    HIP_TRY(hipMemcpyAsync(ctx->d_mol_atoms.p, ctx->mol_atoms.data(), sizeof(int2) * ctx->mol_atoms.size(),
                           hipMemcpyHostToDevice, ctx->st));
    HIP_TRY(hipStreamSynchronize(ctx->st));
    build_shells(ctx);
    if (ctx->shells_ok) {
        rc = upload_shells(ctx);
        if (rc) return rc;
    }
    ctx->pairs_ok = false;
    ctx->schwarz_ok = false;
    ctx->units_ok = false;
    return MMD_OK;
}

int mmd_rhf_gradient(mmd_ctx* ctx, const double* P, const double* W, const int32_t* bf_atom, double* grad)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    const int N = ctx->nbf, natom = ctx->natom;
    if (N == 0) return fail(MMD_ERR_STATE, "no basis set");
    if (natom < 1) return fail(MMD_ERR_STATE, "no nuclei (mmd_set_atoms)");
    if (ctx->nmol > 1) return fail(MMD_ERR_UNSUPPORTED, "the gradient takes one molecule, not a batch");
    if (natom > 4096) return fail(MMD_ERR_UNSUPPORTED, "the gradient takes up to 4096 atoms");
    if (!P || !W || !bf_atom || !grad) return fail(MMD_ERR_ARG, "null pointer");
    for (int i = 0; i < N; ++i)
        if (bf_atom[i] < 0 || bf_atom[i] >= natom) return fail(MMD_ERR_ARG, "basis function %d: atom out of range", i);
    const size_t nn = (size_t)N * N;
    HIP_TRY(ctx->g_PW.reserve(2 * nn + 3 * (size_t)natom));
    HIP_TRY(ctx->g_atom.reserve(N));
    double* dP = ctx->g_PW.p;
    double* dW = dP + nn;
    double* dg = dW + nn;
    HIP_TRY(hipMemcpyAsync(dP, P, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->st));
    HIP_TRY(hipMemcpyAsync(dW, W, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->st));
    HIP_TRY(hipMemcpyAsync(ctx->g_atom.p, bf_atom, sizeof(int32_t) * N, hipMemcpyHostToDevice, ctx->st));
    HIP_TRY(launch_fill(3 * (size_t)natom, dg, 0.0, ctx->st));
    HIP_TRY(hipEventRecord(ctx->ev0, ctx->st));
    HIP_TRY(launch_rhf_gradient(N, ctx->d_bf.p, ctx->d_bf_exp.p, ctx->d_bf_w.p, natom, ctx->d_atoms.p, ctx->d_boys.p,
                                ctx->g_atom.p, dP, dW, dg, ctx->st));
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->st));
    HIP_TRY(hipMemcpyAsync(grad, dg, sizeof(double) * 3 * natom, hipMemcpyDeviceToHost, ctx->st));
    HIP_TRY(hipStreamSynchronize(ctx->st));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_grad_ms = ms;
    return MMD_OK;
}

int mmd_last_gradient_ms(mmd_ctx* ctx, double* ms)
{
    if (!ctx || !ms) return fail(MMD_ERR_ARG, "null pointer");
    *ms = ctx->last_grad_ms;
    return MMD_OK;
}

int mmd_one_electron(mmd_ctx* ctx, const double center[3], double* S, double* T, double* V, double* M, double* L)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (ctx->nbf == 0) return fail(MMD_ERR_STATE, "no basis set");
    if (!center || !S || !T || !V || !M || !L) return fail(MMD_ERR_ARG, "null pointer");
    const size_t nn = (size_t)ctx->nbf * ctx->nbf;
    DevBuf<double>& d = ctx->d_onee;   // kept between calls (no allocation per SCF)
    HIP_TRY(d.reserve(9 * nn));
    const bool several = ctx->nmol > 1;
    if (ctx->shells_ok) {
        if ((rc = one_electron_shells(ctx, center, d.p, d.p + nn, d.p + 2 * nn, d.p + 3 * nn, d.p + 6 * nn)))
            return rc;
    } else {
        HIP_TRY(launch_one_electron(ctx->nbf, ctx->d_bf.p, ctx->d_bf_exp.p, ctx->d_bf_w.p, ctx->natom,
                                    ctx->natom ? ctx->d_atoms.p : nullptr, center[0], center[1], center[2],
                                    ctx->d_boys.p, d.p, d.p + nn, d.p + 2 * nn, d.p + 3 * nn, d.p + 6 * nn, ctx->st,
                                    several ? ctx->d_plist.p : nullptr, ctx->npl,
                                    several ? ctx->d_mol_atoms.p : nullptr));
    }
    HIP_TRY(hipMemcpyAsync(S, d.p, sizeof(double) * nn, hipMemcpyDeviceToHost, ctx->st));
    HIP_TRY(hipMemcpyAsync(T, d.p + nn, sizeof(double) * nn, hipMemcpyDeviceToHost, ctx->st));
    HIP_TRY(hipMemcpyAsync(V, d.p + 2 * nn, sizeof(double) * nn, hipMemcpyDeviceToHost, ctx->st));
    HIP_TRY(hipMemcpyAsync(M, d.p + 3 * nn, sizeof(double) * 3 * nn, hipMemcpyDeviceToHost, ctx->st));
    HIP_TRY(hipMemcpyAsync(L, d.p + 6 * nn, sizeof(double) * 3 * nn, hipMemcpyDeviceToHost, ctx->st));
    HIP_TRY(hipStreamSynchronize(ctx->st));
    return MMD_OK;
}

int mmd_schwarz(mmd_ctx* ctx, double* diag)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (ctx->nbf == 0) return fail(MMD_ERR_STATE, "no basis set");
    const int N = ctx->nbf;
    const size_t nn = (size_t)N * N;
    std::vector<double> h(nn);
    HIP_TRY(ctx->d_schwarz.reserve(nn));
    if ((rc = schwarz_diagonal(ctx, ctx->d_schwarz.p))) return rc;
    HIP_TRY(hipMemcpyAsync(h.data(), ctx->d_schwarz.p, sizeof(double) * nn, hipMemcpyDeviceToHost, ctx->st));
    HIP_TRY(hipStreamSynchronize(ctx->st));
    if (diag) memcpy(diag, h.data(), sizeof(double) * nn);
    return set_schwarz(ctx, h);
}

int mmd_set_schwarz(mmd_ctx* ctx, const double* diag)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (ctx->nbf == 0) return fail(MMD_ERR_STATE, "no basis set");
    if (!diag) return fail(MMD_ERR_ARG, "null pointer");
    std::vector<double> h(diag, diag + (size_t)ctx->nbf * ctx->nbf);
    return set_schwarz(ctx, h);
}

namespace {
// The full (ij|kl) tensor from the shell-pair tables: every quartet class
// with bra pair class >= ket pair class, all canonical shell quartets of it
// (bra pair >= ket pair inside one pair class), every integral stored at its
// eight index permutations.  Every element of dst is written.
hipError_t eri_tensor_shells(mmd_ctx* c, double* dst)
{
    ClassArgs A = base_args(c);
    A.G = dst;
    size_t bps = 0;
    for (int qc = 0; qc < NQCLASS; ++qc) bps = std::max(bps, (size_t)tensor_class_bps(qc));
    DevBuf<double> table;
    hipError_t e = table.reserve(std::max<size_t>(1, (size_t)c->npp_total * bps));
    if (e != hipSuccess) return e;
    for (int qc = 0; qc < NQCLASS; ++qc) {
        int pcb, pck, la, lb, lc, ld;
        class_of(qc, &pcb, &pck);
        qclass_decode(qc, &la, &lb, &lc, &ld);
        const int boff = c->pc_off[pcb], nb = c->pc_off[pcb + 1] - boff;
        const int koff = c->pc_off[pck], nk = c->pc_off[pck + 1] - koff;
        if (nb == 0 || nk == 0) continue;
        const int tri = pcb == pck;
        if (la == 2 && lb == 2 && lc == 2 && ld == 2) {   // (dd|dd): the wavefront tensor kernel
            std::vector<int2> q;
            q.reserve((size_t)nb * (nb + 1) / 2);
            for (int P = 0; P < nb; ++P)
                for (int Q = 0; Q <= P; ++Q) q.push_back(make_int2(boff + P, koff + Q));
            const int n = (int)q.size();
            e = c->d_surv.reserve(q.size());
            if (e == hipSuccess) e = hipMemcpyAsync(c->d_surv.p, q.data(), sizeof(int2) * q.size(),
                                                    hipMemcpyHostToDevice, c->st);
            if (e == hipSuccess) e = hipMemcpyAsync(c->d_count.p + qc, &n, sizeof(int), hipMemcpyHostToDevice, c->st);
            if (e == hipSuccess) e = hipStreamSynchronize(c->st);   // q and n leave scope
            if (e != hipSuccess) return e;
            ClassArgs W = A;
            W.quart = c->d_surv.p;
            W.count = c->d_count.p + qc;
            e = launch_tensor_wave_class(W, n, c->st);
        } else {
            e = launch_tensor_class(qc, A, boff, nb, koff, nk, tri, table.p, dst, c->st);
        }
        if (e != hipSuccess) return e;
    }
    return hipStreamSynchronize(c->st);   // before the table is freed
}
}  // namespace

int mmd_eri_tensor(mmd_ctx* ctx, double* out, int out_on_device)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!out) return fail(MMD_ERR_ARG, "null output");
    if (ctx->nbf == 0) return fail(MMD_ERR_STATE, "no basis set");
    const size_t n4 = (size_t)ctx->nbf * ctx->nbf * ctx->nbf * ctx->nbf;
    DevBuf<double> tmp;
    double* dst = out;
    if (!out_on_device) {
        HIP_TRY(tmp.reserve(n4));
        dst = tmp.p;
    }
    if (ctx->shells_ok) {
        if (!ctx->pairs_ok) {
            rc = build_pairs(ctx, nullptr);
            if (rc) return rc;
        }
        HIP_TRY(eri_tensor_shells(ctx, dst));
    } else {
        HIP_TRY(launch_canonical_quartets(0, ctx->nbf, ctx->d_bf.p, ctx->d_bf_exp.p, ctx->d_bf_w.p, ctx->d_boys.p,
                                          nullptr, nullptr, nullptr, 0.0, dst, ctx->st));
    }
    if (!out_on_device) HIP_TRY(hipMemcpyAsync(out, dst, sizeof(double) * n4, hipMemcpyDeviceToHost, ctx->st));
    HIP_TRY(hipStreamSynchronize(ctx->st));
    return MMD_OK;
}

int mmd_eri_list(mmd_ctx* ctx, int64_t n, const int64_t* ijkl, double* out)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (ctx->nbf == 0) return fail(MMD_ERR_STATE, "no basis set");
    if (n < 0 || (n > 0 && (!ijkl || !out))) return fail(MMD_ERR_ARG, "invalid quartet list");
    if (n == 0) return MMD_OK;
    for (int64_t i = 0; i < 4 * n; ++i)
        if (ijkl[i] < 0 || ijkl[i] >= ctx->nbf) return fail(MMD_ERR_ARG, "basis-function index out of range");
    DevBuf<int64_t> d_idx;
    DevBuf<double> d_out;
    HIP_TRY(d_idx.reserve(4 * n));
    HIP_TRY(d_out.reserve(n));
    HIP_TRY(hipMemcpyAsync(d_idx.p, ijkl, sizeof(int64_t) * 4 * n, hipMemcpyHostToDevice, ctx->st));
    HIP_TRY(launch_eri_bf_list(n, d_idx.p, ctx->d_bf.p, ctx->d_bf_exp.p, ctx->d_bf_w.p, ctx->d_boys.p, d_out.p, ctx->st));
    HIP_TRY(hipMemcpyAsync(out, d_out.p, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->st));
    HIP_TRY(hipStreamSynchronize(ctx->st));
    return MMD_OK;
}

int mmd_fock_delta(mmd_ctx* ctx, const double* dP_re, const double* dP_im, double tol, int raw, double* G_re,
                   double* G_im, int on_device, int64_t* nsurv)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!ctx->schwarz_ok) return fail(MMD_ERR_STATE, "mmd_schwarz must be called before mmd_fock_delta");
    if (!dP_re || !G_re || (dP_im && !G_im)) return fail(MMD_ERR_ARG, "null pointer");
    if (!raw && dP_im) return fail(MMD_ERR_ARG, "the symmetrised build needs a real density difference");
    if (!raw && !ctx->pairs_ok) return fail(MMD_ERR_UNSUPPORTED, "%s", ctx->shells_err.c_str());
    const int N = ctx->nbf;
    const size_t nn = (size_t)N * N;
    const hipMemcpyKind k_in = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const hipMemcpyKind k_out = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    HIP_TRY(hipEventRecord(ctx->ev0, ctx->st));
    HIP_TRY(hipMemcpyAsync(ctx->d_D.p, dP_re, sizeof(double) * nn, k_in, ctx->st));
    if (dP_im) HIP_TRY(hipMemcpyAsync(ctx->d_Dim.p, dP_im, sizeof(double) * nn, k_in, ctx->st));
    HIP_TRY(launch_abs_density((int)nn, ctx->d_D.p, dP_im ? ctx->d_Dim.p : nullptr, ctx->d_absD.p, ctx->st));
    if (!raw) {
        HIP_TRY(launch_shell_block_max((int)ctx->shells.size(), ctx->d_sh_bf0.p, ctx->d_sh_nbf.p, N, ctx->d_absD.p,
                                       ctx->d_Dsh.p, ctx->st));
        ClassArgs A = base_args(ctx);
        A.absD = ctx->d_absD.p;
        A.sqS = ctx->d_sqS.p;
        A.tol = tol;
        A.D = ctx->d_D.p;
        A.G = ctx->d_G.p;
        HIP_TRY(launch_fill(nn, ctx->d_G.p, 0.0, ctx->st));
        HIP_TRY(ctx->d_rowctr.reserve(2 * NQCLASS));
        HIP_TRY(ctx->d_nprim.reserve(NPRIM_SLOTS));
        FockScratch W;
        W.rowctr = ctx->d_rowctr.p;
        W.count = ctx->d_count.p;
        W.nprim = ctx->d_nprim.p;
        W.Dsh = ctx->d_Dsh.p;
        rc = run_fock_classes(ctx, A, tol, W);
        if (rc) return rc;
        HIP_TRY(launch_symmetrize(N, ctx->d_G.p, ctx->d_Gout.p, ctx->st));
        HIP_TRY(hipMemcpyAsync(G_re, ctx->d_Gout.p, sizeof(double) * nn, k_out, ctx->st));
    } else {
        for (int part = 0; part < (dP_im ? 2 : 1); ++part) {
            HIP_TRY(launch_fill(nn, ctx->d_G.p, 0.0, ctx->st));
            HIP_TRY(launch_canonical_quartets(1, N, ctx->d_bf.p, ctx->d_bf_exp.p, ctx->d_bf_w.p, ctx->d_boys.p,
                                              part == 0 ? ctx->d_D.p : ctx->d_Dim.p, ctx->d_absD.p, ctx->d_sqS.p, tol,
                                              ctx->d_G.p, ctx->st));
            HIP_TRY(hipMemcpyAsync(part == 0 ? G_re : G_im, ctx->d_G.p, sizeof(double) * nn, k_out, ctx->st));
        }
    }
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->st));
    if (!raw) {
        // survivors: per-class lists (wave classes) + FP64 and FP32 row-kernel counts
        int h[2 * NQCLASS];
        unsigned long long hq[2 * NQCLASS];
        HIP_TRY(hipMemcpyAsync(h, ctx->d_count.p, sizeof(int) * 2 * NQCLASS, hipMemcpyDeviceToHost, ctx->st));
        HIP_TRY(hipMemcpyAsync(hq, ctx->d_nprim.p + NQCLASS, sizeof(hq), hipMemcpyDeviceToHost, ctx->st));
        HIP_TRY(hipStreamSynchronize(ctx->st));
        int64_t s = 0, s32 = 0;
        for (int i = 0; i < NQCLASS; ++i) {
            s += (int64_t)h[i] + (int64_t)h[NQCLASS + i] + (int64_t)hq[i] + (int64_t)hq[NQCLASS + i];
            s32 += (int64_t)h[NQCLASS + i] + (int64_t)hq[NQCLASS + i];
        }
        ctx->last_nsurv = s;
        ctx->last_nsurv_fp32 = s32;
        if (nsurv) *nsurv = s;
    } else if (nsurv) {
        *nsurv = -1;
    }
    HIP_TRY(hipStreamSynchronize(ctx->st));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_fock_ms = ms;
    if (!raw && ctx->plan_measuring) {
        for (int qc = 0; qc < NQCLASS; ++qc) {
            float cms = 0.f;
            if (ctx->qc_total[qc]) HIP_TRY(hipEventElapsedTime(&cms, ctx->evc[3 * qc], ctx->evc[3 * qc + 2]));
            ctx->plan_ms[qc] = cms > 0.f ? cms : 1e-3;
        }
        ctx->plan_measuring = false;
        ctx->plan_countdown = mmd_ctx::PLAN_EVERY;
    } else if (!raw) {
        --ctx->plan_countdown;
    }
    if (!raw && ctx->tune_measuring) {
        // a build of the implementation choice: record, and decide after both
        for (int qc = 0; qc < NQCLASS; ++qc) {
            float cms = 0.f;
            if (ctx->qc_total[qc]) HIP_TRY(hipEventElapsedTime(&cms, ctx->evc[3 * qc], ctx->evc[3 * qc + 2]));
            ctx->tune_ms[ctx->tune_phase][qc] = cms;
        }
        ctx->tune_measuring = false;
        if (++ctx->tune_phase == 2) {
            for (int qc = 0; qc < NQCLASS; ++qc) {
                int la, lb, lc, ld;
                qclass_decode(qc, &la, &lb, &lc, &ld);
                if (fock_class_has_both(la, lb, lc, ld))
                    ctx->impl_choice[qc] = ctx->tune_ms[1][qc] < ctx->tune_ms[0][qc] ? 1 : 0;
            }
            ctx->plan_countdown = 0;   // plan again on the chosen kernels
        }
    }
    if (ctx->profiling && !raw) return collect_profile(ctx);
    return MMD_OK;
}

// The symmetrised build of mmd_fock_delta (raw = 0, real dP, device
// pointers) with every work buffer in the static `scratch` of this device
// instead of the context's: |dP|, the shell-block maxima, the unsymmetrised
// accumulator, counters and survivor lists.  No allocation on the host or the
// device; the kernels are the same.
int mmd_fock_delta_noalloc(mmd_ctx* ctx, const double* dP, double tol, double* G, int64_t* nsurv)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!ctx->schwarz_ok) return fail(MMD_ERR_STATE, "mmd_schwarz must be called before mmd_fock_delta_noalloc");
    if (!ctx->pairs_ok) return fail(MMD_ERR_UNSUPPORTED, "%s", ctx->shells_err.c_str());
    if (!ctx->units_ok) return fail(MMD_ERR_STATE, "work units missing (set the Schwarz screen again)");
    if (!dP || !G) return fail(MMD_ERR_ARG, "null pointer");
    if (!ctx->noalloc_scratch) return fail(MMD_ERR_STATE, "no scratch address for this device");
    const int N = ctx->nbf;
    const size_t nn = (size_t)N * N;
    const size_t ns = ctx->shells.size();
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    unsigned char* base = static_cast<unsigned char*>(ctx->noalloc_scratch);
    size_t off = 0;
    double* absD = reinterpret_cast<double*>(base + off);
    off += up(sizeof(double) * nn);
    double* Gacc = reinterpret_cast<double*>(base + off);
    off += up(sizeof(double) * nn);
    double* Dsh = reinterpret_cast<double*>(base + off);
    off += up(sizeof(double) * ns * ns);
    int* rowctr = reinterpret_cast<int*>(base + off);
    off += up(sizeof(int) * 2 * NQCLASS);
    int* count = reinterpret_cast<int*>(base + off);
    off += up(sizeof(int) * 2 * NQCLASS);
    unsigned long long* nprim = reinterpret_cast<unsigned long long*>(base + off);
    off += up(sizeof(unsigned long long) * NPRIM_SLOTS);
    if (off > NOALLOC_SCRATCH_BYTES)
        return fail(MMD_ERR_ARG, "a basis of %d functions needs %zu bytes of scratch; the limit is %zu", N, off,
                    NOALLOC_SCRATCH_BYTES);
    FockScratch W;
    W.rowctr = rowctr;
    W.count = count;
    W.nprim = nprim;
    W.Dsh = Dsh;
    W.surv = reinterpret_cast<int2*>(base + off);
    W.surv_cap = (int64_t)((NOALLOC_SCRATCH_BYTES - off) / sizeof(int2));
    W.allow_measure = false;
    HIP_TRY(hipEventRecord(ctx->ev0, ctx->st));
    HIP_TRY(launch_abs_density((int)nn, dP, nullptr, absD, ctx->st));
    HIP_TRY(launch_shell_block_max((int)ns, ctx->d_sh_bf0.p, ctx->d_sh_nbf.p, N, absD, Dsh, ctx->st));
    ClassArgs A = base_args(ctx);
    A.absD = absD;
    A.sqS = ctx->d_sqS.p;
    A.tol = tol;
    A.D = dP;
    A.G = Gacc;
    HIP_TRY(launch_fill(nn, Gacc, 0.0, ctx->st));
    rc = run_fock_classes(ctx, A, tol, W);
    if (rc) return rc;
    HIP_TRY(launch_symmetrize(N, Gacc, G, ctx->st));
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->st));
    int h[2 * NQCLASS];
    unsigned long long hq[2 * NQCLASS];
    HIP_TRY(hipMemcpyAsync(h, count, sizeof(int) * 2 * NQCLASS, hipMemcpyDeviceToHost, ctx->st));
    HIP_TRY(hipMemcpyAsync(hq, nprim + NQCLASS, sizeof(hq), hipMemcpyDeviceToHost, ctx->st));
    HIP_TRY(hipStreamSynchronize(ctx->st));
    int64_t sum = 0, s32 = 0;
    for (int i = 0; i < NQCLASS; ++i) {
        sum += (int64_t)h[i] + (int64_t)h[NQCLASS + i] + (int64_t)hq[i] + (int64_t)hq[NQCLASS + i];
        s32 += (int64_t)h[NQCLASS + i] + (int64_t)hq[NQCLASS + i];
    }
    ctx->last_nsurv = sum;
    ctx->last_nsurv_fp32 = s32;
    if (nsurv) *nsurv = sum;
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_fock_ms = ms;
    return MMD_OK;
}

int mmd_jk_tensor(mmd_ctx* ctx, const double* T, int t_on_device, const double* P_re, const double* P_im,
                  double* J_re, double* J_im, double* K_re, double* K_im)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (ctx->nbf == 0) return fail(MMD_ERR_STATE, "no basis set");
    if (!T || !P_re || !J_re || !K_re || (P_im && (!J_im || !K_im))) return fail(MMD_ERR_ARG, "null pointer");
    const int N = ctx->nbf;
    const size_t nn = (size_t)N * N, n4 = nn * nn;
    DevBuf<double> dT, d;
    const double* t = T;
    if (!t_on_device) {
        HIP_TRY(dT.reserve(n4));
        HIP_TRY(hipMemcpyAsync(dT.p, T, sizeof(double) * n4, hipMemcpyHostToDevice, ctx->st));
        t = dT.p;
    }
    HIP_TRY(d.reserve(6 * nn));
    HIP_TRY(hipMemcpyAsync(d.p, P_re, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->st));
    if (P_im) HIP_TRY(hipMemcpyAsync(d.p + nn, P_im, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->st));
    HIP_TRY(launch_jk_tensor(N, t, d.p, P_im ? d.p + nn : nullptr, d.p + 2 * nn, d.p + 3 * nn, d.p + 4 * nn,
                             d.p + 5 * nn, ctx->st));
    HIP_TRY(hipMemcpyAsync(J_re, d.p + 2 * nn, sizeof(double) * nn, hipMemcpyDeviceToHost, ctx->st));
    HIP_TRY(hipMemcpyAsync(K_re, d.p + 4 * nn, sizeof(double) * nn, hipMemcpyDeviceToHost, ctx->st));
    if (P_im) {
        HIP_TRY(hipMemcpyAsync(J_im, d.p + 3 * nn, sizeof(double) * nn, hipMemcpyDeviceToHost, ctx->st));
        HIP_TRY(hipMemcpyAsync(K_im, d.p + 5 * nn, sizeof(double) * nn, hipMemcpyDeviceToHost, ctx->st));
    }
    HIP_TRY(hipStreamSynchronize(ctx->st));
    return MMD_OK;
}

int mmd_boys(mmd_ctx* ctx, int n, const double* m, const double* T, int table, double* out)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!m || !T || !out))) return fail(MMD_ERR_ARG, "invalid arguments");
    for (int i = 0; i < n; ++i) {
        if (!(m[i] >= 0.0) || !(T[i] >= 0.0)) return fail(MMD_ERR_ARG, "Boys arguments must be non-negative");
        if (table && (m[i] != floor(m[i]) || m[i] > MAX_LTOT))
            return fail(MMD_ERR_ARG, "tabulated Boys path takes integer m <= %d", MAX_LTOT);
    }
    if (n == 0) return MMD_OK;
    DevBuf<double> d;
    HIP_TRY(d.reserve(3 * (size_t)n));
    HIP_TRY(hipMemcpyAsync(d.p, m, sizeof(double) * n, hipMemcpyHostToDevice, ctx->st));
    HIP_TRY(hipMemcpyAsync(d.p + n, T, sizeof(double) * n, hipMemcpyHostToDevice, ctx->st));
    HIP_TRY(launch_boys_eval(n, d.p, d.p + n, table, ctx->d_boys.p, d.p + 2 * n, ctx->st));
    HIP_TRY(hipMemcpyAsync(out, d.p + 2 * n, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->st));
    HIP_TRY(hipStreamSynchronize(ctx->st));
    return MMD_OK;
}

int mmd_set_profiling(mmd_ctx* ctx, int on)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (on && !ctx->evc[0])
        for (int i = 0; i < 3 * NQCLASS; ++i) HIP_TRY(hipEventCreate(&ctx->evc[i]));
    if (on)
        for (auto& a : ctx->psum) a = ProfileSum();
    if (on < 0 || on > 2) return fail(MMD_ERR_ARG, "profiling mode must be 0, 1 or 2");
    ctx->profiling = on;
    return MMD_OK;
}

int mmd_fock_stats(mmd_ctx* ctx, int32_t* lclass, int64_t* nquart, int64_t* nprim, double* ms)
{
    if (!ctx) return fail(MMD_ERR_ARG, "null context");
    for (int qc = 0; qc < NQCLASS; ++qc) {
        if (lclass) qclass_decode(qc, &lclass[4 * qc], &lclass[4 * qc + 1], &lclass[4 * qc + 2], &lclass[4 * qc + 3]);
        if (nquart) nquart[qc] = ctx->st_nquart[qc];
        if (nprim) nprim[qc] = ctx->st_nprim[qc];
        if (ms) ms[qc] = ctx->st_ms[qc];
    }
    return NQCLASS;
}

int mmd_fock_profile_sums(mmd_ctx* ctx, int32_t* lclass, int64_t* nquart, int64_t* nprim, double* ms,
                          int64_t* builds)
{
    if (!ctx) return fail(MMD_ERR_ARG, "null context");
    for (int qc = 0; qc < NQCLASS; ++qc) {
        const ProfileSum& a = ctx->psum[qc];
        if (lclass) qclass_decode(qc, &lclass[4 * qc], &lclass[4 * qc + 1], &lclass[4 * qc + 2], &lclass[4 * qc + 3]);
        for (int t = 0; t < 2; ++t) {
            if (nquart) nquart[2 * qc + t] = a.nquart[t];
            if (nprim) nprim[2 * qc + t] = a.nprim[t];
            if (ms) ms[2 * qc + t] = a.ms[t];
        }
        if (builds) builds[qc] = a.builds;
    }
    return NQCLASS;
}

int mmd_set_ket_slices(mmd_ctx* ctx, const int32_t* slices, int32_t* current)
{
    if (!ctx) return fail(MMD_ERR_ARG, "null context");
    if (slices) {
        for (int qc = 0; qc < NQCLASS; ++qc)
            if (slices[qc] != 0 && (slices[qc] < 1 || slices[qc] > KSLICE_MAX))
                return fail(MMD_ERR_ARG, "class %d: ket slice %d outside 1..%d", qc, (int)slices[qc], KSLICE_MAX);
        for (int qc = 0; qc < NQCLASS; ++qc)
            if (slices[qc] != 0) ctx->ket_slice[qc] = slices[qc];
    }
    if (current)
        for (int qc = 0; qc < NQCLASS; ++qc) current[qc] = ctx->ket_slice[qc];
    return MMD_OK;
}

int mmd_set_class_impl(mmd_ctx* ctx, int mode, int32_t* choice)
{
    if (!ctx) return fail(MMD_ERR_ARG, "null context");
    if (mode < -1 || mode > 2) return fail(MMD_ERR_ARG, "mode must be 0 (row), 1 (matrix cores), 2 (measured) or -1");
    if (mode >= 0 && mode != ctx->impl_mode) {
        ctx->impl_mode = mode;
        for (int qc = 0; qc < NQCLASS; ++qc) ctx->impl_choice[qc] = -1;
        ctx->tune_phase = 0;
        ctx->plan_countdown = 0;
    }
    if (choice)
        for (int qc = 0; qc < NQCLASS; ++qc) {
            int la, lb, lc, ld;
            qclass_decode(qc, &la, &lb, &lc, &ld);
            const bool both = fock_class_has_both(la, lb, lc, ld);
            const bool mf = fock_class_is_wave(la, lb, lc, ld);
            choice[qc] = !mf ? 0 : (!both ? 1 : (ctx->impl_mode == 2 ? ctx->impl_choice[qc] : ctx->impl_mode));
        }
    return MMD_OK;
}

int mmd_set_precision(mmd_ctx* ctx, double fp32_below)
{
    if (!ctx) return fail(MMD_ERR_ARG, "null context");
    if (!(fp32_below >= 0.0) || fp32_below > 1.0) return fail(MMD_ERR_ARG, "fp32_below must be in [0, 1]");
    ctx->fp32_below = fp32_below;
    return MMD_OK;
}

int mmd_set_streams(mmd_ctx* ctx, int n)
{
    if (!ctx) return fail(MMD_ERR_ARG, "null context");
    if (n < 1 || n > 1 + mmd_ctx::NSIDE) return fail(MMD_ERR_ARG, "streams must be in [1, %d]", 1 + mmd_ctx::NSIDE);
    ctx->nstreams = n;
    return MMD_OK;
}

int mmd_set_lds_limit(mmd_ctx* ctx, int64_t bytes)
{
    if (!ctx) return fail(MMD_ERR_ARG, "null context");
    if (bytes < 0) return fail(MMD_ERR_ARG, "LDS limit must be >= 0");
    ctx->max_lds = bytes == 0 ? ctx->max_lds_device : std::min<size_t>((size_t)bytes, ctx->max_lds_device);
    return MMD_OK;
}

int mmd_fock_tiers(mmd_ctx* ctx, int64_t* nquart, int64_t* nquart_fp32)
{
    if (!ctx) return fail(MMD_ERR_ARG, "null context");
    if (nquart) *nquart = ctx->last_nsurv;
    if (nquart_fp32) *nquart_fp32 = ctx->last_nsurv_fp32;
    return MMD_OK;
}

}  // extern "C"

namespace {

// Row-major C (m x n) = op(A) op(B), op = transpose where tA / tB; the
// column-major BLAS computes C^T = op(B)^T op(A)^T.
int gemm_rm(mmd_ctx* c, bool tA, bool tB, int m, int n, int k, const double* A, int lda, const double* B, int ldb,
            double* C, int ldc)
{
    const double one = 1.0, zero = 0.0;
    const rocblas_status s = rocblas_dgemm(c->blas, tB ? rocblas_operation_transpose : rocblas_operation_none,
                                           tA ? rocblas_operation_transpose : rocblas_operation_none, n, m, k, &one,
                                           B, ldb, A, lda, &zero, C, ldc);
    if (s != rocblas_status_success) return fail(MMD_ERR_DEVICE, "rocblas_dgemm: status %d", (int)s);
    return MMD_OK;
}

// Solve the (n x n) system A x = b in place (Gaussian elimination with
// partial pivoting); false if singular.
bool solve_small(int n, std::vector<double>& A, std::vector<double>& b)
{
    for (int k = 0; k < n; ++k) {
        int piv = k;
        for (int i = k + 1; i < n; ++i)
            if (fabs(A[i * n + k]) > fabs(A[piv * n + k])) piv = i;
        if (A[piv * n + k] == 0.0) return false;
        if (piv != k) {
            for (int j = 0; j < n; ++j) std::swap(A[k * n + j], A[piv * n + j]);
            std::swap(b[k], b[piv]);
        }
        for (int i = k + 1; i < n; ++i) {
            const double f = A[i * n + k] / A[k * n + k];
            for (int j = k; j < n; ++j) A[i * n + j] -= f * A[k * n + j];
            b[i] -= f * b[k];
        }
    }
    for (int k = n - 1; k >= 0; --k) {
        double v = b[k];
        for (int j = k + 1; j < n; ++j) v -= A[k * n + j] * b[j];
        b[k] = v / A[k * n + k];
    }
    return true;
}

int scf_check(mmd_ctx* c)
{
    if (c->scf_n == 0 || c->scf_n != c->nbf) return fail(MMD_ERR_STATE, "mmd_scf_begin must be called first");
    return MMD_OK;
}

}  // namespace

extern "C" {

namespace {
// Enqueue G = sym(G(dP)) (the symmetrised build of mmd_fock_delta) on the
// context's stream: no allocation and no host synchronisation, so that a
// batched iteration can be captured into a graph.
int enqueue_fock_sym(mmd_ctx* c, const double* dP, double tol, double* Gout)
{
    const int N = c->nbf;
    const size_t nn = (size_t)N * N;
    const bool blocks = c->nmol > 1;   // a batch: only the molecules' diagonal blocks are live
    if (blocks) {
        HIP_TRY(launch_block_prep(c->nmol, c->d_mol_blocks.p, N, dP, c->d_absD.p, c->d_G.p, c->d_sh_bf0.p,
                                  c->d_sh_nbf.p, (int)c->shells.size(), c->d_Dsh.p, c->d_mol_dmax.p, c->st));
    } else {
        HIP_TRY(launch_abs_density((int)nn, dP, nullptr, c->d_absD.p, c->st));
        HIP_TRY(launch_shell_block_max((int)c->shells.size(), c->d_sh_bf0.p, c->d_sh_nbf.p, N, c->d_absD.p,
                                       c->d_Dsh.p, c->st));
        HIP_TRY(launch_fill(nn, c->d_G.p, 0.0, c->st));
    }
    ClassArgs A = base_args(c);
    A.absD = c->d_absD.p;
    A.sqS = c->d_sqS.p;
    A.tol = tol;
    A.D = dP;
    A.G = c->d_G.p;
    FockScratch W;
    W.rowctr = c->d_rowctr.p;
    W.count = c->d_count.p;
    W.nprim = c->d_nprim.p;
    W.profile_events = true;
    W.Dsh = c->d_Dsh.p;
    W.mol_dmax = blocks ? c->d_mol_dmax.p : nullptr;
    W.surv = c->d_surv.p;
    W.surv_cap = (int64_t)c->d_surv.n;   // reserved by the caller
    W.allow_measure = false;
    int rc = run_fock_classes(c, A, tol, W);
    if (rc) return rc;
    if (blocks)
        HIP_TRY(launch_block_symmetrize(c->nmol, c->d_mol_blocks.p, N, c->d_G.p, Gout, c->st));
    else
        HIP_TRY(launch_symmetrize(N, c->d_G.p, Gout, c->st));
    return MMD_OK;
}

__global__ void batch_axpy_kernel(size_t n, const double* __restrict__ a, const double* __restrict__ b,
                                  double* __restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] + b[i];
}
}  // namespace

int mmd_batch_rhf(mmd_ctx* ctx, const int32_t* nocc, const double* enuc, const double* P0, double conver,
                  double acc2e, int maxiter, int diis, int use_graph, double* energy, int32_t* iterations,
                  int32_t* converged, double* P_out, double* mo_out)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    mmd_ctx* c = ctx;
    if (c->nbf == 0) return fail(MMD_ERR_STATE, "no basis set");
    if (!c->shells_ok) return fail(MMD_ERR_UNSUPPORTED, "%s", c->shells_err.c_str());
    if (!nocc || !enuc || !energy || !iterations || !converged || maxiter < 1)
        return fail(MMD_ERR_ARG, "invalid batch arguments");
    const int M = c->nmol, N = c->nbf;
    const size_t nn = (size_t)N * N;
    std::vector<BatchMol> mols(M);
    size_t wtot = 0, dtot = 0, ptot = 0;
    for (int m = 0; m < M; ++m) {
        const int n = c->mol_n[m];
        if (n > BATCH_MAX_NBF)
            return fail(MMD_ERR_UNSUPPORTED, "molecule %d has %d basis functions; the batch takes up to %d", m, n,
                        BATCH_MAX_NBF);
        if (nocc[m] < 1 || nocc[m] > n) return fail(MMD_ERR_ARG, "molecule %d: invalid occupation", m);
        BatchMol& b = mols[m];
        b.f0 = c->mol_f0[m];
        b.n = n;
        b.nocc = nocc[m];
        b.pad = 0;
        b.wo = wtot;
        b.d8 = dtot;
        b.po = ptot;
        b.enuc = enuc[m];
        wtot += 3 * (size_t)n * n;
        dtot += 16 * (size_t)n * n;
        ptot += (size_t)n * n;
    }
    hipStream_t st = c->st;
    // one-electron integrals of every molecule's own block; H = T + V
    HIP_TRY(c->d_onee.reserve(3 * nn));
    double* S = c->d_onee.p;
    {
        const double origin[3] = {0.0, 0.0, 0.0};
        if ((rc = one_electron_shells(c, origin, S, S + nn, S + 2 * nn, nullptr, nullptr))) return rc;
    }
    for (DevBuf<double>* b : {&c->b_H, &c->b_X, &c->b_U, &c->b_F, &c->b_P, &c->b_Pold, &c->b_dP, &c->b_G})
        HIP_TRY(b->reserve(nn));
    HIP_TRY(c->b_diis.reserve(std::max<size_t>(1, dtot)));
    HIP_TRY(c->b_mo.reserve(N));
    HIP_TRY(c->b_state.reserve(M));
    HIP_TRY(c->b_flag.reserve(M));
    HIP_TRY(c->b_count.reserve(1));
    hipLaunchKernelGGL(batch_axpy_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, nn, S + nn, S + 2 * nn,
                       c->b_H.p);
    HIP_TRY(hipGetLastError());
    // Schwarz diagonal, pair tables and work units of the whole batch
    HIP_TRY(c->d_schwarz.reserve(nn));
    if ((rc = schwarz_diagonal(c, c->d_schwarz.p))) return rc;
    HIP_TRY(c->b_mol.reserve(M));
    HIP_TRY(hipMemcpyAsync(c->b_mol.p, mols.data(), sizeof(BatchMol) * M, hipMemcpyHostToDevice, st));
    BatchArgs B;
    B.nmol = M;
    B.N = N;
    B.maxn = c->rw;
    B.lds_bytes = c->lds_per_block;
    B.mol = c->b_mol.p;
    HIP_TRY(c->b_work.reserve(std::max<size_t>(1, wtot)));   // 3 n^2 per molecule: staging for packed blocks
    {
        // the molecules' diagonal blocks of (pq|pq), packed, for the pair factors
        HIP_TRY(launch_batch_pack(B, c->d_schwarz.p, c->b_work.p, st));
        std::vector<double> h(std::max<size_t>(1, ptot));
        HIP_TRY(hipMemcpyAsync(h.data(), c->b_work.p, sizeof(double) * ptot, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        DiagView v;
        v.p = h.data();
        v.row.resize(N);
        for (int m = 0; m < M; ++m)
            for (int i = 0; i < mols[m].n; ++i)
                v.row[mols[m].f0 + i] = (int64_t)mols[m].po + (int64_t)i * mols[m].n - mols[m].f0;
        if ((rc = finish_schwarz(c, v))) return rc;
    }
    // buffers of the Fock build, reserved before any capture
    HIP_TRY(c->d_rowctr.reserve(2 * NQCLASS));
    HIP_TRY(c->d_nprim.reserve(NPRIM_SLOTS));
    {
        int64_t wave = 0;
        for (int qc = 0; qc < NQCLASS; ++qc) {
            int la, lb, lc, ld;
            qclass_decode(qc, &la, &lb, &lc, &ld);
            if (c->qc_total[qc] && fock_class_is_wave(la, lb, lc, ld)) wave += 2 * c->qc_total[qc];
        }
        // survivor lists of the wavefront classes; row-kernel fallbacks need
        // more and are refused by the capacity check
        HIP_TRY(c->d_surv.reserve((size_t)std::max<int64_t>(wave, 1)));
    }
    // state: core guess F = H, P = P_old = 0; or the given densities
    std::vector<BatchState> st0(M);
    for (auto& x : st0) {
        memset(&x, 0, sizeof(x));
        x.energy = 0.0;
        x.prms = 1e20;
    }
    HIP_TRY(hipMemcpyAsync(c->b_state.p, st0.data(), sizeof(BatchState) * M, hipMemcpyHostToDevice, st));
    HIP_TRY(launch_fill_int(M, c->b_flag.p, 0, st));
    HIP_TRY(hipMemcpyAsync(c->b_F.p, c->b_H.p, sizeof(double) * nn, hipMemcpyDeviceToDevice, st));
    HIP_TRY(launch_fill(nn, c->b_P.p, 0.0, st));
    B.state = c->b_state.p;
    B.S = S;
    B.H = c->b_H.p;
    B.X = c->b_X.p;
    B.U = c->b_U.p;
    B.F = c->b_F.p;
    B.P = c->b_P.p;
    B.Pold = c->b_Pold.p;
    B.dP = c->b_dP.p;
    B.G = c->b_G.p;
    B.work = c->b_work.p;
    B.diis = c->b_diis.p;
    B.mo = c->b_mo.p;
    B.flag = c->b_flag.p;
    HIP_TRY(launch_batch_overlap(B, st));
    if (P0) {
        // an initial density (SAD, learned): F = H + G(P0), P_old = P0
        // (packed blocks through the work area, which the steps use later)
        HIP_TRY(hipMemcpyAsync(c->b_work.p, P0, sizeof(double) * ptot, hipMemcpyHostToDevice, st));
        HIP_TRY(launch_batch_unpack(B, c->b_work.p, c->b_P.p, st));
        HIP_TRY(hipStreamSynchronize(st));
        if ((rc = enqueue_fock_sym(c, c->b_P.p, acc2e, c->b_G.p))) return rc;
        hipLaunchKernelGGL(batch_axpy_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, nn, c->b_H.p,
                           c->b_G.p, c->b_F.p);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipMemcpyAsync(c->b_Pold.p, c->b_P.p, sizeof(double) * nn, hipMemcpyDeviceToDevice, st));
    HIP_TRY(launch_batch_step(B, 0, diis, conver, st));
    {
        std::vector<int> fl(M);
        HIP_TRY(hipMemcpyAsync(fl.data(), c->b_flag.p, sizeof(int) * M, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        for (int m = 0; m < M; ++m)
            if (fl[m] == -1) return fail(MMD_ERR_ARG, "molecule %d: overlap matrix not positive definite", m);
    }
    // iterations 1 .. maxiter-1: Fock build of the whole batch from dP, then
    // the per-molecule step; replayed from a graph when use_graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    auto iteration = [&]() -> int {
        int r = enqueue_fock_sym(c, c->b_dP.p, acc2e, c->b_G.p);
        if (r) return r;
        HIP_TRY(launch_batch_step(B, 1, diis, conver, st));
        return MMD_OK;
    };
    // profiling: per-class events and counters read after every build, so
    // the iterations are launched one by one
    const bool profiled = c->profiling != 0;
    if (use_graph && !profiled) {
        std::lock_guard<std::recursive_mutex> lk(device_alloc_mutex());
        if (hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed) == hipSuccess) {
            const int r = iteration();
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(st, &g);
            if (r == MMD_OK && e == hipSuccess && g &&
                hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) == hipSuccess) {
                graph = g;
            } else {
                if (g) (void)hipGraphDestroy(g);
                exec = nullptr;
                (void)hipGetLastError();
                if (r) return r;
            }
        }
    }
    int left = M;
    for (int it = 1; it < maxiter && left > 0; ++it) {
        if (exec) {
            HIP_TRY(hipGraphLaunch(exec, st));
        } else if ((rc = iteration())) {
            return rc;
        }
        if (profiled && (rc = collect_profile(c))) return rc;
        if (it % 4 == 0 || it == maxiter - 1) {
            HIP_TRY(launch_batch_count(B, c->b_count.p, st));
            HIP_TRY(hipMemcpyAsync(&left, c->b_count.p, sizeof(int), hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
        }
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    std::vector<BatchState> out(M);
    std::vector<int> fl(M);
    HIP_TRY(hipMemcpyAsync(out.data(), c->b_state.p, sizeof(BatchState) * M, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(fl.data(), c->b_flag.p, sizeof(int) * M, hipMemcpyDeviceToHost, st));
    if (mo_out) HIP_TRY(hipMemcpyAsync(mo_out, c->b_mo.p, sizeof(double) * N, hipMemcpyDeviceToHost, st));
    if (P_out) {
        HIP_TRY(launch_batch_pack(B, c->b_P.p, c->b_work.p, st));
        HIP_TRY(hipMemcpyAsync(P_out, c->b_work.p, sizeof(double) * ptot, hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    for (int m = 0; m < M; ++m) {
        if (fl[m] == -2) return fail(MMD_ERR_STATE, "molecule %d: singular DIIS matrix", m);
        energy[m] = out[m].energy;
        iterations[m] = out[m].iterations;
        converged[m] = out[m].converged;
    }
    return MMD_OK;
}

int mmd_scf_begin(mmd_ctx* ctx, int nocc, const double* H, const double* S, const double* X, const double* F,
                  const double* P)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    const int N = ctx->nbf;
    if (N == 0) return fail(MMD_ERR_STATE, "no basis set");
    if (nocc < 1 || nocc > N || !H || !S || !X || !F || !P) return fail(MMD_ERR_ARG, "invalid SCF arguments");
    if (!ctx->blas) {
        std::lock_guard<std::recursive_mutex> lk(device_alloc_mutex());
        if (rocblas_create_handle(&ctx->blas) != rocblas_status_success)
            return fail(MMD_ERR_DEVICE, "rocblas_create_handle failed");
    }
    if (rocblas_set_stream(ctx->blas, ctx->st) != rocblas_status_success)
        return fail(MMD_ERR_DEVICE, "rocblas_set_stream failed");
    const size_t nn = (size_t)N * N;
    for (DevBuf<double>* b : {&ctx->s_H, &ctx->s_S, &ctx->s_X, &ctx->s_F, &ctx->s_P, &ctx->s_Pold, &ctx->s_G,
                              &ctx->s_dP, &ctx->s_T1, &ctx->s_T2, &ctx->s_C})
        HIP_TRY(b->reserve(nn));
    HIP_TRY(ctx->s_Fset.reserve(8 * nn));
    HIP_TRY(ctx->s_Eset.reserve(8 * nn));
    HIP_TRY(ctx->s_stats.reserve(2));
    HIP_TRY(ctx->s_dots.reserve(8));
    const size_t by = sizeof(double) * nn;
    HIP_TRY(hipMemcpyAsync(ctx->s_H.p, H, by, hipMemcpyHostToDevice, ctx->st));
    HIP_TRY(hipMemcpyAsync(ctx->s_S.p, S, by, hipMemcpyHostToDevice, ctx->st));
    HIP_TRY(hipMemcpyAsync(ctx->s_X.p, X, by, hipMemcpyHostToDevice, ctx->st));
    HIP_TRY(hipMemcpyAsync(ctx->s_F.p, F, by, hipMemcpyHostToDevice, ctx->st));
    HIP_TRY(hipMemcpyAsync(ctx->s_P.p, P, by, hipMemcpyHostToDevice, ctx->st));
    HIP_TRY(hipMemcpyAsync(ctx->s_Pold.p, P, by, hipMemcpyHostToDevice, ctx->st));
    HIP_TRY(launch_fill(nn, ctx->s_G.p, 0.0, ctx->st));
    HIP_TRY(hipStreamSynchronize(ctx->st));
    ctx->scf_n = N;
    ctx->scf_nocc = nocc;
    ctx->s_order.clear();
    ctx->s_B.assign(64, 0.0);
    return MMD_OK;
}

int mmd_scf_step(mmd_ctx* ctx, double tol, int build, int diis, double* FO)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if ((rc = scf_check(ctx))) return rc;
    if (!FO) return fail(MMD_ERR_ARG, "null pointer");
    const int N = ctx->nbf;
    const size_t nn = (size_t)N * N;
    hipStream_t st = ctx->st;
    const double* Fuse = ctx->s_F.p;
    if (build) {
        // F <- F + G(P - P_old); P_old <- P   (mmd/scf.py:84-94)
        HIP_TRY(launch_scf_axpby(nn, 1.0, ctx->s_P.p, -1.0, ctx->s_Pold.p, ctx->s_dP.p, st));
        rc = mmd_fock_delta(ctx, ctx->s_dP.p, nullptr, tol, 0, ctx->s_G.p, nullptr, 1, nullptr);
        if (rc) return rc;
        HIP_TRY(launch_scf_axpby(nn, 1.0, ctx->s_F.p, 1.0, ctx->s_G.p, ctx->s_F.p, st));
        HIP_TRY(hipMemcpyAsync(ctx->s_Pold.p, ctx->s_P.p, sizeof(double) * nn, hipMemcpyDeviceToDevice, st));
        if (diis) {
            // error X (FPS - SPF) X of the new F, into the window (mmd/scf.py:146-180)
            int slot;
            std::vector<int>& ord = ctx->s_order;
            std::vector<double>& B = ctx->s_B;
            if ((int)ord.size() == 8) {
                slot = ord.front();
                ord.erase(ord.begin());
                for (int i = 0; i < 7; ++i)
                    for (int j = 0; j < 7; ++j) B[i * 8 + j] = B[(i + 1) * 8 + (j + 1)];
            } else {
                slot = (int)ord.size();
            }
            double* Es = ctx->s_Eset.p + (size_t)slot * nn;
            double* Fs = ctx->s_Fset.p + (size_t)slot * nn;
            if ((rc = gemm_rm(ctx, false, false, N, N, N, ctx->s_P.p, N, ctx->s_S.p, N, ctx->s_T1.p, N))) return rc;
            if ((rc = gemm_rm(ctx, false, false, N, N, N, ctx->s_F.p, N, ctx->s_T1.p, N, ctx->s_T2.p, N))) return rc;
            HIP_TRY(launch_scf_antisym(N, ctx->s_T2.p, ctx->s_T1.p, st));
            if ((rc = gemm_rm(ctx, false, false, N, N, N, ctx->s_X.p, N, ctx->s_T1.p, N, ctx->s_T2.p, N))) return rc;
            if ((rc = gemm_rm(ctx, false, false, N, N, N, ctx->s_T2.p, N, ctx->s_X.p, N, Es, N))) return rc;
            HIP_TRY(hipMemcpyAsync(Fs, ctx->s_F.p, sizeof(double) * nn, hipMemcpyDeviceToDevice, st));
            ord.push_back(slot);
            const int n = (int)ord.size();
            const double* E[8];
            const double* Fk[8];
            for (int k = 0; k < n; ++k) {
                E[k] = ctx->s_Eset.p + (size_t)ord[k] * nn;
                Fk[k] = ctx->s_Fset.p + (size_t)ord[k] * nn;
            }
            double dots[8];
            HIP_TRY(launch_scf_dots(n, nn, E, Es, ctx->s_dots.p, st));
            HIP_TRY(hipMemcpyAsync(dots, ctx->s_dots.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            for (int k = 0; k < n; ++k) B[(n - 1) * 8 + k] = B[k * 8 + (n - 1)] = dots[k];
            // [B -1; -1 0] w = [0 .. 0 -1]
            const int m = n + 1;
            std::vector<double> A((size_t)m * m, 0.0), w(m, 0.0);
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) A[i * m + j] = B[i * 8 + j];
            for (int i = 0; i < n; ++i) A[i * m + n] = A[n * m + i] = -1.0;
            w[n] = -1.0;
            if (!solve_small(m, A, w)) return fail(MMD_ERR_STATE, "singular DIIS matrix");
            double sw = 0.0;
            for (int k = 0; k < n; ++k) sw += w[k];
            if (!(fabs(sw - 1.0) <= 1e-8 + 1e-5)) return fail(MMD_ERR_STATE, "DIIS weights do not sum to one");
            HIP_TRY(launch_scf_combine(nn, n, Fk, w.data(), ctx->s_C.p, st));
            Fuse = ctx->s_C.p;
        }
    }
    // FO = X^T F X
    if ((rc = gemm_rm(ctx, true, false, N, N, N, ctx->s_X.p, N, Fuse, N, ctx->s_T1.p, N))) return rc;
    if ((rc = gemm_rm(ctx, false, false, N, N, N, ctx->s_T1.p, N, ctx->s_X.p, N, ctx->s_T2.p, N))) return rc;
    HIP_TRY(hipMemcpyAsync(FO, ctx->s_T2.p, sizeof(double) * nn, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return MMD_OK;
}

int mmd_scf_density(mmd_ctx* ctx, const double* CO, double* stats)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if ((rc = scf_check(ctx))) return rc;
    if (!CO || !stats) return fail(MMD_ERR_ARG, "null pointer");
    const int N = ctx->nbf;
    const size_t nn = (size_t)N * N;
    hipStream_t st = ctx->st;
    // C = X C'; P = C_occ C_occ^T; E_el = sum (H + F)_pq P_qp; ||P - P_old||^2
    HIP_TRY(hipMemcpyAsync(ctx->s_T1.p, CO, sizeof(double) * nn, hipMemcpyHostToDevice, st));
    if ((rc = gemm_rm(ctx, false, false, N, N, N, ctx->s_X.p, N, ctx->s_T1.p, N, ctx->s_T2.p, N))) return rc;
    if ((rc = gemm_rm(ctx, false, true, N, N, ctx->scf_nocc, ctx->s_T2.p, N, ctx->s_T2.p, N, ctx->s_P.p, N)))
        return rc;
    HIP_TRY(launch_scf_energy(N, ctx->s_H.p, ctx->s_F.p, ctx->s_P.p, ctx->s_Pold.p, ctx->s_stats.p, st));
    double h[2];
    HIP_TRY(hipMemcpyAsync(h, ctx->s_stats.p, sizeof(h), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    stats[0] = h[0];
    stats[1] = sqrt(h[1]);
    return MMD_OK;
}

int mmd_scf_get(mmd_ctx* ctx, double* F, double* P, double* P_old, double* G)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if ((rc = scf_check(ctx))) return rc;
    const size_t by = sizeof(double) * (size_t)ctx->nbf * ctx->nbf;
    if (F) HIP_TRY(hipMemcpyAsync(F, ctx->s_F.p, by, hipMemcpyDeviceToHost, ctx->st));
    if (P) HIP_TRY(hipMemcpyAsync(P, ctx->s_P.p, by, hipMemcpyDeviceToHost, ctx->st));
    if (P_old) HIP_TRY(hipMemcpyAsync(P_old, ctx->s_Pold.p, by, hipMemcpyDeviceToHost, ctx->st));
    if (G) HIP_TRY(hipMemcpyAsync(G, ctx->s_G.p, by, hipMemcpyDeviceToHost, ctx->st));
    HIP_TRY(hipStreamSynchronize(ctx->st));
    return MMD_OK;
}

int mmd_ao2mo(mmd_ctx* ctx, const double* T, int t_on_device, const double* C, double* out, int out_on_device,
              double* gemm_ms)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (ctx->nbf == 0) return fail(MMD_ERR_STATE, "no basis set");
    if (!T || !C || !out) return fail(MMD_ERR_ARG, "null pointer");
    const int N = ctx->nbf;
    const size_t nn = (size_t)N * N, n3 = nn * N, n4 = nn * nn;
    if (n3 > (size_t)INT32_MAX) return fail(MMD_ERR_UNSUPPORTED, "basis too large for the transform");
    DevBuf<double> a, b, dC;
    HIP_TRY(a.reserve(n4));
    HIP_TRY(b.reserve(n4));
    HIP_TRY(dC.reserve(nn));
    HIP_TRY(hipMemcpyAsync(dC.p, C, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->st));
    HIP_TRY(hipMemcpyAsync(a.p, T, sizeof(double) * n4, t_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                           ctx->st));
    // Four quarter transforms.  Each contracts the last index of X (M x N,
    // M = N^3) with C and writes it first: Y[S][pqr] = sum_s C[s][S] X[pqr][s],
    // i.e. Y = C^T X^T; after four passes the index order is (P,Q,R,S) again.
    double* x = a.p;
    double* y = b.p;
    HIP_TRY(hipEventRecord(ctx->ev0, ctx->st));
    for (int pass = 0; pass < 4; ++pass) {
        HIP_TRY(launch_gemm_f64(true, true, N, (int)n3, N, 1.0, dC.p, N, x, N, 0.0, y, (int)n3, ctx->st));
        std::swap(x, y);
    }
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->st));
    HIP_TRY(hipMemcpyAsync(out, x, sizeof(double) * n4, out_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                           ctx->st));
    HIP_TRY(hipStreamSynchronize(ctx->st));
    if (gemm_ms) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        *gemm_ms = ms;
    }
    return MMD_OK;
}

int mmd_overlap_powers(mmd_ctx* ctx, const double* S, double* X, double* U, int* iterations)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (ctx->nbf == 0) return fail(MMD_ERR_STATE, "no basis set");
    if (!X && !U) return fail(MMD_ERR_ARG, "null pointer");
    const int N = ctx->nbf;
    const size_t nn = (size_t)N * N;
    if (!S && ctx->d_onee.n < nn) return fail(MMD_ERR_STATE, "no overlap matrix: call mmd_one_electron or pass S");
    HIP_TRY(ctx->d_ns.reserve(8 * nn + 2));
    double* dS = ctx->d_ns.p;
    double* dX = dS + nn;
    double* dU = dX + nn;
    if (S)
        HIP_TRY(hipMemcpyAsync(dS, S, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->st));
    else   // the overlap of the last mmd_one_electron, still on the device
        HIP_TRY(hipMemcpyAsync(dS, ctx->d_onee.p, sizeof(double) * nn, hipMemcpyDeviceToDevice, ctx->st));
    const int it = overlap_powers_device(N, dS, dX, dU, dU + nn, 100, ctx->st);
    if (it < 0) return fail(MMD_ERR_DEVICE, "overlap powers: kernel launch failed");
    if (it == 0)
        return fail(MMD_ERR_ARG, "overlap matrix is not positive definite (Newton-Schulz did not converge)");
    if (X) HIP_TRY(hipMemcpyAsync(X, dX, sizeof(double) * nn, hipMemcpyDeviceToHost, ctx->st));
    if (U) HIP_TRY(hipMemcpyAsync(U, dU, sizeof(double) * nn, hipMemcpyDeviceToHost, ctx->st));
    HIP_TRY(hipStreamSynchronize(ctx->st));
    if (iterations) *iterations = it;
    return MMD_OK;
}

int mmd_dgemm(mmd_ctx* ctx, int transA, int transB, int M, int N, int K, double alpha, const double* A, int lda,
              const double* B, int ldb, double beta, double* C, int ldc, int on_device)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (M < 0 || N < 0 || K < 0) return fail(MMD_ERR_ARG, "negative dimension");
    if (!A || !B || !C) return fail(MMD_ERR_ARG, "null pointer");
    const int rA = transA ? K : M, cA = transA ? M : K;
    const int rB = transB ? N : K, cB = transB ? K : N;
    if (lda < std::max(1, cA) || ldb < std::max(1, cB) || ldc < std::max(1, N))
        return fail(MMD_ERR_ARG, "leading dimension too small");
    if (on_device) {
        HIP_TRY(launch_gemm_f64(transA, transB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, ctx->st));
        HIP_TRY(hipStreamSynchronize(ctx->st));
        return MMD_OK;
    }
    const size_t nA = (size_t)rA * lda, nB = (size_t)rB * ldb, nC = (size_t)M * ldc;
    DevBuf<double> a, b, c;
    HIP_TRY(a.reserve(nA));
    HIP_TRY(b.reserve(nB));
    HIP_TRY(c.reserve(nC));
    HIP_TRY(hipMemcpyAsync(a.p, A, sizeof(double) * nA, hipMemcpyHostToDevice, ctx->st));
    HIP_TRY(hipMemcpyAsync(b.p, B, sizeof(double) * nB, hipMemcpyHostToDevice, ctx->st));
    HIP_TRY(hipMemcpyAsync(c.p, C, sizeof(double) * nC, hipMemcpyHostToDevice, ctx->st));
    HIP_TRY(launch_gemm_f64(transA, transB, M, N, K, alpha, a.p, lda, b.p, ldb, beta, c.p, ldc, ctx->st));
    HIP_TRY(hipMemcpyAsync(C, c.p, sizeof(double) * nC, hipMemcpyDeviceToHost, ctx->st));
    HIP_TRY(hipStreamSynchronize(ctx->st));
    return MMD_OK;
}

int mmd_last_fock_ms(mmd_ctx* ctx, double* ms)
{
    if (!ctx || !ms) return fail(MMD_ERR_ARG, "null pointer");
    *ms = ctx->last_fock_ms;
    return MMD_OK;
}

}  // extern "C"
