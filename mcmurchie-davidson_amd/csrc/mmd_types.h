// Data layout shared by the host runtime and the gfx950 kernels.
#pragma once
#include <stdint.h>

namespace mmd {

// A contracted shell pair in canonical orientation (la >= lb).  Shell pairs
// are grouped by angular-momentum pair class and, inside a class, sorted by
// primitive-pair count and Schwarz factor (descending) so that the quartets a
// wavefront enumerates have similar cost and similar screening fate.
struct ShellPair {
    int32_t sa, sb;      // shell indices (sa carries la >= lb)
    int32_t bfa, bfb;    // first basis-function index of each shell
    int32_t pp0, npp;    // range in the primitive-pair table
    int32_t same;        // sa == sb
    int32_t mol;         // molecule of the pair (a basis can hold several, block-diagonally)
    int32_t c0;          // first basis function of that molecule
    int32_t pad;
    double q;            // max over the pair's functions of sqrt((ab|ab))
};

// A primitive pair: exponent sum, Gaussian-product prefactor including both
// contraction coefficients, product centre and its offsets from A and B,
// 1/(2p) and 1/p.
struct PrimPair {
    double p, K;
    double Px, Py, Pz;
    double PAx, PAy, PAz;
    double PBx, PBy, PBz;
    double inv2p, invp;
};

// Arguments of the quartet-class kernels (passed by value).
struct ClassArgs {
    const int2* quart;          // surviving (bra pair, ket pair) for this class
    const int* count;           // number of survivors (device)
    const ShellPair* pairs;     // all shell pairs
    const PrimPair* pp;         // primitive pairs
    const double* boys;         // Boys table [BOYS_ROWS][BOYS_NT]
    const float* boysf;         // the same table in float (FP32-tier kernels)
    const double* gcomp;        // per-basis-function component scale
    const double* D;            // density (difference) N x N
    const double* absD;         // |dP| N x N (screening)
    const double* sqS;          // sqrt((ij|ij)) N x N (screening)
    double tol;                 // screening threshold
    int N;
    double* G;                  // Fock accumulation N x N
};

// Ket primitive pairs per lane item of the sliced classes, unless set per
// class (mmd_set_ket_slices): short slices keep the lanes of a wavefront
// even when a batch mixes contraction lengths (a 64-molecule 6-31G batch
// pass measured 369, 341 and 334 ms for slices of 8, 4 and 2; (H2O)8
// cc-pVDZ builds were the same for all three).
constexpr int KSLICE_DEFAULT = 2;
constexpr int KSLICE_MAX = 16;

// Row decomposition of one quartet class for the Fock kernels.
// Primitive pairs per slice of the Fock-build class kernels: a work unit
// takes at most PSLICE bra primitive pairs (a 64-molecule 6-31G batch pass
// measured 372, 336, 313 and 313 ms for 4, 8, 16 and 32; (H2O)8 cc-pVDZ
// builds 1.95, 1.99 and 2.02 ms for 8, 16 and 32).
#ifndef MMD_PSLICE
#define MMD_PSLICE 16
#endif
constexpr int PSLICE = MMD_PSLICE;

// Per work unit: an upper bound of the pair factors q_P q_Q over the unit's
// quartets (rounded up to float) and the molecule of its bra pair, so that a
// unit whose every quartet fails the density-weighted screen is skipped
// before anything else of it is read.
struct UnitBound {
    float q;
    int32_t mol;
};

struct RowArgs {
    int boff, nb;               // bra pairs of the class: pairs[boff .. boff+nb)
    int koff, nk;               // ket pairs of the class
    int tri;                    // same pair class: ket index <= bra index
    int seg;                    // ket pairs per work unit
    int rw;                     // width of the LDS rows: the largest molecule of the basis
    const int4* units;          // work units: {bra pair, first | end << 16 bra primitive pair, first ket, end ket}
    int nunits;
    int maxpp;                  // largest primitive-pair count of the class's bra pairs
    int kj;                     // LDS holds the ket-side Coulomb buffer (classes that have one)
    int kslice;                 // ket primitive pairs per lane item of the sliced classes (1 .. 16)
    int nshell;
    const double* Dsh;          // shell-block maxima of |dP| (nshell x nshell)
    const double* mol_dmax;     // batches: max |dP| of each molecule; or null
    const UnitBound* ubound;    // with mol_dmax: units with 4 q max|dP| < tol are skipped
    double tier_lo, tier_hi;    // this launch takes quartets with tier_lo <= bound < tier_hi
    int* row_counter;           // next work unit (zeroed before the launch)
    unsigned long long* stat_nquart;   // optional statistics
    unsigned long long* stat_nprim;
};


}  // namespace mmd
