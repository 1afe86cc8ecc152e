// Batched RHF driver: per-molecule state and arguments of the molecule
// workgroup kernels (mmd_batch.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace mmd {

// Largest basis of one molecule in a batch (one n x n matrix in LDS).
constexpr int BATCH_MAX_NBF = 88;

struct BatchMol {
    int f0, n, nocc, pad;   // basis functions [f0, f0 + n); occupied orbitals
    size_t wo;              // offset of the molecule's 3 n^2 work doubles
    size_t d8;              // offset of its DIIS window (8 error + 8 Fock matrices, n^2 each)
    size_t po;              // offset of its n x n block in a packed array of all blocks
    double enuc;            // nuclear repulsion
};

struct BatchState {
    int converged, iterations, ndiis;
    int warm;               // the eigenvectors of the last step are stored (warm start)
    int order[8];           // DIIS slots, oldest first
    double energy, delta_energy, prms;
    double Bm[64];          // error overlaps tr(e_i^T e_j), window order
};

struct BatchArgs {
    int nmol, N;
    int maxn;               // largest molecule
    size_t lds_bytes;       // LDS a workgroup may allocate on the device
    const BatchMol* mol;
    BatchState* state;
    double *S, *H, *X, *U, *F, *P, *Pold, *dP, *G;   // N x N, block-diagonal
    double* work;
    double* diis;
    double* mo;             // orbital energies, N
    int* flag;              // per molecule: 0, or < 0 for a failure
};

hipError_t launch_batch_overlap(const BatchArgs& B, hipStream_t st);
hipError_t launch_batch_step(const BatchArgs& B, int build, int diis, double conver, hipStream_t st);
// packed (the molecules' n x n blocks one after another) <-> N x N
// block-diagonal; unpack writes the blocks only
hipError_t launch_batch_pack(const BatchArgs& B, const double* full, double* packed, hipStream_t st);
hipError_t launch_batch_unpack(const BatchArgs& B, const double* packed, double* full, hipStream_t st);
hipError_t launch_batch_count(const BatchArgs& B, int* out, hipStream_t st);

}  // namespace mmd
