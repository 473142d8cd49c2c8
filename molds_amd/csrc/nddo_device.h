// Shared host/device data layout of the NDDO SCF engine.
//
// A batch holds J jobs (molecules).  Everything is flattened over the batch:
//   atoms  : AtomDev[natoms_total]      (job-local AO offsets)
//   pairs  : PairDev[npairs_total]      one record per unordered atom pair A<B
//   slots  : pair partial buffer indexed by  job.pair0 + tri(A,B)
//   arena  : per job (ARENA_FIXED + 2k) * n^2 doubles, see ArenaKind
// Matrices are row-major n x n with leading dimension n.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define NDDO_MAX_DIIS 16

namespace nddo {

// AO order inside an atom: s, py, pz, px (MolDS src/base/Enums.h OrbitalType).
enum { ORB_S = 0, ORB_PY = 1, ORB_PZ = 2, ORB_PX = 3 };

struct ElemDev {
    double core;            // core charge Z
    double U[2];            // U_ss, U_pp (Hartree)
    double zeta[2];         // Slater exponents s, p
    double beta[2];         // resonance beta_s, beta_p (Hartree)
    double alpha;           // core-repulsion alpha (1/bohr)
    double Gss, Gpp, Gsp, Gpp2, Hsp, Hpp;
    double D[3];            // multipole separations D0 D1 D2 (bohr)
    double rho[3];          // additive terms rho0 rho1 rho2 (bohr)
    double K[4], L[4], M[4];// AM1/PM3 Gaussian core terms (a.u.)
    int Z, nao, nshell, pad;
};

struct AtomDev {
    double x, y, z;
    int elem;     // index into the ElemDev table
    int job;      // job index in the batch
    int ao;       // job-local first AO
    int local;    // job-local atom index
};

struct PairDev {
    int a, b;          // global atom indices, job-local a < b
    int job;
    int slot;          // global pair slot (job.pair0 + tri(a_local, b_local))
    long long ioff;    // offset of the packed integrals (doubles)
};

// Row a of a job's pair triangle (pairs (a, b), b > a, in order): where its
// entries go, from which k_build_pairs writes the pair list on the device.
// A pair as the Fock kernels read it, in the class-sorted order of the pair
// list (written by k_build_pairs): everything but the job's own state and
// matrices in one record, so a pair's density and integrals are one dependent
// load away. ao_a / ao_b: the atom's job-local first AO, its AO count in bits 24+.
struct FockPairDev {
    int job;
    int ao_a, ao_b;
    int slot;          // global pair slot (Coulomb partials)
    long long ioff;    // offset of the packed integrals (doubles)
};
__host__ __device__ inline int fp_ao(int x) { return x & 0xFFFFFF; }
__host__ __device__ inline int fp_nao(int x) { return x >> 24; }

struct AtomRowDev {
    long long slot0;    // global slot of pair (a, a+1)
    long long ioff0;    // integral offset (doubles) of that pair
    long long cpos[3];  // its position in the class-sorted pair list: H-H, heavy-H, heavy-heavy
};

// Arena layout per job, in units of n*n doubles.
enum ArenaKind {
    AR_F = 0,      // Fock matrix
    AR_P0 = 1,     // density ping
    AR_P1 = 2,     // density pong
    AR_C = 3,      // eigenvectors (rows); rank order in perm[]
    AR_W = 4,      // eigensolver work matrix (Householder reflectors)
    AR_H = 5,      // core Hamiltonian F(P = 0), full symmetric
    AR_Q2 = 6,     // divide-and-conquer ping-pong partner of AR_C
    AR_U = 7,      // divide-and-conquer merge eigenvectors
    AR_RING = 8    // DIIS ring: k stored densities then k stored errors
};
#define ARENA_FIXED 8

// Eigensolver workspace per job (kernels_eig.hip): d, e, tau and the compact-WY
// T blocks, EIG_WS_PER_AO * n + EIG_WS_PER_JOB doubles.
#define EIG_WS_PER_AO 19
#define EIG_WS_PER_JOB 256

enum JobPhase { PH_ACTIVE = 0, PH_ENERGY = 1, PH_DONE = 2, PH_IDLE = 3 };

struct JobDev {
    int natoms, nao, nocc, theory;
    int atom0;            // first global atom
    int npairs;           // natoms*(natoms-1)/2
    long long pair0;      // first global pair slot
    long long arena;      // offset (doubles) of this job's arena
    long long vec0;       // offset (doubles) of per-job vectors of length n (eps)
    long long ivec0;      // offset (ints) of per-job int vectors of length n (perm)
    int tile0, ntiles;    // density tiles of this job in the tile list
    int tiles_per_dim;
    // SCF options (MolDS Parameters SCF knobs)
    int max_iter, diis_k, requires_guess, precision;
    int xl;               // 1 + position in the batch's large-molecule list (0: not large)
    int want_c;           // 1: back-transform every eigenvector; 0: only the nocc occupied ones
                          // (all the density needs; the caller asked for no C)
    int want_f;           // 1: analytic forces of the converged density (k_pair_grad)
    double rms_thresh, damp_thresh, damp_weight, diis_start, diis_end;
    long long traj0;      // offset (doubles) into the trajectory buffer: max_iter * 4
};

struct JobState {
    int phase, iter, status, diis_on;
    int damp_on, ring_start, warm, sweeps;
    int fp32;          // mixed-precision job still in its FP32 iterations
    int switch_iter;   // first FP64 (polish) iteration of a mixed-precision job
    double rms, diis_error, e_el, e_core;
    double prof[4];   // eigensolver phase times of the last solve (device clock ticks)
    double B[(NDDO_MAX_DIIS + 1) * (NDDO_MAX_DIIS + 1)];  // DIIS bordered matrix
};

struct TileDev { int job, i0, j0, pad; };

// Slot of pair (a, b), a < b job-local, in the job's pair triangle.
__host__ __device__ inline long long tri_slot(int a, int b, int N) {
    return (long long)a * N - (long long)a * (a + 1) / 2 + (b - a - 1);
}

// DIIS partial records: NDDO_MAX_DIIS + 1 doubles (the error dot products in
// logical ring order, then max |e|) per density tile (k_density) or, for large
// molecules, per chunk (k_xl_dots, with the energy partials of k_xl_energy);
// k_step leaves the DIIS coefficients and what to apply for k_apply /
// k_xl_apply in XL_COEF doubles per job.
#define XL_REC (NDDO_MAX_DIIS + 1)
#define XL_COEF (NDDO_MAX_DIIS + 4)
#define XL_COEF_FLAGS (NDDO_MAX_DIIS + 1)   // bit 0: DIIS extrapolation, bit 1: damping
#define XL_COEF_ITER (NDDO_MAX_DIIS + 2)
#define XL_COEF_RS (NDDO_MAX_DIIS + 3)

}  // namespace nddo
