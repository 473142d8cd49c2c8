/*
 * nddo.h -- C ABI of the MI355X NDDO (MNDO / AM1 / PM3) ground-state SCF engine.
 *
 * Drop-in boundary for MolDS's ElectronicStructure path
 *   src/base/ElectronicStructure.h:23-53   (SetMolecule, DoSCF, Get* accessors)
 *   src/base/factories/ElectronicStructureFactory.cpp:58-95 (MNDO :70, AM1 :73, PM3 :79)
 *   src/cndo/Cndo2.cpp:585-729             (Cndo2::DoSCF, the loop this library runs)
 * Plain pointers and sizes only; no exception crosses the boundary.  Inputs are
 * copied at submit; outputs are copied into caller-owned buffers.  One engine
 * drives one HIP device.
 */
#ifndef MOLDS_AMD_NDDO_H
#define MOLDS_AMD_NDDO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Theory selector (MolDS TheoryType MNDO / AM1 / PM3, src/base/Enums.h). */
enum { NDDO_MNDO = 0, NDDO_AM1 = 1, NDDO_PM3 = 2 };
/* Precision of the SCF iterations: FP64 everywhere, or mixed: FP32 two-electron
 * Fock contraction (FP32 copy of the integrals) and split-FP32 MFMA density while
 * the density RMS is above max(rms_density, 1e-5), then FP64 polish iterations to
 * convergence; the eigensolve, DIIS, energy and heat of formation stay FP64. */
enum { NDDO_FP64 = 0, NDDO_MIXED = 1 };

/* Status codes.  The reference throws MolDSException for the same conditions:
 * non-convergence (Cndo2.cpp:700-704), odd electron count (Cndo2.cpp:351-357),
 * unsupported atom (Cndo2.cpp:359-376), LAPACK info != 0 (Lapack.cpp:152-159),
 * heap cap (MallocerFreer.cpp:62-73). */
enum {
    NDDO_OK = 0,
    NDDO_NOT_CONVERGED = -1,
    NDDO_ODD_ELECTRONS = -2,
    NDDO_UNSUPPORTED_ELEMENT = -3,
    NDDO_ADMISSION_DENIED = -4,
    NDDO_EIGENSOLVER_FAILED = -5,
    NDDO_HIP_ERROR = -6,
    NDDO_BAD_ARG = -7,
    NDDO_PENDING = 1
};

/* SCF knobs of MolDS's Parameters (src/base/Parameters.cpp:132-144 defaults;
 * input directives max_iter, rms_density, damping_thresh, damping_weight,
 * diis_num_error_vect, diis_start_error, diis_end_error, InputParser.cpp:646-734). */
typedef struct {
    int max_iter;              /* maxIterationsSCF, default 100 */
    double rms_density;        /* thresholdSCF, default 1e-8 */
    double damping_thresh;     /* default 1.0 */
    double damping_weight;     /* default 0.8 */
    int diis_num_error_vect;   /* default 5 (0 disables DIIS), at most 16 */
    double diis_start_error;   /* default 1e-2 */
    double diis_end_error;     /* default 1e-8 */
} nddo_scf_opts;

/* Fills *o with the reference defaults. */
void nddo_scf_opts_default(nddo_scf_opts* o);

/* One SCF job: the molecule (atomic numbers H, C, N, O, F; coordinates in bohr,
 * xyz interleaved) and its SCF options.  requires_guess = 1 is DoSCF(true):
 * iteration 0 uses the guess Fock (Mndo.cpp:434-435, 511-513); requires_guess = 0
 * is DoSCF(false): iteration 0 builds the Fock matrix from P0 (n*n, row-major),
 * or from a zero density when P0 is NULL. */
typedef struct {
    int theory;
    int precision;
    int natoms;
    const int* Z;
    const double* xyz_bohr;
    nddo_scf_opts opts;
    int requires_guess;
    const double* P0;
} nddo_job_desc;

/* Results.  Scalars are always filled; array pointers are caller-owned and may
 * be NULL.  C holds the MO coefficients as rows (row i = MO i, ascending
 * energy, AO order s, py, pz, px per atom), each row's sign chosen so that its
 * element sum is >= 0, as Lapack::Dsyevd leaves MolDS's fockMatrix
 * (Lapack.cpp:117-150). */
typedef struct {
    int status;
    int nao;
    int nocc;
    int iterations;          /* SCF iterations run, including the converged one */
    int fp64_iterations;     /* of these, iterations run in FP64 (NDDO_MIXED: the polish iterations) */
    int reserved;
    double energy_total;     /* elecSCFEnergy: electronic + core repulsion (Hartree) */
    double energy_elec;
    double energy_core;      /* coreRepulsionEnergy (Hartree) */
    double heat_of_formation;/* Hartree: E_total - sum E_isol,atom + sum dHf_atom.  MNDO as
                                Mndo::CalcHeatsFormation (Mndo.cpp:395-404); AM1 / PM3 (MolDS leaves it
                                unset, Am1.cpp:261-263) with E_isol from the one-centre parameters
                                (nddo_element_eisol) */
    double final_rms;
    double latency_s;        /* submit -> retire (served jobs), or wall time of the solve */
    double* eps;             /* [nao] MO energies (Hartree), ascending */
    double* C;               /* [nao*nao] MO rows */
    double* P;               /* [nao*nao] density */
    double* trajectory;      /* [max_iter*4]: rms, diis_error, diis_on, damping_on per iteration */
    double* mulliken;        /* [natoms] Mulliken charges of the final density: core charge minus the
                                atom's sum of P_mu,mu (Cndo2.cpp:733, 1187-1199, 1733-1747) */
    double* dipole;          /* [9] ground-state dipole moments (a.u.): total x y z, electronic x y z, core
                                x y z, about the centre of core masses, with the STO-6G position integrals
                                MolDS prints (Cndo2.cpp:748-753, 1126-1184, 1871-2131); filled by
                                nddo_scf_batch / nddo_scf_once for converged jobs, left alone by served jobs */
    double* forces;          /* [natoms*3] ground-state forces -dE/dR (Hartree/bohr) of the converged
                                density, analytic (Mndo::CalcForce, Mndo.cpp:2711-2986; GetForce(0),
                                ElectronicStructure.h:36); filled by nddo_scf_batch / nddo_scf_once for
                                converged jobs (NaN otherwise), left alone by served jobs */
} nddo_result;

typedef struct nddo_engine nddo_engine;
typedef int64_t nddo_job_id;

typedef struct {
    size_t hbm_budget_bytes;  /* device bytes for job state: the served jobs' resident total and the largest
                                 synchronous batch (larger batches are split); 0 = 90% of free HBM */
    int max_batch_jobs;       /* served batch capacity; 0 = 4096 */
    int flags;                /* NDDO_ENGINE_* bits, 0 = defaults */
} nddo_engine_opts;

/* Eigensolver placement.  By default a molecule of up to 1200 AOs gets one
 * workgroup (Householder + divide and conquer), which gives the most solves/s
 * when many molecules share the GPU; larger ones use the device-wide solver
 * (the same steps spread over the GPU, several molecules side by side).  A
 * batch with at most 64 molecules above 512 AOs, and every served request
 * above 512 AOs, sends those to the device-wide solver too.  With this flag
 * every molecule above 512 AOs uses the device-wide solver: lower latency when
 * only a few large molecules are in flight. */
#define NDDO_ENGINE_DEVICE_WIDE_EIGEN 1
/* Launch every SCF iteration stage by stage instead of replaying the
 * iteration's hipGraph (captured once per batch or served cohort after its
 * first iteration; batches with device-wide eigensolves always launch eagerly). */
#define NDDO_ENGINE_NO_GRAPHS 2
/* Keep the default launch shapes instead of timing the candidates at engine
 * creation (see nddo_engine_tuning). */
#define NDDO_ENGINE_NO_AUTOTUNE 4
/* Large molecules (the device-wide eigensolver class) use two-sided block
 * Jacobi on FP64 MFMA instead of the default device-wide Householder
 * tridiagonalization + divide and conquer + compact-WY back-transformation. */
#define NDDO_ENGINE_BLOCK_JACOBI 8

int nddo_engine_create(int hip_device, const nddo_engine_opts* opts, nddo_engine** out);
int nddo_engine_destroy(nddo_engine* e);

/* Synchronous single solve (what a DoSCF adapter calls).  Returns the job status. */
int nddo_scf_once(nddo_engine* e, const nddo_job_desc* job, nddo_result* res);
/* Synchronous batch: all jobs share one launch per SCF stage.  Returns NDDO_OK
 * when the batch ran (per-job status in res[i].status) or a negative error.
 * The batch's device buffers are carved out of one engine-owned workspace that
 * grows to the largest batch seen and is reused by later calls.  A batch that
 * does not fit the device's free memory runs as smaller batches one after the
 * other; a job that does not fit on its own ends NDDO_ADMISSION_DENIED (the
 * role of MolDS's heap cap, MallocerFreer.cpp:62-73). */
int nddo_scf_batch(nddo_engine* e, int njobs, const nddo_job_desc* jobs, nddo_result* res);
/* nddo_scf_once without any device allocation: no counterpart in MolDS; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Every device buffer of
 * the solve is carved out of a fixed file-scope scratch array of
 * nddo_noalloc_scratch_bytes() (64 MiB) per GPU, which the library holds until
 * the process ends; a job that needs more (about 500 AOs for water clusters, or
 * any molecule of the device-wide eigensolver class) ends NDDO_ADMISSION_DENIED
 * with a message naming both sizes.  The iterations launch eagerly (no graph
 * instantiation).  Calls that overlap on one GPU share its array and race; the
 * engine serialises the calls made through it. */
int nddo_scf_once_noalloc(nddo_engine* e, const nddo_job_desc* job, nddo_result* res);
size_t nddo_noalloc_scratch_bytes(void);
/* Returns the batch workspace and the stream-ordered pool's unused reserve to
 * the device (no counterpart in MolDS). */
int nddo_engine_trim(nddo_engine* e);
/* Device memory as the engine sees it: free and total bytes (hipMemGetInfo),
 * the batch workspace's size, and the bytes of served results that retired but
 * were not yet fetched with nddo_get_result (no counterpart in MolDS). */
int nddo_engine_memory(nddo_engine* e, size_t* free_bytes, size_t* total_bytes, size_t* workspace_bytes,
                       size_t* served_result_bytes);

/* Serving: continuous batching.  submit copies the job and returns at once;
 * jobs are admitted as device memory frees up and retired at convergence. */
int nddo_submit(nddo_engine* e, const nddo_job_desc* job, nddo_job_id* id);
int nddo_poll(nddo_engine* e, nddo_job_id id);                      /* NDDO_PENDING or the final status */
int nddo_wait(nddo_engine* e, nddo_job_id id, double timeout_s);    /* blocks; NDDO_PENDING on timeout */
int nddo_get_result(nddo_engine* e, nddo_job_id id, nddo_result* res);
int nddo_release(nddo_engine* e, nddo_job_id id);

/* Per-stage device timing (no counterpart in MolDS, whose only timer is the
 * omp_get_wtime around DoSCF, Cndo2.cpp:587, 724-728).  With profiling on, the
 * engine records hipEvents around every stage launch on its stream and
 * accumulates milliseconds and launch counts per stage:
 * 0 setup, 1 fock, 2 eigen, 3 density, 4 step (RMS/DIIS/damping), 5 energy;
 * and the parts of the eigen stage: 6 tridiagonalization of the matrices kept
 * in LDS, 7 tridiagonalization of the matrices in device memory, 8 divide and
 * conquer, 9 back-transformation (these four up to 512 AOs), 10 the
 * device-wide eigensolver of the larger molecules (Householder + divide and
 * conquer + back-transformation, or block Jacobi under
 * NDDO_ENGINE_BLOCK_JACOBI).  In a batch
 * (nddo_scf_batch) with matrices of both kinds, the two classes' eigensolves
 * run concurrently on two streams: stage 6 then times the whole eigensolve of
 * the LDS class on its stream, and stages 8 and 9 the device-memory class. */
enum { NDDO_STAGE_SETUP = 0, NDDO_STAGE_FOCK, NDDO_STAGE_EIGEN, NDDO_STAGE_DENSITY, NDDO_STAGE_STEP,
       NDDO_STAGE_ENERGY, NDDO_STAGE_TRIDIAG_LDS, NDDO_STAGE_TRIDIAG_GLB, NDDO_STAGE_DC,
       NDDO_STAGE_BACKTRANSFORM, NDDO_STAGE_XL_EIGEN, NDDO_STAGE_COUNT };
int nddo_engine_profile(nddo_engine* e, int enable);
int nddo_engine_stage_times(nddo_engine* e, double* ms, long long* launches, int reset);

/* Device facts the engine plans with: LDS bytes per workgroup, compute units,
 * the HBM admission budget (bytes), and the largest AO count whose matrix the
 * eigensolver keeps in LDS for its tridiagonalization. */
int nddo_engine_info(nddo_engine* e, size_t* lds_bytes, int* num_cus, size_t* hbm_budget, int* lds_tridiag_max_n);

/* Text of the last error on the calling thread. */
const char* nddo_last_error(void);

/* Stage-level hooks (no counterpart in MolDS's public API; they expose the
 * per-stage results the parity tests compare):
 *   integrals: dense (mu nu|lam sig) [N][N][4][4][4][4] as MolDS's twoElecsTwoAtomCores
 *              (Mndo.cpp:113-126), overlap S [n*n] (Cndo2.cpp:4030-4146), core Hamiltonian
 *              H = F(P=0) [n*n], core repulsion energy.  Any output may be NULL.
 *   fock:      F [n*n] from a given P (guess = 1 for the iteration-0 guess Fock).
 *   eigh:      eigenvalues [n] ascending and eigenvector rows [n*n] of a symmetric F.
 *   density:   P = 2 C_occ^T C_occ from MO rows C [n*n] (first nocc rows occupied). */
int nddo_dbg_integrals(nddo_engine* e, int theory, int natoms, const int* Z, const double* xyz_bohr,
                       double* g_dense, double* S, double* H, double* e_core);
int nddo_dbg_fock(nddo_engine* e, int theory, int natoms, const int* Z, const double* xyz_bohr,
                  const double* P, int guess, double* F);
int nddo_dbg_eigh(nddo_engine* e, int n, const double* F, double* eps, double* C, int* sweeps);
int nddo_dbg_density(nddo_engine* e, int n, int nocc, const double* C, double* P);
/* The engine's parameter table (host only, no device call): the values
 * MolDS's Atom accessors return (Atom.cpp:263-1109; tables Hatom.cpp, Catom.cpp,
 * Natom.cpp, Oatom.cpp, Fatom.cpp) for element Z and theory, in atomic units:
 * core charge, U_ss, U_pp, zeta_s, zeta_p, beta_s, beta_p, alpha, Gss, Gpp, Gsp,
 * Gpp2, Hsp, Hpp, D0..D2, rho0..rho2, K[4], L[4], M[4], and the MNDO atomic
 * term dHf_atom - E_isol,atom (Mndo.cpp:395-404); 33 doubles. */
int nddo_element_params(int Z, int theory, double* out);
/* Isolated-atom electronic energy (eV) of element Z in a theory, from its
 * one-centre parameters (Dewar & Thiel 1977; MOPAC's EISOL): the term the
 * AM1 / PM3 heat of formation subtracts.  No counterpart in MolDS. */
int nddo_element_eisol(int Z, int theory, double* eisol_ev);
/* Serving scheduler's host wall time in seconds: admission + cohort setup,
 * launches, waiting for the device, retirement; and cohort iterations run. */
int nddo_dbg_server_stats(nddo_engine* e, double* seconds4, long long* iterations, int reset);
/* Device-clock ticks of the last nddo_dbg_eigh phases: tridiagonalization,
 * tridiagonal divide and conquer, (unused, 0), back-transformation; tick rate in Hz. */
int nddo_dbg_eigh_profile(nddo_engine* e, double* ticks4, double* tick_hz);
/* Device-clock ticks of the divide-and-conquer phases of the last nddo_dbg_eigh,
 * summed over the merge levels: merge order + z, deflation, rotations + secular
 * roots, Loewner z + ranks, column norms, merge matrix U, merge products. */
int nddo_dbg_eigh_dc_phases(nddo_engine* e, double* ticks7);
/* Test hook: iteration graphs captured and replayed so far by this engine. */
/* The launch shapes the engine's autotuner picked for its device (timed on
 * synthetic work at engine creation): atom pairs per workgroup of the pair
 * Fock kernel (1, 2, 4 or 8) and the back-transformation rows per workgroup
 * (32 or 64) for n <= 128 / 256 / 384 / 512.  No counterpart in MolDS. */
int nddo_engine_tuning(nddo_engine* e, int* fock_wpb, int* bt_rows4);
/* Test hook: force the launch shapes. */
int nddo_dbg_set_tuning(nddo_engine* e, int fock_wpb, const int* bt_rows4);
/* Test hook: seconds nddo_scf_batch spent so far in its host phases: build and
 * classify, device allocation and upload, the SCF loop, result copies, and
 * returning the batch's device memory. */
int nddo_dbg_host_times(nddo_engine* e, double* t5, int reset);
int nddo_dbg_graph_stats(nddo_engine* e, long long* captured, long long* replays);
/* Test hook: workgroups of the large-molecule tridiagonalization (0 = the
 * engine's choice by size). */
int nddo_dbg_set_xl_grid(nddo_engine* e, int tri_workgroups);
/* Test hook: the one-workgroup tridiagonalization of matrices in device memory
 * (138 to 512 AOs) on this engine's device: blocked, panels of 8 reflectors
 * with a rank-16 FP64 MFMA update per panel (1, the default), or one rank-2
 * update per reflector (0).  Both store the same reflectors and tridiagonal. */
int nddo_dbg_set_tridiag(nddo_engine* e, int blocked);
/* Test hook: the device-wide tridiagonalization of large molecules on this
 * engine's device: blocked (panels of 8 reflectors whose columns only read the
 * trailing block, one rank-16 FP64 MFMA update per panel) from 5600 AOs and
 * one rank-2 update per reflector below (1, the default), always blocked (2),
 * or never (0).  Both store the same reflectors and tridiagonal. */
int nddo_dbg_set_xl_tridiag(nddo_engine* e, int mode);

#ifdef __cplusplus
}
#endif
#endif /* MOLDS_AMD_NDDO_H */
