// Host-side launchers of the engine's HIP kernels (one translation unit per
// kernel family; the engine only calls these).
#pragma once
#include <hip/hip_runtime.h>

#include "nddo_device.h"

namespace nddo {

// Launch shapes picked by the engine's autotuner (engine.cpp), per HIP device;
// none of them changes the arithmetic of a molecule.
struct Tuning {
    int fock_wpb = 4;                      // atom pairs (wavefronts) per workgroup of the pair Fock kernel
    int bt_rows[4] = {64, 64, 64, 64};     // back-transformation rows per workgroup, n <= 128 / 256 / 384 / 512
    int xl_tri_wg = 0;                     // workgroups of the device-wide tridiagonalization (0: by size)
    int tri_blocked = 1;                   // device-memory tridiagonalization up to 512 AOs: blocked (1) or not (0)
    int xl_tri_blocked = 1;                // device-wide tridiagonalization: by size (1), always blocked (2), never (0)
};
Tuning get_tuning();               // of the current device
void set_tuning(const Tuning& t);  // for the current device

int upload_overlap_table(const double* table /* 32*25 */);

// The batch's pair list and its class-sorted index from the per-atom row starts.
void launch_build_pairs(hipStream_t st, const AtomDev* atoms, const AtomRowDev* rows, int natoms, const JobDev* jobs,
                        const ElemDev* elems, PairDev* pairs, int* plist, FockPairDev* fpairs);
void launch_setup(hipStream_t st, const PairDev* pairs, const int* plist, int nlight, int nheavy,
                  const AtomDev* atoms, int natoms, const ElemDev* elems, const JobDev* jobs, int njobs, double* ints,
                  double* arena, double* part, double* ecore, JobState* state, double* sdbg);

// FP32 copy of the packed integrals for mixed-precision jobs.
void launch_ints_f32(hipStream_t st, const double* ints, float* ints32, long long nints);

// mode 0: SCF Fock for PH_ACTIVE jobs (guess on iteration 0 when requested;
// jobs in their FP32 iterations contract ints32, which may be null when no job
// of the batch is mixed-precision);
// mode 1: non-guess FP64 Fock of the final density for PH_ENERGY jobs;
// mode 2: both in one launch (what the SCF loop uses).
// fp: the pairs to build, as pair-list ranges: the H-H and the heavy-H pairs
// (one lane each), the heavy-heavy pairs (one wavefront each); atoms: the
// batch's atoms (the pairs index them globally), of which atoms
// [atom0, atom0 + natoms) get their diagonal blocks reduced (a job's pairs and
// atoms are all in or all out).
// st2 (optional): the light pairs run on it, forked from st by fork_ev and
// joined by join_ev before the per-atom reduction on st.
struct FockPairs {
    const FockPairDev* light[2] = {nullptr, nullptr};
    int nlight[2] = {0, 0};
    const FockPairDev* heavy = nullptr;
    int nheavy = 0;
};
void launch_fock(hipStream_t st, int mode, const FockPairs& fp, const AtomDev* atoms, int atom0,
                 int natoms, const ElemDev* elems, const JobDev* jobs, const JobState* state, const double* ints,
                 const float* ints32, double* arena, double* part, hipStream_t st2 = nullptr,
                 hipEvent_t fork_ev = nullptr, hipEvent_t join_ev = nullptr);

// Warm start: F <- VT F VT^T for PH_ACTIVE jobs whose VT holds a previous solve.
void launch_warm(hipStream_t st, const TileDev* tiles, int ntiles, const JobDev* jobs, const JobState* state,
                 double* arena);

// Eigen-decomposition of F for PH_ACTIVE jobs in joblist, one workgroup per job:
// Householder + divide and conquer (kernels_eig.hip), n <= 1200; tri is the
// eigensolver workspace (EIG_WS_PER_AO, EIG_WS_PER_JOB); a_in_lds: the matrices
// fit the LDS budget (heig_lds_bytes).
size_t heig_lds_bytes(int n, bool a_in_lds);
// Raises a kernel's dynamic-LDS limit (hipFuncAttributeMaxDynamicSharedMemorySize)
// on the current device to at least `bytes`; the limit is never lowered, so a
// launch recorded in a graph with a larger request stays valid after smaller
// launches of the same kernel from other batches.
void raise_dynamic_lds(const void* func, size_t bytes);
bool heig_supported(int n);
void launch_heig_class(hipStream_t st, const JobDev* jobs, JobState* state, const int* joblist, int nlist, int max_n,
                       double* arena, double* eps, int* perm, double* tri, bool a_in_lds);
// One part of it: 0 tridiagonalization, 1 divide and conquer, 2 back-transformation.
void launch_heig_part(int part, hipStream_t st, const JobDev* jobs, JobState* state, const int* joblist, int nlist,
                      int max_n, double* arena, double* eps, int* perm, double* tri, bool a_in_lds);
void launch_eigen(hipStream_t st, const JobDev* jobs, JobState* state, const int* joblist, int nlist,
                  int max_n, double* arena, double* eps, int* perm, double* tri, bool a_in_lds);

// Density of the upper tiles, mirrored; rmspart: one RMS partial per tile;
// tpart: XL_REC doubles per tile (DIIS dot products and max |e|).
void launch_density(hipStream_t st, const TileDev* tiles, int ntiles, const JobDev* jobs,
                    const JobState* state, double* arena, const int* perm, double* rmspart, double* tpart);

// RMS test and DIIS / damping decisions per job from the tile partials
// (xlpart / nchunk: the chunk partials of the large molecules, JobDev::xl > 0);
// leaves XL_COEF doubles per job in xlcoef for launch_apply / launch_xl_apply.
void launch_step(hipStream_t st, const JobDev* jobs, int njobs, JobState* state, const double* rmspart,
                 const double* tpart, double* traj, int* remaining, const double* xlpart, double* xlcoef,
                 int nchunk, int job0 = 0);

// DIIS extrapolation and damping of the jobs that are not large, per upper tile.
void launch_apply(hipStream_t st, const TileDev* tiles, int ntiles, const JobDev* jobs, double* arena,
                  const double* coef);

void launch_energy(hipStream_t st, const JobDev* jobs, int njobs, JobState* state, double* arena,
                   int* remaining, const double* xlpart, int nchunk, int job0 = 0);

// Analytic forces of the converged jobs with want_f (kernels_setup.hip): per
// pair slot grad[4 slot + 0..2] = dE_AB/d(R_B - R_A) and grad[4 slot + 3] =
// E_AB (the pair's two-centre energy), then forces[3 g + c] = -dE/dx_gc for
// every atom g of those jobs (Hartree/bohr).
void launch_forces(hipStream_t st, const PairDev* pairs, long long nslots, const AtomDev* atoms, int natoms,
                   const ElemDev* elems, const JobDev* jobs, const JobState* state, const double* arena, double* grad,
                   double* forces);

// Per-iteration readback of small device records (the jobs' scalar state, the
// batch's remaining count) into pinned host memory by a kernel on stream st.
// A copy-engine transfer (hipMemcpyAsync to the host) would queue behind the
// copies of other streams, each waiting for its own stream's kernels: a served
// cohort of small molecules then advanced one SCF iteration per iteration of
// the slowest resident cohort (measured: 19 ms instead of 0.5 ms per
// iteration beside an 864-AO request).  bytes and both strides are multiples
// of 4; dst_dev is the device address of the pinned buffer
// (hipHostGetDevicePointer).
void launch_publish(hipStream_t st, const void* src, size_t src_stride, void* dst_dev, size_t dst_stride,
                    size_t bytes, int count);

// Ground-state dipole partial sums of converged jobs (kernels_props.hip): per
// AO row mu of job joblist[i], 4 doubles at part[doff[i] + 4 mu]: sum over the
// row's upper part of P r_x, P r_y, P r_z and P S (off-diagonal terms doubled).
void launch_dipole(hipStream_t st, const JobDev* jobs, const JobState* state, const int* joblist,
                   const long long* doff, int nlist, int max_n, const AtomDev* atoms, const ElemDev* elems,
                   const double* arena, double* part);

// Compact-WY T blocks of the Householder reflectors in W (kernels_eig.hip), any n.
void launch_wy_t(hipStream_t st, const JobDev* jobs, JobState* state, const int* joblist, int nlist, int max_n,
                 double* arena, double* tri);

// Large molecules (kernels_xlh.hip): device-wide Householder tridiagonalization,
// divide and conquer and back-transformation, persistent launches of `grid`
// workgroups (at most one per CU) over the jobs of joblist one after another;
// xloff[i] is job i's offset in xlws (xlh_ws_doubles(n) each); bar: two
// unsigned words, reset by the launches themselves.  n <= 6144 (xlh_supported).
bool xlh_supported(int n);
size_t xlh_bar_bytes();   // the barrier words the device-wide launches need
long long xlh_ws_doubles(int n);
int xlh_grid(int num_cus);
// Workgroups of k_xtri for the largest n of a launch (the grid barrier's cost
// grows with the workgroup count, the row work shrinks with it).
int xlh_tri_grid(int num_cus, int max_n, int nlist);
void launch_xlh_eigen(hipStream_t st, int grid, const JobDev* jobs, JobState* state, const int* joblist,
                      const long long* xloff, int nlist, int max_n, double* arena, double* eps, int* perm,
                      double* tri, double* xlws, unsigned* bar);

// Large molecules (kernels_xl.hip): device-wide block-Jacobi eigensolver, one
// persistent launch of `grid` workgroups (at most one per CU) over up to
// xl_max_jobs() jobs; xloff[i] is job i's offset in xlws (xl_ws_doubles(n)
// each); bar: two unsigned words, reset by the launch itself.
long long xl_ws_doubles(int n);
int xl_grid(int num_cus, const int* ns, int nlist);
int xl_max_jobs();
// Offset in a job's workspace of 7 doubles of phase clock ticks of its last solve.
long long xl_dbg_offset(int n);
int xl_nchunk(int max_n);
void launch_xl_eigen(hipStream_t st, int grid, const JobDev* jobs, JobState* state, const int* joblist,
                     const long long* xloff, int nlist, double* arena, double* eps, int* perm, double* xlws,
                     unsigned* bar);
// Chunked DIIS dot products (before k_step) and DIIS / damping update (after
// it), E_el partials (before k_energy); xlpart holds nlist * nchunk * XL_REC doubles.
void launch_xl_dots(hipStream_t st, const JobDev* jobs, const JobState* state, const int* joblist, int nlist,
                    double* arena, double* xlpart, int nchunk);
void launch_xl_apply(hipStream_t st, const JobDev* jobs, const int* joblist, int nlist, double* arena,
                     const double* xlcoef, int nchunk);
void launch_xl_energy(hipStream_t st, const JobDev* jobs, const JobState* state, const int* joblist, int nlist,
                      double* arena, double* xlpart, int nchunk);

}  // namespace nddo
