// SCF-loop kernels: Fock build, Jacobi eigensolver, density, convergence/DIIS
// step and the electronic energy.  Every kernel works on a whole batch of jobs
// and skips jobs whose phase does not match, so one launch per stage serves
// molecules of uneven size at different points of their SCF.
//
// Reference path (MolDS): Cndo2::DoSCF loop (src/cndo/Cndo2.cpp:624-705),
// CalcFockMatrix (:1526-1636) with Mndo::GetFockDiagElement /
// GetFockOffDiagElement (src/mndo/Mndo.cpp:422-580), Lapack::Dsyevd
// (src/wrappers/Lapack.cpp:102-161), CalcOrbitalElectronPopulation
// (Cndo2.cpp:1702-1731), SatisfyConvergenceCriterion (:1472-1517),
// DoDIIS (:920-1028), DoDamp (:1030-1061), CalcElecSCFEnergy (:1324-1416).
#include <hip/hip_runtime.h>
#include <math.h>

#include "launch.h"
#include "nddo_device.h"

namespace nddo {

__device__ inline int pk_index(int a, int b) {
    if (a > b) { int t = a; a = b; b = t; }
    return a == 0 ? b : (a == 1 ? 3 + b : (a == 2 ? 5 + b : 9));
}

namespace {
constexpr int kMaxDevices = 64;
Tuning g_tuning[kMaxDevices];
int current_device() {
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess || d < 0 || d >= kMaxDevices) d = 0;
    return d;
}
}  // namespace

Tuning get_tuning() { return g_tuning[current_device()]; }
void set_tuning(const Tuning& t) { g_tuning[current_device()] = t; }

__device__ inline double* mat(double* arena, const JobDev& jb, int kind) {
    return arena + jb.arena + (long long)kind * jb.nao * jb.nao;
}


// One-centre Coulomb g and exchange h integrals (Mndo.cpp:583-639).
__device__ inline double g1c(const ElemDev& E, int a, int b) {
    if (a == 0 && b == 0) return E.Gss;
    if (a == 0 || b == 0) return E.Gsp;
    return a == b ? E.Gpp : E.Gpp2;
}
__device__ inline double h1c(const ElemDev& E, int a, int b) {
    if (a == b) return g1c(E, a, b);
    if (a == 0 || b == 0) return E.Hsp;
    return E.Hpp;
}

// ------------------------------------------------------------------ Fock
// Per pair (A<B): the off-diagonal block F_AB = beta_bar S - 1/2 sum P_AB (mu s|lam l)
// written to both triangles, and the Coulomb partials of the pair to part[slot]:
//   J_A[mu nu] = sum_{lam sig in B} P_lam,sig (mu nu|lam sig)   (slots 0..9)
//   J_B[lam sig] = sum_{mu nu in A} P_mu,nu (mu nu|lam sig)      (slots 10..19)
// Jobs in their FP32 (mixed-precision) iterations read the FP32 copy of the
// integrals and contract in FP32; F_AB = H_AB - 1/2 K is assembled in FP64.
// WPB wavefronts (pairs) per workgroup, picked by the autotuner; the per-pair
// work does not depend on it.
template <int WPB>
__global__ void __launch_bounds__(64 * WPB) k_fock_pairs(
    int mode, const FockPairDev* __restrict__ fpl, int npairs, const JobDev* __restrict__ jobs,
    const JobState* __restrict__ state, const double* __restrict__ ints, const float* __restrict__ ints32,
    double* __restrict__ arena, double* __restrict__ part) {
    __shared__ double sI[WPB][100], sPab[WPB][16], sWa[WPB][10], sWb[WPB][10];
    __shared__ float fI[WPB][100], fPab[WPB][16], fWa[WPB][10], fWb[WPB][10];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int pidx = blockIdx.x * WPB + wave;
    bool act = pidx < npairs;
    FockPairDev pr;
    JobDev jb;
    int phase = -1, iter = 0;
    if (act) {
        pr = fpl[pidx];
        jb = jobs[pr.job];
        phase = state[pr.job].phase;
        iter = state[pr.job].iter;
        act = (mode != 1 && phase == PH_ACTIVE) || (mode != 0 && phase == PH_ENERGY);
    }
    bool guess = false, f32 = false;
    int nA = 0, nB = 0, nap = 0, nbp = 0, aoA = 0, aoB = 0, n = 0;
    const double* P = nullptr;
    if (act) {
        const bool scf = phase == PH_ACTIVE;   // else the final-density Fock of the energy
        f32 = scf && ints32 != nullptr && state[pr.job].fp32 != 0;
        nA = fp_nao(pr.ao_a);
        nB = fp_nao(pr.ao_b);
        nap = nA == 4 ? 10 : 1;
        nbp = nB == 4 ? 10 : 1;
        aoA = fp_ao(pr.ao_a);
        aoB = fp_ao(pr.ao_b);
        n = jb.nao;
        guess = scf && iter == 0 && jb.requires_guess;
        int pk = scf ? (AR_P0 + (iter & 1)) : (AR_P0 + ((iter + 1) & 1));
        P = mat(arena, jb, pk);
        if (!guess && f32) {
            for (int e = lane; e < nap * nbp; e += 64) fI[wave][e] = ints32[pr.ioff + e];
            if (lane < 16) {
                int i = lane >> 2, j = lane & 3;
                fPab[wave][lane] = (i < nA && j < nB) ? (float)P[(long long)(aoA + i) * n + aoB + j] : 0.0f;
            } else if (lane < 26) {
                int q = lane - 16;
                if (q < nap) {
                    int mu = q == 0 ? 0 : (q < 4 ? 0 : (q < 7 ? 1 : (q < 9 ? 2 : 3)));
                    int nu = q < 4 ? q : (q < 7 ? q - 3 : (q < 9 ? q - 5 : 3));
                    fWa[wave][q] = (float)((mu == nu ? 1.0 : 2.0) * P[(long long)(aoA + mu) * n + aoA + nu]);
                }
            } else if (lane >= 32 && lane < 42) {
                int q = lane - 32;
                if (q < nbp) {
                    int mu = q == 0 ? 0 : (q < 4 ? 0 : (q < 7 ? 1 : (q < 9 ? 2 : 3)));
                    int nu = q < 4 ? q : (q < 7 ? q - 3 : (q < 9 ? q - 5 : 3));
                    fWb[wave][q] = (float)((mu == nu ? 1.0 : 2.0) * P[(long long)(aoB + mu) * n + aoB + nu]);
                }
            }
        } else if (!guess) {
            for (int e = lane; e < nap * nbp; e += 64) sI[wave][e] = ints[pr.ioff + e];
            if (lane < 16) {
                int i = lane >> 2, j = lane & 3;
                sPab[wave][lane] = (i < nA && j < nB) ? P[(long long)(aoA + i) * n + aoB + j] : 0.0;
            } else if (lane < 26) {
                int q = lane - 16;
                if (q < nap) {
                    int mu = q == 0 ? 0 : (q < 4 ? 0 : (q < 7 ? 1 : (q < 9 ? 2 : 3)));
                    int nu = q < 4 ? q : (q < 7 ? q - 3 : (q < 9 ? q - 5 : 3));
                    sWa[wave][q] = (mu == nu ? 1.0 : 2.0) * P[(long long)(aoA + mu) * n + aoA + nu];
                }
            } else if (lane >= 32 && lane < 42) {
                int q = lane - 32;
                if (q < nbp) {
                    int mu = q == 0 ? 0 : (q < 4 ? 0 : (q < 7 ? 1 : (q < 9 ? 2 : 3)));
                    int nu = q < 4 ? q : (q < 7 ? q - 3 : (q < 9 ? q - 5 : 3));
                    sWb[wave][q] = (mu == nu ? 1.0 : 2.0) * P[(long long)(aoB + mu) * n + aoB + nu];
                }
            }
        }
    }
    __syncthreads();
    if (!act) return;
    double* F = mat(arena, jb, AR_F);
    const double* H = mat(arena, jb, AR_H);
    if (f32 && !guess) {
        if (lane < 10) {
            if (lane < nap) {
                float s = 0.0f;
                for (int j = 0; j < nbp; ++j) s += fI[wave][lane * nbp + j] * fWb[wave][j];
                part[(long long)pr.slot * 20 + lane] = (double)s;
            }
        } else if (lane >= 16 && lane < 26) {
            int j = lane - 16;
            if (j < nbp) {
                float s = 0.0f;
                for (int i = 0; i < nap; ++i) s += fI[wave][i * nbp + j] * fWa[wave][i];
                part[(long long)pr.slot * 20 + 10 + j] = (double)s;
            }
        } else if (lane >= 32 && lane < 48) {
            int mu = (lane - 32) >> 2, la = (lane - 32) & 3;
            if (mu < nA && la < nB) {
                long long ij = (long long)(aoA + mu) * n + aoB + la;
                float k = 0.0f;
                for (int nu = 0; nu < nA; ++nu)
                    for (int sg = 0; sg < nB; ++sg)
                        k += fPab[wave][nu * 4 + sg] * fI[wave][pk_index(mu, nu) * nbp + pk_index(la, sg)];
                const double v = H[ij] - 0.5 * (double)k;
                F[ij] = v;
                F[(long long)(aoB + la) * n + aoA + mu] = v;
            }
        }
        return;
    }
    if (lane < 10) {
        if (!guess && lane < nap) {
            double s = 0.0;
            for (int j = 0; j < nbp; ++j) s += sI[wave][lane * nbp + j] * sWb[wave][j];
            part[(long long)pr.slot * 20 + lane] = s;
        }
    } else if (lane >= 16 && lane < 26) {
        int j = lane - 16;
        if (!guess && j < nbp) {
            double s = 0.0;
            for (int i = 0; i < nap; ++i) s += sI[wave][i * nbp + j] * sWa[wave][i];
            part[(long long)pr.slot * 20 + 10 + j] = s;
        }
    } else if (lane >= 32 && lane < 48) {
        int mu = (lane - 32) >> 2, la = (lane - 32) & 3;
        if (mu < nA && la < nB) {
            long long ij = (long long)(aoA + mu) * n + aoB + la;
            double v = H[ij];
            if (!guess) {
                double k = 0.0;
                for (int nu = 0; nu < nA; ++nu)
                    for (int sg = 0; sg < nB; ++sg)
                        k += sPab[wave][nu * 4 + sg] * sI[wave][pk_index(mu, nu) * nbp + pk_index(la, sg)];
                v -= 0.5 * k;
            }
            F[ij] = v;
            F[(long long)(aoB + la) * n + aoA + mu] = v;
        }
    }
}

// H-H and heavy-H pairs (1 and 10 packed integrals, about 80% of the pairs of
// an organic or water molecule), one lane per pair: the same arithmetic, in
// the same order, as k_fock_pairs, whose wavefront per pair these pairs would
// leave mostly idle.
template <typename R>
__device__ inline R wpack(const double* P, int n, int ao, int q) {   // packed one-centre density, off-diagonal doubled
    const int mu = q == 0 ? 0 : (q < 4 ? 0 : (q < 7 ? 1 : (q < 9 ? 2 : 3)));
    const int nu = q < 4 ? q : (q < 7 ? q - 3 : (q < 9 ? q - 5 : 3));
    return (R)((mu == nu ? 1.0 : 2.0) * P[(long long)(ao + mu) * n + ao + nu]);
}

// The three light cases spelled out, so that every array index is a constant
// (registers, no scratch).  Sums start from zero and run in the order of
// k_fock_pairs.
template <typename R>
__device__ inline void fock_light_pair(int nA, int nB, const R* I, const double* P, int n, int aoA, int aoB,
                                       const double* H, double* F, double* pp, bool guess) {
    if (nA == 1 && nB == 1) {   // H-H: one integral
        const long long ij = (long long)aoA * n + aoB;
        double v = H[ij];
        if (!guess) {
            const R wa = wpack<R>(P, n, aoA, 0), wb = wpack<R>(P, n, aoB, 0);
            R s = (R)0;
            s += I[0] * wb;
            pp[0] = (double)s;
            s = (R)0;
            s += I[0] * wa;
            pp[10] = (double)s;
            R k = (R)0;
            k += (R)P[ij] * I[0];
            v -= 0.5 * (double)k;
        }
        F[ij] = v;
        F[(long long)aoB * n + aoA] = v;
        return;
    }
    if (nA == 4) {   // heavy A, hydrogen B: I[q], q packed on A
        if (!guess) {
            const R wb = wpack<R>(P, n, aoB, 0);
            R wa[10];
#pragma unroll
            for (int q = 0; q < 10; ++q) wa[q] = wpack<R>(P, n, aoA, q);
#pragma unroll
            for (int q = 0; q < 10; ++q) {
                R s = (R)0;
                s += I[q] * wb;
                pp[q] = (double)s;
            }
            R s = (R)0;
#pragma unroll
            for (int i = 0; i < 10; ++i) s += I[i] * wa[i];
            pp[10] = (double)s;
        }
        R pab[4];
#pragma unroll
        for (int nu = 0; nu < 4; ++nu) pab[nu] = (R)P[(long long)(aoA + nu) * n + aoB];
#pragma unroll
        for (int mu = 0; mu < 4; ++mu) {
            const long long ij = (long long)(aoA + mu) * n + aoB;
            double v = H[ij];
            if (!guess) {
                R k = (R)0;
#pragma unroll
                for (int nu = 0; nu < 4; ++nu) k += pab[nu] * I[pk_index(mu, nu)];
                v -= 0.5 * (double)k;
            }
            F[ij] = v;
            F[(long long)aoB * n + aoA + mu] = v;
        }
        return;
    }
    // hydrogen A, heavy B: I[j], j packed on B
    if (!guess) {
        const R wa = wpack<R>(P, n, aoA, 0);
        R wb[10];
#pragma unroll
        for (int j = 0; j < 10; ++j) wb[j] = wpack<R>(P, n, aoB, j);
        R s = (R)0;
#pragma unroll
        for (int j = 0; j < 10; ++j) s += I[j] * wb[j];
        pp[0] = (double)s;
#pragma unroll
        for (int j = 0; j < 10; ++j) {
            R t = (R)0;
            t += I[j] * wa;
            pp[10 + j] = (double)t;
        }
    }
    R pab[4];
#pragma unroll
    for (int sg = 0; sg < 4; ++sg) pab[sg] = (R)P[(long long)aoA * n + aoB + sg];
#pragma unroll
    for (int la = 0; la < 4; ++la) {
        const long long ij = (long long)aoA * n + aoB + la;
        double v = H[ij];
        if (!guess) {
            R k = (R)0;
#pragma unroll
            for (int sg = 0; sg < 4; ++sg) k += pab[sg] * I[pk_index(la, sg)];
            v -= 0.5 * (double)k;
        }
        F[ij] = v;
        F[(long long)(aoB + la) * n + aoA] = v;
    }
}

__global__ void __launch_bounds__(256) k_fock_light(
    int mode, const FockPairDev* __restrict__ fpl, int nlight, const JobDev* __restrict__ jobs,
    const JobState* __restrict__ state, const double* __restrict__ ints, const float* __restrict__ ints32,
    double* __restrict__ arena, double* __restrict__ part) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= nlight) return;
    const FockPairDev pr = fpl[idx];
    const JobState& st = state[pr.job];
    const int phase = st.phase, iter = st.iter;
    if (!((mode != 1 && phase == PH_ACTIVE) || (mode != 0 && phase == PH_ENERGY))) return;
    const bool scf = phase == PH_ACTIVE;   // else the final-density Fock of the energy
    const JobDev& jb = jobs[pr.job];
    const int nA = fp_nao(pr.ao_a), nB = fp_nao(pr.ao_b), aoA = fp_ao(pr.ao_a), aoB = fp_ao(pr.ao_b);
    const bool guess = scf && iter == 0 && jb.requires_guess;
    const bool f32 = scf && ints32 != nullptr && st.fp32 != 0;
    const double* P = mat(arena, jb, scf ? (AR_P0 + (iter & 1)) : (AR_P0 + ((iter + 1) & 1)));
    double* F = mat(arena, jb, AR_F);
    const double* H = mat(arena, jb, AR_H);
    double* pp = part + (long long)pr.slot * 20;
    const int ni = (nA == 4 ? 10 : 1) * (nB == 4 ? 10 : 1);
    if (f32 && !guess) {
        float I[10];
#pragma unroll
        for (int e = 0; e < 10; ++e) I[e] = e < ni ? ints32[pr.ioff + e] : 0.0f;
        fock_light_pair<float>(nA, nB, I, P, jb.nao, aoA, aoB, H, F, pp, false);
    } else {
        double I[10];
#pragma unroll
        for (int e = 0; e < 10; ++e) I[e] = (e < ni && !guess) ? ints[pr.ioff + e] : 0.0;
        fock_light_pair<double>(nA, nB, I, P, jb.nao, aoA, aoB, H, F, pp, guess);
    }
}

// Per atom: F_AA = H_AA + one-centre G(P_AA) + sum_B J partials (fixed-order
// reduction over B); on the guess iteration F_AA = diag(U) (Mndo.cpp:422-488, 509-514).
__global__ void __launch_bounds__(256) k_fock_atoms(
    int mode, const AtomDev* __restrict__ atoms, int natoms, const ElemDev* __restrict__ elems,
    const JobDev* __restrict__ jobs, const JobState* __restrict__ state, double* __restrict__ arena,
    const double* __restrict__ part) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ai = blockIdx.x * 4 + wave;
    if (ai >= natoms) return;
    const AtomDev at = atoms[ai];
    const JobState& st = state[at.job];
    const int phase = st.phase, iter = st.iter;
    const bool scf = phase == PH_ACTIVE;   // else the final-density Fock of the energy
    if (!((mode != 1 && scf) || (mode != 0 && phase == PH_ENERGY))) return;
    const JobDev jb = jobs[at.job];
    const ElemDev& E = elems[at.elem];
    const int n = jb.nao, N = jb.natoms, A = at.local, na = E.nao, np = na == 4 ? 10 : 1;
    const bool guess = scf && iter == 0 && jb.requires_guess;
    double* F = mat(arena, jb, AR_F);
    const double* H = mat(arena, jb, AR_H);
    const double* P = mat(arena, jb, scf ? (AR_P0 + (iter & 1)) : (AR_P0 + ((iter + 1) & 1)));
    if (guess) {
        if (lane < 16) {
            int mu = lane >> 2, nu = lane & 3;
            if (mu < na && nu < na)
                F[(long long)(at.ao + mu) * n + at.ao + nu] = mu == nu ? E.U[mu == 0 ? 0 : 1] : 0.0;
        }
        return;
    }
    double acc[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) acc[i] = 0.0;
    for (int Bl = lane; Bl < N; Bl += 64) {
        if (Bl == A) continue;
        const double* p = A < Bl ? part + (jb.pair0 + tri_slot(A, Bl, N)) * 20
                                 : part + (jb.pair0 + tri_slot(Bl, A, N)) * 20 + 10;
#pragma unroll
        for (int i = 0; i < 10; ++i)
            if (i < np) acc[i] += p[i];
    }
#pragma unroll
    for (int i = 0; i < 10; ++i)
        for (int off = 32; off > 0; off >>= 1) acc[i] += __shfl_xor(acc[i], off);
    if (lane < 16) {
        int mu = lane >> 2, nu = lane & 3;
        if (mu < na && nu < na) {
            int pi = pk_index(mu, nu);
            double j = 0.0;
#pragma unroll
            for (int i = 0; i < 10; ++i)
                if (i == pi && i < np) j = acc[i];
            const long long base = (long long)at.ao * n + at.ao;
            double g;
            if (mu == nu) {
                g = 0.0;
                for (int k = 0; k < na; ++k)
                    g += P[base + (long long)k * n + k] * (g1c(E, mu, k) - 0.5 * h1c(E, mu, k));
            } else {
                g = (1.5 * h1c(E, mu, nu) - 0.5 * g1c(E, mu, nu)) * P[base + (long long)mu * n + nu];
            }
            const long long ij = base + (long long)mu * n + nu;
            F[ij] = H[ij] + g + j;
        }
    }
}

void launch_fock(hipStream_t st, int mode, const FockPairs& fp, const AtomDev* atoms, int atom0,
                 int natoms, const ElemDev* elems, const JobDev* jobs, const JobState* state, const double* ints,
                 const float* ints32, double* arena, double* part, hipStream_t st2, hipEvent_t fork_ev,
                 hipEvent_t join_ev) {
    // the light-pair and heavy-pair kernels write disjoint F blocks and pair
    // slots: with a second stream they run concurrently, joined before the
    // per-atom reduction
    const int nl = fp.nlight[0] + fp.nlight[1];
    const bool fork = st2 && nl > 0 && fp.nheavy > 0;
    if (fork) {
        hipEventRecord(fork_ev, st);
        hipStreamWaitEvent(st2, fork_ev, 0);
    }
    for (int c = 0; c < 2; ++c)
        if (fp.nlight[c] > 0)
            hipLaunchKernelGGL(k_fock_light, dim3((fp.nlight[c] + 255) / 256), dim3(256), 0, fork ? st2 : st, mode,
                               fp.light[c], fp.nlight[c], jobs, state, ints, ints32, arena, part);
    if (fork) hipEventRecord(join_ev, st2);
    if (fp.nheavy > 0) {
        const int npairs = fp.nheavy;
        const FockPairDev* hl = fp.heavy;
        const int w = get_tuning().fock_wpb;
#define NDDO_FOCK_PAIRS(W)                                                                                 \
    hipLaunchKernelGGL(k_fock_pairs<W>, dim3((npairs + (W)-1) / (W)), dim3(64 * (W)), 0, st, mode, hl, npairs, \
                       jobs, state, ints, ints32, arena, part)
        if (w == 1) NDDO_FOCK_PAIRS(1);
        else if (w == 2) NDDO_FOCK_PAIRS(2);
        else if (w == 8) NDDO_FOCK_PAIRS(8);
        else NDDO_FOCK_PAIRS(4);
#undef NDDO_FOCK_PAIRS
    }
    if (fork) hipStreamWaitEvent(st, join_ev, 0);
    if (natoms > 0)
        hipLaunchKernelGGL(k_fock_atoms, dim3((natoms + 3) / 4), dim3(256), 0, st, mode, atoms + atom0, natoms, elems,
                           jobs, state, arena, part);
}

// ------------------------------------------------------------------ eigensolver dispatch
// The one-workgroup Householder + divide-and-conquer solver (kernels_eig.hip)
// for the LDS and device-memory classes; molecules beyond it are solved by the
// device-wide block Jacobi (kernels_xl.hip), launched by the engine.
void launch_eigen(hipStream_t st, const JobDev* jobs, JobState* state, const int* joblist, int nlist, int max_n,
                  double* arena, double* eps, int* perm, double* tri, bool a_in_lds) {
    if (nlist <= 0 || !heig_supported(max_n)) return;
    launch_heig_class(st, jobs, state, joblist, nlist, max_n, arena, eps, perm, tri, a_in_lds);
}

// ------------------------------------------------------------------ warm start
// Rotates F into the previous MO basis before the Jacobi solve:
//   pass 0:  W = F VT^T        pass 1:  F <- VT W      (= VT F VT^T)
// Once the SCF settles, F in that basis is nearly diagonal and the cyclic
// Jacobi needs one or two sweeps instead of a cold solve.  64x64 output tile
// per workgroup, 16-deep K slabs staged in LDS, 4x4 outputs per thread.
template <int PASS>
__global__ void __launch_bounds__(256) k_warm_gemm(const TileDev* __restrict__ tiles, const JobDev* __restrict__ jobs,
                                                   const JobState* __restrict__ state, double* __restrict__ arena) {
    const TileDev tl = tiles[blockIdx.x];
    const JobState& st = state[tl.job];
    if (st.phase != PH_ACTIVE || !st.warm) return;
    const JobDev jb = jobs[tl.job];
    const int n = jb.nao;
    double* F = mat(arena, jb, AR_F);
    double* W = mat(arena, jb, AR_W);
    const double* VT = mat(arena, jb, AR_C);
    // PASS 0: out W,  a(i,k) = F[i][k],  b(k,j) = VT[j][k]
    // PASS 1: out F,  a(i,k) = VT[i][k], b(k,j) = W[k][j]
    const double* Am = PASS == 0 ? F : VT;
    const double* Bm = PASS == 0 ? VT : W;
    double* Out = PASS == 0 ? W : F;
    __shared__ double sA[16][68], sB[16][68];
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    double acc[4][4];
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
    for (int k0 = 0; k0 < n; k0 += 16) {
        for (int e = tid; e < 16 * 64; e += 256) {
            // a(i,k): rows i of a row-major matrix, contiguous in k
            const int ii = e >> 4, kk = e & 15, i = tl.i0 + ii, k = k0 + kk;
            sA[kk][ii] = (i < n && k < n) ? Am[(long long)i * n + k] : 0.0;
            if (PASS == 0) {
                const int j = tl.j0 + ii;
                sB[kk][ii] = (j < n && k < n) ? Bm[(long long)j * n + k] : 0.0;
            }
        }
        if (PASS == 1) {
            for (int e = tid; e < 16 * 64; e += 256) {
                const int kk = e >> 6, jj = e & 63, k = k0 + kk, j = tl.j0 + jj;
                sB[kk][jj] = (j < n && k < n) ? Bm[(long long)k * n + j] : 0.0;
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            double av[4], bv[4];
            for (int a = 0; a < 4; ++a) av[a] = sA[kk][ty * 4 + a];
            for (int b = 0; b < 4; ++b) bv[b] = sB[kk][tx * 4 + b];
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b) acc[a][b] += av[a] * bv[b];
        }
        __syncthreads();
    }
    for (int a = 0; a < 4; ++a) {
        const int i = tl.i0 + ty * 4 + a;
        if (i >= n) continue;
        for (int b = 0; b < 4; ++b) {
            const int j = tl.j0 + tx * 4 + b;
            if (j < n) Out[(long long)i * n + j] = acc[a][b];
        }
    }
}

void launch_warm(hipStream_t st, const TileDev* tiles, int ntiles, const JobDev* jobs, const JobState* state,
                 double* arena) {
    if (ntiles <= 0) return;
    hipLaunchKernelGGL(k_warm_gemm<0>, dim3(ntiles), dim3(256), 0, st, tiles, jobs, state, arena);
    hipLaunchKernelGGL(k_warm_gemm<1>, dim3(ntiles), dim3(256), 0, st, tiles, jobs, state, arena);
}

// ------------------------------------------------------------------ density
// P = 2 C_occ^T C_occ (rows of C are MOs; occupied = the nocc lowest after the
// eigenvalue sort). One workgroup per upper 64x64 tile (I <= J) computes the
// tile and writes it and its mirror (J, I), so P is exactly symmetric, as
// dsyrk's upper triangle mirrored to the lower (Blas.cpp:716-740). Each tile
// also emits
//  - its partial of sum (P - P_old)^2 over both tiles (the RMS test);
//  - for a DIIS job (not a large one, not the guess iteration), the new ring
//    slot (P and e = P - P_old) on the upper tile and the tile's share of
//    <e_i, e_new> for every ring slot and of max |e_new| (Cndo2.cpp:955-988):
//    e is symmetric, so an off-diagonal tile counts twice for its mirror and
//    the ring holds the upper tiles only.

typedef float flt4 __attribute__((ext_vector_type(4)));

constexpr int TLD = 65;   // LDS row stride of a staged 64x64 tile (transposed reads without bank conflicts)

// Mixed precision: every operand split as x = hi + lo (hi = fp32(x),
// lo = fp32(x - hi)) and P ~ 2 (A_hi B_hi + A_hi B_lo + A_lo B_hi) on
// v_mfma_f32_16x16x4f32 with FP32 accumulation; each wavefront owns a 32x32
// quadrant of the tile (2 x 2 MFMA blocks). The tile is staged in T.
__device__ void density_tile_split_f32(const TileDev& tl, const JobDev& jb, const double* VT, const int* perm,
                                       double (*T)[TLD]) {
    const int n = jb.nao, nocc = jb.nocc;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, lr = lane & 15, lk = lane >> 4;
    const int qi = 32 * (wave >> 1), qj = 32 * (wave & 1);
    const int ib = tl.i0 + qi, jbb = tl.j0 + qj;
    flt4 acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = flt4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int o0 = 0; o0 < nocc; o0 += 4) {
        const int o = o0 + lk;
        const long long row = o < nocc ? (long long)perm[jb.ivec0 + o] * n : -1;
        float ah[2], al[2], bh[2], bl[2];
#pragma unroll
        for (int x = 0; x < 2; ++x) {
            const int i = ib + 16 * x + lr, j = jbb + 16 * x + lr;
            const double a = (row >= 0 && i < n) ? VT[row + i] : 0.0;
            const double b = (row >= 0 && j < n) ? VT[row + j] : 0.0;
            ah[x] = (float)a;
            al[x] = (float)(a - (double)ah[x]);
            bh[x] = (float)b;
            bl[x] = (float)(b - (double)bh[x]);
        }
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[x], bh[y], acc[x][y], 0, 0, 0);
                acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[x], bl[y], acc[x][y], 0, 0, 0);
                acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(al[x], bh[y], acc[x][y], 0, 0, 0);
            }
    }
    // D[i][j] of a 16x16 FP32 block sits in lane j + 16 (i / 4), element i % 4
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) T[qi + 16 * x + 4 * lk + v][qj + 16 * y + lr] = 2.0 * (double)acc[x][y][v];
}

__device__ inline double wave_sum_d(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ inline double wave_max_d(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    return v;
}

// Writes the staged tile T (upper tile, value (r, c) = P(i0 + r, j0 + c); a
// diagonal tile takes its upper half for both halves) and its mirror; the RMS
// and DIIS partials of the tile go to rmspart[tile] and tpart[tile * XL_REC ..]
// (k dot products in logical ring order, then max |e_new|).
__device__ void density_store(const TileDev& tl, const JobDev& jb, const JobState& st, const double (*T)[TLD],
                              double* __restrict__ arena, double (*red)[NDDO_MAX_DIIS + 3],
                              double* __restrict__ rmspart, double* __restrict__ tpart) {
    const int n = jb.nao, k = jb.diis_k, iter = st.iter;
    const int tid = threadIdx.x, c = tid & 63, lane = tid & 63, wave = tid >> 6;
    const long long nn = (long long)n * n;
    const double* Pin = mat(arena, jb, AR_P0 + (iter & 1));
    double* Pout = mat(arena, jb, AR_P0 + ((iter + 1) & 1));
    const bool diag = tl.i0 == tl.j0;
    const bool ring = k > 0 && !jb.xl && !(iter == 0 && jb.requires_guess);
    const int rs = ring ? (st.ring_start + 1) % k : 0, slot_new = ring ? (rs + k - 1) % k : 0;
    double* RP = mat(arena, jb, AR_RING);
    double* RE = RP + (long long)k * nn;
    // the older slots (logical 0 .. k-2) with compile-time indices, so the
    // accumulators stay in registers
    double part[NDDO_MAX_DIIS - 1];
    long long roff[NDDO_MAX_DIIS - 1];
#pragma unroll
    for (int q = 0; q < NDDO_MAX_DIIS - 1; ++q) {
        part[q] = 0.0;
        roff[q] = (ring && q < k - 1) ? (long long)((rs + q) % k) * nn : 0;
    }
    double loc = 0.0, self = 0.0, amax = 0.0;
    for (int r = wave; r < 64; r += 4) {
        const int i = tl.i0 + r, j = tl.j0 + c;
        if (i >= n || j >= n) continue;
        const double v = (diag && c < r) ? T[c][r] : T[r][c];
        const long long ij = (long long)i * n + j;
        const double d = v - Pin[ij];
        Pout[ij] = v;
        loc += d * d;
        if (ring) {
            RP[(long long)slot_new * nn + ij] = v;
            RE[(long long)slot_new * nn + ij] = d;
            amax = fmax(amax, fabs(d));
            self += d * d;
#pragma unroll
            for (int q = 0; q < NDDO_MAX_DIIS - 1; ++q)
                if (q < k - 1) part[q] += RE[roff[q] + ij] * d;
        }
    }
    if (!diag) {
        for (int r = wave; r < 64; r += 4) {
            const int i = tl.j0 + r, j = tl.i0 + c;
            if (i >= n || j >= n) continue;
            const double v = T[c][r];
            const long long ij = (long long)i * n + j;
            const double d = v - Pin[ij];
            Pout[ij] = v;
            loc += d * d;
        }
    }
    // fixed-order reductions: butterfly within each wavefront, then the four
    // wavefronts in order
    loc = wave_sum_d(loc);
    if (ring) {
        self = wave_sum_d(self);
        amax = wave_max_d(amax);
#pragma unroll
        for (int q = 0; q < NDDO_MAX_DIIS - 1; ++q)
            if (q < k - 1) part[q] = wave_sum_d(part[q]);
    }
    if (lane == 0) {
        red[wave][0] = loc;
        if (ring) {
#pragma unroll
            for (int q = 0; q < NDDO_MAX_DIIS - 1; ++q)
                if (q < k - 1) red[wave][1 + q] = part[q];
            red[wave][k] = self;
            red[wave][k + 1] = amax;
        }
    }
    __syncthreads();
    if (tid <= (ring ? k + 1 : 0)) {
        const int q = tid;
        double s = red[0][q];
        for (int w = 1; w < 4; ++w) s = q == k + 1 ? fmax(s, red[w][q]) : s + red[w][q];
        const double wt = diag ? 1.0 : 2.0;
        if (q == 0) rmspart[blockIdx.x] = s;
        else if (q <= k) tpart[(long long)blockIdx.x * XL_REC + q - 1] = wt * s;
        else tpart[(long long)blockIdx.x * XL_REC + k] = s;
    }
}

__global__ void __launch_bounds__(256) k_density(const TileDev* __restrict__ tiles, const JobDev* __restrict__ jobs,
                                                 const JobState* __restrict__ state, double* __restrict__ arena,
                                                 const int* __restrict__ perm, double* __restrict__ rmspart,
                                                 double* __restrict__ tpart) {
    const TileDev tl = tiles[blockIdx.x];
    if (tl.i0 > tl.j0) return;   // lower tiles: written by their upper mirror
    const JobState& st = state[tl.job];
    if (st.phase != PH_ACTIVE) return;
    const JobDev jb = jobs[tl.job];
    const int n = jb.nao, nocc = jb.nocc;
    const double* VT = mat(arena, jb, AR_C);
    // the staged tile; the FP64 path's operand panels alias its start
    __shared__ double smem[64 * TLD];
    __shared__ double red[4][NDDO_MAX_DIIS + 3];
    double (*T)[TLD] = reinterpret_cast<double (*)[TLD]>(smem);
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    if (st.fp32) {
        density_tile_split_f32(tl, jb, VT, perm, T);
    } else {
        double (*sA)[68] = reinterpret_cast<double (*)[68]>(smem);
        double (*sB)[68] = sA + 16;
        double acc[4][4];
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
        for (int o0 = 0; o0 < nocc; o0 += 16) {
            for (int e = tid; e < 16 * 64; e += 256) {
                const int kk = e >> 6, ii = e & 63, o = o0 + kk;
                const int row = o < nocc ? perm[jb.ivec0 + o] : -1;
                sA[kk][ii] = (row >= 0 && tl.i0 + ii < n) ? VT[(long long)row * n + tl.i0 + ii] : 0.0;
                sB[kk][ii] = (row >= 0 && tl.j0 + ii < n) ? VT[(long long)row * n + tl.j0 + ii] : 0.0;
            }
            __syncthreads();
#pragma unroll
            for (int kk = 0; kk < 16; ++kk) {
                double av[4], bv[4];
                for (int a = 0; a < 4; ++a) av[a] = sA[kk][ty * 4 + a];
                for (int b = 0; b < 4; ++b) bv[b] = sB[kk][tx * 4 + b];
                for (int a = 0; a < 4; ++a)
                    for (int b = 0; b < 4; ++b) acc[a][b] += av[a] * bv[b];
            }
            __syncthreads();
        }
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) T[ty * 4 + a][tx * 4 + b] = 2.0 * acc[a][b];
    }
    __syncthreads();
    density_store(tl, jb, st, T, arena, red, rmspart, tpart);
}

void launch_density(hipStream_t st, const TileDev* tiles, int ntiles, const JobDev* jobs, const JobState* state,
                    double* arena, const int* perm, double* rmspart, double* tpart) {
    if (ntiles > 0)
        hipLaunchKernelGGL(k_density, dim3(ntiles), dim3(256), 0, st, tiles, jobs, state, arena, perm, rmspart, tpart);
}

// DIIS extrapolation and damping of the ordinary (not large) jobs after k_step
// (Cndo2.cpp:1016-1061), one workgroup per upper tile, written with its mirror
// like the density; k_step leaves the coefficients, the ring start and what
// to apply in coef[job * XL_COEF ..]. Large jobs: k_xl_apply.
__global__ void __launch_bounds__(256) k_apply(const TileDev* __restrict__ tiles, const JobDev* __restrict__ jobs,
                                               double* __restrict__ arena, const double* __restrict__ coef) {
    const TileDev tl = tiles[blockIdx.x];
    if (tl.i0 > tl.j0) return;
    const double* cf = coef + (long long)tl.job * XL_COEF;
    const int flags = (int)cf[XL_COEF_FLAGS];
    if (!flags) return;
    const JobDev jb = jobs[tl.job];
    if (jb.xl) return;
    const int n = jb.nao, k = jb.diis_k, iter = (int)cf[XL_COEF_ITER], rs = (int)cf[XL_COEF_RS];
    const long long nn = (long long)n * n;
    const double* Pin = mat(arena, jb, AR_P0 + (iter & 1));
    double* Pout = mat(arena, jb, AR_P0 + ((iter + 1) & 1));
    const double* RP = mat(arena, jb, AR_RING);
    const double w = jb.damp_weight;
    double cq[NDDO_MAX_DIIS];
    long long so[NDDO_MAX_DIIS];
#pragma unroll
    for (int q = 0; q < NDDO_MAX_DIIS; ++q) {
        cq[q] = q < k ? cf[q] : 0.0;
        so[q] = q < k ? (long long)((rs + q) % k) * nn : 0;
    }
    __shared__ double T[64][TLD];
    const int tid = threadIdx.x, c = tid & 63, wave = tid >> 6;
    const bool diag = tl.i0 == tl.j0;
    for (int r = wave; r < 64; r += 4) {
        const int i = tl.i0 + r, j = tl.j0 + c;
        if (i >= n || j >= n) continue;
        const long long ij = (long long)i * n + j;
        double v;
        if (flags & 1) {
            v = 0.0;
#pragma unroll
            for (int q = 0; q < NDDO_MAX_DIIS; ++q)
                if (q < k) v += cq[q] * RP[so[q] + ij];
        } else {
            v = Pout[ij];
        }
        if (flags & 2) v = v * (1.0 - w) + w * Pin[ij];
        T[r][c] = v;
    }
    __syncthreads();
    for (int r = wave; r < 64; r += 4) {
        const int i = tl.i0 + r, j = tl.j0 + c;
        if (i >= n || j >= n) continue;
        Pout[(long long)i * n + j] = (diag && c < r) ? T[c][r] : T[r][c];
    }
    if (!diag)
        for (int r = wave; r < 64; r += 4) {
            const int i = tl.j0 + r, j = tl.i0 + c;
            if (i >= n || j >= n) continue;
            Pout[(long long)i * n + j] = T[c][r];
        }
}

void launch_apply(hipStream_t st, const TileDev* tiles, int ntiles, const JobDev* jobs, double* arena,
                  const double* coef) {
    if (ntiles > 0) hipLaunchKernelGGL(k_apply, dim3(ntiles), dim3(256), 0, st, tiles, jobs, arena, coef);
}

// ------------------------------------------------------------------ step
// RMS test, density-DIIS decisions and damping decisions, one job per
// workgroup (Cndo2.cpp:663-695, 920-1061), from the tile (or chunk) partials;
// k_apply / k_xl_apply then update the density. The DIIS ring keeps the
// reference's logical order (oldest first) over physical slots starting at
// ring_start.
__device__ bool small_solve(int m, double* a /* m*m row-major, destroyed */, double* b) {
    for (int c = 0; c < m; ++c) {
        int piv = c;
        double best = fabs(a[c * m + c]);
        for (int r = c + 1; r < m; ++r)
            if (fabs(a[r * m + c]) > best) { best = fabs(a[r * m + c]); piv = r; }
        if (best == 0.0) return false;
        if (piv != c) {
            for (int k = 0; k < m; ++k) { double t = a[c * m + k]; a[c * m + k] = a[piv * m + k]; a[piv * m + k] = t; }
            double t = b[c]; b[c] = b[piv]; b[piv] = t;
        }
        for (int r = c + 1; r < m; ++r) {
            double f = a[r * m + c] / a[c * m + c];
            if (f == 0.0) continue;
            for (int k = c; k < m; ++k) a[r * m + k] -= f * a[c * m + k];
            b[r] -= f * b[c];
        }
    }
    for (int r = m - 1; r >= 0; --r) {
        double s = b[r];
        for (int k = r + 1; k < m; ++k) s -= a[r * m + k] * b[k];
        b[r] = s / a[r * m + r];
    }
    return true;
}

__global__ void __launch_bounds__(256) k_step(const JobDev* __restrict__ jobs, JobState* __restrict__ state,
                                              const double* __restrict__ rmspart, const double* __restrict__ tpart,
                                              double* __restrict__ traj, int* __restrict__ remaining,
                                              const double* __restrict__ xlpart, double* __restrict__ xlcoef,
                                              int nchunk, int job0) {
    const int job = job0 + (int)blockIdx.x;
    JobState* st = &state[job];
    const JobDev jb = jobs[job];
    const int tid = threadIdx.x, T = blockDim.x;
    // the element-wise work runs elsewhere: the ring update and the dot products
    // in k_density's tiles (k_xl_dots' chunks for large molecules), the
    // extrapolation and damping in k_apply (k_xl_apply) from the coefficients
    // and flags left here; nothing to apply unless the decisions below say so
    double* xc = xlcoef + (long long)job * XL_COEF;
    if (tid == 0) xc[XL_COEF_FLAGS] = 0.0;
    if (st->phase != PH_ACTIVE) return;
    const int n = jb.nao, iter = st->iter, k = jb.diis_k, tpd = jb.tiles_per_dim;
    const long long nn = (long long)n * n;
    __shared__ double sh_rms;
    __shared__ double red[NDDO_MAX_DIIS + 1][256];
    __shared__ double coef[NDDO_MAX_DIIS + 1];
    __shared__ double sa[(NDDO_MAX_DIIS + 1) * (NDDO_MAX_DIIS + 1)], srhs[NDDO_MAX_DIIS + 1];
    __shared__ int diis_on;
    if (tid == 0) {
        // the upper tiles (I <= J) carry the RMS partials of their mirrors too
        double s = 0.0;
        for (int t = 0; t < jb.ntiles; ++t)
            if (t / tpd <= t % tpd) s += rmspart[jb.tile0 + t];
        sh_rms = sqrt(s / (double)nn);
        double* tr = traj + jb.traj0 + 4LL * iter;
        tr[0] = sh_rms;
        tr[1] = st->diis_error;
        tr[2] = st->diis_on;
        tr[3] = st->damp_on;
    }
    __syncthreads();
    if (st->sweeps < 0) {   // the eigensolver rejected this job's Fock matrix
        if (tid == 0) {
            st->phase = PH_DONE;
            st->status = -5;   // NDDO_EIGENSOLVER_FAILED
            st->rms = sh_rms;
            atomicSub(remaining, 1);
        }
        return;
    }
    const double rms = sh_rms;
    const bool in32 = st->fp32 != 0;
    if (rms < jb.rms_thresh && !in32) {
        if (tid == 0) { st->rms = rms; st->phase = PH_ENERGY; }
        return;
    }
    // a mixed-precision job leaves its FP32 iterations once the density RMS is
    // below max(rms_density, 1e-5) (well above the FP32 noise floor of the Fock
    // contraction) or after half its iteration budget; the FP64 polish
    // iterations that follow decide convergence
    if (in32 && tid == 0 && (rms < fmax(jb.rms_thresh, 1e-5) || 2 * (iter + 1) >= jb.max_iter)) {
        st->fp32 = 0;
        st->switch_iter = iter + 1;
    }
    const bool guess = iter == 0 && jb.requires_guess;
    if (!guess) {
        if (k > 0) {
            const int rs = (st->ring_start + 1) % k;
            // per-slot accumulators with compile-time indices, so they stay in
            // registers; partial records in a fixed order per thread
            double part[NDDO_MAX_DIIS];
#pragma unroll
            for (int i = 0; i < NDDO_MAX_DIIS; ++i) part[i] = 0.0;
            double amax = 0.0;
            if (jb.xl) {
                for (int c = tid; c < nchunk; c += T) {
                    const double* rec = xlpart + ((long long)(jb.xl - 1) * nchunk + c) * XL_REC;
#pragma unroll
                    for (int i = 0; i < NDDO_MAX_DIIS; ++i)
                        if (i < k) part[i] += rec[i];
                    amax = fmax(amax, rec[k]);
                }
            } else {
                for (int t = tid; t < jb.ntiles; t += T) {
                    if (t / tpd > t % tpd) continue;
                    const double* rec = tpart + (long long)(jb.tile0 + t) * XL_REC;
#pragma unroll
                    for (int i = 0; i < NDDO_MAX_DIIS; ++i)
                        if (i < k) part[i] += rec[i];
                    amax = fmax(amax, rec[k]);
                }
            }
#pragma unroll
            for (int i = 0; i < NDDO_MAX_DIIS; ++i)
                if (i < k) red[i][tid] = part[i];
            red[k][tid] = amax;
            __syncthreads();
            for (int s = T / 2; s > 0; s >>= 1) {
                if (tid < s) {
                    for (int i = 0; i < k; ++i) red[i][tid] += red[i][tid + s];
                    red[k][tid] = fmax(red[k][tid], red[k][tid + s]);
                }
                __syncthreads();
            }
            if (tid == 0) {
                const int ld = NDDO_MAX_DIIS + 1;
                double* B = st->B;
                for (int a = 0; a < k - 1; ++a)
                    for (int b = 0; b < k - 1; ++b) B[a * ld + b] = B[(a + 1) * ld + b + 1];
                for (int i = 0; i < k; ++i) {
                    B[(k - 1) * ld + i] = red[i][0];
                    B[i * ld + (k - 1)] = red[i][0];
                    B[i * ld + k] = -1.0;
                    B[k * ld + i] = -1.0;
                }
                B[k * ld + k] = 0.0;
                const double err = red[k][0];
                st->diis_error = err;
                int on = 0;
                if (k <= iter && jb.diis_end < err && err < jb.diis_start) {
                    for (int i = 0; i <= k; ++i) {
                        srhs[i] = i == k ? -1.0 : 0.0;
                        for (int j = 0; j <= k; ++j) sa[i * (k + 1) + j] = B[i * ld + j];
                    }
                    on = small_solve(k + 1, sa, srhs) ? 1 : 0;
                    for (int i = 0; i <= k; ++i) coef[i] = srhs[i];
                }
                diis_on = on;
                st->diis_on = on;
                st->ring_start = rs;
            }
            __syncthreads();
            if (tid == 0) {
                for (int i = 0; i < k; ++i) xc[i] = coef[i];
                xc[XL_COEF_FLAGS] = diis_on ? 1.0 : 0.0;
                xc[XL_COEF_ITER] = (double)iter;
                xc[XL_COEF_RS] = (double)rs;
            }
        }
        const double w = jb.damp_weight;
        const int don = (0.0 < w && jb.damp_thresh < rms) ? 1 : 0;
        if (tid == 0) {
            st->damp_on = don;
            if (don) {
                xc[XL_COEF_FLAGS] += 2.0;
                xc[XL_COEF_ITER] = (double)iter;
                if (k <= 0) xc[XL_COEF_RS] = 0.0;
            }
        }
    }
    if (tid == 0) {
        st->iter = iter + 1;
        if (iter + 1 >= jb.max_iter) {
            st->phase = PH_DONE;
            st->status = -1;   // NDDO_NOT_CONVERGED
            st->rms = rms;
            atomicSub(remaining, 1);
        }
    }
}

void launch_step(hipStream_t st, const JobDev* jobs, int njobs, JobState* state, const double* rmspart,
                 const double* tpart, double* traj, int* remaining, const double* xlpart, double* xlcoef, int nchunk,
                 int job0) {
    if (njobs > 0)
        hipLaunchKernelGGL(k_step, dim3(njobs), dim3(256), 0, st, jobs, state, rmspart, tpart, traj, remaining, xlpart,
                           xlcoef, nchunk, job0);
}

// ------------------------------------------------------------------ energy
// E_el = 1/2 sum_ij P_ij (F_ij + H_ij) with F built from the final P
// (Cndo2.cpp:1348-1381); compensated summation.
__global__ void __launch_bounds__(256) k_energy(const JobDev* __restrict__ jobs, JobState* __restrict__ state,
                                                double* __restrict__ arena, int* __restrict__ remaining,
                                                const double* __restrict__ xlpart, int nchunk, int job0) {
    const int job = job0 + (int)blockIdx.x;
    JobState* st = &state[job];
    if (st->phase != PH_ENERGY) return;
    const JobDev jb = jobs[job];
    const int n = jb.nao, iter = st->iter, tid = threadIdx.x;
    const long long nn = (long long)n * n;
    const double* P = mat(arena, jb, AR_P0 + ((iter + 1) & 1));
    const double* F = mat(arena, jb, AR_F);
    const double* H = mat(arena, jb, AR_H);
    double s = 0.0, c = 0.0;
    if (jb.xl) {   // compensated chunk partials of k_xl_energy, fixed order
        for (int ch = tid; ch < nchunk; ch += blockDim.x) {
            const double* rec = xlpart + ((long long)(jb.xl - 1) * nchunk + ch) * XL_REC;
            const double y = (rec[0] - rec[1]) - c;
            const double t = s + y;
            c = (t - s) - y;
            s = t;
        }
    }
    for (long long e = jb.xl ? nn : tid; e < nn; e += blockDim.x) {
        const double y = P[e] * (F[e] + H[e]) - c;
        const double t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __shared__ double ss[256], sc[256];
    ss[tid] = s;
    sc[tid] = c;
    __syncthreads();
    if (tid == 0) {
        double S = 0.0, C = 0.0;
        for (int i = 0; i < (int)blockDim.x; ++i) {
            const double y = ss[i] - sc[i] - C;
            const double t = S + y;
            C = (t - S) - y;
            S = t;
        }
        st->e_el = 0.5 * S;
        st->phase = PH_DONE;
        st->status = 0;
        atomicSub(remaining, 1);
    }
}

void launch_energy(hipStream_t st, const JobDev* jobs, int njobs, JobState* state, double* arena, int* remaining,
                   const double* xlpart, int nchunk, int job0) {
    if (njobs > 0)
        hipLaunchKernelGGL(k_energy, dim3(njobs), dim3(256), 0, st, jobs, state, arena, remaining, xlpart, nchunk,
                           job0);
}

// count records of `words` 32-bit words, src_stride / dst_stride bytes apart,
// from device memory into mapped pinned host memory, with plain vector stores.
__global__ void __launch_bounds__(256) k_publish(const unsigned* __restrict__ src, long long src_stride_w,
                                                 unsigned* __restrict__ dst, long long dst_stride_w, int words,
                                                 int count) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)words * count) return;
    const long long r = i / words, w = i - r * words;
    dst[r * dst_stride_w + w] = src[r * src_stride_w + w];
}

void launch_publish(hipStream_t st, const void* src, size_t src_stride, void* dst_dev, size_t dst_stride,
                    size_t bytes, int count) {
    if (count <= 0 || bytes == 0) return;
    const int words = (int)(bytes / 4);
    const long long total = (long long)words * count;
    hipLaunchKernelGGL(k_publish, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, (const unsigned*)src,
                       (long long)(src_stride / 4), (unsigned*)dst_dev, (long long)(dst_stride / 4), words, count);
}

}  // namespace nddo
