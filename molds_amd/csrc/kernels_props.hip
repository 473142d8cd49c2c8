// SCF properties of converged jobs: the ground-state dipole moment.
//
// MolDS evaluates the position matrix <mu| r |nu> between the Slater valence
// AOs expanded in STO-6G Gaussians (Cndo2::CalcCartesianMatrixByGTOExpansion,
// src/cndo/Cndo2.cpp:1946-2131) and takes the electronic dipole about the
// centre of core masses as -sum P r + centre * sum P S
// (CalcElectronicTransitionDipoleMoment, Cndo2.cpp:1899-1934).  Here the
// position integrals between normalized primitive Gaussians come from the
// Obara-Saika 1-D overlap recurrence (Obara & Saika, J. Chem. Phys. 84, 3963
// (1986)), with x = (x - A_x) + A_x raising the bra's angular momentum; the
// off-atom overlap is taken from the core Hamiltonian's resonance block,
// S_mu,nu = H_mu,nu / (beta_mu + beta_nu) * 2, and S = 1 on the diagonal.
//
// k_dipole: one workgroup per (AO row mu, job); its threads take the columns
// nu >= mu and write four partial sums per row (sum P x, sum P y, sum P z,
// sum P S over the row's upper part, off-diagonal terms doubled); the host
// adds the rows in order.
#include <hip/hip_runtime.h>
#include <math.h>

#include "launch.h"
#include "nddo_device.h"

namespace nddo {

namespace {

// STO-6G fits of unit-exponent Slater orbitals (Stewart, J. Chem. Phys. 52,
// 431 (1970)): 1s, 2s, 2p exponents and coefficients of normalized primitives.
__constant__ double c_sto_exp[3][6] = {
    {23.10303149, 4.235915534, 1.185056519, 0.4070988982, 0.1580884151, 0.06510953954},
    {27.68496241, 5.077140627, 1.42678605, 0.2040335729, 0.09260298399, 0.04416183978},
    {5.868285913, 1.530329631, 0.5475665231, 0.2288932733, 0.1046655969, 0.04948220127}};
__constant__ double c_sto_coef[3][6] = {
    {0.00916359628, 0.04936149294, 0.1685383049, 0.3705627997, 0.4164915298, 0.1303340841},
    {-0.004151277819, -0.02067024148, -0.05150303337, 0.3346271174, 0.5621061301, 0.1712994697},
    {0.007924233646, 0.05144104825, 0.189840006, 0.4049863191, 0.4012362861, 0.1051855189}};

struct AoDesc {
    int atom;        // global atom index
    int shell;       // 0: 1s, 1: 2s, 2: 2p
    int dir;         // Cartesian direction of a p AO (0 x, 1 y, 2 z), -1 for s
    double zeta;
};

// AO k of an atom in MolDS's order s, py, pz, px.
__device__ inline AoDesc ao_desc(int atom, const ElemDev& el, int k) {
    AoDesc d;
    d.atom = atom;
    d.shell = el.nao == 1 ? 0 : (k == 0 ? 1 : 2);
    d.dir = k == 0 ? -1 : (k == 1 ? 1 : (k == 2 ? 2 : 0));
    d.zeta = el.zeta[k == 0 ? 0 : 1];
    return d;
}

// <a| r_c |b> (c = 0, 1, 2, absolute coordinates) between two contracted AOs.
__device__ void position_integrals(const AoDesc& a, const double* A, const AoDesc& b, const double* B, double out[3]) {
    const double pi = 3.14159265358979323846;
    double R2 = 0.0;
    for (int c = 0; c < 3; ++c) R2 += (A[c] - B[c]) * (A[c] - B[c]);
    out[0] = out[1] = out[2] = 0.0;
    for (int i = 0; i < 6; ++i) {
        const double ea = a.zeta * a.zeta * c_sto_exp[a.shell][i];
        const double na = pow(2.0 * ea / pi, 0.75) * (a.dir >= 0 ? 2.0 * sqrt(ea) : 1.0) * c_sto_coef[a.shell][i];
        for (int j = 0; j < 6; ++j) {
            const double eb = b.zeta * b.zeta * c_sto_exp[b.shell][j];
            const double nb = pow(2.0 * eb / pi, 0.75) * (b.dir >= 0 ? 2.0 * sqrt(eb) : 1.0) * c_sto_coef[b.shell][j];
            const double p = ea + eb, h = 0.5 / p;
            const double pre = sqrt(pi / p);
            const double k3 = exp(-ea * eb / p * R2);
            // per direction: E[i][j] for i <= 2, j <= 1 of (x - A)^i, (x - B)^j (the common
            // factor exp(-mu R^2) applied once)
            double E[3][3][2];
            for (int c = 0; c < 3; ++c) {
                const double P = (ea * A[c] + eb * B[c]) / p;
                const double xa = P - A[c], xb = P - B[c];
                const double e00 = pre;
                const double e10 = xa * e00, e01 = xb * e00;
                const double e20 = xa * e10 + h * e00;
                E[c][0][0] = e00;
                E[c][1][0] = e10;
                E[c][0][1] = e01;
                E[c][1][1] = xb * e10 + h * e00;
                E[c][2][0] = e20;
                E[c][2][1] = xb * e20 + 2.0 * h * e10;
            }
            int la[3], lb[3];
            for (int c = 0; c < 3; ++c) { la[c] = c == a.dir; lb[c] = c == b.dir; }
            const double ov = E[0][la[0]][lb[0]] * E[1][la[1]][lb[1]] * E[2][la[2]][lb[2]];
            const double w = na * nb * k3;
            for (int c = 0; c < 3; ++c) {
                const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
                const double raised = E[c][la[c] + 1][lb[c]] * E[c1][la[c1]][lb[c1]] * E[c2][la[c2]][lb[c2]];
                out[c] += w * (raised + A[c] * ov);
            }
        }
    }
}

// The atom holding AO nu of a job (atoms of a job are in AO order).
__device__ inline int atom_of_ao(const AtomDev* atoms, int atom0, int natoms, int nu) {
    int lo = 0, hi = natoms - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (atoms[atom0 + mid].ao <= nu) lo = mid;
        else hi = mid - 1;
    }
    return atom0 + lo;
}

__global__ void __launch_bounds__(256) k_dipole(const JobDev* __restrict__ jobs, const JobState* __restrict__ state,
                                                const int* __restrict__ joblist, const long long* __restrict__ doff,
                                                const AtomDev* __restrict__ atoms, const ElemDev* __restrict__ elems,
                                                const double* __restrict__ arena, double* __restrict__ part) {
    __shared__ double red[4][4];
    const int job = joblist[blockIdx.y];
    const JobDev jb = jobs[job];
    const int n = jb.nao, mu = blockIdx.x;
    if (mu >= n) return;
    const long long nn = (long long)n * n;
    const double* P = arena + jb.arena + (long long)(AR_P0 + ((state[job].iter + 1) & 1)) * nn;
    const double* H = arena + jb.arena + (long long)AR_H * nn;
    const int am = atom_of_ao(atoms, jb.atom0, jb.natoms, mu);
    const AtomDev& Am = atoms[am];
    const ElemDev& Em = elems[Am.elem];
    const AoDesc dm = ao_desc(am, Em, mu - Am.ao);
    const double Apos[3] = {Am.x, Am.y, Am.z};
    const double bm = Em.beta[mu == Am.ao ? 0 : 1];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int nu = mu + threadIdx.x; nu < n; nu += blockDim.x) {
        const int an = atom_of_ao(atoms, jb.atom0, jb.natoms, nu);
        const AtomDev& An = atoms[an];
        const ElemDev& En = elems[An.elem];
        const AoDesc dn = ao_desc(an, En, nu - An.ao);
        const double Bpos[3] = {An.x, An.y, An.z};
        double r[3];
        position_integrals(dm, Apos, dn, Bpos, r);
        const double pmn = P[(long long)mu * n + nu];
        const double wgt = nu == mu ? 1.0 : 2.0;
        double s = 0.0;
        if (nu == mu) s = 1.0;
        else if (an != am) s = H[(long long)mu * n + nu] / (0.5 * (bm + En.beta[nu == An.ao ? 0 : 1]));
        for (int c = 0; c < 3; ++c) acc[c] += wgt * pmn * r[c];
        acc[3] += wgt * pmn * s;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = 0; k < 4; ++k) {
        double v = acc[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) red[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double v = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) v += red[w][threadIdx.x];
        part[doff[blockIdx.y] + 4LL * mu + threadIdx.x] = v;
    }
}

}  // namespace

void launch_dipole(hipStream_t st, const JobDev* jobs, const JobState* state, const int* joblist,
                   const long long* doff, int nlist, int max_n, const AtomDev* atoms, const ElemDev* elems,
                   const double* arena, double* part) {
    if (nlist <= 0 || max_n <= 0) return;
    hipLaunchKernelGGL(k_dipole, dim3(max_n, nlist), dim3(256), 0, st, jobs, state, joblist, doff, atoms, elems, arena,
                       part);
}

}  // namespace nddo
