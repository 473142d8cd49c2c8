// Per-geometry setup kernels: two-centre two-electron integrals, overlap,
// core Hamiltonian and core repulsion: one lane per pair with an H atom, one
// wavefront per pair of heavy atoms.
//
// Reference path (MolDS): Mndo::CalcTwoElecsTwoAtomCores (src/mndo/Mndo.cpp:3703-3820)
// -> CalcDiatomicTwoElecsTwoCores (:3948-4025) -> GetNddoRepulsionIntegral
// (:4886-5238) -> GetSemiEmpiricalMultipoleInteraction (:6592-6834) and the
// rotation RotateDiatomicTwoElecsTwoCoresToSpaceFrame (:4223-4293);
// Cndo2::CalcOverlapAOs (src/cndo/Cndo2.cpp:4030-4146, 5801-5869, 6091-6149);
// the core terms Mndo::GetElectronCoreAttraction (Mndo.cpp:643-650) and the
// core repulsion (Mndo.cpp:223-240, 319-327; src/am1/Am1.cpp:145-170).
//
// Multipoles are evaluated as their point-charge configurations (Dewar and
// Thiel, Theor. Chim. Acta 46, 89 (1977)): every multipole of the NDDO
// expansion is a set of charges +-1/2 or +-1/4 at offsets of D, and the
// Klopman-Ohno interaction of two multipoles is
//     sum_ij q_i q_j / sqrt(|R_B + d_j - d_i|^2 + (rho_A + rho_B)^2),
// which equals the closed forms (52)-(66) term by term.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <math.h>

#include "nddo_device.h"
#include "launch.h"

namespace nddo {

// Overlap polynomial coefficients (angular factor folded in), 32 cases of
// (na-1, la, nb-1, lb, m) times 5x5 powers (xi^i eta^j).  Filled by the host.
__constant__ double c_ovl[32 * 25];

enum { MQ = 0, MUX, MUY, MUZ, QXX, QYY, QZZ, QXZ, QYZ };

// Local-frame charge distribution of packed AO pair p (s,py,pz,px order):
// up to two multipoles; -1 = unused; p = 6 (py px) is the special case.
__constant__ int c_dist[10][2] = {
    {MQ, -1}, {MUY, -1}, {MUZ, -1}, {MUX, -1}, {MQ, QYY},
    {QYZ, -1}, {-1, -1}, {MQ, QZZ}, {QXZ, -1}, {MQ, QXX}};
// packed index -> (mu, nu)
__constant__ int c_pk[10][2] = {{0, 0}, {0, 1}, {0, 2}, {0, 3}, {1, 1},
                                {1, 2}, {1, 3}, {2, 2}, {2, 3}, {3, 3}};

__device__ inline int mp_kind(int t) { return t == MQ ? 0 : (t <= MUZ ? 1 : 2); }
__device__ inline int parity_x(int t) { return (t == MUX || t == QXZ) ? 1 : 0; }
__device__ inline int parity_y(int t) { return (t == MUY || t == QYZ) ? 1 : 0; }

struct Charge { double q, x, y, z; };

__device__ int charges(int t, double D, Charge* c) {
    switch (t) {
    case MQ: c[0] = {1.0, 0, 0, 0}; return 1;
    case MUX: c[0] = {0.5, D, 0, 0}; c[1] = {-0.5, -D, 0, 0}; return 2;
    case MUY: c[0] = {0.5, 0, D, 0}; c[1] = {-0.5, 0, -D, 0}; return 2;
    case MUZ: c[0] = {0.5, 0, 0, D}; c[1] = {-0.5, 0, 0, -D}; return 2;
    case QXX: c[0] = {0.25, 2 * D, 0, 0}; c[1] = {0.25, -2 * D, 0, 0}; c[2] = {-0.5, 0, 0, 0}; return 3;
    case QYY: c[0] = {0.25, 0, 2 * D, 0}; c[1] = {0.25, 0, -2 * D, 0}; c[2] = {-0.5, 0, 0, 0}; return 3;
    case QZZ: c[0] = {0.25, 0, 0, 2 * D}; c[1] = {0.25, 0, 0, -2 * D}; c[2] = {-0.5, 0, 0, 0}; return 3;
    case QXZ:
        c[0] = {0.25, D, 0, D}; c[1] = {-0.25, -D, 0, D};
        c[2] = {-0.25, D, 0, -D}; c[3] = {0.25, -D, 0, -D}; return 4;
    case QYZ:
        c[0] = {0.25, 0, D, D}; c[1] = {-0.25, 0, -D, D};
        c[2] = {-0.25, 0, D, -D}; c[3] = {0.25, 0, -D, -D}; return 4;
    }
    return 0;
}

__device__ __forceinline__ double multipole_pair(const ElemDev& A, const ElemDev& B, int ta, int tb, double r) {
    if (parity_x(ta) != parity_x(tb) || parity_y(ta) != parity_y(tb)) return 0.0;
    int ka = mp_kind(ta), kb = mp_kind(tb);
    double a = A.rho[ka] + B.rho[kb];
    double a2 = a * a;
    // four charge slots each, unused ones zero: the loops then have fixed trip
    // counts (registers, not scratch) and the zero terms add exact zeros
    Charge ca[4], cb[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) ca[i] = cb[i] = Charge{0.0, 0.0, 0.0, 0.0};
    const int na = charges(ta, A.D[ka], ca), nb = charges(tb, B.D[kb], cb);
    double v = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (i >= na || j >= nb) continue;
            double dx = cb[j].x - ca[i].x, dy = cb[j].y - ca[i].y, dz = r + cb[j].z - ca[i].z;
            v += ca[i].q * cb[j].q / sqrt(dz * dz + dx * dx + dy * dy + a2);
        }
    return v;
}

__device__ __forceinline__ double local_integral_sum(const ElemDev& A, const ElemDev& B, int i, int j, double r) {
    double v = 0.0;
    for (int u = 0; u < 2; ++u) {
        int ta = c_dist[i][u];
        if (ta < 0) continue;
        for (int w = 0; w < 2; ++w) {
            int tb = c_dist[j][w];
            if (tb < 0) continue;
            v += multipole_pair(A, B, ta, tb, r);
        }
    }
    return v;
}

// Diatomic-frame (mu nu|lam sig) for packed indices i (on A), j (on B).
__device__ __forceinline__ double local_integral(const ElemDev& A, const ElemDev& B, int i, int j, double r) {
    if (i == 6 || j == 6) {
        if (i == 6 && j == 6)   // (px py|px py) = 1/2[(px px|px px) - (px px|py py)]
            return 0.5 * (local_integral_sum(A, B, 9, 9, r) - local_integral_sum(A, B, 9, 4, r));
        return 0.0;
    }
    return local_integral_sum(A, B, i, j, r);
}

// Rotation R[new][old] for the p block with local z along A->B, gamma = 0
// (src/cndo/Cndo2.cpp:5498-5531; Euler angles of src/base/EularAngle.cpp:38-61).
__device__ void rotation(double dx, double dy, double dz, double R[4][4]) {
    double rxy = sqrt(dx * dx + dy * dy), r = sqrt(dx * dx + dy * dy + dz * dz);
    double cb = dz / r, sb = rxy / r, ca = 1.0, sa = 0.0;
    if (rxy > 0.0) { ca = dx / rxy; sa = dy / rxy; }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) R[i][j] = 0.0;
    R[ORB_S][ORB_S] = 1.0;
    R[ORB_PY][ORB_PY] = ca; R[ORB_PY][ORB_PZ] = sa * sb; R[ORB_PY][ORB_PX] = sa * cb;
    R[ORB_PZ][ORB_PY] = 0.0; R[ORB_PZ][ORB_PZ] = cb; R[ORB_PZ][ORB_PX] = -sb;
    R[ORB_PX][ORB_PY] = -sa; R[ORB_PX][ORB_PZ] = ca * sb; R[ORB_PX][ORB_PX] = ca * cb;
}

// A_k(p) = int_1^inf x^k e^{-px} dx and B_k(q) = int_-1^1 x^k e^{-qx} dx
// (Pople-Beveridge B.22/B.23; src/cndo/Cndo2.cpp:6385-6447).
__device__ double aux_A(int k, double p) {
    double s = 0.0, fk = 1.0, pm = 1.0;
    for (int i = 2; i <= k; ++i) fk *= i;
    for (int mu = 1; mu <= k + 1; ++mu) {
        pm *= p;
        double fac = 1.0;
        for (int i = 2; i <= k - mu + 1; ++i) fac *= i;
        s += 1.0 / (pm * fac);
    }
    return exp(-p) * fk * s;
}

__device__ double aux_B(int k, double q) {
    if (fabs(q) > 0.0) {
        double t1 = 0.0, t2 = 0.0, fk = 1.0, qm = 1.0;
        for (int i = 2; i <= k; ++i) fk *= i;
        for (int mu = 1; mu <= k + 1; ++mu) {
            qm *= q;
            double fac = 1.0;
            for (int i = 2; i <= k - mu + 1; ++i) fac *= i;
            double c = fk / fac;
            double sg = ((k - mu) % 2 == 0) ? 1.0 : -1.0;
            t1 += c / qm;
            t2 += c * sg / qm;
        }
        return -exp(-q) * t1 - exp(q) * t2;
    }
    return (k % 2 == 0) ? 2.0 / (1.0 + k) : 0.0;
}

__device__ __forceinline__ double overlap_local(const ElemDev& A, const ElemDev& B, int la, int lb, int m, double r) {
    double za = A.zeta[la], zb = B.zeta[lb];
    int na = A.nshell, nb = B.nshell;
    int cid = ((((na - 1) * 2 + la) * 2 + (nb - 1)) * 2 + lb) * 2 + m;
    const double* c = &c_ovl[cid * 25];
    // alpha = zeta_a R and beta = zeta_b R are rounded separately (no FMA), so
    // equal exponents give q == 0 exactly and take the B_k(0) closed form
    const double al = __dmul_rn(za, r), be = __dmul_rn(zb, r);
    double p = 0.5 * (al + be), q = 0.5 * (al - be);
    double Ak[5], Bk[5];
    for (int k = 0; k < 5; ++k) { Ak[k] = aux_A(k, p); Bk[k] = aux_B(k, q); }
    double s = 0.0;
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j)
            if (c[i * 5 + j] != 0.0) s += c[i * 5 + j] * Ak[i] * Bk[j];
    double fa = 1.0, fb = 1.0;
    for (int i = 2; i <= 2 * na; ++i) fa *= i;
    for (int i = 2; i <= 2 * nb; ++i) fb *= i;
    double pre = pow(2.0 * za, na + 0.5) * pow(2.0 * zb, nb + 0.5) / sqrt(fa * fb);
    pre *= pow(r / 2.0, na + nb + 1.0);
    return pre * s;
}

// Pair core repulsion (Mndo.cpp:223-240, 319-327; Am1.cpp:145-170) from the
// rotated (ss|ss) integral gss.
__device__ __forceinline__ double pair_core_repulsion(const ElemDev& A, const ElemDev& B, double r, double gss, int theory) {
    const double ang = 1.0 / 0.5291772;
    double f;
    if (A.Z == 1 && (B.Z == 7 || B.Z == 8))
        f = 1.0 + (r / ang) * exp(-B.alpha * r) + exp(-A.alpha * r);
    else if (B.Z == 1 && (A.Z == 7 || A.Z == 8))
        f = 1.0 + (r / ang) * exp(-A.alpha * r) + exp(-B.alpha * r);
    else
        f = 1.0 + exp(-A.alpha * r) + exp(-B.alpha * r);
    double e = A.core * B.core * gss * f;
    if (theory != 0) {
        double t = 0.0;
        for (int k = 0; k < 4; ++k) {
            t += A.K[k] * exp(-A.L[k] * (r - A.M[k]) * (r - A.M[k]));
            t += B.K[k] * exp(-B.L[k] * (r - B.M[k]) * (r - B.M[k]));
        }
        e += A.core * B.core * t * ang / r;
    }
    return e;
}

// One wavefront per pair (the X-X pairs, plist[nlight ..]).  Writes:
//   ints[ioff + i*nbp + j]  rotated (mu nu|lam sig), packed mu<=nu on A, lam<=sig on B
//   H (arena AR_H) off-diagonal block  beta_bar * S  and its transpose
//   part[slot*20 + 0..9]   -Z_B (mu nu|s s) for A,  part[slot*20+10..19] -Z_A (s s|lam sig) for B
//   ecore[slot]            pair core repulsion
__global__ void __launch_bounds__(256) k_pair_setup(
    const PairDev* __restrict__ pairs, const int* __restrict__ plist, int npairs, const AtomDev* __restrict__ atoms,
    const ElemDev* __restrict__ elems, const JobDev* __restrict__ jobs,
    double* __restrict__ ints, double* __restrict__ arena, double* __restrict__ part,
    double* __restrict__ ecore, double* __restrict__ sdbg) {
    __shared__ double sL[4][100], sT[4][100], sX[4][100];
    __shared__ double sR[4][16], sS[4][16], sSl[4][16];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int pidx = blockIdx.x * 4 + wave;
    const bool live = pidx < npairs;
    PairDev pr = pairs[plist[live ? pidx : 0]];
    const AtomDev aA = atoms[pr.a], aB = atoms[pr.b];
    // parameters read in place: a private copy indexed at run time would live in scratch
    const ElemDev& A = elems[aA.elem];
    const ElemDev& B = elems[aB.elem];
    const int nap = A.nao == 4 ? 10 : 1, nbp = B.nao == 4 ? 10 : 1;
    const double dx = aB.x - aA.x, dy = aB.y - aA.y, dz = aB.z - aA.z;
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    double* L = sL[wave];
    double* T = sT[wave];
    double* X = sX[wave];

    // local-frame integrals and the packed rotation T[i][k]
    for (int e = lane; e < nap * nbp; e += 64) L[e] = local_integral(A, B, e / nbp, e % nbp, r);
    if (lane == 0) rotation(dx, dy, dz, reinterpret_cast<double(*)[4]>(sR[wave]));
    __syncthreads();
    for (int e = lane; e < 100; e += 64) {
        int i = e / 10, k = e % 10;
        int mu = c_pk[i][0], nu = c_pk[i][1], al = c_pk[k][0], be = c_pk[k][1];
        const double* R = sR[wave];
        double t = R[mu * 4 + al] * R[nu * 4 + be];
        if (al != be) t += R[mu * 4 + be] * R[nu * 4 + al];
        T[e] = t;
    }
    __syncthreads();
    // X = T_A L ; I = X T_B^T   (H atoms have the 1x1 identity rotation)
    for (int e = lane; e < nap * nbp; e += 64) {
        int i = e / nbp, j = e % nbp;
        double s = 0.0;
        if (nap == 1) s = L[e];
        else for (int k = 0; k < 10; ++k) s += T[i * 10 + k] * L[k * nbp + j];
        X[e] = s;
    }
    __syncthreads();
    for (int e = lane; e < nap * nbp; e += 64) {
        int i = e / nbp, j = e % nbp;
        double s = 0.0;
        if (nbp == 1) s = X[e];
        else for (int l = 0; l < 10; ++l) s += X[i * nbp + l] * T[j * 10 + l];
        if (live) ints[pr.ioff + e] = s;
        L[e] = s;   // reuse as the rotated block
    }
    // overlap in the diatomic frame: m-matched pairs (s|s),(s|pz),(pz|s),(pz|pz),(px|px),(py|py)
    if (lane < 16) {
        int i = lane >> 2, j = lane & 3;
        double v = 0.0;
        if (i < A.nao && j < B.nao) {
            int la = i == ORB_S ? 0 : 1, lb = j == ORB_S ? 0 : 1;
            int ma = i == ORB_PX ? 1 : (i == ORB_PY ? -1 : 0);
            int mb = j == ORB_PX ? 1 : (j == ORB_PY ? -1 : 0);
            if (ma == mb) v = overlap_local(A, B, la, lb, ma != 0 ? 1 : 0, r);
        }
        sSl[wave][lane] = v;
    }
    __syncthreads();
    if (lane < 16) {   // S = R S_loc R^T
        int i = lane >> 2, j = lane & 3;
        const double* R = sR[wave];
        const double* Sl = sSl[wave];
        double s = 0.0;
        for (int k = 0; k < 4; ++k)
            for (int l = 0; l < 4; ++l) s += R[i * 4 + k] * Sl[k * 4 + l] * R[j * 4 + l];
        sS[wave][lane] = s;
    }
    __syncthreads();
    if (!live) return;
    const JobDev jb = jobs[pr.job];
    const int n = jb.nao;
    double* H = arena + jb.arena + (long long)AR_H * n * n;
    if (lane < 16) {
        int i = lane >> 2, j = lane & 3;
        if (i < A.nao && j < B.nao) {
            double bb = 0.5 * (A.beta[i == ORB_S ? 0 : 1] + B.beta[j == ORB_S ? 0 : 1]);
            double h = bb * sS[wave][lane];
            int mu = aA.ao + i, nu = aB.ao + j;
            H[(long long)mu * n + nu] = h;
            H[(long long)nu * n + mu] = h;
            if (sdbg) {
                sdbg[jb.arena + (long long)mu * n + nu] = sS[wave][lane];
                sdbg[jb.arena + (long long)nu * n + mu] = sS[wave][lane];
            }
        }
    }
    // electron-core attraction partials and the core repulsion
    if (lane < nap) part[(long long)pr.slot * 20 + lane] = -B.core * L[lane * nbp + 0];
    if (lane >= 16 && lane < 16 + nbp) part[(long long)pr.slot * 20 + 10 + (lane - 16)] = -A.core * L[lane - 16];
    if (lane == 63) ecore[pr.slot] = pair_core_repulsion(A, B, r, L[0], jb.theory);
}

// One lane per light pair (H-H, heavy-H, H-heavy: plist[0 .. nlight)), with
// the operations of k_pair_setup in the same order (this file is compiled
// without contraction), so every value is bitwise the same.  Terms that are
// products with an exact zero of the rotation (the H rows and columns) are
// left out: adding them changes at most the sign of a zero.
__device__ inline double sdot_ri(const double R[4][4], const double* Sl, int i) {
    // (R S_loc R^T)[i][s] for an H on B: sum_k R[i][k] S_loc[k][s]
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) s += R[i][k] * Sl[k];
    return s;
}

__global__ void __launch_bounds__(256) k_setup_light(
    const PairDev* __restrict__ pairs, const int* __restrict__ plist, int nlight, const AtomDev* __restrict__ atoms,
    const ElemDev* __restrict__ elems, const JobDev* __restrict__ jobs, double* __restrict__ ints,
    double* __restrict__ arena, double* __restrict__ part, double* __restrict__ ecore, double* __restrict__ sdbg) {
    const int id = blockIdx.x * 256 + threadIdx.x;
    if (id >= nlight) return;
    const PairDev pr = pairs[plist[id]];
    const AtomDev aA = atoms[pr.a], aB = atoms[pr.b];
    const ElemDev& A = elems[aA.elem];
    const ElemDev& B = elems[aB.elem];
    const double dx = aB.x - aA.x, dy = aB.y - aA.y, dz = aB.z - aA.z;
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    const JobDev jb = jobs[pr.job];
    const int n = jb.nao;
    double* H = arena + jb.arena + (long long)AR_H * n * n;
    double* pp = part + (long long)pr.slot * 20;
    auto put_h = [&](int i, int j, double sv) {
        const double bb = 0.5 * (A.beta[i == ORB_S ? 0 : 1] + B.beta[j == ORB_S ? 0 : 1]);
        const int mu = aA.ao + i, nu = aB.ao + j;
        H[(long long)mu * n + nu] = bb * sv;
        H[(long long)nu * n + mu] = bb * sv;
        if (sdbg) {
            sdbg[jb.arena + (long long)mu * n + nu] = sv;
            sdbg[jb.arena + (long long)nu * n + mu] = sv;
        }
    };
    if (A.nao == 1 && B.nao == 1) {   // H-H: one integral, the rotation-invariant (s|s) overlap
        const double g = local_integral(A, B, 0, 0, r);
        ints[pr.ioff] = g;
        put_h(0, 0, overlap_local(A, B, 0, 0, 0, r));
        pp[0] = -B.core * g;
        pp[10] = -A.core * g;
        ecore[pr.slot] = pair_core_repulsion(A, B, r, g, jb.theory);
        return;
    }
    double R[4][4];
    rotation(dx, dy, dz, R);
    // the lane's rotation, element-major in LDS, for the runtime-indexed reads below
    __shared__ double sR[16 * 256];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) sR[(a * 4 + b) * 256 + threadIdx.x] = R[a][b];
    const double* Rl = sR + threadIdx.x;
    double I[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) I[i] = 0.0;
    if (A.nao == 4) {   // heavy on A, H on B: I[i] = sum_k T_A[i][k] L[k] (10 x 1)
        for (int k = 0; k < 10; ++k) {
            const double Lk = local_integral(A, B, k, 0, r);
            const int al = c_pk[k][0], be = c_pk[k][1];
#pragma unroll
            for (int i = 0; i < 10; ++i) {
                const int mu = c_pk[i][0], nu = c_pk[i][1];
                double t = Rl[(mu * 4 + al) * 256] * Rl[(nu * 4 + be) * 256];
                if (al != be) t += Rl[(mu * 4 + be) * 256] * Rl[(nu * 4 + al) * 256];
                I[i] += t * Lk;
            }
        }
        const double Sl[4] = {overlap_local(A, B, 0, 0, 0, r), 0.0, overlap_local(A, B, 1, 0, 0, r), 0.0};
#pragma unroll
        for (int i = 0; i < 4; ++i) put_h(i, 0, sdot_ri(R, Sl, i));
#pragma unroll
        for (int e = 0; e < 10; ++e) {
            ints[pr.ioff + e] = I[e];
            pp[e] = -B.core * I[e];
        }
        pp[10] = -A.core * I[0];
    } else {            // H on A, heavy on B: I[j] = sum_l L[l] T_B[j][l] (1 x 10)
        for (int l = 0; l < 10; ++l) {
            const double Ll = local_integral(A, B, 0, l, r);
            const int al = c_pk[l][0], be = c_pk[l][1];
#pragma unroll
            for (int jj = 0; jj < 10; ++jj) {
                const int mu = c_pk[jj][0], nu = c_pk[jj][1];
                double t = Rl[(mu * 4 + al) * 256] * Rl[(nu * 4 + be) * 256];
                if (al != be) t += Rl[(mu * 4 + be) * 256] * Rl[(nu * 4 + al) * 256];
                I[jj] += Ll * t;
            }
        }
        // (R S_loc R^T)[s][j] = sum_l S_loc[s][l] R[j][l]
        const double Sl[4] = {overlap_local(A, B, 0, 0, 0, r), 0.0, overlap_local(A, B, 0, 1, 0, r), 0.0};
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
            double sv = 0.0;
#pragma unroll
            for (int l = 0; l < 4; ++l) sv += Sl[l] * R[jj][l];
            put_h(0, jj, sv);
        }
#pragma unroll
        for (int e = 0; e < 10; ++e) {
            ints[pr.ioff + e] = I[e];
            pp[10 + e] = -A.core * I[e];
        }
        pp[0] = -B.core * I[0];
    }
    ecore[pr.slot] = pair_core_repulsion(A, B, r, I[0], jb.theory);
}

__device__ inline long long tri_slot(int a, int b, int N) {   // a < b, job-local
    return (long long)a * N - (long long)a * (a + 1) / 2 + (b - a - 1);
}

// Core Hamiltonian diagonal blocks: H_AA = diag(U) + sum_B V_AB  (one wavefront per atom,
// deterministic fixed-order reduction over B).
__global__ void __launch_bounds__(256) k_atom_hcore(
    const AtomDev* __restrict__ atoms, int natoms, const ElemDev* __restrict__ elems,
    const JobDev* __restrict__ jobs, const double* __restrict__ part, double* __restrict__ arena) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ai = blockIdx.x * 4 + wave;
    if (ai >= natoms) return;
    const AtomDev at = atoms[ai];
    const JobDev jb = jobs[at.job];
    const ElemDev E = elems[at.elem];
    const int np = E.nao == 4 ? 10 : 1, N = jb.natoms, A = at.local;
    double acc[10];
    for (int i = 0; i < 10; ++i) acc[i] = 0.0;
    for (int Bl = lane; Bl < N; Bl += 64) {
        if (Bl == A) continue;
        const double* p = A < Bl ? part + (jb.pair0 + tri_slot(A, Bl, N)) * 20
                                 : part + (jb.pair0 + tri_slot(Bl, A, N)) * 20 + 10;
        for (int i = 0; i < np; ++i) acc[i] += p[i];
    }
    for (int i = 0; i < np; ++i)
        for (int off = 32; off > 0; off >>= 1) acc[i] += __shfl_xor(acc[i], off);
    const int n = jb.nao;
    double* H = arena + jb.arena + (long long)AR_H * n * n;
    if (lane < 16) {
        int mu = lane >> 2, nu = lane & 3;
        if (mu < E.nao && nu < E.nao) {
            int a = mu < nu ? mu : nu, b = mu < nu ? nu : mu;
            int pi = np == 1 ? 0 : (a == 0 ? b : (a == 1 ? 3 + b : (a == 2 ? 5 + b : 9)));
            double v = 0.0;
            for (int i = 0; i < np; ++i) if (i == pi) v = acc[i];
            if (mu == nu) v += E.U[mu == 0 ? 0 : 1];
            H[(long long)(at.ao + mu) * n + at.ao + nu] = v;
        }
    }
}

// E_core per job: fixed-order compensated sum over the job's pair slots.
__global__ void __launch_bounds__(256) k_ecore_reduce(const JobDev* __restrict__ jobs,
                                                      const double* __restrict__ ecore,
                                                      JobState* __restrict__ state) {
    const JobDev jb = jobs[blockIdx.x];
    __shared__ double ss[256], sc[256];
    double s = 0.0, c = 0.0;
    for (int i = threadIdx.x; i < jb.npairs; i += blockDim.x) {
        double y = ecore[jb.pair0 + i] - c;
        double t = s + y;
        c = (t - s) - y;
        s = t;
    }
    ss[threadIdx.x] = s;
    sc[threadIdx.x] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        double S = 0.0, C = 0.0;
        for (int i = 0; i < (int)blockDim.x; ++i) {
            double y = ss[i] - sc[i] - C;
            double t = S + y;
            C = (t - S) - y;
            S = t;
        }
        state[blockIdx.x].e_core = S;
    }
}

// One thread per atom a: the pairs (a, b), b > a, of its job, at the slots,
// integral offsets and class-list positions its row starts at.
__global__ void __launch_bounds__(256) k_build_pairs(const AtomDev* __restrict__ atoms,
                                                     const AtomRowDev* __restrict__ rows, int natoms,
                                                     const JobDev* __restrict__ jobs,
                                                     const ElemDev* __restrict__ elems, PairDev* __restrict__ pairs,
                                                     int* __restrict__ plist) {
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= natoms) return;
    const AtomDev at = atoms[g];
    const int nat = jobs[at.job].natoms, a = at.local, atom0 = g - a;
    const AtomRowDev rw = rows[g];
    long long slot = rw.slot0, io = rw.ioff0, c0 = rw.cpos[0], c1 = rw.cpos[1], c2 = rw.cpos[2];
    const bool ha = elems[at.elem].nao == 1;
    for (int b = a + 1; b < nat; ++b, ++slot) {
        const bool hb = elems[atoms[atom0 + b].elem].nao == 1;
        pairs[slot] = PairDev{g, atom0 + b, at.job, (int)slot, io};
        io += (ha ? 1 : 10) * (hb ? 1 : 10);
        const long long pos = (ha && hb) ? c0++ : ((ha || hb) ? c1++ : c2++);
        plist[pos] = (int)slot;
    }
}

void launch_build_pairs(hipStream_t st, const AtomDev* atoms, const AtomRowDev* rows, int natoms, const JobDev* jobs,
                        const ElemDev* elems, PairDev* pairs, int* plist) {
    if (natoms > 0)
        hipLaunchKernelGGL(k_build_pairs, dim3((natoms + 255) / 256), dim3(256), 0, st, atoms, rows, natoms, jobs,
                           elems, pairs, plist);
}

void launch_setup(hipStream_t st, const PairDev* pairs, const int* plist, int nlight, int nheavy,
                  const AtomDev* atoms, int natoms, const ElemDev* elems, const JobDev* jobs, int njobs, double* ints,
                  double* arena, double* part, double* ecore, JobState* state, double* sdbg) {
    if (nlight > 0)
        hipLaunchKernelGGL(k_setup_light, dim3((nlight + 255) / 256), dim3(256), 0, st, pairs, plist, nlight, atoms,
                           elems, jobs, ints, arena, part, ecore, sdbg);
    if (nheavy > 0)
        hipLaunchKernelGGL(k_pair_setup, dim3((nheavy + 3) / 4), dim3(256), 0, st, pairs, plist + nlight, nheavy,
                           atoms, elems, jobs, ints, arena, part, ecore, sdbg);
    hipLaunchKernelGGL(k_atom_hcore, dim3((natoms + 3) / 4), dim3(256), 0, st, atoms, natoms, elems, jobs,
                       part, arena);
    hipLaunchKernelGGL(k_ecore_reduce, dim3(njobs), dim3(256), 0, st, jobs, ecore, state);
}

__global__ void __launch_bounds__(256) k_ints_f32(const double* __restrict__ ints, float* __restrict__ ints32,
                                                 long long nints) {
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < nints; i += (long long)gridDim.x * 256)
        ints32[i] = (float)ints[i];
}

void launch_ints_f32(hipStream_t st, const double* ints, float* ints32, long long nints) {
    if (nints <= 0) return;
    const long long blocks = std::min<long long>((nints + 255) / 256, 65536);
    hipLaunchKernelGGL(k_ints_f32, dim3((unsigned)blocks), dim3(256), 0, st, ints, ints32, nints);
}

int upload_overlap_table(const double* table /* 32*25 */) {
    return hipMemcpyToSymbol(HIP_SYMBOL(c_ovl), table, sizeof(double) * 32 * 25) == hipSuccess ? 0 : -1;
}

}  // namespace nddo
