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

#include "dual.h"
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

// The distance-dependent functions below are templates on the type of r:
// double for the integrals, D3 for their analytic derivatives (k_pair_grad).
template <class T>
__device__ __forceinline__ T multipole_pair(const ElemDev& A, const ElemDev& B, int ta, int tb, T r) {
    if (parity_x(ta) != parity_x(tb) || parity_y(ta) != parity_y(tb)) return T(0.0);
    int ka = mp_kind(ta), kb = mp_kind(tb);
    double a = A.rho[ka] + B.rho[kb];
    double a2 = a * a;
    // four charge slots each, unused ones zero: the loops then have fixed trip
    // counts (registers, not scratch) and the zero terms add exact zeros
    Charge ca[4], cb[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) ca[i] = cb[i] = Charge{0.0, 0.0, 0.0, 0.0};
    const int na = charges(ta, A.D[ka], ca), nb = charges(tb, B.D[kb], cb);
    T v = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (i >= na || j >= nb) continue;
            const double dx = cb[j].x - ca[i].x, dy = cb[j].y - ca[i].y;
            const T dz = r + cb[j].z - ca[i].z;
            v += ca[i].q * cb[j].q / sqrt(dz * dz + dx * dx + dy * dy + a2);
        }
    return v;
}

template <class T>
__device__ __forceinline__ T local_integral_sum(const ElemDev& A, const ElemDev& B, int i, int j, T r) {
    T v = 0.0;
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
template <class T>
__device__ __forceinline__ T local_integral(const ElemDev& A, const ElemDev& B, int i, int j, T r) {
    if (i == 6 || j == 6) {
        if (i == 6 && j == 6)   // (px py|px py) = 1/2[(px px|px px) - (px px|py py)]
            return 0.5 * (local_integral_sum(A, B, 9, 9, r) - local_integral_sum(A, B, 9, 4, r));
        return T(0.0);
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
template <class T>
__device__ T aux_A(int k, T p) {
    T s = 0.0, pm = 1.0;
    double fk = 1.0;
    for (int i = 2; i <= k; ++i) fk *= i;
    for (int mu = 1; mu <= k + 1; ++mu) {
        pm *= p;
        double fac = 1.0;
        for (int i = 2; i <= k - mu + 1; ++i) fac *= i;
        s += 1.0 / (pm * fac);
    }
    return exp(-p) * fk * s;
}

template <class T>
__device__ T aux_B(int k, T q) {
    if (fabs(val(q)) > 0.0) {
        T t1 = 0.0, t2 = 0.0, qm = 1.0;
        double fk = 1.0;
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
    return T((k % 2 == 0) ? 2.0 / (1.0 + k) : 0.0);
}

template <class T>
__device__ __forceinline__ T overlap_local(const ElemDev& A, const ElemDev& B, int la, int lb, int m, T r) {
    double za = A.zeta[la], zb = B.zeta[lb];
    int na = A.nshell, nb = B.nshell;
    int cid = ((((na - 1) * 2 + la) * 2 + (nb - 1)) * 2 + lb) * 2 + m;
    const double* c = &c_ovl[cid * 25];
    // alpha = zeta_a R and beta = zeta_b R are rounded separately (no FMA), so
    // equal exponents give q == 0 exactly and take the B_k(0) closed form
    const T al = mul_rn(za, r), be = mul_rn(zb, r);
    const T p = 0.5 * (al + be), q = 0.5 * (al - be);
    T Ak[5], Bk[5];
    for (int k = 0; k < 5; ++k) { Ak[k] = aux_A(k, p); Bk[k] = aux_B(k, q); }
    T s = 0.0;
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j)
            if (c[i * 5 + j] != 0.0) s += c[i * 5 + j] * Ak[i] * Bk[j];
    double fa = 1.0, fb = 1.0;
    for (int i = 2; i <= 2 * na; ++i) fa *= i;
    for (int i = 2; i <= 2 * nb; ++i) fb *= i;
    const double pre = pow(2.0 * za, na + 0.5) * pow(2.0 * zb, nb + 0.5) / sqrt(fa * fb);
    const T prer = pre * pow(r / 2.0, na + nb + 1.0);
    return prer * s;
}

// Pair core repulsion (Mndo.cpp:223-240, 319-327; Am1.cpp:145-170) from the
// rotated (ss|ss) integral gss.
template <class T>
__device__ __forceinline__ T pair_core_repulsion(const ElemDev& A, const ElemDev& B, T r, T gss, int theory) {
    const double ang = 1.0 / 0.5291772;
    T f;
    if (A.Z == 1 && (B.Z == 7 || B.Z == 8))
        f = 1.0 + (r / ang) * exp(-B.alpha * r) + exp(-A.alpha * r);
    else if (B.Z == 1 && (A.Z == 7 || A.Z == 8))
        f = 1.0 + (r / ang) * exp(-A.alpha * r) + exp(-B.alpha * r);
    else
        f = 1.0 + exp(-A.alpha * r) + exp(-B.alpha * r);
    T e = A.core * B.core * gss * f;
    if (theory != 0) {
        T t = 0.0;
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
                                                     int* __restrict__ plist, FockPairDev* __restrict__ fpairs) {
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= natoms) return;
    const AtomDev at = atoms[g];
    const int nat = jobs[at.job].natoms, a = at.local, atom0 = g - a;
    const AtomRowDev rw = rows[g];
    long long slot = rw.slot0, io = rw.ioff0, c0 = rw.cpos[0], c1 = rw.cpos[1], c2 = rw.cpos[2];
    const bool ha = elems[at.elem].nao == 1;
    const int aoa = at.ao | (elems[at.elem].nao << 24);
    for (int b = a + 1; b < nat; ++b, ++slot) {
        const AtomDev atb = atoms[atom0 + b];
        const int nb = elems[atb.elem].nao;
        const bool hb = nb == 1;
        pairs[slot] = PairDev{g, atom0 + b, at.job, (int)slot, io};
        const long long pos = (ha && hb) ? c0++ : ((ha || hb) ? c1++ : c2++);
        plist[pos] = (int)slot;
        fpairs[pos] = FockPairDev{at.job, aoa, atb.ao | (nb << 24), (int)slot, io};
        io += (ha ? 1 : 10) * (hb ? 1 : 10);
    }
}

void launch_build_pairs(hipStream_t st, const AtomDev* atoms, const AtomRowDev* rows, int natoms, const JobDev* jobs,
                        const ElemDev* elems, PairDev* pairs, int* plist, FockPairDev* fpairs) {
    if (natoms > 0)
        hipLaunchKernelGGL(k_build_pairs, dim3((natoms + 255) / 256), dim3(256), 0, st, atoms, rows, natoms, jobs,
                           elems, pairs, plist, fpairs);
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

// ---------------------------------------------------------------------------
// Analytic SCF forces.  At convergence the SCF energy is stationary in the
// density, and the NDDO secular equation has no overlap matrix, so the force
// is the gradient of the energy at fixed P (no density response and no
// energy-weighted-density term).  The geometry enters only the two-centre
// terms of each pair A < B, through R = R_B - R_A:
//   E_AB = 2 sum_{mu on A, nu on B} P_mu,nu beta_bar_mu,nu S_mu,nu
//        + sum_{i on A} w_i P_A(i) (-Z_B) (i|s_B s_B) + sum_{j on B} w_j P_B(j) (-Z_A) (s_A s_A|j)
//        + sum_{i on A, j on B} [w_i w_j P_A(i) P_B(j) - 1/2 K_ij] (i|j)
//        + E_core(A, B),
// with i, j packed AO pairs, w = 1 for mu = nu and 2 otherwise, and K_ij the
// sum over the orderings (a, b) of i and (c, d) of j of P_ac P_bd (exchange).
// k_pair_grad evaluates E_AB with the integral, overlap and core-repulsion
// functions above in D3 dual numbers seeded on R, so dE_AB/dR is exact to
// rounding; k_atom_forces sums F_A = -dE/dR_A over the pairs in a fixed order.
// Reference: Mndo::CalcForce (src/mndo/Mndo.cpp:2711-2986), which builds the
// same gradient from derivative integrals (GetNddoRepulsionIntegral1stDerivative,
// :5245-5661) and the overlap and rotation derivatives of Cndo2.

// Rotation R[space][local] for the gradient: local z along A->B, local x the
// normalized part of a fixed space axis e orthogonal to z (e the axis of the
// smallest |z_c|, picked from the values and then held fixed, so the frame is
// a smooth function of R around the current geometry), local y = z x x.  The
// local-frame integrals and overlaps are symmetric under rotations about z (the
// (px py|px py) identity above), so every space-frame quantity equals that of
// the setup kernels' Euler-angle frame, whose azimuth is singular for R along
// the space z axis.
__device__ void rotation_frame(const D3 u[3], D3 R[4][4]) {
    const int sp[3] = {ORB_PX, ORB_PY, ORB_PZ};   // space x, y, z -> AO slot
    int c0 = 0;
    for (int c = 1; c < 3; ++c)
        if (fabs(u[c].v) < fabs(u[c0].v)) c0 = c;
    D3 x[3];
    for (int c = 0; c < 3; ++c) x[c] = (c == c0 ? 1.0 : 0.0) - u[c0] * u[c];   // e - (e.u) u
    const D3 nx = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    for (int c = 0; c < 3; ++c) x[c] = x[c] / nx;
    const D3 y[3] = {u[1] * x[2] - u[2] * x[1], u[2] * x[0] - u[0] * x[2], u[0] * x[1] - u[1] * x[0]};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) R[i][j] = D3(0.0);
    R[ORB_S][ORB_S] = D3(1.0);
    for (int c = 0; c < 3; ++c) {
        R[sp[c]][ORB_PZ] = u[c];
        R[sp[c]][ORB_PX] = x[c];
        R[sp[c]][ORB_PY] = y[c];
    }
}

__device__ inline void wave_sum4(double& a, double& b, double& c, double& d) {
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
        c += __shfl_xor(c, off);
        d += __shfl_xor(d, off);
    }
}

// One wavefront per pair slot of the batch; pairs of jobs without want_f are
// skipped.  grad[4 slot + 0..2] = dE_AB/dR, grad[4 slot + 3] = E_AB.
__global__ void __launch_bounds__(256) k_pair_grad(const PairDev* __restrict__ pairs, long long nslots,
                                                   const AtomDev* __restrict__ atoms,
                                                   const ElemDev* __restrict__ elems, const JobDev* __restrict__ jobs,
                                                   const JobState* __restrict__ state,
                                                   const double* __restrict__ arena, double* __restrict__ grad) {
    // D3 has constructors, so its LDS arrays are declared as plain storage
    __shared__ double sLraw[4][400], sTraw[4][400], sXraw[4][400], sRraw[4][64], sSlraw[4][64];
    __shared__ double sG[4][100], sP[4][3][16];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long long slot = (long long)blockIdx.x * 4 + wave;
    const PairDev pr = pairs[slot < nslots ? slot : 0];
    const JobDev jb = jobs[pr.job];
    const bool live = slot < nslots && jb.want_f;
    const AtomDev aA = atoms[pr.a], aB = atoms[pr.b];
    const ElemDev& A = elems[aA.elem];
    const ElemDev& B = elems[aB.elem];
    const int na = A.nao, nb = B.nao, nap = na == 4 ? 10 : 1, nbp = nb == 4 ? 10 : 1;
    const int n = jb.nao;
    // the converged density (the one the energy was evaluated with)
    const double* P = arena + jb.arena + (long long)(AR_P0 + ((state[pr.job].iter + 1) & 1)) * n * n;
    double* PAA = sP[wave][0];
    double* PBB = sP[wave][1];
    double* PAB = sP[wave][2];
    if (lane < 16) {
        const int i = lane >> 2, j = lane & 3;
        PAA[lane] = live && i < na && j < na ? P[(long long)(aA.ao + i) * n + aA.ao + j] : 0.0;
        PBB[lane] = live && i < nb && j < nb ? P[(long long)(aB.ao + i) * n + aB.ao + j] : 0.0;
        PAB[lane] = live && i < na && j < nb ? P[(long long)(aA.ao + i) * n + aB.ao + j] : 0.0;
    }
    const D3 dx(aB.x - aA.x, 1.0, 0.0, 0.0), dy(aB.y - aA.y, 0.0, 1.0, 0.0), dz(aB.z - aA.z, 0.0, 0.0, 1.0);
    const D3 r = sqrt(dx * dx + dy * dy + dz * dz);
    D3* L = reinterpret_cast<D3*>(sLraw[wave]);
    D3* T = reinterpret_cast<D3*>(sTraw[wave]);
    D3* X = reinterpret_cast<D3*>(sXraw[wave]);
    D3* Rw = reinterpret_cast<D3*>(sRraw[wave]);
    D3* Slw = reinterpret_cast<D3*>(sSlraw[wave]);
    double* G = sG[wave];
    if (live)
        for (int e = lane; e < nap * nbp; e += 64) L[e] = local_integral(A, B, e / nbp, e % nbp, r);
    if (lane == 0) {
        const D3 u[3] = {dx / r, dy / r, dz / r};
        rotation_frame(u, reinterpret_cast<D3(*)[4]>(Rw));
    }
    __syncthreads();
    // weights of the rotated integrals (i|j): Coulomb, exchange, electron-core
    for (int e = lane; e < nap * nbp; e += 64) {
        const int i = e / nbp, j = e % nbp;
        const int mu = nap == 1 ? 0 : c_pk[i][0], nu = nap == 1 ? 0 : c_pk[i][1];
        const int la = nbp == 1 ? 0 : c_pk[j][0], si = nbp == 1 ? 0 : c_pk[j][1];
        const double wa = mu == nu ? 1.0 : 2.0, wb = la == si ? 1.0 : 2.0;
        const double pa = PAA[mu * 4 + nu], pb = PBB[la * 4 + si];
        double k = PAB[mu * 4 + la] * PAB[nu * 4 + si];
        if (la != si) k += PAB[mu * 4 + si] * PAB[nu * 4 + la];
        if (mu != nu) {
            k += PAB[nu * 4 + la] * PAB[mu * 4 + si];
            if (la != si) k += PAB[nu * 4 + si] * PAB[mu * 4 + la];
        }
        double g = wa * wb * pa * pb - 0.5 * k;
        if (j == 0) g -= B.core * wa * pa;
        if (i == 0) g -= A.core * wb * pb;
        G[e] = g;
    }
    for (int e = lane; e < 100; e += 64) {
        const int i = e / 10, k = e % 10;
        const int mu = c_pk[i][0], nu = c_pk[i][1], al = c_pk[k][0], be = c_pk[k][1];
        const D3* R = Rw;
        D3 t = R[mu * 4 + al] * R[nu * 4 + be];
        if (al != be) t += R[mu * 4 + be] * R[nu * 4 + al];
        T[e] = t;
    }
    // overlap in the diatomic frame (m-matched pairs), as k_pair_setup
    if (lane < 16) {
        const int i = lane >> 2, j = lane & 3;
        D3 v = 0.0;
        if (live && i < na && j < nb) {
            const int la = i == ORB_S ? 0 : 1, lb = j == ORB_S ? 0 : 1;
            const int ma = i == ORB_PX ? 1 : (i == ORB_PY ? -1 : 0);
            const int mb = j == ORB_PX ? 1 : (j == ORB_PY ? -1 : 0);
            if (ma == mb) v = overlap_local(A, B, la, lb, ma != 0 ? 1 : 0, r);
        }
        Slw[lane] = v;
    }
    __syncthreads();
    // X = T_A L
    for (int e = lane; e < nap * nbp; e += 64) {
        const int i = e / nbp, j = e % nbp;
        D3 s = 0.0;
        if (nap == 1) s = L[e];
        else for (int k = 0; k < 10; ++k) s += T[i * 10 + k] * L[k * nbp + j];
        X[e] = s;
    }
    __syncthreads();
    D3 acc = 0.0;
    // (i|j) = (X T_B^T)_ij against its weight
    for (int e = lane; e < nap * nbp; e += 64) {
        const int i = e / nbp, j = e % nbp;
        D3 s = 0.0;
        if (nbp == 1) s = X[e];
        else for (int l = 0; l < 10; ++l) s += X[i * nbp + l] * T[j * 10 + l];
        acc += G[e] * s;
    }
    // resonance: 2 P_AB beta_bar S with S = R S_loc R^T
    if (lane < 16) {
        const int i = lane >> 2, j = lane & 3;
        if (i < na && j < nb) {
            const D3* R = Rw;
            const D3* Sl = Slw;
            D3 sv = 0.0;
            for (int k = 0; k < 4; ++k)
                for (int l = 0; l < 4; ++l) sv += R[i * 4 + k] * Sl[k * 4 + l] * R[j * 4 + l];
            const double bb = 0.5 * (A.beta[i == ORB_S ? 0 : 1] + B.beta[j == ORB_S ? 0 : 1]);
            acc += (2.0 * bb * PAB[i * 4 + j]) * sv;
        }
    }
    if (lane == 63 && live) acc += pair_core_repulsion(A, B, r, L[0], jb.theory);
    double g0 = acc.d[0], g1 = acc.d[1], g2 = acc.d[2], ev = acc.v;
    wave_sum4(g0, g1, g2, ev);
    if (live && lane == 0) {
        double* o = grad + 4 * slot;
        o[0] = g0;
        o[1] = g1;
        o[2] = g2;
        o[3] = ev;
    }
}

// F_A = sum_B -dE/dR_A: +dE_AB/dR for the pairs (A, B), -dE_BA/dR for (B, A),
// B in order.  forces[3 g + c] for every atom g of a job with want_f.
__global__ void __launch_bounds__(256) k_atom_forces(const AtomDev* __restrict__ atoms, int natoms,
                                                     const JobDev* __restrict__ jobs, const double* __restrict__ grad,
                                                     double* __restrict__ forces) {
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= natoms) return;
    const AtomDev at = atoms[g];
    const JobDev jb = jobs[at.job];
    if (!jb.want_f) return;
    const int A = at.local, N = jb.natoms;
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    for (int b = 0; b < N; ++b) {
        if (b == A) continue;
        const double* gp = grad + 4 * (jb.pair0 + (A < b ? tri_slot(A, b, N) : tri_slot(b, A, N)));
        const double sg = A < b ? 1.0 : -1.0;
        f0 += sg * gp[0];
        f1 += sg * gp[1];
        f2 += sg * gp[2];
    }
    forces[3LL * g + 0] = f0;
    forces[3LL * g + 1] = f1;
    forces[3LL * g + 2] = f2;
}

void launch_forces(hipStream_t st, const PairDev* pairs, long long nslots, const AtomDev* atoms, int natoms,
                   const ElemDev* elems, const JobDev* jobs, const JobState* state, const double* arena, double* grad,
                   double* forces) {
    if (nslots > 0)
        hipLaunchKernelGGL(k_pair_grad, dim3((unsigned)((nslots + 3) / 4)), dim3(256), 0, st, pairs, nslots, atoms,
                           elems, jobs, state, arena, grad);
    if (natoms > 0)
        hipLaunchKernelGGL(k_atom_forces, dim3((natoms + 255) / 256), dim3(256), 0, st, atoms, natoms, jobs, grad,
                           forces);
}

int upload_overlap_table(const double* table /* 32*25 */) {
    return hipMemcpyToSymbol(HIP_SYMBOL(c_ovl), table, sizeof(double) * 32 * 25) == hipSuccess ? 0 : -1;
}

}  // namespace nddo
