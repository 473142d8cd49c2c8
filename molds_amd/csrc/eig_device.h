// Device helpers shared by the one-workgroup eigensolver (kernels_eig.hip)
// and the device-wide one for large molecules (kernels_xlh.hip): job matrix
// and workspace addressing, the secular-equation root finder and the total
// order of the divide-and-conquer merges.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "nddo_device.h"

namespace nddo {

typedef double dbl4 __attribute__((ext_vector_type(4)));

__device__ inline double* amat(double* arena, const JobDev& jb, int kind) {
    return arena + jb.arena + (long long)kind * jb.nao * jb.nao;
}

constexpr double kEps = 1.1102230246251565e-16;        // 2^-53, unit roundoff
constexpr double kEps32 = 5.9604644775390625e-08;      // 2^-24, FP32 unit roundoff

// Per-job eigensolver workspace (EIG_WS_PER_AO doubles per AO plus
// EIG_WS_PER_JOB per job): d[n], e[n], tau[n], then the compact-WY T blocks
// (16 x 16 per block of 16 reflectors, at most 16 (n + 14) doubles).
__device__ inline double* eig_ws(double* ws, const JobDev& jb, int job) {
    return ws + (long long)EIG_WS_PER_AO * jb.vec0 + (long long)EIG_WS_PER_JOB * job;
}

// Block-wide sum; every thread gets the result.
__device__ inline double block_sum(double v, double* red, int tid, int nw) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    __syncthreads();
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += red[w];
    return s;
}

// 1/x from the hardware estimate (v_rcp_f64) refined by two Newton-Raphson
// steps r <- r + r (1 - x r); the arguments here are differences of
// eigenvalues, far from the denormal and overflow ranges.
__device__ inline double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// Sums of the secular function at x = dl[o] + tau: psi over poles t <= j,
// phi over t > j, their derivatives and sum |terms| (all scaled by rho).
struct SecSums {
    double psi, dpsi, phi, dphi, asum;
};

// sum_{t0 <= t < t1} z_t^2 / ((dl_t - dO) - tau) and its derivative sum, with
// sum |terms|, as four interleaved partial sums so that four terms' reciprocal
// chains are in flight at once.
__device__ inline void dc_sums_range(const double* dl, const double* zl, int t0, int t1, double dO, double tau,
                                     double& sum, double& dsum, double& asum, int sub = 0, int g = 1) {
    // poles t0 + sub, t0 + sub + g, ... (a lane's share when g lanes split the sum)
    double s[4] = {0.0, 0.0, 0.0, 0.0}, ds[4] = {0.0, 0.0, 0.0, 0.0}, as[4] = {0.0, 0.0, 0.0, 0.0};
    int t = t0 + sub;
    for (; t + 3 * g < t1; t += 4 * g) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double z = zl[t + u * g];
            const double inv = rcp_nr((dl[t + u * g] - dO) - tau);
            const double term = z * z * inv;
            s[u] += term;
            ds[u] += term * inv;
            as[u] += fabs(term);
        }
    }
    for (; t < t1; t += g) {
        const double z = zl[t];
        const double inv = rcp_nr((dl[t] - dO) - tau);
        const double term = z * z * inv;
        s[0] += term;
        ds[0] += term * inv;
        as[0] += fabs(term);
    }
    sum = (s[0] + s[1]) + (s[2] + s[3]);
    dsum = (ds[0] + ds[1]) + (ds[2] + ds[3]);
    asum = (as[0] + as[1]) + (as[2] + as[3]);
}

// Sum over the g (a power of two, at most 64) aligned lanes of a group by an
// xor butterfly: every lane of the group gets bitwise the same sum.
__device__ inline double group_allsum(double v, int g) {
    for (int off = g >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ inline int group_allsum(int v, int g) {
    for (int off = g >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}
__device__ inline double group_allprod(double v, int g) {
    for (int off = g >> 1; off > 0; off >>= 1) v *= __shfl_xor(v, off);
    return v;
}

// With g > 1 the g lanes of a group split the poles (lane sub takes every
// g-th) and all get the same sums.
__device__ inline SecSums dc_sums(const double* dl, const double* zl, int kk, double rho, int j, double dO,
                                  double tau, int sub = 0, int g = 1) {
    double psi, dpsi, phi, dphi, a1, a2;
    dc_sums_range(dl, zl, 0, j + 1, dO, tau, psi, dpsi, a1, sub, g);
    dc_sums_range(dl, zl, j + 1, kk, dO, tau, phi, dphi, a2, sub, g);
    if (g > 1) {
        psi = group_allsum(psi, g);
        dpsi = group_allsum(dpsi, g);
        phi = group_allsum(phi, g);
        dphi = group_allsum(dphi, g);
        a1 = group_allsum(a1, g);
        a2 = group_allsum(a2, g);
    }
    return SecSums{rho * psi, rho * dpsi, rho * phi, rho * dphi, a1 + a2};
}

// Root j (0-based) of the secular equation 1 + rho sum_t z_t^2 / (dl_t - x) = 0
// for sorted dl[0..kk), rho > 0: x = dl[o] + tau with o the nearer pole, so
// that dl_t - x = (dl_t - dl_o) - tau keeps its relative accuracy.  The first
// evaluation, at the middle of (dl_j, dl_j+1), both picks the pole and starts
// the iteration.
// ue: the unit roundoff the root is resolved to (FP64, or FP32 for a job in
// its FP32 iterations, whose tridiagonal is only that accurate).
// sub, g: as dc_sums (g lanes solve the root together and all return it).
__device__ inline void dc_secular(const double* dl, const double* zl, int kk, double rho, int j, int* orig,
                           double* taup, double ue, int sub = 0, int g = 1) {
    int o = kk - 1;
    double lo = 0.0, hi = 0.0, tau = 0.0, dO = dl[kk - 1];
    // the first evaluation point: the middle of (dl_j, dl_j+1) seen from dl_j,
    // or for the last root half of rho |z|^2 above the last pole
    const bool inner = j < kk - 1;
    const double mid = inner ? 0.5 * (dl[j + 1] - dl[j]) : 0.0;
    double evO = inner ? dl[j] : dl[kk - 1], evT = mid;
    if (!inner) {
        double zz = 0.0;
        for (int t = sub; t < kk; t += g) zz += zl[t] * zl[t];
        if (g > 1) zz = group_allsum(zz, g);
        hi = rho * zz;
        tau = 0.5 * hi;
        evT = tau;
    }
    for (int it = 0; it < 100; ++it) {
        const SecSums S = dc_sums(dl, zl, kk, rho, j, evO, evT, sub, g);   // the one evaluation site
        if (it == 0 && inner) {   // the first value picks the nearer pole as origin
            if (1.0 + S.psi + S.phi >= 0.0) { o = j; lo = 0.0; hi = mid; tau = mid; }
            else { o = j + 1; lo = -mid; hi = 0.0; tau = -mid; }
            dO = dl[o];
        }
        const double f = 1.0 + S.psi + S.phi;
        if (fabs(f) <= 8.0 * ue * (1.0 + rho * S.asum)) break;
        if (f < 0.0) lo = tau;
        else hi = tau;
        double eta = 0.0;
        bool have = false;
        if (inner) {
            // psi ~ a1 + b1/(p_j - x), phi ~ a2 + b2/(p_j+1 - x), matched in value and slope at tau
            const double u1 = (dl[j] - dO) - tau, u2 = (dl[j + 1] - dO) - tau;
            const double b1 = u1 * u1 * S.dpsi, b2 = u2 * u2 * S.dphi;
            const double c = 1.0 + (S.psi - b1 / u1) + (S.phi - b2 / u2);
            const double qa = c, qb = -(c * (u1 + u2) + b1 + b2), qc = c * u1 * u2 + b1 * u2 + b2 * u1;
            if (qa != 0.0) {
                const double disc = qb * qb - 4.0 * qa * qc;
                if (disc >= 0.0) {
                    const double q = -0.5 * (qb + copysign(sqrt(disc), qb));
                    const double e2 = q / qa;
                    const double e1 = q != 0.0 ? qc / q : e2;
                    eta = fabs(e1) < fabs(e2) ? e1 : e2;
                    have = true;
                }
            } else if (qb != 0.0) {
                eta = -qc / qb;
                have = true;
            }
        } else {
            const double u1 = -tau;   // the last pole is the origin
            const double b1 = u1 * u1 * S.dpsi;
            const double c = f - b1 / u1;
            if (c != 0.0) { eta = u1 + b1 / c; have = true; }
        }
        if (have && fabs(eta) <= 2.0 * ue * fabs(tau)) { tau += eta; break; }
        const double nt = (have && lo < tau + eta && tau + eta < hi) ? tau + eta : 0.5 * (lo + hi);
        if (fabs(nt - tau) <= 2.0 * ue * fabs(nt) || hi - lo <= 2.0 * ue * fmax(fabs(lo), fabs(hi))) {
            tau = nt;
            break;
        }
        tau = nt;
        evO = dO;
        evT = tau;
    }
    *orig = o;
    *taup = tau;
}

// Strict total order on (value, index) with NaN placed after +inf, so that the
// positions computed from it are always a permutation of the block.
__device__ inline bool dc_before(double u, int iu, double v, int iv) {
    const double ku = isnan(u) ? INFINITY : u, kv = isnan(v) ? INFINITY : v;
    return ku < kv || (ku == kv && iu < iv);
}

}  // namespace nddo
