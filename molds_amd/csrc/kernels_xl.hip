// Device-wide kernels for single large molecules (more than 512 AOs), where one
// workgroup per molecule would leave the chip idle:
//
//   k_bjac         FP64 symmetric eigensolver (replaces LAPACKE_dsyevd,
//                  src/wrappers/Lapack.cpp:102-161): two-sided block Jacobi
//                  with the round-robin ordering of 32-row blocks, one
//                  persistent launch per solve.  Each round diagonalizes the
//                  64 x 64 pair blocks exactly (scalar cyclic Jacobi in LDS,
//                  Golub & Van Loan, Matrix Computations, Alg. 8.5.3) and
//                  applies the rotations to the rest of the matrix and to the
//                  eigenvector rows as 64 x 64 x 64 products on FP64 MFMA
//                  (v_mfma_f64_16x16x4f64).  In the SCF loop F arrives already
//                  rotated into the previous eigenvectors (k_warm_gemm), so
//                  after the first iterations two or three sweeps suffice.
//                  Convergence: a sweep that rotates nothing, with the same
//                  threshold as the one-workgroup Jacobi (|a_pq| <= 1e-15 ||F||_F).
//   k_xl_dots      DIIS ring update and error dot products (Cndo2.cpp:920-1028)
//   k_xl_apply     DIIS extrapolation and damping (Cndo2.cpp:1016-1061)
//   k_xl_energy    E_el partial sums (Cndo2.cpp:1349-1381)
//                  chunked over the whole grid; k_step and k_energy reduce the
//                  chunk partials in a fixed order, so results do not depend on
//                  scheduling.
//
// Grid-wide synchronisation of k_bjac: a counter/generation barrier in device
// memory with agent-scope release/acquire, all workgroups resident (the grid
// is at most one workgroup per CU); a bounded wait that marks the solve failed
// instead of waiting forever.
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>

#include "launch.h"

namespace nddo {

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

constexpr int BJ_B = 32;              // rows per block
constexpr int BJ_P = 64;              // rows per pair block
constexpr int BJ_LD = 65;             // LDS leading dimension of a 64 x 64 tile
constexpr int BJ_MAXSWEEP = 50;
constexpr int BJ_INNER_MAXSWEEP = 40;
constexpr int BJ_MAXJOBS = 32;        // jobs per launch
constexpr int BJ_MAXGRID = 1024;
constexpr int BJ_T = 256;             // threads per workgroup

__device__ inline double* matx(double* arena, const JobDev& jb, int kind) {
    return arena + jb.arena + (long long)kind * jb.nao * jb.nao;
}

// Per-job geometry of the block ordering.
struct BjGeom {
    int n, m, mp, np, rounds, nct, ntri, nvt;
};

__device__ inline BjGeom bj_geom(int n) {
    BjGeom g;
    g.n = n;
    g.m = (n + BJ_B - 1) / BJ_B;
    g.mp = g.m + (g.m & 1);
    g.np = g.mp / 2;
    g.rounds = g.mp - 1;
    g.nct = (n + BJ_P - 1) / BJ_P;
    g.ntri = g.np * (g.np - 1) / 2;
    g.nvt = g.np * g.nct;
    return g;
}

// Workspace of one job (doubles): GT[np][64][64], rotated flag per pair,
// diagonal, Frobenius partials per workgroup, rotated flag per sweep.
__host__ __device__ inline long long bj_ws_gt(int) { return 0; }
__host__ __device__ inline long long bj_ws_flag(int np) { return (long long)np * BJ_P * BJ_P; }
__host__ __device__ inline long long bj_ws_diag(int np) { return bj_ws_flag(np) + np; }
__host__ __device__ inline long long bj_ws_part(int np, int n) { return bj_ws_diag(np) + n; }
__host__ __device__ inline long long bj_ws_sweep(int np, int n) { return bj_ws_part(np, n) + BJ_MAXGRID; }
__host__ __device__ inline long long bj_ws_size(int np, int n) { return bj_ws_sweep(np, n) + BJ_MAXSWEEP + 14; }

// Blocks of pair k in round r (round-robin / circle method over mp blocks).
__device__ inline void bj_pair(int mp, int r, int k, int* p, int* q) {
    int a, b;
    if (k == 0) { a = r; b = mp - 1; }
    else { a = (r + k) % (mp - 1); b = (r - k + mp - 1) % (mp - 1); }
    *p = min(a, b);
    *q = max(a, b);
}

// Matrix row of local index a (0..63) of the pair block (p, q); >= n is padding.
__device__ inline int bj_row(int p, int q, int a) { return a < BJ_B ? p * BJ_B + a : q * BJ_B + (a - BJ_B); }

// Grid barrier.  bar[0] counts arrivals, bar[1] is the generation.  Returns
// false when the wait ran out (the caller then abandons the solve).
__device__ bool grid_sync(unsigned* bar, unsigned nblk, unsigned* gen, int* sh_ok) {
    __threadfence();   // every wave publishes its own stores (agent scope)
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        const unsigned target = *gen + 1;
        const unsigned prev = __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (prev == nblk - 1) {
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&bar[1], target, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const unsigned long long t0 = wall_clock64();
            while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != target) {
                __builtin_amdgcn_s_sleep(2);
                // 100 MHz constant clock: give up after 20 s
                if (wall_clock64() - t0 > 2000000000ull) { ok = 0; break; }
            }
        }
        *gen = target;
        *sh_ok = ok;
    }
    __syncthreads();
    __threadfence();   // and sees everyone else's
    return *sh_ok != 0;
}

// C (64 x 64, registers, D layout) = A * B with A rows in LDS (ld BJ_LD) and
// B[k][c] = Bm[k * BJ_LD + c] (BT = false) or Bm[c * BJ_LD + k] (BT = true).
// Wave w computes rows 16w .. 16w + 15; acc[cb][v] is C[16w + 4v + lane/16][16cb + lane%16].
template <bool BT>
__device__ inline void bj_gemm(const double* Am, const double* Bm, dbl4 acc[4]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) acc[cb] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int k0 = 0; k0 < BJ_P; k0 += 4) {
        const double a = Am[(16 * w + li) * BJ_LD + k0 + lk];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
            const double b = BT ? Bm[(16 * cb + li) * BJ_LD + k0 + lk] : Bm[(k0 + lk) * BJ_LD + 16 * cb + li];
            acc[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[cb], 0, 0, 0);
        }
    }
}

__device__ inline void bj_store(double* Cm, const dbl4 acc[4]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int v = 0; v < 4; ++v) Cm[(16 * w + 4 * v + lk) * BJ_LD + 16 * cb + li] = acc[cb][v];
}

struct BjShared {
    double sc[32], ss[32], st[32], su[32];   // c, s, t = s / c, u = s / (1 + c)
    int sp[32], sq[32], srot[32];
    double red[BJ_T];
    double thr[BJ_MAXJOBS];
    int nsweep[BJ_MAXJOBS];
    int ok;
};

// Phase 1: exact eigen-decomposition of the 64 x 64 pair block k (round r):
// S <- GT S GT^T diagonal; S written back to the matrix, GT to the workspace.
__device__ void bj_inner(double* S, double* G, BjShared& sh, double* W, int n, int p, int q, double thr, double* gt,
                         double* flag, double* sweepflag) {
    const int tid = threadIdx.x;
    double offmax = 0.0;
    for (int e = tid; e < BJ_P * BJ_P; e += BJ_T) {
        const int a = e >> 6, b = e & 63;
        const int ra = bj_row(p, q, a), rb = bj_row(p, q, b);
        const double v = (ra < n && rb < n) ? W[(long long)ra * n + rb] : 0.0;
        S[a * BJ_LD + b] = v;
        G[a * BJ_LD + b] = a == b ? 1.0 : 0.0;
        if (a != b) offmax = fmax(offmax, fabs(v));
    }
    // the whole block is already diagonal to the threshold: nothing to rotate
    // (a NaN entry also lands here: fmax ignores it and the solve fails later
    // on the non-finite norm)
    if (!__syncthreads_or(offmax > thr)) {
        if (tid == 0) *flag = 0.0;
        return;
    }
    for (int isw = 0; isw < BJ_INNER_MAXSWEEP; ++isw) {
        int any = 0;
        for (int r = 0; r < 63; ++r) {
            int rot = 0;
            if (tid < 32) {
                int pp, qq;
                if (tid == 0) { pp = r; qq = 63; }
                else { pp = (r + tid) % 63; qq = (r - tid + 63) % 63; }
                if (pp > qq) { const int t = pp; pp = qq; qq = t; }
                double c = 1.0, s = 0.0, t = 0.0, u = 0.0;
                const double apq = S[pp * BJ_LD + qq];
                if (fabs(apq) > thr) {
                    const double app = S[pp * BJ_LD + pp], aqq = S[qq * BJ_LD + qq];
                    const double tau = (aqq - app) / (2.0 * apq);
                    t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                    c = 1.0 / sqrt(1.0 + t * t);
                    s = t * c;
                    u = s / (1.0 + c);
                    rot = 1;
                }
                sh.sp[tid] = pp; sh.sq[tid] = qq; sh.sc[tid] = c; sh.ss[tid] = s; sh.st[tid] = t; sh.su[tid] = u;
                sh.srot[tid] = rot;
            }
            if (!__syncthreads_or(rot)) continue;
            any = 1;
            // 2 x 2 blocks (rows {p1,q1}, columns {p2,q2}) of the upper block triangle
            for (int b = tid; b < 528; b += BJ_T) {
                int hi = (int)((sqrtf(8.0f * (float)b + 1.0f) - 1.0f) * 0.5f);
                while (hi * (hi + 1) / 2 > b) --hi;
                while ((hi + 1) * (hi + 2) / 2 <= b) ++hi;
                const int ki = b - hi * (hi + 1) / 2, kj = hi;
                const int ri = sh.srot[ki], rj = sh.srot[kj];
                if (!ri && !rj) continue;
                const int p1 = sh.sp[ki], q1 = sh.sq[ki], p2 = sh.sp[kj], q2 = sh.sq[kj];
                if (ki == kj) {
                    const double t = sh.st[ki], apq = S[p1 * BJ_LD + q1];
                    S[p1 * BJ_LD + p1] -= t * apq;
                    S[q1 * BJ_LD + q1] += t * apq;
                    S[p1 * BJ_LD + q1] = 0.0;
                    S[q1 * BJ_LD + p1] = 0.0;
                    continue;
                }
                // rotations in Rutishauser's form x - s (y + u x), y + s (x - u y), which
                // keeps them orthogonal to rounding also when 1 + t^2 rounds to 1
                const double s1 = sh.ss[ki], u1 = sh.su[ki], s2 = sh.ss[kj], u2 = sh.su[kj];
                const double a = S[p1 * BJ_LD + p2], bq = S[p1 * BJ_LD + q2];
                const double cq = S[q1 * BJ_LD + p2], d = S[q1 * BJ_LD + q2];
                const double a1 = a - s1 * (cq + u1 * a), b1 = bq - s1 * (d + u1 * bq);
                const double c1r = cq + s1 * (a - u1 * cq), d1 = d + s1 * (bq - u1 * d);
                const double a2 = a1 - s2 * (b1 + u2 * a1), b2 = b1 + s2 * (a1 - u2 * b1);
                const double c2r = c1r - s2 * (d1 + u2 * c1r), d2 = d1 + s2 * (c1r - u2 * d1);
                S[p1 * BJ_LD + p2] = a2; S[p2 * BJ_LD + p1] = a2;
                S[p1 * BJ_LD + q2] = b2; S[q2 * BJ_LD + p1] = b2;
                S[q1 * BJ_LD + p2] = c2r; S[p2 * BJ_LD + q1] = c2r;
                S[q1 * BJ_LD + q2] = d2; S[q2 * BJ_LD + q1] = d2;
            }
            // rotation rows of GT: GT <- J^T GT
            for (int e = tid; e < 32 * BJ_P; e += BJ_T) {
                const int k = e >> 6, c = e & 63;
                if (!sh.srot[k]) continue;
                const int pp = sh.sp[k], qq = sh.sq[k];
                const double sn = sh.ss[k], u = sh.su[k];
                const double gp = G[pp * BJ_LD + c], gq = G[qq * BJ_LD + c];
                G[pp * BJ_LD + c] = gp - sn * (gq + u * gp);
                G[qq * BJ_LD + c] = gq + sn * (gp - u * gq);
            }
            __syncthreads();
        }
        if (!any) break;
    }
    for (int e = tid; e < BJ_P * BJ_P; e += BJ_T) {
        const int a = e >> 6, b = e & 63;
        const int ra = bj_row(p, q, a), rb = bj_row(p, q, b);
        if (ra < n && rb < n) W[(long long)ra * n + rb] = S[a * BJ_LD + b];
        gt[e] = G[a * BJ_LD + b];
    }
    if (tid == 0) {
        *flag = 1.0;
        *sweepflag = 1.0;
    }
}

// Phase 2, off-diagonal tile (i, j), i < j:  A_ij <- GT_i A_ij GT_j^T, both triangles.
__device__ void bj_tile(double* As, double* Bs, double* W, int n, int pi, int qi, int pj, int qj, bool fi, bool fj,
                        const double* gti, const double* gtj) {
    const int tid = threadIdx.x;
    for (int e = tid; e < BJ_P * BJ_P; e += BJ_T) {
        const int a = e >> 6, b = e & 63;
        const int ra = bj_row(pi, qi, a), rb = bj_row(pj, qj, b);
        As[a * BJ_LD + b] = (ra < n && rb < n) ? W[(long long)ra * n + rb] : 0.0;
        if (fj) Bs[a * BJ_LD + b] = gtj[e];
    }
    __syncthreads();
    dbl4 acc[4];
    if (fj) {   // X = A GT_j^T
        bj_gemm<true>(As, Bs, acc);
        __syncthreads();
        bj_store(As, acc);
    }
    if (fi) {   // T = GT_i X
        for (int e = tid; e < BJ_P * BJ_P; e += BJ_T) Bs[(e >> 6) * BJ_LD + (e & 63)] = gti[e];
        __syncthreads();
        bj_gemm<false>(Bs, As, acc);
        __syncthreads();
        bj_store(As, acc);
    }
    __syncthreads();
    for (int e = tid; e < BJ_P * BJ_P; e += BJ_T) {
        const int a = e >> 6, b = e & 63;
        const int ra = bj_row(pi, qi, a), rb = bj_row(pj, qj, b);
        if (ra < n && rb < n) W[(long long)ra * n + rb] = As[a * BJ_LD + b];
        // transposed copy: row (pj,qj)[a], column (pi,qi)[b]
        const int ta = bj_row(pj, qj, a), tb = bj_row(pi, qi, b);
        if (ta < n && tb < n) W[(long long)ta * n + tb] = As[b * BJ_LD + a];
    }
    __syncthreads();
}

// Phase 2, eigenvector tile: rows (pi, qi) x columns c0 .. c0 + 63 of VT <- GT_i VT.
__device__ void bj_vt_tile(double* As, double* Bs, double* VT, int n, int pi, int qi, int c0, const double* gti) {
    const int tid = threadIdx.x;
    for (int e = tid; e < BJ_P * BJ_P; e += BJ_T) {
        const int a = e >> 6, b = e & 63;
        const int ra = bj_row(pi, qi, a), cb = c0 + b;
        As[a * BJ_LD + b] = (ra < n && cb < n) ? VT[(long long)ra * n + cb] : 0.0;
        Bs[a * BJ_LD + b] = gti[e];
    }
    __syncthreads();
    dbl4 acc[4];
    bj_gemm<false>(Bs, As, acc);
    const int lane = tid & 63, w = tid >> 6, li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int a = 16 * w + 4 * v + lk, c = c0 + 16 * cb + li;
            const int ra = bj_row(pi, qi, a);
            if (ra < n && c < n) VT[(long long)ra * n + c] = acc[cb][v];
        }
    __syncthreads();
}

__device__ inline double block_sum_bj(double v, double* red) {
    const int tid = threadIdx.x;
    red[tid] = v;
    __syncthreads();
    for (int s = BJ_T / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(BJ_T) k_bjac(const JobDev* __restrict__ jobs, JobState* __restrict__ state,
                                               const int* __restrict__ joblist, const long long* __restrict__ xloff,
                                               int nlist, double* __restrict__ arena, double* __restrict__ eps,
                                               int* __restrict__ perm, double* __restrict__ xlws,
                                               unsigned* __restrict__ bar) {
    extern __shared__ double dyn[];
    __shared__ BjShared sh;
    double* As = dyn;
    double* Bs = dyn + BJ_P * BJ_LD;
    const int G = gridDim.x, wg = blockIdx.x, tid = threadIdx.x;
    const long long gt0 = (long long)wg * BJ_T + tid, gstride = (long long)G * BJ_T;
    unsigned gen = 0;   // bar[] is reset to zero before every launch
    bool alive = true;   // false once a barrier wait ran out

    // jobs solved in this launch
    unsigned live = 0;
    for (int jl = 0; jl < nlist && jl < BJ_MAXJOBS; ++jl)
        if (state[joblist[jl]].phase == PH_ACTIVE) live |= 1u << jl;
    const unsigned started = live;

    // W <- F, VT <- I on a cold start, Frobenius norms
    for (int jl = 0; jl < nlist && jl < BJ_MAXJOBS; ++jl) {
        if (!(started >> jl & 1)) continue;
        const int job = joblist[jl];
        const JobDev jb = jobs[job];
        const int n = jb.nao;
        const long long nn = (long long)n * n;
        const BjGeom g = bj_geom(n);
        double* ws = xlws + xloff[jl];
        const double* F = matx(arena, jb, AR_F);
        double* W = matx(arena, jb, AR_W);
        double* VT = matx(arena, jb, AR_C);
        const bool warm = state[job].warm != 0;
        double fro = 0.0;
        for (long long e = gt0; e < nn; e += gstride) {
            const double v = F[e];
            W[e] = v;
            fro += v * v;
            if (!warm) {
                const long long i = e / n, j = e - i * n;
                VT[e] = i == j ? 1.0 : 0.0;
            }
        }
        fro = block_sum_bj(fro, sh.red);
        if (tid == 0) ws[bj_ws_part(g.np, n) + wg] = fro;
        if (wg == 0)
            for (int s = tid; s < BJ_MAXSWEEP; s += BJ_T) ws[bj_ws_sweep(g.np, n) + s] = 0.0;
    }
    alive = grid_sync(bar, G, &gen, &sh.ok);
    for (int jl = 0; jl < nlist && jl < BJ_MAXJOBS; ++jl) {
        if (!(started >> jl & 1)) continue;
        const int n = jobs[joblist[jl]].nao;
        const BjGeom g = bj_geom(n);
        const double* part = xlws + xloff[jl] + bj_ws_part(g.np, n);
        if (tid == 0) {
            double s = 0.0;
            for (int w = 0; w < G; ++w) s += part[w];
            sh.thr[jl] = 1e-15 * sqrt(s);
        }
    }
    __syncthreads();
    unsigned failed = 0;
    for (int jl = 0; jl < nlist && jl < BJ_MAXJOBS; ++jl)
        if ((started >> jl & 1) && !isfinite(sh.thr[jl])) failed |= 1u << jl;
    live &= ~failed;

    if (tid < BJ_MAXJOBS) sh.nsweep[tid] = -1;
    __syncthreads();
    for (int sweep = 0; sweep < BJ_MAXSWEEP && live && alive; ++sweep) {
        int R = 0;
        for (int jl = 0; jl < nlist; ++jl)
            if (live >> jl & 1) R = max(R, bj_geom(jobs[joblist[jl]].nao).rounds);
        for (int r = 0; r < R && alive; ++r) {
            // phase 1: pair blocks
            int c = 0;
            for (int jl = 0; jl < nlist; ++jl) {
                if (!(live >> jl & 1)) continue;
                const JobDev jb = jobs[joblist[jl]];
                const int n = jb.nao;
                const BjGeom g = bj_geom(n);
                if (r >= g.rounds) continue;
                double* ws = xlws + xloff[jl];
                for (int k = 0; k < g.np; ++k, ++c) {
                    if (c % G != wg) continue;
                    int p, q;
                    bj_pair(g.mp, r, k, &p, &q);
                    bj_inner(As, Bs, sh, matx(arena, jb, AR_W), n, p, q, sh.thr[jl], ws + (long long)k * BJ_P * BJ_P,
                             ws + bj_ws_flag(g.np) + k, ws + bj_ws_sweep(g.np, n) + sweep);
                    __syncthreads();
                }
            }
            alive = grid_sync(bar, G, &gen, &sh.ok);
            if (!alive) break;
            // phase 2: off-diagonal tiles and eigenvector tiles
            c = 0;
            for (int jl = 0; jl < nlist; ++jl) {
                if (!(live >> jl & 1)) continue;
                const JobDev jb = jobs[joblist[jl]];
                const int n = jb.nao;
                const BjGeom g = bj_geom(n);
                if (r >= g.rounds) continue;
                const double* ws = xlws + xloff[jl];
                const double* fl = ws + bj_ws_flag(g.np);
                double* W = matx(arena, jb, AR_W);
                double* VT = matx(arena, jb, AR_C);
                const int nt = g.ntri + g.nvt;
                int t = (wg - c % G + G) % G;
                for (; t < nt; t += G) {
                    if (t < g.ntri) {
                        int j = (int)((1.0 + sqrt(1.0 + 8.0 * (double)t)) * 0.5);
                        while (j * (j - 1) / 2 > t) --j;
                        while ((j + 1) * j / 2 <= t) ++j;
                        const int i = t - j * (j - 1) / 2;
                        const bool fi = fl[i] != 0.0, fj = fl[j] != 0.0;
                        if (!fi && !fj) continue;
                        int pi, qi, pj, qj;
                        bj_pair(g.mp, r, i, &pi, &qi);
                        bj_pair(g.mp, r, j, &pj, &qj);
                        bj_tile(As, Bs, W, n, pi, qi, pj, qj, fi, fj, ws + (long long)i * BJ_P * BJ_P,
                                ws + (long long)j * BJ_P * BJ_P);
                    } else {
                        const int u = t - g.ntri, i = u / g.nct, ct = u - i * g.nct;
                        if (fl[i] == 0.0) continue;
                        int pi, qi;
                        bj_pair(g.mp, r, i, &pi, &qi);
                        bj_vt_tile(As, Bs, VT, n, pi, qi, ct * BJ_P, ws + (long long)i * BJ_P * BJ_P);
                    }
                }
                c += nt;
            }
            alive = grid_sync(bar, G, &gen, &sh.ok);
        }
        if (!alive) break;
        for (int jl = 0; jl < nlist; ++jl) {
            if (!(live >> jl & 1)) continue;
            const int n = jobs[joblist[jl]].nao;
            const BjGeom g = bj_geom(n);
            if (xlws[xloff[jl] + bj_ws_sweep(g.np, n) + sweep] == 0.0) {
                live &= ~(1u << jl);
                if (tid == 0) sh.nsweep[jl] = sweep + 1;
            }
        }
    }
    if (!alive) live = started;   // abandoned: every job fails
    // eigenvalues ascending (NaN after +inf, ties by index: always a permutation)
    for (int jl = 0; jl < nlist && jl < BJ_MAXJOBS; ++jl) {
        if (!(started >> jl & 1)) continue;
        const JobDev jb = jobs[joblist[jl]];
        const int n = jb.nao;
        const double* W = matx(arena, jb, AR_W);
        double* dg = xlws + xloff[jl] + bj_ws_diag(bj_geom(n).np);
        for (long long p = gt0; p < n; p += gstride) dg[p] = W[p * n + p];
    }
    if (alive) alive = grid_sync(bar, G, &gen, &sh.ok);
    for (int jl = 0; jl < nlist && jl < BJ_MAXJOBS; ++jl) {
        if (!(started >> jl & 1)) continue;
        const int job = joblist[jl];
        const JobDev jb = jobs[job];
        const int n = jb.nao;
        const bool bad = ((live | failed) >> jl & 1) != 0;
        const double* dg = xlws + xloff[jl] + bj_ws_diag(bj_geom(n).np);
        for (long long p = gt0; p < n; p += gstride) {
            if (bad) {
                eps[jb.vec0 + p] = 0.0;
                perm[jb.ivec0 + p] = (int)p;
                continue;
            }
            const double d = dg[p], kd = isnan(d) ? INFINITY : d;
            int rank = 0;
            for (int q = 0; q < n; ++q) {
                const double e = dg[q], ke = isnan(e) ? INFINITY : e;
                rank += (ke < kd) || (ke == kd && q < p);
            }
            eps[jb.vec0 + rank] = d;
            perm[jb.ivec0 + rank] = (int)p;
        }
        if (wg == 0 && tid == 0) {
            state[job].sweeps = bad ? -1 : sh.nsweep[jl];
            state[job].warm = bad ? 0 : 1;
        }
    }
}

// ------------------------------------------------------------------ DIIS / energy partials
__device__ inline void xl_chunk(long long nn, int nchunk, int c, long long* lo, long long* hi) {
    const long long len = ((nn + nchunk - 1) / nchunk + 255) / 256 * 256;
    *lo = min(nn, (long long)c * len);
    *hi = min(nn, *lo + len);
}

__global__ void __launch_bounds__(256) k_xl_dots(const JobDev* __restrict__ jobs, const JobState* __restrict__ state,
                                                 const int* __restrict__ joblist, double* __restrict__ arena,
                                                 double* __restrict__ xlpart, int nchunk) {
    const int job = joblist[blockIdx.y];
    const JobState& st = state[job];
    if (st.phase != PH_ACTIVE || st.sweeps < 0) return;
    const JobDev jb = jobs[job];
    const int iter = st.iter, k = jb.diis_k;
    if ((iter == 0 && jb.requires_guess) || k <= 0) return;
    const int n = jb.nao, tid = threadIdx.x;
    const long long nn = (long long)n * n;
    long long lo, hi;
    xl_chunk(nn, nchunk, blockIdx.x, &lo, &hi);
    const double* Pin = matx(arena, jb, AR_P0 + (iter & 1));
    const double* Pout = matx(arena, jb, AR_P0 + ((iter + 1) & 1));
    const int rs = (st.ring_start + 1) % k, slot_new = (rs + k - 1) % k;
    double* RP = matx(arena, jb, AR_RING);
    double* RE = RP + (long long)k * nn;
    double part[NDDO_MAX_DIIS];
    for (int i = 0; i < k; ++i) part[i] = 0.0;
    double amax = 0.0;
    for (long long e = lo + tid; e < hi; e += 256) {
        const double pv = Pout[e], ev = pv - Pin[e];
        RP[(long long)slot_new * nn + e] = pv;
        RE[(long long)slot_new * nn + e] = ev;
        amax = fmax(amax, fabs(ev));
        for (int i = 0; i < k; ++i) {
            const int sl = (rs + i) % k;
            const double x = sl == slot_new ? ev : RE[(long long)sl * nn + e];
            part[i] += x * ev;
        }
    }
    __shared__ double red[256];
    double* rec = xlpart + ((long long)blockIdx.y * nchunk + blockIdx.x) * XL_REC;
    for (int i = 0; i <= k; ++i) {
        red[tid] = i < k ? part[i] : amax;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (tid < s) red[tid] = i < k ? red[tid] + red[tid + s] : fmax(red[tid], red[tid + s]);
            __syncthreads();
        }
        if (tid == 0) rec[i] = red[0];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_xl_apply(const JobDev* __restrict__ jobs, const int* __restrict__ joblist,
                                                  double* __restrict__ arena, const double* __restrict__ xlcoef,
                                                  int nchunk) {
    const int job = joblist[blockIdx.y];
    const double* cf = xlcoef + (long long)job * XL_COEF;
    const int flags = (int)cf[XL_COEF_FLAGS];
    if (!flags) return;
    const JobDev jb = jobs[job];
    const int n = jb.nao, k = jb.diis_k, tid = threadIdx.x;
    const int iter = (int)cf[XL_COEF_ITER], rs = (int)cf[XL_COEF_RS];
    const long long nn = (long long)n * n;
    long long lo, hi;
    xl_chunk(nn, nchunk, blockIdx.x, &lo, &hi);
    const double* Pin = matx(arena, jb, AR_P0 + (iter & 1));
    double* Pout = matx(arena, jb, AR_P0 + ((iter + 1) & 1));
    const double* RP = matx(arena, jb, AR_RING);
    const double w = jb.damp_weight;
    for (long long e = lo + tid; e < hi; e += 256) {
        double v = Pout[e];
        if (flags & 1) {
            double s = 0.0;
            for (int i = 0; i < k; ++i) s += cf[i] * RP[(long long)((rs + i) % k) * nn + e];
            v = s;
        }
        if (flags & 2) v = v * (1.0 - w) + w * Pin[e];
        Pout[e] = v;
    }
}

__global__ void __launch_bounds__(256) k_xl_energy(const JobDev* __restrict__ jobs, const JobState* __restrict__ state,
                                                   const int* __restrict__ joblist, double* __restrict__ arena,
                                                   double* __restrict__ xlpart, int nchunk) {
    const int job = joblist[blockIdx.y];
    const JobState& st = state[job];
    if (st.phase != PH_ENERGY) return;
    const JobDev jb = jobs[job];
    const int n = jb.nao, iter = st.iter, tid = threadIdx.x;
    const long long nn = (long long)n * n;
    long long lo, hi;
    xl_chunk(nn, nchunk, blockIdx.x, &lo, &hi);
    const double* P = matx(arena, jb, AR_P0 + ((iter + 1) & 1));
    const double* F = matx(arena, jb, AR_F);
    const double* H = matx(arena, jb, AR_H);
    double s = 0.0, c = 0.0;
    for (long long e = lo + tid; e < hi; e += 256) {
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
        for (int i = 0; i < 256; ++i) {
            const double y = ss[i] - sc[i] - C;
            const double t = S + y;
            C = (t - S) - y;
            S = t;
        }
        double* rec = xlpart + ((long long)blockIdx.y * nchunk + blockIdx.x) * XL_REC;
        rec[0] = S;
        rec[1] = C;
    }
}

__global__ void k_bar_reset(unsigned* bar) {
    bar[0] = 0u;
    bar[1] = 0u;
}

}  // namespace

long long xl_ws_doubles(int n) {
    const int m = (n + BJ_B - 1) / BJ_B, np = (m + (m & 1)) / 2;
    return (bj_ws_size(np, n) + 31) / 32 * 32;
}

int xl_grid(int num_cus, const int* ns, int nlist) {
    long long np = 0, tiles = 0;
    for (int i = 0; i < nlist; ++i) {
        const int m = (ns[i] + BJ_B - 1) / BJ_B, npi = (m + (m & 1)) / 2, nct = (ns[i] + BJ_P - 1) / BJ_P;
        np += npi;
        tiles += (long long)npi * (npi - 1) / 2 + (long long)npi * nct;
    }
    long long g = std::max<long long>(np, (tiles + 1) / 2);
    g = std::min<long long>(g, std::min(num_cus, BJ_MAXGRID));
    return (int)std::max<long long>(g, 1);
}

int xl_max_jobs() { return BJ_MAXJOBS; }

int xl_nchunk(int max_n) {
    const long long nn = (long long)max_n * max_n;
    return (int)std::max<long long>(1, std::min<long long>(1024, (nn + 16383) / 16384));
}

void launch_xl_eigen(hipStream_t st, int grid, const JobDev* jobs, JobState* state, const int* joblist,
                     const long long* xloff, int nlist, double* arena, double* eps, int* perm, double* xlws,
                     unsigned* bar) {
    if (nlist <= 0) return;
    const size_t lds = (size_t)2 * BJ_P * BJ_LD * sizeof(double);
    hipFuncSetAttribute((const void*)k_bjac, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_bar_reset, dim3(1), dim3(1), 0, st, bar);
    hipLaunchKernelGGL(k_bjac, dim3(grid), dim3(BJ_T), lds, st, jobs, state, joblist, xloff, nlist, arena, eps, perm,
                       xlws, bar);
}

void launch_xl_dots(hipStream_t st, const JobDev* jobs, const JobState* state, const int* joblist, int nlist,
                    double* arena, double* xlpart, int nchunk) {
    if (nlist <= 0) return;
    hipLaunchKernelGGL(k_xl_dots, dim3(nchunk, nlist), dim3(256), 0, st, jobs, state, joblist, arena, xlpart, nchunk);
}

void launch_xl_apply(hipStream_t st, const JobDev* jobs, const int* joblist, int nlist, double* arena,
                     const double* xlcoef, int nchunk) {
    if (nlist <= 0) return;
    hipLaunchKernelGGL(k_xl_apply, dim3(nchunk, nlist), dim3(256), 0, st, jobs, joblist, arena, xlcoef, nchunk);
}

void launch_xl_energy(hipStream_t st, const JobDev* jobs, const JobState* state, const int* joblist, int nlist,
                      double* arena, double* xlpart, int nchunk) {
    if (nlist <= 0) return;
    hipLaunchKernelGGL(k_xl_energy, dim3(nchunk, nlist), dim3(256), 0, st, jobs, state, joblist, arena, xlpart,
                       nchunk);
}

}  // namespace nddo
