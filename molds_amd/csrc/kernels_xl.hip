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
constexpr int BJ_LD = 68;             // LDS leading dimension of a 64 x 64 tile (2-way bank
                                      // aliasing at most for the MFMA operand reads)
constexpr int BJ_MAXSWEEP = 50;
// One cyclic sweep per pair block and round: diagonalizing the pair blocks
// exactly buys no fewer outer sweeps (measured on random and near-diagonal
// matrices) and costs several times more.
constexpr int BJ_INNER_MAXSWEEP = 1;
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

// Newton-Raphson refinements of the hardware reciprocal and reciprocal square
// root estimates (v_rcp_f64, v_rsq_f64): two steps each, for arguments well
// inside the normal range.
__device__ inline double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
__device__ inline double rsq_nr(double x) {
    double r = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    double e = fma(-h, r * r, 0.5);
    r = fma(r, e, r);
    e = fma(-h, r * r, 0.5);
    return fma(r, e, r);
}

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
            // relaxed polls, one acquire afterwards (one cache invalidation per barrier)
            while (__hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != target) {
                __builtin_amdgcn_s_sleep(2);
                // 100 MHz constant clock: give up after 20 s
                if (wall_clock64() - t0 > 2000000000ull) { ok = 0; break; }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
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
    unsigned long long it[4];   // workgroup 0: pair-block load, rotations, write-back, phase-1 barrier wait
    unsigned char bki[528], bkj[528];   // (row pair, column pair) of the 2 x 2 blocks, kj >= ki
    int any[2];   // by round parity: a wave may still read round r's flag while wave 0 writes r + 1's
    int ok;
};

// Phase 1: exact eigen-decomposition of the 64 x 64 pair block k (round r):
// S <- GT S GT^T diagonal; S written back to the matrix, GT to the workspace.
// The 64 x 64 tile loops below have a fixed trip count per thread (thread t
// owns column t % 64 of rows t / 64 + 4 i), fully unrolled so that all of a
// thread's loads are in flight at once.
#define BJ_ROWS 16   // rows per thread of a 64 x 64 tile

// Phase 1: one cyclic Jacobi sweep of the 64 x 64 pair block (p, q):
// S <- GT S GT^T written back to the matrix, GT to the workspace.
__device__ void bj_inner(double* S, double* G, BjShared& sh, double* W, int n, int p, int q, double thr, double* gt,
                         double* flag, double* sweepflag) {
    const int tid = threadIdx.x, b = tid & 63, a0 = tid >> 6;
    const bool clk = blockIdx.x == 0 && tid == 0;
    unsigned long long c0 = clk ? wall_clock64() : 0ull;
    const int rb = bj_row(p, q, b);
    double v[BJ_ROWS];
#pragma unroll
    for (int i = 0; i < BJ_ROWS; ++i) {
        const int ra = bj_row(p, q, a0 + 4 * i);
        v[i] = (ra < n && rb < n) ? W[(long long)ra * n + rb] : 0.0;
    }
    double offmax = 0.0;
#pragma unroll
    for (int i = 0; i < BJ_ROWS; ++i) {
        const int a = a0 + 4 * i;
        S[a * BJ_LD + b] = v[i];
        G[a * BJ_LD + b] = a == b ? 1.0 : 0.0;
        if (a != b) offmax = fmax(offmax, fabs(v[i]));
    }
    // the whole block is already diagonal to the threshold: nothing to rotate
    // (a NaN entry also lands here: fmax ignores it and the solve fails later
    // on the non-finite norm)
    const int need = __syncthreads_or(offmax > thr);
    if (clk) {
        const unsigned long long c1 = wall_clock64();
        sh.it[0] += c1 - c0;
        c0 = c1;
    }
    if (!need) {
        if (tid == 0) *flag = 0.0;
        return;
    }
    for (int isw = 0; isw < BJ_INNER_MAXSWEEP; ++isw) {
        int any = 0;
        for (int r = 0; r < 63; ++r) {
            if (tid < 64) {   // wave 0: the 32 rotations of this round
                int rot = 0;
                int pp, qq;
                const int kk = tid & 31;
                if (kk == 0) { pp = r; qq = 63; }
                else { pp = (r + kk) % 63; qq = (r - kk + 63) % 63; }
                if (pp > qq) { const int t = pp; pp = qq; qq = t; }
                double c = 1.0, s = 0.0, t = 0.0, u = 0.0;
                const double apq = S[pp * BJ_LD + qq];
                if (tid < 32 && fabs(apq) > thr) {
                    // t: the smaller root of t^2 + 2 tau t - 1 = 0, tau = (a_qq - a_pp) / (2 a_pq),
                    // as 2 a_pq sign(d) / (|d| + sqrt(d^2 + 4 a_pq^2)) with d = a_qq - a_pp
                    const double app = S[pp * BJ_LD + pp], aqq = S[qq * BJ_LD + qq];
                    const double d = aqq - app, h = fma(d, d, 4.0 * apq * apq);
                    if (h > 1e-290 && h < 1e290) {
                        t = (d >= 0.0 ? 2.0 * apq : -2.0 * apq) * rcp_nr(fabs(d) + h * rsq_nr(h));
                        c = rsq_nr(fma(t, t, 1.0));
                        s = t * c;
                        u = s * rcp_nr(1.0 + c);
                    } else {   // extreme scales: the same rotation with IEEE division and square root
                        const double tau = d / (2.0 * apq);
                        t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                        c = 1.0 / sqrt(1.0 + t * t);
                        s = t * c;
                        u = s / (1.0 + c);
                    }
                    rot = 1;
                }
                if (tid < 32) {
                    sh.sp[tid] = pp; sh.sq[tid] = qq; sh.sc[tid] = c; sh.ss[tid] = s; sh.st[tid] = t; sh.su[tid] = u;
                    sh.srot[tid] = rot;
                }
                const bool anyr = __ballot(tid < 32 && rot) != 0ull;
                if (tid == 0) sh.any[r & 1] = anyr ? 1 : 0;
            }
            __syncthreads();
            if (!sh.any[r & 1]) continue;
            any = 1;
            // 2 x 2 blocks (rows {p1,q1}, columns {p2,q2}) of the upper block triangle,
            // three per thread; the blocks are disjoint, so every read goes first
            int P1[3], Q1[3], P2[3], Q2[3], KI[3], KJ[3];
            bool act[3];
            double x0[3], x1[3], x2[3], x3[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const int bb = tid + BJ_T * i;
                const int ki = bb < 528 ? sh.bki[bb] : 0, kj = bb < 528 ? sh.bkj[bb] : 0;
                KI[i] = ki;
                KJ[i] = kj;
                act[i] = bb < 528 && (sh.srot[min(ki, 31)] || sh.srot[min(kj, 31)]);
                P1[i] = sh.sp[min(ki, 31)]; Q1[i] = sh.sq[min(ki, 31)];
                P2[i] = sh.sp[min(kj, 31)]; Q2[i] = sh.sq[min(kj, 31)];
                x0[i] = x1[i] = x2[i] = x3[i] = 0.0;
                if (act[i]) {
                    if (ki == kj) {
                        x0[i] = S[P1[i] * BJ_LD + Q1[i]];
                    } else {
                        x0[i] = S[P1[i] * BJ_LD + P2[i]];
                        x1[i] = S[P1[i] * BJ_LD + Q2[i]];
                        x2[i] = S[Q1[i] * BJ_LD + P2[i]];
                        x3[i] = S[Q1[i] * BJ_LD + Q2[i]];
                    }
                }
            }
            double aP[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) aP[i] = (act[i] && KI[i] == KJ[i]) ? S[P1[i] * BJ_LD + P1[i]] : 0.0;
            double aQ[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) aQ[i] = (act[i] && KI[i] == KJ[i]) ? S[Q1[i] * BJ_LD + Q1[i]] : 0.0;
            // GT rows of the rotations: GT <- J^T GT, eight (pair, column) entries per thread
            double gp[8], gq[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int k = a0 + 4 * i;
                gp[i] = G[sh.sp[k] * BJ_LD + b];
                gq[i] = G[sh.sq[k] * BJ_LD + b];
            }
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                if (!act[i]) continue;
                const int ki = KI[i], kj = KJ[i];
                const int p1 = P1[i], q1 = Q1[i], p2 = P2[i], q2 = Q2[i];
                if (ki == kj) {
                    const double t = sh.st[ki], apq = x0[i];
                    S[p1 * BJ_LD + p1] = aP[i] - t * apq;
                    S[q1 * BJ_LD + q1] = aQ[i] + t * apq;
                    S[p1 * BJ_LD + q1] = 0.0;
                    S[q1 * BJ_LD + p1] = 0.0;
                    continue;
                }
                // rotations in Rutishauser's form x - s (y + u x), y + s (x - u y), which
                // keeps them orthogonal to rounding also when 1 + t^2 rounds to 1
                const double s1 = sh.ss[ki], u1 = sh.su[ki], s2 = sh.ss[kj], u2 = sh.su[kj];
                const double a = x0[i], bq = x1[i], cq = x2[i], d = x3[i];
                const double a1 = a - s1 * (cq + u1 * a), b1 = bq - s1 * (d + u1 * bq);
                const double c1r = cq + s1 * (a - u1 * cq), d1 = d + s1 * (bq - u1 * d);
                const double a2 = a1 - s2 * (b1 + u2 * a1), b2 = b1 + s2 * (a1 - u2 * b1);
                const double c2r = c1r - s2 * (d1 + u2 * c1r), d2 = d1 + s2 * (c1r - u2 * d1);
                S[p1 * BJ_LD + p2] = a2; S[p2 * BJ_LD + p1] = a2;
                S[p1 * BJ_LD + q2] = b2; S[q2 * BJ_LD + p1] = b2;
                S[q1 * BJ_LD + p2] = c2r; S[p2 * BJ_LD + q1] = c2r;
                S[q1 * BJ_LD + q2] = d2; S[q2 * BJ_LD + q1] = d2;
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int k = a0 + 4 * i;
                if (!sh.srot[k]) continue;
                const double sn = sh.ss[k], u = sh.su[k];
                G[sh.sp[k] * BJ_LD + b] = gp[i] - sn * (gq[i] + u * gp[i]);
                G[sh.sq[k] * BJ_LD + b] = gq[i] + sn * (gp[i] - u * gq[i]);
            }
            __syncthreads();
        }
        if (!any) break;
        // a sweep that would rotate nothing leaves S as it is, so checking every
        // off-diagonal entry against the threshold now decides the same thing
        double om = 0.0;
#pragma unroll
        for (int i = 0; i < BJ_ROWS; ++i) {
            const int a = a0 + 4 * i;
            if (a != b) om = fmax(om, fabs(S[a * BJ_LD + b]));
        }
        if (!__syncthreads_or(om > thr)) break;
    }
    if (clk) {
        const unsigned long long c1 = wall_clock64();
        sh.it[1] += c1 - c0;
        c0 = c1;
    }
#pragma unroll
    for (int i = 0; i < BJ_ROWS; ++i) {
        const int a = a0 + 4 * i;
        const int ra = bj_row(p, q, a);
        if (ra < n && rb < n) W[(long long)ra * n + rb] = S[a * BJ_LD + b];
        gt[a * BJ_P + b] = G[a * BJ_LD + b];
    }
    if (tid == 0) {
        *flag = 1.0;
        *sweepflag = 1.0;
    }
    if (clk) sh.it[2] += wall_clock64() - c0;
}

// Phase 2, off-diagonal tile (i, j), i < j:  A_ij <- GT_i A_ij GT_j^T, both triangles.
__device__ void bj_tile(double* As, double* Bs, double* W, int n, int pi, int qi, int pj, int qj, bool fi, bool fj,
                        const double* gti, const double* gtj) {
    const int tid = threadIdx.x, b = tid & 63, a0 = tid >> 6;
    {
        const int rb = bj_row(pj, qj, b);
        double v[BJ_ROWS], g[BJ_ROWS];
#pragma unroll
        for (int i = 0; i < BJ_ROWS; ++i) {
            const int ra = bj_row(pi, qi, a0 + 4 * i);
            v[i] = (ra < n && rb < n) ? W[(long long)ra * n + rb] : 0.0;
            g[i] = fj ? gtj[(a0 + 4 * i) * BJ_P + b] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < BJ_ROWS; ++i) {
            As[(a0 + 4 * i) * BJ_LD + b] = v[i];
            Bs[(a0 + 4 * i) * BJ_LD + b] = g[i];
        }
    }
    __syncthreads();
    dbl4 acc[4];
    if (fj) {   // X = A GT_j^T
        bj_gemm<true>(As, Bs, acc);
        __syncthreads();
        bj_store(As, acc);
    }
    if (fi) {   // T = GT_i X
        double g[BJ_ROWS];
#pragma unroll
        for (int i = 0; i < BJ_ROWS; ++i) g[i] = gti[(a0 + 4 * i) * BJ_P + b];
#pragma unroll
        for (int i = 0; i < BJ_ROWS; ++i) Bs[(a0 + 4 * i) * BJ_LD + b] = g[i];
        __syncthreads();
        bj_gemm<false>(Bs, As, acc);
        __syncthreads();
        bj_store(As, acc);
    }
    __syncthreads();
    {
        const int rbj = bj_row(pj, qj, b), rbi = bj_row(pi, qi, b);
#pragma unroll
        for (int i = 0; i < BJ_ROWS; ++i) {
            const int a = a0 + 4 * i;
            const int ra = bj_row(pi, qi, a);
            if (ra < n && rbj < n) W[(long long)ra * n + rbj] = As[a * BJ_LD + b];
            // transposed copy: row (pj,qj)[a], column (pi,qi)[b]
            const int ta = bj_row(pj, qj, a);
            if (ta < n && rbi < n) W[(long long)ta * n + rbi] = As[b * BJ_LD + a];
        }
    }
    __syncthreads();
}

// Phase 2, eigenvector tile: rows (pi, qi) x columns c0 .. c0 + 63 of VT <- GT_i VT.
__device__ void bj_vt_tile(double* As, double* Bs, double* VT, int n, int pi, int qi, int c0, const double* gti) {
    const int tid = threadIdx.x, b = tid & 63, a0 = tid >> 6;
    {
        const int cb = c0 + b;
        double v[BJ_ROWS], g[BJ_ROWS];
#pragma unroll
        for (int i = 0; i < BJ_ROWS; ++i) {
            const int ra = bj_row(pi, qi, a0 + 4 * i);
            v[i] = (ra < n && cb < n) ? VT[(long long)ra * n + cb] : 0.0;
            g[i] = gti[(a0 + 4 * i) * BJ_P + b];
        }
#pragma unroll
        for (int i = 0; i < BJ_ROWS; ++i) {
            As[(a0 + 4 * i) * BJ_LD + b] = v[i];
            Bs[(a0 + 4 * i) * BJ_LD + b] = g[i];
        }
    }
    __syncthreads();
    dbl4 acc[4];
    bj_gemm<false>(Bs, As, acc);
    const int lane = tid & 63, w = tid >> 6, li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int vv = 0; vv < 4; ++vv) {
            const int a = 16 * w + 4 * vv + lk, c = c0 + 16 * cb + li;
            const int ra = bj_row(pi, qi, a);
            if (ra < n && c < n) VT[(long long)ra * n + c] = acc[cb][vv];
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
    // device clock ticks seen by workgroup 0: phase 1 (pair blocks, up to the
    // barrier), phase 2 (tiles, up to the barrier), the rest; rounds run
    unsigned long long tk[3] = {0, 0, 0}, tl = wall_clock64(), nrounds = 0;
    const unsigned long long cyc0 = clock64(), wal0 = tl;
    if (tid < 4) sh.it[tid] = 0;
    for (int hi = 0; hi < 32; ++hi)
        for (int ki = tid; ki <= hi; ki += BJ_T) {
            sh.bki[hi * (hi + 1) / 2 + ki] = (unsigned char)ki;
            sh.bkj[hi * (hi + 1) / 2 + ki] = (unsigned char)hi;
        }
    __syncthreads();
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
            const unsigned long long tb = wall_clock64();
            alive = grid_sync(bar, G, &gen, &sh.ok);
            if (wg == 0 && tid == 0) sh.it[3] += wall_clock64() - tb;
            if (wg == 0 && tid == 0) {
                const unsigned long long t = wall_clock64();
                tk[0] += t - tl;
                tl = t;
                ++nrounds;
            }
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
            if (wg == 0 && tid == 0) {
                const unsigned long long t = wall_clock64();
                tk[1] += t - tl;
                tl = t;
            }
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
            const unsigned long long t = wall_clock64();
            state[job].prof[0] = (double)tk[0];
            state[job].prof[1] = (double)tk[1];
            state[job].prof[2] = (double)(tk[2] + (t - tl));
            state[job].prof[3] = (double)nrounds;
            // the same and the parts of workgroup 0's phase 1, for the profiling hook
            double* dt = xlws + xloff[jl] + bj_ws_sweep(bj_geom(n).np, n) + BJ_MAXSWEEP;
            for (int k = 0; k < 4; ++k) dt[k] = (double)sh.it[k];
            dt[4] = (double)tk[0];
            dt[5] = (double)tk[1];
            dt[6] = (double)nrounds;
            dt[7] = (double)(clock64() - cyc0);    // shader clock cycles over the solve
            dt[8] = (double)(wall_clock64() - wal0);
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

long long xl_dbg_offset(int n) {
    const int m = (n + BJ_B - 1) / BJ_B, np = (m + (m & 1)) / 2;
    return bj_ws_sweep(np, n) + BJ_MAXSWEEP;
}

int xl_nchunk(int max_n) {
    const long long nn = (long long)max_n * max_n;
    return (int)std::max<long long>(1, std::min<long long>(1024, (nn + 16383) / 16384));
}

void launch_xl_eigen(hipStream_t st, int grid, const JobDev* jobs, JobState* state, const int* joblist,
                     const long long* xloff, int nlist, double* arena, double* eps, int* perm, double* xlws,
                     unsigned* bar) {
    if (nlist <= 0) return;
    const size_t lds = (size_t)2 * BJ_P * BJ_LD * sizeof(double);
    raise_dynamic_lds((const void*)k_bjac, lds);
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
