// Device-wide FP64 symmetric eigensolver for large molecules (replaces
// LAPACKE_dsyevd, src/wrappers/Lapack.cpp:102-161, where one workgroup per
// molecule would leave the chip idle and its tables no longer fit LDS):
//
//   k_xtri   Householder tridiagonalization T = Q^T F Q (Golub & Van Loan,
//            Alg. 8.3.1) by the whole grid: rows of the trailing block are
//            dealt to the wavefronts, one row per wavefront at a time; step k
//            updates every row by the rank-2 correction of reflector k and, in
//            the same pass over the row, forms its part of A v_{k+1} for the
//            next reflector, so the trailing block is read and written once
//            per step.  The pivot row of the next reflector is rebuilt by every
//            workgroup from its stored copy and the previous step's v and w
//            (so nobody writes it while others read it), with the same
//            fixed-order reductions everywhere, hence bitwise the same v in
//            every workgroup.  One grid barrier per step.
//   k_xtri_blk  the same reduction in panels of 8 reflectors (from 5600 AOs):
//            read-only row passes and one rank-16 FP64 MFMA update per panel.
//   k_xdc    Cuppen divide and conquer on T, bottom-up from 1 x 1 leaves, all
//            merges of a level at once, as the one-workgroup k_dc
//            (kernels_eig.hip) with its per-row tables in device memory and
//            every phase spread over the grid: deflation (Dongarra-Sorensen),
//            secular roots (rational model with bisection safeguard), Loewner
//            eigenvector components (Gu-Eisenstat) and the merge products on
//            FP64 MFMA; the per-element loops over a merge block give each
//            element a group of lanes.  One grid barrier per phase.
//   k_xbt    eigenvector rows <- rows * H_{n-3} ... H_0 in compact WY blocks of
//            16 reflectors (T from k_wy_t), two blocks per pass, one workgroup
//            per 16 rows streaming them through FP64 MFMA; no synchronisation
//            between workgroups.
//
// Output as the one-workgroup path: rows of AR_C are the eigenvectors in
// ascending eigenvalue order, eps[] ascending, perm[] = identity.
//
// The persistent kernels use a counter/generation barrier in device memory
// with agent-scope release/acquire; all their workgroups are resident (at most
// one per CU, and the engine runs one such launch at a time); a wait that runs
// out (20 s) marks the solve failed instead of hanging.
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>

#include "eig_device.h"
#include "launch.h"

namespace nddo {

namespace {

constexpr int XT = 256;             // threads per workgroup
constexpr int XW = XT / 64;         // wavefronts per workgroup
constexpr int XGRID_MAX = 1024;     // workgroups per persistent launch
constexpr int XWAVE_MAX = XGRID_MAX * XW;
constexpr int kXlhMaxN = 6144;      // k_xtri(_blk) keep v, w and the next v in LDS: 3 n doubles

// Grid barrier in two levels, so that no counter sees more than 16 atomic
// arrivals: the workgroups of each group of XB count on the group's counter,
// the last of a group counts on the top counter, the last group releases the
// generation word every workgroup waits on.  Words 128 bytes apart:
// bar[0] generation, bar[32] top counter, bar[64 + 32 g] group g's counter
// (XBAR_WORDS per barrier, XMAX_GROUPS barriers, cleared by k_xbar_reset
// before each launch).  Only
// thread 0 makes agent-scope atomics: its release (L2 write-back) and acquire
// (cache invalidation) act on the caches its workgroup shares, after the
// workgroup barrier has made every wavefront's stores visible there.  Returns
// false when the wait ran out (20 s).
constexpr int XB = 16;
constexpr int XBAR_WORDS = 64 + 32 * (XGRID_MAX / XB);   // one barrier
constexpr int XMAX_GROUPS = 32;                           // molecules solved side by side per launch
constexpr int XMIN_GROUP = 8;                             // fewest workgroups per molecule
// blk: the workgroup's index among the nblk of its barrier
__device__ bool xsync(unsigned* bar, unsigned nblk, unsigned blk, unsigned* gen, int* sh_ok) {
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        const unsigned target = *gen + 1;
        const unsigned grp = blk / XB, ngrp = (nblk + XB - 1) / XB;
        const unsigned gsize = min((unsigned)XB, nblk - grp * XB);
        unsigned* cnt = bar + 64 + 32 * grp;
        bool release = false;
        // Arrivals are release RMWs; whoever goes on acquires with one fence
        // afterwards (the RMWs on a counter form one release sequence), so the
        // cache invalidation of an acquire happens once per barrier and
        // workgroup, not once per poll of the generation word.
        if (__hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT) == gsize - 1) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (__hip_atomic_fetch_add(bar + 32, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT) == ngrp - 1) {
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                __hip_atomic_store(bar + 32, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(bar, target, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                release = true;
            }
        }
        if (!release) {
            const unsigned long long t0 = wall_clock64();
            while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != target) {
                __builtin_amdgcn_s_sleep(1);
                if (wall_clock64() - t0 > 2000000000ull) { ok = 0; break; }   // 20 s of the 100 MHz clock
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        *gen = target;
        *sh_ok = ok;
    }
    __syncthreads();
    return *sh_ok != 0;
}

// The molecules of a launch are solved side by side: the grid is split into
// ngroups = min(nlist, grid / XMIN_GROUP) groups of consecutive workgroups,
// group g takes the jobs g, g + ngroups, ... one after another, with a
// barrier of its own.
struct XGroup {
    int ngroups, grp, G, wg;
    unsigned* bar;
};
__device__ inline XGroup x_group(int nlist, unsigned* bar) {
    XGroup x;
    const int grid = (int)gridDim.x;
    x.ngroups = max(1, min(min(nlist, XMAX_GROUPS), grid / XMIN_GROUP));
    const int gsz = grid / x.ngroups;
    x.grp = min((int)blockIdx.x / gsz, x.ngroups - 1);
    x.G = x.grp == x.ngroups - 1 ? grid - x.grp * gsz : gsz;
    x.wg = (int)blockIdx.x - x.grp * gsz;
    x.bar = bar + (long long)x.grp * XBAR_WORDS;
    return x;
}

// Sum of one double per wavefront of the grid (q[0 .. nwaves)) in a fixed
// order, the same in every workgroup; every thread gets it.
__device__ inline double grid_partials_sum(const double* q, int nwaves, double* sh) {
    const int tid = threadIdx.x, lane = tid & 63;
    if (tid < 64) {
        // four loads in flight per lane (one pass up to 256 partials)
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        for (int g = lane; g < nwaves; g += 256) {
            s0 += q[g];
            if (g + 64 < nwaves) s1 += q[g + 64];
            if (g + 128 < nwaves) s2 += q[g + 128];
            if (g + 192 < nwaves) s3 += q[g + 192];
        }
        double s = (s0 + s1) + (s2 + s3);
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (lane == 0) *sh = s;
    }
    __syncthreads();
    return *sh;
}

constexpr int XNB = 8;              // reflectors per panel of the blocked k_xtri_blk
constexpr int XSR = 6 * XNB;        // k_xtri_blk's per-workgroup record of panel dot partials
// Size from which k_xtri_blk is the default.  Measured on MI355X (ms per
// tridiagonalization, per-reflector / blocked): n = 4608: 149 / 165; n = 5200:
// 198 / 203; n = 6000: 293 / 273.  The blocked row pass reads half the bytes
// (1.75x faster at n = 6000), but the panel update and the larger pivot phase
// add a fixed cost per column that smaller trailing blocks do not repay.
constexpr int kXtriBlkMinN = 5600;

// Workspace of one job in xlws (doubles): p[2][n], per-wave partials [2][XWAVE_MAX],
// the divide-and-conquer tables (12 n doubles, 9 n ints), then k_xtri_blk's
// w vectors of the current panel (XNB rows of n, by global index), its next
// pivot row [2][n] and its panel dot partials [2][XSR][XGRID_MAX].
__host__ __device__ inline long long xw_p(int) { return 0; }
__host__ __device__ inline long long xw_q(int n) { return 2LL * n; }
__host__ __device__ inline long long xw_dc(int n) { return 2LL * n + 2LL * XWAVE_MAX; }
__host__ __device__ inline long long xw_wp(int n) { return xw_dc(n) + 12LL * n + (9LL * n + 1) / 2 + 64; }
__host__ __device__ inline long long xw_xq(int n) { return xw_wp(n) + (long long)XNB * n; }
__host__ __device__ inline long long xw_xs(int n) { return xw_xq(n) + 2LL * n; }
__host__ __device__ inline long long xw_size(int n) { return xw_xs(n) + 2LL * XGRID_MAX * XSR; }

// ------------------------------------------------------------------ tridiagonalization
constexpr int TT = 1024;            // threads per workgroup of k_xtri (16 wavefronts: memory parallelism)
constexpr int XU = 16;              // row elements loaded at once per lane in k_xtri's row pass
constexpr int XP = 8;               // pivot-row and p elements loaded at once per thread (n <= 8192 in one)
static_assert(kXlhMaxN <= TT * XP, "one pass of XP loads per thread covers the pivot row and p");
constexpr int TW = TT / 64;

__global__ void __launch_bounds__(TT) k_xtri(const JobDev* __restrict__ jobs, JobState* __restrict__ state,
                                             const int* __restrict__ joblist, const long long* __restrict__ xloff,
                                             int nlist, double* __restrict__ arena, double* __restrict__ tri,
                                             double* __restrict__ xlws, unsigned* __restrict__ bar) {
    extern __shared__ double dyn[];
    __shared__ double red[TW], red2[TW], bc[2];
    __shared__ int sh_ok;
    const XGroup xg = x_group(nlist, bar);
    const int G = xg.G, wg = xg.wg, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int gw = wg * TW + wave, GW = G * TW;
    unsigned gen = 0;
    bool alive = true;
    for (int jl = xg.grp; jl < nlist; jl += xg.ngroups) {
        const int job = joblist[jl];
        if (state[job].phase != PH_ACTIVE) continue;   // read-only here: the same in every workgroup
        const unsigned long long t0 = wall_clock64();
        const JobDev jb = jobs[job];
        const int n = jb.nao;
        const long long ln = n;
        double* A = amat(arena, jb, AR_U);   // the trailing block, full and symmetric
        double* Wm = amat(arena, jb, AR_W);
        const double* F = amat(arena, jb, AR_F);
        double* d = eig_ws(tri, jb, job);
        double* e = d + n;
        double* tau = e + n;
        double* ws = xlws + xloff[jl];
        double* pb[2] = {ws + xw_p(n), ws + xw_p(n) + n};
        // one partial of p.v per workgroup (its wavefronts' partials summed in a fixed order)
        double* qb[2] = {ws + xw_q(n), ws + xw_q(n) + XGRID_MAX};
        double* vcur = dyn;
        double* vnext = dyn + n;
        double* wv = dyn + 2 * n;
        if (!alive) {
            if (wg == 0 && tid == 0) state[job].sweeps = -1;
            continue;
        }
        if (wg == 0)
            for (int i = tid; i < n; i += TT) { e[i] = 0.0; tau[i] = 0.0; }
        if (n < 3) {
            if (wg == 0 && tid == 0) {
                d[0] = F[0];
                if (n == 2) { e[0] = F[1]; d[1] = F[3]; }
                state[job].prof[0] = (double)(wall_clock64() - t0);
            }
            continue;
        }
        // the workgroup's partial of p.v: its wavefronts' sums added in order
        auto wg_partial = [&](double pacc, double* q) {
            if (lane == 0) red2[wave] = pacc;
            __syncthreads();
            if (tid == 0) {
                double s = 0.0;
                for (int w = 0; w < TW; ++w) s += red2[w];
                q[wg] = s;
            }
        };
        // reflector of row 0 from F, then A = F (rows 1 .. n-1) and p = tau_0 A v_0
        double tk;
        {
            double ss = 0.0;
            for (int t = 2 + tid; t < n; t += TT) { const double x = F[t]; ss += x * x; }
            ss = block_sum(ss, red, tid, TW);
            const double alpha = F[1];
            double beta = alpha, scale = 0.0;
            tk = 0.0;
            if (ss > 0.0) {
                beta = -copysign(sqrt(alpha * alpha + ss), alpha);
                tk = (beta - alpha) / beta;
                scale = 1.0 / (alpha - beta);
            }
            for (int t = 1 + tid; t < n - 1; t += TT) vcur[t] = ss > 0.0 ? F[1 + t] * scale : 0.0;
            if (tid == 0) vcur[0] = 1.0;
            __syncthreads();
            if (wg == 0) {
                for (int t = tid; t < n - 1; t += TT) Wm[1 + t] = vcur[t];
                if (tid == 0) { d[0] = F[0]; e[0] = beta; tau[0] = tk; }
            }
            double pv = 0.0;
            for (int r = 1 + (gw - 1 + GW) % GW; r < n; r += GW) {   // row r with wavefront r mod GW
                double s = 0.0;
                int c = 1 + lane;
                for (; c + 64 * 7 < n; c += 64 * 8) {
                    double a[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) a[u] = F[(long long)r * ln + c + 64 * u];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        A[(long long)r * ln + c + 64 * u] = a[u];
                        s += a[u] * vcur[c + 64 * u - 1];
                    }
                }
                if (c < n) {   // the last (partial) chunk
                    double a[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) a[u] = c + 64 * u < n ? F[(long long)r * ln + c + 64 * u] : 0.0;
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        if (c + 64 * u < n) {
                            A[(long long)r * ln + c + 64 * u] = a[u];
                            s += a[u] * vcur[c + 64 * u - 1];
                        }
                    }
                }
                for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
                const double pn = tk * s;
                if (lane == 0) pb[0][r - 1] = pn;
                pv += pn * vcur[r - 1];
            }
            wg_partial(pv, qb[0]);
        }
        alive = xsync(xg.bar, G, wg, &gen, &sh_ok);
        int cur = 0;
        for (int k = 0; alive && k <= n - 3; ++k) {
            const int m = n - 1 - k;   // length of v_k (indices t <-> row / column k + 1 + t)
            // this thread's p and pivot-row elements (t = tid + TT u, m <= TT XP),
            // loaded before the wait for the partials so the three loads overlap
            const double* rowp = A + (long long)(k + 1) * ln + (k + 1);
            double pr[XP], ar[XP];
#pragma unroll
            for (int u = 0; u < XP; ++u) {
                const int t = tid + TT * u;
                pr[u] = t < m ? pb[cur][t] : 0.0;
                ar[u] = t < m ? rowp[t] : 0.0;
            }
            const double pv = grid_partials_sum(qb[cur], G, &bc[0]);
            const double K = -0.5 * tk * pv;   // w = p + K v
#pragma unroll
            for (int u = 0; u < XP; ++u) {
                const int t = tid + TT * u;
                if (t < m) wv[t] = pr[u] + K * vcur[t];
            }
            __syncthreads();
            const double w0 = wv[0];
            if (k == n - 3) {   // last trailing 2 x 2 block (rows / columns n-2, n-1)
                if (wg == 0 && tid == 0) {
                    const double w1 = wv[1];
                    d[n - 2] = A[(long long)(n - 2) * ln + n - 2] - 2.0 * vcur[0] * w0;
                    e[n - 2] = A[(long long)(n - 2) * ln + n - 1] - (vcur[0] * w1 + w0 * vcur[1]);
                    d[n - 1] = A[(long long)(n - 1) * ln + n - 1] - 2.0 * vcur[1] * w1;
                }
                break;
            }
            // row k+1 after the update of step k; its tail defines reflector k+1
            double ss = 0.0;
#pragma unroll
            for (int u = 0; u < XP; ++u) {
                const int t = tid + TT * u;
                if (t < m) {
                    const double a = ar[u] - (vcur[0] * wv[t] + w0 * vcur[t]);
                    if (t >= 2) ss += a * a;
                    else if (t == 1) bc[1] = a;
                    else if (wg == 0) d[k + 1] = a;
                    if (t >= 1) vnext[t - 1] = a;   // scaled below
                }
            }
            ss = block_sum(ss, red, tid, TW);   // its barriers also publish bc[1] and vnext
            const double alpha = bc[1];
            double beta = alpha, tn = 0.0, scale = 0.0;
            if (ss > 0.0) {
                beta = -copysign(sqrt(alpha * alpha + ss), alpha);
                tn = (beta - alpha) / beta;
                scale = 1.0 / (alpha - beta);
            }
            const int m2 = m - 1;   // length of v_{k+1}
            for (int t = 1 + tid; t < m2; t += TT) vnext[t] = ss > 0.0 ? vnext[t] * scale : 0.0;
            if (tid == 0) vnext[0] = 1.0;
            __syncthreads();
            if (wg == 0) {
                double* wr = Wm + (long long)(k + 1) * ln + (k + 2);
                for (int t = tid; t < m2; t += TT) wr[t] = vnext[t];
                if (tid == 0) { e[k + 1] = beta; tau[k + 1] = tn; }
            }
            // rows r = k+2 .. n-1: rank-2 update of step k and (A v_{k+1})_r in one
            // pass, XU independent loads in flight per lane (the pass is bound by
            // the bytes a CU keeps in flight)
            const int nx = cur ^ 1;
            double pacc = 0.0;
            // row r stays with wavefront r mod GW at every step (its workgroup's
            // L2 keeps the rows it rewrites)
            for (int r = k + 2 + (gw - (k + 2) % GW + GW) % GW; r < n; r += GW) {
                const int tr = r - k - 1;
                const double vr = vcur[tr], wr = wv[tr];
                double* row = A + (long long)r * ln;
                double s = 0.0;
                int c = k + 2 + lane;
                for (; c + 64 * (XU - 1) < n; c += 64 * XU) {
                    double a[XU];
#pragma unroll
                    for (int u = 0; u < XU; ++u) a[u] = row[c + 64 * u];
#pragma unroll
                    for (int u = 0; u < XU; ++u) {
                        const int tc = c + 64 * u - k - 1;
                        a[u] -= vr * wv[tc] + wr * vcur[tc];
                        row[c + 64 * u] = a[u];
                        s += a[u] * vnext[tc - 1];
                    }
                }
                if (c < n) {   // the last (partial) chunk: its loads in flight together
                    double a[XU];
#pragma unroll
                    for (int u = 0; u < XU; ++u) a[u] = c + 64 * u < n ? row[c + 64 * u] : 0.0;
#pragma unroll
                    for (int u = 0; u < XU; ++u) {
                        if (c + 64 * u < n) {
                            const int tc = c + 64 * u - k - 1;
                            a[u] -= vr * wv[tc] + wr * vcur[tc];
                            row[c + 64 * u] = a[u];
                            s += a[u] * vnext[tc - 1];
                        }
                    }
                }
                for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
                const double pn = tn * s;
                if (lane == 0) pb[nx][tr - 1] = pn;
                pacc += pn * vnext[tr - 1];
            }
            wg_partial(pacc, qb[nx]);
            alive = xsync(xg.bar, G, wg, &gen, &sh_ok);
            double* tmp = vcur; vcur = vnext; vnext = tmp;
            cur = nx;
            tk = tn;
        }
        if (!alive) {
            if (wg == 0 && tid == 0) state[job].sweeps = -1;
            continue;
        }
        // keep the grid together before the next job reuses the LDS vectors
        alive = xsync(xg.bar, G, wg, &gen, &sh_ok);
        if (wg == 0 && tid == 0) state[job].prof[0] = (double)(wall_clock64() - t0);
    }
}

// ------------------------------------------------------------------ blocked tridiagonalization
// k_xtri_blk: the reduction of k_xtri in panels of XNB reflectors (the
// two-sided rank-2 form of LAPACK dsytrd / dlatrd).  Within a panel that
// starts at reflector k0 the stored trailing block stays A_p, the matrix after
// reflector k0 - 1, and the current one is
// A_p - sum_{k0 <= j <= k} (v_j w_j^T + w_j v_j^T).  Step k:
//   - every workgroup forms w_k = p_k + K v_k, then the pivot row k+1 from
//     x'[c] (row k+1 of A_p less the panel's terms before reflector k, by
//     symmetry column k+1, formed by the row owners in the previous step's
//     row pass) and the rank-2 term of reflector k;
//   - the panel's dot products with the new reflector, w_j . v_{k+1} and
//     v_j . v_{k+1}, come from grid partials of v_j . x', v_j . v_k and
//     v_j . p_k (and the same for w_j) that the row owners also accumulated,
//     plus the two of reflector k from LDS;
//   - the row pass only reads A_p: (A v_{k+1})_r = (A_p v_{k+1})_r
//     - sum_j (v_j[r] (w_j . v_{k+1}) + w_j[r] (v_j . v_{k+1})), with one
//     panel vector per lane (lane l: v_l, lane XNB + l: w_l);
//   - after the panel's last reflector a pass over the trailing block applies
//     the panel's rank-2 XNB update on FP64 MFMA, then a grid barrier, then
//     the row pass against the new A_p.
// A column thus reads the trailing block (8 m^2 bytes) where k_xtri reads and
// writes it (16 m^2), plus one read and write per panel.  v_j are rows of AR_W
// (stored there for the back-transform anyway), w_j rows of the panel buffer
// in the job's workspace; the newest w stays in LDS.  Every workgroup makes the
// same fixed-order sums, so all hold bitwise the same reflectors.

// Raw buffer loads of the panel vectors: the element's byte offset within a
// row in a VGPR, the row's in a scalar register (no 64-bit address per row).
typedef unsigned int xu32;
__device__ inline __amdgpu_buffer_rsrc_t xrsrc(const double* base, long long ndoubles) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(base), (short)0, (int)(ndoubles * 8), 0x00020000);
}
__device__ inline double xbuf_load(__amdgpu_buffer_rsrc_t r, xu32 voff, xu32 soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

// Three block-wide sums at once (fixed order); every thread gets them in out[].
__device__ inline void block_sum3(double a, double b, double c, double* red3, double* out, int tid) {
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
        c += __shfl_xor(c, off);
    }
    if ((tid & 63) == 0) {
        red3[3 * (tid >> 6)] = a;
        red3[3 * (tid >> 6) + 1] = b;
        red3[3 * (tid >> 6) + 2] = c;
    }
    __syncthreads();
    if (tid < 3) {
        double s = 0.0;
        for (int w = 0; w < TW; ++w) s += red3[3 * w + tid];
        out[tid] = s;
    }
    __syncthreads();
}

// A[r][c] -= sum_{l < XNB} (V_l[r] W_l[c] + W_l[r] V_l[c]) for r, c in [r0, n):
// V_l = row k0 + l of Wm, W_l = row l of Wp (both by global index), the last
// W_l from LDS (wl[t] <-> index r0 - 1 + t).  A wavefront takes two 16 x 16
// tiles side by side, K = 2 XNB = 16 in four 16x16x4 FP64 MFMAs each: A operand row rb + lane % 16,
// k = lane / 16 of the step (V, then W); B operand column cb + lane % 16 (W,
// then V); the accumulator holds rows rb + 4 v + lane / 16.
__device__ inline void xtri_panel_update(double* __restrict__ A, const double* __restrict__ Wm,
                                         const double* __restrict__ Wp, const double* wl, int n, int k0, int r0,
                                         int gw, int GW, int lane) {
    static_assert(2 * XNB == 16, "four MFMA k-steps");
    const long long ln = n;
    const int nt = (n - r0 + 15) >> 4;
    const int lr = lane & 15, lk = lane >> 4;
    // two side-by-side tiles (I, 2 J2) and (I, 2 J2 + 1) per pass: 256-byte row
    // segments, the row operands loaded once for both, all loads in flight together
    const int nt2 = (nt + 1) >> 1, npairs = nt * nt2;
    for (int p = gw; p < npairs; p += GW) {
        const int I = p / nt2, J2 = p - I * nt2;
        const int rb = r0 + 16 * I, ra = rb + lr;
        int cs[2];
        dbl4 acc[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            cs[h] = r0 + 16 * (2 * J2 + h) + lr;
            const int c = cs[h];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = rb + 4 * v + lk;
                acc[h][v] = (r < n && c < n) ? A[(long long)r * ln + c] : 0.0;
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int kk = 4 * q + lk;
            double a = 0.0;
            if (q < 2) {   // V_kk (rows) against W_kk (columns)
                if (ra < n) a = Wm[(long long)(k0 + kk) * ln + ra];
            } else {       // W_l (rows) against V_l (columns)
                const int l = kk - XNB;
                if (ra < n) a = l == XNB - 1 ? wl[ra - r0 + 1] : Wp[(long long)l * ln + ra];
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int c = cs[h];
                double b = 0.0;
                if (c < n) {
                    if (q < 2) b = kk == XNB - 1 ? wl[c - r0 + 1] : Wp[(long long)kk * ln + c];
                    else b = Wm[(long long)(k0 + kk - XNB) * ln + c];
                }
                acc[h] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a, b, acc[h], 0, 0, 0);
            }
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int c = cs[h];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = rb + 4 * v + lk;
                if (r < n && c < n) A[(long long)r * ln + c] = acc[h][v];
            }
        }
    }
}

__global__ void __launch_bounds__(TT) k_xtri_blk(const JobDev* __restrict__ jobs, JobState* __restrict__ state,
                                                 const int* __restrict__ joblist, const long long* __restrict__ xloff,
                                                 int nlist, double* __restrict__ arena, double* __restrict__ tri,
                                                 double* __restrict__ xlws, unsigned* __restrict__ bar) {
    extern __shared__ double dyn[];
    __shared__ double red[TW], red2[TW], bc[2], red3[3 * TW], sum3[3];
    // v_j and w_j of the panel's earlier reflectors at indices k+1 and k+2
    __shared__ double s1v[XNB], s1w[XNB], s2v[XNB], s2w[XNB];
    // al[l] = w_l . v_{k+1}, be[l] = v_l . v_{k+1} over indices >= k+2 (l <= i)
    __shared__ double al[XNB], be[XNB];
    // grid sums of the panel partials (per panel lane L: P_L . x', P_L . v_k,
    // P_L . p_k at L, 2 XNB + L, 4 XNB + L) and their per-wavefront terms
    __shared__ double sS[XSR], redS[TW * XSR];
    __shared__ int sh_ok;
    const XGroup xg = x_group(nlist, bar);
    const int G = xg.G, wg = xg.wg, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int gw = wg * TW + wave, GW = G * TW;
    unsigned gen = 0;
    bool alive = true;
    for (int jl = xg.grp; jl < nlist; jl += xg.ngroups) {
        const int job = joblist[jl];
        if (state[job].phase != PH_ACTIVE) continue;   // read-only here: the same in every workgroup
        const unsigned long long t0 = wall_clock64();
        const JobDev jb = jobs[job];
        const int n = jb.nao;
        const long long ln = n;
        double* A = amat(arena, jb, AR_U);   // A_p, full and symmetric
        double* Wm = amat(arena, jb, AR_W);
        const double* F = amat(arena, jb, AR_F);
        double* d = eig_ws(tri, jb, job);
        double* e = d + n;
        double* tau = e + n;
        double* ws = xlws + xloff[jl];
        double* pb[2] = {ws + xw_p(n), ws + xw_p(n) + n};
        double* qb[2] = {ws + xw_q(n), ws + xw_q(n) + XGRID_MAX};
        double* xq[2] = {ws + xw_xq(n), ws + xw_xq(n) + n};
        double* xs[2] = {ws + xw_xs(n), ws + xw_xs(n) + (long long)XGRID_MAX * XSR};
        double* Wp = ws + xw_wp(n);
        double* vcur = dyn;
        double* vnext = dyn + n;
        double* wv = dyn + 2 * n;
        if (!alive) {
            if (wg == 0 && tid == 0) state[job].sweeps = -1;
            continue;
        }
        if (wg == 0)
            for (int i = tid; i < n; i += TT) { e[i] = 0.0; tau[i] = 0.0; }
        if (n < 3) {
            if (wg == 0 && tid == 0) {
                d[0] = F[0];
                if (n == 2) { e[0] = F[1]; d[1] = F[3]; }
                state[job].prof[0] = (double)(wall_clock64() - t0);
            }
            continue;
        }
        // reflector of row 0 from F, then A_p = F (rows 1 .. n-1), p = tau_0 A v_0
        // and x' = column 1 of F
        double tk;
        {
            double ss = 0.0;
            for (int t = 2 + tid; t < n; t += TT) { const double x = F[t]; ss += x * x; }
            ss = block_sum(ss, red, tid, TW);
            const double alpha = F[1];
            double beta = alpha, scale = 0.0;
            tk = 0.0;
            if (ss > 0.0) {
                beta = -copysign(sqrt(alpha * alpha + ss), alpha);
                tk = (beta - alpha) / beta;
                scale = 1.0 / (alpha - beta);
            }
            for (int t = 1 + tid; t < n - 1; t += TT) vcur[t] = ss > 0.0 ? F[1 + t] * scale : 0.0;
            if (tid == 0) vcur[0] = 1.0;
            __syncthreads();
            if (wg == 0) {
                for (int t = tid; t < n - 1; t += TT) Wm[1 + t] = vcur[t];
                if (tid == 0) { d[0] = F[0]; e[0] = beta; tau[0] = tk; }
            }
            double pv = 0.0;
            for (int r = 1 + (gw - 1 + GW) % GW; r < n; r += GW) {
                double s = 0.0;
                int c = 1 + lane;
                for (; c + 64 * 7 < n; c += 64 * 8) {
                    double a[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) a[u] = F[(long long)r * ln + c + 64 * u];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        A[(long long)r * ln + c + 64 * u] = a[u];
                        s += a[u] * vcur[c + 64 * u - 1];
                    }
                }
                if (c < n) {
                    double a[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) a[u] = c + 64 * u < n ? F[(long long)r * ln + c + 64 * u] : 0.0;
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        if (c + 64 * u < n) {
                            A[(long long)r * ln + c + 64 * u] = a[u];
                            s += a[u] * vcur[c + 64 * u - 1];
                        }
                    }
                }
                for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
                const double pn = tk * s;
                if (lane == 0) {
                    pb[0][r - 1] = pn;
                    xq[0][r] = F[(long long)r * ln + 1];
                }
                pv += pn * vcur[r - 1];
            }
            if (lane == 0) red2[wave] = pv;
            __syncthreads();
            if (tid == 0) {
                double s = 0.0;
                for (int w = 0; w < TW; ++w) s += red2[w];
                qb[0][wg] = s;
            }
        }
        alive = xsync(xg.bar, G, wg, &gen, &sh_ok);
        int cur = 0, k0 = 0;
        for (int k = 0; alive && k <= n - 3; ++k) {
            const int m = n - 1 - k;   // length of v_k (indices t <-> row / column k + 1 + t)
            const int i = k - k0;      // earlier reflectors of the panel: k0 .. k-1
            // the panel's v_l, w_l at indices k+1, k+2 (l < i), and the grid
            // sums of its dot partials (wavefront 2 + l: v_l, 9 + l: w_l),
            // loaded beside the partials of p . v
            if (tid >= 64 && tid < 64 + i) {
                const int l = tid - 64;
                const long long vrow = (long long)(k0 + l) * ln, wrow = (long long)l * ln;
                s1v[l] = Wm[vrow + k + 1];
                s1w[l] = Wp[wrow + k + 1];
                s2v[l] = Wm[vrow + k + 2];
                s2w[l] = Wp[wrow + k + 2];
            }
            if (wave >= 2) {
                const int l = wave < 9 ? wave - 2 : wave - 9;
                const int L = wave < 9 ? l : XNB + l;
                if (l < i) {
                    double a1 = 0.0, a2 = 0.0, a3 = 0.0, b1 = 0.0, b2 = 0.0, b3 = 0.0;
                    const double* q1 = xs[cur] + (long long)L * XGRID_MAX;
                    const double* q2 = xs[cur] + (long long)(2 * XNB + L) * XGRID_MAX;
                    const double* q3 = xs[cur] + (long long)(4 * XNB + L) * XGRID_MAX;
                    for (int g = lane; g < G; g += 128) {   // two partials of each in flight per lane
                        a1 += q1[g];
                        a2 += q2[g];
                        a3 += q3[g];
                        if (g + 64 < G) {
                            b1 += q1[g + 64];
                            b2 += q2[g + 64];
                            b3 += q3[g + 64];
                        }
                    }
                    a1 += b1;
                    a2 += b2;
                    a3 += b3;
                    for (int off = 32; off > 0; off >>= 1) {
                        a1 += __shfl_xor(a1, off);
                        a2 += __shfl_xor(a2, off);
                        a3 += __shfl_xor(a3, off);
                    }
                    if (lane == 0) {
                        sS[L] = a1;
                        sS[2 * XNB + L] = a2;
                        sS[4 * XNB + L] = a3;
                    }
                }
            }
            // this thread's p and x' elements (t = tid + TT u, m <= TT XP), loaded
            // before the wait for the partials so the loads overlap
            double pr[XP], xr[XP];
#pragma unroll
            for (int u = 0; u < XP; ++u) {
                const int t = tid + TT * u;
                pr[u] = t < m ? pb[cur][t] : 0.0;
                xr[u] = t < m ? xq[cur][k + 1 + t] : 0.0;
            }
            const double pv = grid_partials_sum(qb[cur], G, &bc[0]);   // its barrier publishes the above
            const double K = -0.5 * tk * pv;   // w = p + K v
#pragma unroll
            for (int u = 0; u < XP; ++u) {
                const int t = tid + TT * u;
                if (t < m) wv[t] = pr[u] + K * vcur[t];
            }
            __syncthreads();
            if (wg == 0)   // w_k for the later steps of the panel (read after the next grid barrier)
                for (int t = tid; t < m; t += TT) Wp[(long long)i * ln + k + 1 + t] = wv[t];
            if (k == n - 3) {   // last trailing 2 x 2 block (rows / columns n-2, n-1)
                if (wg == 0 && tid == 0) {
                    double d0 = A[(long long)(n - 2) * ln + n - 2];
                    double e0 = A[(long long)(n - 2) * ln + n - 1];
                    double d1 = A[(long long)(n - 1) * ln + n - 1];
                    for (int l = 0; l < i; ++l) {
                        d0 -= 2.0 * s1v[l] * s1w[l];
                        e0 -= s1v[l] * s2w[l] + s1w[l] * s2v[l];
                        d1 -= 2.0 * s2v[l] * s2w[l];
                    }
                    d[n - 2] = d0 - 2.0 * vcur[0] * wv[0];
                    e[n - 2] = e0 - (vcur[0] * wv[1] + wv[0] * vcur[1]);
                    d[n - 1] = d1 - 2.0 * vcur[1] * wv[1];
                }
                break;
            }
            // row k+1 of the current matrix: x' less reflector k's rank-2 term;
            // its tail norm and its dots with v_k and w_k past index k+2
            const double v0 = vcur[0], w0 = wv[0];
            double ss = 0.0, dv = 0.0, dw = 0.0;
#pragma unroll
            for (int u = 0; u < XP; ++u) {
                const int t = tid + TT * u;
                if (t < m) {
                    const double vt = vcur[t], wt = wv[t];
                    const double x = xr[u] - (v0 * wt + w0 * vt);
                    if (t >= 2) {
                        ss += x * x;
                        dv += vt * x;
                        dw += wt * x;
                    } else if (t == 1) {
                        bc[1] = x;
                    } else if (wg == 0) {
                        d[k + 1] = x;
                    }
                    if (t >= 1) vnext[t - 1] = x;   // scaled below
                }
            }
            block_sum3(ss, dv, dw, red3, sum3, tid);   // its barriers also publish bc[1] and vnext
            const double alpha = bc[1];
            double beta = alpha, tn = 0.0, scale = 0.0;
            if (sum3[0] > 0.0) {
                beta = -copysign(sqrt(alpha * alpha + sum3[0]), alpha);
                tn = (beta - alpha) / beta;
                scale = 1.0 / (alpha - beta);
            }
            const int m2 = m - 1;   // length of v_{k+1}
            for (int t = 1 + tid; t < m2; t += TT) vnext[t] = sum3[0] > 0.0 ? vnext[t] * scale : 0.0;
            if (tid == 0) vnext[0] = 1.0;
            if (tid < XNB) {
                // P . x = P . x' - v0 (P . p_k + K P . v_k) - w0 P . v_k over indices >= k+3
                const int l = tid;
                if (l < i) {
                    const double xv = sS[l] - v0 * (sS[4 * XNB + l] + K * sS[2 * XNB + l]) - w0 * sS[2 * XNB + l];
                    const double xw = sS[XNB + l] - v0 * (sS[5 * XNB + l] + K * sS[3 * XNB + l]) -
                                      w0 * sS[3 * XNB + l];
                    al[l] = s2w[l] + scale * xw;
                    be[l] = s2v[l] + scale * xv;
                } else if (l == i) {
                    al[l] = wv[1] + scale * sum3[2];
                    be[l] = vcur[1] + scale * sum3[1];
                }
            }
            __syncthreads();
            if (wg == 0) {
                double* wr = Wm + (long long)(k + 1) * ln + (k + 2);
                for (int t = tid; t < m2; t += TT) wr[t] = vnext[t];
                if (tid == 0) { e[k + 1] = beta; tau[k + 1] = tn; }
            }
            // the panel's last reflector: A_p <- A_p - V W^T - W V^T on the
            // trailing block, which every workgroup's row pass then reads
            const bool pend = i == XNB - 1;
            if (pend) {
                xtri_panel_update(A, Wm, Wp, wv, n, k0, k + 2, gw, GW, lane);
                alive = xsync(xg.bar, G, wg, &gen, &sh_ok);
                if (!alive) break;
            }
            // rows r = k+2 .. n-1: (A v_{k+1})_r from A_p and the panel, x'_r
            // for the next pivot row (A_p[r][k+2] less the panel's terms) and
            // this lane's panel vector's dot partials; XU loads in flight per lane
            const __amdgpu_buffer_rsrc_t Vr = xrsrc(Wm + (long long)k0 * ln, (long long)XNB * ln);
            const __amdgpu_buffer_rsrc_t Wr = xrsrc(Wp, (long long)XNB * ln);
            const xu32 nbytes = (xu32)n * 8u;
            const int pl = lane & (XNB - 1);
            const bool isv = lane < XNB;
            const bool plane = !pend && lane < 2 * XNB && pl <= i;
            // lane v_l: -v_l[r] (al[l], w_l[k+2]); lane w_l: -w_l[r] (be[l], v_l[k+2])
            double cs = 0.0, cx = 0.0;
            if (plane) {
                cs = isv ? al[pl] : be[pl];
                cx = isv ? (pl < i ? s2w[pl] : wv[1]) : (pl < i ? s2v[pl] : vcur[1]);
            }
            const int nx = cur ^ 1;
            double pacc = 0.0, S1 = 0.0, S2 = 0.0, S3 = 0.0;
            for (int r = k + 2 + (gw - (k + 2) % GW + GW) % GW; r < n; r += GW) {
                const int tr = r - k - 1;
                const double* row = A + (long long)r * ln;
                double P = 0.0;
                if (plane) {
                    if (pl == i) P = isv ? vcur[tr] : wv[tr];
                    else if (isv) P = xbuf_load(Vr, (xu32)r * 8u, (xu32)pl * nbytes);
                    else P = xbuf_load(Wr, (xu32)r * 8u, (xu32)pl * nbytes);
                }
                double s = -P * cs, xp = -P * cx;
                if (lane == 0) xp += row[k + 2];
                int c = k + 2 + lane;
                for (; c + 64 * (XU - 1) < n; c += 64 * XU) {
                    double a[XU];
#pragma unroll
                    for (int u = 0; u < XU; ++u) a[u] = row[c + 64 * u];
#pragma unroll
                    for (int u = 0; u < XU; ++u) s += a[u] * vnext[c + 64 * u - k - 2];
                }
                if (c < n) {
                    double a[XU];
#pragma unroll
                    for (int u = 0; u < XU; ++u) a[u] = c + 64 * u < n ? row[c + 64 * u] : 0.0;
#pragma unroll
                    for (int u = 0; u < XU; ++u)
                        if (c + 64 * u < n) s += a[u] * vnext[c + 64 * u - k - 2];
                }
                for (int off = 32; off > 0; off >>= 1) {
                    s += __shfl_xor(s, off);
                    xp += __shfl_xor(xp, off);
                }
                const double pn = tn * s;
                const double vr = vnext[tr - 1];
                if (lane == 0) {
                    pb[nx][tr - 1] = pn;
                    xq[nx][r] = xp;
                }
                pacc += pn * vr;
                if (r >= k + 4) {
                    S1 += P * xp;
                    S2 += P * vr;
                    S3 += P * pn;
                }
            }
            if (lane == 0) red2[wave] = pacc;
            if (lane < 2 * XNB) {
                redS[wave * XSR + lane] = S1;
                redS[wave * XSR + 2 * XNB + lane] = S2;
                redS[wave * XSR + 4 * XNB + lane] = S3;
            }
            __syncthreads();
            if (tid == 0) {
                double sp = 0.0;
                for (int w = 0; w < TW; ++w) sp += red2[w];
                qb[nx][wg] = sp;
            }
            if (!pend && tid < XSR) {
                double sp = 0.0;
                for (int w = 0; w < TW; ++w) sp += redS[w * XSR + tid];
                xs[nx][(long long)tid * XGRID_MAX + wg] = sp;
            }
            alive = xsync(xg.bar, G, wg, &gen, &sh_ok);
            double* tmp = vcur; vcur = vnext; vnext = tmp;
            cur = nx;
            tk = tn;
            if (pend) k0 = k + 1;
        }
        if (!alive) {
            if (wg == 0 && tid == 0) state[job].sweeps = -1;
            continue;
        }
        alive = xsync(xg.bar, G, wg, &gen, &sh_ok);
        if (wg == 0 && tid == 0) state[job].prof[0] = (double)(wall_clock64() - t0);
    }
}

// ------------------------------------------------------------------ divide and conquer
// Lane t's value of v, as a wave-uniform value (v_readlane on both halves).
__device__ inline double lane_value(double v, int t) {
    const long long x = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane((int)(x & 0xffffffffLL), t);
    const int hi = __builtin_amdgcn_readlane((int)(x >> 32), t);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}

// The one-workgroup k_dc (kernels_eig.hip) over the whole grid: gt / GT index
// the grid's threads, gwv / GWV its wavefronts, and every __syncthreads of
// k_dc between phases is a grid barrier.  Tables live in device memory.
__global__ void __launch_bounds__(XT) k_xdc(const JobDev* __restrict__ jobs, JobState* __restrict__ state,
                                            const int* __restrict__ joblist, const long long* __restrict__ xloff,
                                            int nlist, double* __restrict__ arena, const double* __restrict__ tri,
                                            double* __restrict__ eps, int* __restrict__ perm,
                                            double* __restrict__ xlws, unsigned* __restrict__ bar) {
    __shared__ int sh_ok;
    const XGroup xg = x_group(nlist, bar);
    const int G = xg.G, wg = xg.wg, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int gt = wg * XT + tid, GT = G * XT;
    const int gwv = wg * XW + wave, GWV = G * XW;
    unsigned gen = 0;
    bool alive = true;
    for (int jl = xg.grp; jl < nlist; jl += xg.ngroups) {
        const int job = joblist[jl];
        if (state[job].phase != PH_ACTIVE) continue;
        if (state[job].sweeps < 0) continue;   // the tridiagonalization gave up (set before this launch)
        const unsigned long long t0 = wall_clock64();
        const JobDev jb = jobs[job];
        const int n = jb.nao;
        const double ue = kEps;
        double* tb = xlws + xloff[jl] + xw_dc(n);
        double* Dc = tb;
        double* ds = Dc + n;
        double* zs = ds + n;
        double* dl = zs + n;
        double* zl = dl + n;
        double* vals = zl + n;
        double* rtau = vals + n;
        double* zhat = rtau + n;
        double* cnorm = zhat + n;
        double* rc = cnorm + n;
        double* rs = rc + n;
        double* rhoq = rs + n;
        int* col = (int*)(rhoq + n);
        int* flag = col + n;
        int* cl = flag + n;
        int* rorig = cl + n;
        int* rank = rorig + n;
        int* r1 = rank + n;
        int* r2 = r1 + n;
        int* kkq = r2 + n;
        int* nrot = kkq + n;
        const double* d = eig_ws(const_cast<double*>(tri), jb, job);
        const double* e = d + n;
        double* QC = amat(arena, jb, AR_C);
        double* QX = amat(arena, jb, AR_Q2);
        double* U = amat(arena, jb, AR_U);
        if (!alive) {
            if (wg == 0 && tid == 0) state[job].sweeps = -1;
            continue;
        }
        // a non-finite tridiagonal fails the solve (every workgroup checks all of it)
        int bad = 0;
        for (int i = tid; i < n; i += XT) bad |= !isfinite(d[i]) || (i < n - 1 && !isfinite(e[i]));
        if (__syncthreads_or(bad)) {
            for (int i = gt; i < n; i += GT) {
                eps[jb.vec0 + i] = 0.0;
                perm[jb.ivec0 + i] = i;
            }
            if (wg == 0 && tid == 0) state[job].sweeps = -1;
            continue;
        }
        int L = 0;
        while ((1 << L) < n) ++L;
        // lanes per secular root: the largest power of two up to 64 with n of
        // them in the group's threads
        int sg = 1;
        while (sg < 64 && (long long)n * sg * 2 <= GT) sg *= 2;
        {
            double* Q0 = (L & 1) ? QX : QC;
            for (int i = gt; i < n; i += GT) {
                const double el = i > 0 ? fabs(e[i - 1]) : 0.0, er = i < n - 1 ? fabs(e[i]) : 0.0;
                Dc[i] = d[i] - el - er;
                Q0[(long long)i * n + i] = 1.0;
            }
        }
        alive = xsync(xg.bar, G, wg, &gen, &sh_ok);
        for (int l = 1; alive && l <= L; ++l) {
            const int s = 1 << l, h = s >> 1, nm = (n + s - 1) >> l;
            double* QO = ((L - l + 1) & 1) ? QX : QC;
            double* QN = ((L - l) & 1) ? QX : QC;
            // A: z and the stable merge order of the two sorted blocks (sg lanes
            // per element split the count)
            for (int i = gt / sg; i < n; i += GT / sg) {
                const int sub = gt & (sg - 1);
                const int a = (i >> l) << l, m = min(a + h, n), b = min(a + s, n);
                const long long ri = (long long)i * n;
                if (m == b) {
                    for (int c = a + sub; c < b; c += sg) QN[ri + c] = QO[ri + c];
                    continue;
                }
                const double di = Dc[i];
                int pos = 0;
                for (int t = a + sub; t < b; t += sg) pos += dc_before(Dc[t], t, di, i);
                pos = group_allsum(pos, sg);
                if (sub == 0) {
                    double z = i < m ? QO[ri + m - 1] : QO[ri + m];
                    if (i >= m && e[m - 1] < 0.0) z = -z;
                    z *= 0.70710678118654752440;
                    ds[a + pos] = di;
                    zs[a + pos] = z;
                    col[a + pos] = i;
                }
            }
            // z comes from the diagonal blocks of QO; only its off-diagonal blocks are cleared here
            for (long long x = gt; x < ((long long)n << (l - 1)); x += GT) {
                const int i = (int)(x >> (l - 1)), c0 = (int)(x & (h - 1));
                const int a = (i >> l) << l, m = min(a + h, n), b = min(a + s, n);
                if (m == b) continue;
                const int c = i < m ? m + c0 : a + c0;
                if (c < (i < m ? b : m)) QO[(long long)i * n + c] = 0.0;
            }
            alive = xsync(xg.bar, G, wg, &gen, &sh_ok);
            if (!alive) break;
            // B: deflation, one wavefront per merge (the scan on lane 0)
            for (int q = gwv; q < nm; q += GWV) {
                const int a = q << l, m = min(a + h, n), b = min(a + s, n);
                if (m == b) {
                    if (lane == 0) kkq[q] = -1;
                    continue;
                }
                const double rho = 2.0 * fabs(e[m - 1]);
                double dmax = 0.0, zmax = 0.0;
                for (int jj = a + lane; jj < b; jj += 64) { dmax = fmax(dmax, fabs(ds[jj])); zmax = fmax(zmax, fabs(zs[jj])); }
                for (int off = 32; off > 0; off >>= 1) {
                    dmax = fmax(dmax, __shfl_xor(dmax, off));
                    zmax = fmax(zmax, __shfl_xor(zmax, off));
                }
                const double tol = 8.0 * ue * fmax(dmax, zmax);
                if (rho * zmax <= tol) {
                    for (int jj = a + lane; jj < b; jj += 64) flag[jj] = 1;
                    if (lane == 0) nrot[q] = 0;
                } else {
                    // the scan, wave-uniform: each chunk of 64 poles is loaded by
                    // the wavefront at once and read lane by lane (readlane), so
                    // no load waits behind the scan's own stores; lane 0 stores
                    int pj = -1, nr = 0, ps = 0, cp = 0;
                    double zp = 0.0, dp = 0.0;
                    for (int c0 = a; c0 < b; c0 += 64) {
                        const int jl = c0 + lane;
                        const double zc = jl < b ? zs[jl] : 0.0, dc = jl < b ? ds[jl] : 0.0;
                        const int cc = jl < b ? col[jl] : 0;
                        const int cnt = min(64, b - c0);
                        for (int t = 0; t < cnt; ++t) {
                            const int jj = c0 + t;
                            double zj = lane_value(zc, t), dj = lane_value(dc, t);
                            const int cj = __builtin_amdgcn_readlane(cc, t);
                            int js = cj < m ? 1 : 2;
                            if (rho * fabs(zj) <= tol) {
                                if (lane == 0) flag[jj] = 1;
                                continue;
                            }
                            int fj = 0;
                            if (pj >= 0) {
                                const double tt = dj - dp, zz = zj * zj + zp * zp;
                                if (fabs(tt * zj * zp) <= tol * zz) {
                                    const double tz = sqrt(zz);
                                    const double C = zj / tz, S = -zp / tz;
                                    const double dpn = dp * C * C + dj * S * S;
                                    dj = dp * S * S + dj * C * C;
                                    if (lane == 0) {
                                        zs[jj] = tz;
                                        zs[pj] = 0.0;
                                        r1[a + nr] = cp;
                                        r2[a + nr] = cj;
                                        rc[a + nr] = C;
                                        rs[a + nr] = S;
                                        ds[pj] = dpn;
                                        ds[jj] = dj;
                                        flag[pj] = 1;
                                    }
                                    ++nr;
                                    zj = tz;
                                    js |= ps;
                                    if (js == 3) fj = 2;
                                }
                            }
                            if (lane == 0) flag[jj] = fj;
                            pj = jj;
                            ps = js;
                            zp = zj;
                            dp = dj;
                            cp = cj;
                        }
                    }
                    if (lane == 0) nrot[q] = nr;
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
                __builtin_amdgcn_wave_barrier();
                int k = 0;
                for (int c0 = a; c0 < b; c0 += 64) {
                    const int jj = c0 + lane;
                    const int fj = jj < b ? flag[jj] : 1;
                    const bool keep = fj != 1;
                    const unsigned long long mk = __ballot(keep);
                    const int pre = __popcll(mk & ((1ull << lane) - 1ull));
                    if (keep) {
                        dl[a + k + pre] = ds[jj];
                        zl[a + k + pre] = zs[jj];
                        cl[a + k + pre] = col[jj];
                        rorig[a + k + pre] = fj;
                    }
                    k += __popcll(mk);
                }
                int t2 = k;
                for (int c0 = a; c0 < b; c0 += 64) {
                    const int jj = c0 + lane;
                    const bool dfl = jj < b && flag[jj] == 1;
                    const unsigned long long mk = __ballot(dfl);
                    const int pre = __popcll(mk & ((1ull << lane) - 1ull));
                    if (dfl) { dl[a + t2 + pre] = ds[jj]; zl[a + t2 + pre] = 0.0; cl[a + t2 + pre] = col[jj]; }
                    t2 += __popcll(mk);
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
                __builtin_amdgcn_wave_barrier();
                int kl = -1, ke = 0;
                if (l >= 5) {
                    int nl = 0, nr = 0;
                    for (int c0 = 0; c0 < k; c0 += 64) {
                        const int t = c0 + lane;
                        const int sd = t < k ? (rorig[a + t] == 2 ? 3 : (cl[a + t] < m ? 1 : 2)) : 0;
                        nl += __popcll(__ballot(sd & 1));
                        nr += __popcll(__ballot(sd & 2));
                    }
                    if (nl + nr <= b - a) {
                        kl = 0;
                        ke = nl;
                        for (int c0 = 0; c0 < k; c0 += 64) {
                            const int t = c0 + lane;
                            const int sd = t < k ? (rorig[a + t] == 2 ? 3 : (cl[a + t] < m ? 1 : 2)) : 0;
                            const unsigned long long ml = __ballot(sd & 1), mr = __ballot(sd & 2);
                            const unsigned long long below = (1ull << lane) - 1ull;
                            if (sd & 1) flag[a + kl + __popcll(ml & below)] = t;
                            if (sd & 2) flag[a + ke + __popcll(mr & below)] = t;
                            kl += __popcll(ml);
                            ke += __popcll(mr);
                        }
                    }
                }
                if (lane == 0) {
                    kkq[q] = k;
                    rhoq[q] = rho;
                    if (l >= 5) {
                        kkq[(n >> 1) + q] = kl;
                        kkq[(n >> 1) + (n >> 2) + q] = ke;
                    }
                }
            }
            alive = xsync(xg.bar, G, wg, &gen, &sh_ok);
            if (!alive) break;
            // C: deflation rotations on the rows of QO (thread = column)
            for (int i = gt; i < n; i += GT) {
                const int q = i >> l, a = q << l;
                const int kk = kkq[q];
                if (kk < 0) continue;
                const int nr = nrot[q];
                for (int t = 0; t < nr; ++t) {
                    const long long p1 = (long long)r1[a + t] * n + i, p2 = (long long)r2[a + t] * n + i;
                    const double x = QO[p1], y = QO[p2], C = rc[a + t], S = rs[a + t];
                    QO[p1] = C * x + S * y;
                    QO[p2] = C * y - S * x;
                }
            }
            // D: secular roots, sg aligned lanes per root splitting its pole sums,
            // so the roots of the top merges spread over the whole grid instead
            // of the first n / XT workgroups
            for (int i = gt / sg; i < n; i += GT / sg) {
                const int sub = gt & (sg - 1);
                const int q = i >> l, a = q << l;
                const int kk = kkq[q];
                if (kk < 0) continue;
                const int j = i - a;
                if (j < kk) {
                    int o;
                    double tau;
                    dc_secular(dl + a, zl + a, kk, rhoq[q], j, &o, &tau, ue, sub, sg);
                    if (sub == 0) {
                        rorig[i] = a + o;
                        rtau[i] = tau;
                        vals[i] = dl[a + o] + tau;
                    }
                } else if (sub == 0) {
                    vals[i] = dl[i];
                }
            }
            alive = xsync(xg.bar, G, wg, &gen, &sh_ok);
            if (!alive) break;
            // E: Loewner z-hat and the rank of every new eigenvalue in its block
            // (sg lanes per element split the product and the count)
            for (int i = gt / sg; i < n; i += GT / sg) {
                const int sub = gt & (sg - 1);
                const int q = i >> l, a = q << l;
                const int kk = kkq[q];
                if (kk < 0) continue;
                const int b = min(a + s, n), j = i - a;
                if (j < kk) {
                    const double di = dl[i];
                    double wv = sub == 0 ? (di - dl[rorig[i]]) - rtau[i] : 1.0;
                    for (int t = sub; t < kk; t += sg) {
                        if (t == j) continue;
                        const int ct = a + t;
                        wv *= ((di - dl[rorig[ct]]) - rtau[ct]) / (di - dl[ct]);
                    }
                    wv = group_allprod(wv, sg);
                    if (sub == 0) zhat[i] = copysign(sqrt(fmax(-wv, 0.0)), zl[i]);
                }
                const double v = vals[i];
                int rk = 0;
                for (int t = a + sub; t < b; t += sg) rk += dc_before(vals[t], t, v, i);
                rk = group_allsum(rk, sg);
                if (sub == 0) rank[i] = rk;
            }
            alive = xsync(xg.bar, G, wg, &gen, &sh_ok);
            if (!alive) break;
            // F: inverse norms of the columns of U (sg lanes per column)
            for (int i = gt / sg; i < n; i += GT / sg) {
                const int sub = gt & (sg - 1);
                const int q = i >> l, a = q << l;
                const int kk = kkq[q];
                if (i - a >= kk) continue;
                const double dO = dl[rorig[i]], tj = rtau[i];
                double s2 = 0.0;
                for (int t = sub; t < kk; t += sg) {
                    const double u = zhat[a + t] / ((dl[a + t] - dO) - tj);
                    s2 += u * u;
                }
                s2 = group_allsum(s2, sg);
                if (sub == 0) cnorm[i] = 1.0 / sqrt(s2);
            }
            alive = xsync(xg.bar, G, wg, &gen, &sh_ok);
            if (!alive) break;
            // G: U[t][j] = zhat_t / (dl_t - lambda_j), columns normalized
            const long long tot = (long long)n << l;
            for (long long x = gt; x < tot; x += GT) {
                const int i = (int)(x >> l), j = (int)(x & (s - 1));
                const int q = i >> l, a = q << l, kk = kkq[q];
                if (i - a >= kk || j >= kk) continue;
                const int cj = a + j;
                U[(long long)i * n + cj] = zhat[i] / ((dl[i] - dl[rorig[cj]]) - rtau[cj]) * cnorm[cj];
            }
            alive = xsync(xg.bar, G, wg, &gen, &sh_ok);
            if (!alive) break;
            // H: new eigenvector rows QN[a + rank] = U^T QO[cl]; deflated rows copied
            const bool small = l < 5;
            for (long long x = gt; x < tot; x += GT) {
                const int u = (int)(x >> l), rr = (int)(x & (s - 1));
                const int q = u >> l, a = q << l, kk = kkq[q];
                if (kk < 0) continue;
                const int b = min(a + s, n), r = a + rr, j = u - a;
                if (r >= b) continue;
                double v;
                if (j < kk) {
                    if (!small) continue;
                    v = 0.0;
                    for (int t = 0; t < kk; ++t) v += U[(long long)(a + t) * n + u] * QO[(long long)cl[a + t] * n + r];
                } else {
                    v = QO[(long long)cl[u] * n + r];
                }
                QN[(long long)(a + rank[u]) * n + r] = v;
            }
            if (!small) {
                const int tpm = s >> 5;
                const long long ntile = (long long)nm * tpm * tpm;
                const int lr = lane & 15, lk = lane >> 4;
                for (long long ti = gwv; ti < ntile; ti += GWV) {
                    const int q = (int)(ti / ((long long)tpm * tpm));
                    const int rem = (int)(ti - (long long)q * tpm * tpm);
                    const int jt = rem / tpm, rt = rem - jt * tpm;
                    const int kk = kkq[q];
                    const int a = q << l, b = min(a + s, n), m = min(a + (s >> 1), n);
                    const int j0 = jt * 32, r0 = a + rt * 32;
                    if (kk <= 0 || j0 >= kk || r0 >= b) continue;
                    const int kl = kkq[(n >> 1) + q], ke = kkq[(n >> 1) + (n >> 2) + q];
                    const bool lh = kl >= 0 && r0 + 32 <= m, rh = kl >= 0 && r0 >= m;
                    const int i0 = rh ? kl : 0, i1 = lh ? kl : (rh ? ke : kk);
                    dbl4 acc[2][2];
#pragma unroll
                    for (int x = 0; x < 2; ++x)
#pragma unroll
                        for (int y = 0; y < 2; ++y) acc[x][y] = dbl4{0.0, 0.0, 0.0, 0.0};
                    for (int k0 = i0; k0 < i1; k0 += 4) {
                        const int ii = k0 + lk;
                        const bool tv = ii < i1;
                        const int t = (lh || rh) ? (tv ? flag[a + ii] : 0) : ii;
                        const long long urow = (long long)(a + t) * n + a;
                        const long long qrow = tv ? (long long)cl[a + t] * n : 0;
                        double av[2], bv[2];
#pragma unroll
                        for (int x = 0; x < 2; ++x) {
                            const int j = j0 + 16 * x + lr;
                            av[x] = (tv && j < kk) ? U[urow + j] : 0.0;
                            const int r = r0 + 16 * x + lr;
                            bv[x] = (tv && r < b) ? QO[qrow + r] : 0.0;
                        }
#pragma unroll
                        for (int x = 0; x < 2; ++x)
#pragma unroll
                            for (int y = 0; y < 2; ++y)
                                acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[x], bv[y], acc[x][y], 0, 0, 0);
                    }
#pragma unroll
                    for (int x = 0; x < 2; ++x)
#pragma unroll
                        for (int y = 0; y < 2; ++y)
#pragma unroll
                            for (int v = 0; v < 4; ++v) {
                                const int j = j0 + 16 * x + 4 * v + lk, r = r0 + 16 * y + lr;
                                if (j < kk && r < b) QN[(long long)(a + rank[a + j]) * n + r] = acc[x][y][v];
                            }
                }
            }
            // the new eigenvalues in block order (vals / rank of this level are read only here)
            for (int i = gt; i < n; i += GT) {
                const int q = i >> l, a = q << l;
                if (kkq[q] >= 0) Dc[a + rank[i]] = vals[i];
            }
            alive = xsync(xg.bar, G, wg, &gen, &sh_ok);
        }
        if (!alive) {
            if (wg == 0 && tid == 0) state[job].sweeps = -1;
            continue;
        }
        for (int i = gt; i < n; i += GT) {
            eps[jb.vec0 + i] = Dc[i];
            perm[jb.ivec0 + i] = i;
        }
        if (wg == 0 && tid == 0) {
            state[job].sweeps = 1;
            state[job].warm = 0;
            state[job].prof[1] = (double)(wall_clock64() - t0);
            state[job].prof[2] = 0.0;
        }
        // the next job reuses nothing of this one's tables, but its phase A
        // writes QN rows of its own arena only: no barrier needed before it
    }
}

// ------------------------------------------------------------------ back-transform
// One workgroup (4 wavefronts) per (job, 16 eigenvector rows): Z <- Z H_{n-3}
// ... H_0 applied as compact-WY blocks of 16 reflectors from the last block
// down, Z <- Z - (Z V) T^T V^T.  Per block: Y = Z V (16 x 16, each wavefront a
// quarter of the columns, partials summed in LDS), Y2 = -Y T^T, then the
// column tiles Z[:, j] += Y2 V^T on MFMA, read and written in place.
// V[j][c] = W[k0 + c][j] for j > k0 + c (v[0] = 1 is stored).
constexpr int kXbt = 16;
// Two compact-WY blocks per pass over the workgroup's 16 rows: with block a
// (applied first) and block b = a - 1, Z B_a B_b = Z - Y_a T_a^T V_a^T -
// Y_b T_b^T V_b^T where Y_a = Z V_a and Y_b = Z V_b - Y_a T_a^T (V_a^T V_b), so
// one read pass forms Z V_a, Z V_b and V_a^T V_b, and one K = 32 update pass
// writes Z: half the row traffic of a pass per block.
__global__ void __launch_bounds__(XT) k_xbt(const JobDev* __restrict__ jobs, JobState* __restrict__ state,
                                            const int* __restrict__ joblist, double* __restrict__ arena,
                                            const double* __restrict__ tri) {
    __shared__ double Yp[3][XW][16 * 16];   // per-wavefront partials of Z V_a, Z V_b, V_a^T V_b
    __shared__ double Ya[16 * 16], Xb[16 * 16], Gs[16 * 16], Ta[16 * 16], Tb[16 * 16];
    __shared__ double Us[16 * 32];          // [Y2_a | Y2_b], row-major 16 x 32
    const int job = joblist[blockIdx.y];
    const JobDev jb = jobs[job];
    const JobState& st = state[job];
    if (st.phase != PH_ACTIVE || st.sweeps < 0) return;
    const int n = jb.nao, nref = n - 2, r0 = blockIdx.x * 16;
    const int nrow = jb.want_c ? n : jb.nocc;
    if (nref <= 0 || r0 >= nrow) return;
    const unsigned long long t0 = wall_clock64();
    const long long ln = n;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, li = lane & 15, lk = lane >> 4;
    double* Z = amat(arena, jb, AR_C);
    const double* Wm = amat(arena, jb, AR_W);
    const double* Tall = eig_ws(const_cast<double*>(tri), jb, job) + 3 * n;
    const int nblk = (nref + kXbt - 1) / kXbt;
    const bool rowok = r0 + li < nrow;
    const long long zrow = (long long)(r0 + li) * ln;
    for (int ba = nblk - 1; ba >= 0; ba -= 2) {
        const int bb = ba - 1;   // < 0: block a alone
        const bool pair = bb >= 0;
        const int ka = ba * kXbt, nba = min(kXbt, nref - ka);
        const int kb = pair ? bb * kXbt : ka;
        const int jlo = kb + 1;
        if (tid < 256) {
            Ta[tid] = Tall[(long long)ba * 256 + tid];
            Tb[tid] = pair ? Tall[(long long)bb * 256 + tid] : 0.0;
        }
        // read pass: Y_a = Z V_a, X_b = Z V_b (A = Z: i = row li, k = column;
        // B = V: k = column, j = reflector li) and G = V_a^T V_b (A = V_a^T, the
        // same register as V_a's B operand)
        dbl4 accA = dbl4{0.0, 0.0, 0.0, 0.0}, accB = accA, accG = accA;
        for (int j0 = jlo + 4 * wave; j0 < n; j0 += 4 * XW) {
            const int j = j0 + lk;
            const double z = (rowok && j < n) ? Z[zrow + j] : 0.0;
            const double va = (li < nba && j < n && j > ka + li) ? Wm[(long long)(ka + li) * ln + j] : 0.0;
            const double vb = (pair && j < n && j > kb + li) ? Wm[(long long)(kb + li) * ln + j] : 0.0;
            accA = __builtin_amdgcn_mfma_f64_16x16x4f64(z, va, accA, 0, 0, 0);
            accB = __builtin_amdgcn_mfma_f64_16x16x4f64(z, vb, accB, 0, 0, 0);
            accG = __builtin_amdgcn_mfma_f64_16x16x4f64(va, vb, accG, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            Yp[0][wave][(4 * v + lk) * 16 + li] = accA[v];
            Yp[1][wave][(4 * v + lk) * 16 + li] = accB[v];
            Yp[2][wave][(4 * v + lk) * 16 + li] = accG[v];
        }
        __syncthreads();
        const int r = tid >> 4, c = tid & 15;   // XT = 256: one (r, c) per thread
        {
            double y = 0.0, x = 0.0, g = 0.0;
#pragma unroll
            for (int w = 0; w < XW; ++w) {
                y += Yp[0][w][r * 16 + c];
                x += Yp[1][w][r * 16 + c];
                g += Yp[2][w][r * 16 + c];
            }
            Ya[r * 16 + c] = y;
            Xb[r * 16 + c] = x;
            Gs[r * 16 + c] = g;
        }
        __syncthreads();
        // Y2_a[r][c] = -sum_{c2 >= c} T_a[c][c2] Y_a[r][c2]
        double y2a = 0.0;
        for (int c2 = c; c2 < kXbt; ++c2) y2a -= Ta[c * 16 + c2] * Ya[r * 16 + c2];
        Us[r * 32 + c] = y2a;
        __syncthreads();
        // Y_b = X_b + Y2_a G, then Y2_b[r][c] = -sum_{c2 >= c} T_b[c][c2] Y_b[r][c2]
        {
            double yb = Xb[r * 16 + c];
            for (int c2 = 0; c2 < kXbt; ++c2) yb += Us[r * 32 + c2] * Gs[c2 * 16 + c];
            Ya[r * 16 + c] = yb;   // Y_a is no longer needed
        }
        __syncthreads();
        {
            double y2b = 0.0;
            for (int c2 = c; c2 < kXbt; ++c2) y2b -= Tb[c * 16 + c2] * Ya[r * 16 + c2];
            Us[r * 32 + 16 + c] = pair ? y2b : 0.0;
        }
        __syncthreads();
        // Z[r][j] += sum_c U[r][c] [V_a | V_b][j][c] over the column tiles j >= jlo
        const int tlo = jlo >> 4, ntj = (n + 15) >> 4;
        for (int tj = tlo + wave; tj < ntj; tj += XW) {
            dbl4 u = dbl4{0.0, 0.0, 0.0, 0.0};
            const int j = 16 * tj + li;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int cc = 4 * q + lk;
                const double aa = Us[li * 32 + cc];   // A[i = row li][k = cc]
                double bv = 0.0;
                if (q < 4) {
                    if (cc < nba && j < n && j > ka + cc) bv = Wm[(long long)(ka + cc) * ln + j];
                } else {
                    const int cb = cc - 16;
                    if (pair && j < n && j > kb + cb) bv = Wm[(long long)(kb + cb) * ln + j];
                }
                u = __builtin_amdgcn_mfma_f64_16x16x4f64(aa, bv, u, 0, 0, 0);
            }
            // D[i][j] sits in lane j + 16 (i % 4), element i / 4
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int rr = r0 + 4 * v + lk;
                if (rr < nrow && j < n && j >= jlo) {
                    double* zp = Z + (long long)rr * ln + j;
                    *zp += u[v];
                }
            }
        }
        __syncthreads();
    }
    if (tid == 0 && blockIdx.x == 0) state[job].prof[3] = (double)(wall_clock64() - t0);
}

__global__ void k_xbar_reset(unsigned* bar) {
    for (int i = threadIdx.x; i < XBAR_WORDS * XMAX_GROUPS; i += blockDim.x) bar[i] = 0u;
}

}  // namespace

bool xlh_supported(int n) { return n <= kXlhMaxN; }

size_t xlh_bar_bytes() { return (size_t)XBAR_WORDS * XMAX_GROUPS * sizeof(unsigned); }

long long xlh_ws_doubles(int n) { return (xw_size(n) + 31) / 32 * 32; }

int xlh_grid(int num_cus) { return std::max(1, std::min(num_cus, XGRID_MAX)); }

// Measured on MI355X (256 CUs; ms per cold tridiagonalization of one molecule
// at 32 / 64 / 128 / 256 workgroups): n = 1201: 15.3 / 13.2 / 14.0 / 18.0;
// n = 1536: 23.9 / 20.2 / 19.4 / 24.2; n = 3000: 105 / 73 / 62 / 64;
// n = 6000: 676 / 409 / 296 / 291.  More workgroups stream the rows faster
// while the per-step grid barrier costs more the more workgroups take part.
// Several molecules of a launch share the grid, each with its own group.
int xlh_tri_grid(int num_cus, int max_n, int nlist) {
    const int forced = get_tuning().xl_tri_wg;
    const int per = max_n > 4096 ? 256 : (max_n > 1400 ? 128 : (nlist > 1 ? 32 : 64));
    const int g = forced > 0 ? forced : per * std::max(1, nlist);
    return std::max(1, std::min(g, std::min(num_cus, XGRID_MAX)));
}

void launch_xlh_eigen(hipStream_t st, int grid, const JobDev* jobs, JobState* state, const int* joblist,
                      const long long* xloff, int nlist, int max_n, double* arena, double* eps, int* perm,
                      double* tri, double* xlws, unsigned* bar) {
    if (nlist <= 0) return;
    const size_t lt = (size_t)3 * max_n * sizeof(double);
    const int mode = get_tuning().xl_tri_blocked;
    const bool blocked = mode == 2 || (mode == 1 && max_n >= kXtriBlkMinN);
    const void* ktri = blocked ? (const void*)k_xtri_blk : (const void*)k_xtri;
    raise_dynamic_lds(ktri, lt);
    hipLaunchKernelGGL(k_xbar_reset, dim3(1), dim3(256), 0, st, bar);
    if (blocked)
        hipLaunchKernelGGL(k_xtri_blk, dim3(xlh_tri_grid(grid, max_n, nlist)), dim3(TT), lt, st, jobs, state, joblist,
                           xloff, nlist, arena, tri, xlws, bar);
    else
        hipLaunchKernelGGL(k_xtri, dim3(xlh_tri_grid(grid, max_n, nlist)), dim3(TT), lt, st, jobs, state, joblist,
                           xloff, nlist, arena, tri, xlws, bar);
    hipLaunchKernelGGL(k_xbar_reset, dim3(1), dim3(256), 0, st, bar);
    hipLaunchKernelGGL(k_xdc, dim3(grid), dim3(XT), 0, st, jobs, state, joblist, xloff, nlist, arena, tri, eps, perm,
                       xlws, bar);
    launch_wy_t(st, jobs, state, joblist, nlist, max_n, arena, tri);
    hipLaunchKernelGGL(k_xbt, dim3((max_n + 15) / 16, nlist), dim3(XT), 0, st, jobs, state, joblist, arena, tri);
}

}  // namespace nddo
