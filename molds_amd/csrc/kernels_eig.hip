// FP64 symmetric eigensolver of the SCF loop (replaces LAPACKE_dsyevd,
// src/wrappers/Lapack.cpp:102-161) as three batched kernels, one workgroup per
// molecule each, for matrices up to 1200 AOs:
//   1. k_tridiag        Householder tridiagonalization T = Q^T F Q (Golub & Van
//                       Loan, Alg. 8.3.1) with the matrix in LDS (n <= 137);
//                       each step's rank-2 update is fused with the next step's
//                       matrix-vector product, so the trailing block is read
//                       and written once per step.
//      k_tridiag_sym    the same for matrices in device memory, on the upper
//                       triangle only.
//   2. k_dc             divide and conquer on T (Cuppen, Numer. Math. 36, 177
//                       (1981)), bottom-up from 1x1 leaves, all merges of a
//                       level at once: deflation as in Dongarra & Sorensen
//                       (SIAM J. Sci. Stat. Comput. 8, s139 (1987)), secular
//                       roots by a two-pole rational model with a bisection
//                       safeguard, eigenvector components by the Loewner
//                       formula of Gu & Eisenstat (SIAM J. Matrix Anal. Appl.
//                       16, 172 (1995)), and the merge products Q_new = Q U on
//                       FP64 MFMA (v_mfma_f64_16x16x4f64).
//   3. k_wy_t, k_bt_wy  eigenvector rows <- rows * H_{n-3} ... H_0 in compact
//                       WY blocks on FP64 MFMA, one workgroup per 64 rows.
// Output: rows of AR_C are the eigenvectors, already in ascending eigenvalue
// order; eps[] ascending and perm[] = identity, as the density kernel reads them.
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>
#include <map>
#include <mutex>

#include "eig_device.h"
#include "launch.h"
#include "nddo_device.h"

namespace nddo {

// Threads of k_tridiag for molecules of at most max_n AOs.
static int tridiag_threads(int max_n) { return max_n <= 44 ? 256 : (max_n <= 96 ? 512 : 1024); }

size_t tridiag_lds_bytes(int n, bool a_in_lds) {
    size_t d = 4 * (size_t)n;   // vcur vnext p pnext
    // the matrix and the segment partials of A v (at most max(threads, n))
    if (a_in_lds) d += (size_t)n * n + (size_t)std::max(tridiag_threads(n), n);
    return d * 8;
}

size_t dc_lds_bytes(int n) { return (size_t)n * (12 * 8 + 9 * 4); }

// Largest matrix of the one-workgroup path: the divide-and-conquer kernel keeps
// 132 bytes of LDS per row (158 KB at 1200 AOs).
constexpr int kHeigMaxN = 1200;

// ------------------------------------------------------------------ 1. tridiagonalize
// Reflector k acts on indices k+1..n-1: H_k = I - tau_k v v^T with v[0] = 1,
// kept in row k of W (columns k+1..n-1); d, e, tau go to the job's workspace.
// Up to 64 VGPRs (8 wavefronts per SIMD), so that eight 256-thread or four
// 512-thread workgroups of small molecules share a CU.
__global__ void __launch_bounds__(1024, 8) k_tridiag(const JobDev* __restrict__ jobs, JobState* __restrict__ state,
                                                     const int* __restrict__ joblist, double* __restrict__ arena,
                                                     double* __restrict__ tri) {
    extern __shared__ double dyn[];
    __shared__ double red[16], red2[16], sh_alpha;
    const int job = joblist[blockIdx.x];
    const JobDev jb = jobs[job];
    if (state[job].phase != PH_ACTIVE) return;
    const unsigned long long t0 = wall_clock64();
    const int n = jb.nao;
    const int T = blockDim.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = T >> 6;
    double* vcur = dyn;
    double* vnext = vcur + n;
    double* p = vnext + n;
    double* pnext = p + n;
    double* W = amat(arena, jb, AR_W);
    double* A = pnext + n;
    // segment partials of A v after A (S * m2 <= max(blockDim.x, m2) doubles)
    double* part = A + (long long)n * n;
    const double* F = amat(arena, jb, AR_F);
    double* d = eig_ws(tri, jb, job);
    double* e = d + n;
    double* tau = e + n;
    const long long nn = (long long)n * n;
    for (long long x = tid; x < nn; x += T) A[x] = F[x];
    for (int i = tid; i < n; i += T) { e[i] = 0.0; tau[i] = 0.0; }
    __syncthreads();
    if (n < 3) {
        if (tid == 0) {
            d[0] = A[0];
            if (n == 2) { e[0] = A[1]; d[1] = A[3]; }
        }
        if (tid == 0) state[job].prof[0] = (double)(wall_clock64() - t0);
        return;
    }
    // Each step k keeps v_k (vcur, indices k+1..n-1), p = tau_k A v_k (p), and
    // per-wave partial sums of p.v in red; four barriers per step.
    double tk;
    {   // reflector of row 0, then p = tau_0 A[1:, 1:] v_0
        double ss = 0.0;
        for (int t = 2 + tid; t < n; t += T) { const double x = A[t]; ss += x * x; }
        for (int off = 32; off > 0; off >>= 1) ss += __shfl_xor(ss, off);
        if (lane == 0) red2[wave] = ss;
        __syncthreads();
        ss = 0.0;
        for (int w = 0; w < nw; ++w) ss += red2[w];
        const double alpha = A[1];
        double beta = alpha, scale = 0.0;
        tk = 0.0;
        if (ss > 0.0) {
            beta = -copysign(sqrt(alpha * alpha + ss), alpha);
            tk = (beta - alpha) / beta;
            scale = 1.0 / (alpha - beta);
        }
        for (int t = 1 + tid; t < n - 1; t += T) vcur[t] = ss > 0.0 ? A[1 + t] * scale : 0.0;
        if (tid == 0) { vcur[0] = 1.0; d[0] = A[0]; e[0] = beta; tau[0] = tk; }
        __syncthreads();
        for (int t = 1 + tid; t < n - 1; t += T) A[1 + t] = vcur[t];
        if (tid == 0) A[1] = 1.0;
        double pv = 0.0;
        for (int i = 1 + wave; i < n; i += nw) {
            double s = 0.0;
            for (int j = 1 + lane; j < n; j += 64) s += A[(long long)i * n + j] * vcur[j - 1];
            for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
            if (lane == 0) { p[i - 1] = tk * s; pv += tk * s * vcur[i - 1]; }
        }
        if (lane == 0) red[wave] = pv;
        __syncthreads();
    }
    for (int k = 0; k <= n - 3; ++k) {
        const int m = n - 1 - k;   // length of v_k
        double pv = 0.0;
        for (int w = 0; w < nw; ++w) pv += red[w];
        const double K = -0.5 * tk * pv;   // w = p + K v
        const double w0 = p[0] + K * vcur[0];
        if (k == n - 3) {   // last trailing 2x2 block (rows/cols n-2, n-1)
            if (tid == 0) {
                const double w1 = p[1] + K * vcur[1];
                d[n - 2] = A[(long long)(n - 2) * n + n - 2] - 2.0 * vcur[0] * w0;
                e[n - 2] = A[(long long)(n - 2) * n + n - 1] - (vcur[0] * w1 + w0 * vcur[1]);
                d[n - 1] = A[(long long)(n - 1) * n + n - 1] - 2.0 * vcur[1] * w1;
            }
            break;
        }
        // row k+1 of the updated block; its tail x = A[k+1][k+2 ..] defines the next reflector
        const long long r1 = (long long)(k + 1) * n + k + 1;
        double ss = 0.0;
        for (int t = tid; t < m; t += T) {
            const double a = A[r1 + t] - (vcur[0] * (p[t] + K * vcur[t]) + w0 * vcur[t]);
            A[r1 + t] = a;
            if (t >= 2) ss += a * a;
            else if (t == 1) sh_alpha = a;
            else d[k + 1] = a;
        }
        for (int off = 32; off > 0; off >>= 1) ss += __shfl_xor(ss, off);
        if (lane == 0) red2[wave] = ss;
        __syncthreads();   // 1
        ss = 0.0;
        for (int w = 0; w < nw; ++w) ss += red2[w];
        const double alpha = sh_alpha;
        double beta = alpha, tn = 0.0, scale = 0.0;
        if (ss > 0.0) {
            beta = -copysign(sqrt(alpha * alpha + ss), alpha);
            tn = (beta - alpha) / beta;
            scale = 1.0 / (alpha - beta);
        }
        const int m2 = m - 1;   // length of v_{k+1}
        for (int t = 1 + tid; t < m2; t += T) {
            const double v = ss > 0.0 ? A[r1 + 1 + t] * scale : 0.0;
            vnext[t] = v;
            A[r1 + 1 + t] = v;
        }
        if (tid == 0) { vnext[0] = 1.0; A[r1 + 1] = 1.0; e[k + 1] = beta; tau[k + 1] = tn; }
        __syncthreads();   // 2
        // columns i = k+2 .. n-1 (A is kept symmetric, so column i is row i): rank-2
        // update of step k and the partial sums of (A v_{k+1})_i over S segments of j
        {
            // S segments of j per column; past T columns every thread takes whole
            // columns, several of them
            const int S = max(1, min(16, T / m2)), L = (m2 + S - 1) / S;
            for (int cc = tid; cc < S * m2; cc += T) {
                const int seg = cc / m2, ci = cc - seg * m2;
                const int i = k + 2 + ci, ti = ci + 1;
                const double vi = vcur[ti], wi = p[ti] + K * vi;
                const int jb = k + 2 + seg * L, je = min(jb + L, n);
                double s = 0.0;
                int j = jb;
                for (; j + 3 < je; j += 4) {
                    double a[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) a[u] = A[(long long)(j + u) * n + i];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int tj = j + u - k - 1;
                        const double vj = vcur[tj];
                        a[u] -= vj * wi + (p[tj] + K * vj) * vi;
                        A[(long long)(j + u) * n + i] = a[u];
                        s += a[u] * vnext[tj - 1];
                    }
                }
                for (; j < je; ++j) {
                    const int tj = j - k - 1;
                    const double vj = vcur[tj];
                    const double a = A[(long long)j * n + i] - (vj * wi + (p[tj] + K * vj) * vi);
                    A[(long long)j * n + i] = a;
                    s += a * vnext[tj - 1];
                }
                part[seg * m2 + ci] = s;
            }
            __syncthreads();   // 3
            pv = 0.0;
            for (int t = tid; t < m2; t += T) {
                double sp = 0.0;
                for (int sg = 0; sg < S; ++sg) sp += part[sg * m2 + t];
                const double pn = tn * sp;
                pnext[t] = pn;
                pv += pn * vnext[t];
            }
            for (int off = 32; off > 0; off >>= 1) pv += __shfl_xor(pv, off);
            if (lane == 0) red[wave] = pv;
        }
        __syncthreads();   // 4
        double* tmp = vcur; vcur = vnext; vnext = tmp;
        tmp = p; p = pnext; pnext = tmp;
        tk = tn;
    }
    // keep the reflectors (upper rows) for the back-transformation
    __syncthreads();
    for (long long x = tid; x < nn; x += T) W[x] = A[x];
    if (tid == 0) state[job].prof[0] = (double)(wall_clock64() - t0);
}

// Upper-triangle variant for matrices in device memory (n > 137): the same
// reflectors and fused rank-2 update / matrix-vector product as k_tridiag, but
// only the elements r <= c of the trailing block are read and written, which
// halves the HBM traffic of a step. A stored element a_rc (r < c) adds a_rc v_r
// to (A v)_c and a_rc v_c to (A v)_r. Work items are bands of BR rows by 64
// columns; a lane owns one column, so its column sum stays in a register and
// the row sums of each 4-row slice are reduced across the wave by a butterfly.
// Per-item partial sums go to colpart[band][t] and rowpart[tile][t] (t = index
// in the trailing block), in LDS when PG is false and in the job's AR_U matrix
// (unused until divide and conquer) when it is true, and are added in a fixed
// order, so the result does not depend on timing.
constexpr int kSymMaxLdsN = 512;   // largest n whose partial sums fit in LDS

__device__ inline int sym_col_cap(int n) { return ((n + 31) >> 5) * n; }

size_t tridiag_sym_lds_bytes(int n, bool pg) {
    size_t d = 4 * (size_t)n;
    if (!pg) d += (size_t)(((n + 31) >> 5) + ((n + 63) >> 6)) * n;
    return d * 8;
}

// Cross-lane exchanges on doubles without LDS: v_permlane32_swap trades the
// upper 32 lanes of its first operand for the lower 32 lanes of its second,
// v_permlane16_swap the odd 16-lane rows of the first for the even rows of the
// second; DPP row_ror:8 is lane ^ 8 within a row, row_half_mirror pairs lanes
// i and 7 - i of each 8, and the two quad permutations are lane ^ 2 and ^ 1.
typedef unsigned int u32;
__device__ inline void swap32_d(double& a, double& b) {
    const unsigned long long x = __double_as_longlong(a), y = __double_as_longlong(b);
    const auto lo = __builtin_amdgcn_permlane32_swap((u32)x, (u32)y, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((u32)(x >> 32), (u32)(y >> 32), false, false);
    a = __longlong_as_double(((unsigned long long)hi[0] << 32) | lo[0]);
    b = __longlong_as_double(((unsigned long long)hi[1] << 32) | lo[1]);
}
__device__ inline void swap16_d(double& a, double& b) {
    const unsigned long long x = __double_as_longlong(a), y = __double_as_longlong(b);
    const auto lo = __builtin_amdgcn_permlane16_swap((u32)x, (u32)y, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((u32)(x >> 32), (u32)(y >> 32), false, false);
    a = __longlong_as_double(((unsigned long long)hi[0] << 32) | lo[0]);
    b = __longlong_as_double(((unsigned long long)hi[1] << 32) | lo[1]);
}
template <int CTRL>
__device__ inline double dpp_d(double v) {
    const unsigned long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(u32)x, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(u32)(x >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((unsigned long long)(u32)hi << 32) | (u32)lo);
}
constexpr int kDppRor8 = 0x128, kDppHalfMirror = 0x141, kDppXor2 = 0x4E, kDppXor1 = 0xB1;

// Sum over aligned groups of G lanes (G = 1, 2, 4, ..., 64, wave-uniform);
// every lane of a group gets the group's sum.
__device__ inline double group_sum(double v, int G) {
    if (G >= 64) { double a = v, b = v; swap32_d(a, b); v = a + b; }
    if (G >= 32) { double a = v, b = v; swap16_d(a, b); v = a + b; }
    if (G >= 16) v += dpp_d<kDppRor8>(v);
    if (G >= 8) v += dpp_d<kDppHalfMirror>(v);
    if (G >= 4) v += dpp_d<kDppXor2>(v);
    if (G >= 2) v += dpp_d<kDppXor1>(v);
    return v;
}

// x0..x3 hold one value per row of a 4-row slice; afterwards every lane holds
// the sum over the 64 lanes of row (lane >> 4) & 3.
__device__ inline double rows4_reduce(double x0, double x1, double x2, double x3) {
    // lanes 0-31 keep rows 0-1 and lanes 32-63 rows 2-3
    swap32_d(x0, x2); swap32_d(x1, x3);
    double y0 = x0 + x2, y1 = x1 + x3;
    // even 16-lane rows keep y0 and odd rows y1
    swap16_d(y0, y1);
    double s = y0 + y1;
    s += dpp_d<kDppRor8>(s);
    s += dpp_d<kDppHalfMirror>(s);
    s += dpp_d<kDppXor2>(s);
    s += dpp_d<kDppXor1>(s);
    return s;
}

// Raw buffer access to one molecule's matrix (base, n*n*8 bytes, the gfx9
// raw-buffer descriptor word 0x00020000): a lane supplies a 32-bit byte
// offset and the row offset comes from a scalar register, so a slice of rows
// costs one VGPR of addressing. Every access issued stays inside the matrix.
typedef u32 u32x2 __attribute__((ext_vector_type(2)));
__device__ inline double buf_load_d(__amdgpu_buffer_rsrc_t r, u32 voff, u32 soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ inline void buf_store_d(double v, __amdgpu_buffer_rsrc_t r, u32 voff, u32 soff) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 0);
}

// Storage of the trailing matrix: FP64, or FP32 for mixed-precision jobs in
// their FP32 iterations (half the HBM bytes; every product and sum in FP64).
template <typename TS>
__device__ inline double sym_load(__amdgpu_buffer_rsrc_t r, u32 voff, u32 soff);
template <>
__device__ inline double sym_load<double>(__amdgpu_buffer_rsrc_t r, u32 voff, u32 soff) {
    return buf_load_d(r, voff, soff);
}
template <>
__device__ inline double sym_load<float>(__amdgpu_buffer_rsrc_t r, u32 voff, u32 soff) {
    return (double)__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <typename TS>
__device__ inline void sym_store(double v, __amdgpu_buffer_rsrc_t r, u32 voff, u32 soff);
template <>
__device__ inline void sym_store<double>(double v, __amdgpu_buffer_rsrc_t r, u32 voff, u32 soff) {
    buf_store_d(v, r, voff, soff);
}
template <>
__device__ inline void sym_store<float>(double v, __amdgpu_buffer_rsrc_t r, u32 voff, u32 soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(u32, (float)v), r, voff, soff, 0);
}

// The reduction of one job. As: the trailing matrix (upper triangle) in storage
// TS; W: the reflector rows (row k holds v_k in columns k+1..n-1).
template <bool PG, typename TS>
__device__ void tridiag_sym_body(TS* __restrict__ As, double* __restrict__ W, const double* __restrict__ F,
                                 double* colpart, double* dyn, int n, int job, double* d, double* e,
                                 double* tau, double* red, double* red2) {
    const int T = blockDim.x, tid = threadIdx.x, lane = tid & 63, nw = T >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    constexpr u32 ES = sizeof(TS);
    double* vcur = dyn;
    double* vnext = vcur + n;
    double* p = vnext + n;
    double* pnext = p + n;
    // the pivot row between its update and its reflector (free between the
    // reduction of one step and the band pass of the next)
    double* xrow = colpart;
    const __amdgpu_buffer_rsrc_t Ab = __builtin_amdgcn_make_buffer_rsrc(As, (short)0, n * n * (int)ES, 0x00020000);
    const int ccap = sym_col_cap(n);
    for (int r = wave; r < n; r += nw)
        for (int c = r + lane; c < n; c += 64) As[(long long)r * n + c] = (TS)F[(long long)r * n + c];
    __syncthreads();
    double tk;
    {   // reflector of row 0, then p = tau_0 F[1:, 1:] v_0 from the full (symmetric) rows of F
        double ss = 0.0;
        for (int t = 2 + tid; t < n; t += T) { const double x = F[t]; ss += x * x; }
        ss = group_sum(ss, 64);
        if (lane == 0) red2[wave] = ss;
        __syncthreads();
        ss = 0.0;
        for (int w = 0; w < nw; ++w) ss += red2[w];
        const double alpha = F[1];
        double beta = alpha, scale = 0.0;
        tk = 0.0;
        if (ss > 0.0) {
            beta = -copysign(sqrt(alpha * alpha + ss), alpha);
            tk = (beta - alpha) / beta;
            scale = 1.0 / (alpha - beta);
        }
        for (int t = 1 + tid; t < n - 1; t += T) vcur[t] = ss > 0.0 ? F[1 + t] * scale : 0.0;
        if (tid == 0) { vcur[0] = 1.0; d[0] = F[0]; e[0] = beta; tau[0] = tk; }
        __syncthreads();
        for (int t = 1 + tid; t < n - 1; t += T) W[1 + t] = vcur[t];
        if (tid == 0) W[1] = 1.0;
        double pv = 0.0;
        for (int i = 1 + wave; i < n; i += nw) {
            double s = 0.0;
            for (int j = 1 + lane; j < n; j += 64) s += F[(long long)i * n + j] * vcur[j - 1];
            s = group_sum(s, 64);
            if (lane == 0) { p[i - 1] = tk * s; pv += tk * s * vcur[i - 1]; }
        }
        if (lane == 0) red[wave] = pv;
        __syncthreads();
    }
    for (int k = 0; k <= n - 3; ++k) {
        const int m = n - 1 - k;   // length of v_k
        double pv = 0.0;
        for (int w = 0; w < nw; ++w) pv += red[w];
        const double K = -0.5 * tk * pv;   // w = p + K v
        const double w0 = p[0] + K * vcur[0];
        if (k == n - 3) {   // last trailing 2x2 block (rows/cols n-2, n-1)
            if (tid == 0) {
                const double w1 = p[1] + K * vcur[1];
                d[n - 2] = (double)As[(long long)(n - 2) * n + n - 2] - 2.0 * vcur[0] * w0;
                e[n - 2] = (double)As[(long long)(n - 2) * n + n - 1] - (vcur[0] * w1 + w0 * vcur[1]);
                d[n - 1] = (double)As[(long long)(n - 1) * n + n - 1] - 2.0 * vcur[1] * w1;
            }
            break;
        }
        // row k+1 of the updated block (upper part) into xrow; its tail defines the next reflector
        const long long r1 = (long long)(k + 1) * n + k + 1;
        double ss = 0.0;
        for (int t = tid; t < m; t += T) {
            const double a = (double)As[r1 + t] - (vcur[0] * (p[t] + K * vcur[t]) + w0 * vcur[t]);
            xrow[t] = a;
            if (t >= 2) ss += a * a;
        }
        ss = group_sum(ss, 64);
        if (lane == 0) red2[wave] = ss;
        __syncthreads();   // 1
        ss = 0.0;
        for (int w = 0; w < nw; ++w) ss += red2[w];
        const double alpha = xrow[1];
        double beta = alpha, tn = 0.0, scale = 0.0;
        if (ss > 0.0) {
            beta = -copysign(sqrt(alpha * alpha + ss), alpha);
            tn = (beta - alpha) / beta;
            scale = 1.0 / (alpha - beta);
        }
        const int m2 = m - 1;   // length of v_{k+1}
        for (int t = 1 + tid; t < m2; t += T) {
            const double v = ss > 0.0 ? xrow[1 + t] * scale : 0.0;
            vnext[t] = v;
            W[r1 + 1 + t] = v;
        }
        if (tid == 0) { vnext[0] = 1.0; W[r1 + 1] = 1.0; d[k + 1] = xrow[0]; e[k + 1] = beta; tau[k + 1] = tn; }
        __syncthreads();   // 2
        // upper triangle of rows/columns base .. n-1: rank-2 update of step k and
        // the partial sums of A v_{k+1}
        const int base = k + 2;
        const int nct = (m2 + 63) >> 6;
        // smallest band height whose column partials fit in ccap doubles
        const int BR = ((m2 + 7) >> 3) * m2 <= ccap ? 8 : (((m2 + 15) >> 4) * m2 <= ccap ? 16 : 32);
        const int per_tile = 64 / BR, nb = (m2 + BR - 1) / BR;
        double* rowpart = colpart + ccap;
        int nitems = 0;
        for (int ct = 0; ct < nct; ++ct) nitems += min(per_tile * (ct + 1), nb);
        for (int it = wave; it < nitems; it += nw) {
            int ct = 0, b = it;
            while (b >= min(per_tile * (ct + 1), nb)) { b -= min(per_tile * (ct + 1), nb); ++ct; }
            const int c = base + 64 * ct + lane;
            const bool cv = c < n;
            const u32 cbyte = (u32)(cv ? c : 0) * ES;
            const int tc = c - k - 1;
            const double vi = cv ? vcur[tc] : 0.0;
            const double wi = cv ? p[tc] + K * vi : 0.0;
            const double vnc = cv ? vnext[tc - 1] : 0.0;
            double colacc = 0.0;
            const int r0 = base + b * BR, rend = min(r0 + BR, n);
            // element (r, c) at byte r * n * ES + cbyte; only lanes on or above
            // the diagonal load (the lower triangle is never touched). With FP32
            // storage the next 4-row slice is loaded while the current one is
            // processed (four more registers); FP64 storage has no room for it.
            constexpr bool PF = sizeof(TS) == 4;
            TS nx[4];
            if constexpr (PF) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int r = r0 + u;
                    nx[u] = (cv && r <= c && r < rend) ? (TS)sym_load<TS>(Ab, (u32)r0 * (u32)n * ES + cbyte,
                                                                          (u32)u * (u32)n * ES)
                                                        : (TS)0;
                }
            }
            for (int rs = r0; rs < rend; rs += 4) {
                const u32 voff = (u32)rs * (u32)n * ES + cbyte;
                double a[4], x[4];
                if constexpr (PF) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) a[u] = (double)nx[u];
                    const int rn = rs + 4;
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int r = rn + u;
                        nx[u] = (cv && r <= c && r < rend) ? (TS)sym_load<TS>(Ab, (u32)rn * (u32)n * ES + cbyte,
                                                                              (u32)u * (u32)n * ES)
                                                            : (TS)0;
                    }
                } else {
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int r = rs + u;
                        a[u] = (cv && r <= c) ? sym_load<TS>(Ab, voff, (u32)u * (u32)n * ES) : 0.0;
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int r = rs + u;
                    const bool ok = cv && r <= c;
                    const int tj = min(r, n - 1) - k - 1;
                    const double vj = r < n ? vcur[tj] : 0.0;
                    const double wj = r < n ? p[tj] + K * vj : 0.0;
                    const double vnr = r < n ? vnext[tj - 1] : 0.0;
                    const double an = a[u] - (vj * wi + wj * vi);
                    if (ok) sym_store<TS>(an, Ab, voff, (u32)u * (u32)n * ES);
                    colacc += ok ? an * vnr : 0.0;
                    x[u] = (ok && r < c) ? an * vnc : 0.0;
                }
                const double s4 = rows4_reduce(x[0], x[1], x[2], x[3]);
                const int rr = rs + ((lane >> 4) & 3);
                if ((lane & 15) == 0 && rr < n) rowpart[ct * m2 + (rr - base)] = s4;
            }
            if (cv) colpart[b * m2 + (c - base)] = colacc;
        }
        __syncthreads();   // 3
        // (A v_{k+1})_t = sum of column partials of bands 0 .. t/BR and row
        // partials of column tiles t/64 .. nct-1, S lanes per t
        int S = 1;
        while (S < 64 && 2 * S * m2 <= T) S *= 2;
        pv = 0.0;
        for (int t0s = 0; t0s < m2; t0s += T / S) {
            const int t = t0s + tid / S, sl = tid & (S - 1);
            double q = 0.0;
            if (t < m2) {
                for (int bb = sl; bb <= t / BR; bb += S) q += colpart[bb * m2 + t];
                for (int cc = (t >> 6) + sl; cc < nct; cc += S) q += rowpart[cc * m2 + t];
            }
            q = group_sum(q, S);
            if (sl == 0 && t < m2) {
                const double pn = tn * q;
                pnext[t] = pn;
                pv += pn * vnext[t];
            }
        }
        pv = group_sum(pv, 64);
        if (lane == 0) red[wave] = pv;
        __syncthreads();   // 4
        double* tmp = vcur; vcur = vnext; vnext = tmp;
        tmp = p; p = pnext; pnext = tmp;
        tk = tn;
    }
}

template <bool PG>
__global__ void __launch_bounds__(1024, 8) k_tridiag_sym(const JobDev* __restrict__ jobs, JobState* __restrict__ state,
                                                      const int* __restrict__ joblist, double* __restrict__ arena,
                                                      double* __restrict__ tri) {
    extern __shared__ double dyn[];
    __shared__ double red[16], red2[16];
    const int job = joblist[blockIdx.x];
    const JobDev jb = jobs[job];
    if (state[job].phase != PH_ACTIVE) return;
    const unsigned long long t0 = wall_clock64();
    const int n = jb.nao, T = blockDim.x, tid = threadIdx.x;
    double* W = amat(arena, jb, AR_W);
    const double* F = amat(arena, jb, AR_F);
    double* d = eig_ws(tri, jb, job);
    double* e = d + n;
    double* tau = e + n;
    for (int i = tid; i < n; i += T) { e[i] = 0.0; tau[i] = 0.0; }
    if (n < 3) {
        __syncthreads();
        if (tid == 0) {
            d[0] = F[0];
            if (n == 2) { e[0] = F[1]; d[1] = F[3]; }
            state[job].prof[0] = (double)(wall_clock64() - t0);
        }
        return;
    }
    double* colpart = PG ? amat(arena, jb, AR_U) : dyn + 4 * n;
    // a mixed-precision job in its FP32 iterations keeps the trailing matrix in
    // FP32 in its AR_U matrix (free until divide and conquer)
    if (!PG && state[job].fp32)
        tridiag_sym_body<false, float>(reinterpret_cast<float*>(amat(arena, jb, AR_U)), W, F, colpart, dyn, n, job,
                                       d, e, tau, red, red2);
    else
        tridiag_sym_body<PG, double>(W, W, F, colpart, dyn, n, job, d, e, tau, red, red2);
    if (tid == 0) state[job].prof[0] = (double)(wall_clock64() - t0);
}

// Blocked (deferred-update) variant for matrices in device memory, n <= 512:
// LAPACK's dsytrd / dlatrd scheme (Dongarra, Hammarling & Sorensen, J. Comput.
// Appl. Math. 27, 215 (1989)) on the upper triangle. The reflectors form
// panels of kTriNb. Within a panel the stored trailing matrix A_p (the matrix
// at the panel's first step) is only read: step k forms row k of the current
// matrix A_k = A_p - sum_l (V_l W_l^T + W_l V_l^T) over the panel's earlier
// reflectors l, the reflector v_k from it, then
//   p = tau_k (A_p v_k - sum_l V_l (W_l . v_k) + W_l (V_l . v_k)),
//   w_k = p - (tau_k / 2) (p . v_k) v_k,
// and keeps V_l = v_l, W_l = w_l (full length, zero before the reflector's
// first index) in the job's AR_Q2 matrix, free until divide and conquer. At
// the next panel's first step one rank-2 kTriNb update of the upper triangle
// on FP64 MFMA brings the trailing block to A_{p + kTriNb}. A step reads the
// trailing triangle once and writes nothing to it, against one read and one
// write per step in k_tridiag_sym; the reflectors, d, e and tau are stored as
// there, so divide and conquer and the back-transform are unchanged.
constexpr int kTriNb = 8;

// Trailing block k0..n-1 of the upper triangle minus sum_l (V_l W_l^T + W_l V_l^T)
// over the kTriNb vectors of VW (V_0..V_7 then W_0..W_7, n doubles each); one
// wave per 16 x 16 tile, the rank-16 product as four v_mfma_f64_16x16x4f64 with
// the tile itself as the accumulator (element v of a lane is row 4 v + lane / 16,
// column lane % 16 of the tile).
template <typename TS>
__device__ inline void tridiag_panel_update(__amdgpu_buffer_rsrc_t Ab, double* __restrict__ VW, int n, int k0,
                                            int wave, int nw, int lane) {
    constexpr u32 ES = sizeof(TS);
    const int m0 = n - k0, nt = (m0 + 15) >> 4, ntiles = nt * (nt + 1) / 2;
    const int lr = lane & 15, lk = lane >> 4;
    // V/W element (kk, j) at byte (kk n + j) 8: the lane's part (lk n + j) 8 in
    // a VGPR, the k-step's 4 q n 8 in a scalar offset
    const __amdgpu_buffer_rsrc_t Vb =
        __builtin_amdgcn_make_buffer_rsrc(VW, (short)0, 2 * kTriNb * n * (int)sizeof(double), 0x00020000);
    const u32 nbytes = (u32)n * 8u;
    for (int tile = wave; tile < ntiles; tile += nw) {
        // tile = J (J + 1) / 2 + I with 0 <= I <= J < nt
        int J = (int)((sqrtf(8.0f * (float)tile + 1.0f) - 1.0f) * 0.5f);
        while (J * (J + 1) / 2 > tile) --J;
        while ((J + 1) * (J + 2) / 2 <= tile) ++J;
        const int I = tile - J * (J + 1) / 2;
        const int rb = k0 + 16 * I, cb = k0 + 16 * J, c = cb + lr;
        const u32 cbyte = (u32)(c < n ? c : 0) * ES;
        dbl4 acc;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = rb + 4 * v + lk;
            acc[v] = (c < n && r <= c) ? sym_load<TS>(Ab, (u32)r * (u32)n * ES + cbyte, 0) : 0.0;
        }
        const int ra = rb + lr;
        const u32 va = (u32)lk * nbytes + (u32)(ra < n ? ra : 0) * 8u, vb = (u32)lk * nbytes + (u32)(c < n ? c : 0) * 8u;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            // k = 4 q + lk: V_k (k < 8) or W_{k-8} against W_k or V_{k-8}
            const double a = buf_load_d(Vb, va, (u32)(4 * q) * nbytes);
            const double b = buf_load_d(Vb, vb, (u32)((4 * q + kTriNb) & (2 * kTriNb - 1)) * nbytes);
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ra < n ? -a : 0.0, c < n ? b : 0.0, acc, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = rb + 4 * v + lk;
            if (c < n && r <= c) sym_store<TS>(acc[v], Ab, (u32)r * (u32)n * ES + cbyte, 0);
        }
    }
}

template <typename TS>
__device__ void tridiag_blk_body(TS* __restrict__ As, double* __restrict__ W, const double* __restrict__ F,
                                 double* __restrict__ VW, double* dyn, int n, double* d, double* e, double* tau,
                                 double* red, double* red2, double* dots) {
    const int T = blockDim.x, tid = threadIdx.x, lane = tid & 63, nw = T >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    constexpr u32 ES = sizeof(TS);
    // v and p of the current step (indexed from the reflector's first index)
    // and of the previous one, ping-pong
    double* vb0 = dyn;
    double* vb1 = dyn + n;
    double* pb0 = dyn + 2 * n;
    double* pb1 = dyn + 3 * n;
    double* colpart = dyn + 4 * n;
    const int ccap = sym_col_cap(n);
    double* rowpart = colpart + ccap;
    // row k of the current matrix between its formation and its reflector
    double* xrow = colpart;
    const __amdgpu_buffer_rsrc_t Ab = __builtin_amdgcn_make_buffer_rsrc(As, (short)0, n * n * (int)ES, 0x00020000);
    for (int r = wave; r < n; r += nw)
        for (int c = r + lane; c < n; c += 64) As[(long long)r * n + c] = (TS)F[(long long)r * n + c];
    __syncthreads();
    double Kp = 0.0;   // -tau/2 (p . v) of the previous step: w = p + Kp v
    for (int k = 0; k <= n - 3; ++k) {
        const int i = k % kTriNb;
        double* vcur = (k & 1) ? vb1 : vb0;
        double* pcur = (k & 1) ? pb1 : pb0;
        const double* vprev = (k & 1) ? vb0 : vb1;   // v_{k-1}, indices k .. n-1
        const double* pprev = (k & 1) ? pb0 : pb1;
        if (k > 0 && i == 0) {
            // last W of the panel, then the panel's rank-2 kTriNb update
            for (int c = k + tid; c < n; c += T)
                VW[(long long)(2 * kTriNb - 1) * n + c] = pprev[c - k] + Kp * vprev[c - k];
            __syncthreads();
            tridiag_panel_update<TS>(Ab, VW, n, k, wave, nw, lane);
            __syncthreads();
        }
        // row k of A_k, columns k .. n-1; W_{i-1} is formed here from p and v
        const int m0 = n - k;
        double ss = 0.0;
        for (int t = tid; t < m0; t += T) {
            const int c = k + t;
            double x = (double)As[(long long)k * n + c];
            for (int l = 0; l + 1 < i; ++l) {
                const double* V = VW + (long long)l * n;
                const double* Wl = VW + (long long)(kTriNb + l) * n;
                x -= V[k] * Wl[c] + Wl[k] * V[c];
            }
            if (i > 0) {
                const double wc = pprev[t] + Kp * vprev[t], wk = pprev[0] + Kp * vprev[0];
                VW[(long long)(kTriNb + i - 1) * n + c] = wc;
                x -= vprev[0] * wc + wk * vprev[t];
            }
            xrow[t] = x;
            if (t >= 2) ss += x * x;
        }
        ss = group_sum(ss, 64);
        if (lane == 0) red2[wave] = ss;
        __syncthreads();   // 1
        ss = 0.0;
        for (int w = 0; w < nw; ++w) ss += red2[w];
        const double alpha = xrow[1];
        double beta = alpha, tn = 0.0, scale = 0.0;
        if (ss > 0.0) {
            beta = -copysign(sqrt(alpha * alpha + ss), alpha);
            tn = (beta - alpha) / beta;
            scale = 1.0 / (alpha - beta);
        }
        const int m = m0 - 1;   // length of v_k (indices k+1 .. n-1)
        const long long wrow = (long long)k * n + k + 1;
        double* Vi = VW + (long long)i * n + k + 1;
        for (int t = 1 + tid; t < m; t += T) {
            const double v = ss > 0.0 ? xrow[1 + t] * scale : 0.0;
            vcur[t] = v;
            W[wrow + t] = v;
            Vi[t] = v;
        }
        if (tid == 0) {
            vcur[0] = 1.0;
            W[wrow] = 1.0;
            Vi[0] = 1.0;
            d[k] = xrow[0];
            e[k] = beta;
            tau[k] = tn;
        }
        __syncthreads();   // 2
        const int base = k + 1;
        // W_l . v_k and V_l . v_k of the panel's earlier reflectors, one wave each
        if (wave < 2 * i) {
            const int l = wave >> 1;
            const double* src = VW + (long long)((wave & 1) ? l : kTriNb + l) * n + base;
            double s = 0.0;
            for (int t = lane; t < m; t += 64) s += src[t] * vcur[t];
            s = group_sum(s, 64);
            if (lane == 0) dots[wave] = s;
        }
        // A_p v_k over the upper triangle of rows/columns base .. n-1 (read only)
        const int nct = (m + 63) >> 6;
        const int BR = ((m + 7) >> 3) * m <= ccap ? 8 : (((m + 15) >> 4) * m <= ccap ? 16 : 32);
        const int per_tile = 64 / BR, nb = (m + BR - 1) / BR;
        int nitems = 0;
        for (int ct = 0; ct < nct; ++ct) nitems += min(per_tile * (ct + 1), nb);
        for (int it = wave; it < nitems; it += nw) {
            int ct = 0, b = it;
            while (b >= min(per_tile * (ct + 1), nb)) { b -= min(per_tile * (ct + 1), nb); ++ct; }
            const int c = base + 64 * ct + lane;
            const bool cv = c < n;
            const u32 cbyte = (u32)(cv ? c : 0) * ES;
            const double vc = cv ? vcur[c - base] : 0.0;
            double colacc = 0.0;
            const int r0 = base + b * BR, rend = min(r0 + BR, n);
            // two 4-row slices per pass, their eight loads issued together (BR is
            // a multiple of 8, so a pass never crosses into the next band)
            for (int rs = r0; rs < rend; rs += 8) {
                const u32 voff = (u32)rs * (u32)n * ES + cbyte;
                double a[8];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    a[u] = (cv && rs + u <= c && rs + u < rend) ? sym_load<TS>(Ab, voff, (u32)u * (u32)n * ES)
                                                                : 0.0;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int rh = rs + 4 * h;
                    if (rh >= rend) break;
                    double x[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int r = rh + u;
                        const double vr = r < n ? vcur[r - base] : 0.0;
                        colacc += a[4 * h + u] * vr;
                        x[u] = r < c ? a[4 * h + u] * vc : 0.0;
                    }
                    const double s4 = rows4_reduce(x[0], x[1], x[2], x[3]);
                    const int rr = rh + ((lane >> 4) & 3);
                    if ((lane & 15) == 0 && rr < n) rowpart[ct * m + (rr - base)] = s4;
                }
            }
            if (cv) colpart[b * m + (c - base)] = colacc;
        }
        __syncthreads();   // 3
        int S = 1;
        while (S < 64 && 2 * S * m <= T) S *= 2;
        double pv = 0.0;
        for (int t0s = 0; t0s < m; t0s += T / S) {
            const int t = t0s + tid / S, sl = tid & (S - 1);
            double q = 0.0;
            if (t < m) {
                for (int bb = sl; bb <= t / BR; bb += S) q += colpart[bb * m + t];
                for (int cc = (t >> 6) + sl; cc < nct; cc += S) q += rowpart[cc * m + t];
            }
            q = group_sum(q, S);
            if (sl == 0 && t < m) {
                double corr = 0.0;
                for (int l = 0; l < i; ++l)
                    corr += VW[(long long)l * n + base + t] * dots[2 * l] +
                            VW[(long long)(kTriNb + l) * n + base + t] * dots[2 * l + 1];
                const double pn = tn * (q - corr);
                pcur[t] = pn;
                pv += pn * vcur[t];
            }
        }
        pv = group_sum(pv, 64);
        if (lane == 0) red[wave] = pv;
        __syncthreads();   // 4
        pv = 0.0;
        for (int w = 0; w < nw; ++w) pv += red[w];
        Kp = -0.5 * tn * pv;
        if (k == n - 3) {
            // the last 2 x 2 block: rows n-2, n-1 of A_{n-2}
            if (tid == 0) {
                double a00 = (double)As[(long long)(n - 2) * n + n - 2];
                double a01 = (double)As[(long long)(n - 2) * n + n - 1];
                double a11 = (double)As[(long long)(n - 1) * n + n - 1];
                for (int l = 0; l < i; ++l) {
                    const double* V = VW + (long long)l * n;
                    const double* Wl = VW + (long long)(kTriNb + l) * n;
                    a00 -= 2.0 * V[n - 2] * Wl[n - 2];
                    a01 -= V[n - 2] * Wl[n - 1] + Wl[n - 2] * V[n - 1];
                    a11 -= 2.0 * V[n - 1] * Wl[n - 1];
                }
                const double w0 = pcur[0] + Kp * vcur[0], w1 = pcur[1] + Kp * vcur[1];
                a00 -= 2.0 * vcur[0] * w0;
                a01 -= vcur[0] * w1 + w0 * vcur[1];
                a11 -= 2.0 * vcur[1] * w1;
                d[n - 2] = a00;
                e[n - 2] = a01;
                d[n - 1] = a11;
            }
        }
    }
}

__global__ void __launch_bounds__(1024, 8) k_tridiag_blk(const JobDev* __restrict__ jobs,
                                                       JobState* __restrict__ state, const int* __restrict__ joblist,
                                                       double* __restrict__ arena, double* __restrict__ tri) {
    extern __shared__ double dyn[];
    __shared__ double red[16], red2[16], dots[2 * kTriNb];
    const int job = joblist[blockIdx.x];
    const JobDev jb = jobs[job];
    if (state[job].phase != PH_ACTIVE) return;
    const unsigned long long t0 = wall_clock64();
    const int n = jb.nao, T = blockDim.x, tid = threadIdx.x;
    double* W = amat(arena, jb, AR_W);
    const double* F = amat(arena, jb, AR_F);
    double* VW = amat(arena, jb, AR_Q2);
    double* d = eig_ws(tri, jb, job);
    double* e = d + n;
    double* tau = e + n;
    for (int i = tid; i < n; i += T) { e[i] = 0.0; tau[i] = 0.0; }
    if (n < 3) {
        __syncthreads();
        if (tid == 0) {
            d[0] = F[0];
            if (n == 2) { e[0] = F[1]; d[1] = F[3]; }
            state[job].prof[0] = (double)(wall_clock64() - t0);
        }
        return;
    }
    if (n < 2 * kTriNb) {
        // the panel vectors need 16 n <= n^2 doubles of AR_Q2; molecules this
        // small take the LDS class on gfx950 (160 KB of LDS), so this is not
        // reached: a NaN diagonal makes divide and conquer fail the job loudly
        if (tid == 0) d[0] = __builtin_nan("");
    } else {
        // FP64 storage also for mixed-precision jobs in their FP32 iterations:
        // this kernel's steps are bound by load round trips, not bytes, and an
        // FP32-storage form of it ran slower (328 against 294 ms of
        // tridiagonalization per 4096-molecule pass)
        tridiag_blk_body<double>(W, W, F, VW, dyn, n, d, e, tau, red, red2, dots);
    }
    if (tid == 0) state[job].prof[0] = (double)(wall_clock64() - t0);
}

// ------------------------------------------------------------------ 2. divide and conquer
// Rows of QT (AR_C / AR_Q2, ping-pong) are eigenvectors of the current blocks;
// the block [a, b) of level l only touches rows and columns a..b-1.  U (AR_U)
// holds each merge's kk x kk eigenvector matrix in its diagonal block.
__global__ void __launch_bounds__(1024) k_dc(const JobDev* __restrict__ jobs, JobState* __restrict__ state,
                                             const int* __restrict__ joblist, double* __restrict__ arena,
                                             const double* __restrict__ tri, double* __restrict__ eps,
                                             int* __restrict__ perm) {
    extern __shared__ double dyn[];
    const int job = joblist[blockIdx.x];
    const JobDev jb = jobs[job];
    if (state[job].phase != PH_ACTIVE) return;
    // deflation and secular roots to FP32 accuracy for a mixed-precision job in
    // its FP32 iterations (its tridiagonal comes from FP32-stored data)
    const double ue = state[job].fp32 ? kEps32 : kEps;
    const unsigned long long t0 = wall_clock64();
    const int n = jb.nao;
    const int T = blockDim.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = T >> 6;
    double* Dc = dyn;            // sorted eigenvalues of each current block
    double* ds = Dc + n;         // merged d, z in sorted order (per merge at [a, b))
    double* zs = ds + n;
    double* dl = zs + n;         // non-deflated first (kk of them), then deflated
    double* zl = dl + n;
    double* vals = zl + n;       // new eigenvalues: kk roots, then the deflated values
    double* rtau = vals + n;     // root offset from its pole
    double* zhat = rtau + n;
    double* cnorm = zhat + n;    // inverse column norms of U
    double* rc = cnorm + n;      // deflation rotations (c, s)
    double* rs = rc + n;
    double* rhoq = rs + n;       // per merge
    int* col = (int*)(rhoq + n); // QT row of each sorted entry
    int* flag = col + n;
    int* cl = flag + n;          // QT row of each dl entry
    int* rorig = cl + n;         // pole index (into dl) of each root
    int* rank = rorig + n;       // position of each vals entry in the sorted block
    int* r1 = rank + n;          // deflation rotation rows
    int* r2 = r1 + n;
    int* kkq = r2 + n;           // per merge: kk (-1: block carried to the next level)
    int* nrot = kkq + n;
    const double* d = eig_ws(const_cast<double*>(tri), jb, job);
    const double* e = d + n;
    double* QC = amat(arena, jb, AR_C);
    double* QX = amat(arena, jb, AR_Q2);
    double* U = amat(arena, jb, AR_U);
    // a non-finite tridiagonal (from a non-finite Fock matrix) fails the solve
    int bad = 0;
    for (int i = tid; i < n; i += T) bad |= !isfinite(d[i]) || (i < n - 1 && !isfinite(e[i]));
    if (__syncthreads_or(bad)) {
        for (int i = tid; i < n; i += T) {
            eps[jb.vec0 + i] = 0.0;
            perm[jb.ivec0 + i] = i;
        }
        if (tid == 0) state[job].sweeps = -1;
        return;
    }
    int L = 0;
    while ((1 << L) < n) ++L;
    {
        double* Q0 = (L & 1) ? QX : QC;
        for (int i = tid; i < n; i += T) {
            // T = diag(T_1, T_2) + |e| u u^T at every cut: both diagonal ends lose |e|
            const double el = i > 0 ? fabs(e[i - 1]) : 0.0, er = i < n - 1 ? fabs(e[i]) : 0.0;
            Dc[i] = d[i] - el - er;
            Q0[(long long)i * n + i] = 1.0;
        }
    }
    __syncthreads();
    // per-phase device clock ticks, summed over the levels (thread 0), kept in
    // the spare tail of the job's eigensolver workspace for the profiling hook
    unsigned long long tick[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    unsigned long long tlast = wall_clock64();
#define DC_TICK(k)                                   \
    if (tid == 0) {                                  \
        const unsigned long long tnow = wall_clock64(); \
        tick[k] += tnow - tlast;                     \
        tlast = tnow;                                \
    }
    for (int l = 1; l <= L; ++l) {
        const int s = 1 << l, h = s >> 1, nm = (n + s - 1) >> l;
        double* QO = ((L - l + 1) & 1) ? QX : QC;
        double* QN = ((L - l) & 1) ? QX : QC;
        // A: z = [last row of Q_1; sign(e) first row of Q_2] / sqrt(2); stable merge of the two sorted blocks
        for (int i = tid; i < n; i += T) {
            const int a = (i >> l) << l, m = min(a + h, n), b = min(a + s, n);
            const long long ri = (long long)i * n;
            if (m == b) {   // unpaired last block: carried as it is
                for (int c = a; c < b; ++c) QN[ri + c] = QO[ri + c];
                continue;
            }
            double z = i < m ? QO[ri + m - 1] : QO[ri + m];
            if (i >= m && e[m - 1] < 0.0) z = -z;
            z *= 0.70710678118654752440;
            const double di = Dc[i];
            // position in the merged order: the stable merge of the two sorted blocks
            // (left before right on ties), counted directly so it is a permutation
            int pos = 0;
            for (int t = a; t < b; ++t) pos += dc_before(Dc[t], t, di, i);
            ds[a + pos] = di;
            zs[a + pos] = z;
            col[a + pos] = i;
        }
        // the merged block of QO is diag(Q_1, Q_2): clear its two off-diagonal blocks
        for (long long x = tid; x < ((long long)n << (l - 1)); x += T) {
            const int i = (int)(x >> (l - 1)), c0 = (int)(x & (h - 1));
            const int a = (i >> l) << l, m = min(a + h, n), b = min(a + s, n);
            if (m == b) continue;
            const int c = i < m ? m + c0 : a + c0;
            if (c < (i < m ? b : m)) QO[(long long)i * n + c] = 0.0;
        }
        __syncthreads();
        DC_TICK(0)
        // B: deflation (small |z| and close d pairs), one wavefront per merge: the
        // maxima and the compaction are wave-parallel, the scan itself (each pair
        // depends on the previous decision) runs on lane 0 with the running pair
        // in registers
        for (int q = wave; q < nm; q += nw) {
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
            } else if (lane == 0) {
                int pj = -1, nr = 0, ps = 0;
                double zp = 0.0, dp = 0.0;
                // the next entry's (z, d) is loaded one iteration ahead: entries at
                // or after jj are not written before their own iteration
                double zn = zs[a], dn = ds[a];
                int cn = col[a];
                for (int jj = a; jj < b; ++jj) {
                    double zj = zn, dj = dn;
                    // halves of QO the row spans (1 left, 2 right, 3 both after a rotation)
                    int js = cn < m ? 1 : 2;
                    if (jj + 1 < b) { zn = zs[jj + 1]; dn = ds[jj + 1]; cn = col[jj + 1]; }
                    if (rho * fabs(zj) <= tol) { flag[jj] = 1; continue; }
                    flag[jj] = 0;
                    if (pj >= 0) {
                        // |t c s| <= tol with c = zj / tz, s = -zp / tz, tz^2 = zj^2 + zp^2
                        const double t = dj - dp, zz = zj * zj + zp * zp;
                        if (fabs(t * zj * zp) <= tol * zz) {
                            const double tz = sqrt(zz);
                            const double C = zj / tz, S = -zp / tz;
                            zs[jj] = tz;
                            zs[pj] = 0.0;
                            r1[a + nr] = col[pj];
                            r2[a + nr] = col[jj];
                            rc[a + nr] = C;
                            rs[a + nr] = S;
                            ++nr;
                            ds[pj] = dp * C * C + dj * S * S;
                            dj = dp * S * S + dj * C * C;
                            ds[jj] = dj;
                            zj = tz;
                            flag[pj] = 1;
                            js |= ps;
                            // kept, with support in both halves (2: kept, mixed)
                            if (js == 3) flag[jj] = 2;
                        }
                    }
                    pj = jj;
                    ps = js;
                    zp = zj;
                    dp = dj;
                }
                nrot[q] = nr;
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            // compaction: non-deflated first (kk of them), then the deflated ones
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
                    rorig[a + k + pre] = fj;  // mixed mark until the secular roots
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
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            // the non-deflated rows by the halves of QO = diag(Q1, Q2) they span:
            // flag[a ..) lists the rows t that meet the left half, then those that
            // meet the right half (a row rotated with one from the other half is in
            // both); the merge products of a column tile within one half run over
            // that half's list only (the other rows meet zeros there). Only at the
            // levels whose products run on MFMA, l >= 5: there nm <= n / 32, so
            // kkq[n/2 + q] and kkq[3n/4 + q] are clear of the counts kkq[0 .. nm).
            // Lists longer than the merge's flag slots leave the full range (-1).
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
                    // kl ends at nl, ke at nl + nr
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
        __syncthreads();
        DC_TICK(1)
        // C: deflation rotations on the rows of QO (thread = column); D: secular roots
        for (int i = tid; i < n; i += T) {
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
            const int j = i - a;
            if (j < kk) {
                int o;
                double tau;
                dc_secular(dl + a, zl + a, kk, rhoq[q], j, &o, &tau, ue);
                rorig[i] = a + o;
                rtau[i] = tau;
                vals[i] = dl[a + o] + tau;
            } else {
                vals[i] = dl[i];
            }
        }
        __syncthreads();
        DC_TICK(2)
        // E: Loewner z-hat (Gu & Eisenstat) and the rank of every new eigenvalue in its block
        for (int i = tid; i < n; i += T) {
            const int q = i >> l, a = q << l;
            const int kk = kkq[q];
            if (kk < 0) continue;
            const int b = min(a + s, n), j = i - a;
            if (j < kk) {
                const double di = dl[i];
                double wv = (di - dl[rorig[i]]) - rtau[i];
                for (int t = 0; t < kk; ++t) {
                    if (t == j) continue;
                    const int ct = a + t;
                    wv *= ((di - dl[rorig[ct]]) - rtau[ct]) / (di - dl[ct]);
                }
                zhat[i] = copysign(sqrt(fmax(-wv, 0.0)), zl[i]);
            }
            const double v = vals[i];
            int rk = 0;
            for (int t = a; t < b; ++t) rk += dc_before(vals[t], t, v, i);
            rank[i] = rk;
        }
        __syncthreads();
        DC_TICK(3)
        // F: inverse norms of the columns of U
        for (int i = tid; i < n; i += T) {
            const int q = i >> l, a = q << l;
            const int kk = kkq[q];
            if (i - a >= kk) continue;
            const double dO = dl[rorig[i]], tj = rtau[i];
            double s2 = 0.0;
            for (int t = 0; t < kk; ++t) {
                const double u = zhat[a + t] / ((dl[a + t] - dO) - tj);
                s2 += u * u;
            }
            cnorm[i] = 1.0 / sqrt(s2);
        }
        __syncthreads();
        DC_TICK(4)
        // G: U[t][j] = zhat_t / (dl_t - lambda_j), columns normalized
        const long long tot = (long long)n << l;
        for (long long x = tid; x < tot; x += T) {
            const int i = (int)(x >> l), j = (int)(x & (s - 1));
            const int q = i >> l, a = q << l, kk = kkq[q];
            if (i - a >= kk || j >= kk) continue;
            const int cj = a + j;
            U[(long long)i * n + cj] = zhat[i] / ((dl[i] - dl[rorig[cj]]) - rtau[cj]) * cnorm[cj];
        }
        __syncthreads();
        DC_TICK(5)
        // H: new eigenvector rows QN[a + rank] = U^T QO[cl]; deflated rows copied; sorted eigenvalues
        const bool small = l < 5;
        for (long long x = tid; x < tot; x += T) {
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
            // 32 x 32 output tiles (eigenvector index j x column r), one per wavefront,
            // as 2 x 2 v_mfma_f64_16x16x4f64 blocks; k runs over the kk non-deflated rows
            const int tpm = s >> 5;
            const int ntile = nm * tpm * tpm;
            const int lr = lane & 15, lk = lane >> 4;
            for (int ti = wave; ti < ntile; ti += nw) {
                const int q = ti / (tpm * tpm), rem = ti - q * tpm * tpm;
                const int jt = rem / tpm, rt = rem - jt * tpm;
                const int kk = kkq[q];
                const int a = q << l, b = min(a + s, n), m = min(a + (s >> 1), n);
                const int j0 = jt * 32, r0 = a + rt * 32;
                if (kk <= 0 || j0 >= kk || r0 >= b) continue;
                // rows that meet this tile's columns: one half's list, or all kk
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
                // D[i][c] of a 16x16 FP64 block sits in lane c + 16 (i % 4), element i / 4
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
        for (int i = tid; i < n; i += T) {
            const int q = i >> l, a = q << l;
            if (kkq[q] >= 0) Dc[a + rank[i]] = vals[i];
        }
        __syncthreads();
        DC_TICK(6)
    }
#undef DC_TICK
    for (int i = tid; i < n; i += T) {
        eps[jb.vec0 + i] = Dc[i];
        perm[jb.ivec0 + i] = i;
    }
    if (tid == 0) {
        state[job].sweeps = 1;
        state[job].warm = 0;
        state[job].prof[1] = (double)(wall_clock64() - t0);
        state[job].prof[2] = 0.0;
        double* ph = eig_ws(const_cast<double*>(tri), jb, job) + 19 * n + 224;
        for (int k = 0; k < 7; ++k) ph[k] = (double)tick[k];
    }
}

// ------------------------------------------------------------------ 3. back-transform
// Eigenvector rows x^T = y^T H_{n-3} ... H_0.  Reflectors are taken in blocks of
// kBt (k0 .. k0+kBt-1); in compact WY form (Schreiber & Van Loan, SIAM J. Sci.
// Stat. Comput. 10, 53 (1989)) H_k0 ... H_k0+kBt-1 = I - V T V^T with T upper
// triangular, so a row block Z is updated as Z <- Z - (Z V) T^T V^T, applying
// the blocks from the last one down.  V[j][c] = W[k0 + c][j] for j > k0 + c.
constexpr int kBt = 16;

// T of every reflector block: one wavefront per (job, block); T goes to the
// job's eigensolver workspace.
__global__ void __launch_bounds__(64) k_wy_t(const JobDev* __restrict__ jobs, const JobState* __restrict__ state,
                                             const int* __restrict__ joblist, const double* __restrict__ arena,
                                             double* __restrict__ tri) {
    __shared__ double G[kBt * kBt];
    __shared__ double Ts[kBt * kBt];
    const int job = joblist[blockIdx.x];
    const JobDev jb = jobs[job];
    if (state[job].phase != PH_ACTIVE) return;
    const int n = jb.nao, nref = n - 2, k0 = blockIdx.y * kBt;
    if (nref <= 0 || k0 >= nref) return;
    const int lane = threadIdx.x, lr = lane & 15, lk = lane >> 4;
    const double* W = arena + jb.arena + (long long)AR_W * n * n;
    double* ws = eig_ws(tri, jb, job);
    const double* tau = ws + 2 * n;
    const int nb = min(kBt, nref - k0);
    // G = V^T V (16 x 16, K = j) on MFMA; rows/columns beyond nb stay zero
    dbl4 acc = dbl4{0.0, 0.0, 0.0, 0.0};
    const int c = k0 + lr;   // reflector of this lane's M (and N) index
    for (int j0 = k0 + 1; j0 < n; j0 += 4) {
        const int j = j0 + lk;
        const double v = (lr < nb && j < n && j > c) ? W[(long long)c * n + j] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(v, v, acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) G[(4 * v + lk) * kBt + lr] = acc[v];
    for (int x = lane; x < kBt * kBt; x += 64) Ts[x] = 0.0;
    __syncthreads();
    // T[i][i] = tau_i; T[0:i, i] = -tau_i T[0:i, 0:i] G[0:i, i]
    for (int i = 0; i < nb; ++i) {
        const double ti = tau[k0 + i];
        double t = 0.0;
        if (lane < i) {
            for (int cc = lane; cc < i; ++cc) t += Ts[lane * kBt + cc] * G[cc * kBt + i];
            t *= -ti;
        }
        __syncthreads();
        if (lane < i) Ts[lane * kBt + i] = t;
        if (lane == i) Ts[i * kBt + i] = ti;
        __syncthreads();
    }
    double* Tout = ws + 3 * n + (long long)blockIdx.y * kBt * kBt;
    for (int x = lane; x < kBt * kBt; x += 64) Tout[x] = Ts[x];
}

// One workgroup (8 wavefronts) per (job, 16 RT eigenvector rows).  The rows are
// held transposed, Z^T[j][r], as v_mfma_f64_16x16x4f64 accumulator tiles: wave
// w owns the 16-column tiles tj = w, w + 8, ... for all RT 16-row tiles.  In
// that layout element v of a lane holds Z^T[16 tj + 4 v + lk][r0 + 16 rt + lr],
// which is exactly the B operand of the k-step v of Y^T = V^T Z^T, so Z never
// leaves registers.  Per block: Y^T = V^T Z^T (partials per wave, summed in
// LDS), Y2^T = -T Y^T, Z^T += V Y2^T.  The block's 16 reflectors are staged in
// LDS (VG = false, up to 512 AOs: RT = 4, 64 rows) or read from device memory
// (VG = true, larger molecules: RT = 2, 32 rows, which keeps the row block in
// registers and the LDS footprint small).
template <int JT, int RT, bool VG>
__global__ void __launch_bounds__(512) k_bt_wy(const JobDev* __restrict__ jobs, JobState* __restrict__ state,
                                               const int* __restrict__ joblist, double* __restrict__ arena,
                                               double* __restrict__ tri) {
    constexpr int R = 16 * RT;   // rows per workgroup
    extern __shared__ double dyn[];
    const int job = joblist[blockIdx.x];
    const JobDev jb = jobs[job];
    if (state[job].phase != PH_ACTIVE) return;
    const int n = jb.nao, nref = n - 2, r0 = blockIdx.y * R;
    // rows (eigenvectors) wanted: all, or the occupied ones when the caller takes no C
    const int nrow = jb.want_c ? n : jb.nocc;
    if (nref <= 0 || r0 >= nrow) return;
    const unsigned long long t0 = wall_clock64();
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, lr = lane & 15, lk = lane >> 4;
    const int ld = n + 1;
    const int ntj = (n + 15) >> 4;
    double* Vs = dyn;                           // 16 x ld: Vs[c][j] = V[j][c] (VG = false)
    double* Yp = VG ? dyn : Vs + 16 * ld;       // 8 x 16 x R partial Y^T
    double* Ys = Yp + 8 * 16 * R;               // 16 x R: Y^T, then Y2^T
    double* Ts = Ys + 16 * R;                   // 16 x 16
    double* ZT = amat(arena, jb, AR_C);
    const double* W = amat(arena, jb, AR_W);
    const double* Tall = eig_ws(tri, jb, job) + 3 * n;
    dbl4 acc[JT][RT];
#pragma unroll
    for (int x = 0; x < JT; ++x) {
        const int tj = wave + 8 * x;
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) {
            const int r = r0 + 16 * rt + lr;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int j = 16 * tj + 4 * v + lk;
                acc[x][rt][v] = (tj < ntj && r < n && j < n) ? ZT[(long long)r * n + j] : 0.0;
            }
        }
    }
    const int nblk = (nref + kBt - 1) / kBt;
    for (int bi = nblk - 1; bi >= 0; --bi) {
        const int k0 = bi * kBt, nb = min(kBt, nref - k0);
        const int tlo = (k0 + 1) >> 4;   // first column tile touched by this block
        // V[j][c] of the block: row k0 + c of W from column k0 + c + 1 on (v[0] = 1 is stored)
        auto vget = [&](int c, int j) -> double {
            if (VG) return (c < nb && j > k0 + c && j < n) ? W[(long long)(k0 + c) * n + j] : 0.0;
            return j < n ? Vs[c * ld + j] : 0.0;
        };
        if (!VG) {
            const int jlo = tlo << 4, span = n - jlo;
            for (int x = tid; x < 16 * span; x += 512) {
                const int cc = x / span, j = jlo + (x - cc * span);
                Vs[cc * ld + j] = (cc < nb && j > k0 + cc) ? W[(long long)(k0 + cc) * n + j] : 0.0;
            }
        }
        if (tid < kBt * kBt) Ts[tid] = Tall[(long long)bi * kBt * kBt + tid];
        __syncthreads();
        // partial Y^T[c][r] over this wave's column tiles
        dbl4 y[RT];
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) y[rt] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int x = 0; x < JT; ++x) {
            const int tj = wave + 8 * x;
            if (tj < tlo || tj >= ntj) continue;
            double a[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) a[v] = vget(lr, 16 * tj + 4 * v + lk);
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int rt = 0; rt < RT; ++rt)
                    y[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[v], acc[x][rt][v], y[rt], 0, 0, 0);
        }
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
            for (int v = 0; v < 4; ++v) Yp[wave * 16 * R + (4 * v + lk) * R + 16 * rt + lr] = y[rt][v];
        __syncthreads();
        for (int x = tid; x < 16 * R; x += 512) {
            double sum = 0.0;
#pragma unroll
            for (int w = 0; w < 8; ++w) sum += Yp[w * 16 * R + x];
            Ys[x] = sum;
        }
        __syncthreads();
        constexpr int NY = (16 * R + 511) / 512;
        double y2[NY];
#pragma unroll
        for (int u = 0; u < NY; ++u) {   // Y2^T[c][r] = -sum_{c2 >= c} T[c][c2] Y^T[c2][r]
            const int x = tid + 512 * u, cc = x / R, r = x % R;
            double sum = 0.0;
            if (x < 16 * R)
                for (int c2 = cc; c2 < kBt; ++c2) sum += Ts[cc * kBt + c2] * Ys[c2 * R + r];
            y2[u] = -sum;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < NY; ++u)
            if (tid + 512 * u < 16 * R) Ys[tid + 512 * u] = y2[u];
        __syncthreads();
        // Z^T[j][r] += sum_c V[j][c] Y2^T[c][r]
#pragma unroll
        for (int x = 0; x < JT; ++x) {
            const int tj = wave + 8 * x;
            if (tj < tlo || tj >= ntj) continue;
            const int j = 16 * tj + lr;
            double a[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) a[q] = vget(4 * q + lk, j);
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int rt = 0; rt < RT; ++rt) {
                    const double b = Ys[(4 * q + lk) * R + 16 * rt + lr];
                    acc[x][rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], b, acc[x][rt], 0, 0, 0);
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int x = 0; x < JT; ++x) {
        const int tj = wave + 8 * x;
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) {
            const int r = r0 + 16 * rt + lr;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int j = 16 * tj + 4 * v + lk;
                if (tj < ntj && r < n && j < n) ZT[(long long)r * n + j] = acc[x][rt][v];
            }
        }
    }
    if (tid == 0 && blockIdx.y == 0) state[job].prof[3] = (double)(wall_clock64() - t0);
}

// ------------------------------------------------------------------ host side
void raise_dynamic_lds(const void* func, size_t bytes) {
    static std::mutex mu;
    static std::map<std::pair<int, const void*>, size_t> limit;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    std::lock_guard<std::mutex> lk(mu);
    size_t& cur = limit[{dev, func}];
    if (bytes <= cur) return;
    if (hipFuncSetAttribute(func, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) == hipSuccess) cur = bytes;
}

static size_t bt_lds_bytes(int n, int rt, bool vg) {
    const size_t r = 16 * (size_t)rt;
    return ((vg ? 0 : (size_t)16 * (n + 1)) + 8 * 16 * r + 16 * r + 256) * 8;
}

template <int JT, int RT, bool VG>
static void launch_bt(hipStream_t st, const JobDev* jobs, JobState* state, const int* joblist, int nlist, int max_n,
                      double* arena, double* tri) {
    const size_t lb = bt_lds_bytes(max_n, RT, VG);
    raise_dynamic_lds((const void*)k_bt_wy<JT, RT, VG>, lb);
    hipLaunchKernelGGL((k_bt_wy<JT, RT, VG>), dim3(nlist, (max_n + 16 * RT - 1) / (16 * RT)), dim3(512), lb, st, jobs,
                       state, joblist, arena, tri);
}

bool heig_supported(int n) { return n <= kHeigMaxN; }

size_t heig_lds_bytes(int n, bool a_in_lds) {
    return a_in_lds ? tridiag_lds_bytes(n, true) : tridiag_sym_lds_bytes(n, n > kSymMaxLdsN);
}

void launch_heig_part(int part, hipStream_t st, const JobDev* jobs, JobState* state, const int* joblist, int nlist,
                      int max_n, double* arena, double* eps, int* perm, double* tri, bool a_in_lds) {
    if (nlist <= 0) return;
    if (part == 0) {
        const int threads = tridiag_threads(max_n);
        const size_t lt = tridiag_lds_bytes(max_n, a_in_lds);
        if (a_in_lds) {
            raise_dynamic_lds((const void*)k_tridiag, lt);
            hipLaunchKernelGGL(k_tridiag, dim3(nlist), dim3(threads), lt, st, jobs, state, joblist, arena, tri);
        } else if (max_n <= kSymMaxLdsN && get_tuning().tri_blocked) {
            const size_t ls = tridiag_sym_lds_bytes(max_n, false);
            raise_dynamic_lds((const void*)k_tridiag_blk, ls);
            hipLaunchKernelGGL(k_tridiag_blk, dim3(nlist), dim3(1024), ls, st, jobs, state, joblist, arena, tri);
        } else if (max_n <= kSymMaxLdsN) {
            const size_t ls = tridiag_sym_lds_bytes(max_n, false);
            raise_dynamic_lds((const void*)k_tridiag_sym<false>, ls);
            hipLaunchKernelGGL(k_tridiag_sym<false>, dim3(nlist), dim3(1024), ls, st, jobs, state, joblist, arena, tri);
        } else {
            const size_t ls = tridiag_sym_lds_bytes(max_n, true);
            raise_dynamic_lds((const void*)k_tridiag_sym<true>, ls);
            hipLaunchKernelGGL(k_tridiag_sym<true>, dim3(nlist), dim3(1024), ls, st, jobs, state, joblist, arena, tri);
        }
    } else if (part == 1) {
        // 512 threads above 64 AOs: the secular and deflation phases are latency
        // bound, and three 512-thread workgroups (LDS permitting) share a CU where
        // two 1024-thread ones did (208 vs 300 ms per 4096-molecule pass)
        const int tdc = max_n <= 64 ? 256 : 512;
        const size_t ld = dc_lds_bytes(max_n);
        raise_dynamic_lds((const void*)k_dc, ld);
        hipLaunchKernelGGL(k_dc, dim3(nlist), dim3(tdc), ld, st, jobs, state, joblist, arena, tri, eps, perm);
    } else if (max_n >= 3) {
        const int nblk = (max_n - 2 + kBt - 1) / kBt;
        hipLaunchKernelGGL(k_wy_t, dim3(nlist, nblk), dim3(64), 0, st, jobs, state, joblist, arena, tri);
        const Tuning tu = get_tuning();
        const int cls = max_n <= 128 ? 0 : (max_n <= 256 ? 1 : (max_n <= 384 ? 2 : 3));
        const bool r32 = max_n <= 512 && tu.bt_rows[cls] == 32;
        if (max_n <= 128) {
            if (r32) launch_bt<1, 2, false>(st, jobs, state, joblist, nlist, max_n, arena, tri);
            else launch_bt<1, 4, false>(st, jobs, state, joblist, nlist, max_n, arena, tri);
        } else if (max_n <= 256) {
            if (r32) launch_bt<2, 2, false>(st, jobs, state, joblist, nlist, max_n, arena, tri);
            else launch_bt<2, 4, false>(st, jobs, state, joblist, nlist, max_n, arena, tri);
        } else if (max_n <= 384) {
            if (r32) launch_bt<3, 2, false>(st, jobs, state, joblist, nlist, max_n, arena, tri);
            else launch_bt<3, 4, false>(st, jobs, state, joblist, nlist, max_n, arena, tri);
        } else if (max_n <= 512) {
            if (r32) launch_bt<4, 2, false>(st, jobs, state, joblist, nlist, max_n, arena, tri);
            else launch_bt<4, 4, false>(st, jobs, state, joblist, nlist, max_n, arena, tri);
        }
        else if (max_n <= 768) launch_bt<6, 2, true>(st, jobs, state, joblist, nlist, max_n, arena, tri);
        else if (max_n <= 1024) launch_bt<8, 2, true>(st, jobs, state, joblist, nlist, max_n, arena, tri);
        else launch_bt<10, 1, true>(st, jobs, state, joblist, nlist, max_n, arena, tri);
    }
}

void launch_wy_t(hipStream_t st, const JobDev* jobs, JobState* state, const int* joblist, int nlist, int max_n,
                 double* arena, double* tri) {
    if (nlist <= 0 || max_n < 3) return;
    const int nblk = (max_n - 2 + kBt - 1) / kBt;
    hipLaunchKernelGGL(k_wy_t, dim3(nlist, nblk), dim3(64), 0, st, jobs, state, joblist, arena, tri);
}

void launch_heig_class(hipStream_t st, const JobDev* jobs, JobState* state, const int* joblist, int nlist, int max_n,
                       double* arena, double* eps, int* perm, double* tri, bool a_in_lds) {
    for (int part = 0; part < 3; ++part)
        launch_heig_part(part, st, jobs, state, joblist, nlist, max_n, arena, eps, perm, tri, a_in_lds);
}

}  // namespace nddo
