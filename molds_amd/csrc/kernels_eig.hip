// FP64 symmetric eigensolver for the SCF loop, one workgroup per molecule:
//   1. Householder tridiagonalization T = Q^T F Q (Golub & Van Loan, Alg. 8.3.1),
//      with each step's rank-2 update fused with the next step's matrix-vector
//      product, so the trailing block is read and written once per step;
//   2. implicit QL with Wilkinson shifts on T (Bowdler, Martin, Reinsch &
//      Wilkinson, Numer. Math. 11, 293 (1968)): one thread runs the scalar
//      recurrence and records the Givens rotations of each step in an LDS ring,
//      then every thread applies them to its own column of the eigenvector
//      rows (coalesced: thread r owns element r of every row);
//   3. back-transformation of the eigenvector rows by the stored reflectors,
//      staged through LDS in blocks.
// It replaces LAPACKE_dsyevd (src/wrappers/Lapack.cpp:102-161).  Output: rows of
// AR_C are eigenvectors (unsorted), eps[] ascending and perm[] the row of each
// rank, exactly as the density kernel consumes them.  Matrices up to the LDS
// budget are tridiagonalized in LDS; larger ones in the arena (AR_W).
#include <hip/hip_runtime.h>
#include <math.h>

#include "launch.h"
#include "nddo_device.h"

namespace nddo {

namespace {

__device__ inline double* amat(double* arena, const JobDev& jb, int kind) {
    return arena + jb.arena + (long long)kind * jb.nao * jb.nao;
}

constexpr int kRec = 2048;       // rotation record ring (doubles)
constexpr int kBlk = 16;         // reflectors staged per back-transform block

// Block-wide sum; every thread gets the result.  red has >= 17 doubles.
__device__ inline double block_sum(double v, double* red, int tid, int nw) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    __syncthreads();
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += red[w];
    return s;
}

}  // namespace

size_t heig_lds_bytes(int n, bool a_in_lds) {
    size_t d = 8 * (size_t)n + kRec;        // d e tau vcur vnext w p pnext + rotation ring
    if (a_in_lds) d += (size_t)n * n;
    else d += (size_t)kBlk * n;             // reflector staging
    return d * 8;
}

template <bool ALDS, int NPL>
__global__ void __launch_bounds__(1024) k_heig(const JobDev* __restrict__ jobs, JobState* __restrict__ state,
                                               const int* __restrict__ joblist, double* __restrict__ arena,
                                               double* __restrict__ eps, int* __restrict__ perm) {
    extern __shared__ double dyn[];
    __shared__ double red[32];
    __shared__ int sh[4];
    const int job = joblist[blockIdx.x];
    const JobDev jb = jobs[job];
    if (state[job].phase != PH_ACTIVE) return;
    const int n = jb.nao;
    const int T = blockDim.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = T >> 6;
    double* d = dyn;
    double* e = d + n;
    double* tau = e + n;
    double* vcur = tau + n;
    double* vnext = vcur + n;
    double* w = vnext + n;
    double* p = w + n;
    double* pnext = p + n;
    double* rec = pnext + n;
    double* stage = rec + kRec;
    double* A = ALDS ? stage : amat(arena, jb, AR_W);
    const double* F = amat(arena, jb, AR_F);
    double* ZT = amat(arena, jb, AR_C);
    const long long nn = (long long)n * n;

    const unsigned long long t0 = wall_clock64();
    for (long long x = tid; x < nn; x += T) {
        A[x] = F[x];
        const int i = (int)(x / n), j = (int)(x - (long long)i * n);
        ZT[x] = i == j ? 1.0 : 0.0;
    }
    for (int i = tid; i < n; i += T) e[i] = 0.0;
    __syncthreads();

    // ---------------------------------------------------------------- 1. tridiagonalize
    // Reflector k acts on indices k+1..n-1: H_k = I - tau_k v v^T, v[0] = 1,
    // stored in row k of A (columns k+1..n-1).  Vectors below are indexed from
    // k+1 (t = index - k - 1).
    if (n >= 3) {
        // reflector of row 0 and p = tau B v for B = A[1:, 1:]
        {
            const int m = n - 1;
            double ss = 0.0;
            for (int t = 1 + tid; t < m; t += T) { const double x = A[1 + t]; ss += x * x; }
            ss = block_sum(ss, red, tid, nw);
            const double alpha = A[1];
            __syncthreads();   // every wave has read alpha before A[1] is overwritten below
            double beta = alpha, tk = 0.0, scale = 0.0;
            if (ss > 0.0) {
                beta = -copysign(sqrt(alpha * alpha + ss), alpha);
                tk = (beta - alpha) / beta;
                scale = 1.0 / (alpha - beta);
            }
            for (int t = tid; t < m; t += T) {
                const double v = t == 0 ? 1.0 : (ss > 0.0 ? A[1 + t] * scale : 0.0);
                vcur[t] = v;
                A[1 + t] = v;
            }
            if (tid == 0) { d[0] = A[0]; e[0] = beta; tau[0] = tk; }
            __syncthreads();
            for (int i = 1 + wave; i < n; i += nw) {
                double s = 0.0;
                for (int j = 1 + lane; j < n; j += 64) s += A[(long long)i * n + j] * vcur[j - 1];
                for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
                if (lane == 0) p[i - 1] = tk * s;
            }
            __syncthreads();
        }
        for (int k = 0; k <= n - 3; ++k) {
            const int m = n - 1 - k;
            const double tk = tau[k];
            double pv = 0.0;
            for (int t = tid; t < m; t += T) pv += p[t] * vcur[t];
            pv = block_sum(pv, red, tid, nw);
            const double K = -0.5 * tk * pv;
            for (int t = tid; t < m; t += T) w[t] = p[t] + K * vcur[t];
            __syncthreads();
            if (k == n - 3) {
                // last trailing 2x2 block (rows/cols n-2, n-1)
                if (tid == 0) {
                    const double a00 = A[(long long)(n - 2) * n + n - 2] - 2.0 * vcur[0] * w[0];
                    const double a01 = A[(long long)(n - 2) * n + n - 1] - (vcur[0] * w[1] + w[0] * vcur[1]);
                    const double a11 = A[(long long)(n - 1) * n + n - 1] - 2.0 * vcur[1] * w[1];
                    d[n - 2] = a00;
                    e[n - 2] = a01;
                    d[n - 1] = a11;
                }
                __syncthreads();
                break;
            }
            // row k+1 (t = 0) of the updated trailing block
            const long long r1 = (long long)(k + 1) * n;
            for (int t = tid; t < m; t += T) A[r1 + k + 1 + t] -= vcur[0] * w[t] + w[0] * vcur[t];
            __syncthreads();
            // reflector of row k+1: x = A[k+1][k+2 .. n-1]
            const int m2 = m - 1;
            double ss = 0.0;
            for (int t = 1 + tid; t < m2; t += T) { const double x = A[r1 + k + 2 + t]; ss += x * x; }
            ss = block_sum(ss, red, tid, nw);
            const double alpha = A[r1 + k + 2];
            __syncthreads();   // every wave has read alpha before it is overwritten below
            double beta = alpha, tn = 0.0, scale = 0.0;
            if (ss > 0.0) {
                beta = -copysign(sqrt(alpha * alpha + ss), alpha);
                tn = (beta - alpha) / beta;
                scale = 1.0 / (alpha - beta);
            }
            for (int t = tid; t < m2; t += T) {
                const double v = t == 0 ? 1.0 : (ss > 0.0 ? A[r1 + k + 2 + t] * scale : 0.0);
                vnext[t] = v;
                A[r1 + k + 2 + t] = v;
            }
            if (tid == 0) { d[k + 1] = A[r1 + k + 1]; e[k + 1] = beta; tau[k + 1] = tn; }
            __syncthreads();
            // fused pass over rows k+2 .. n-1: rank-2 update of step k, then p for step k+1
            for (int i = k + 2 + wave; i < n; i += nw) {
                const int ti = i - k - 1;
                const double vi = vcur[ti], wi = w[ti];
                const long long ri = (long long)i * n;
                double s = 0.0;
                for (int j = k + 2 + lane; j < n; j += 64) {
                    const int tj = j - k - 1;
                    const double a = A[ri + j] - (vi * w[tj] + wi * vcur[tj]);
                    A[ri + j] = a;
                    s += a * vnext[tj - 1];
                }
                for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
                if (lane == 0) pnext[i - k - 2] = tn * s;
            }
            __syncthreads();
            double* tmp = vcur; vcur = vnext; vnext = tmp;
            tmp = p; p = pnext; pnext = tmp;
        }
    } else {
        if (tid == 0) {
            d[0] = A[0];
            if (n == 2) { e[0] = A[1]; d[1] = A[3]; }
        }
        __syncthreads();
    }

    const unsigned long long t1 = wall_clock64();
    unsigned long long tql = 0, tapply = 0;
    // ---------------------------------------------------------------- 2. implicit QL
    // e[i] couples i and i+1; e[n-1] = 0.  Thread 0 runs the recurrence and
    // records (m, count, c, s, ...) per QL step; the block applies a full ring.
    if (tid == 0) { sh[0] = 0; sh[1] = 0; sh[2] = 0; sh[3] = 0; }   // l, done, failed, qlsteps
    __syncthreads();
    int l_state = 0;
    while (true) {
        const unsigned long long ta = wall_clock64();
        if (tid == 0) {
            int pos = 0;
            int l = sh[0];
            bool failed = false;
            int steps = sh[3];
            while (l < n) {
                int it = 0;
                bool deflated = false;
                while (true) {
                    int m = l;
                    for (; m < n - 1; ++m) {
                        const double dd = fabs(d[m]) + fabs(d[m + 1]);
                        if (fabs(e[m]) <= 2.220446049250313e-16 * dd) break;
                    }
                    if (m == l) { deflated = true; break; }
                    // the ring's last slot holds its used length; n <= 512 keeps a step < kRec - 1
                    if (pos + 2 + 2 * (m - l) > kRec - 1) break;     // ring full: apply first
                    if (++it > 60) { failed = true; break; }
                    ++steps;
                    double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                    double r = hypot(g, 1.0);
                    g = d[m] - d[l] + e[l] / (g + copysign(r, g));
                    double s = 1.0, c = 1.0, pp = 0.0;
                    int i = m - 1, cnt = 0;
                    const int hdr = pos;
                    pos += 2;
                    bool underflow = false;
                    // d[i+1] rides in a register (it is d[i] of the previous index);
                    // e[i], d[i] are fetched one index ahead of their use
                    double dip1 = d[m], ei = e[i], di = d[i];
                    for (; i >= l; --i) {
                        const double ein = i > l ? e[i - 1] : 0.0, din = i > l ? d[i - 1] : 0.0;
                        const double f = s * ei, b = c * ei;
                        r = sqrt(f * f + g * g);
                        e[i + 1] = r;
                        if (r == 0.0) { d[i + 1] = dip1 - pp; e[m] = 0.0; underflow = true; break; }
                        const double ir = 1.0 / r;
                        s = f * ir;
                        c = g * ir;
                        g = dip1 - pp;
                        r = (di - g) * s + 2.0 * c * b;
                        pp = s * r;
                        d[i + 1] = g + pp;
                        g = c * r - b;
                        rec[pos] = c;
                        rec[pos + 1] = s;
                        pos += 2;
                        ++cnt;
                        dip1 = di;
                        ei = ein;
                        di = din;
                    }
                    rec[hdr] = (double)m;
                    rec[hdr + 1] = (double)cnt;
                    if (underflow) continue;
                    d[l] -= pp;
                    e[l] = g;
                    e[m] = 0.0;
                }
                if (failed || !deflated) break;
                ++l;
            }
            sh[0] = l;
            sh[1] = (l >= n || failed) ? 1 : 0;
            sh[2] = failed ? 1 : 0;
            sh[3] = steps;
            rec[kRec - 1] = (double)pos;   // used length (the ring never fills its last slot)
        }
        __syncthreads();
        const unsigned long long tb = wall_clock64();
        tql += tb - ta;
        const int used = (int)rec[kRec - 1];
        // apply the recorded rotations: thread r owns column r of the eigenvector rows
        for (int r = tid; r < n; r += T) {
            int q = 0;
            while (q < used) {
                const int m = (int)rec[q], cnt = (int)rec[q + 1];
                q += 2;
                double val = ZT[(long long)m * n + r];
                double zi = cnt > 0 ? ZT[(long long)(m - 1) * n + r] : 0.0;
                for (int t = 0; t < cnt; ++t) {
                    const double c = rec[q + 2 * t], s = rec[q + 2 * t + 1];
                    const int i = m - 1 - t;
                    // fetch the next row's element before this step's store
                    const double zn = t + 1 < cnt ? ZT[(long long)(i - 1) * n + r] : 0.0;
                    ZT[(long long)(i + 1) * n + r] = s * zi + c * val;
                    val = c * zi - s * val;
                    zi = zn;
                }
                ZT[(long long)(m - cnt) * n + r] = val;
                q += 2 * cnt;
            }
        }
        __syncthreads();
        tapply += wall_clock64() - tb;
        const int done = sh[1];
        __syncthreads();   // all threads read the flag before thread 0 starts the next segment
        if (done) break;
        ++l_state;
    }
    const bool failed = sh[2] != 0;

    // ---------------------------------------------------------------- 3. back-transform
    // row r of the eigenvectors <- row r * H_{n-3} * ... * H_0, reflectors in
    // blocks of kBlk staged in LDS (or read in place when A lives in LDS);
    // each wavefront holds one row in registers.
    if (n >= 3) {
        const int nref = n - 2;
        for (int r0 = 0; r0 < n; r0 += nw) {
            const int r = r0 + wave;
            double row[NPL];
#pragma unroll
            for (int u = 0; u < NPL; ++u) {
                const int j = lane + 64 * u;
                row[u] = (r < n && j < n) ? ZT[(long long)r * n + j] : 0.0;
            }
            for (int kb = nref - 1; kb >= 0; kb -= kBlk) {
                const int kl = kb - kBlk + 1 > 0 ? kb - kBlk + 1 : 0;   // block [kl, kb], applied kb first
                const double* vb;
                if (ALDS) {
                    vb = A;
                } else {
                    __syncthreads();
                    for (int x = tid; x < (kb - kl + 1) * n; x += T) {
                        const int kk = kl + x / n, j = x % n;
                        stage[x] = j > kk ? A[(long long)kk * n + j] : 0.0;
                    }
                    __syncthreads();
                    vb = stage;
                }
                for (int k = kb; k >= kl; --k) {
                    const double* v = ALDS ? vb + (long long)k * n : vb + (long long)(k - kl) * n;
                    double s = 0.0;
#pragma unroll
                    for (int u = 0; u < NPL; ++u) {
                        const int j = lane + 64 * u;
                        if (j > k && j < n) s += row[u] * v[j];
                    }
                    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
                    s *= tau[k];
#pragma unroll
                    for (int u = 0; u < NPL; ++u) {
                        const int j = lane + 64 * u;
                        if (j > k && j < n) row[u] -= s * v[j];
                    }
                }
            }
            if (r < n) {
#pragma unroll
                for (int u = 0; u < NPL; ++u) {
                    const int j = lane + 64 * u;
                    if (j < n) ZT[(long long)r * n + j] = row[u];
                }
            }
        }
    }
    __syncthreads();
    // eigenvalues ascending (ties by index) and the row permutation
    for (int pidx = tid; pidx < n; pidx += T) {
        const double dv = d[pidx];
        int rank = 0;
        for (int q = 0; q < n; ++q) {
            const double ev = d[q];
            rank += (ev < dv) || (ev == dv && q < pidx);
        }
        eps[jb.vec0 + rank] = dv;
        perm[jb.ivec0 + rank] = pidx;
    }
    if (tid == 0) {
        state[job].sweeps = failed ? -1 : (sh[3] > 0 ? sh[3] : 1);
        state[job].warm = 0;
        const unsigned long long t3 = wall_clock64();
        state[job].prof[0] = (double)(t1 - t0);
        state[job].prof[1] = (double)tql;
        state[job].prof[2] = (double)tapply;
        state[job].prof[3] = (double)(t3 - t1 - tql - tapply);
    }
}

template <bool ALDS, int NPL>
static void launch_heig(hipStream_t st, const JobDev* jobs, JobState* state, const int* joblist, int nlist,
                        int threads, double* arena, double* eps, int* perm, size_t lds) {
    hipFuncSetAttribute((const void*)k_heig<ALDS, NPL>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((k_heig<ALDS, NPL>), dim3(nlist), dim3(threads), lds, st, jobs, state, joblist, arena, eps,
                       perm);
}

bool heig_supported(int n) { return n <= 512; }

void launch_heig_class(hipStream_t st, const JobDev* jobs, JobState* state, const int* joblist, int nlist, int max_n,
                       double* arena, double* eps, int* perm, bool a_in_lds) {
    if (nlist <= 0) return;
    const int threads = max_n <= 32 ? 256 : (max_n <= 96 ? 512 : 1024);
    const size_t lds = heig_lds_bytes(max_n, a_in_lds);
    if (a_in_lds) {
        if (max_n <= 64) launch_heig<true, 1>(st, jobs, state, joblist, nlist, threads, arena, eps, perm, lds);
        else launch_heig<true, 2>(st, jobs, state, joblist, nlist, threads, arena, eps, perm, lds);
    } else {
        if (max_n <= 128) launch_heig<false, 2>(st, jobs, state, joblist, nlist, threads, arena, eps, perm, lds);
        else if (max_n <= 256) launch_heig<false, 4>(st, jobs, state, joblist, nlist, threads, arena, eps, perm, lds);
        else if (max_n <= 384) launch_heig<false, 6>(st, jobs, state, joblist, nlist, threads, arena, eps, perm, lds);
        else launch_heig<false, 8>(st, jobs, state, joblist, nlist, threads, arena, eps, perm, lds);
    }
}

}  // namespace nddo
