// Host side of the NDDO SCF engine: batch assembly, the SCF driver loop over
// batched kernel launches, result extraction and the extern "C" boundary
// declared in include/nddo.h.
//
// The driver mirrors Cndo2::DoSCF (src/cndo/Cndo2.cpp:585-729): per-geometry
// setup (overlap, two-centre integrals, core repulsion), then per iteration
// Fock -> eigensolve -> density -> convergence test -> DIIS -> damping, and at
// convergence the electronic energy from a Fock matrix rebuilt with the final
// density (Cndo2.cpp:732-754, 1324-1416).  All of it runs on the device; the
// host only reads one counter per iteration to know when every job is done.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <cstdlib>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "nddo.h"
#include "engine_internal.h"
#include "launch.h"
#include "nddo_device.h"
#include "params.h"

using namespace nddo;

namespace {
thread_local std::string g_err;
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
}

// Device scratch of nddo_scf_once_noalloc: every buffer of its one-molecule
// batch is carved out of this array, so that entry point allocates nothing.
// Each GPU the library runs on holds its own copy until the process ends; host
// code uses only the address hipGetSymbolAddress gives for the current device
// (Engine::d_scratch), never the array's name.
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

void nddo::set_error(const std::string& s) { g_err = s; }
std::string nddo::get_error() { return g_err; }

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            nddo::set_error(std::string("HIP error ") + hipGetErrorString(e_) + " in " #x); \
            return NDDO_HIP_ERROR;                                                         \
        }                                                                                  \
    } while (0)

namespace nddo {

static const int kElemZ[kNumElem] = {1, 6, 7, 8, 9};

int elem_slot(int Z) {
    for (int i = 0; i < kNumElem; ++i)
        if (kElemZ[i] == Z) return i;
    return -1;
}

// Overlap integrand polynomials in prolate spheroidal coordinates
// (Pople & Beveridge, Approximate MO Theory, App. B): for orbitals
// (n_a, l_a) on A at z = -R/2 and (n_b, l_b) on B at z = +R/2, same |m|,
//   r_a = R/2 (xi+eta), r_b = R/2 (xi-eta), z - z_A = R/2 (1 + xi eta),
//   z - z_B = R/2 (xi eta - 1), rho^2 = (R/2)^2 (xi^2-1)(1-eta^2),
//   dV = (R/2)^3 (xi^2 - eta^2) dxi deta dphi,
// so S = N_a N_b (R/2)^(na+nb+1) * ang * sum_ij c_ij A_i(p) B_j(q).
static void poly_mul(double a[5][5], const double b[5][5]) {
    double r[5][5] = {};
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j)
            if (a[i][j] != 0.0)
                for (int k = 0; i + k < 5; ++k)
                    for (int l = 0; j + l < 5; ++l) r[i + k][j + l] += a[i][j] * b[k][l];
    std::memcpy(a, r, sizeof(r));
}

static void overlap_table(double* tab) {
    std::memset(tab, 0, sizeof(double) * 32 * 25);
    const double pi = 3.14159265358979323846;
    double plus[5][5] = {}, minus[5][5] = {}, zA[5][5] = {}, zB[5][5] = {}, perp[5][5] = {}, vol[5][5] = {};
    plus[1][0] = 1; plus[0][1] = 1;
    minus[1][0] = 1; minus[0][1] = -1;
    zA[0][0] = 1; zA[1][1] = 1;
    zB[1][1] = 1; zB[0][0] = -1;
    perp[2][0] = 1; perp[2][2] = -1; perp[0][0] = -1; perp[0][2] = 1;
    vol[2][0] = 1; vol[0][2] = -1;
    for (int na = 1; na <= 2; ++na)
        for (int la = 0; la < na; ++la)
            for (int nb = 1; nb <= 2; ++nb)
                for (int lb = 0; lb < nb; ++lb)
                    for (int m = 0; m <= std::min(la, lb); ++m) {
                        double p[5][5] = {};
                        p[0][0] = 1.0;
                        for (int i = 0; i < na - 1 - la; ++i) poly_mul(p, plus);
                        for (int i = 0; i < nb - 1 - lb; ++i) poly_mul(p, minus);
                        double ang = (la ? std::sqrt(3.0 / (4 * pi)) : 1.0 / std::sqrt(4 * pi)) *
                                     (lb ? std::sqrt(3.0 / (4 * pi)) : 1.0 / std::sqrt(4 * pi));
                        if (m == 0) {
                            if (la) poly_mul(p, zA);
                            if (lb) poly_mul(p, zB);
                            ang *= 2 * pi;
                        } else {
                            poly_mul(p, perp);
                            ang *= pi;
                        }
                        poly_mul(p, vol);
                        int cid = ((((na - 1) * 2 + la) * 2 + (nb - 1)) * 2 + lb) * 2 + m;
                        for (int i = 0; i < 5; ++i)
                            for (int j = 0; j < 5; ++j) tab[cid * 25 + i * 5 + j] = ang * p[i][j];
                    }
}

size_t eig_lds_bytes(int n, bool a_in_lds) { return heig_supported(n) ? heig_lds_bytes(n, a_in_lds) : 0; }

// ------------------------------------------------------------------ batch
Batch::~Batch() { release(); }

void Batch::release() {
    if (gexec) hipGraphExecDestroy(gexec);
    gexec = nullptr;
    iterations_launched = 0;
    graph_off = false;
    std::vector<std::pair<void**, size_t>> req;
    plan(true, true, &req);
    for (auto& r : req) {
        void** p = r.first;
        if (*p && !carved) {
            if (alloc_stream) hipFreeAsync(*p, alloc_stream);
            else hipFree(*p);
        }
        *p = nullptr;
    }
    carved = false;
}

// Stream-ordered allocations (served cohorts) do not synchronize the device.
int Batch::dalloc(void** p, size_t bytes) {
    const hipError_t err = alloc_stream ? hipMallocAsync(p, bytes, alloc_stream) : hipMalloc(p, bytes);
    if (err != hipSuccess) {
        set_error(std::string("HIP error ") + hipGetErrorString(err) + " in device allocation");
        *p = nullptr;
        return NDDO_HIP_ERROR;
    }
    return NDDO_OK;
}

int validate_job(const nddo_job_desc& d, int* nao, int* nocc) {
    if (d.natoms <= 0 || !d.Z || !d.xyz_bohr || d.theory < 0 || d.theory > 2) {
        set_error("bad job descriptor");
        return NDDO_BAD_ARG;
    }
    if (d.opts.diis_num_error_vect < 0 || d.opts.diis_num_error_vect > NDDO_MAX_DIIS || d.opts.max_iter <= 0) {
        set_error("bad SCF options (diis_num_error_vect must be 0..16, max_iter > 0)");
        return NDDO_BAD_ARG;
    }
    int n = 0, ne = 0;
    for (int a = 0; a < d.natoms; ++a) {
        if (elem_slot(d.Z[a]) < 0) {
            set_error("Error in cndo::Cndo2::CheckEnableAtomType: Non available atom is contained.\n(Z=" +
                      std::to_string(d.Z[a]) + ")");
            return NDDO_UNSUPPORTED_ELEMENT;
        }
        ElemDev e;
        elem_params(d.Z[a], d.theory, &e);
        n += e.nao;
        ne += (int)e.core;
    }
    if (ne % 2 != 0) {
        set_error("Error in cndo::Cndo2::SetMolecule: Total number of valence electrons is odd. "
                  "totalNumberValenceElectrons=" + std::to_string(ne));
        return NDDO_ODD_ELECTRONS;
    }
    for (int a = 0; a < 3 * d.natoms; ++a)
        if (!std::isfinite(d.xyz_bohr[a])) {
            set_error("non-finite coordinate");
            return NDDO_BAD_ARG;
        }
    // coincident atoms have no diatomic frame (R = 0 in every pair formula)
    for (int a = 0; a < d.natoms; ++a)
        for (int b = a + 1; b < d.natoms; ++b) {
            const double dx = d.xyz_bohr[3 * a] - d.xyz_bohr[3 * b], dy = d.xyz_bohr[3 * a + 1] - d.xyz_bohr[3 * b + 1],
                         dz = d.xyz_bohr[3 * a + 2] - d.xyz_bohr[3 * b + 2];
            if (dx * dx + dy * dy + dz * dz < 1e-12) {
                set_error("atoms " + std::to_string(a) + " and " + std::to_string(b) + " coincide");
                return NDDO_BAD_ARG;
            }
        }
    *nao = n;
    *nocc = ne / 2;
    return NDDO_OK;
}

size_t job_footprint_bytes(int natoms, int nao, int diis_k, int max_iter) {
    const size_t n2 = (size_t)nao * nao;
    const size_t npairs = (size_t)natoms * (natoms - 1) / 2;
    size_t b = (ARENA_FIXED + 2 * (size_t)diis_k) * n2 * 8;   // arena
    // integrals (upper bound), partials, ecore, the pair records
    b += npairs * (100 * 8 + 20 * 8 + 8 + sizeof(PairDev) + sizeof(int) + sizeof(FockPairDev));
    const size_t tpd = ((size_t)nao + 63) / 64;
    b += tpd * tpd * (sizeof(TileDev) + (1 + XL_REC) * 8);           // density tiles and their partials
    b += (size_t)nao * 12 + (size_t)max_iter * 32 + sizeof(JobDev) + sizeof(JobState) + XL_COEF * 8;
    return b;
}

// Two passes: the per-job sizes and offsets (sequential, cheap), then the
// atom, pair-row and tile records of the jobs in parallel threads (the pairs
// themselves are written on the device).
int Batch::build(const nddo_job_desc* descs, int nd, const std::vector<int>& which, const int* elem_index_base) {
    (void)elem_index_base;
    (void)nd;
    atoms.clear(); arows.clear(); npairs_all = 0; jobs.clear(); tiles.clear(); list_lds.clear(); list_glb.clear();
    list_xl.clear(); xl_off.clear();
    src = which;
    nslots = 0; nints = 0; arena_size = 0; nvec = 0; ntraj = 0;
    maxn_lds = 0; maxn_glb = 0; maxn_xl = 0; max_iter_all = 0;
    const int nj = (int)which.size();
    struct Off { long long atom0, pair0, ints0, tile0, cls[3]; };
    std::vector<Off> off(nj);
    long long natoms_all = 0, ntiles_all = 0, ncls[3] = {0, 0, 0};
    jobs.resize(nj);
    for (int jj = 0; jj < nj; ++jj) {
        const nddo_job_desc& d = descs[which[jj]];
        JobDev& jb = jobs[jj];
        std::memset(&jb, 0, sizeof(jb));
        long long nh = 0, nx = 0;
        int n = 0, ne = 0;
        for (int a = 0; a < d.natoms; ++a) {
            const int z = d.Z[a];
            const bool h = z == 1;
            nh += h;
            nx += !h;
            n += h ? 1 : 4;
            ne += h ? 1 : z - 2;   // valence electrons: H 1, C 4, N 5, O 6, F 7
        }
        jb.natoms = d.natoms;
        jb.theory = d.theory;
        jb.atom0 = (int)natoms_all;
        jb.nao = n;
        jb.nocc = ne / 2;
        jb.npairs = d.natoms * (d.natoms - 1) / 2;
        jb.pair0 = nslots;
        off[jj].atom0 = natoms_all;
        off[jj].pair0 = nslots;
        off[jj].ints0 = nints;
        off[jj].tile0 = ntiles_all;
        const long long c3[3] = {nh * (nh - 1) / 2, nh * nx, nx * (nx - 1) / 2};
        for (int c = 0; c < 3; ++c) {
            off[jj].cls[c] = ncls[c];
            ncls[c] += c3[c];
        }
        natoms_all += d.natoms;
        nslots += jb.npairs;
        nints += c3[0] + 10 * c3[1] + 100 * c3[2];
        jb.diis_k = d.opts.diis_num_error_vect;
        jb.want_c = 1;
        jb.arena = arena_size;
        arena_size += (long long)(ARENA_FIXED + 2 * jb.diis_k) * n * n;
        jb.vec0 = nvec;
        jb.ivec0 = nvec;
        nvec += n;
        jb.tile0 = (int)ntiles_all;
        jb.tiles_per_dim = (n + 63) / 64;
        jb.ntiles = jb.tiles_per_dim * jb.tiles_per_dim;
        ntiles_all += jb.ntiles;
        jb.max_iter = d.opts.max_iter;
        jb.requires_guess = d.requires_guess ? 1 : 0;
        jb.precision = d.precision;
        jb.rms_thresh = d.opts.rms_density;
        jb.damp_thresh = d.opts.damping_thresh;
        jb.damp_weight = d.opts.damping_weight;
        jb.diis_start = d.opts.diis_start_error;
        jb.diis_end = d.opts.diis_end_error;
        jb.traj0 = ntraj;
        ntraj += 4LL * d.opts.max_iter;
        max_iter_all = std::max(max_iter_all, d.opts.max_iter);
    }
    atoms.resize(natoms_all);
    arows.resize(natoms_all);
    npairs_all = nslots;
    tiles.resize(ntiles_all);
    n_light = (int)(ncls[0] + ncls[1]);
    n_heavy = (int)ncls[2];
    const long long cbase[3] = {0, ncls[0], ncls[0] + ncls[1]};
    for (int c = 0; c < 3; ++c) {
        cls_pos[c].resize(nj + 1);
        for (int jj = 0; jj < nj; ++jj) cls_pos[c][jj] = cbase[c] + off[jj].cls[c];
        cls_pos[c][nj] = cbase[c] + ncls[c];
    }
    auto fill = [&](int j0, int j1) {
        for (int jj = j0; jj < j1; ++jj) {
            const nddo_job_desc& d = descs[which[jj]];
            const JobDev& jb = jobs[jj];
            const Off& o = off[jj];
            int ao = 0;
            for (int a = 0; a < d.natoms; ++a) {
                AtomDev& at = atoms[o.atom0 + a];
                at.x = d.xyz_bohr[3 * a];
                at.y = d.xyz_bohr[3 * a + 1];
                at.z = d.xyz_bohr[3 * a + 2];
                at.elem = d.theory * kNumElem + elem_slot(d.Z[a]);
                at.job = jj;
                at.ao = ao;
                at.local = a;
                ao += d.Z[a] == 1 ? 1 : 4;
            }
            // where each pair row starts; the pairs (a, b) themselves are written on
            // the device (k_build_pairs), in the same order: slots row by row,
            // integrals packed 1 / 10 / 100 per H-H / heavy-H / heavy-heavy pair,
            // and each class of the pair list in (a, b) order
            int nh = 0;
            for (int a = 0; a < d.natoms; ++a) nh += d.Z[a] == 1;
            const int nx = d.natoms - nh;
            long long io = o.ints0, cpos[3] = {cbase[0] + o.cls[0], cbase[1] + o.cls[1], cbase[2] + o.cls[2]};
            int h_upto = 0, x_upto = 0;
            const long long N = d.natoms;
            for (int a = 0; a < d.natoms; ++a) {
                const bool ha = d.Z[a] == 1;
                h_upto += ha;
                x_upto += !ha;
                const long long h_after = nh - h_upto, x_after = nx - x_upto;
                AtomRowDev& rw = arows[o.atom0 + a];
                rw.slot0 = o.pair0 + a * N - (long long)a * (a + 1) / 2;
                rw.ioff0 = io;
                for (int c = 0; c < 3; ++c) rw.cpos[c] = cpos[c];
                io += (ha ? 1 : 10) * (h_after + 10 * x_after);
                if (ha) { cpos[0] += h_after; cpos[1] += x_after; }
                else { cpos[1] += h_after; cpos[2] += x_after; }
            }
            long long t = o.tile0;
            for (int i0 = 0; i0 < jb.nao; i0 += 64)
                for (int j0 = 0; j0 < jb.nao; j0 += 64) tiles[t++] = {jj, i0, j0, 0};
        }
    };
    const int nthreads = nj < 64 ? 1 : (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
    if (nthreads <= 1) {
        fill(0, nj);
    } else {
        std::vector<std::thread> th;
        const int per = (nj + nthreads - 1) / nthreads;
        for (int t = 0; t < nthreads; ++t) {
            const int j0 = t * per, j1 = std::min(nj, j0 + per);
            if (j0 < j1) th.emplace_back(fill, j0, j1);
        }
        for (auto& x : th) x.join();
    }
    return NDDO_OK;
}

// Eigensolver classes: the one-workgroup solvers with the matrix in LDS, the
// same in device memory, and the device-wide solver (kernels_xlh.hip, or the
// block Jacobi of kernels_xl.hip) for molecules of at least xl_min_nao AOs.
// With route_few_large, a batch with few molecules above 512 AOs (at most
// kFewLarge) sends those to the device-wide solver whatever xl_min_nao says: one workgroup would hold
// each for 50-190 ms per iteration on one CU while the rest of the GPU idles
// (measured on the 400 requests of the served 5-600-atom mix, 61 of them
// above 512 AOs: 14.4 s as one batch with one workgroup each up to 1200 AOs,
// 8.8 s with the device-wide solver above 512).  Many such molecules fill the
// GPU one workgroup each, which gives more solves/s.
constexpr int kFewLarge = 64;
// LDS kept free beside the dynamic LDS of k_tridiag (its static reduction words)
constexpr size_t kLdsStaticReserve = 1024;
bool lds_eig_class(int n, size_t lds_max) {
    return n <= 512 && lds_max > kLdsStaticReserve && eig_lds_bytes(n, true) <= lds_max - kLdsStaticReserve;
}
// Largest molecule of each LDS-class span but the last: with k_tridiag's LDS
// (n^2 + 4 n + max(threads, n) doubles) and threads (256 up to 44 AOs, 512 up
// to 96, else 1024), spans up to 44 / 64 / 96 AOs fit 8 / 4 / 2 workgroups per
// CU, the last span (up to 137 AOs on 160 KB) one.
constexpr int kLdsSpanMax[] = {44, 64, 96};
int Batch::classify(size_t lds_max, int num_cus, int xl_min_nao, bool route_few_large) {
    list_lds.clear();
    list_glb.clear();
    list_xl.clear();
    xl_off.clear();
    maxn_lds = maxn_glb = maxn_xl = 0;
    xlws_size = 0;
    std::vector<int> order(jobs.size());
    for (size_t i = 0; i < jobs.size(); ++i) order[i] = (int)i;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return jobs[a].nao > jobs[b].nao; });
    if (route_few_large && xl_min_nao > 513) {
        int nlarge = 0;
        for (const JobDev& jb : jobs) nlarge += jb.nao > 512;
        if (nlarge <= kFewLarge) xl_min_nao = 513;
    }
    for (int j : order) {
        int n = jobs[j].nao;
        // the Jacobi kernels also hold up to 8 KB of static LDS (reduction buffers)
        jobs[j].xl = 0;
        if (lds_eig_class(n, lds_max)) {
            list_lds.push_back(j);
            maxn_lds = std::max(maxn_lds, n);
        } else if (heig_supported(n) && n < xl_min_nao) {
            list_glb.push_back(j);
            maxn_glb = std::max(maxn_glb, n);
        } else {
            list_xl.push_back(j);
            jobs[j].xl = (int)list_xl.size();
            xl_off.push_back(xlws_size);
            xlws_size += std::max(xl_ws_doubles(n), xlh_ws_doubles(n));
            maxn_xl = std::max(maxn_xl, n);
        }
    }
    lds_spans.clear();
    for (int i = 0; i < (int)list_lds.size(); ++i) {
        const int n = jobs[list_lds[i]].nao;
        int cap = INT_MAX;
        for (int c : kLdsSpanMax)
            if (n <= c) { cap = c; break; }
        if (lds_spans.empty() || cap != lds_spans.back().maxn) {
            if (!lds_spans.empty()) lds_spans.back().maxn = jobs[list_lds[lds_spans.back().start]].nao;
            lds_spans.push_back(Span{i, 0, cap});
        }
        ++lds_spans.back().count;
    }
    if (!lds_spans.empty()) lds_spans.back().maxn = jobs[list_lds[lds_spans.back().start]].nao;
    xl_chunks = xl_nchunk(maxn_xl);
    xl_householder = use_xl_householder && xlh_supported(maxn_xl);
    split_job = 0;
    if (split_classes && !list_lds.empty() && list_lds.size() < jobs.size()) {
        bool prefix = true;   // the LDS-class jobs are exactly the first ones
        for (int j : list_lds) prefix &= j < (int)list_lds.size();
        if (prefix) {
            split_job = (int)list_lds.size();
            split_atom = jobs[split_job].atom0;
            split_tile = jobs[split_job].tile0;
        }
    }
    xl_grid = 1;
    for (size_t g = 0; g < list_xl.size(); g += xl_max_jobs()) {
        std::vector<int> ns;
        for (size_t i = g; i < std::min(list_xl.size(), g + xl_max_jobs()); ++i) ns.push_back(jobs[list_xl[i]].nao);
        xl_grid = std::max(xl_grid, ::nddo::xl_grid(num_cus, ns.data(), (int)ns.size()));
    }
    return NDDO_OK;
}

// Every device buffer of the batch with its size in bytes, in one fixed order.
// all = true lists the optional buffers too (release() clears every pointer).
void Batch::plan(bool want_sdbg, bool all, std::vector<std::pair<void**, size_t>>* req) {
    req->clear();
    auto add = [&](void* p, size_t bytes) { req->emplace_back((void**)p, bytes); };
    const size_t nj = std::max<size_t>(jobs.size(), 1);
    add(&d_atoms, std::max<size_t>(atoms.size(), 1) * sizeof(AtomDev));
    add(&d_pairs, std::max<long long>(npairs_all, 1) * sizeof(PairDev));
    add(&d_plist, std::max<long long>(npairs_all, 1) * sizeof(int));
    add(&d_fpairs, std::max<long long>(npairs_all, 1) * sizeof(FockPairDev));
    add(&d_arows, std::max<size_t>(arows.size(), 1) * sizeof(AtomRowDev));
    add(&d_jobs, nj * sizeof(JobDev));
    add(&d_state, nj * sizeof(JobState));
    add(&d_tiles, std::max<size_t>(tiles.size(), 1) * sizeof(TileDev));
    add(&d_ints, std::max<long long>(nints, 1) * sizeof(double));
    bool any_mixed = false;
    for (const JobDev& jb : jobs) any_mixed |= jb.precision == NDDO_MIXED;
    if (any_mixed || all) add(&d_ints32, std::max<long long>(nints, 1) * sizeof(float));
    add(&d_arena, std::max<long long>(arena_size, 1) * sizeof(double));
    add(&d_part, std::max<long long>(nslots, 1) * 20 * sizeof(double));
    add(&d_ecore, std::max<long long>(nslots, 1) * sizeof(double));
    add(&d_eps, std::max<long long>(nvec, 1) * sizeof(double));
    add(&d_perm, std::max<long long>(nvec, 1) * sizeof(int));
    add(&d_tri, (EIG_WS_PER_AO * std::max<long long>(nvec, 1) + EIG_WS_PER_JOB * (long long)nj) * sizeof(double));
    add(&d_rmspart, std::max<size_t>(tiles.size(), 1) * sizeof(double));
    add(&d_tpart, std::max<size_t>(tiles.size(), 1) * XL_REC * sizeof(double));
    add(&d_xlcoef, nj * XL_COEF * sizeof(double));
    add(&d_traj, std::max<long long>(ntraj, 1) * sizeof(double));
    add(&d_remaining, sizeof(int));
    add(&d_list_lds, std::max<size_t>(list_lds.size(), 1) * sizeof(int));
    add(&d_list_glb, std::max<size_t>(list_glb.size(), 1) * sizeof(int));
    if (!list_xl.empty() || all) {
        add(&d_list_xl, std::max<size_t>(list_xl.size(), 1) * sizeof(int));
        add(&d_xl_off, std::max<size_t>(list_xl.size(), 1) * sizeof(long long));
        add(&d_xlws, std::max<long long>(xlws_size, 1) * sizeof(double));
        add(&d_xlpart, std::max<size_t>(list_xl.size(), 1) * (size_t)xl_chunks * XL_REC * sizeof(double));
        add(&d_bar, std::max<size_t>(2 * sizeof(unsigned), xlh_bar_bytes()));
    }
    if (want_sdbg || all) add(&d_sdbg, std::max<long long>(arena_size, 1) * sizeof(double));
    if (!list_dip.empty() || all) {
        add(&d_list_dip, std::max<size_t>(list_dip.size(), 1) * sizeof(int));
        add(&d_dip_off, std::max<size_t>(list_dip.size(), 1) * sizeof(long long));
        add(&d_dippart, std::max<long long>(dip_size, 1) * sizeof(double));
    }
    if (any_forces || all) {
        add(&d_grad, std::max<long long>(nslots, 1) * 4 * sizeof(double));
        add(&d_forces, std::max<size_t>(atoms.size(), 1) * 3 * sizeof(double));
    }
}

void Batch::set_force_jobs(const std::vector<int>& jj) {
    for (JobDev& jb : jobs) jb.want_f = 0;
    for (int j : jj) jobs[j].want_f = 1;
    any_forces = !jj.empty();
}

void Batch::launch_properties(Engine* e, hipStream_t st) {
    if (!list_dip.empty())
        launch_dipole(st, d_jobs, d_state, d_list_dip, d_dip_off, (int)list_dip.size(), maxn_dip, d_atoms, e->d_elems,
                      d_arena, d_dippart);
    if (any_forces)
        launch_forces(st, d_pairs, nslots, d_atoms, (int)atoms.size(), e->d_elems, d_jobs, d_state, d_arena, d_grad,
                      d_forces);
}

// The jobs (batch positions) whose caller wants the dipole moment.
void Batch::set_dipole_jobs(const std::vector<int>& jj) {
    list_dip = jj;
    dip_off.clear();
    dip_size = 0;
    maxn_dip = 0;
    for (int j : jj) {
        dip_off.push_back(dip_size);
        dip_size += 4LL * jobs[j].nao;
        maxn_dip = std::max(maxn_dip, jobs[j].nao);
    }
}

size_t Batch::device_bytes(bool want_sdbg) {
    std::vector<std::pair<void**, size_t>> req;
    plan(want_sdbg, false, &req);
    size_t b = 0;
    for (auto& r : req) b += align_up(r.second);
    return b;
}

// With `region` set the buffers are carved out of it in plan order (nothing is
// allocated or freed); otherwise each comes from hipMalloc, or from the
// stream-ordered pool when alloc_stream is set.
int Batch::allocate(bool want_sdbg) {
    release();
    std::vector<std::pair<void**, size_t>> req;
    plan(want_sdbg, false, &req);
    if (region) {
        size_t off = 0;
        for (auto& r : req) {
            const size_t b = align_up(r.second);
            if (off + b > region_cap) {
                for (auto& q : req) *q.first = nullptr;
                set_error("the batch needs " + std::to_string(device_bytes(want_sdbg)) +
                          " device bytes; its region holds " + std::to_string(region_cap));
                return NDDO_ADMISSION_DENIED;
            }
            *r.first = region + off;
            off += b;
        }
        carved = true;
        return NDDO_OK;
    }
    for (auto& r : req) {
        const int rc = dalloc(r.first, r.second);
        if (rc != NDDO_OK) return rc;
    }
    return NDDO_OK;
}

// Staging bytes upload() needs; independent of classify(): every job sits in
// exactly one of the three eigensolver lists.
size_t Batch::upload_bytes() const {
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    return up(atoms.size() * sizeof(AtomDev)) + up(arows.size() * sizeof(AtomRowDev)) +
           up(jobs.size() * sizeof(JobDev)) +
           up(tiles.size() * sizeof(TileDev)) + 3 * up(jobs.size() * sizeof(int)) +
           up(jobs.size() * sizeof(long long)) + up(jobs.size() * sizeof(JobState)) + 256;
}

// With a pinned staging buffer (stage, at least upload_bytes()), every copy is
// asynchronous for the host; without one the copies come from pageable memory.
int Batch::upload(hipStream_t st, const nddo_job_desc* descs, char* stage) {
    size_t off = 0;
    const size_t cap = stage ? upload_bytes() : 0;
    auto h2d = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
        if (!bytes) return hipSuccess;
        if (!stage) return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
        if (off + bytes > cap) return hipErrorInvalidValue;   // staging smaller than upload_bytes()
        std::memcpy(stage + off, src, bytes);
        const hipError_t err = hipMemcpyAsync(dst, stage + off, bytes, hipMemcpyHostToDevice, st);
        off += (bytes + 255) & ~(size_t)255;
        return err;
    };
    HIPCHK(h2d(d_atoms, atoms.data(), atoms.size() * sizeof(AtomDev)));
    HIPCHK(h2d(d_arows, arows.data(), arows.size() * sizeof(AtomRowDev)));
    HIPCHK(h2d(d_jobs, jobs.data(), jobs.size() * sizeof(JobDev)));
    HIPCHK(h2d(d_tiles, tiles.data(), tiles.size() * sizeof(TileDev)));
    HIPCHK(h2d(d_list_lds, list_lds.data(), list_lds.size() * sizeof(int)));
    HIPCHK(h2d(d_list_glb, list_glb.data(), list_glb.size() * sizeof(int)));
    if (!list_xl.empty()) {
        HIPCHK(h2d(d_list_xl, list_xl.data(), list_xl.size() * sizeof(int)));
        HIPCHK(h2d(d_xl_off, xl_off.data(), xl_off.size() * sizeof(long long)));
    }
    if (!list_dip.empty()) {
        HIPCHK(h2d(d_list_dip, list_dip.data(), list_dip.size() * sizeof(int)));
        HIPCHK(h2d(d_dip_off, dip_off.data(), dip_off.size() * sizeof(long long)));
    }
    HIPCHK(hipMemsetAsync(d_arena, 0, std::max<long long>(arena_size, 1) * sizeof(double), st));
    HIPCHK(hipMemsetAsync(d_traj, 0, std::max<long long>(ntraj, 1) * sizeof(double), st));
    if (d_sdbg) HIPCHK(hipMemsetAsync(d_sdbg, 0, std::max<long long>(arena_size, 1) * sizeof(double), st));
    std::vector<JobState> init(jobs.size());
    std::memset(init.data(), 0, init.size() * sizeof(JobState));
    for (size_t j = 0; j < jobs.size(); ++j) {
        init[j].phase = PH_ACTIVE;
        init[j].fp32 = jobs[j].precision == NDDO_MIXED ? 1 : 0;
        const nddo_job_desc& d = descs[src[j]];
        if (!d.requires_guess && d.P0) {
            const long long n = jobs[j].nao;
            HIPCHK(hipMemcpyAsync(d_arena + jobs[j].arena + (long long)AR_P0 * n * n, d.P0, n * n * sizeof(double),
                                  hipMemcpyHostToDevice, st));
        }
    }
    HIPCHK(h2d(d_state, init.data(), init.size() * sizeof(JobState)));
    const int rem = (int)jobs.size();
    HIPCHK(h2d(d_remaining, &rem, sizeof(int)));
    if (!stage) HIPCHK(hipStreamSynchronize(st));   // pageable sources live on this stack frame
    return NDDO_OK;
}

void Batch::launch_setup_stage(Engine* e, hipStream_t st) {
    launch_build_pairs(st, d_atoms, d_arows, (int)atoms.size(), d_jobs, e->d_elems, d_pairs, d_plist, d_fpairs);
    launch_setup(st, d_pairs, d_plist, n_light, n_heavy, d_atoms, (int)atoms.size(), e->d_elems, d_jobs,
                 (int)jobs.size(), d_ints, d_arena, d_part, d_ecore, d_state, d_sdbg);
    if (d_ints32) launch_ints_f32(st, d_ints, d_ints32, nints);
}

// The pairs of jobs [j0, j1) as pair-list ranges.
FockPairs Batch::fock_pairs(int j0, int j1) const {
    FockPairs fp;
    if (jobs.empty() || j1 <= j0) return fp;
    for (int c = 0; c < 2; ++c) {
        fp.light[c] = d_fpairs + cls_pos[c][j0];
        fp.nlight[c] = (int)(cls_pos[c][j1] - cls_pos[c][j0]);
    }
    fp.heavy = d_fpairs + cls_pos[2][j0];
    fp.nheavy = (int)(cls_pos[2][j1] - cls_pos[2][j0]);
    return fp;
}

void Batch::launch_lds_eigen(int part, hipStream_t st) {
    // the tridiagonalization span by span (its LDS grows as n^2); divide and
    // conquer and the back-transform once over the class
    if (part <= 0)
        for (const Span& sp : lds_spans)
            launch_heig_part(0, st, d_jobs, d_state, d_list_lds + sp.start, sp.count, sp.maxn, d_arena, d_eps, d_perm,
                             d_tri, true);
    for (int p = 1; p < 3; ++p)
        if (part < 0 || part == p)
            launch_heig_part(p, st, d_jobs, d_state, d_list_lds, (int)list_lds.size(), maxn_lds, d_arena, d_eps,
                             d_perm, d_tri, true);
}

bool Batch::split_ready() const { return split_job > 0 && fork_stream && aux_stream && aux_fork && aux_join; }

// One iteration with the LDS-class jobs [0, split_job) on fork_stream and the
// others on st, every stage restricted to its part's pairs, atoms, jobs and
// tiles (a job's are contiguous), joined once at the end: neither part waits
// for the other between stages.
void Batch::launch_iteration_split(Engine* e, hipStream_t st) {
    const int nj = (int)jobs.size(), js = split_job, nxl = (int)list_xl.size();
    const int nt = (int)tiles.size(), na = (int)atoms.size();
    hipStream_t s2 = fork_stream;
    hipEventRecord(fork_ev, st);
    hipStreamWaitEvent(s2, fork_ev, 0);
    launch_fock(s2, 2, fock_pairs(0, js), d_atoms, 0, split_atom, e->d_elems, d_jobs, d_state, d_ints, d_ints32,
                d_arena, d_part);
    launch_energy(s2, d_jobs, js, d_state, d_arena, d_remaining, d_xlpart, xl_chunks, 0);
    launch_lds_eigen(-1, s2);
    launch_density(s2, d_tiles, split_tile, d_jobs, d_state, d_arena, d_perm, d_rmspart, d_tpart);
    launch_step(s2, d_jobs, js, d_state, d_rmspart, d_tpart, d_traj, d_remaining, d_xlpart, d_xlcoef, xl_chunks, 0);
    launch_apply(s2, d_tiles, split_tile, d_jobs, d_arena, d_xlcoef);
    hipEventRecord(join_ev, s2);

    launch_fock(st, 2, fock_pairs(js, nj), d_atoms, split_atom, na - split_atom, e->d_elems, d_jobs, d_state,
                d_ints, d_ints32, d_arena, d_part, aux_stream, aux_fork, aux_join);
    launch_xl_energy(st, d_jobs, d_state, d_list_xl, nxl, d_arena, d_xlpart, xl_chunks);
    launch_energy(st, d_jobs, nj - js, d_state, d_arena, d_remaining, d_xlpart, xl_chunks, js);
    if (nxl && !xl_householder)
        launch_warm(st, d_tiles + split_tile, nt - split_tile, d_jobs, d_state, d_arena);
    launch_eigen(st, d_jobs, d_state, d_list_glb, (int)list_glb.size(), maxn_glb, d_arena, d_eps, d_perm, d_tri, false);
    launch_xl_eigen_all(e, st);
    launch_density(st, d_tiles + split_tile, nt - split_tile, d_jobs, d_state, d_arena, d_perm, d_rmspart + split_tile,
                   d_tpart + (long long)split_tile * XL_REC);
    launch_xl_dots(st, d_jobs, d_state, d_list_xl, nxl, d_arena, d_xlpart, xl_chunks);
    launch_step(st, d_jobs, nj - js, d_state, d_rmspart, d_tpart, d_traj, d_remaining, d_xlpart, d_xlcoef, xl_chunks,
                js);
    launch_apply(st, d_tiles + split_tile, nt - split_tile, d_jobs, d_arena, d_xlcoef);
    launch_xl_apply(st, d_jobs, d_list_xl, nxl, d_arena, d_xlcoef, xl_chunks);
    hipStreamWaitEvent(st, join_ev, 0);
}

// One SCF iteration of every job: the Fock launch also rebuilds F from the final
// density for jobs that converged in the previous iteration, whose energy is
// then evaluated (Cndo2.cpp:1324-1416) before the eigensolve of the others.
void Batch::launch_iteration(Engine* e, hipStream_t st) {
    if (split_ready()) {
        launch_iteration_split(e, st);
        return;
    }
    const int nxl = (int)list_xl.size();
    launch_fock(st, 2, fock_pairs(0, (int)jobs.size()), d_atoms, 0, (int)atoms.size(), e->d_elems, d_jobs, d_state,
                d_ints, d_ints32, d_arena, d_part, fork_stream, fork_ev2, join_ev2);
    launch_xl_energy(st, d_jobs, d_state, d_list_xl, nxl, d_arena, d_xlpart, xl_chunks);
    launch_energy(st, d_jobs, (int)jobs.size(), d_state, d_arena, d_remaining, d_xlpart, xl_chunks);
    // the warm-start rotation is for the block-Jacobi (large) jobs only
    if (nxl && !xl_householder) launch_warm(st, d_tiles, (int)tiles.size(), d_jobs, d_state, d_arena);
    const bool fork = fork_stream && !list_lds.empty() && !list_glb.empty();
    if (fork) {
        hipEventRecord(fork_ev, st);
        hipStreamWaitEvent(fork_stream, fork_ev, 0);
    }
    launch_lds_eigen(-1, fork ? fork_stream : st);
    if (fork) hipEventRecord(join_ev, fork_stream);
    launch_eigen(st, d_jobs, d_state, d_list_glb, (int)list_glb.size(), maxn_glb, d_arena, d_eps, d_perm, d_tri, false);
    launch_xl_eigen_all(e, st);
    if (fork) hipStreamWaitEvent(st, join_ev, 0);
    launch_density(st, d_tiles, (int)tiles.size(), d_jobs, d_state, d_arena, d_perm, d_rmspart, d_tpart);
    launch_xl_dots(st, d_jobs, d_state, d_list_xl, nxl, d_arena, d_xlpart, xl_chunks);
    launch_step(st, d_jobs, (int)jobs.size(), d_state, d_rmspart, d_tpart, d_traj, d_remaining, d_xlpart, d_xlcoef,
                xl_chunks);
    launch_apply(st, d_tiles, (int)tiles.size(), d_jobs, d_arena, d_xlcoef);
    launch_xl_apply(st, d_jobs, d_list_xl, nxl, d_arena, d_xlcoef, xl_chunks);
}

void Batch::iterate(Engine* e, hipStream_t st) {
    const bool eager = !e->use_graphs || !list_xl.empty() || graph_off || eager_only || iterations_launched == 0;
    ++iterations_launched;
    if (eager) {
        launch_iteration(e, st);
        return;
    }
    if (!gexec) {
        hipGraph_t g = nullptr;
        bool ok;
        {
            std::lock_guard<std::recursive_mutex> cap(e->capture_mu);
            if (served && e->sync_busy.load() > 0) {   // a synchronous call runs: capture at a later iteration
                launch_iteration(e, st);
                return;
            }
            ok = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess;
            if (ok) {
                launch_iteration(e, st);
                ok = hipStreamEndCapture(st, &g) == hipSuccess && g;
            }
        }
        if (ok) ok = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0) == hipSuccess;
        if (g) hipGraphDestroy(g);
        if (ok) e->graphs_captured++;
        if (!ok) {   // capture unavailable: this batch launches eagerly from now on
            if (gexec) hipGraphExecDestroy(gexec);
            gexec = nullptr;
            graph_off = true;
            hipGetLastError();
            launch_iteration(e, st);
            return;
        }
    }
    if (hipGraphLaunch(gexec, st) != hipSuccess) {   // replay refused: this batch launches eagerly from now on
        hipGetLastError();
        hipGraphExecDestroy(gexec);
        gexec = nullptr;
        graph_off = true;
        launch_iteration(e, st);
        return;
    }
    e->graph_replays++;
}

// The large-molecule eigensolves, at most xl_max_jobs() per persistent launch,
// ordered after the previous such launch on any stream of this engine.
void Batch::launch_xl_eigen_all(Engine* e, hipStream_t st) {
    if (list_xl.empty()) return;
    if (e->xl_chain) hipStreamWaitEvent(st, e->xl_chain, 0);
    for (size_t g = 0; g < list_xl.size(); g += xl_max_jobs()) {
        const int cnt = (int)std::min<size_t>(xl_max_jobs(), list_xl.size() - g);
        if (xl_householder)
            launch_xlh_eigen(st, xlh_grid(e->num_cus), d_jobs, d_state, d_list_xl + g, d_xl_off + g, cnt, maxn_xl,
                             d_arena, d_eps, d_perm, d_tri, d_xlws, d_bar);
        else
            launch_xl_eigen(st, xl_grid, d_jobs, d_state, d_list_xl + g, d_xl_off + g, cnt, d_arena, d_eps, d_perm,
                            d_xlws, d_bar);
    }
    if (e->xl_chain) hipEventRecord(e->xl_chain, st);
}

int Batch::run_scf(Engine* e, hipStream_t st) {
    const bool prof = e->profile;
    if (prof) HIPCHK(hipEventRecord(e->ev[0], st));
    launch_setup_stage(e, st);
    HIPCHK(hipGetLastError());
    if (prof) {
        HIPCHK(hipEventRecord(e->ev[1], st));
        HIPCHK(hipEventSynchronize(e->ev[1]));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e->ev[0], e->ev[1]));
        e->stage_ms[NDDO_STAGE_SETUP] += ms;
        e->stage_launches[NDDO_STAGE_SETUP] += 1;
    }
    int* h_rem = e->h_remaining;
    const bool fork = fork_stream && !list_lds.empty() && !list_glb.empty();
    for (int it = 0; it <= max_iter_all + 1; ++it) {
        if (!prof) {
            iterate(e, st);
        } else {
            // the same launches as launch_iteration, with an event between stages
            // and between the parts of the eigensolver
            HIPCHK(hipEventRecord(e->ev[0], st));
            launch_fock(st, 2, fock_pairs(0, (int)jobs.size()), d_atoms, 0, (int)atoms.size(), e->d_elems, d_jobs,
                        d_state, d_ints, d_ints32, d_arena, d_part, fork_stream, fork_ev2, join_ev2);
            HIPCHK(hipEventRecord(e->ev[9], st));
            launch_xl_energy(st, d_jobs, d_state, d_list_xl, (int)list_xl.size(), d_arena, d_xlpart, xl_chunks);
            launch_energy(st, d_jobs, (int)jobs.size(), d_state, d_arena, d_remaining, d_xlpart, xl_chunks);
            HIPCHK(hipEventRecord(e->ev[1], st));
            if (!list_xl.empty() && !xl_householder) launch_warm(st, d_tiles, (int)tiles.size(), d_jobs, d_state, d_arena);
            if (fork) {
                // the LDS class on the second stream, timed as one stage (all three parts)
                HIPCHK(hipEventRecord(fork_ev, st));
                HIPCHK(hipStreamWaitEvent(fork_stream, fork_ev, 0));
                HIPCHK(hipEventRecord(e->ev[11], fork_stream));
                launch_lds_eigen(-1, fork_stream);
                HIPCHK(hipEventRecord(e->ev[12], fork_stream));
                HIPCHK(hipEventRecord(join_ev, fork_stream));
            } else {
                launch_lds_eigen(0, st);
            }
            HIPCHK(hipEventRecord(e->ev[2], st));
            launch_heig_part(0, st, d_jobs, d_state, d_list_glb, (int)list_glb.size(), maxn_glb, d_arena, d_eps, d_perm,
                             d_tri, false);
            HIPCHK(hipEventRecord(e->ev[3], st));
            for (int part = 1; part < 3; ++part) {
                if (!fork) launch_lds_eigen(part, st);
                launch_heig_part(part, st, d_jobs, d_state, d_list_glb, (int)list_glb.size(), maxn_glb, d_arena, d_eps,
                                 d_perm, d_tri, false);
                HIPCHK(hipEventRecord(e->ev[3 + part], st));
            }
            launch_xl_eigen_all(e, st);
            HIPCHK(hipEventRecord(e->ev[13], st));
            if (fork) HIPCHK(hipStreamWaitEvent(st, join_ev, 0));
            HIPCHK(hipEventRecord(e->ev[10], st));
            launch_density(st, d_tiles, (int)tiles.size(), d_jobs, d_state, d_arena, d_perm, d_rmspart, d_tpart);
            HIPCHK(hipEventRecord(e->ev[6], st));
            launch_xl_dots(st, d_jobs, d_state, d_list_xl, (int)list_xl.size(), d_arena, d_xlpart, xl_chunks);
            launch_step(st, d_jobs, (int)jobs.size(), d_state, d_rmspart, d_tpart, d_traj, d_remaining, d_xlpart,
                        d_xlcoef, xl_chunks);
            launch_apply(st, d_tiles, (int)tiles.size(), d_jobs, d_arena, d_xlcoef);
            launch_xl_apply(st, d_jobs, d_list_xl, (int)list_xl.size(), d_arena, d_xlcoef, xl_chunks);
            HIPCHK(hipEventRecord(e->ev[7], st));
        }
        HIPCHK(hipGetLastError());
        launch_publish(st, d_remaining, sizeof(int), e->d_h_remaining, sizeof(int), sizeof(int), 1);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
        if (prof) {
            // (stage, first event, last event)
            // forked: the LDS-class stage is its whole eigensolve on the second
            // stream, and D&C / back-transform time the device-memory class
            const int l0 = fork ? 11 : 1, l1 = fork ? 12 : 2;
            const int map[10][3] = {{NDDO_STAGE_FOCK, 0, 9},          {NDDO_STAGE_EIGEN, 1, 10},
                                    {NDDO_STAGE_TRIDIAG_LDS, l0, l1}, {NDDO_STAGE_TRIDIAG_GLB, 2, 3},
                                    {NDDO_STAGE_DC, 3, 4},            {NDDO_STAGE_BACKTRANSFORM, 4, 5},
                                    {NDDO_STAGE_XL_EIGEN, 5, 13},     {NDDO_STAGE_DENSITY, 10, 6},
                                    {NDDO_STAGE_STEP, 6, 7},          {NDDO_STAGE_ENERGY, 9, 1}};
            for (int k = 0; k < 10; ++k) {
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, e->ev[map[k][1]], e->ev[map[k][2]]));
                e->stage_ms[map[k][0]] += ms;
                e->stage_launches[map[k][0]] += 1;
            }
        }
        if (*h_rem <= 0) break;
    }
    return NDDO_OK;
}

// Scalar results of a job from its device state (shared by the synchronous
// fetch and the serving scheduler).
void fill_scalars(const JobDev& jb, const JobState& s, int theory, double hof_offset, nddo_result* res) {
    const int n = jb.nao;
    res->nao = n;
    res->nocc = jb.nocc;
    res->status = s.phase == PH_DONE ? s.status : NDDO_NOT_CONVERGED;
    if (s.sweeps < 0 && res->status == NDDO_OK) res->status = NDDO_EIGENSOLVER_FAILED;
    res->iterations = s.status == 0 && s.phase == PH_DONE ? s.iter + 1 : s.iter;
    res->fp64_iterations = jb.precision == NDDO_MIXED ? (s.fp32 ? 0 : res->iterations - s.switch_iter)
                                                       : res->iterations;
    res->energy_core = s.e_core;
    res->energy_elec = s.e_el;
    res->energy_total = s.e_el + s.e_core;
    (void)theory;
    res->heat_of_formation = res->energy_total + hof_offset;
    res->final_rms = s.rms;
    if (res->status != NDDO_OK) {
        res->energy_elec = NAN;
        res->energy_total = NAN;
        res->heat_of_formation = NAN;
    }
}

// C rows in rank order with MolDS's sign rule (Lapack.cpp:139-150): each row's element sum >= 0.
void order_and_phase(int n, const double* VT, const int* perm, double* C) {
    for (int r = 0; r < n; ++r) {
        const double* src = &VT[(long long)perm[r] * n];
        double sum = 0.0;
        for (int i = 0; i < n; ++i) sum += src[i];
        const double sg = sum < 0.0 ? -1.0 : 1.0;
        for (int i = 0; i < n; ++i) C[(long long)r * n + i] = sg * src[i];
    }
}

// Copies job jj's results into res (status already set by the caller on failure).
// All jobs' results with bulk copies of the scalar state, eigenvalues and
// orbital order; per-job copies only for the requested arrays (C, P, trajectory).
// Mulliken charges of a final density: the atom's core charge (valence
// electrons: H 1, C 4, N 5, O 6, F 7) minus its sum of P_mu,mu
// (Cndo2.cpp:733, 1187-1199, 1733-1747); pdiag[k * stride] = P_kk.
void mulliken_charges(int natoms, const int* Z, const double* pdiag, long long stride, double* q) {
    long long mu = 0;
    for (int a = 0; a < natoms; ++a) {
        const bool h = Z[a] == 1;
        double pop = 0.0;
        for (int k = 0; k < (h ? 1 : 4); ++k, ++mu) pop += pdiag[mu * stride];
        q[a] = (h ? 1.0 : (double)(Z[a] - 2)) - pop;
    }
}

// Dipole moments (a.u.) from the per-row partial sums of k_dipole, as MolDS
// prints them (Cndo2.cpp:748-753, 1871-1934): about the centre of core masses
// (Molecule::CalcXyzCOC: atomic mass in electron masses minus the valence
// electrons), electronic = -sum P r + centre sum P S, core = sum_A Z_A (R_A -
// centre); out = total, electronic, core (x, y, z each).
void dipole_moments(const nddo_job_desc& d, const double* rows, int n, double* out) {
    const double gmol2au = 1.0e5 / (6.0221415 * 9.1095);   // Parameters.cpp:52
    double centre[3] = {0.0, 0.0, 0.0}, msum = 0.0;
    for (int a = 0; a < d.natoms; ++a) {
        const double core = d.Z[a] == 1 ? 1.0 : (double)(d.Z[a] - 2);
        const double m = atomic_mass_gmol(d.Z[a]) * gmol2au - core;
        msum += m;
        for (int c = 0; c < 3; ++c) centre[c] += m * d.xyz_bohr[3 * a + c];
    }
    for (int c = 0; c < 3; ++c) centre[c] /= msum;
    double pr[3] = {0.0, 0.0, 0.0}, ps = 0.0;
    for (int mu = 0; mu < n; ++mu) {
        for (int c = 0; c < 3; ++c) pr[c] += rows[4 * mu + c];
        ps += rows[4 * mu + 3];
    }
    for (int c = 0; c < 3; ++c) {
        double core = 0.0;
        for (int a = 0; a < d.natoms; ++a)
            core += (d.Z[a] == 1 ? 1.0 : (double)(d.Z[a] - 2)) * (d.xyz_bohr[3 * a + c] - centre[c]);
        const double el = -pr[c] + centre[c] * ps;
        out[c] = el + core;
        out[3 + c] = el;
        out[6 + c] = core;
    }
}

int Batch::fetch_all(hipStream_t st, const nddo_job_desc* descs, nddo_result* const* res, const double* hof_offsets) {
    const size_t nj = jobs.size();
    std::vector<JobState> s(nj);
    std::vector<double> eps(std::max<long long>(nvec, 1));
    std::vector<int> perm(std::max<long long>(nvec, 1));
    HIPCHK(hipMemcpyAsync(s.data(), d_state, nj * sizeof(JobState), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(eps.data(), d_eps, nvec * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(perm.data(), d_perm, nvec * sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<std::vector<double>> VT(nj), PD(nj);
    for (size_t jj = 0; jj < nj; ++jj) {
        const JobDev& jb = jobs[jj];
        const nddo_job_desc& d = descs[src[jj]];
        nddo_result* r = res[jj];
        fill_scalars(jb, s[jj], d.theory, hof_offsets[jj], r);
        const int n = jb.nao;
        if (r->eps) std::memcpy(r->eps, eps.data() + jb.vec0, n * sizeof(double));
        const double* base = d_arena + jb.arena;
        const long long nn = (long long)n * n;
        if (r->C) {
            VT[jj].resize(nn);
            HIPCHK(hipMemcpyAsync(VT[jj].data(), base + (long long)AR_C * nn, nn * sizeof(double), hipMemcpyDeviceToHost,
                                  st));
        }
        const int pk = AR_P0 + ((s[jj].iter + 1) & 1);
        if (r->P) {
            HIPCHK(hipMemcpyAsync(r->P, base + (long long)pk * nn, nn * sizeof(double), hipMemcpyDeviceToHost, st));
        } else if (r->mulliken) {   // only the diagonal of P
            PD[jj].resize(n);
            HIPCHK(hipMemcpy2DAsync(PD[jj].data(), sizeof(double), base + (long long)pk * nn, (n + 1) * sizeof(double),
                                    sizeof(double), n, hipMemcpyDeviceToHost, st));
        }
        if (r->trajectory)
            HIPCHK(hipMemcpyAsync(r->trajectory, d_traj + jb.traj0, 4LL * jb.max_iter * sizeof(double),
                                  hipMemcpyDeviceToHost, st));
    }
    std::vector<double> dip(std::max<long long>(dip_size, 1));
    if (!list_dip.empty())
        HIPCHK(hipMemcpyAsync(dip.data(), d_dippart, dip_size * sizeof(double), hipMemcpyDeviceToHost, st));
    std::vector<double> frc;
    if (any_forces) {
        frc.resize(3 * atoms.size());
        HIPCHK(hipMemcpyAsync(frc.data(), d_forces, frc.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    HIPCHK(hipStreamSynchronize(st));
    for (size_t jj = 0; jj < nj && any_forces; ++jj) {
        nddo_result* r = res[jj];
        if (!r->forces || !jobs[jj].want_f) continue;
        const int na = jobs[jj].natoms;
        if (r->status == NDDO_OK)
            std::memcpy(r->forces, frc.data() + 3LL * jobs[jj].atom0, 3 * (size_t)na * sizeof(double));
        else
            for (int k = 0; k < 3 * na; ++k) r->forces[k] = std::nan("");
    }
    for (size_t k = 0; k < list_dip.size(); ++k) {
        const int jj = list_dip[k];
        nddo_result* r = res[jj];
        if (!r->dipole || r->status != NDDO_OK) continue;
        dipole_moments(descs[src[jj]], dip.data() + dip_off[k], jobs[jj].nao, r->dipole);
    }
    for (size_t jj = 0; jj < nj; ++jj) {
        nddo_result* r = res[jj];
        if (r->C) order_and_phase(jobs[jj].nao, VT[jj].data(), perm.data() + jobs[jj].ivec0, r->C);
        if (r->mulliken) {
            const nddo_job_desc& d = descs[src[jj]];
            if (r->P) mulliken_charges(d.natoms, d.Z, r->P, jobs[jj].nao + 1, r->mulliken);
            else mulliken_charges(d.natoms, d.Z, PD[jj].data(), 1, r->mulliken);
        }
    }
    return NDDO_OK;
}

size_t device_available(int device) {
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) != hipSuccess) return 0;
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
        uint64_t reserved = 0, used = 0;
        if (hipMemPoolGetAttribute(pool, hipMemPoolAttrReservedMemCurrent, &reserved) == hipSuccess &&
            hipMemPoolGetAttribute(pool, hipMemPoolAttrUsedMemCurrent, &used) == hipSuccess && reserved > used)
            f += (size_t)(reserved - used);
    }
    return f;
}

void trim_pool(int device) {
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) hipMemPoolTrimTo(pool, 0);
}

void free_workspace(Engine* e) {
    if (e->d_ws) hipFree(e->d_ws);
    e->d_ws = nullptr;
    e->ws_cap = 0;
}

// Grow-only: a batch that fits the current workspace reuses it as it is; a
// larger one replaces it with one of the new size plus 1/16 (so a slightly
// larger batch next time does not reallocate), capped by what the device can
// give.  1 GiB of the device's free memory is always left to the runtime.
int ensure_workspace(Engine* e, size_t bytes) {
    if (bytes <= e->ws_cap) return NDDO_OK;
    if (bytes > e->hbm_budget) {
        set_error("a batch needs " + std::to_string(bytes) + " device bytes; the engine's budget is " +
                  std::to_string(e->hbm_budget));
        return NDDO_ADMISSION_DENIED;
    }
    const size_t margin = (size_t)1 << 30;
    free_workspace(e);
    size_t avail = device_available(e->device);
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) == hipSuccess && avail > f) {   // the pool's unused reserve goes back first
        trim_pool(e->device);
        if (hipMemGetInfo(&f, &t) == hipSuccess) avail = f;
    }
    if (bytes + margin > avail) {
        set_error("a batch needs " + std::to_string(bytes) + " device bytes; " + std::to_string(avail) +
                  " are free");
        return NDDO_ADMISSION_DENIED;
    }
    const size_t want = std::min(std::min(bytes + bytes / 16, avail - margin), std::max(bytes, e->hbm_budget));
    void* p = nullptr;
    const hipError_t err = hipMalloc(&p, want);
    if (err != hipSuccess) {
        hipGetLastError();
        set_error(std::string("workspace allocation of ") + std::to_string(want) + " bytes: " + hipGetErrorString(err));
        return NDDO_ADMISSION_DENIED;
    }
    e->d_ws = (char*)p;
    e->ws_cap = want;
    ++e->ws_grows;
    return NDDO_OK;
}

}  // namespace nddo

// ------------------------------------------------------------------ C ABI
extern "C" {

const char* nddo_last_error(void) { return g_err.c_str(); }

void nddo_scf_opts_default(nddo_scf_opts* o) {
    // src/base/Parameters.cpp:132-144
    o->max_iter = 100;
    o->rms_density = 1.0e-8;
    o->damping_thresh = 1.0;
    o->damping_weight = 0.8;
    o->diis_num_error_vect = 5;
    o->diis_start_error = 1.0e-2;
    o->diis_end_error = 1.0e-8;
}

static int autotune(Engine* e);

int nddo_engine_create(int hip_device, const nddo_engine_opts* opts, nddo_engine** out) {
    if (!out) { set_error("null engine pointer"); return NDDO_BAD_ARG; }
    *out = nullptr;
    // The serving scheduler runs up to 8 cohorts on streams of their own; ask
    // HIP for 16 hardware queues (its default is 4) unless the caller chose.
    // Only effective when this is the process's first HIP call.
    setenv("GPU_MAX_HW_QUEUES", "16", 0);
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (hip_device < 0 || hip_device >= ndev) {
        set_error("no such HIP device");
        return NDDO_BAD_ARG;
    }
    HIPCHK(hipSetDevice(hip_device));
    Engine* e = new Engine();
    e->device = hip_device;
    if (opts) e->opts = *opts;
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, hip_device));
    e->lds_max = prop.sharedMemPerBlock;
    e->num_cus = prop.multiProcessorCount;
    e->xl_min_nao = (e->opts.flags & NDDO_ENGINE_DEVICE_WIDE_EIGEN) ? 513 : 1201;
    e->use_graphs = !(e->opts.flags & NDDO_ENGINE_NO_GRAPHS);
    e->xl_householder = !(e->opts.flags & NDDO_ENGINE_BLOCK_JACOBI);
    HIPCHK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&e->api_stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&e->eig_stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&e->eig_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&e->eig_join, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&e->fock_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&e->fock_join, hipEventDisableTiming));
    HIPCHK(hipStreamCreateWithFlags(&e->aux_stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&e->aux_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&e->aux_join, hipEventDisableTiming));
    std::vector<ElemDev> el(3 * kNumElem);
    for (int t = 0; t < 3; ++t)
        for (int i = 0; i < kNumElem; ++i) elem_params(kElemZ[i], t, &el[t * kNumElem + i]);
    HIPCHK(hipMalloc(&e->d_elems, el.size() * sizeof(ElemDev)));
    HIPCHK(hipMemcpy(e->d_elems, el.data(), el.size() * sizeof(ElemDev), hipMemcpyHostToDevice));
    double tab[32 * 25];
    overlap_table(tab);
    if (upload_overlap_table(tab) != 0) {
        set_error("overlap table upload failed");
        return NDDO_HIP_ERROR;
    }
    // fine-grained (coherent) pinned memory: written by a kernel each iteration
    HIPCHK(hipHostMalloc((void**)&e->h_remaining, sizeof(int), hipHostMallocCoherent));
    HIPCHK(hipHostGetDevicePointer((void**)&e->d_h_remaining, e->h_remaining, 0));
    for (int i = 0; i < 16; ++i) HIPCHK(hipEventCreate(&e->ev[i]));
    HIPCHK(hipEventCreateWithFlags(&e->xl_chain, hipEventDisableTiming));
    {
        void* ptr = nullptr;
        const hipError_t se = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (se != hipSuccess || !ptr) {
            set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(se));
            return NDDO_HIP_ERROR;
        }
        e->d_scratch = (char*)ptr;
    }
    size_t freeb = 0, totb = 0;
    HIPCHK(hipMemGetInfo(&freeb, &totb));
    e->hbm_budget = e->opts.hbm_budget_bytes ? e->opts.hbm_budget_bytes : (size_t)(0.9 * freeb);
    if (!(e->opts.flags & NDDO_ENGINE_NO_AUTOTUNE) && autotune(e) != NDDO_OK) {
        set_tuning(Tuning());   // keep the defaults when the tuning runs fail
        hipGetLastError();
    }
    e->start_server();
    *out = reinterpret_cast<nddo_engine*>(e);
    return NDDO_OK;
}

int nddo_engine_destroy(nddo_engine* eh) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e) return NDDO_BAD_ARG;
    e->stop_server();
    hipSetDevice(e->device);
    free_workspace(e);
    if (e->d_elems) hipFree(e->d_elems);
    if (e->h_remaining) hipHostFree(e->h_remaining);
    for (int i = 0; i < 16; ++i)
        if (e->ev[i]) hipEventDestroy(e->ev[i]);
    if (e->xl_chain) hipEventDestroy(e->xl_chain);
    if (e->stream) hipStreamDestroy(e->stream);
    if (e->api_stream) hipStreamDestroy(e->api_stream);
    if (e->eig_fork) hipEventDestroy(e->eig_fork);
    if (e->eig_join) hipEventDestroy(e->eig_join);
    if (e->fock_fork) hipEventDestroy(e->fock_fork);
    if (e->fock_join) hipEventDestroy(e->fock_join);
    if (e->eig_stream) hipStreamDestroy(e->eig_stream);
    if (e->aux_fork) hipEventDestroy(e->aux_fork);
    if (e->aux_join) hipEventDestroy(e->aux_join);
    if (e->aux_stream) hipStreamDestroy(e->aux_stream);
    delete e;
    return NDDO_OK;
}

// Runs the jobs `which` (indices into descs) as one batch carved out of the
// engine's workspace.  When they do not fit the device's free memory they run
// as two halves, one after the other (a job's results do not depend on the
// batch it shares); a single job that does not fit ends ADMISSION_DENIED.
static int run_batch(Engine* e, const nddo_job_desc* descs, int njobs, nddo_result* res, const std::vector<int>& which,
                     std::chrono::steady_clock::time_point t0) {
    using Clock = std::chrono::steady_clock;
    auto lap = [](Clock::time_point& t) {
        const auto now = Clock::now();
        const double dt = std::chrono::duration<double>(now - t).count();
        t = now;
        return dt;
    };
    Clock::time_point tp = Clock::now();
    Batch b;
    b.build(descs, njobs, which, nullptr);
    // jobs whose caller takes no C back-transform only the occupied eigenvectors
    for (size_t jj = 0; jj < which.size(); ++jj) b.jobs[jj].want_c = res[which[jj]].C != nullptr ? 1 : 0;
    b.use_xl_householder = e->xl_householder;
    b.split_classes = true;
    b.classify(e->lds_max, e->num_cus, e->xl_min_nao, true);
    {
        std::vector<int> dj, fj;
        for (size_t jj = 0; jj < which.size(); ++jj) {
            if (res[which[jj]].dipole) dj.push_back((int)jj);
            if (res[which[jj]].forces) fj.push_back((int)jj);
        }
        b.set_dipole_jobs(dj);
        b.set_force_jobs(fj);
    }
    const size_t need = b.device_bytes(false);
    int rc = ensure_workspace(e, need);
    if (rc == NDDO_ADMISSION_DENIED) {
        if (which.size() == 1) {
            res[which[0]].status = NDDO_ADMISSION_DENIED;
            return NDDO_OK;
        }
        const size_t h = which.size() / 2;
        const std::vector<int> lo(which.begin(), which.begin() + h), hi(which.begin() + h, which.end());
        rc = run_batch(e, descs, njobs, res, lo, t0);
        if (rc) return rc;
        return run_batch(e, descs, njobs, res, hi, t0);
    }
    if (rc) return rc;
    b.region = e->d_ws;
    b.region_cap = e->ws_cap;
    b.fork_stream = e->eig_stream;
    b.fork_ev = e->eig_fork;
    b.join_ev = e->eig_join;
    b.fork_ev2 = e->fock_fork;
    b.join_ev2 = e->fock_join;
    b.aux_stream = e->aux_stream;
    b.aux_fork = e->aux_fork;
    b.aux_join = e->aux_join;
    e->host_s[0] += lap(tp);
    rc = b.allocate(false);
    if (rc) return rc;
    rc = b.upload(e->stream, descs);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(e->stream));
    e->host_s[1] += lap(tp);
    rc = b.run_scf(e, e->stream);
    if (rc) return rc;
    b.launch_properties(e, e->stream);
    HIPCHK(hipGetLastError());
    e->host_s[2] += lap(tp);
    const double wall = std::chrono::duration<double>(Clock::now() - t0).count();
    std::vector<nddo_result*> rp(which.size());
    std::vector<double> hof(which.size(), 0.0);
    for (size_t jj = 0; jj < which.size(); ++jj) {
        const nddo_job_desc& d = descs[which[jj]];
        for (int a = 0; a < d.natoms; ++a) hof[jj] += atom_hof_offset(d.Z[a], d.theory);
        rp[jj] = &res[which[jj]];
    }
    rc = b.fetch_all(e->stream, descs, rp.data(), hof.data());
    if (rc) return rc;
    for (size_t jj = 0; jj < which.size(); ++jj) res[which[jj]].latency_s = wall;
    e->host_s[3] += lap(tp);
    return NDDO_OK;
}

int nddo_scf_batch(nddo_engine* eh, int njobs, const nddo_job_desc* descs, nddo_result* res) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || njobs < 0 || (njobs > 0 && (!descs || !res))) { set_error("bad arguments"); return NDDO_BAD_ARG; }
    auto t0 = std::chrono::steady_clock::now();
    SyncRun lk(e);
    HIPCHK(hipSetDevice(e->device));
    std::vector<int> ok, rest;
    for (int i = 0; i < njobs; ++i) {
        int n = 0, no = 0;
        int s = validate_job(descs[i], &n, &no);
        res[i].status = s;
        res[i].nao = n;
        res[i].nocc = no;
        // the jobs of the LDS eigensolver class first (the split iteration)
        if (s == NDDO_OK) (lds_eig_class(n, e->lds_max) ? ok : rest).push_back(i);
    }
    ok.insert(ok.end(), rest.begin(), rest.end());
    if (ok.empty()) return NDDO_OK;
    const int rc = run_batch(e, descs, njobs, res, ok, t0);
    for (int i : ok)
        if (res[i].status == NDDO_ADMISSION_DENIED)
            set_error("a job's device footprint exceeds the free device memory");
    return rc;
}

// nddo_scf_once on buffers carved out of the fixed scratch array: no device
// allocation or free, no graph instantiation (the iterations launch eagerly).
// Calls that overlap on one device share that device's array and race.
int nddo_scf_once_noalloc(nddo_engine* eh, const nddo_job_desc* job, nddo_result* res) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || !job || !res) { set_error("bad arguments"); return NDDO_BAD_ARG; }
    auto t0 = std::chrono::steady_clock::now();
    SyncRun lk(e);
    HIPCHK(hipSetDevice(e->device));
    int n = 0, no = 0;
    const int s = validate_job(*job, &n, &no);
    res->status = s;
    res->nao = n;
    res->nocc = no;
    if (s != NDDO_OK) return s;
    Batch b;
    const std::vector<int> which{0};
    b.build(job, 1, which, nullptr);
    b.jobs[0].want_c = res->C != nullptr ? 1 : 0;
    b.use_xl_householder = e->xl_householder;
    b.classify(e->lds_max, e->num_cus, e->xl_min_nao);
    if (res->dipole) b.set_dipole_jobs(std::vector<int>{0});
    if (res->forces) b.set_force_jobs(std::vector<int>{0});
    const size_t need = b.device_bytes(false);
    if (need > kNoAllocScratchBytes || !b.list_xl.empty()) {
        set_error("nddo_scf_once_noalloc: the job needs " + std::to_string(need) +
                  " device bytes; the fixed scratch holds " + std::to_string(kNoAllocScratchBytes));
        res->status = NDDO_ADMISSION_DENIED;
        return NDDO_ADMISSION_DENIED;
    }
    b.region = e->d_scratch;
    b.region_cap = kNoAllocScratchBytes;
    int rc = b.allocate(false);
    if (rc) return rc;
    b.eager_only = true;
    rc = b.upload(e->stream, job);   // clears the arena and trajectory the last call left
    if (rc) return rc;
    rc = b.run_scf(e, e->stream);
    if (rc) return rc;
    b.launch_properties(e, e->stream);
    HIPCHK(hipGetLastError());
    double hof = 0.0;
    for (int a = 0; a < job->natoms; ++a) hof += atom_hof_offset(job->Z[a], job->theory);
    nddo_result* rp = res;
    rc = b.fetch_all(e->stream, job, &rp, &hof);
    if (rc) return rc;
    res->latency_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (res->status == NDDO_NOT_CONVERGED)
        set_error("Error in cndo::Cndo2::DoSCF: SCF did not met convergence criterion. maxIterationsSCF=" +
                  std::to_string(job->opts.max_iter));
    return res->status;
}

size_t nddo_noalloc_scratch_bytes(void) { return kNoAllocScratchBytes; }

int nddo_element_eisol(int Z, int theory, double* eisol_ev_out) {
    if (!eisol_ev_out || !element_supported(Z) || theory < 0 || theory > 2) {
        set_error("unsupported element or theory");
        return NDDO_BAD_ARG;
    }
    *eisol_ev_out = eisol_ev(Z, theory);
    return NDDO_OK;
}

int nddo_element_params(int Z, int theory, double* out) {
    ElemDev e;
    if (!out || elem_params(Z, theory, &e) != 0) { set_error("unsupported element or theory"); return NDDO_BAD_ARG; }
    const double v[33] = {e.core, e.U[0], e.U[1], e.zeta[0], e.zeta[1], e.beta[0], e.beta[1], e.alpha,
                          e.Gss, e.Gpp, e.Gsp, e.Gpp2, e.Hsp, e.Hpp, e.D[0], e.D[1], e.D[2], e.rho[0], e.rho[1],
                          e.rho[2], e.K[0], e.K[1], e.K[2], e.K[3], e.L[0], e.L[1], e.L[2], e.L[3], e.M[0], e.M[1],
                          e.M[2], e.M[3], mndo_atom_hof_offset(Z)};
    std::memcpy(out, v, sizeof(v));
    return NDDO_OK;
}

int nddo_engine_trim(nddo_engine* eh) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e) return NDDO_BAD_ARG;
    SyncRun lk(e);
    HIPCHK(hipSetDevice(e->device));
    free_workspace(e);
    trim_pool(e->device);
    return NDDO_OK;
}

int nddo_engine_memory(nddo_engine* eh, size_t* free_bytes, size_t* total_bytes, size_t* workspace_bytes,
                       size_t* served_result_bytes) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e) return NDDO_BAD_ARG;
    HIPCHK(hipSetDevice(e->device));
    size_t f = 0, t = 0;
    {
        std::lock_guard<std::recursive_mutex> cap(e->capture_mu);
        HIPCHK(hipMemGetInfo(&f, &t));
    }
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    if (workspace_bytes) {
        std::lock_guard<std::mutex> lk(e->run_mu);
        *workspace_bytes = e->ws_cap;
    }
    if (served_result_bytes) *served_result_bytes = (size_t)std::max(0LL, e->result_resident.load());
    return NDDO_OK;
}

int nddo_engine_info(nddo_engine* eh, size_t* lds_bytes, int* num_cus, size_t* hbm_budget, int* lds_tridiag_max_n) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e) return NDDO_BAD_ARG;
    if (lds_bytes) *lds_bytes = e->lds_max;
    if (num_cus) *num_cus = e->num_cus;
    if (hbm_budget) *hbm_budget = e->hbm_budget;
    if (lds_tridiag_max_n) {
        int n = 0;
        while (n < 512 && lds_eig_class(n + 1, e->lds_max)) ++n;
        *lds_tridiag_max_n = n;
    }
    return NDDO_OK;
}

int nddo_engine_profile(nddo_engine* eh, int enable) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e) return NDDO_BAD_ARG;
    SyncRun lk(e);
    e->profile = enable != 0;
    return NDDO_OK;
}

int nddo_engine_stage_times(nddo_engine* eh, double* ms, long long* launches, int reset) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e) return NDDO_BAD_ARG;
    SyncRun lk(e);
    for (int i = 0; i < NDDO_STAGE_COUNT; ++i) {
        if (ms) ms[i] = e->stage_ms[i];
        if (launches) launches[i] = e->stage_launches[i];
        if (reset) { e->stage_ms[i] = 0.0; e->stage_launches[i] = 0; }
    }
    return NDDO_OK;
}

int nddo_scf_once(nddo_engine* e, const nddo_job_desc* job, nddo_result* res) {
    int rc = nddo_scf_batch(e, 1, job, res);
    if (rc) return rc;
    if (res->status == NDDO_NOT_CONVERGED)
        set_error("Error in cndo::Cndo2::DoSCF: SCF did not met convergence criterion. maxIterationsSCF=" +
                  std::to_string(job->opts.max_iter));
    return res->status;
}

// ------------------------------------------------------------------ stage hooks
static int one_job(Engine* e, Batch& b, nddo_job_desc& d, int theory, int natoms, const int* Z, const double* xyz,
                   bool sdbg) {
    d.theory = theory;
    d.precision = NDDO_FP64;
    d.natoms = natoms;
    d.Z = Z;
    d.xyz_bohr = xyz;
    nddo_scf_opts_default(&d.opts);
    d.opts.diis_num_error_vect = 0;
    d.requires_guess = 1;
    d.P0 = nullptr;
    int n = 0, no = 0;
    int s = validate_job(d, &n, &no);
    if (s) return s;
    std::vector<int> which{0};
    b.build(&d, 1, which, nullptr);
    b.use_xl_householder = e->xl_householder;
    b.classify(e->lds_max, e->num_cus, e->xl_min_nao);
    int rc = b.allocate(sdbg);
    if (rc) return rc;
    return b.upload(e->stream, &d);
}

int nddo_dbg_integrals(nddo_engine* eh, int theory, int natoms, const int* Z, const double* xyz_bohr,
                       double* g_dense, double* S, double* H, double* e_core) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e) return NDDO_BAD_ARG;
    SyncRun lk(e);
    HIPCHK(hipSetDevice(e->device));
    Batch b;
    nddo_job_desc d;
    int rc = one_job(e, b, d, theory, natoms, Z, xyz_bohr, S != nullptr);
    if (rc) return rc;
    b.launch_setup_stage(e, e->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(e->stream));
    const int n = b.jobs[0].nao;
    const long long nn = (long long)n * n;
    if (H) HIPCHK(hipMemcpy(H, b.d_arena + (long long)AR_H * nn, nn * sizeof(double), hipMemcpyDeviceToHost));
    if (S) {
        HIPCHK(hipMemcpy(S, b.d_sdbg, nn * sizeof(double), hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) S[(long long)i * n + i] = 1.0;
    }
    if (e_core) {
        JobState s;
        HIPCHK(hipMemcpy(&s, b.d_state, sizeof(JobState), hipMemcpyDeviceToHost));
        *e_core = s.e_core;
    }
    if (g_dense) {
        std::vector<double> ints(std::max<long long>(b.nints, 1));
        HIPCHK(hipMemcpy(ints.data(), b.d_ints, b.nints * sizeof(double), hipMemcpyDeviceToHost));
        const int N = natoms;
        std::memset(g_dense, 0, sizeof(double) * (size_t)N * N * 256);
        static const int pk[10][2] = {{0, 0}, {0, 1}, {0, 2}, {0, 3}, {1, 1}, {1, 2}, {1, 3}, {2, 2}, {2, 3}, {3, 3}};
        std::vector<PairDev> pairs(std::max<long long>(b.npairs_all, 1));   // as k_build_pairs wrote them
        HIPCHK(hipMemcpy(pairs.data(), b.d_pairs, b.npairs_all * sizeof(PairDev), hipMemcpyDeviceToHost));
        pairs.resize(b.npairs_all);
        for (const PairDev& pr : pairs) {
            const int a = pr.a, bb = pr.b;
            const int nap = b.atoms[a].elem % kNumElem == 0 ? 1 : 10;
            const int nbp = b.atoms[bb].elem % kNumElem == 0 ? 1 : 10;
            for (int i = 0; i < nap; ++i)
                for (int j = 0; j < nbp; ++j) {
                    const double v = ints[pr.ioff + i * nbp + j];
                    const int mu = pk[i][0], nu = pk[i][1], la = pk[j][0], sg = pk[j][1];
                    auto G = [&](int A, int B, int p, int q, int r, int s) -> double& {
                        return g_dense[((((((size_t)A * N + B) * 4 + p) * 4 + q) * 4 + r) * 4 + s)];
                    };
                    G(a, bb, mu, nu, la, sg) = v; G(a, bb, nu, mu, la, sg) = v;
                    G(a, bb, mu, nu, sg, la) = v; G(a, bb, nu, mu, sg, la) = v;
                    G(bb, a, la, sg, mu, nu) = v; G(bb, a, sg, la, mu, nu) = v;
                    G(bb, a, la, sg, nu, mu) = v; G(bb, a, sg, la, nu, mu) = v;
                }
        }
    }
    return NDDO_OK;
}

int nddo_dbg_fock(nddo_engine* eh, int theory, int natoms, const int* Z, const double* xyz_bohr, const double* P,
                  int guess, double* F) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || !F) return NDDO_BAD_ARG;
    SyncRun lk(e);
    HIPCHK(hipSetDevice(e->device));
    Batch b;
    nddo_job_desc d;
    int rc = one_job(e, b, d, theory, natoms, Z, xyz_bohr, false);
    if (rc) return rc;
    const int n = b.jobs[0].nao;
    const long long nn = (long long)n * n;
    if (!guess) {
        b.jobs[0].requires_guess = 0;
        HIPCHK(hipMemcpy(b.d_jobs, b.jobs.data(), sizeof(JobDev), hipMemcpyHostToDevice));
    }
    if (P) HIPCHK(hipMemcpy(b.d_arena + (long long)AR_P0 * nn, P, nn * sizeof(double), hipMemcpyHostToDevice));
    b.launch_setup_stage(e, e->stream);
    launch_fock(e->stream, 0, b.fock_pairs(0, (int)b.jobs.size()), b.d_atoms, 0, (int)b.atoms.size(), e->d_elems,
                b.d_jobs, b.d_state, b.d_ints, b.d_ints32, b.d_arena, b.d_part);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(e->stream));
    HIPCHK(hipMemcpy(F, b.d_arena + (long long)AR_F * nn, nn * sizeof(double), hipMemcpyDeviceToHost));
    return NDDO_OK;
}

// A pseudo job of n AOs and no atoms, used by the eigensolver and density hooks.
// K jobs of one n x n matrix each (the stage-level test hooks and the autotuner).
static int matrix_batch(Engine* e, Batch& b, int n, int nocc, int K) {
    b.atoms.clear(); b.arows.clear(); b.npairs_all = 0; b.jobs.clear(); b.tiles.clear();
    for (int k = 0; k < K; ++k) {
        JobDev jb;
        std::memset(&jb, 0, sizeof(jb));
        jb.nao = n;
        jb.nocc = nocc;
        jb.want_c = 1;   // the hooks read every eigenvector
        jb.max_iter = 1;
        jb.rms_thresh = 0.0;
        jb.arena = (long long)k * ARENA_FIXED * n * n;
        jb.vec0 = (long long)k * n;
        jb.ivec0 = (long long)k * n;
        jb.traj0 = 4LL * k;
        jb.tile0 = (int)b.tiles.size();
        jb.tiles_per_dim = (n + 63) / 64;
        for (int i0 = 0; i0 < n; i0 += 64)
            for (int j0 = 0; j0 < n; j0 += 64) b.tiles.push_back({k, i0, j0, 0});
        jb.ntiles = (int)b.tiles.size() - jb.tile0;
        b.jobs.push_back(jb);
    }
    b.nslots = 0; b.nints = 0; b.nvec = (long long)K * n; b.ntraj = 4LL * K;
    b.arena_size = (long long)K * ARENA_FIXED * n * n;
    b.src.assign(K, 0);
    b.max_iter_all = 1;
    b.use_xl_householder = e->xl_householder;
    b.classify(e->lds_max, e->num_cus, e->xl_min_nao);
    int rc = b.allocate(false);
    if (rc) return rc;
    HIPCHK(hipMemcpy(b.d_jobs, b.jobs.data(), K * sizeof(JobDev), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b.d_tiles, b.tiles.data(), b.tiles.size() * sizeof(TileDev), hipMemcpyHostToDevice));
    if (!b.list_lds.empty())
        HIPCHK(hipMemcpy(b.d_list_lds, b.list_lds.data(), b.list_lds.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!b.list_glb.empty())
        HIPCHK(hipMemcpy(b.d_list_glb, b.list_glb.data(), b.list_glb.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!b.list_xl.empty()) {
        HIPCHK(hipMemcpy(b.d_list_xl, b.list_xl.data(), b.list_xl.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(b.d_xl_off, b.xl_off.data(), b.xl_off.size() * sizeof(long long), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemset(b.d_arena, 0, b.arena_size * sizeof(double)));
    std::vector<JobState> s(K);
    std::memset(s.data(), 0, K * sizeof(JobState));
    for (auto& x : s) x.phase = PH_ACTIVE;
    HIPCHK(hipMemcpy(b.d_state, s.data(), K * sizeof(JobState), hipMemcpyHostToDevice));
    return NDDO_OK;
}

static int matrix_job(Engine* e, Batch& b, int n, int nocc) { return matrix_batch(e, b, n, nocc, 1); }

// ------------------------------------------------------------------ autotuner
// Picks the launch shapes of Tuning (launch.h) when the engine is created, by
// timing every candidate with hipEvents on synthetic work: the workgroup size
// of the pair Fock kernel on water lattices (all three atom-pair classes, a
// non-guess iteration), and the row block of the back-transformation on
// random symmetric matrices of each size class.  Every candidate computes the
// same numbers; the choice only moves time.
static int time_ms(Engine* e, const std::function<void()>& f, int reps, float* best) {
    *best = 1e30f;
    for (int r = 0; r < reps; ++r) {
        HIPCHK(hipEventRecord(e->ev[14], e->stream));
        f();
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(e->ev[15], e->stream));
        HIPCHK(hipEventSynchronize(e->ev[15]));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e->ev[14], e->ev[15]));
        *best = std::min(*best, ms);
    }
    return NDDO_OK;
}

static int autotune(Engine* e) {
    Tuning t;
    set_tuning(t);
    const double ang = 1.0 / 0.52917721092;
    {   // pair Fock: 48 cubic water lattices of 27, 64 and 125 molecules (2.9 A apart)
        std::vector<std::vector<int>> Z(48);
        std::vector<std::vector<double>> X(48);
        std::vector<nddo_job_desc> d(48);
        for (int k = 0; k < 48; ++k) {
            const int m = 3 + k % 3;
            for (int i = 0; i < m * m * m; ++i) {
                const double o[3] = {2.9 * (i % m), 2.9 * ((i / m) % m), 2.9 * (i / (m * m))};
                const double h[2][3] = {{0.757, 0.586, 0.0}, {-0.757, 0.586, 0.0}};
                Z[k].push_back(8);
                for (int c = 0; c < 3; ++c) X[k].push_back(o[c] * ang);
                for (int a = 0; a < 2; ++a) {
                    Z[k].push_back(1);
                    for (int c = 0; c < 3; ++c) X[k].push_back((o[c] + h[a][c]) * ang);
                }
            }
            std::memset(&d[k], 0, sizeof(nddo_job_desc));
            d[k].theory = NDDO_PM3;
            d[k].precision = NDDO_FP64;
            d[k].natoms = (int)Z[k].size();
            d[k].Z = Z[k].data();
            d[k].xyz_bohr = X[k].data();
            d[k].requires_guess = 1;
            nddo_scf_opts_default(&d[k].opts);
        }
        std::vector<int> which(48);
        for (int k = 0; k < 48; ++k) which[k] = k;
        Batch b;
        int rc = b.build(d.data(), 48, which, nullptr);
        b.use_xl_householder = e->xl_householder;
        if (rc == NDDO_OK) rc = b.classify(e->lds_max, e->num_cus, e->xl_min_nao);
        if (rc == NDDO_OK) rc = b.allocate(false);
        if (rc == NDDO_OK) rc = b.upload(e->stream, d.data());
        if (rc != NDDO_OK) return rc;
        b.launch_setup_stage(e, e->stream);
        std::vector<JobState> st(48);
        std::memset(st.data(), 0, st.size() * sizeof(JobState));
        for (auto& x : st) { x.phase = PH_ACTIVE; x.iter = 1; }   // a regular (non-guess) Fock build
        HIPCHK(hipMemcpyAsync(b.d_state, st.data(), st.size() * sizeof(JobState), hipMemcpyHostToDevice, e->stream));
        // only the heavy-pair kernel (and the per-atom reduction) depends on the
        // setting: time them without the light pairs, the candidates
        // interleaved over two rounds, the fastest of each kept
        const int cand[4] = {1, 2, 4, 8};
        float ms_of[4] = {1e30f, 1e30f, 1e30f, 1e30f};
        for (int round = 0; round < 2; ++round)
            for (int ci = 0; ci < 4; ++ci) {
                Tuning c = t;
                c.fock_wpb = cand[ci];
                set_tuning(c);
                float ms = 0.f;
                FockPairs hp = b.fock_pairs(0, (int)b.jobs.size());
                hp.nlight[0] = hp.nlight[1] = 0;
                rc = time_ms(e, [&] {
                    for (int r = 0; r < 4; ++r)
                        launch_fock(e->stream, 0, hp, b.d_atoms, 0, (int)b.atoms.size(), e->d_elems,
                                    b.d_jobs, b.d_state, b.d_ints, b.d_ints32, b.d_arena, b.d_part);
                }, 3, &ms);
                if (rc != NDDO_OK) return rc;
                ms_of[ci] = std::min(ms_of[ci], ms);
            }
        // 4 pairs per workgroup unless another setting is clearly (5%) faster:
        // the candidates time within noise of each other on this lattice set,
        // while on the 4096-molecule PM3 mix 4 is the fastest (Fock 109 ms per
        // pass against 115, 124 and 124 ms for 2, 1 and 8)
        int bi = 2;
        for (int ci = 0; ci < 4; ++ci)
            if (ms_of[ci] < 0.95f * ms_of[bi]) bi = ci;
        t.fock_wpb = cand[bi];
        set_tuning(t);
    }
    // back-transformation row block per size class: 32 random symmetric matrices each
    const int sizes[4] = {128, 256, 384, 512};
    std::vector<double> F;
    for (int cls = 0; cls < 4; ++cls) {
        const int n = sizes[cls], K = 32;
        Batch b;
        int rc = matrix_batch(e, b, n, n / 2, K);
        if (rc != NDDO_OK) return rc;
        F.assign((size_t)n * n, 0.0);
        unsigned long long x = 0x9e3779b97f4a7c15ull * (cls + 1);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j <= i; ++j) {
                x ^= x << 13; x ^= x >> 7; x ^= x << 17;   // xorshift64
                const double v = (double)(x >> 11) * (1.0 / 9007199254740992.0) - 0.5;
                F[(size_t)i * n + j] = F[(size_t)j * n + i] = v;
            }
        for (int k = 0; k < K; ++k)
            HIPCHK(hipMemcpy(b.d_arena + b.jobs[k].arena + (long long)AR_F * n * n, F.data(), F.size() * sizeof(double),
                             hipMemcpyHostToDevice));
        const bool lds = !b.list_lds.empty();
        const int* list = lds ? b.d_list_lds : b.d_list_glb;
        const int nl = (int)(lds ? b.list_lds.size() : b.list_glb.size());
        for (int part = 0; part < 2; ++part)
            launch_heig_part(part, e->stream, b.d_jobs, b.d_state, list, nl, n, b.d_arena, b.d_eps, b.d_perm, b.d_tri,
                             lds);
        float best = 1e30f;
        for (int rows : {64, 32}) {
            Tuning c = t;
            c.bt_rows[cls] = rows;
            set_tuning(c);
            float ms = 0.f;
            rc = time_ms(e, [&] {
                launch_heig_part(2, e->stream, b.d_jobs, b.d_state, list, nl, n, b.d_arena, b.d_eps, b.d_perm, b.d_tri,
                                 lds);
            }, 3, &ms);
            if (rc != NDDO_OK) return rc;
            if (ms < best) { best = ms; t.bt_rows[cls] = rows; }
        }
        set_tuning(t);
    }
    return NDDO_OK;
}

int nddo_engine_tuning(nddo_engine* eh, int* fock_wpb, int* bt_rows4) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e) return NDDO_BAD_ARG;
    HIPCHK(hipSetDevice(e->device));
    const Tuning t = get_tuning();
    if (fock_wpb) *fock_wpb = t.fock_wpb;
    if (bt_rows4)
        for (int i = 0; i < 4; ++i) bt_rows4[i] = t.bt_rows[i];
    return NDDO_OK;
}

int nddo_dbg_set_tuning(nddo_engine* eh, int fock_wpb, const int* bt_rows4) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || !bt_rows4 || !(fock_wpb == 1 || fock_wpb == 2 || fock_wpb == 4 || fock_wpb == 8)) return NDDO_BAD_ARG;
    for (int i = 0; i < 4; ++i)
        if (bt_rows4[i] != 32 && bt_rows4[i] != 64) return NDDO_BAD_ARG;
    SyncRun lk(e);
    HIPCHK(hipSetDevice(e->device));
    Tuning t = get_tuning();
    t.fock_wpb = fock_wpb;
    for (int i = 0; i < 4; ++i) t.bt_rows[i] = bt_rows4[i];
    set_tuning(t);
    return NDDO_OK;
}

int nddo_dbg_set_tridiag(nddo_engine* eh, int blocked) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || (blocked != 0 && blocked != 1)) return NDDO_BAD_ARG;
    SyncRun lk(e);
    HIPCHK(hipSetDevice(e->device));
    Tuning t = get_tuning();
    t.tri_blocked = blocked;
    set_tuning(t);
    return NDDO_OK;
}

int nddo_dbg_set_xl_tridiag(nddo_engine* eh, int mode) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || mode < 0 || mode > 2) return NDDO_BAD_ARG;
    SyncRun lk(e);
    HIPCHK(hipSetDevice(e->device));
    Tuning t = get_tuning();
    t.xl_tri_blocked = mode;
    set_tuning(t);
    return NDDO_OK;
}

int nddo_dbg_set_xl_grid(nddo_engine* eh, int tri_workgroups) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || tri_workgroups < 0) return NDDO_BAD_ARG;
    SyncRun lk(e);
    HIPCHK(hipSetDevice(e->device));
    Tuning t = get_tuning();
    t.xl_tri_wg = tri_workgroups;
    set_tuning(t);
    return NDDO_OK;
}

int nddo_dbg_eigh(nddo_engine* eh, int n, const double* F, double* eps, double* C, int* sweeps) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || n <= 0 || !F) return NDDO_BAD_ARG;
    SyncRun lk(e);
    HIPCHK(hipSetDevice(e->device));
    Batch b;
    int rc = matrix_job(e, b, n, 0);
    if (rc) return rc;
    const long long nn = (long long)n * n;
    HIPCHK(hipMemcpy(b.d_arena + (long long)AR_F * nn, F, nn * sizeof(double), hipMemcpyHostToDevice));
    b.launch_lds_eigen(-1, e->stream);
    launch_eigen(e->stream, b.d_jobs, b.d_state, b.d_list_glb, (int)b.list_glb.size(), b.maxn_glb, b.d_arena, b.d_eps,
                 b.d_perm, b.d_tri, false);
    b.launch_xl_eigen_all(e, e->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(e->stream));
    std::vector<int> perm(n);
    std::vector<double> VT(nn);
    if (eps) HIPCHK(hipMemcpy(eps, b.d_eps, n * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(perm.data(), b.d_perm, n * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(VT.data(), b.d_arena + (long long)AR_C * nn, nn * sizeof(double), hipMemcpyDeviceToHost));
    JobState s;
    HIPCHK(hipMemcpy(&s, b.d_state, sizeof(s), hipMemcpyDeviceToHost));
    if (sweeps) *sweeps = s.sweeps;
    for (int i = 0; i < 4; ++i) e->last_eig_prof[i] = s.prof[i];
    if (b.list_xl.empty())
        HIPCHK(hipMemcpy(e->last_dc_phase, b.d_tri + 19LL * n + 224, 7 * sizeof(double), hipMemcpyDeviceToHost));
    else
        HIPCHK(hipMemcpy(e->last_dc_phase, b.d_xlws + xl_dbg_offset(n), 9 * sizeof(double), hipMemcpyDeviceToHost));
    if (C) {
        for (int r = 0; r < n; ++r) {
            const double* src = &VT[(long long)perm[r] * n];
            double sum = 0.0;
            for (int i = 0; i < n; ++i) sum += src[i];
            const double sg = sum < 0.0 ? -1.0 : 1.0;
            for (int i = 0; i < n; ++i) C[(long long)r * n + i] = sg * src[i];
        }
    }
    if (s.sweeps < 0) { set_error("Jacobi did not converge"); return NDDO_EIGENSOLVER_FAILED; }
    return NDDO_OK;
}

int nddo_dbg_eigh_profile(nddo_engine* eh, double* ticks4, double* tick_hz) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || !ticks4) return NDDO_BAD_ARG;
    for (int i = 0; i < 4; ++i) ticks4[i] = e->last_eig_prof[i];
    if (tick_hz) {
        int khz = 0;
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, e->device);
        *tick_hz = 1e3 * khz;
    }
    return NDDO_OK;
}

int nddo_dbg_host_times(nddo_engine* eh, double* t5, int reset) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || !t5) return NDDO_BAD_ARG;
    for (int i = 0; i < 5; ++i) {
        t5[i] = e->host_s[i];
        if (reset) e->host_s[i] = 0.0;
    }
    return NDDO_OK;
}

int nddo_dbg_graph_stats(nddo_engine* eh, long long* captured, long long* replays) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e) return NDDO_BAD_ARG;
    if (captured) *captured = e->graphs_captured.load();
    if (replays) *replays = e->graph_replays.load();
    return NDDO_OK;
}

int nddo_dbg_eigh_dc_phases(nddo_engine* eh, double* ticks7) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || !ticks7) return NDDO_BAD_ARG;
    for (int i = 0; i < 7; ++i) ticks7[i] = e->last_dc_phase[i];
    return NDDO_OK;
}

int nddo_dbg_density(nddo_engine* eh, int n, int nocc, const double* C, double* P) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || n <= 0 || nocc < 0 || nocc > n || !C || !P) return NDDO_BAD_ARG;
    SyncRun lk(e);
    HIPCHK(hipSetDevice(e->device));
    Batch b;
    int rc = matrix_job(e, b, n, nocc);
    if (rc) return rc;
    const long long nn = (long long)n * n;
    std::vector<int> perm(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    HIPCHK(hipMemcpy(b.d_perm, perm.data(), n * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b.d_arena + (long long)AR_C * nn, C, nn * sizeof(double), hipMemcpyHostToDevice));
    launch_density(e->stream, b.d_tiles, (int)b.tiles.size(), b.d_jobs, b.d_state, b.d_arena, b.d_perm, b.d_rmspart,
                   b.d_tpart);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(e->stream));
    HIPCHK(hipMemcpy(P, b.d_arena + (long long)AR_P1 * nn, nn * sizeof(double), hipMemcpyDeviceToHost));
    return NDDO_OK;
}

}  // extern "C"
