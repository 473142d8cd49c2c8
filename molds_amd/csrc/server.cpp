// Serving loop of the NDDO engine: continuous batching at SCF-iteration
// granularity.
//
// nddo_submit copies a request and queues it.  One scheduler thread owns a
// stream of its own and a set of resident cohorts; a cohort is a Batch (the
// same device layout and kernels as nddo_scf_batch) of requests admitted
// together.  Every turn of the scheduler
//   1. admits queued requests as a new cohort when the HBM budget (the sum of
//      job_footprint_bytes over resident jobs) and the cohort limit allow,
//      runs its setup stage (integrals, H, core repulsion),
//   2. launches one SCF iteration of every resident cohort (each job keeps
//      its geometry, integrals, density and DIIS ring resident in its
//      cohort's arena),
//   3. reads back every job's phase (one strided copy per cohort) and
//      retires each job the turn it finishes: its results are copied out,
//      its latency (submit -> retire) recorded and waiters woken,
//   4. frees a cohort's device memory once its last job has retired.
// Requests that arrive while others are on the GPU therefore start at the
// next iteration boundary and run interleaved with the resident ones; a
// finished job costs its cohort nothing more than an early exit per kernel.
//
// MolDS has no serving layer (one DoSCF per ElectronicStructure object,
// src/cndo/Cndo2.cpp:585-729); admission control plays the role of the host
// heap cap of src/base/MallocerFreer.cpp:62-73.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>

#include "engine_internal.h"
#include "launch.h"
#include "params.h"

namespace nddo {

using Clock = std::chrono::steady_clock;

struct ServedJob {
    long long id = 0;
    nddo_job_desc desc{};
    std::vector<int> Z;
    std::vector<double> xyz, P0;
    size_t footprint = 0;
    double hof_offset = 0.0;
    Clock::time_point t_submit, t_done;
    bool done = false;
    bool released = false;   // released by the caller before it finished
    nddo_result res{};
    std::vector<double> eps, C, P, traj;
};

struct Cohort {
    Batch b;
    std::vector<nddo_job_desc> descs;
    std::vector<ServedJob*> members;
    std::vector<char> retired;
    int* h_phase = nullptr;   // pinned, one int per job
    int live = 0;
    size_t bytes = 0;
    int turns = 0;
};

namespace {

constexpr int kMaxCohorts = 8;

int admit_limit(const Engine* e) { return e->opts.max_batch_jobs > 0 ? e->opts.max_batch_jobs : 4096; }

void retire(Engine* e, Cohort* c, int jj, hipStream_t st) {
    ServedJob* j = c->members[jj];
    nddo_result& r = j->res;
    r.eps = j->eps.data();
    r.C = j->C.data();
    r.P = j->P.data();
    r.trajectory = j->traj.data();
    const int rc = c->b.fetch(st, jj, c->descs[jj], &r, j->hof_offset);
    if (rc != NDDO_OK) r.status = rc;
    std::lock_guard<std::mutex> lk(e->smu);
    j->t_done = Clock::now();
    r.latency_s = std::chrono::duration<double>(j->t_done - j->t_submit).count();
    j->done = true;
    c->retired[jj] = 1;
    --c->live;
    if (j->released) {
        e->served.erase(j->id);
        delete j;
        c->members[jj] = nullptr;
    }
    e->scv.notify_all();
}

// Builds, uploads and sets up a cohort from the admitted jobs.
int start_cohort(Engine* e, Cohort* c, hipStream_t st) {
    std::vector<int> which(c->members.size());
    for (size_t i = 0; i < which.size(); ++i) which[i] = (int)i;
    c->descs.resize(c->members.size());
    for (size_t i = 0; i < c->members.size(); ++i) c->descs[i] = c->members[i]->desc;
    c->b.build(c->descs.data(), (int)c->descs.size(), which, nullptr);
    c->b.classify(e->lds_max);
    int rc = c->b.allocate(false);
    if (rc) return rc;
    rc = c->b.upload(st, c->descs.data());
    if (rc) return rc;
    c->b.launch_setup_stage(e, st);
    if (hipGetLastError() != hipSuccess) return NDDO_HIP_ERROR;
    if (hipHostMalloc((void**)&c->h_phase, c->members.size() * sizeof(int), hipHostMallocDefault) != hipSuccess)
        return NDDO_HIP_ERROR;
    c->retired.assign(c->members.size(), 0);
    c->live = (int)c->members.size();
    return NDDO_OK;
}

void fail_cohort(Engine* e, Cohort* c, int rc) {
    std::lock_guard<std::mutex> lk(e->smu);
    for (size_t i = 0; i < c->members.size(); ++i) {
        ServedJob* j = c->members[i];
        if (!j || (i < c->retired.size() && c->retired[i])) continue;
        j->res.status = rc;
        j->t_done = Clock::now();
        j->res.latency_s = std::chrono::duration<double>(j->t_done - j->t_submit).count();
        j->done = true;
        if (j->released) {
            e->served.erase(j->id);
            delete j;
        }
    }
    e->scv.notify_all();
}

void free_cohort(Cohort* c) {
    if (c->h_phase) hipHostFree(c->h_phase);
    c->h_phase = nullptr;
    c->b.release();
}

}  // namespace

void Engine::start_server() {
    stopping = false;
    server = std::thread([this]() {
        hipSetDevice(device);
        hipStream_t st = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return;
        std::vector<std::unique_ptr<Cohort>> cohorts;
        size_t resident = 0;
        for (;;) {
            // 1. admission
            std::unique_ptr<Cohort> fresh;
            {
                std::unique_lock<std::mutex> lk(smu);
                if (cohorts.empty()) scv.wait(lk, [&] { return stopping || !queue.empty(); });
                if (stopping) break;
                if (!queue.empty() && (int)cohorts.size() < kMaxCohorts) {
                    fresh.reset(new Cohort());
                    const int cap = admit_limit(this);
                    while (!queue.empty() && (int)fresh->members.size() < cap) {
                        ServedJob* j = queue.front();
                        if (resident + fresh->bytes + j->footprint > hbm_budget) break;
                        queue.pop_front();
                        fresh->members.push_back(j);
                        fresh->bytes += j->footprint;
                    }
                    if (fresh->members.empty()) fresh.reset();
                }
            }
            if (fresh) {
                const int rc = start_cohort(this, fresh.get(), st);
                if (rc != NDDO_OK) {
                    fail_cohort(this, fresh.get(), rc);
                    free_cohort(fresh.get());
                } else {
                    resident += fresh->bytes;
                    cohorts.push_back(std::move(fresh));
                }
            }
            if (cohorts.empty()) continue;
            // 2. one SCF iteration of every resident cohort, then the job phases
            for (auto& c : cohorts) {
                c->b.launch_iteration(this, st);
                hipMemcpy2DAsync(c->h_phase, sizeof(int), c->b.d_state, sizeof(JobState), sizeof(int),
                                 c->members.size(), hipMemcpyDeviceToHost, st);
                ++c->turns;
            }
            const hipError_t err = hipStreamSynchronize(st);
            // 3. retire finished jobs; 4. free finished cohorts
            for (auto& c : cohorts) {
                if (err != hipSuccess) {
                    fail_cohort(this, c.get(), NDDO_HIP_ERROR);
                    c->live = 0;
                    continue;
                }
                for (size_t jj = 0; jj < c->members.size(); ++jj)
                    if (!c->retired[jj] && c->h_phase[jj] == PH_DONE) retire(this, c.get(), (int)jj, st);
                // every job ends within max_iter + 1 turns (k_step marks NOT_CONVERGED)
                if (c->live > 0 && c->turns > c->b.max_iter_all + 2) {
                    for (size_t jj = 0; jj < c->members.size(); ++jj)
                        if (!c->retired[jj]) retire(this, c.get(), (int)jj, st);
                }
            }
            for (size_t i = 0; i < cohorts.size();) {
                if (cohorts[i]->live <= 0) {
                    resident -= cohorts[i]->bytes;
                    free_cohort(cohorts[i].get());
                    cohorts.erase(cohorts.begin() + i);
                } else {
                    ++i;
                }
            }
        }
        for (auto& c : cohorts) {
            fail_cohort(this, c.get(), NDDO_BAD_ARG);
            free_cohort(c.get());
        }
        hipStreamDestroy(st);
    });
}

void Engine::stop_server() {
    {
        std::lock_guard<std::mutex> lk(smu);
        stopping = true;
    }
    scv.notify_all();
    if (server.joinable()) server.join();
    std::lock_guard<std::mutex> lk(smu);
    for (auto& kv : served) delete kv.second;
    served.clear();
    queue.clear();
}

}  // namespace nddo

using nddo::Engine;
using nddo::ServedJob;

extern "C" {

int nddo_submit(nddo_engine* eh, const nddo_job_desc* d, nddo_job_id* id) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || !d || !id) { nddo::set_error("bad arguments"); return NDDO_BAD_ARG; }
    int n = 0, no = 0;
    const int s = nddo::validate_job(*d, &n, &no);
    ServedJob* j = new ServedJob();
    j->t_submit = nddo::Clock::now();
    j->Z.assign(d->Z ? d->Z : nullptr, d->Z ? d->Z + std::max(d->natoms, 0) : nullptr);
    j->xyz.assign(d->xyz_bohr ? d->xyz_bohr : nullptr, d->xyz_bohr ? d->xyz_bohr + 3 * std::max(d->natoms, 0) : nullptr);
    j->desc = *d;
    j->desc.Z = j->Z.data();
    j->desc.xyz_bohr = j->xyz.data();
    if (s == NDDO_OK && !d->requires_guess && d->P0) {
        j->P0.assign(d->P0, d->P0 + (size_t)n * n);
        j->desc.P0 = j->P0.data();
    } else {
        j->desc.P0 = nullptr;
    }
    j->res.status = s;
    j->res.nao = n;
    j->res.nocc = no;
    if (s == NDDO_OK) {
        j->footprint = nddo::job_footprint_bytes(d->natoms, n, d->opts.diis_num_error_vect, d->opts.max_iter);
        for (int a = 0; a < d->natoms; ++a) j->hof_offset += nddo::mndo_atom_hof_offset(d->Z[a]);
        j->eps.resize(n);
        j->C.resize((size_t)n * n);
        j->P.resize((size_t)n * n);
        j->traj.assign(4 * (size_t)d->opts.max_iter, 0.0);
        j->res.status = NDDO_PENDING;
    }
    std::lock_guard<std::mutex> lk(e->smu);
    j->id = e->next_id++;
    *id = j->id;
    e->served[j->id] = j;
    if (s == NDDO_OK && j->footprint > e->hbm_budget) {
        j->res.status = NDDO_ADMISSION_DENIED;
        nddo::set_error("job footprint exceeds the engine's HBM budget");
    }
    if (j->res.status == NDDO_PENDING) {
        e->queue.push_back(j);
        e->scv.notify_all();
    } else {
        j->done = true;
        j->t_done = j->t_submit;
    }
    return NDDO_OK;
}

int nddo_poll(nddo_engine* eh, nddo_job_id id) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e) return NDDO_BAD_ARG;
    std::lock_guard<std::mutex> lk(e->smu);
    auto it = e->served.find(id);
    if (it == e->served.end() || it->second->released) { nddo::set_error("unknown job id"); return NDDO_BAD_ARG; }
    return it->second->done ? it->second->res.status : NDDO_PENDING;
}

int nddo_wait(nddo_engine* eh, nddo_job_id id, double timeout_s) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e) return NDDO_BAD_ARG;
    std::unique_lock<std::mutex> lk(e->smu);
    auto it = e->served.find(id);
    if (it == e->served.end() || it->second->released) { nddo::set_error("unknown job id"); return NDDO_BAD_ARG; }
    ServedJob* j = it->second;
    const auto pred = [&] { return j->done; };
    if (timeout_s < 0) e->scv.wait(lk, pred);
    else if (!e->scv.wait_for(lk, std::chrono::duration<double>(timeout_s), pred)) return NDDO_PENDING;
    return j->res.status;
}

int nddo_get_result(nddo_engine* eh, nddo_job_id id, nddo_result* res) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || !res) return NDDO_BAD_ARG;
    std::lock_guard<std::mutex> lk(e->smu);
    auto it = e->served.find(id);
    if (it == e->served.end() || it->second->released) { nddo::set_error("unknown job id"); return NDDO_BAD_ARG; }
    ServedJob* j = it->second;
    if (!j->done) return NDDO_PENDING;
    double* eps = res->eps;
    double* C = res->C;
    double* P = res->P;
    double* tr = res->trajectory;
    *res = j->res;
    res->eps = eps;
    res->C = C;
    res->P = P;
    res->trajectory = tr;
    const size_t n = (size_t)std::max(j->res.nao, 0);
    const bool have = j->res.status != NDDO_ODD_ELECTRONS && j->res.status != NDDO_UNSUPPORTED_ELEMENT &&
                      j->res.status != NDDO_BAD_ARG && j->res.status != NDDO_ADMISSION_DENIED && !j->eps.empty();
    if (have) {
        if (eps) std::memcpy(eps, j->eps.data(), n * sizeof(double));
        if (C) std::memcpy(C, j->C.data(), n * n * sizeof(double));
        if (P) std::memcpy(P, j->P.data(), n * n * sizeof(double));
        if (tr) std::memcpy(tr, j->traj.data(), j->traj.size() * sizeof(double));
    }
    return j->res.status;
}

int nddo_release(nddo_engine* eh, nddo_job_id id) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e) return NDDO_BAD_ARG;
    std::lock_guard<std::mutex> lk(e->smu);
    auto it = e->served.find(id);
    if (it == e->served.end() || it->second->released) { nddo::set_error("unknown job id"); return NDDO_BAD_ARG; }
    ServedJob* j = it->second;
    auto q = std::find(e->queue.begin(), e->queue.end(), j);
    if (q != e->queue.end()) {   // not admitted yet: cancel
        e->queue.erase(q);
        e->served.erase(it);
        delete j;
        return NDDO_OK;
    }
    if (!j->done) {   // running: dropped when it retires
        j->released = true;
        return NDDO_OK;
    }
    e->served.erase(it);
    delete j;
    return NDDO_OK;
}

}  // extern "C"
