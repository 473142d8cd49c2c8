// Serving loop of the NDDO engine: continuous batching at SCF-iteration
// granularity.
//
// nddo_submit copies a request and queues it.  One scheduler thread keeps a
// set of resident cohorts; a cohort is a Batch (the same device layout and
// kernels as nddo_scf_batch) of requests admitted together, grouped by size,
// with a stream, an event and pinned buffers from the scheduler's pools, so
// the kernels of different cohorts overlap on the device.  The loop:
//   1. admits queued requests as new cohorts (one per size class) when the HBM
//      budget (the sum of job_footprint_bytes over resident jobs), the pool and
//      the cohort slots allow (slots kept free for small molecules, see
//      kSmallReserve); uploads through pinned staging and runs the setup stage;
//   2. launches the next SCF iteration of every cohort that is not running one
//      (each job keeps its geometry, integrals, density and DIIS ring resident
//      in its cohort's arena), then copies the jobs' scalar state (phase,
//      iteration, RMS, energies) to pinned memory and records the cohort event;
//   3. polls the events; for a cohort whose iteration finished, every job that
//      reached its end is retired at once: its arrays (eps, C, P, trajectory)
//      are copied device-to-device into a result buffer from the
//      stream-ordered pool, its scalars come from the pinned copy, and its
//      latency (submit -> retire) is recorded -- the scheduler never waits on
//      a host copy;
//   4. returns a cohort's memory to the pool once its last job has retired.
// nddo_get_result copies a retired job's arrays to the host in the caller's
// thread.  A request that arrives while others are on the GPU starts at once
// in a new cohort and runs beside the resident ones; a finished job costs its
// cohort nothing more than an early exit per kernel.
//
// MolDS has no serving layer (one DoSCF per ElectronicStructure object,
// src/cndo/Cndo2.cpp:585-729); admission control plays the role of the host
// heap cap of src/base/MallocerFreer.cpp:62-73.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <memory>

#include "engine_internal.h"
#include "launch.h"
#include "params.h"

namespace nddo {

using Clock = std::chrono::steady_clock;

namespace {

// Leading part of JobState the scheduler reads every iteration: the int
// fields and rms, diis_error, e_el, e_core.
constexpr size_t kScalarBytes = offsetof(JobState, prof);
static_assert(kScalarBytes % 4 == 0 && sizeof(JobState) % 4 == 0, "launch_publish copies 32-bit words");

struct EventRef {
    hipEvent_t ev = nullptr;
    ~EventRef() {
        if (ev) hipEventDestroy(ev);
    }
};

}  // namespace

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
    // arrays of a retired job: on the device until the first nddo_get_result
    std::mutex mu;
    double* d_res = nullptr;   // eps[n] | VT[n*n] | P[n*n] | traj[4*max_iter] | perm[n] (int)
    std::shared_ptr<EventRef> res_ev;
    bool on_host = false;
    std::vector<double> eps, C, P, traj;
    std::string err;   // why the job failed, reported to the caller by nddo_get_result / nddo_wait
};

struct Cohort {
    Batch b;
    int slot = -1;   // index into the scheduler's stream / event / pinned-buffer pools
    int cls = 0;     // size class of its members
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    char* h_state = nullptr;   // pinned, fine-grained: kScalarBytes per job
    char* d_hstate = nullptr;  // its device address (written by launch_publish)
    bool inflight = false;
    std::vector<nddo_job_desc> descs;
    std::vector<ServedJob*> members;
    std::vector<char> retired;
    int live = 0;
    size_t bytes = 0;
    int turns = 0;
};

namespace {

double secs_since(Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); }

// Cohorts run on streams of their own, and HIP places streams on a few
// hardware queues (GPU_MAX_HW_QUEUES, 4 by default), where a kernel waits for
// every kernel queued before it. A cohort of large molecules holds its slot for
// seconds (up to max_iter iterations of tens of ms), one of small molecules
// for milliseconds, so at most kMaxCohorts - kSmallReserve resident cohorts
// may hold molecules above 137 AOs, the rest of the slots are kept for the
// small size classes, and the two kinds run on different hardware queues
// (large_slot). Measured on the bench's Poisson
// mix at 500 requests/s (mixed precision): 16 slots, no reserve: 33.8 solves/s,
// p50 2.26 s; 12 slots, 4 reserved: 34.6, 0.87 s; 8 slots, 4 reserved: 38.3,
// 0.40 s (requests up to 137 AOs 0.14 s); 32 slots, 8 reserved: 27.5, 3.2 s.
constexpr int kMaxCohorts = 8;
constexpr int kSmallReserve = 4;
constexpr int kFirstLargeClass = 2;
// HIP gives a new stream the least-used hardware queue, so the slot streams,
// made one after another, take the queues in turn: slots k and k + 4 share
// one, slots k .. k + 3 are on four.  Cohorts of the large size classes take
// the slots of two queues ({0, 1, 4, 5}), those of the small classes the
// other two ({2, 3, 6, 7}), so a small cohort's kernels never queue behind a
// large cohort's device-wide launches; a small cohort borrows a large slot
// only while no large cohort is resident.
inline bool large_slot(int k) { return (k & 2) == 0; }
// AO count <= 64, <= 137 (eigensolver matrix in LDS), <= 256, <= 512, up to
// the device-wide eigensolver's threshold (one workgroup per molecule; empty
// in served mode, whose threshold is 513), larger (device-wide eigensolver,
// whose launches of different cohorts run one after another; at most
// xl_max_jobs() molecules per cohort, solved side by side)
constexpr int kSizeClasses = 6;

// Served requests above 512 AOs take the device-wide eigensolver: one
// workgroup per molecule would hold such a request for 50-190 ms per SCF
// iteration on one CU, the device-wide solver takes about 20-30 ms with the
// molecules of a cohort (at most xl_max_jobs()) side by side.
constexpr int kServedXlMin = 513;
int served_xl_min(const Engine* e) { return std::min(e->xl_min_nao, kServedXlMin); }

int size_class(int nao, int xl_min_nao) {
    if (nao >= xl_min_nao) return 5;
    return nao <= 64 ? 0 : (nao <= 137 ? 1 : (nao <= 256 ? 2 : (nao <= 512 ? 3 : 4)));
}

int admit_limit(const Engine* e) { return e->opts.max_batch_jobs > 0 ? e->opts.max_batch_jobs : 4096; }

}  // namespace

size_t result_bytes(int n, int max_iter) {
    return ((size_t)n + 2 * (size_t)n * n + 4 * (size_t)max_iter) * sizeof(double) + (size_t)n * sizeof(int);
}

namespace {

// Retires the finished jobs `done` of cohort c: device-to-device copies of
// their arrays, one event after them, then the scalars and the wake-up.
void retire_jobs(Engine* e, Cohort* c, const std::vector<int>& done) {
    std::shared_ptr<EventRef> ev;
    std::vector<double*> bufs(done.size(), nullptr);
    for (size_t k = 0; k < done.size(); ++k) {
        const int jj = done[k];
        ServedJob* j = c->members[jj];
        if (!j || j->released) continue;
        JobState s;
        std::memcpy(&s, c->h_state + (size_t)jj * kScalarBytes, kScalarBytes);
        const JobDev& jb = c->b.jobs[jj];
        const long long n = jb.nao, nn = n * n;
        double* d = nullptr;
        if (hipMallocAsync((void**)&d, result_bytes((int)n, jb.max_iter), c->st) != hipSuccess) continue;
        e->result_resident += (long long)result_bytes((int)n, jb.max_iter);
        const double* base = c->b.d_arena + jb.arena;
        const int pk = AR_P0 + ((s.iter + 1) & 1);
        hipMemcpyAsync(d, c->b.d_eps + jb.vec0, n * sizeof(double), hipMemcpyDeviceToDevice, c->st);
        hipMemcpyAsync(d + n, base + (long long)AR_C * nn, nn * sizeof(double), hipMemcpyDeviceToDevice, c->st);
        hipMemcpyAsync(d + n + nn, base + (long long)pk * nn, nn * sizeof(double), hipMemcpyDeviceToDevice, c->st);
        hipMemcpyAsync(d + n + 2 * nn, c->b.d_traj + jb.traj0, 4LL * jb.max_iter * sizeof(double),
                       hipMemcpyDeviceToDevice, c->st);
        hipMemcpyAsync(d + n + 2 * nn + 4LL * jb.max_iter, c->b.d_perm + jb.ivec0, n * sizeof(int),
                       hipMemcpyDeviceToDevice, c->st);
        bufs[k] = d;
        if (!ev) {
            ev = std::make_shared<EventRef>();
            if (hipEventCreateWithFlags(&ev->ev, hipEventDisableTiming) != hipSuccess) ev->ev = nullptr;
        }
    }
    if (ev && ev->ev) hipEventRecord(ev->ev, c->st);
    else if (ev) hipStreamSynchronize(c->st);   // no event: the copies are complete before the jobs show done
    std::lock_guard<std::mutex> lk(e->smu);
    const auto now = Clock::now();
    for (size_t k = 0; k < done.size(); ++k) {
        const int jj = done[k];
        ServedJob* j = c->members[jj];
        c->retired[jj] = 1;
        --c->live;
        if (!j) continue;
        if (j->released) {   // dropped by the caller while it ran
            if (bufs[k]) {
                hipFreeAsync(bufs[k], c->st);
                e->result_resident -= (long long)result_bytes(c->b.jobs[jj].nao, c->b.jobs[jj].max_iter);
            }
            e->served.erase(j->id);
            delete j;
            c->members[jj] = nullptr;
            continue;
        }
        JobState s;
        std::memcpy(&s, c->h_state + (size_t)jj * kScalarBytes, kScalarBytes);
        nddo_result& r = j->res;
        fill_scalars(c->b.jobs[jj], s, j->desc.theory, j->hof_offset, &r);
        if (!bufs[k] && r.status == NDDO_OK) {
            r.status = NDDO_HIP_ERROR;
            j->err = "result buffer allocation failed";
        }
        r.latency_s = std::chrono::duration<double>(now - j->t_submit).count();
        j->d_res = bufs[k];
        j->res_ev = ev;
        j->t_done = now;
        j->done = true;
    }
    e->scv.notify_all();
}

// Pinned upload staging of one scheduler slot; grows on demand.
struct Staging {
    char* p = nullptr;
    size_t cap = 0;
    int ensure(size_t need, hipStream_t st) {
        if (need <= cap) return NDDO_OK;
        const size_t want = std::max(need, 2 * cap);
        if (p) {
            hipStreamSynchronize(st);   // the slot's previous upload has been consumed
            hipHostFree(p);
        }
        p = nullptr;
        cap = 0;
        if (hipHostMalloc((void**)&p, want, hipHostMallocDefault) != hipSuccess) return NDDO_HIP_ERROR;
        cap = want;
        return NDDO_OK;
    }
};

// Builds, uploads and sets up a cohort from the admitted jobs on its own
// stream; its device memory comes from the stream-ordered pool and its upload
// goes through the slot's pinned staging.
int start_cohort(Engine* e, Cohort* c, Staging& stage) {
    hipStream_t st = c->st;
    c->b.alloc_stream = st;
    c->b.served = true;
    std::vector<int> which(c->members.size());
    for (size_t i = 0; i < which.size(); ++i) which[i] = (int)i;
    c->descs.resize(c->members.size());
    for (size_t i = 0; i < c->members.size(); ++i) c->descs[i] = c->members[i]->desc;
    c->b.build(c->descs.data(), (int)c->descs.size(), which, nullptr);
    c->b.use_xl_householder = e->xl_householder;
    c->b.classify(e->lds_max, e->num_cus, served_xl_min(e));
    int rc = stage.ensure(c->b.upload_bytes(), st);
    if (rc) return rc;
    rc = c->b.allocate(false);
    if (rc) return rc;
    rc = c->b.upload(st, c->descs.data(), stage.p);
    if (rc) return rc;
    c->b.launch_setup_stage(e, st);
    const hipError_t le = hipGetLastError();
    if (le != hipSuccess) {
        nddo::set_error(std::string("HIP error ") + hipGetErrorString(le) + " after the cohort's setup launches");
        return NDDO_HIP_ERROR;
    }
    c->retired.assign(c->members.size(), 0);
    c->live = (int)c->members.size();
    return NDDO_OK;
}

void fail_cohort(Engine* e, Cohort* c, int rc, const std::string& why) {
    std::lock_guard<std::mutex> lk(e->smu);
    for (size_t i = 0; i < c->members.size(); ++i) {
        ServedJob* j = c->members[i];
        if (!j || (i < c->retired.size() && c->retired[i])) continue;
        j->res.status = rc;
        j->err = why;
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

// Returns the cohort's device memory to the pool (stream-ordered frees); the
// stream, event and pinned buffers go back to the scheduler's pools.
void free_cohort(Cohort* c) { c->b.release(); }

// Copies a retired job's arrays to the host (once), applying MolDS's row order
// and sign rule to C.
// Runs on the caller's thread: it waits on the retire event and copies and
// frees on the engine's API stream, never on the cohort stream.
int materialize(Engine* e, ServedJob* j) {
    if (j->on_host || !j->d_res) return NDDO_OK;
    const int n = j->res.nao, mi = j->desc.opts.max_iter;
    const size_t nn = (size_t)n * n;
    // wait for the cohort's result copy by polling its event, each query (and
    // the copy below) outside any capture window of the scheduler thread
    if (j->res_ev && j->res_ev->ev) {
        for (;;) {
            hipError_t q;
            {
                std::lock_guard<std::recursive_mutex> cap(e->capture_mu);
                q = hipEventQuery(j->res_ev->ev);
            }
            if (q == hipSuccess) break;
            if (q != hipErrorNotReady) {
                nddo::set_error(std::string("result event: ") + hipGetErrorString(q));
                return NDDO_HIP_ERROR;
            }
            std::this_thread::sleep_for(std::chrono::microseconds(50));
        }
    }
    std::vector<char> h(result_bytes(n, mi));
    {
        std::lock_guard<std::recursive_mutex> cap(e->capture_mu);
        hipError_t err = hipMemcpyAsync(h.data(), j->d_res, h.size(), hipMemcpyDeviceToHost, e->api_stream);
        if (err == hipSuccess) err = hipStreamSynchronize(e->api_stream);
        if (err == hipSuccess) err = hipFreeAsync(j->d_res, e->api_stream);
        if (err != hipSuccess) {
            nddo::set_error(std::string("result copy: ") + hipGetErrorString(err));
            return NDDO_HIP_ERROR;
        }
    }
    const double* hd = (const double*)h.data();
    j->eps.assign(hd, hd + n);
    j->C.resize(nn);
    order_and_phase(n, hd + n, (const int*)(hd + n + 2 * nn + 4 * (size_t)mi), j->C.data());
    j->P.assign(hd + n + nn, hd + n + 2 * nn);
    j->traj.assign(hd + n + 2 * nn, hd + n + 2 * nn + 4 * (size_t)mi);
    j->d_res = nullptr;
    e->result_resident -= (long long)result_bytes(n, mi);
    j->res_ev.reset();
    j->on_host = true;
    return NDDO_OK;
}

}  // namespace

void Engine::start_server() {
    stopping = false;
    server = std::thread([this]() {
        hipSetDevice(device);
        // keep freed cohort memory in the stream-ordered pool, where later
        // cohorts reuse it: device memory that an instantiated iteration graph
        // has used did not come back to the device when it was freed (measured
        // with hipMemGetInfo), so memory released by the pool would be lost to
        // the process.  The reserve goes back only on request (trim_pool: the
        // synchronous batch path when it needs the room, nddo_engine_trim).
        hipMemPool_t pool;
        if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
            uint64_t keep = UINT64_MAX;
            hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
        const int cap = admit_limit(this);
        std::vector<hipStream_t> spool(kMaxCohorts, nullptr);
        std::vector<hipEvent_t> epool(kMaxCohorts, nullptr);
        std::vector<char*> hstate(kMaxCohorts, nullptr), hstate_dev(kMaxCohorts, nullptr);
        std::vector<Staging> stage(kMaxCohorts);
        std::vector<int> free_slots;
        for (int k = kMaxCohorts - 1; k >= 0; --k) {
            if (hipStreamCreateWithFlags(&spool[k], hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&epool[k], hipEventDisableTiming) != hipSuccess ||
                hipHostMalloc((void**)&hstate[k], (size_t)cap * kScalarBytes, hipHostMallocCoherent) != hipSuccess ||
                hipHostGetDevicePointer((void**)&hstate_dev[k], hstate[k], 0) != hipSuccess)
                continue;
            free_slots.push_back(k);
        }
        std::vector<std::unique_ptr<Cohort>> cohorts;
        size_t resident = 0;
        for (;;) {
            auto t_phase = Clock::now();
            // 1. admission: queued requests are grouped by size (AO count) so that a
            //    cohort's iteration time is set by molecules of like size
            std::vector<std::unique_ptr<Cohort>> fresh;
            bool starved = false;   // requests wait and nothing is resident to free memory
            {
                std::unique_lock<std::mutex> lk(smu);
                if (cohorts.empty()) scv.wait(lk, [&] { return stopping || !queue.empty(); });
                if (stopping) break;
                std::unique_ptr<Cohort> bucket[kSizeClasses];
                size_t admitted = 0;
                // the budget, the device memory free right now (a synchronous
                // batch's workspace or other users may hold some) and the bytes
                // of retired results not yet fetched, less 1 GiB for the runtime
                const size_t results = (size_t)std::max(0LL, result_resident.load());
                const size_t margin = (size_t)1 << 30;
                const size_t avail = queue.empty() ? 0 : device_available(device);
                const size_t room_dev = avail > margin ? avail - margin : 0;
                const size_t used = resident + results;
                const size_t room_budget = hbm_budget > used ? hbm_budget - used : 0;
                const size_t room = std::min(room_dev, room_budget);
                int nlarge = 0;   // resident and new cohorts of the large size classes
                for (auto& c : cohorts) nlarge += c->cls >= kFirstLargeClass;
                int free_l = 0, free_s = 0;   // free slots of the large / small classes' queues
                for (int k : free_slots) (large_slot(k) ? free_l : free_s) += 1;
                for (auto it = queue.begin(); it != queue.end();) {
                    ServedJob* j = *it;
                    const int k = size_class(j->res.nao, served_xl_min(this));
                    if (!bucket[k]) {
                        if (k >= kFirstLargeClass) {
                            if (free_l == 0 || nlarge >= kMaxCohorts - kSmallReserve) { ++it; continue; }
                            --free_l;
                        } else if (free_s > 0) {
                            --free_s;
                        } else if (free_l > 0 && nlarge == 0) {
                            --free_l;
                        } else {
                            ++it;
                            continue;
                        }
                        bucket[k].reset(new Cohort());
                        bucket[k]->cls = k;
                        nlarge += k >= kFirstLargeClass;
                    }
                    Cohort* c = bucket[k].get();
                    const int kcap = k == kSizeClasses - 1 ? std::min(cap, xl_max_jobs()) : cap;
                    if ((int)c->members.size() >= kcap || admitted + j->footprint > room) {
                        ++it;
                        continue;
                    }
                    c->members.push_back(j);
                    c->bytes += j->footprint;
                    admitted += j->footprint;
                    it = queue.erase(it);
                }
                for (auto& b : bucket)
                    if (b && !b->members.empty()) fresh.push_back(std::move(b));
                starved = fresh.empty() && cohorts.empty() && !queue.empty();
            }
            if (starved) {
                // nothing resident will free memory: return an idle synchronous
                // batch workspace and the pool's reserve, then try again
                std::unique_lock<std::mutex> rl(run_mu, std::try_to_lock);
                if (rl.owns_lock() && d_ws) free_workspace(this);
                if (rl.owns_lock()) trim_pool(device);
                std::this_thread::sleep_for(std::chrono::milliseconds(1));
                continue;
            }
            // large cohorts first on their queues' slots, then the small ones on
            // theirs and, when admission allowed it, on a large queue's slot
            std::stable_sort(fresh.begin(), fresh.end(), [](const std::unique_ptr<Cohort>& a,
                                                            const std::unique_ptr<Cohort>& b) {
                return (a->cls >= kFirstLargeClass) > (b->cls >= kFirstLargeClass);
            });
            for (auto& f : fresh) {
                const bool lg = f->cls >= kFirstLargeClass;
                auto pick = free_slots.end();
                for (auto sl = free_slots.begin(); sl != free_slots.end(); ++sl)
                    if (large_slot(*sl) == lg) { pick = sl; break; }
                if (pick == free_slots.end()) pick = free_slots.begin();   // a small cohort on a large slot
                const int k = *pick;
                free_slots.erase(pick);
                f->slot = k;
                f->st = spool[k];
                f->ev = epool[k];
                f->h_state = hstate[k];
                f->d_hstate = hstate_dev[k];
                int rc;
                {
                    std::lock_guard<std::recursive_mutex> cap(capture_mu);
                    rc = start_cohort(this, f.get(), stage[k]);
                }
                if (rc != NDDO_OK) {
                    fail_cohort(this, f.get(), rc, "cohort start: " + nddo::get_error());
                    {
                        std::lock_guard<std::recursive_mutex> cap(capture_mu);
                        free_cohort(f.get());
                    }
                    free_slots.push_back(k);
                } else {
                    resident += f->bytes;
                    cohorts.push_back(std::move(f));
                }
            }
            server_s[0] += secs_since(t_phase);
            t_phase = Clock::now();
            if (cohorts.empty()) continue;
            // 2. the next SCF iteration of every idle cohort, then its jobs' scalar state
            for (auto& c : cohorts) {
                if (c->inflight || c->live <= 0) continue;
                std::lock_guard<std::recursive_mutex> cap(capture_mu);
                c->b.iterate(this, c->st);
                launch_publish(c->st, c->b.d_state, sizeof(JobState), c->d_hstate, kScalarBytes, kScalarBytes,
                               (int)c->members.size());
                hipEventRecord(c->ev, c->st);
                c->inflight = true;
                ++c->turns;
            }
            server_s[1] += secs_since(t_phase);
            t_phase = Clock::now();
            // 3. wait until some cohort finished its iteration (or new requests can be admitted)
            std::vector<Cohort*> ready;
            for (;;) {
                for (auto& c : cohorts) {
                    if (!c->inflight) continue;
                    hipError_t q;
                    {
                        std::lock_guard<std::recursive_mutex> cap(capture_mu);
                        q = hipEventQuery(c->ev);
                    }
                    if (q == hipSuccess) {
                        ready.push_back(c.get());
                    } else if (q != hipErrorNotReady) {
                        fail_cohort(this, c.get(), NDDO_HIP_ERROR,
                                    std::string("HIP error ") + hipGetErrorString(q) + " in a cohort iteration");
                        c->live = 0;
                        c->inflight = false;
                    }
                }
                if (!ready.empty()) break;
                {
                    std::lock_guard<std::mutex> lk(smu);
                    if (stopping || (!queue.empty() && !free_slots.empty())) break;
                }
                std::this_thread::sleep_for(std::chrono::microseconds(20));
            }
            server_s[2] += secs_since(t_phase);
            t_phase = Clock::now();
            server_turns += (long long)ready.size();
            for (Cohort* c : ready) {
                c->inflight = false;
                if (c->live <= 0) continue;
                std::vector<int> done;
                for (size_t jj = 0; jj < c->members.size(); ++jj) {
                    if (c->retired[jj]) continue;
                    int phase;
                    std::memcpy(&phase, c->h_state + jj * kScalarBytes + offsetof(JobState, phase), sizeof(int));
                    // every job ends within max_iter + 1 iterations (k_step marks NOT_CONVERGED)
                    if (phase == PH_DONE || c->turns > c->b.max_iter_all + 2) done.push_back((int)jj);
                }
                if (!done.empty()) {
                    std::lock_guard<std::recursive_mutex> cap(capture_mu);
                    retire_jobs(this, c, done);
                }
            }
            server_s[3] += secs_since(t_phase);
            // 4. free finished cohorts
            for (size_t i = 0; i < cohorts.size();) {
                if (cohorts[i]->live <= 0 && !cohorts[i]->inflight) {
                    resident -= cohorts[i]->bytes;
                    {
                        std::lock_guard<std::recursive_mutex> cap(capture_mu);
                        free_cohort(cohorts[i].get());
                    }
                    free_slots.push_back(cohorts[i]->slot);
                    cohorts.erase(cohorts.begin() + i);
                } else {
                    ++i;
                }
            }
        }
        for (auto& c : cohorts) {
            hipStreamSynchronize(c->st);
            fail_cohort(this, c.get(), NDDO_BAD_ARG, "engine closed before the job finished");
            free_cohort(c.get());
        }
        for (int k = 0; k < kMaxCohorts; ++k)
            if (spool[k]) hipStreamSynchronize(spool[k]);
        {   // results not yet copied to the host go back to the pool now
            std::lock_guard<std::mutex> lk(smu);
            for (auto& kv : served)
                if (kv.second->d_res) {
                    hipFree(kv.second->d_res);
                    result_resident -= (long long)result_bytes(kv.second->res.nao, kv.second->desc.opts.max_iter);
                    kv.second->d_res = nullptr;
                }
        }
        for (int k = 0; k < kMaxCohorts; ++k) {
            if (spool[k]) hipStreamDestroy(spool[k]);
            if (epool[k]) hipEventDestroy(epool[k]);
            if (hstate[k]) hipHostFree(hstate[k]);
            if (stage[k].p) hipHostFree(stage[k].p);
        }
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

int nddo_dbg_server_stats(nddo_engine* eh, double* seconds4, long long* iterations, int reset) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e) return NDDO_BAD_ARG;
    for (int i = 0; i < 4; ++i) {
        if (seconds4) seconds4[i] = e->server_s[i];
        if (reset) e->server_s[i] = 0.0;
    }
    if (iterations) *iterations = e->server_turns;
    if (reset) e->server_turns = 0;
    return NDDO_OK;
}

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
        j->footprint = nddo::job_footprint_bytes(d->natoms, n, d->opts.diis_num_error_vect, d->opts.max_iter) +
                       nddo::result_bytes(n, d->opts.max_iter);
        for (int a = 0; a < d->natoms; ++a) j->hof_offset += nddo::atom_hof_offset(d->Z[a], d->theory);
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
    if (j->res.status < 0 && !j->err.empty()) nddo::set_error(j->err);
    return j->res.status;
}

int nddo_get_result(nddo_engine* eh, nddo_job_id id, nddo_result* res) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || !res) return NDDO_BAD_ARG;
    ServedJob* j = nullptr;
    {
        std::lock_guard<std::mutex> lk(e->smu);
        auto it = e->served.find(id);
        if (it == e->served.end() || it->second->released) { nddo::set_error("unknown job id"); return NDDO_BAD_ARG; }
        j = it->second;
        if (!j->done) return NDDO_PENDING;
        if (j->res.status < 0 && !j->err.empty()) nddo::set_error(j->err);
    }
    std::lock_guard<std::mutex> jl(j->mu);
    hipSetDevice(e->device);
    if (nddo::materialize(e, j) != NDDO_OK) {
        nddo::set_error("result copy failed: " + nddo::get_error());
        return NDDO_HIP_ERROR;
    }
    double* eps = res->eps;
    double* C = res->C;
    double* P = res->P;
    double* tr = res->trajectory;
    double* q = res->mulliken;
    double* dip = res->dipole;
    double* frc = res->forces;
    *res = j->res;
    res->eps = eps;
    res->C = C;
    res->P = P;
    res->trajectory = tr;
    res->mulliken = q;
    res->dipole = dip;
    res->forces = frc;
    const size_t n = (size_t)std::max(j->res.nao, 0);
    if (j->on_host) {
        if (eps) std::memcpy(eps, j->eps.data(), n * sizeof(double));
        if (C) std::memcpy(C, j->C.data(), n * n * sizeof(double));
        if (P) std::memcpy(P, j->P.data(), n * n * sizeof(double));
        if (tr) std::memcpy(tr, j->traj.data(), j->traj.size() * sizeof(double));
        if (q && j->P.size() == n * n && n > 0)
            nddo::mulliken_charges((int)j->Z.size(), j->Z.data(), j->P.data(), (long long)n + 1, q);
    }
    return j->res.status;
}

int nddo_release(nddo_engine* eh, nddo_job_id id) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e) return NDDO_BAD_ARG;
    ServedJob* j = nullptr;
    {
        std::lock_guard<std::mutex> lk(e->smu);
        auto it = e->served.find(id);
        if (it == e->served.end() || it->second->released) { nddo::set_error("unknown job id"); return NDDO_BAD_ARG; }
        j = it->second;
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
    }
    {
        std::lock_guard<std::mutex> jl(j->mu);
        if (j->d_res) {   // on the API stream, after the retire copies (never on the cohort stream)
            hipSetDevice(e->device);
            std::lock_guard<std::recursive_mutex> cap(e->capture_mu);
            if (j->res_ev && j->res_ev->ev) hipStreamWaitEvent(e->api_stream, j->res_ev->ev, 0);
            hipFreeAsync(j->d_res, e->api_stream);
            e->result_resident -= (long long)nddo::result_bytes(j->res.nao, j->desc.opts.max_iter);
            j->d_res = nullptr;
        }
    }
    delete j;
    return NDDO_OK;
}

}  // extern "C"
