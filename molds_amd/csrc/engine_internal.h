// Engine and batch internals shared by the engine and the serving loop.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "launch.h"
#include "nddo.h"
#include "nddo_device.h"

namespace nddo {

constexpr int kNumElem = 5;   // H, C, N, O, F

inline size_t align_up(size_t b) { return (b + 255) & ~(size_t)255; }

void set_error(const std::string& s);
std::string get_error();   // the calling thread's last message
int elem_slot(int Z);
int validate_job(const nddo_job_desc& d, int* nao, int* nocc);
size_t eig_lds_bytes(int n, bool a_in_lds);   // LDS of the one-workgroup eigensolver
// Device bytes a job keeps resident (admission control; MolDS's MallocerFreer heap
// cap, src/base/MallocerFreer.cpp:62-73, plays the same role on the host).
size_t job_footprint_bytes(int natoms, int nao, int diis_k, int max_iter);
// Device bytes a retired served job keeps for its arrays until nddo_get_result.
size_t result_bytes(int n, int max_iter);

struct Engine;

// Device bytes an allocation can get now: hipMemGetInfo's free bytes plus what
// the default stream-ordered pool holds reserved but unused.
size_t device_available(int device);
// Returns the default pool's unused reserve to the device.
void trim_pool(int device);
// Whether n AOs take the one-workgroup eigensolver with the matrix in LDS.
bool lds_eig_class(int n, size_t lds_max);
// Makes the engine's batch workspace at least `bytes` (caller holds run_mu);
// NDDO_ADMISSION_DENIED when the device cannot provide it.
int ensure_workspace(Engine* e, size_t bytes);
void free_workspace(Engine* e);

void fill_scalars(const JobDev& jb, const JobState& s, int theory, double hof_offset, nddo_result* res);
void order_and_phase(int n, const double* VT, const int* perm, double* C);

// A set of jobs sharing one launch per SCF stage.
struct Batch {
    std::vector<AtomDev> atoms;
    std::vector<AtomRowDev> arows;   // per atom: the pair row it starts (pairs are built on the device)
    long long npairs_all = 0;
    std::vector<JobDev> jobs;
    std::vector<TileDev> tiles;
    std::vector<int> list_lds, list_glb, list_xl, src;
    // d_plist: pair indices, the light (H-H, heavy-H) pairs first, then the heavy-heavy ones
    int n_light = 0, n_heavy = 0;
    std::vector<long long> xl_off;   // per list_xl entry: offset in d_xlws
    int maxn_lds = 0, maxn_glb = 0, maxn_xl = 0, max_iter_all = 0;
    // The LDS class in spans of similar size (list_lds is sorted by size,
    // largest first): each span's launches size their LDS and threads for its
    // own largest molecule, so small molecules share a CU several at a time.
    struct Span {
        int start, count, maxn;
    };
    std::vector<Span> lds_spans;
    int xl_grid = 1, xl_chunks = 1;
    // large molecules: device-wide Householder + divide and conquer (kernels_xlh.hip),
    // or block Jacobi (kernels_xl.hip); classify() decides from use_xl_householder
    bool use_xl_householder = true, xl_householder = true;
    long long xlws_size = 0;
    long long nslots = 0, nints = 0, arena_size = 0, nvec = 0, ntraj = 0;

    AtomDev* d_atoms = nullptr;
    PairDev* d_pairs = nullptr;
    AtomRowDev* d_arows = nullptr;
    JobDev* d_jobs = nullptr;
    JobState* d_state = nullptr;
    TileDev* d_tiles = nullptr;
    double* d_ints = nullptr;
    float* d_ints32 = nullptr;   // FP32 copy of the integrals (only when a job is mixed-precision)
    double* d_arena = nullptr;
    double* d_part = nullptr;
    double* d_ecore = nullptr;
    double* d_eps = nullptr;
    int* d_perm = nullptr;
    double* d_tri = nullptr;     // eigensolver workspace (EIG_WS_PER_AO / EIG_WS_PER_JOB)
    double* d_rmspart = nullptr;
    double* d_tpart = nullptr;   // per density tile: DIIS dot products and max |e| (XL_REC doubles)
    double* d_traj = nullptr;
    int* d_remaining = nullptr;
    int* d_list_lds = nullptr;
    int* d_list_glb = nullptr;
    int* d_plist = nullptr;
    FockPairDev* d_fpairs = nullptr;   // the pairs in d_plist order, as the Fock kernels read them
    // large molecules (kernels_xl.hip)
    int* d_list_xl = nullptr;
    long long* d_xl_off = nullptr;
    double* d_xlws = nullptr;
    double* d_xlpart = nullptr;
    double* d_xlcoef = nullptr;  // per job: DIIS coefficients and what to apply (XL_COEF doubles)
    unsigned* d_bar = nullptr;
    double* d_sdbg = nullptr;
    // jobs whose caller asked for the dipole moment (set before allocate)
    std::vector<int> list_dip;
    std::vector<long long> dip_off;   // per list_dip entry: offset of its 4 n row partials
    long long dip_size = 0;
    int maxn_dip = 0;
    int* d_list_dip = nullptr;
    long long* d_dip_off = nullptr;
    double* d_dippart = nullptr;
    bool any_forces = false;
    double* d_grad = nullptr;     // 4 doubles per pair slot: dE_AB/dR, E_AB
    double* d_forces = nullptr;   // 3 doubles per atom
    hipStream_t alloc_stream = nullptr;   // set: stream-ordered allocation on this stream
    // set: every device buffer is carved out of [region, region + region_cap)
    // (the engine's batch workspace, or the fixed scratch array of
    // nddo_scf_once_noalloc); nothing is allocated or freed per batch
    char* region = nullptr;
    size_t region_cap = 0;
    bool carved = false;
    // set: the two one-workgroup eigensolver classes run concurrently, the LDS
    // class on this stream (forked from and joined to the iteration's stream),
    // and so do the light-pair and heavy-pair Fock kernels (fork_ev2, join_ev2)
    hipStream_t fork_stream = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr, fork_ev2 = nullptr, join_ev2 = nullptr;
    // set with split_classes (synchronous batches whose jobs come with the LDS
    // eigensolver class first): the whole iteration of those jobs runs on
    // fork_stream beside the others' on the iteration's stream, joined once;
    // the others' light-pair Fock kernels then fork to aux_stream
    bool split_classes = false;
    hipStream_t aux_stream = nullptr;
    hipEvent_t aux_fork = nullptr, aux_join = nullptr;
    int split_job = 0, split_atom = 0, split_tile = 0;   // first job / atom / tile of the second part
    // per job, the pair-list position of its first pair of each class (H-H,
    // heavy-H, heavy-heavy); nj + 1 entries
    std::vector<long long> cls_pos[3];
    FockPairs fock_pairs(int j0, int j1) const;
    bool split_ready() const;
    void launch_iteration_split(Engine* e, hipStream_t st);

    ~Batch();
    int dalloc(void** p, size_t bytes);
    void release();
    // Every device buffer allocate() makes and its size (all: the optional ones too).
    void plan(bool want_sdbg, bool all, std::vector<std::pair<void**, size_t>>* req);
    // Bytes allocate() takes, each buffer rounded up to 256 bytes.
    size_t device_bytes(bool want_sdbg);
    int build(const nddo_job_desc* descs, int nd, const std::vector<int>& which, const int* unused);
    // xl_min_nao: smallest molecule of the device-wide eigensolver class
    // route_few_large: molecules above 512 AOs take the device-wide solver
    // when the batch holds few of them (nddo_scf_batch / nddo_scf_once)
    int classify(size_t lds_max, int num_cus, int xl_min_nao, bool route_few_large = false);
    void launch_xl_eigen_all(Engine* e, hipStream_t st);
    // The LDS-class eigensolver span by span: part 0 / 1 / 2 of launch_heig_part,
    // or -1 for all three parts of each span.
    void launch_lds_eigen(int part, hipStream_t st);
    int allocate(bool want_sdbg);
    size_t upload_bytes() const;
    int upload(hipStream_t st, const nddo_job_desc* descs, char* stage = nullptr);
    void launch_setup_stage(Engine* e, hipStream_t st);
    void launch_iteration(Engine* e, hipStream_t st);
    // One SCF iteration: the first eagerly, the next ones as replays of a
    // hipGraph captured from the second (unless graphs are off or the batch
    // has device-wide eigensolves, whose launches are chained across streams).
    void iterate(Engine* e, hipStream_t st);
    hipGraphExec_t gexec = nullptr;
    int iterations_launched = 0;
    bool graph_off = false;
    bool eager_only = false;   // never capture (nddo_scf_once_noalloc: instantiation may allocate)
    bool served = false;       // a served cohort: no capture while a synchronous call runs
    int run_scf(Engine* e, hipStream_t st);
    int fetch_all(hipStream_t st, const nddo_job_desc* descs, nddo_result* const* res, const double* hof_offsets);
    void set_dipole_jobs(const std::vector<int>& jj);
    // Jobs (batch positions) whose caller wants forces: sets JobDev::want_f
    // (before upload()).
    void set_force_jobs(const std::vector<int>& jj);
    // After run_scf: the dipole partials and the forces the jobs asked for.
    void launch_properties(Engine* e, hipStream_t st);
};

// Total, electronic and core dipole moments of a job from k_dipole's row partials.
void dipole_moments(const nddo_job_desc& d, const double* rows, int n, double* out);

struct ServedJob;

// Mulliken charges from the diagonal of a final density, pdiag[k * stride] = P_kk.
void mulliken_charges(int natoms, const int* Z, const double* pdiag, long long stride, double* q);

struct Engine {
    int device = 0;
    nddo_engine_opts opts{0, 0, 0};
    hipStream_t stream = nullptr;
    // the API threads' own stream for served-result copies and frees: work
    // issued from a caller's thread never goes to a cohort stream, which the
    // scheduler thread may be capturing into a graph at that moment
    hipStream_t api_stream = nullptr;
    // batch runs: the LDS-class eigensolves and the light-pair Fock kernel run
    // on eig_stream beside the rest of the iteration on the batch stream (fork
    // and join events of each)
    hipStream_t eig_stream = nullptr;
    hipEvent_t eig_fork = nullptr, eig_join = nullptr, fock_fork = nullptr, fock_join = nullptr;
    // the third stream of a split iteration (the second part's light-pair Fock)
    hipStream_t aux_stream = nullptr;
    hipEvent_t aux_fork = nullptr, aux_join = nullptr;
    ElemDev* d_elems = nullptr;
    size_t lds_max = 65536;
    int num_cus = 0;
    int xl_min_nao = 1201;   // 513 with NDDO_ENGINE_DEVICE_WIDE_EIGEN
    bool use_graphs = true;  // false with NDDO_ENGINE_NO_GRAPHS
    bool xl_householder = true;   // false with NDDO_ENGINE_BLOCK_JACOBI
    std::atomic<long long> graphs_captured{0}, graph_replays{0};
    size_t hbm_budget = 0;
    int* h_remaining = nullptr;     // pinned, fine-grained
    int* d_h_remaining = nullptr;   // its device address
    std::mutex run_mu;   // one synchronous batch at a time
    // Held while a stream captures an iteration graph, and by every HIP call
    // that may run beside another thread's capture: the served scheduler's
    // launches, queries, copies and frees, and the caller threads' served-result
    // copies and frees.  The synchronous entry points instead keep the
    // scheduler from opening a capture while they run (sync_busy, SyncRun).
    // Recursive: the scheduler holds it around a cohort iteration whose
    // launch may itself begin a capture.
    std::recursive_mutex capture_mu;
    std::atomic<int> sync_busy{0};
    bool profile = false;
    double stage_ms[NDDO_STAGE_COUNT] = {};
    long long stage_launches[NDDO_STAGE_COUNT] = {};
    hipEvent_t ev[16] = {};
    // completion of the last large-molecule eigensolve launched on any stream:
    // those persistent launches run one at a time, so their workgroups are
    // always resident together
    hipEvent_t xl_chain = nullptr;
    double last_eig_prof[4] = {};
    double last_dc_phase[9] = {};
    // nddo_scf_batch host phases (s): build + classify, allocate + upload, SCF, results, release
    double host_s[5] = {};
    // nddo_scf_batch's device buffers: one allocation, grown when a batch needs
    // more and reused by every later batch (guarded by run_mu)
    char* d_ws = nullptr;
    size_t ws_cap = 0;
    long long ws_grows = 0;
    // device bytes of retired served results not yet copied to the host
    std::atomic<long long> result_resident{0};
    // device address of this device's copy of the file-scope scratch array of
    // nddo_scf_once_noalloc (engine.cpp), from hipGetSymbolAddress at creation
    char* d_scratch = nullptr;

    // serving state (server.cpp)
    std::mutex smu;
    std::condition_variable scv;
    std::thread server;
    bool stopping = false;
    long long next_id = 1;
    std::map<long long, ServedJob*> served;
    std::deque<ServedJob*> queue;
    // scheduler thread wall time: admission + cohort setup, launches, waiting, retirement
    double server_s[4] = {};
    long long server_turns = 0;   // cohort iterations completed

    void start_server();
    void stop_server();
};

// A synchronous entry point (nddo_scf_batch, nddo_scf_once, the stage hooks):
// one at a time (run_mu); while it runs the served scheduler opens no graph
// capture, so none of its HIP calls on the caller's thread falls inside
// another thread's capture window, and a capture already open ends before it
// starts.
struct SyncRun {
    Engine* e;
    std::lock_guard<std::mutex> run;
    explicit SyncRun(Engine* eng) : e(eng), run(eng->run_mu) {
        std::lock_guard<std::recursive_mutex> cap(e->capture_mu);
        e->sync_busy.fetch_add(1);
    }
    ~SyncRun() { e->sync_busy.fetch_sub(1); }
    SyncRun(const SyncRun&) = delete;
    SyncRun& operator=(const SyncRun&) = delete;
};

}  // namespace nddo
