// capi.hip -- host orchestration behind the C ABI (include/risgraph_amd.h).
//
// One HIP stream per graph.  Store batches run as two kernels (expand+link,
// apply); each labels instance then runs its engine: a persistent
// single-workgroup launch while the work of a pass is small, grid-wide
// passes otherwise, with the control block copied back after each launch to
// pick the next one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cerrno>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <exception>
#include <fcntl.h>
#include <unistd.h>
#include <functional>
#include <limits>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/risgraph_amd.h"
#include "host_batch.h"
#include "rg_internal.h"

using namespace rg;

namespace {

thread_local std::string g_err;

struct RgError {
    int code;
    std::string msg;
};

void hipchk(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw RgError{RG_EDEVICE, std::string(what) + ": " + hipGetErrorString(e)};
}
#define HC(x) hipchk((x), #x)

template <class T>
T *dalloc(uint64_t n)
{
    void *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(&p, n * sizeof(T));
    if (e != hipSuccess)
        throw RgError{RG_ENOMEM, "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed"};
    return (T *)p;
}

template <class T>
T *halloc(uint64_t n)
{
    void *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipHostMalloc(&p, n * sizeof(T), hipHostMallocDefault);
    if (e != hipSuccess) throw RgError{RG_ENOMEM, "hipHostMalloc failed"};
    return (T *)p;
}

template <class T>
void dfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

template <class T>
void hfree(T *&p)
{
    if (p) (void)hipHostFree(p);
    p = nullptr;
}

int guard(const std::function<void()> &f)
{
    try {
        f();
        return RG_OK;
    } catch (const RgError &e) {
        g_err = e.msg;
        return e.code;
    } catch (const std::exception &e) {
        g_err = e.what();
        return RG_EDEVICE;
    }
}

constexpr uint64_t MAX_V = 0x7FFFFFFFull;

}  // namespace

struct rg_graph {
    uint64_t V = 0, hint = 0;
    int symmetric = 0, dual = 0, weighted = 0;
    hipStream_t st = nullptr;
    Slots slots{};
    List lists[2]{};
    Pool *pool = nullptr;
    uint64_t slot_cap = 0, hash_cap = 0;
    uint32_t fcap = 0;                // entries per size-class free list of the slab pools
    uint32_t lstride = 4;             // list-op slots per edge op of the current batch
    EdgeOps ops{};
    ListOps lops{};
    EdgeOps uops{};                   // a stream batch's unsafe ops (the engine's input after the split)
    uint32_t *ucount = nullptr;       // device: their count
    uint32_t op_cap = 0;
    uint32_t *err = nullptr;          // device: [0] bad ids
    uint64_t *free_off = nullptr, *pend_off = nullptr;   // slab free lists
    uint32_t *cnt_tmp[2]{};           // bulk reservation counters
    uint32_t *h_status = nullptr;     // pinned: err, oom, then the static path's bad-id and bad-weight flags
    unsigned char *scratch = nullptr; // this device's copy of the static scratch array (no-alloc entry point)
    // pinned staging for host-side batches
    uint32_t *h_src = nullptr, *h_dst = nullptr, *h_w = nullptr;
    OpResult *h_res = nullptr;
    uint8_t *h_kind = nullptr, *h_flag = nullptr;
    uint8_t *d_flag = nullptr;
    uint32_t h_cap = 0;
    hipEvent_t ev[4]{};
    double last_ms = 0;
    // the safe/unsafe split's per-op flags reach the host (pinned) when
    // ev_safe completes, while the repair runs
    hipEvent_t ev_safe = nullptr;
    uint64_t last_unsafe = 0;         // unsafe ops of the last split batch
    uint64_t version = 0;             // stream batches applied (the history's version clock)
    // captured pipelines of stream batches (window -> store -> split), one
    // per batch-size bucket, tree set and window source (rg_graph_capture);
    // keyed by the argument values the captured launches hold
    struct Captured {
        uint32_t bucket = 0;
        SplitParams P{};
        const uint64_t *src = nullptr, *dst = nullptr;
        const uint32_t *w = nullptr;
        uint64_t len = 0;
        uint8_t kind_ins = 0, kind_del = 0;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<Captured> captured;
    int capture = -1;                 // -1: RG_GRAPH decides (default on)
    // engine knobs per batch-size bucket, picked by rg_autotune: the team
    // cap (x 1/4) and the team weights of the four algorithms
    struct TuneEntry {
        uint32_t cap_x4 = 4;
        uint32_t weight[4] = {0, 0, 0, 0};   // 0: the algorithm default
    };
    TuneEntry tune_tab[RG_TUNE_BUCKETS];
    const TuneEntry *run_tune = nullptr;   // the current run's entry
    uint64_t replays = 0;             // batches run through a captured pipeline
    BatchDesc *h_desc = nullptr;      // pinned
    BatchDesc *d_desc = nullptr;
    int wal_fd = -1;                  // write-ahead log of transaction batches (rg_wal_open)
    int wal_sync = 1;                 // fdatasync after every logged batch (group commit)
    std::recursive_mutex mu;
    // kernel timing (rg_profile_enable)
    struct ProfRec {
        int kind;
        hipEvent_t a, b;
        uint64_t units;
        int phase;   // grid-wide passes: the engine phase the step ran (else -1)
    };
    int prof = 0;   // bit k: time kernels of kind k
    std::vector<hipEvent_t> ev_free;
    std::vector<ProfRec> prof_pending;
    // kinds 0-2, then 3 + phase: grid-wide passes of each engine phase
    // 0 store apply, 1 team, 2 grid, 3 + p grid passes of phase p, then the
    // one-workgroup store kernel of small batches
    static constexpr int PKINDS = 4 + NPHASE;
    static constexpr int KIND_STORE_SMALL = 3 + NPHASE;
    uint64_t p_launch[PKINDS]{}, p_units[PKINDS]{};
    double p_ms[PKINDS]{};
};

struct rg_labels {
    rg_graph *g = nullptr;
    int algo = 0;
    uint32_t root = 0;
    int trace = 0;
    VRec *vr = nullptr;            // label word, mark, frontier stamp per vertex
    uint32_t *init = nullptr;
    uint32_t *mstamp = nullptr;
    uint32_t *q[3]{};
    ChunkItem *cq = nullptr;
    uint32_t *bnd = nullptr;             // SSWP deletion bounds (device)
    uint32_t cq_cap = 0;
    uint32_t *mlist = nullptr;
    Ctl *ctl = nullptr;
    Ctl *hc = nullptr;     // pinned mirror
    unsigned long long *ticks = nullptr;   // per-phase block-mode time (device)
    uint64_t dense_passes = 0;             // pull passes of its builds so far
    uint32_t *bar = nullptr;               // team barrier sub-counters (device)
    hipStream_t st = nullptr;      // this tree's grid-wide passes (concurrent with other trees)
    EngineArgs *dargs = nullptr;   // device copy of the engine arguments (team launches)
    EngineArgs *hargs = nullptr;   // pinned staging of the same
    Tuning tune{};
    uint32_t team_w = 0;           // workgroups per tree in the team launch (0: by batch size)
    uint32_t block_max = RG_AUTO;  // largest pass (vertices, or ops/16) run by the team
    uint32_t chunk_max = RG_AUTO;  // most chunk items of a pass run by the team
    uint32_t run_w = 8;            // team width of the current run
    uint32_t team_weight = 0;      // share of the team's XCD slots (0: by algorithm)
    uint32_t *init_active = nullptr;
    // device history (rg_history_*): ring of value records per batch version
    struct Hist {
        bool on = false;
        HistArgs a{};
        uint64_t top = 0;                 // next logical record index
        uint64_t base_version = 0;        // base holds the values at this version
        uint32_t *dcount = nullptr;       // device: records appended by the last append
        uint32_t *hcount = nullptr;       // pinned copy
        std::deque<std::pair<uint64_t, uint64_t>> batches;   // (version, first record) in ring order
    } hist;
};

struct rg_stream {
    uint64_t n = 0;
    uint64_t *src = nullptr, *dst = nullptr;
    uint32_t *w = nullptr;
    hipStream_t st = nullptr;
};

namespace {

void alloc_list(List &l, uint64_t V, hipStream_t st)
{
    l.h = dalloc<ListHdr>(V);
    launch_list_init(l, (uint32_t)V, st);
    HC(hipGetLastError());
}

void free_list(List &l) { dfree(l.h); }

// live degrees of every list -> host array (gathered on the device, one copy)
void copy_degrees(const List &l, uint32_t *out, uint64_t V, hipStream_t st)
{
    uint32_t *tmp = dalloc<uint32_t>(V);
    launch_list_degrees(l, (uint32_t)V, tmp, st);
    HC(hipGetLastError());
    HC(hipMemcpyAsync(out, tmp, V * 4, hipMemcpyDeviceToHost, st));
    HC(hipStreamSynchronize(st));
    dfree(tmp);
}

void free_ops(rg_graph *g)
{
    dfree(g->ops.src); dfree(g->ops.dst); dfree(g->ops.w); dfree(g->ops.kind);
    dfree(g->ops.res);
    dfree(g->lops.op); dfree(g->lops.segbuf); dfree(g->lops.heavy); dfree(g->lops.owner);
    dfree(g->d_flag);
    dfree(g->uops.src); dfree(g->uops.dst); dfree(g->uops.w); dfree(g->uops.kind); dfree(g->uops.res);
    dfree(g->ucount);
    g->op_cap = 0;
}

// captured pipelines hold the op buffers' and pinned flags' addresses
void captured_clear(rg_graph *g)
{
    for (auto &c : g->captured)
        if (c.exec) (void)hipGraphExecDestroy(c.exec);
    g->captured.clear();
}

void ensure_ops(rg_graph *g, uint64_t n)
{
    if (n > 0xFFFFFFFFull / 4) throw RgError{RG_EINVAL, "batch too large"};
    if (n <= g->op_cap) return;
    captured_clear(g);
    uint32_t cap = std::max<uint32_t>((uint32_t)n, 1024);
    HC(hipStreamSynchronize(g->st));
    free_ops(g);
    g->ops.src = dalloc<uint32_t>(cap);
    g->ops.dst = dalloc<uint32_t>(cap);
    g->ops.w = dalloc<uint32_t>(cap);
    g->ops.kind = dalloc<uint8_t>(cap);
    g->ops.res = dalloc<OpResult>(cap);
    uint64_t lc = 4ull * cap;
    g->lops.op = dalloc<LOp>(lc);
    g->lops.segbuf = dalloc<uint32_t>(lc);
    g->lops.heavy = dalloc<uint32_t>(lc);
    g->lops.owner = dalloc<LOp>(lc);
    g->d_flag = dalloc<uint8_t>(cap);
    g->uops.src = dalloc<uint32_t>(cap);
    g->uops.dst = dalloc<uint32_t>(cap);
    g->uops.w = dalloc<uint32_t>(cap);
    g->uops.kind = dalloc<uint8_t>(cap);
    g->uops.res = dalloc<OpResult>(cap);
    g->ucount = dalloc<uint32_t>(1);
    g->op_cap = cap;
}

void ensure_host(rg_graph *g, uint64_t n)
{
    if (n <= g->h_cap) return;
    captured_clear(g);
    uint32_t cap = std::max<uint32_t>((uint32_t)n, 1024);
    hfree(g->h_src); hfree(g->h_dst); hfree(g->h_w); hfree(g->h_res); hfree(g->h_kind); hfree(g->h_flag);
    g->h_src = halloc<uint32_t>(cap);
    g->h_dst = halloc<uint32_t>(cap);
    g->h_w = halloc<uint32_t>(cap);
    g->h_res = halloc<OpResult>(cap);
    g->h_kind = halloc<uint8_t>(cap);
    g->h_flag = halloc<uint8_t>(cap);
    g->h_cap = cap;
}

StoreArgs store_args(rg_graph *g)
{
    StoreArgs a{};
    a.V = (uint32_t)g->V;
    a.symmetric = g->symmetric;
    a.weighted = g->weighted;
    a.slots = g->slots;
    a.lists[0] = g->lists[0];
    a.lists[1] = g->lists[1];
    a.pool = g->pool;
    a.ops = g->ops;
    a.lops = g->lops;
    a.err = g->err;
    a.hstatus = g->h_status;   // the batch's last kernel publishes err / oom here
    a.free_off = g->free_off;
    a.pend_off = g->pend_off;
    a.cnt_tmp[0] = g->cnt_tmp[0];
    a.cnt_tmp[1] = g->cnt_tmp[1];
    a.fcap = g->fcap;
    a.lstride = g->lstride;
    return a;
}

// upload host edges as edge ops [0, n) with the given kind; validates ids
void upload_edges(rg_graph *g, const rg_edge *e, uint64_t n, uint8_t kind)
{
    ensure_ops(g, n);
    ensure_host(g, n);
    for (uint64_t i = 0; i < n; i++) {
        if (e[i].src >= g->V || e[i].dst >= g->V) throw RgError{RG_EINVAL_VERTEX, "VertexId error."};
        if (g->weighted && e[i].data > 0xFFFFFFFFull) throw RgError{RG_EINVAL, "edge weight exceeds 32 bits"};
        g->h_src[i] = (uint32_t)e[i].src;
        g->h_dst[i] = (uint32_t)e[i].dst;
        g->h_w[i] = g->weighted ? (uint32_t)e[i].data : 0;
        g->h_kind[i] = kind;
    }
    HC(hipMemcpyAsync(g->ops.src, g->h_src, n * 4, hipMemcpyHostToDevice, g->st));
    HC(hipMemcpyAsync(g->ops.dst, g->h_dst, n * 4, hipMemcpyHostToDevice, g->st));
    HC(hipMemcpyAsync(g->ops.w, g->h_w, n * 4, hipMemcpyHostToDevice, g->st));
    HC(hipMemcpyAsync(g->ops.kind, g->h_kind, n, hipMemcpyHostToDevice, g->st));
    g->ops.n = (uint32_t)n;
    g->lops.n = (uint32_t)(4 * n);
    g->lstride = 4;   // callers may rewrite the kinds op by op
}

void store_status_check(rg_graph *g)
{
    if (g->h_status[0]) {
        uint32_t e = g->h_status[0];
        g->h_status[0] = 0;
        HC(hipMemsetAsync(g->err, 0, 4, g->st));
        if (e & 1) throw RgError{RG_EINVAL_VERTEX, "VertexId error."};
        if (e & 2) throw RgError{RG_EDEVICE, "an edge op needed more list-op slots than its batch's stride"};
        throw RgError{RG_EDEVICE, "store op chain corrupt"};
    }
    if (g->h_status[1]) throw RgError{RG_ENOMEM, "adjacency slot or hash pool exhausted"};
}

// ---------------------------------------------------------------------
// Scratch of the entry point that allocates nothing
// (rg_edge_update_batch_static): the store batch's edge-op and list-op
// arrays for up to STATIC_MAX_OPS edge ops live in one file-scope device
// array instead of buffers the graph allocates and grows.  Each device holds
// its own copy; host code reaches it only through the address
// hipGetSymbolAddress returns on that device (taken once per device, with
// the device current, when a graph is created there).  Calls on a device
// share its copy, and each call leaves its contents behind; a process-wide
// lock keeps the calls of this process from overlapping.
// ---------------------------------------------------------------------
constexpr uint32_t STATIC_MAX_OPS = RG_STATIC_MAX_OPS;
constexpr size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t SC_SRC = 0;
constexpr size_t SC_DST = SC_SRC + al256(4ull * STATIC_MAX_OPS);
constexpr size_t SC_W = SC_DST + al256(4ull * STATIC_MAX_OPS);
constexpr size_t SC_KIND = SC_W + al256(4ull * STATIC_MAX_OPS);
constexpr size_t SC_RES = SC_KIND + al256(1ull * STATIC_MAX_OPS);
constexpr size_t SC_LOP = SC_RES + al256(sizeof(OpResult) * STATIC_MAX_OPS);
constexpr size_t SC_SEG = SC_LOP + al256(sizeof(LOp) * 4ull * STATIC_MAX_OPS);
constexpr size_t SC_OWNER = SC_SEG + al256(16ull * STATIC_MAX_OPS);
constexpr size_t SC_HEAVY = SC_OWNER + al256(sizeof(LOp) * 4ull * STATIC_MAX_OPS);
constexpr size_t SCRATCH_BYTES = SC_HEAVY + al256(16ull * STATIC_MAX_OPS);

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

std::mutex g_static_mu;            // one static-path call at a time in this process
unsigned char *g_scratch_dev[64];  // per-device address of `scratch`, set once
std::mutex g_scratch_mu;

// this device's address of the scratch array (called with the device current)
unsigned char *scratch_address()
{
    int dev = 0;
    HC(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) throw RgError{RG_EDEVICE, "device index out of range"};
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw RgError{RG_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
        g_scratch_dev[dev] = (unsigned char *)p;
    }
    return g_scratch_dev[dev];
}

// rg_edge (device) -> edge ops in the scratch; ids and weights validated:
// bad[0] = 1 for an id >= V, bad[1] = 1 for a weight over 32 bits
__global__ void k_edges_to_ops(const rg_edge *e, uint32_t n, uint8_t kind, uint64_t V, int weighted, EdgeOps ops,
                               uint32_t *bad)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const rg_edge x = e[i];
    if (x.src >= V || x.dst >= V) bad[0] = 1u;
    if (weighted && x.data > 0xFFFFFFFFull) bad[1] = 1u;
    ops.src[i] = (uint32_t)x.src;
    ops.dst[i] = (uint32_t)x.dst;
    ops.w[i] = weighted ? (uint32_t)x.data : 0u;
    ops.kind[i] = kind;
}

__global__ void k_old_out(const OpResult *res, uint32_t n, uint64_t *out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = res[i].old;
}

hipEvent_t prof_event(rg_graph *g)
{
    if (!g->ev_free.empty()) {
        hipEvent_t e = g->ev_free.back();
        g->ev_free.pop_back();
        return e;
    }
    hipEvent_t e;
    HC(hipEventCreate(&e));
    return e;
}

// bracket the next launch on stream s with events when profiling
void prof_begin(rg_graph *g, int kind, uint64_t units, hipStream_t s, hipEvent_t *a, hipEvent_t *b,
                int phase = -1)
{
    *a = *b = nullptr;
    const int bit = kind == rg_graph::KIND_STORE_SMALL ? 0 : kind;
    if (!((g->prof >> bit) & 1)) return;
    *a = prof_event(g);
    *b = prof_event(g);
    g->prof_pending.push_back({kind, *a, *b, units, phase});
    (void)s;
}

// fold completed event pairs into the counters (call after the streams synced)
void prof_collect(rg_graph *g)
{
    for (auto &r : g->prof_pending) {
        float ms = 0;
        HC(hipEventSynchronize(r.b));
        HC(hipEventElapsedTime(&ms, r.a, r.b));
        g->p_launch[r.kind]++;
        g->p_ms[r.kind] += ms;
        g->p_units[r.kind] += r.units;
        if (r.phase >= 0 && r.phase < NPHASE) {
            g->p_launch[3 + r.phase]++;
            g->p_ms[3 + r.phase] += ms;
            g->p_units[3 + r.phase] += r.units;
        }
        g->ev_free.push_back(r.a);
        g->ev_free.push_back(r.b);
    }
    g->prof_pending.clear();
}

void run_store(rg_graph *g)
{
    StoreArgs a = store_args(g);
    hipEvent_t ea, eb;
    prof_begin(g, g->ops.n <= STORE_SMALL_OPS ? rg_graph::KIND_STORE_SMALL : 0, g->ops.n, g->st, &ea, &eb);
    if (!launch_store_batch(a, g->st, ea, eb))
        launch_pool_recycle(a, g->st);   // also publishes err/oom to g->h_status
    HC(hipGetLastError());
}

EngineArgs engine_args(rg_labels *L, bool split = false)
{
    rg_graph *g = L->g;
    EngineArgs a{};
    a.V = (uint32_t)g->V;
    a.algo = L->algo;
    a.trace = L->trace;
    a.slots = g->slots;
    a.lists[0] = g->lists[0];
    a.lists[1] = g->symmetric ? g->lists[0] : g->lists[1];
    if (g->symmetric) {
        a.undirected = 0;
        a.npush = 1; a.push_l[0] = 0;
        a.npull = 1; a.pull_l[0] = 0;
    } else if (L->algo == WCC) {
        // WCC over the undirected view of a directed store
        a.undirected = 1;
        a.npush = 2; a.push_l[0] = 0; a.push_l[1] = 1;
        a.npull = 2; a.pull_l[0] = 1; a.pull_l[1] = 0;
    } else {
        a.undirected = 0;
        a.npush = 1; a.push_l[0] = 0;
        a.npull = 1; a.pull_l[0] = 1;
    }
    a.vr = L->vr;
    a.init = L->init;
    a.init_active = L->init_active;
    a.root = L->root;

    a.mstamp = L->mstamp;
    a.q[0] = L->q[0]; a.q[1] = L->q[1]; a.q[2] = L->q[2];
    a.cq = L->cq;
    a.bnd = L->bnd;
    a.cq_cap = L->cq_cap;
    a.mlist = L->mlist;
    a.ctl = L->ctl;
    a.hctl = L->hc;
    a.ticks = L->ticks;
    a.bar = L->bar;
    a.ops = split ? g->uops : g->ops;
    a.ops.n = 0;   // the engine takes the batch size from the control block;
                   // keeping it out lets the device copy of the arguments stay valid
    a.tune = L->tune;
    return a;
}

uint64_t host_work(const Ctl &h, uint64_t V)
{
    switch (h.phase) {
    case PH_SEED:
    case PH_DEL_ROOTS: return h.n_ops / 16;   // same units as the device's phase_work
    case PH_DEL_PULL: return h.n_tree + (h.seed ? h.n_ops / 16 : 0);
    case PH_BUILD_INIT: return V;
    case PH_PUSH: return (h.flags & F_DENSE) ? V : h.n_cur;
    default: return h.n_cur;
    }
}

// Team width of a run: 8 workgroups per tree for small batches (cheaper
// barriers), growing with the batch to 64 (4 trees x 64 workgroups of 1024
// threads, one per CU of a 256-CU device; the launcher narrows a team that
// would not be co-resident, which the team barrier needs).
uint32_t team_width(const rg_labels *L, uint32_t n_ops)
{
    if (L->team_w) return L->team_w;
    uint32_t w = 8;
    while (w < 64 && (uint64_t)w * 2048 < n_ops) w *= 2;
    return w;
}

uint32_t team_cap(const rg_labels *L)
{
    if (L->block_max != RG_AUTO) return L->block_max;
    const uint32_t x4 = L->g->run_tune ? L->g->run_tune->cap_x4 : 4u;
    return (uint32_t)std::min<uint64_t>(0xFFFFFFF0ull, 256ull * L->run_w * x4);
}

// Chunk items a team pass keeps: 1024 per workgroup, scaled like team_cap;
// a pass with more hands its chunk stage to grid mode.  (A grid-wide chunk
// stage cannot use the CUs the other trees' team workgroups hold, since a
// team workgroup takes a CU's registers whole, so the hand-over mostly adds
// a host round trip and a second team launch: on R-MAT scale 24, 1024
// instead of 256 items per workgroup made batches of 0.5-2M updates 10-17%
// shorter (256k: 2%), and no cap at all gained at most 4% more.)
uint32_t team_chunk_cap(const rg_labels *L)
{
    if (L->chunk_max != RG_AUTO) return L->chunk_max;
    const uint32_t x4 = L->g->run_tune ? L->g->run_tune->cap_x4 : 4u;
    return (uint32_t)std::min<uint64_t>(0xFFFFFFF0ull, 256ull * L->run_w * x4);
}

// batch-size bucket of the tuning table: up to 32k, 128k, 512k edge ops, more
int tune_bucket(uint64_t n_ops)
{
    return n_ops <= (32u << 10) ? 0 : n_ops <= (128u << 10) ? 1 : n_ops <= (512u << 10) ? 2 : 3;
}

// A tree's share of the team launch.  Measured per-tree repair times of
// all_inc batches at R-MAT scale 24 with equal shares: SSSP 0.76 ms, SSWP
// 0.61, WCC 0.38, BFS 0.29 (the launch waits for the slowest).
uint32_t algo_weight(const rg_graph *g, int algo)
{
    const auto *t = g->run_tune;
    if (t && algo >= 0 && algo < 4 && t->weight[algo]) return t->weight[algo];
    switch (algo) {
    case RG_SSSP: return 3;
    case RG_SSWP: return 2;
    case RG_WCC: return 2;
    default: return 1;
    }
}

uint32_t team_weight(const rg_labels *L) { return L->team_weight ? L->team_weight : algo_weight(L->g, L->algo); }

// The eight XCD slots of a team launch over trees L[0..n): one slot each at
// least, the rest by weight (largest remainder).
void split_slots(rg_labels *const *L, int n, uint32_t *k)
{
    double q[MAXT], tot = 0;
    for (int i = 0; i < n; i++) tot += team_weight(L[i]);
    uint32_t used = 0;
    for (int i = 0; i < n; i++) {
        q[i] = XCD_SLOTS * team_weight(L[i]) / tot;
        k[i] = std::max<uint32_t>(1, (uint32_t)q[i]);
        used += k[i];
    }
    while (used > XCD_SLOTS) {   // only when n trees forced a minimum slot
        int j = -1;
        for (int i = 0; i < n; i++)
            if (k[i] > 1 && (j < 0 || q[i] - k[i] < q[j] - k[j])) j = i;
        k[j]--;
        used--;
    }
    while (used < XCD_SLOTS) {
        int j = 0;
        for (int i = 1; i < n; i++)
            if (q[i] - k[i] > q[j] - k[j]) j = i;
        k[j]++;
        used++;
    }
}

// A tree's next pass runs in the persistent team launch when it is not a
// full rebuild and fits the team's work cap; otherwise it is one grid step.
bool team_ok(const rg_labels *L, const Ctl &h)
{
    return h.phase != PH_BUILD_INIT && !(h.phase == PH_PUSH && (h.flags & F_DENSE)) && h.stage == 0 &&
           host_work(h, L->g->V) <= team_cap(L);
}

// The first pass of a split run works on the batch's unsafe ops, whose count
// the split kernel publishes on the device after the host has routed the
// pass: the host mirror still holds the whole batch's op count then, so an
// op pass (seed or deletion roots) of such a run starts in the team.  The
// team checks the published count itself and hands a pass that is still too
// large to grid mode before running it.
bool team_ok_first(const rg_labels *L, const Ctl &h, bool split)
{
    if (split && h.stage == 0 && (h.phase == PH_SEED || h.phase == PH_DEL_ROOTS)) return true;
    return team_ok(L, h);
}

// seed: a mixed batch starts with its deletions and seeds its inserts in the
// pull pass (or right after the roots pass when no tree edge was deleted)
void engine_init(rg_labels *L, uint32_t phase0, bool seed, uint32_t n_ops, bool new_idx)
{
    Ctl &h = *L->hc;
    h.phase = phase0;
    h.stage = 0;
    h.level = 0;
    h.seed = seed ? 1u : 0u;
    h.n_cur = h.n_next = h.n_tree = h.n_chunk = 0;
    h.n_ops = n_ops;
    h.sel_cur = 0; h.sel_next = 1; h.sel_tree = 2;
    h.epoch += 1;
    if (new_idx || phase0 == PH_DEL_ROOTS) {
        // invalidation marks of earlier runs stay in place and never match a
        // fresh index; when the index wraps the marks are cleared
        h.idx += 1;
        if (h.idx >= 0x3FFFFFF0) {   // the device's valid stamps need idx < 2^30
            h.idx = 1;
            launch_clear_marks(L->vr, (uint32_t)L->g->V, L->g->st);
        }
    }
    h.flags &= F_ALL_MODIFIED;
    if (phase0 == PH_BUILD_INIT && L->tune.dense) h.flags |= F_BUILD;
    h.e_next = h.e_live = 0;
    h.dense = 0;
    h.bar_count = h.bar_gen = h.bar_exit = 0;
    h.rbound = L->algo == SSWP ? 0u : 0xFFFFFFFFu;
    h.work_max = team_cap(L);
    h.levels = 0;
}

// grid steps enqueued per host round trip (RG_GRID_SPEC overrides): the
// steps after the first are speculative and turn into no-ops once the
// tree's passes fit the team again
int grid_spec()
{
    static const int v = [] {
        const char *e = getenv("RG_GRID_SPEC");
        const int x = e && *e ? atoi(e) : 1;
        return std::max(1, std::min(8, x));
    }();
    return v;
}

// Run the engines of labels L[0..n) from `phase0` to completion, after the
// store batch already queued on the graph stream.  Trees whose passes fit the
// team cap share one persistent launch on the graph stream (W workgroups per
// tree, so the trees repair concurrently on different CUs); larger passes go
// grid-wide on the tree's own stream, concurrently with the team launch and
// with other trees' grid-wide passes.  Every round ends with all streams
// drained, so the host's view of each control block is current.
// Hooks of a run: `before_launch` is called once every tree's control
// block is initialised in its pinned mirror and before anything of the run
// is enqueued (the split kernels go there: they write n_ops into the
// mirrors); `after_first_enqueue` once the first round's launches are
// enqueued, before the host waits for them.
struct RunHooks {
    std::function<void()> before_launch, after_first_enqueue;
};

void engine_run_multi(rg_labels **L, int n, uint32_t phase0, bool seed, uint32_t n_ops, bool new_idx,
                      rg_batch_stats *stats, bool split = false, const RunHooks *hooks = nullptr)
{
    if (n <= 0) return;
    rg_graph *g = L[0]->g;
    hipStream_t st = g->st;
    std::vector<EngineArgs> args(n);
    std::vector<char> active(n, 1), pending_init(n, 0);
    // the run's team layout: per-tree XCD slots, workgroups per slot sized
    // by the batch (an even split gives every tree team_width workgroups)
    uint32_t kslots[MAXT] = {};
    const bool one_launch = n <= MAXT;
    g->run_tune = &g->tune_tab[tune_bucket(n_ops)];
    if (one_launch) split_slots(L, n, kslots);
    const uint32_t rounds =
        std::max<uint32_t>(1, std::min<uint32_t>((uint32_t)n * team_width(L[0], n_ops) / XCD_SLOTS,
                                                 team_resident_max() / XCD_SLOTS));
    for (int i = 0; i < n; i++) {
        L[i]->run_w = one_launch ? rounds * kslots[i] : team_width(L[i], n_ops);
        engine_init(L[i], phase0, seed, n_ops, new_idx);
        // a run that starts in the team hands its initial control block to
        // the kernel through the pinned mirror; otherwise it is copied below
        if (team_ok_first(L[i], *L[i]->hc, split)) pending_init[i] = 1;
    }
    if (hooks && hooks->before_launch) hooks->before_launch();
    for (int i = 0; i < n; i++) {
        if (!pending_init[i]) HC(hipMemcpyAsync(L[i]->ctl, L[i]->hc, sizeof(Ctl), hipMemcpyHostToDevice, st));
        args[i] = engine_args(L[i], split);
        // the team kernel reads the arguments from device memory; refresh
        // them when they changed (op buffers regrown, tuning)
        if (memcmp(L[i]->hargs, &args[i], sizeof(EngineArgs)) != 0) {
            memcpy(L[i]->hargs, &args[i], sizeof(EngineArgs));
            HC(hipMemcpyAsync(L[i]->dargs, L[i]->hargs, sizeof(EngineArgs), hipMemcpyHostToDevice, st));
        }
    }
    // grid-wide passes of different trees run concurrently, each on its
    // tree's stream; the first ones wait for what the graph stream still
    // has queued (the store batch); later rounds start with every stream
    // drained by the host
    std::vector<char> on_own(n, 0);
    bool forked = false;
    int left = n;
    uint64_t guard_iter = 0;
    std::vector<int> team;
    // progress guard: every round must move some tree's state machine
    auto state_of = [](const Ctl &h) {
        return std::array<uint64_t, 4>{h.phase | (uint64_t)h.stage << 32, h.epoch, h.n_cur, h.levels};
    };
    std::vector<std::array<uint64_t, 4>> last(n);
    for (int i = 0; i < n; i++) last[i] = state_of(*L[i]->hc);
    int stalled = 0;
    while (left) {
        team.clear();
        for (int i = 0; i < n; i++) {
            on_own[i] = 0;
            if (!active[i]) continue;
            const Ctl &h = *L[i]->hc;
            if (guard_iter == 0 ? team_ok_first(L[i], h, split) : team_ok(L[i], h)) {
                team.push_back(i);
            } else {
                // a grid-wide step on the tree's own stream; it turns into a
                // no-op once the tree's passes fit the team again (or it is
                // done)
                hipStream_t ts = L[i]->st;
                if (guard_iter == 0) {
                    if (!forked) HC(hipEventRecord(g->ev[2], st));
                    forked = true;
                    HC(hipStreamWaitEvent(ts, g->ev[2], 0));
                }
                hipEvent_t ea, eb;
                prof_begin(g, 2, host_work(h, g->V), ts, &ea, &eb, (int)h.phase);
                if (ea) HC(hipEventRecord(ea, ts));
                launch_engine_grid_steps(args[i], h, grid_spec(), ts);
                if (eb) HC(hipEventRecord(eb, ts));
                on_own[i] = 1;
                if (stats) stats->grid_steps++;
            }
        }
        // the team trees share one launch on their XCD slots (a run of more
        // than MAXT trees: one launch per tree, all eight slots each)
        for (size_t b = 0; b < team.size(); b += one_launch ? team.size() : 1) {
            TeamLaunch tl{};
            const size_t e = one_launch ? team.size() : b + 1;
            for (size_t k = b; k < e; k++) {
                const int i = team[k], t = one_launch ? i : 0;
                tl.a[t] = L[i]->dargs;
                tl.k[t] = one_launch ? kslots[i] : XCD_SLOTS;
                tl.work_max[t] = team_cap(L[i]);
                tl.chunk_max[t] = team_chunk_cap(L[i]);
                if (pending_init[i]) tl.init_mask |= 1u << t;
                pending_init[i] = 0;
            }
            tl.rounds = one_launch ? rounds : std::max<uint32_t>(1, L[team[b]]->run_w / XCD_SLOTS);
            hipEvent_t ea, eb;
            prof_begin(g, 1, (uint64_t)(e - b), st, &ea, &eb);
            if (ea) HC(hipEventRecord(ea, st));
            launch_engine_team(tl, st);
            if (eb) HC(hipEventRecord(eb, st));
        }
        team.clear();
        HC(hipGetLastError());
        if (guard_iter == 0 && hooks && hooks->after_first_enqueue) hooks->after_first_enqueue();
        HC(hipStreamSynchronize(st));
        for (int i = 0; i < n; i++)
            if (on_own[i]) HC(hipStreamSynchronize(L[i]->st));
        for (int i = 0; i < n; i++) {
            if (!active[i]) continue;
            Ctl &h = *L[i]->hc;   // published by the last kernel
            if (h.flags & F_QUEUE_OVERFLOW) throw RgError{RG_EDEVICE, "engine queue overflow"};
            if (h.flags & F_ABORT) throw RgError{RG_EDEVICE, "engine team barrier timed out"};
            if (h.phase == PH_DONE) {
                active[i] = 0;
                left--;
                if (stats) stats->levels += h.levels;
                L[i]->dense_passes += h.dense;
                // RG_LOG_SUBTREE=n: report runs whose deletion subtree exceeded n vertices
                static const long log_subtree = [] {
                    const char *e = getenv("RG_LOG_SUBTREE");
                    return e ? atol(e) : 0L;
                }();
                if (log_subtree > 0 && phase0 == PH_DEL_ROOTS && h.n_tree > (uint64_t)log_subtree)
                    fprintf(stderr, "[rg] subtree: algo %d vertices %u levels %llu grid steps so far %llu\n", L[i]->algo,
                            h.n_tree, (unsigned long long)h.levels,
                            (unsigned long long)(stats ? stats->grid_steps : 0));
            } else if (h.phase == PH_REBUILD) {
                // deletion fallback (graph.hpp:1252-1321): full recompute
                static const bool log_rebuild = [] {
                    const char *e = getenv("RG_LOG_REBUILD");
                    return e && *e && *e != '0';
                }();
                if (log_rebuild)
                    fprintf(stderr, "[rg] rebuild: algo %d level %u subtree %u frontier %u flags %#x\n", L[i]->algo,
                            h.level, h.n_tree, h.n_cur, h.flags);
                h.phase = PH_BUILD_INIT;
                h.stage = 0;
                h.seed = 0;
                h.n_cur = h.n_next = h.n_tree = h.n_chunk = 0;
                h.sel_cur = 0; h.sel_next = 1; h.sel_tree = 2;
                h.epoch += 1;
                h.flags = (h.flags & ~(F_NEED_REBUILD | F_DENSE)) | F_ALL_MODIFIED | (L[i]->tune.dense ? F_BUILD : 0u);
                h.e_next = h.e_live = 0;
                // the rebuild runs grid-wide on the tree's stream
                HC(hipMemcpyAsync(L[i]->ctl, &h, sizeof(Ctl), hipMemcpyHostToDevice, L[i]->st));
                if (stats) stats->rebuilds++;
            }
        }
        bool moved = false;
        for (int i = 0; i < n; i++) {
            const auto now = state_of(*L[i]->hc);
            if (now != last[i]) moved = true;
            last[i] = now;
        }
        stalled = moved ? 0 : stalled + 1;
        if (stalled >= 3) throw RgError{RG_EDEVICE, "engine made no progress"};
        if (++guard_iter > (1ull << 26)) throw RgError{RG_EDEVICE, "engine did not converge"};
    }
    // end of the run: every stream is drained, the graph stream is the join point
    HC(hipEventRecord(g->ev[3], st));
}

void engine_run(rg_labels *L, uint32_t phase0, bool seed, uint32_t n_ops, bool new_idx,
                rg_batch_stats *stats)
{
    engine_run_multi(&L, 1, phase0, seed, n_ops, new_idx, stats);
}

void hist_free(rg_labels *L)
{
    auto &h = L->hist;
    dfree(h.a.rec); dfree(h.a.head); dfree(h.a.base); dfree(h.dcount);
    hfree(h.hcount);
    h = rg_labels::Hist{};
}

// fold the oldest batches of L's history until `need` more records fit,
// or every batch with version <= `upto`
void hist_fold(rg_labels *L, uint64_t need, uint64_t upto)
{
    auto &h = L->hist;
    rg_graph *g = L->g;
    uint64_t new_tail = h.a.tail;
    uint64_t base_version = h.base_version;
    while (!h.batches.empty() &&
           ((need && h.top + need - new_tail > h.a.cap) || (!need && h.batches.front().first <= upto))) {
        base_version = h.batches.front().first;
        h.batches.pop_front();
        new_tail = h.batches.empty() ? h.top : h.batches.front().second;
    }
    if (need && h.top + need - new_tail > h.a.cap)
        throw RgError{RG_ENOMEM, "history ring smaller than one batch's changed vertices"};
    if (new_tail == h.a.tail && base_version == h.base_version) return;
    launch_hist_fold(h.a, new_tail, g->st);
    HC(hipGetLastError());
    h.a.tail = new_tail;
    h.base_version = base_version;
}

// After a batch's repair: one record per vertex whose value changed
// (the modified set, or every vertex after a full rebuild), then the
// modified set starts over for the next batch.
void hist_commit(rg_labels *L, uint64_t version)
{
    auto &h = L->hist;
    rg_graph *g = L->g;
    Ctl &c = *L->hc;
    const bool all = (c.flags & F_ALL_MODIFIED) != 0;
    const uint64_t n = all ? g->V : std::min<uint64_t>(c.n_mod, g->V);
    if (n) {
        hist_fold(L, n, 0);
        HC(hipMemsetAsync(h.dcount, 0, 4, g->st));
        launch_hist_append(h.a, L->vr, all ? nullptr : L->mlist, (uint32_t)n, (uint32_t)version, h.top, h.dcount,
                           g->st);
        HC(hipGetLastError());
        HC(hipMemcpyAsync(h.hcount, h.dcount, 4, hipMemcpyDeviceToHost, g->st));
        HC(hipStreamSynchronize(g->st));
        if (*h.hcount) {
            h.batches.emplace_back(version, h.top);
            h.top += *h.hcount;
        }
    }
    c.n_mod = 0;
    c.mepoch += 1;
    c.flags &= ~F_ALL_MODIFIED;
}

void labels_free(rg_labels *L)
{
    hist_free(L);
    dfree(L->vr); dfree(L->init); dfree(L->mstamp);
    dfree(L->q[0]); dfree(L->q[1]); dfree(L->q[2]); dfree(L->cq); dfree(L->bnd); dfree(L->mlist); dfree(L->ctl);
    dfree(L->init_active);
    dfree(L->ticks);
    dfree(L->bar);
    dfree(L->dargs);
    hfree(L->hargs);
    hfree(L->hc);
    if (L->st) (void)hipStreamDestroy(L->st);
    L->st = nullptr;
}

}  // namespace

// =====================================================================
// C ABI
// =====================================================================
extern "C" {

const char *rg_last_error(void) { return g_err.c_str(); }

int rg_graph_create(uint64_t vertices, uint64_t hint_edges, int symmetric, int dual, int weighted,
                    rg_graph **out)
{
    if (!out) return RG_EINVAL;
    *out = nullptr;
    rg_graph *g = new rg_graph();
    int rc = guard([&] {
        if (vertices == 0 || vertices > MAX_V) throw RgError{RG_EINVAL_VERTEX, "vertex count must be in [1, 2^31)"};
        g->V = vertices;
        g->hint = hint_edges ? hint_edges : vertices * 16;   // graph.hpp:64
        g->symmetric = symmetric;
        g->dual = dual;
        g->weighted = weighted;
        HC(hipStreamCreateWithFlags(&g->st, hipStreamNonBlocking));
        for (auto &e : g->ev) HC(hipEventCreate(&e));
        HC(hipEventCreateWithFlags(&g->ev_safe, hipEventDisableTiming));
        g->h_desc = halloc<BatchDesc>(1);
        g->d_desc = dalloc<BatchDesc>(1);
        alloc_list(g->lists[0], vertices, g->st);
        if (!symmetric) alloc_list(g->lists[1], vertices, g->st);
        else g->lists[1] = g->lists[0];
        const char *env = getenv("RG_SLOT_FACTOR");
        double f = env ? atof(env) : 6.0;
        g->slot_cap = std::max<uint64_t>(1ull << 20, (uint64_t)(f * (double)g->hint)) + 4 * vertices;
        g->hash_cap = std::max<uint64_t>(1ull << 20, 8 * g->hint);
        // a batch frees at most a few blocks per list op; the class lists hold
        // one per 1,024 pool entries (4,096 to 524,288)
        g->fcap = 4096;
        while (g->fcap < (1u << 19) && (uint64_t)g->fcap * 1024 < g->slot_cap) g->fcap <<= 1;
        uint64_t nfree = 2ull * NCLS * g->fcap;
        g->free_off = dalloc<uint64_t>(nfree);
        g->pend_off = dalloc<uint64_t>(nfree);
        g->slots.nbr = dalloc<uint32_t>(g->slot_cap);
        g->slots.cnt = dalloc<uint32_t>(g->slot_cap);
        g->slots.wt = weighted ? dalloc<uint32_t>(g->slot_cap) : nullptr;
        g->slots.hb = dalloc<HBucket>(g->hash_cap);
        g->pool = dalloc<Pool>(1);
        Pool p{};
        p.slot_cap = g->slot_cap;
        p.hash_cap = g->hash_cap;
        HC(hipMemcpyAsync(g->pool, &p, sizeof(Pool), hipMemcpyHostToDevice, g->st));
        g->err = dalloc<uint32_t>(4);
        HC(hipMemsetAsync(g->err, 0, 16, g->st));
        g->h_status = halloc<uint32_t>(4);
        memset(g->h_status, 0, 16);
        g->scratch = scratch_address();   // set-up of the no-alloc entry point on this device
        HC(hipStreamSynchronize(g->st));
    });
    if (rc != RG_OK) {
        rg_graph_destroy(g);
        return rc;
    }
    *out = g;
    return RG_OK;
}

int rg_graph_destroy(rg_graph *g)
{
    if (!g) return RG_OK;
    if (g->st) (void)hipStreamSynchronize(g->st);
    if (g->wal_fd >= 0) (void)close(g->wal_fd);
    free_list(g->lists[0]);
    if (!g->symmetric) free_list(g->lists[1]);
    dfree(g->slots.nbr); dfree(g->slots.cnt); dfree(g->slots.wt); dfree(g->slots.hb);
    dfree(g->pool); dfree(g->err);
    dfree(g->free_off); dfree(g->pend_off); dfree(g->cnt_tmp[0]); dfree(g->cnt_tmp[1]);
    free_ops(g);
    hfree(g->h_src); hfree(g->h_dst); hfree(g->h_w); hfree(g->h_res); hfree(g->h_kind); hfree(g->h_flag);
    hfree(g->h_status);
    for (auto &e : g->ev) if (e) (void)hipEventDestroy(e);
    if (g->ev_safe) (void)hipEventDestroy(g->ev_safe);
    captured_clear(g);
    hfree(g->h_desc);
    dfree(g->d_desc);
    for (auto &r : g->prof_pending) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto &e : g->ev_free) (void)hipEventDestroy(e);
    if (g->st) (void)hipStreamDestroy(g->st);
    delete g;
    return RG_OK;
}

int rg_graph_info(rg_graph *g, uint64_t *vertices, uint64_t *slots_used, uint64_t *slots_cap,
                  uint64_t *hash_used, uint64_t *hash_cap)
{
    if (!g) return RG_EINVAL;
    return guard([&] {
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        Pool p{};
        HC(hipMemcpyAsync(&p, g->pool, sizeof(Pool), hipMemcpyDeviceToHost, g->st));
        HC(hipStreamSynchronize(g->st));
        if (vertices) *vertices = g->V;
        if (slots_used) *slots_used = std::min<uint64_t>(p.slot_top, p.slot_cap);
        if (slots_cap) *slots_cap = p.slot_cap;
        if (hash_used) *hash_used = std::min<uint64_t>(p.hash_top, p.hash_cap);
        if (hash_cap) *hash_cap = p.hash_cap;
    });
}

int rg_graph_pool_info(rg_graph *g, uint64_t *out)
{
    if (!g || !out) return RG_EINVAL;
    return guard([&] {
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        Pool p{};
        HC(hipMemcpyAsync(&p, g->pool, sizeof(Pool), hipMemcpyDeviceToHost, g->st));
        HC(hipStreamSynchronize(g->st));
        for (int h = 0; h < 2; h++) {
            out[h] = p.recycled[h];
            out[2 + h] = p.dropped[h];
            out[4 + h] = p.dropped_entries[h];
        }
    });
}

int rg_edge_update_batch(rg_graph *g, const rg_edge *edges, uint64_t n, int delta, int directed,
                         uint64_t *old_out)
{
    if (!g || (n && !edges) || (delta != 1 && delta != -1)) return RG_EINVAL;
    return guard([&] {
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        if (n == 0) return;
        uint8_t kind = (delta > 0 ? K_INS : K_DEL) | (directed ? 0 : K_UNDIR);
        upload_edges(g, edges, n, kind);
        run_store(g);
        HC(hipMemcpyAsync(g->h_res, g->ops.res, n * sizeof(OpResult), hipMemcpyDeviceToHost, g->st));
        HC(hipStreamSynchronize(g->st));
        store_status_check(g);
        if (old_out)
            for (uint64_t i = 0; i < n; i++) old_out[i] = g->h_res[i].old;
    });
}

int rg_edge_update(rg_graph *g, uint64_t src, uint64_t dst, uint64_t data, int delta, int directed,
                   uint64_t *old)
{
    rg_edge e{src, dst, data};
    return rg_edge_update_batch(g, &e, 1, delta, directed, old);
}

int rg_edge_update_batch_static(rg_graph *g, const rg_edge *d_edges, uint64_t n, int delta, int directed,
                                uint64_t *d_old)
{
    if (!g || (n && !d_edges) || (delta != 1 && delta != -1)) return RG_EINVAL;
    if (n > STATIC_MAX_OPS) {
        g_err = "batch exceeds the static scratch (" + std::to_string(STATIC_MAX_OPS) + " edge ops at most)";
        return RG_EINVAL;
    }
    std::lock_guard<std::recursive_mutex> lk(g->mu);
    std::lock_guard<std::mutex> sl(g_static_mu);
    try {
        if (n == 0) return RG_OK;
        if (!g->scratch) throw RgError{RG_EDEVICE, "static scratch not set up on this device"};
        unsigned char *sc = g->scratch;
        EdgeOps ops{};
        ops.src = (uint32_t *)(sc + SC_SRC);
        ops.dst = (uint32_t *)(sc + SC_DST);
        ops.w = (uint32_t *)(sc + SC_W);
        ops.kind = (uint8_t *)(sc + SC_KIND);
        ops.res = (OpResult *)(sc + SC_RES);
        ops.n = (uint32_t)n;
        ListOps lops{};
        lops.op = (LOp *)(sc + SC_LOP);
        lops.segbuf = (uint32_t *)(sc + SC_SEG);
        lops.owner = (LOp *)(sc + SC_OWNER);
        lops.heavy = (uint32_t *)(sc + SC_HEAVY);
        lops.n = (uint32_t)(4 * n);
        // validate and convert first: a batch with a bad id or weight is
        // rejected whole, like the host-buffer entry point
        g->h_status[2] = g->h_status[3] = 0;
        const uint8_t kind = (uint8_t)((delta > 0 ? K_INS : K_DEL) | (directed ? 0 : K_UNDIR));
        const dim3 grid((uint32_t)((n + 255) / 256)), blk(256);
        hipLaunchKernelGGL(k_edges_to_ops, grid, blk, 0, g->st, d_edges, (uint32_t)n, kind, g->V, g->weighted, ops,
                           g->h_status + 2);
        HC(hipGetLastError());
        HC(hipStreamSynchronize(g->st));
        if (g->h_status[2]) throw RgError{RG_EINVAL_VERTEX, "VertexId error."};
        if (g->h_status[3]) throw RgError{RG_EINVAL, "edge weight exceeds 32 bits"};
        StoreArgs a = store_args(g);
        a.ops = ops;
        a.lops = lops;
        a.lstride = 4;
        if (!launch_store_batch(a, g->st)) launch_pool_recycle(a, g->st);
        if (d_old) hipLaunchKernelGGL(k_old_out, grid, blk, 0, g->st, ops.res, (uint32_t)n, d_old);
        HC(hipGetLastError());
        HC(hipStreamSynchronize(g->st));
        store_status_check(g);
        return RG_OK;
    } catch (const RgError &e) {
        g_err = e.msg;
        return e.code;
    }
}

int rg_degree(rg_graph *g, uint64_t v, int incoming, uint64_t *out)
{
    if (!g || !out) return RG_EINVAL;
    return guard([&] {
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        if (v >= g->V) throw RgError{RG_EINVAL_VERTEX, "VertexId error."};
        const List &l = (incoming && !g->symmetric) ? g->lists[1] : g->lists[0];
        uint32_t d = 0;
        HC(hipMemcpyAsync(&d, &l.h[v].deg, 4, hipMemcpyDeviceToHost, g->st));
        HC(hipStreamSynchronize(g->st));
        *out = d;
    });
}

int rg_edge_num(rg_graph *g, uint64_t src, uint64_t dst, uint64_t data, uint64_t *out)
{
    if (!g || !out) return RG_EINVAL;
    return guard([&] {
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        if (src >= g->V || dst >= g->V) throw RgError{RG_EINVAL_VERTEX, "VertexId error."};
        ListHdr hd{};
        HC(hipMemcpyAsync(&hd, &g->lists[0].h[src], sizeof(ListHdr), hipMemcpyDeviceToHost, g->st));
        HC(hipStreamSynchronize(g->st));
        const uint64_t off = hd.off;
        const uint32_t len = hd.len;
        std::vector<uint32_t> nb(len), cnt(len), w(len);
        if (len) {
            HC(hipMemcpyAsync(nb.data(), g->slots.nbr + off, len * 4ull, hipMemcpyDeviceToHost, g->st));
            HC(hipMemcpyAsync(cnt.data(), g->slots.cnt + off, len * 4ull, hipMemcpyDeviceToHost, g->st));
            if (g->weighted)
                HC(hipMemcpyAsync(w.data(), g->slots.wt + off, len * 4ull, hipMemcpyDeviceToHost, g->st));
            HC(hipStreamSynchronize(g->st));
        }
        *out = 0;
        for (uint32_t i = 0; i < len; i++)
            if ((nb[i] & NBR_MASK) == dst && (!g->weighted || w[i] == data)) *out = cnt[i];
    });
}

int rg_labels_create(rg_graph *g, int algo, uint64_t root, const uint64_t *init_values, int trace_modified,
                     rg_labels **out)
{
    if (!g || !out) return RG_EINVAL;
    if (algo < RG_BFS || algo > RG_WCC) {
        g_err = "unknown algorithm";
        return RG_EALGO;
    }
    *out = nullptr;
    rg_labels *L = new rg_labels();
    int rc = guard([&] {
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        if (algo != RG_WCC && root >= g->V) throw RgError{RG_EINVAL_VERTEX, "root out of range"};
        uint64_t V = g->V;
        L->g = g;
        L->algo = algo;
        L->root = (uint32_t)root;
        L->trace = trace_modified;
        L->vr = dalloc<VRec>(V);
        L->init = dalloc<uint32_t>(V);
        L->mstamp = dalloc<uint32_t>(V);
        for (auto &q : L->q) q = dalloc<uint32_t>(V);
        L->cq_cap = (uint32_t)std::min<uint64_t>(0xFFFFFFF0ull,
                                                  g->slot_cap / CHUNK + 2 * std::min<uint64_t>(V, g->slot_cap / LONG_LIST) + 1024);
        L->cq = dalloc<ChunkItem>(L->cq_cap);
        if (algo == RG_SSWP) L->bnd = dalloc<uint32_t>(V);
        L->mlist = dalloc<uint32_t>(V);
        L->ctl = dalloc<Ctl>(1);
        L->ticks = dalloc<unsigned long long>(NTICK * NPHASE);
        HC(hipMemsetAsync(L->ticks, 0, NTICK * NPHASE * 8, g->st));
        L->bar = dalloc<uint32_t>(BAR_SUBS * BAR_STRIDE);
        HC(hipMemsetAsync(L->bar, 0, BAR_SUBS * BAR_STRIDE * 4, g->st));
        L->hc = halloc<Ctl>(1);
        memset(L->hc, 0, sizeof(Ctl));
        L->dargs = dalloc<EngineArgs>(1);
        L->hargs = halloc<EngineArgs>(1);
        memset(L->hargs, 0, sizeof(EngineArgs));
        L->hc->mepoch = 1;
        HC(hipStreamCreateWithFlags(&L->st, hipStreamNonBlocking));
        HC(hipMemsetAsync(L->vr, 0, V * sizeof(VRec), g->st));
        HC(hipMemsetAsync(L->mstamp, 0, V * 4, g->st));
        // labels start as (init, empty parent) like alloc_vertex_tree_array
        if (algo == RG_WCC) {
            L->init_active = nullptr;
            if (init_values) {
                std::vector<uint32_t> tmp(V);
                for (uint64_t v = 0; v < V; v++) {
                    if (init_values[v] > 0xFFFFFFFEull) throw RgError{RG_EINVAL, "WCC init label exceeds 32 bits"};
                    tmp[v] = (uint32_t)init_values[v];
                }
                HC(hipMemcpyAsync(L->init, tmp.data(), V * 4, hipMemcpyHostToDevice, g->st));
                HC(hipStreamSynchronize(g->st));
            } else {
                EngineArgs a = engine_args(L);
                launch_engine_fill_init(a, g->st);
            }
            // every vertex starts active (wcc_inc.cpp:94-97)
            L->init_active = dalloc<uint32_t>(V);
            HC(hipMemsetAsync(L->init_active, 0xFF, V * 4, g->st));
        } else {
            EngineArgs a = engine_args(L);
            launch_engine_fill_init(a, g->st);
        }
        // BFS / SSSP rescue on the first deletion level only: deeper rescue
        // scans cost more passes than the subtree they save on this device.
        // WCC rescues on every level: nearly every vertex of an invalidated
        // subtree has a neighbour outside it with the same component label,
        // so rescue stops the expansion at once, where expanding can
        // invalidate a large part of a component and force a full
        // recompute.  SSWP rescues on every level too: each scan also raises
        // the vertex's lower bound, under which the expand leaves children
        // alone, so the large widest-path subtrees stay small.  (The
        // reference's policy, level <= 5 or frontier < 16384, is one
        // rg_labels_set_tuning call away; values do not depend on the
        // policy.)
        L->tune.rescue_level = 0;
        L->tune.rescue_frontier = (algo == RG_WCC || algo == RG_SSWP) ? 0xFFFFFFFFu : 0u;
        L->tune.recompute_level = 10;
        L->tune.recompute_abs = 2000000;
        L->tune.recompute_frac = 0.05f;
        L->tune.walk_limit = (uint32_t)std::max<uint64_t>(64, std::min<uint64_t>(V / 10, 1u << 16));
        // builds stay push-only unless asked (rg_labels_dense, RG_DENSE=1):
        // on this GPU the reference's pull passes made scale-24 builds
        // 25-80% slower (a pull pass reads every in-edge; pushes from the
        // frontier are cheap atomics here)
        L->tune.dense = 0;
        if (const char *de = getenv("RG_DENSE"); de && *de == '1') L->tune.dense = 1;
        // environment overrides; empty values are ignored
        if (const char *tw = getenv("RG_TEAM_W"); tw && *tw)
            L->team_w = (uint32_t)std::max(1, std::min(64, atoi(tw)));
        if (const char *bm = getenv("RG_BLOCK_MAX"); bm && *bm) L->block_max = (uint32_t)strtoul(bm, nullptr, 0);
        if (const char *cm = getenv("RG_CHUNK_MAX"); cm && *cm) L->chunk_max = (uint32_t)strtoul(cm, nullptr, 0);
        HC(hipMemcpyAsync(L->ctl, L->hc, sizeof(Ctl), hipMemcpyHostToDevice, g->st));
        HC(hipStreamSynchronize(g->st));
    });
    if (rc != RG_OK) {
        labels_free(L);
        delete L;
        return rc;
    }
    *out = L;
    return RG_OK;
}

int rg_labels_destroy(rg_labels *L)
{
    if (!L) return RG_OK;
    if (L->g && L->g->st) (void)hipStreamSynchronize(L->g->st);
    labels_free(L);
    delete L;
    return RG_OK;
}

int rg_labels_set_tuning(rg_labels *L, uint32_t rescue_level, uint32_t rescue_frontier,
                         uint32_t recompute_level, uint32_t recompute_abs, double recompute_frac,
                         uint32_t walk_limit, uint32_t block_max)
{
    if (!L || recompute_frac < 0) return RG_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(L->g->mu);
    L->tune.rescue_level = rescue_level;
    L->tune.rescue_frontier = rescue_frontier;
    L->tune.recompute_level = recompute_level;
    L->tune.recompute_abs = recompute_abs;
    L->tune.recompute_frac = (float)recompute_frac;
    L->tune.walk_limit = walk_limit ? walk_limit
                                    : (uint32_t)std::max<uint64_t>(64, std::min<uint64_t>(L->g->V / 10, 1u << 16));
    L->block_max = block_max;
    return RG_OK;
}

int rg_build(rg_labels *L, uint64_t *result)
{
    if (!L) return RG_EINVAL;
    return guard([&] {
        std::lock_guard<std::recursive_mutex> lk(L->g->mu);
        L->hc->flags |= F_ALL_MODIFIED;
        rg_batch_stats st{};
        engine_run(L, PH_BUILD_INIT, false, 0, false, &st);
        if (result) *result = st.levels;
    });
}

int rg_inc_add(rg_labels *L, const rg_edge *edges, uint64_t n, int directed, uint64_t *result)
{
    if (!L || (n && !edges)) return RG_EINVAL;
    return guard([&] {
        rg_graph *g = L->g;
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        if (result) *result = 0;
        if (!n) return;
        upload_edges(g, edges, n, K_INS | K_FORCE | K_NOSTORE | (directed ? 0 : K_UNDIR));
        rg_batch_stats st{};
        engine_run(L, PH_SEED, false, (uint32_t)n, false, &st);
        if (result) *result = st.levels;
    });
}

int rg_inc_del(rg_labels *L, const rg_edge *edges, uint64_t n, int directed, uint64_t *result)
{
    if (!L || (n && !edges)) return RG_EINVAL;
    return guard([&] {
        rg_graph *g = L->g;
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        if (result) *result = 0;
        if (!n) return;
        upload_edges(g, edges, n, K_DEL | K_FORCE | K_NOSTORE | (directed ? 0 : K_UNDIR));
        rg_batch_stats st{};
        engine_run(L, PH_DEL_ROOTS, false, (uint32_t)n, true, &st);
        if (result) *result = st.levels;
    });
}

int rg_classify(rg_labels *L, const rg_edge *edges, const int8_t *kind, uint64_t n, int directed,
                uint8_t *unsafe)
{
    if (!L || (n && (!edges || !kind || !unsafe))) return RG_EINVAL;
    return guard([&] {
        rg_graph *g = L->g;
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        if (!n) return;
        upload_edges(g, edges, n, K_NOSTORE | (directed ? 0 : K_UNDIR));
        for (uint64_t i = 0; i < n; i++)
            g->h_kind[i] = (uint8_t)(K_NOSTORE | (directed ? 0 : K_UNDIR) | (kind[i] > 0 ? K_INS : K_DEL));
        HC(hipMemcpyAsync(g->ops.kind, g->h_kind, n, hipMemcpyHostToDevice, g->st));
        EngineArgs a = engine_args(L);
        launch_classify(a, (uint32_t)n, g->d_flag, g->st);
        HC(hipGetLastError());
        HC(hipMemcpyAsync(g->h_flag, g->d_flag, n, hipMemcpyDeviceToHost, g->st));
        HC(hipStreamSynchronize(g->st));
        memcpy(unsafe, g->h_flag, n);
    });
}

int rg_graph_version(rg_graph *g, uint64_t *version)
{
    if (!g || !version) return RG_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(g->mu);
    *version = g->version;
    return RG_OK;
}

int rg_history_enable(rg_labels *L, uint64_t records)
{
    if (!L || records == 0) return RG_EINVAL;
    return guard([&] {
        rg_graph *g = L->g;
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        if (!L->trace) throw RgError{RG_EINVAL, "history needs labels created with trace_modified"};
        hist_free(L);
        auto &h = L->hist;
        h.a.cap = records;
        h.a.rec = dalloc<HRec>(records);
        h.a.head = dalloc<uint64_t>(g->V);
        h.a.base = dalloc<uint32_t>(g->V);
        h.a.tail = 0;
        h.dcount = dalloc<uint32_t>(1);
        h.hcount = halloc<uint32_t>(1);
        h.top = 0;
        h.base_version = g->version;
        launch_hist_init(h.a, L->vr, (uint32_t)g->V, g->st);
        HC(hipGetLastError());
        // the snapshot is the base; changes from here on are recorded
        Ctl &c = *L->hc;
        c.n_mod = 0;
        c.mepoch += 1;
        c.flags &= ~F_ALL_MODIFIED;
        HC(hipMemcpyAsync(L->ctl, L->hc, sizeof(Ctl), hipMemcpyHostToDevice, g->st));
        HC(hipStreamSynchronize(g->st));
        h.on = true;
    });
}

int rg_history_read(rg_labels *L, uint64_t version, uint64_t first, uint64_t count, uint64_t *values)
{
    if (!L) return RG_EINVAL;
    return guard([&] {
        rg_graph *g = L->g;
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        auto &h = L->hist;
        if (!h.on) throw RgError{RG_EINVAL, "history is not enabled on these labels"};
        if (first + count > g->V || first + count < first) throw RgError{RG_EINVAL_VERTEX, "range out of bounds"};
        if (version > g->version) throw RgError{RG_EINVAL, "version is in the future"};
        if (version < h.base_version) throw RgError{RG_EINVAL, "version was folded (older than the history keeps)"};
        if (!count) return;
        uint64_t *dv = dalloc<uint64_t>(count);
        launch_hist_read(h.a, (uint32_t)first, (uint32_t)count, (uint32_t)version, dv, g->st);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(values, dv, count * 8, hipMemcpyDeviceToHost, g->st);
        if (e == hipSuccess) e = hipStreamSynchronize(g->st);
        dfree(dv);
        hipchk(e, "history read");
    });
}

int rg_history_gc(rg_labels *L, uint64_t version)
{
    if (!L) return RG_EINVAL;
    return guard([&] {
        rg_graph *g = L->g;
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        if (!L->hist.on) throw RgError{RG_EINVAL, "history is not enabled on these labels"};
        hist_fold(L, 0, version);
        HC(hipStreamSynchronize(g->st));
    });
}

int rg_history_info(rg_labels *L, uint64_t *oldest_version, uint64_t *records, uint64_t *capacity)
{
    if (!L) return RG_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(L->g->mu);
    const auto &h = L->hist;
    if (!h.on) {
        g_err = "history is not enabled on these labels";
        return RG_EINVAL;
    }
    if (oldest_version) *oldest_version = h.base_version;
    if (records) *records = h.top - h.a.tail;
    if (capacity) *capacity = h.a.cap;
    return RG_OK;
}

int rg_labels_read(rg_labels *L, uint64_t first, uint64_t count, uint64_t *values, uint64_t *parents)
{
    if (!L) return RG_EINVAL;
    return guard([&] {
        rg_graph *g = L->g;
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        if (first + count > g->V || first + count < first) throw RgError{RG_EINVAL_VERTEX, "range out of bounds"};
        if (!count) return;
        std::vector<uint64_t> tmp(count);
        uint64_t *dv = dalloc<uint64_t>(count);
        launch_read_labels(L->vr + first, (uint32_t)count, dv, g->st);
        HC(hipGetLastError());
        HC(hipMemcpyAsync(tmp.data(), dv, count * 8, hipMemcpyDeviceToHost, g->st));
        HC(hipStreamSynchronize(g->st));
        dfree(dv);
        for (uint64_t i = 0; i < count; i++) {
            uint32_t p = (uint32_t)tmp[i];
            if (values) values[i] = tmp[i] >> 32;
            if (parents) parents[i] = p == NILP ? g->V : p;
        }
    });
}

int rg_modified_read(rg_labels *L, uint64_t *vids, uint64_t cap, uint64_t *n, int *all)
{
    if (!L || !n) return RG_EINVAL;
    return guard([&] {
        rg_graph *g = L->g;
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        uint64_t cnt = std::min<uint64_t>(L->hc->n_mod, g->V);
        if (all) *all = (L->hc->flags & F_ALL_MODIFIED) ? 1 : 0;
        uint64_t k = std::min(cnt, cap);
        if (k && vids) {
            std::vector<uint32_t> tmp(k);
            HC(hipMemcpyAsync(tmp.data(), L->mlist, k * 4, hipMemcpyDeviceToHost, g->st));
            HC(hipStreamSynchronize(g->st));
            for (uint64_t i = 0; i < k; i++) vids[i] = tmp[i];
        }
        *n = vids ? k : cnt;
    });
}

int rg_modified_clear(rg_labels *L)
{
    if (!L) return RG_EINVAL;
    return guard([&] {
        rg_graph *g = L->g;
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        L->hc->n_mod = 0;
        L->hc->mepoch += 1;
        L->hc->flags &= ~F_ALL_MODIFIED;
        HC(hipMemcpyAsync(L->ctl, L->hc, sizeof(Ctl), hipMemcpyHostToDevice, g->st));
        HC(hipStreamSynchronize(g->st));
    });
}

int rg_stream_create(int kind, uint64_t scale_or_v, uint64_t n_edges, uint64_t seed, double a, double b,
                     double c, uint64_t wseed, rg_stream **out)
{
    if (!out || (kind != 0 && kind != 1)) return RG_EINVAL;
    *out = nullptr;
    rg_stream *s = new rg_stream();
    int rc = guard([&] {
        if (kind == 0 && (scale_or_v < 1 || scale_or_v > 31)) throw RgError{RG_EINVAL, "R-MAT scale must be 1..31"};
        if (kind == 1 && (scale_or_v < 1 || scale_or_v > MAX_V)) throw RgError{RG_EINVAL, "vertex count out of range"};
        HC(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
        s->n = n_edges;
        s->src = dalloc<uint64_t>(n_edges);
        s->dst = dalloc<uint64_t>(n_edges);
        if (kind == 0) launch_gen_rmat(s->src, s->dst, n_edges, 0, (int)scale_or_v, seed, a, b, c, s->st);
        else launch_gen_er(s->src, s->dst, n_edges, 0, scale_or_v, seed, s->st);
        if (wseed) {
            s->w = dalloc<uint32_t>(n_edges);
            launch_gen_weights(s->w, n_edges, 0, wseed, s->st);
        }
        HC(hipGetLastError());
        HC(hipStreamSynchronize(s->st));
    });
    if (rc != RG_OK) {
        rg_stream_destroy(s);
        return rc;
    }
    *out = s;
    return RG_OK;
}

int rg_stream_destroy(rg_stream *s)
{
    if (!s) return RG_OK;
    if (s->st) (void)hipStreamSynchronize(s->st);
    dfree(s->src); dfree(s->dst); dfree(s->w);
    if (s->st) (void)hipStreamDestroy(s->st);
    delete s;
    return RG_OK;
}

int rg_stream_read(rg_stream *s, uint64_t first, uint64_t n, uint64_t *src, uint64_t *dst, uint32_t *w)
{
    if (!s) return RG_EINVAL;
    return guard([&] {
        if (first + n > s->n) throw RgError{RG_EINVAL, "range out of bounds"};
        if (src) HC(hipMemcpy(src, s->src + first, n * 8, hipMemcpyDeviceToHost));
        if (dst) HC(hipMemcpy(dst, s->dst + first, n * 8, hipMemcpyDeviceToHost));
        if (w) {
            if (!s->w) throw RgError{RG_EINVAL, "stream has no weights"};
            HC(hipMemcpy(w, s->w + first, n * 4, hipMemcpyDeviceToHost));
        }
    });
}

int rg_stream_max_degree_vertex(rg_stream *s, uint64_t n, uint64_t V, uint64_t *out)
{
    if (!s || !out) return RG_EINVAL;
    return guard([&] {
        if (n > s->n) throw RgError{RG_EINVAL, "range out of bounds"};
        uint32_t *deg = dalloc<uint32_t>(V);
        unsigned long long *best = dalloc<unsigned long long>(1);
        launch_max_degree(s->src, n, V, deg, best, s->st);
        unsigned long long b = 0;
        hipError_t e = hipMemcpyAsync(&b, best, 8, hipMemcpyDeviceToHost, s->st);
        if (e == hipSuccess) e = hipStreamSynchronize(s->st);
        dfree(deg);
        dfree(best);
        hipchk(e, "max degree");
        *out = 0xFFFFFFFFull - (b & 0xFFFFFFFFull);
    });
}

int rg_stream_load(rg_graph *g, rg_stream *s, uint64_t first, uint64_t n, int directed)
{
    if (!g || !s) return RG_EINVAL;
    return guard([&] {
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        if (first + n > s->n) throw RgError{RG_EINVAL, "range out of bounds"};
        HC(hipStreamSynchronize(s->st));
        const uint64_t B = 1ull << 24;
        uint8_t kind = K_INS | (directed ? 0 : K_UNDIR);
        const uint32_t *w = g->weighted ? s->w : nullptr;
        if (g->weighted && !w) throw RgError{RG_EINVAL, "weighted graph needs a weighted stream"};
        // size every list (and hub index) once for the whole load
        for (int L = 0; L < (g->symmetric ? 1 : 2); L++) {
            if (!g->cnt_tmp[L]) {
                captured_clear(g);   // captured store launches hold the buffer set
                g->cnt_tmp[L] = dalloc<uint32_t>(g->V);
                HC(hipMemsetAsync(g->cnt_tmp[L], 0, g->V * 4, g->st));
            }
        }
        for (uint64_t b = 0; b < n; b += B) {
            uint32_t m = (uint32_t)std::min(B, n - b);
            ensure_ops(g, m);
            launch_ops_from_stream(s->src, s->dst, w, first + b, m, kind, g->ops, 0, g->st);
            g->ops.n = m;
            launch_store_count(store_args(g), g->st);
        }
        launch_store_reserve(store_args(g), g->st);
        launch_pool_recycle(store_args(g), g->st);
        HC(hipGetLastError());
        for (uint64_t b = 0; b < n; b += B) {
            uint32_t m = (uint32_t)std::min(B, n - b);
            ensure_ops(g, m);
            launch_ops_from_stream(s->src, s->dst, w, first + b, m, kind, g->ops, 0, g->st);
            g->ops.n = m;
            g->lops.n = 4 * m;
            g->lstride = 4;
            run_store(g);
        }
        HC(hipStreamSynchronize(g->st));
        store_status_check(g);
    });
}

// What the safe/unsafe split of one batch reports to the serving loop.
struct SplitOut {
    bool flags = false;         // copy the per-op unsafe flags to the host
    bool done = false;          // the batch was split
    std::chrono::steady_clock::time_point t_safe{};   // the safe ops completed (host clock)
    uint64_t unsafe = 0;        // unsafe ops of the batch
    const uint8_t *flag = nullptr;   // pinned, one per op (ops [0, n_ins) inserts, then deletes)
    // called once the flags are on the host, while the repair still runs
    std::function<void(const uint8_t *flag)> on_safe;
};

// batches of more than this many edge ops are split into safe and unsafe
// ops before the repair (smaller ones run the repair on all their ops)
constexpr uint64_t SPLIT_MIN_OPS = STORE_SMALL_OPS;

// The repair of a store batch already enqueued on the graph stream (the
// batch's n edge ops in g->ops, bracketed by ev[0] / ev[1]): the split of the
// ops into safe and unsafe ones (when `split`), the engine run of every tree
// on the unsafe ops, the store status, the version clock and the histories.
static bool capture_on(const rg_graph *g)
{
    if (g->capture >= 0) return g->capture != 0;
    static const bool env = [] {
        const char *e = getenv("RG_GRAPH");
        return !(e && *e == '0');
    }();
    return env;
}

// The captured window -> store -> split pipeline for a batch-size bucket:
// the descriptor copy (pinned -> device), the window's edge ops, the store
// kernels and the pool recycle, the split and its count publication, every
// kernel sized for `bucket` edge ops and taking the batch's size from the
// device descriptor.  Captured once per (bucket, trees, stream, kinds) and
// replayed; null when capture is unavailable (the caller runs the launches).
static hipGraphExec_t captured_pipeline(rg_graph *g, rg_labels **L, int nL, rg_stream *s, uint32_t bucket,
                                        uint8_t kind_ins, uint8_t kind_del, bool flags)
{
    const uint32_t *w = g->weighted ? s->w : nullptr;
    StoreArgs a = store_args(g);
    a.dyn_n = &g->d_desc->n;
    SplitParams P{};
    for (int i = 0; i < nL; i++) {
        const EngineArgs e = engine_args(L[i]);
        P.vr[i] = e.vr;
        P.algo[i] = e.algo;
        P.undirected[i] = e.undirected;
        P.hctl[i] = L[i]->hc;
    }
    P.ntrees = nL;
    P.in = g->ops;
    P.out = g->uops;
    P.unsafe = flags ? g->h_flag : nullptr;
    P.count = g->ucount;
    P.dyn_n = &g->d_desc->n;
    P.in.n = P.out.n = 0;   // the count comes from the descriptor
    auto same_split = [nL](const SplitParams &x, const SplitParams &y) {
        bool eq = x.ntrees == y.ntrees && x.unsafe == y.unsafe && x.count == y.count && x.dyn_n == y.dyn_n;
        for (int i = 0; eq && i < nL; i++)
            eq = x.vr[i] == y.vr[i] && x.algo[i] == y.algo[i] && x.undirected[i] == y.undirected[i] &&
                 x.hctl[i] == y.hctl[i];
        for (const EdgeOps *e : {&x.in, &x.out}) {
            const EdgeOps &f = e == &x.in ? y.in : y.out;
            eq = eq && e->src == f.src && e->dst == f.dst && e->w == f.w && e->kind == f.kind && e->res == f.res;
        }
        return eq;
    };
    for (const auto &c : g->captured)
        if (c.bucket == bucket && same_split(c.P, P) && c.src == s->src && c.dst == s->dst && c.w == w &&
            c.len == s->n && c.kind_ins == kind_ins && c.kind_del == kind_del)
            return c.exec;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (hipStreamBeginCapture(g->st, hipStreamCaptureModeRelaxed) != hipSuccess) {
        g->capture = 0;
        return nullptr;
    }
    (void)hipMemcpyAsync(g->d_desc, g->h_desc, sizeof(BatchDesc), hipMemcpyHostToDevice, g->st);
    launch_ops_from_window_dyn(s->src, s->dst, w, s->n, g->d_desc, bucket, kind_ins, kind_del, g->ops, g->ucount,
                               g->st);
    launch_store_batch_dyn(a, bucket, g->st);
    launch_pool_recycle(a, g->st);
    launch_split(P, bucket, g->st);
    hipError_t e = hipStreamEndCapture(g->st, &graph);
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    if (e != hipSuccess || !exec) {
        // capture is not available here: every later batch launches eagerly
        g->capture = 0;
        return nullptr;
    }
    rg_graph::Captured c;
    c.bucket = bucket;
    c.P = P;
    c.src = s->src;
    c.dst = s->dst;
    c.w = w;
    c.len = s->n;
    c.kind_ins = kind_ins;
    c.kind_del = kind_del;
    c.exec = exec;
    g->captured.push_back(c);
    return exec;
}

// `pipeline` (may be null): a captured window -> store -> split pipeline
// that launches in place of the split (the store is then not yet enqueued).
static void repair_batch(rg_graph *g, rg_labels **L, int nL, uint64_t n, bool has_ins, bool has_del,
                         rg_batch_stats *stats, bool split, SplitOut *so,
                         const std::function<void()> *pipeline = nullptr)
{
    const bool flags = split && so && so->flags;
    uint32_t ph = has_del ? PH_DEL_ROOTS : PH_SEED;
    RunHooks hooks;
    if (split) {
        SplitParams P{};
        for (int i = 0; i < nL; i++) {
            const EngineArgs a = engine_args(L[i]);
            P.vr[i] = a.vr;
            P.algo[i] = a.algo;
            P.undirected[i] = a.undirected;
            P.hctl[i] = L[i]->hc;
        }
        P.ntrees = nL;
        P.in = g->ops;
        P.out = g->uops;
        // the split writes its flags straight into pinned host memory: a copy
        // on another stream would wait for CUs behind the repair's
        // persistent launch, and the safe updates with it
        P.unsafe = flags ? g->h_flag : nullptr;
        P.count = g->ucount;
        hooks.before_launch = [g, P, n, flags, pipeline] {
            if (pipeline) (*pipeline)();
            else launch_split(P, (uint32_t)n, g->st);
            HC(hipGetLastError());
            if (flags) HC(hipEventRecord(g->ev_safe, g->st));
        };
        if (flags)
            hooks.after_first_enqueue = [g, so] {
                HC(hipEventSynchronize(g->ev_safe));
                so->t_safe = std::chrono::steady_clock::now();
                if (so->on_safe) so->on_safe(g->h_flag);
            };
    }
    engine_run_multi(L, nL, ph, has_ins && has_del, (uint32_t)n, has_del, stats, split, &hooks);
    HC(hipEventSynchronize(g->ev[1]));
    store_status_check(g);
    g->version += 1;
    for (int i = 0; i < nL; i++)
        if (L[i]->hist.on) hist_commit(L[i], g->version);
    if (split) {
        // each tree's control block kept the compacted op count as n_ops
        g->last_unsafe = L[0]->hc->n_ops;
        if (so) {
            so->done = flags;
            so->unsafe = g->last_unsafe;
            so->flag = flags ? g->h_flag : nullptr;
        }
    } else {
        g->last_unsafe = n;
    }
    // device time of the batch: store start to the end of the engine run
    float ms = 0;
    if (nL > 0) {
        HC(hipEventSynchronize(g->ev[3]));
        HC(hipEventElapsedTime(&ms, g->ev[0], g->ev[3]));
    } else {
        HC(hipEventElapsedTime(&ms, g->ev[0], g->ev[1]));
    }
    g->last_ms = ms;
    if (g->prof) prof_collect(g);
    // RG_LOG_SLOW=ms: report batches whose device time exceeded ms
    static const double log_slow = [] {
        const char *e = getenv("RG_LOG_SLOW");
        return e ? atof(e) : 0.0;
    }();
    if (log_slow > 0 && ms > log_slow) {
        float store_ms = 0;
        (void)hipEventElapsedTime(&store_ms, g->ev[0], g->ev[1]);
        fprintf(stderr, "[rg] slow batch: %.2f ms (store %.2f), %llu edge ops, %llu unsafe, grid steps %llu;", ms,
                store_ms, (unsigned long long)n, (unsigned long long)g->last_unsafe,
                (unsigned long long)(stats ? stats->grid_steps : 0));
        for (int i = 0; i < nL; i++)
            fprintf(stderr, " algo %d levels %llu subtree %u dense %u;", L[i]->algo,
                    (unsigned long long)L[i]->hc->levels, L[i]->hc->n_tree, L[i]->hc->dense);
        fprintf(stderr, "\n");
    }
}

static void apply_batch(rg_graph *g, rg_labels **L, int nL, rg_stream *s, uint64_t ins_first, uint64_t n_ins,
                        uint64_t del_first, uint64_t n_del, int directed, rg_batch_stats *stats,
                        SplitOut *so = nullptr)
{
    // ranges wrap around the end of the stream (a sliding window over a
    // circular stream: update k inserts e[(n0 + k) mod n], deletes e[k mod n])
    if (ins_first >= s->n || del_first >= s->n || n_ins > s->n || n_del > s->n)
        throw RgError{RG_EINVAL, "range out of bounds"};
    for (int i = 0; i < nL; i++)
        if (!L[i] || L[i]->g != g) throw RgError{RG_EINVAL, "labels belong to another graph"};
    uint64_t n = n_ins + n_del;
    if (so) so->done = false;
    if (!n) return;
    const bool split = nL > 0 && nL <= MAXT && n > SPLIT_MIN_OPS;
    const bool flags = split && so && so->flags;
    uint8_t ud = directed ? 0 : K_UNDIR;
    // every op of a stream batch has the same direction: its list ops fit
    // this many slots (half the records to group for a directed batch)
    g->lstride = (directed ? 1u : 2u) * (g->symmetric ? 1u : 2u);
    const uint32_t *w = g->weighted ? s->w : nullptr;
    if (g->weighted && !w) throw RgError{RG_EINVAL, "weighted graph needs a weighted stream"};
    // a split batch whose store kernels are not timed one by one runs the
    // window -> store -> split part as one captured graph of its size bucket
    if (split && capture_on(g) && !(g->prof & 1)) {
        uint32_t bucket = 128;
        while (bucket < n) bucket <<= 1;
        ensure_ops(g, bucket);
        if (flags) ensure_host(g, bucket);
        g->ops.n = (uint32_t)n;
        g->lops.n = (uint32_t)(g->lstride * n);
        hipGraphExec_t exec = captured_pipeline(g, L, nL, s, bucket, (uint8_t)(K_INS | ud), (uint8_t)(K_DEL | ud),
                                                flags);
        if (exec) {
            const std::function<void()> pipe = [&] {
                *g->h_desc = BatchDesc{ins_first, del_first, (uint32_t)n_ins, (uint32_t)n_del, (uint32_t)n, 0};
                HC(hipEventRecord(g->ev[0], g->st));
                HC(hipGraphLaunch(exec, g->st));
                g->replays++;
                HC(hipEventRecord(g->ev[1], g->st));
            };
            repair_batch(g, L, nL, n, n_ins > 0, n_del > 0, stats, split, so, &pipe);
            if (stats) {
                stats->seeds += n_ins;
                stats->roots += n_del;
            }
            return;
        }
    }
    ensure_ops(g, n);
    if (flags) ensure_host(g, n);
    launch_ops_from_window(s->src, s->dst, w, s->n, ins_first, (uint32_t)n_ins, del_first, (uint32_t)n_del,
                           (uint8_t)(K_INS | ud), (uint8_t)(K_DEL | ud), g->ops, split ? g->ucount : nullptr, g->st);
    g->ops.n = (uint32_t)n;
    g->lops.n = (uint32_t)(g->lstride * n);
    HC(hipEventRecord(g->ev[0], g->st));
    run_store(g);
    HC(hipEventRecord(g->ev[1], g->st));
    repair_batch(g, L, nL, n, n_ins > 0, n_del > 0, stats, split, so);
    if (stats) {
        stats->seeds += n_ins;
        stats->roots += n_del;
    }
}

}  // extern "C"

namespace {

// ---------------------------------------------------------------------
// Write-ahead log of transaction batches (the WAL of all_inc_rpc.cpp:305-336,
// 572-590: 32-byte entries {type, client, src, dst, data}, group-committed
// before the responses go out).  File: a 32-byte header {magic "RGWAL001",
// V, weighted, symmetric}, then one record per batch: a 24-byte header
// {magic, count, version the batch produces, FNV-1a checksum of the
// entries} and `count` 32-byte entries {type (bit 0 delete, bit 1
// undirected), transaction, src, dst, data}.  A batch is written and
// (with sync) fdatasync'ed before it is applied; replay stops at the first
// record that is cut short or fails its checksum (a batch that was never
// acknowledged).
// ---------------------------------------------------------------------
constexpr char WAL_MAGIC[8] = {'R', 'G', 'W', 'A', 'L', '0', '0', '1'};
constexpr uint32_t WAL_BATCH_MAGIC = 0x31424752u;   // "RGB1"
struct WalFileHdr {
    char magic[8];
    uint64_t V;
    uint32_t weighted, symmetric;
    uint64_t pad;
};
struct WalBatchHdr {
    uint32_t magic, count;
    uint64_t version;
    uint64_t checksum;
};
struct WalEntry {
    uint32_t type, txn;
    uint64_t src, dst, data;
};
static_assert(sizeof(WalFileHdr) == 32 && sizeof(WalBatchHdr) == 24 && sizeof(WalEntry) == 32, "WAL layout");

uint64_t fnv1a(const void *p, size_t n)
{
    const unsigned char *b = (const unsigned char *)p;
    uint64_t h = 0xcbf29ce484222325ull;
    for (size_t i = 0; i < n; i++) h = (h ^ b[i]) * 0x100000001b3ull;
    return h;
}

void write_all(int fd, const void *p, size_t n, const char *what)
{
    const char *c = (const char *)p;
    while (n) {
        ssize_t k = write(fd, c, n);
        if (k < 0) {
            if (errno == EINTR) continue;
            throw RgError{RG_EDEVICE, std::string(what) + ": " + strerror(errno)};
        }
        c += k;
        n -= (size_t)k;
    }
}

bool read_all(int fd, void *p, size_t n)
{
    char *c = (char *)p;
    while (n) {
        ssize_t k = read(fd, c, n);
        if (k < 0 && errno == EINTR) continue;
        if (k <= 0) return false;
        c += k;
        n -= (size_t)k;
    }
    return true;
}

void wal_append(rg_graph *g, const rg_edge *e, const int8_t *kind, const uint32_t *txn_of, uint64_t n, int directed,
                uint64_t version)
{
    std::vector<WalEntry> buf(n);
    for (uint64_t i = 0; i < n; i++)
        buf[i] = WalEntry{(kind[i] < 0 ? 1u : 0u) | (directed ? 0u : 2u), txn_of ? txn_of[i] : (uint32_t)i, e[i].src,
                          e[i].dst, e[i].data};
    WalBatchHdr h{WAL_BATCH_MAGIC, (uint32_t)n, version, fnv1a(buf.data(), n * sizeof(WalEntry))};
    std::vector<char> rec(sizeof(h) + n * sizeof(WalEntry));
    memcpy(rec.data(), &h, sizeof(h));
    memcpy(rec.data() + sizeof(h), buf.data(), n * sizeof(WalEntry));
    write_all(g->wal_fd, rec.data(), rec.size(), "WAL write");
    if (g->wal_sync && fdatasync(g->wal_fd) != 0) throw RgError{RG_EDEVICE, std::string("WAL fdatasync: ") + strerror(errno)};
}

// One device batch of mixed host-buffer updates: logged first when a WAL is
// open (and `log`), then the store batch and the repair of every tree, split
// into safe and unsafe updates (so->flag[i] for update i).
void ops_batch(rg_graph *g, rg_labels **L, int nL, const rg_edge *e, const int8_t *kind, const uint32_t *txn_of,
               uint64_t n, int directed, SplitOut *so, bool log)
{
    if (log && g->wal_fd >= 0) wal_append(g, e, kind, txn_of, n, directed, g->version + 1);
    upload_edges(g, e, n, 0);
    const uint8_t ud = directed ? 0 : K_UNDIR;
    bool has_ins = false, has_del = false;
    for (uint64_t i = 0; i < n; i++) {
        g->h_kind[i] = (uint8_t)((kind[i] > 0 ? K_INS : K_DEL) | ud);
        has_ins |= kind[i] > 0;
        has_del |= kind[i] < 0;
    }
    HC(hipMemcpyAsync(g->ops.kind, g->h_kind, n, hipMemcpyHostToDevice, g->st));
    HC(hipMemsetAsync(g->ucount, 0, 4, g->st));
    HC(hipEventRecord(g->ev[0], g->st));
    run_store(g);
    HC(hipEventRecord(g->ev[1], g->st));
    const bool split = nL > 0 && nL <= MAXT;
    repair_batch(g, L, nL, n, has_ins, has_del, nullptr, split, split ? so : nullptr);
}

}  // namespace

extern "C" {

int rg_txn_apply(rg_graph *g, rg_labels **L, int nL, const rg_edge *edges, const int8_t *kind, uint64_t n,
                 const uint32_t *txn_len, uint64_t ntxn, int directed, uint8_t *txn_unsafe, uint64_t *txn_version)
{
    if (!g || (nL && !L) || nL < 1 || nL > MAXT || (n && (!edges || !kind)) || (ntxn && !txn_len)) return RG_EINVAL;
    return guard([&] {
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        for (int i = 0; i < nL; i++)
            if (!L[i] || L[i]->g != g) throw RgError{RG_EINVAL, "labels belong to another graph"};
        uint64_t tot = 0;
        for (uint64_t t = 0; t < ntxn; t++) {
            if (txn_len[t] < 1 || txn_len[t] > RG_TXN_MAX)
                throw RgError{RG_EINVAL, "a transaction holds 1 to 16 updates"};
            tot += txn_len[t];
        }
        if (tot != n) throw RgError{RG_EINVAL, "transaction lengths do not add up to the update count"};
        for (uint64_t i = 0; i < n; i++)
            if (kind[i] != 1 && kind[i] != -1) throw RgError{RG_EINVAL, "kind must be +1 (add) or -1 (del)"};
        const uint64_t v0 = g->version;
        if (n == 0) return;
        std::vector<uint32_t> txn_of(n);
        for (uint64_t t = 0, k = 0; t < ntxn; t++)
            for (uint32_t j = 0; j < txn_len[t]; j++) txn_of[k++] = (uint32_t)t;
        SplitOut so;
        so.flags = true;
        ops_batch(g, L, nL, edges, kind, txn_of.data(), n, directed, &so, true);
        uint64_t k = 0;
        for (uint64_t t = 0; t < ntxn; t++) {
            bool u = false;
            for (uint32_t j = 0; j < txn_len[t]; j++, k++) u = u || so.flag[k];
            if (txn_unsafe) txn_unsafe[t] = u ? 1 : 0;
            if (txn_version) txn_version[t] = u ? g->version : v0;
        }
    });
}

}  // extern "C"

// One batch of host-buffer updates for the request server (host_batch.h).
int rg::host_batch(rg_graph *g, rg_labels **L, int nL, const rg_edge *e, const int8_t *kind, const uint32_t *txn_of,
                   uint64_t n, int directed, uint8_t *flags, const std::function<void(const uint8_t *)> &on_safe,
                   uint64_t *v_before, uint64_t *v_after, std::string *err)
{
    try {
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        *v_before = *v_after = g->version;
        if (!n) return RG_OK;
        SplitOut so;
        so.flags = true;
        so.on_safe = on_safe;
        ops_batch(g, L, nL, e, kind, txn_of, n, directed, &so, true);
        if (flags) {
            if (so.flag) memcpy(flags, so.flag, n);
            else memset(flags, 1, n);   // no trees to split for: everything waited for the batch
        }
        *v_after = g->version;
        return RG_OK;
    } catch (const RgError &x) {
        if (err) *err = x.msg;
        return x.code;
    } catch (const std::exception &x) {
        if (err) *err = x.what();
        return RG_EDEVICE;
    }
}

extern "C" {

int rg_wal_open(rg_graph *g, const char *path, int sync)
{
    if (!g || !path) return RG_EINVAL;
    return guard([&] {
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        if (g->wal_fd >= 0) throw RgError{RG_EINVAL, "a WAL is already open on this graph"};
        int fd = open(path, O_RDWR | O_CREAT | O_APPEND, 0644);
        if (fd < 0) throw RgError{RG_EINVAL, std::string("WAL open: ") + strerror(errno)};
        off_t size = lseek(fd, 0, SEEK_END);
        if (size == 0) {
            WalFileHdr h{};
            memcpy(h.magic, WAL_MAGIC, 8);
            h.V = g->V;
            h.weighted = (uint32_t)g->weighted;
            h.symmetric = (uint32_t)g->symmetric;
            try {
                write_all(fd, &h, sizeof(h), "WAL header");
            } catch (...) {
                close(fd);
                throw;
            }
            if (sync) (void)fsync(fd);
        } else {
            WalFileHdr h{};
            if (pread(fd, &h, sizeof(h), 0) != (ssize_t)sizeof(h) || memcmp(h.magic, WAL_MAGIC, 8) != 0 ||
                h.V != g->V || h.weighted != (uint32_t)g->weighted) {
                close(fd);
                throw RgError{RG_EINVAL, "WAL file belongs to another graph"};
            }
        }
        g->wal_fd = fd;
        g->wal_sync = sync ? 1 : 0;
    });
}

int rg_wal_close(rg_graph *g)
{
    if (!g) return RG_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(g->mu);
    if (g->wal_fd >= 0) {
        (void)fdatasync(g->wal_fd);
        (void)close(g->wal_fd);
    }
    g->wal_fd = -1;
    return RG_OK;
}

int rg_wal_replay(rg_graph *g, rg_labels **L, int nL, const char *path, uint64_t *batches)
{
    if (!g || !path || (nL && !L) || nL < 0 || nL > MAXT) return RG_EINVAL;
    return guard([&] {
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        if (batches) *batches = 0;
        int fd = open(path, O_RDONLY);
        if (fd < 0) throw RgError{RG_EINVAL, std::string("WAL open: ") + strerror(errno)};
        struct Closer {
            int fd;
            ~Closer() { close(fd); }
        } closer{fd};
        WalFileHdr fh{};
        if (!read_all(fd, &fh, sizeof(fh)) || memcmp(fh.magic, WAL_MAGIC, 8) != 0)
            throw RgError{RG_EINVAL, "not a WAL file"};
        if (fh.V != g->V || fh.weighted != (uint32_t)g->weighted)
            throw RgError{RG_EINVAL, "WAL file belongs to another graph"};
        std::vector<WalEntry> ent;
        std::vector<rg_edge> e;
        std::vector<int8_t> kind;
        for (;;) {
            WalBatchHdr h{};
            if (!read_all(fd, &h, sizeof(h)) || h.magic != WAL_BATCH_MAGIC) break;   // end or torn tail
            ent.resize(h.count);
            if (!read_all(fd, ent.data(), h.count * sizeof(WalEntry))) break;
            if (fnv1a(ent.data(), h.count * sizeof(WalEntry)) != h.checksum) break;
            if (h.version <= g->version) continue;   // already in the snapshot
            if (h.version != g->version + 1) throw RgError{RG_EINVAL, "WAL has a gap after the graph's version"};
            // a logged batch has one direction for all its entries
            const int directed = h.count && (ent[0].type & 2u) ? 0 : 1;
            e.resize(h.count);
            kind.resize(h.count);
            for (uint32_t i = 0; i < h.count; i++) {
                e[i] = rg_edge{ent[i].src, ent[i].dst, ent[i].data};
                kind[i] = (ent[i].type & 1u) ? -1 : 1;
            }
            if (h.count) ops_batch(g, L, nL, e.data(), kind.data(), nullptr, h.count, directed, nullptr, false);
            else g->version += 1;
            if (batches) (*batches)++;
        }
    });
}

int rg_graph_capture(rg_graph *g, int on)
{
    if (!g) return RG_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(g->mu);
    g->capture = on ? 1 : 0;
    return RG_OK;
}

int rg_graph_capture_info(rg_graph *g, int *on, uint32_t *pipelines, uint64_t *replays)
{
    if (!g) return RG_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(g->mu);
    if (on) *on = capture_on(g) ? 1 : 0;
    if (pipelines) *pipelines = (uint32_t)g->captured.size();
    if (replays) *replays = g->replays;
    return RG_OK;
}

int rg_graph_shape(rg_graph *g, int *symmetric, int *dual, int *weighted)
{
    if (!g) return RG_EINVAL;
    if (symmetric) *symmetric = g->symmetric;
    if (dual) *dual = g->dual;
    if (weighted) *weighted = g->weighted;
    return RG_OK;
}

int rg_labels_algo(rg_labels *L, int *algo, uint64_t *root)
{
    if (!L) return RG_EINVAL;
    if (algo) *algo = L->algo;
    if (root) *root = L->root;
    return RG_OK;
}

int rg_stream_from_edges(const rg_edge *edges, uint64_t n, int weighted, rg_stream **out)
{
    if (!out || (n && !edges)) return RG_EINVAL;
    *out = nullptr;
    rg_stream *s = new rg_stream();
    int rc = guard([&] {
        HC(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
        s->n = n;
        s->src = dalloc<uint64_t>(n);
        s->dst = dalloc<uint64_t>(n);
        if (weighted) s->w = dalloc<uint32_t>(n);
        const uint64_t B = 1ull << 22;
        std::vector<uint64_t> a(std::min(B, std::max<uint64_t>(n, 1))), b(a.size());
        std::vector<uint32_t> c(weighted ? a.size() : 0);
        for (uint64_t f = 0; f < n; f += B) {
            const uint64_t m = std::min(B, n - f);
            for (uint64_t i = 0; i < m; i++) {
                const rg_edge &e = edges[f + i];
                if (e.src > MAX_V || e.dst > MAX_V) throw RgError{RG_EINVAL_VERTEX, "VertexId error."};
                if (weighted && e.data > 0xFFFFFFFFull) throw RgError{RG_EINVAL, "edge weight exceeds 32 bits"};
                a[i] = e.src;
                b[i] = e.dst;
                if (weighted) c[i] = (uint32_t)e.data;
            }
            HC(hipMemcpy(s->src + f, a.data(), m * 8, hipMemcpyHostToDevice));
            HC(hipMemcpy(s->dst + f, b.data(), m * 8, hipMemcpyHostToDevice));
            if (weighted) HC(hipMemcpy(s->w + f, c.data(), m * 4, hipMemcpyHostToDevice));
        }
    });
    if (rc != RG_OK) {
        rg_stream_destroy(s);
        return rc;
    }
    *out = s;
    return RG_OK;
}

int rg_labels_assign(rg_labels *L, const uint64_t *values, const uint64_t *parents)
{
    if (!L || !values || !parents) return RG_EINVAL;
    return guard([&] {
        rg_graph *g = L->g;
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        const uint64_t V = g->V;
        std::vector<uint64_t> w(V);
        for (uint64_t v = 0; v < V; v++) {
            if (values[v] > 0xFFFFFFFFull) throw RgError{RG_EINVAL, "label exceeds 32 bits"};
            if (parents[v] > V) throw RgError{RG_EINVAL_VERTEX, "parent out of range"};
            w[v] = (values[v] << 32) | (parents[v] == V ? (uint64_t)NILP : parents[v]);
        }
        uint64_t *dw = dalloc<uint64_t>(V);
        hipError_t e = hipMemcpyAsync(dw, w.data(), V * 8, hipMemcpyHostToDevice, g->st);
        if (e == hipSuccess) {
            launch_write_labels(L->vr, (uint32_t)V, dw, g->st);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(g->st);
        dfree(dw);
        hipchk(e, "labels assign");
        // every vertex may have changed
        L->hc->flags |= F_ALL_MODIFIED;
        HC(hipMemcpyAsync(L->ctl, L->hc, sizeof(Ctl), hipMemcpyHostToDevice, g->st));
        HC(hipStreamSynchronize(g->st));
    });
}

// ---------------------------------------------------------------------
// Snapshots.  File: header {magic "RGSNAP01", V, hint, symmetric, dual,
// weighted, labels count, version, edge count}; the live multiset of the
// out-lists as rg_edge records (one per copy, so reloading them as directed
// inserts rebuilds every list and count); then per labels instance {algo,
// trace, root, WCC init labels flag} + [init labels] + values + parents.
// ---------------------------------------------------------------------
}  // extern "C"

namespace {

constexpr char SNAP_MAGIC[8] = {'R', 'G', 'S', 'N', 'A', 'P', '0', '1'};
struct SnapHdr {
    char magic[8];
    uint64_t V, hint;
    uint32_t symmetric, dual, weighted, nL;
    uint64_t version, n_edges;
};
struct SnapLabelsHdr {
    uint32_t algo, trace;
    uint64_t root;
    uint32_t has_init, pad;
};

struct FileCloser {
    FILE *f;
    ~FileCloser()
    {
        if (f) fclose(f);
    }
};

void fwrite_all(FILE *f, const void *p, size_t n)
{
    if (n && fwrite(p, 1, n, f) != n) throw RgError{RG_EDEVICE, std::string("snapshot write: ") + strerror(errno)};
}

void fread_all(FILE *f, void *p, size_t n)
{
    if (n && fread(p, 1, n, f) != n) throw RgError{RG_EINVAL, "snapshot file is cut short"};
}

}  // namespace

extern "C" {

int rg_snapshot_save(rg_graph *g, rg_labels **L, int nL, const char *path)
{
    if (!g || !path || (nL && !L) || nL < 0) return RG_EINVAL;
    return guard([&] {
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        for (int i = 0; i < nL; i++)
            if (!L[i] || L[i]->g != g) throw RgError{RG_EINVAL, "labels belong to another graph"};
        const uint64_t V = g->V;
        // per-vertex live degrees of the out-lists -> record offsets
        std::vector<uint32_t> deg(V);
        copy_degrees(g->lists[0], deg.data(), V, g->st);
        std::vector<uint64_t> off(V + 1, 0);
        for (uint64_t v = 0; v < V; v++) off[v + 1] = off[v] + deg[v];
        const std::string tmp = std::string(path) + ".tmp";
        FILE *f = fopen(tmp.c_str(), "wb");
        if (!f) throw RgError{RG_EINVAL, std::string("snapshot open: ") + strerror(errno)};
        FileCloser fc{f};
        SnapHdr h{};
        memcpy(h.magic, SNAP_MAGIC, 8);
        h.V = V;
        h.hint = g->hint;
        h.symmetric = (uint32_t)g->symmetric;
        h.dual = (uint32_t)g->dual;
        h.weighted = (uint32_t)g->weighted;
        h.nL = (uint32_t)nL;
        h.version = g->version;
        h.n_edges = off[V];
        fwrite_all(f, &h, sizeof(h));
        // the edges in vertex ranges of at most ~2^24 records
        const uint64_t CH = 1ull << 24;
        uint64_t *doff = dalloc<uint64_t>(V + 1);
        rg_edge_rec *drec = dalloc<rg_edge_rec>(std::max<uint64_t>(1, std::min(CH + 1, off[V])));
        std::vector<rg_edge_rec> hrec;
        try {
            HC(hipMemcpyAsync(doff, off.data(), (V + 1) * 8, hipMemcpyHostToDevice, g->st));
            uint64_t v0 = 0;
            while (v0 < V) {
                uint64_t v1 = v0 + 1;
                while (v1 < V && off[v1 + 1] - off[v0] <= CH) v1++;
                const uint64_t m = off[v1] - off[v0];
                if (m > CH + 1) {
                    // one vertex above the chunk: a buffer of its own size
                    dfree(drec);
                    drec = dalloc<rg_edge_rec>(m);
                }
                launch_list_dump(g->lists[0], g->slots, (uint32_t)v0, (uint32_t)v1, doff + v0, drec, g->st);
                HC(hipGetLastError());
                hrec.resize(m);
                if (m) HC(hipMemcpyAsync(hrec.data(), drec, m * sizeof(rg_edge_rec), hipMemcpyDeviceToHost, g->st));
                HC(hipStreamSynchronize(g->st));
                fwrite_all(f, hrec.data(), m * sizeof(rg_edge_rec));
                v0 = v1;
            }
        } catch (...) {
            dfree(doff);
            dfree(drec);
            throw;
        }
        dfree(doff);
        dfree(drec);
        std::vector<uint64_t> val(V), par(V), init;
        for (int i = 0; i < nL; i++) {
            SnapLabelsHdr lh{};
            lh.algo = (uint32_t)L[i]->algo;
            lh.trace = (uint32_t)L[i]->trace;
            lh.root = L[i]->root;
            lh.has_init = L[i]->algo == RG_WCC ? 1u : 0u;
            fwrite_all(f, &lh, sizeof(lh));
            if (lh.has_init) {
                init.resize(V);
                uint64_t *di = dalloc<uint64_t>(V);
                launch_read_init(L[i]->init, (uint32_t)V, di, g->st);
                hipError_t e = hipGetLastError();
                if (e == hipSuccess) e = hipMemcpyAsync(init.data(), di, V * 8, hipMemcpyDeviceToHost, g->st);
                if (e == hipSuccess) e = hipStreamSynchronize(g->st);
                dfree(di);
                hipchk(e, "snapshot init labels");
                fwrite_all(f, init.data(), V * 8);
            }
            if (rg_labels_read(L[i], 0, V, val.data(), par.data()) != RG_OK) throw RgError{RG_EDEVICE, g_err};
            fwrite_all(f, val.data(), V * 8);
            fwrite_all(f, par.data(), V * 8);
        }
        if (fflush(f) != 0 || fsync(fileno(f)) != 0)
            throw RgError{RG_EDEVICE, std::string("snapshot sync: ") + strerror(errno)};
        fclose(f);
        fc.f = nullptr;
        if (rename(tmp.c_str(), path) != 0) throw RgError{RG_EDEVICE, std::string("snapshot rename: ") + strerror(errno)};
    });
}

int rg_snapshot_load(const char *path, rg_graph **gout, rg_labels **Lout, int cap, int *nL_out)
{
    if (!path || !gout || (cap && !Lout) || cap < 0 || !nL_out) return RG_EINVAL;
    *gout = nullptr;
    *nL_out = 0;
    rg_graph *g = nullptr;
    std::vector<rg_labels *> made;
    int rc = guard([&] {
        FILE *f = fopen(path, "rb");
        if (!f) throw RgError{RG_EINVAL, std::string("snapshot open: ") + strerror(errno)};
        FileCloser fc{f};
        SnapHdr h{};
        fread_all(f, &h, sizeof(h));
        if (memcmp(h.magic, SNAP_MAGIC, 8) != 0) throw RgError{RG_EINVAL, "not a snapshot file"};
        if ((int)h.nL > cap) throw RgError{RG_EINVAL, "more labels in the snapshot than the output array holds"};
        int r = rg_graph_create(h.V, h.hint, (int)h.symmetric, (int)h.dual, (int)h.weighted, &g);
        if (r != RG_OK) throw RgError{r, g_err};
        // the edges: a device stream, loaded as directed inserts
        std::vector<rg_edge> e(h.n_edges);
        fread_all(f, e.data(), h.n_edges * sizeof(rg_edge));
        rg_stream *s = nullptr;
        r = rg_stream_from_edges(e.data(), h.n_edges, (int)h.weighted, &s);
        if (r != RG_OK) throw RgError{r, g_err};
        r = rg_stream_load(g, s, 0, h.n_edges, 1);
        rg_stream_destroy(s);
        if (r != RG_OK) throw RgError{r, g_err};
        e.clear();
        e.shrink_to_fit();
        std::vector<uint64_t> val(h.V), par(h.V), init;
        for (uint32_t i = 0; i < h.nL; i++) {
            SnapLabelsHdr lh{};
            fread_all(f, &lh, sizeof(lh));
            if (lh.has_init) {
                init.resize(h.V);
                fread_all(f, init.data(), h.V * 8);
            }
            fread_all(f, val.data(), h.V * 8);
            fread_all(f, par.data(), h.V * 8);
            rg_labels *L = nullptr;
            r = rg_labels_create(g, (int)lh.algo, lh.root, lh.has_init ? init.data() : nullptr, (int)lh.trace, &L);
            if (r != RG_OK) throw RgError{r, g_err};
            made.push_back(L);
            r = rg_labels_assign(L, val.data(), par.data());
            if (r != RG_OK) throw RgError{r, g_err};
        }
        g->version = h.version;
    });
    if (rc != RG_OK) {
        for (rg_labels *L : made) rg_labels_destroy(L);
        if (g) rg_graph_destroy(g);
        return rc;
    }
    for (size_t i = 0; i < made.size(); i++) Lout[i] = made[i];
    *nL_out = (int)made.size();
    *gout = g;
    return RG_OK;
}

int rg_stream_apply(rg_graph *g, rg_labels **L, int nL, rg_stream *s, uint64_t ins_first, uint64_t n_ins,
                    uint64_t del_first, uint64_t n_del, int directed, rg_batch_stats *stats)
{
    if (!g || !s || (nL && !L) || nL < 0) return RG_EINVAL;
    return guard([&] {
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        HC(hipStreamSynchronize(s->st));
        apply_batch(g, L, nL, s, ins_first, n_ins, del_first, n_del, directed, stats);
    });
}

int rg_autotune(rg_graph *g, rg_labels **L, int nL, rg_stream *s, const rg_autotune_params *p,
                rg_autotune_result *r)
{
    if (!g || !s || !p || !r || !L || nL < 1 || nL > MAXT) return RG_EINVAL;
    return guard([&] {
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        for (int i = 0; i < nL; i++)
            if (!L[i] || L[i]->g != g) throw RgError{RG_EINVAL, "labels belong to another graph"};
        HC(hipStreamSynchronize(s->st));
        memset(r, 0, sizeof(*r));
        const uint64_t V = g->V;
        std::vector<uint32_t> d0(V), d1(V);
        copy_degrees(g->lists[0], d0.data(), V, g->st);
        HC(hipStreamSynchronize(g->st));
        const uint32_t pairs = p->pairs ? p->pairs : 3u;
        static const uint32_t sizes[RG_TUNE_BUCKETS] = {16384u, 65536u, 262144u, 1048576u};
        static const uint32_t caps[] = {2, 8, 16};
        static const uint32_t wsets[][4] = {{1, 1, 1, 1}, {1, 2, 3, 2}, {1, 3, 3, 1}, {1, 4, 2, 1}};
        const uint64_t ins0 = p->ins_first % s->n, del0 = p->del_first % s->n;
        // a forward batch and its inverse; device ms of the pair
        auto run_pair = [&](uint32_t ops) {
            const uint64_t half = ops / 2;
            apply_batch(g, L, nL, s, ins0, half, del0, half, p->directed, nullptr);
            double ms = g->last_ms;
            apply_batch(g, L, nL, s, del0, half, ins0, half, p->directed, nullptr);
            r->batches += 2;
            return ms + g->last_ms;
        };
        auto measure = [&](int b, const rg_graph::TuneEntry &e) {
            g->tune_tab[b] = e;
            run_pair(sizes[b]);   // warm-up (and the bucket's captured pipeline)
            std::vector<double> t;
            for (uint32_t k = 0; k < pairs; k++) t.push_back(run_pair(sizes[b]));
            std::sort(t.begin(), t.end());
            return t[t.size() / 2];
        };
        bool used[4] = {false, false, false, false};
        for (int i = 0; i < nL; i++)
            if (L[i]->algo >= 0 && L[i]->algo < 4) used[L[i]->algo] = true;
        double fwd_ms[RG_TUNE_BUCKETS] = {};
        for (int b = 0; b < RG_TUNE_BUCKETS; b++) {
            if (sizes[b] / 2 > s->n) break;
            r->ops[b] = sizes[b];
            rg_graph::TuneEntry best;
            double best_ms = measure(b, best);
            r->default_ms[b] = best_ms;
            // a setting has to win by 3% to replace the incumbent
            for (uint32_t c : caps) {
                rg_graph::TuneEntry e = best;
                e.cap_x4 = c;
                const double ms = measure(b, e);
                if (ms < 0.97 * best_ms) best = e, best_ms = ms;
            }
            if (nL > 1) {
                for (const auto &w : wsets) {
                    rg_graph::TuneEntry e = best;
                    for (int a = 0; a < 4; a++) e.weight[a] = w[a];
                    const double ms = measure(b, e);
                    if (ms < 0.97 * best_ms) best = e, best_ms = ms;
                }
            }
            g->tune_tab[b] = best;
            g->run_tune = &g->tune_tab[b];
            r->cap_x4[b] = best.cap_x4;
            for (int a = 0; a < 4; a++) r->weight[b][a] = used[a] ? algo_weight(g, a) : 0u;
            r->best_ms[b] = best_ms;
            fwd_ms[b] = best_ms / 2;
        }
        g->run_tune = nullptr;
        if (p->target_ms > 0) {
            // the largest batch whose device time, doubled, fits a quarter of
            // the target (the serving loop's shrink rule): measured sizes,
            // then the line through the two largest.  Doubled because an
            // update of a capped batch also waits for the batch before it:
            // on R-MAT scale 24 the rule gives about 1M updates, and fixed
            // caps of 2^20 and 2^21 served 250M updates/s at p99.9 19.2 and
            // 21.7 ms.
            const double budget = 0.125 * p->target_ms;
            // piecewise-linear device time over the measured sizes (the
            // last segment extended, up to 4x the largest size)
            double cap = 0;
            int m = 0;
            while (m < RG_TUNE_BUCKETS && r->ops[m]) m++;
            for (int b = 0; b < m; b++) {
                if (fwd_ms[b] <= budget) cap = r->ops[b];
                const bool last = b + 1 == m;
                if (b == 0 || fwd_ms[b] <= budget) {
                    if (!last && fwd_ms[b + 1] > budget && fwd_ms[b + 1] > fwd_ms[b])
                        cap = r->ops[b] + (budget - fwd_ms[b]) / (fwd_ms[b + 1] - fwd_ms[b]) *
                                              ((double)r->ops[b + 1] - r->ops[b]);
                }
                if (last && b >= 1 && fwd_ms[b] <= budget && fwd_ms[b] > fwd_ms[b - 1]) {
                    const double slope = (fwd_ms[b] - fwd_ms[b - 1]) / ((double)r->ops[b] - r->ops[b - 1]);
                    cap = std::min(4.0 * r->ops[b], r->ops[b] + (budget - fwd_ms[b]) / slope);
                }
            }
            r->batch_cap = (uint32_t)std::max(cap, 1024.0);
        }
        copy_degrees(g->lists[0], d1.data(), V, g->st);
        HC(hipStreamSynchronize(g->st));
        if (d0 != d1)
            throw RgError{RG_EINVAL, "autotune windows did not restore the graph (deleted edges must be live)"};
    });
}

int rg_stream_plain(rg_graph *g, rg_labels **L, int nL, rg_stream *s, uint64_t ins_first, uint64_t del_first,
                    uint64_t n_steps, int directed, double *lat_ns, rg_batch_stats *stats)
{
    if (!g || !s || (nL && !L) || nL < 0) return RG_EINVAL;
    return guard([&] {
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        HC(hipStreamSynchronize(s->st));
        for (uint64_t k = 0; k < n_steps; k++) {
            auto t0 = std::chrono::steady_clock::now();
            apply_batch(g, L, nL, s, (ins_first + k) % s->n, 1, (del_first + k) % s->n, 1, directed, stats);
            auto t1 = std::chrono::steady_clock::now();
            if (lat_ns) lat_ns[k] = (double)std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count();
        }
    });
}

int rg_last_gpu_ms(rg_graph *g, double *ms)
{
    if (!g || !ms) return RG_EINVAL;
    *ms = g->last_ms;
    return RG_OK;
}

int rg_last_unsafe(rg_graph *g, uint64_t *n)
{
    if (!g || !n) return RG_EINVAL;
    *n = g->last_unsafe;
    return RG_OK;
}

int rg_profile_enable(rg_graph *g, int on)
{
    if (!g) return RG_EINVAL;
    return guard([&] {
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        prof_collect(g);
        g->prof = on & 7;
    });
}

int rg_profile_read(rg_graph *g, int kind, uint64_t *launches, double *ms, uint64_t *units)
{
    if (!g || kind < 0 || kind >= rg_graph::PKINDS) return RG_EINVAL;
    return guard([&] {
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        prof_collect(g);
        if (launches) *launches = g->p_launch[kind];
        if (ms) *ms = g->p_ms[kind];
        if (units) *units = g->p_units[kind];
    });
}

int rg_labels_team_weight(rg_labels *L, uint32_t weight)
{
    if (!L || weight > 64) return RG_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(L->g->mu);
    L->team_weight = weight;
    return RG_OK;
}

int rg_labels_dense(rg_labels *L, int mode, uint64_t *dense_passes)
{
    if (!L || mode < -1 || mode > 1) return RG_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(L->g->mu);
    if (mode >= 0) L->tune.dense = (uint32_t)mode;
    if (dense_passes) *dense_passes = L->dense_passes;
    return RG_OK;
}

int rg_labels_phase_time(rg_labels *L, double *us, uint64_t *passes)
{
    if (!L) return RG_EINVAL;
    return guard([&] {
        rg_graph *g = L->g;
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        unsigned long long t[NTICK * NPHASE];
        HC(hipMemcpyAsync(t, L->ticks, sizeof(t), hipMemcpyDeviceToHost, g->st));
        HC(hipStreamSynchronize(g->st));
        int khz = 0;
        int dev = 0;
        HC(hipGetDevice(&dev));
        HC(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        double per_us = khz > 0 ? khz / 1000.0 : 100.0;
        for (int p = 0; p < NPHASE; p++) {
            if (us) us[p] = (double)t[p] / per_us;
            if (passes) passes[p] = t[NPHASE + p];
        }
    });
}

int rg_labels_phase_detail(rg_labels *L, double *stage_us, uint64_t *work, uint64_t *chunks)
{
    if (!L) return RG_EINVAL;
    return guard([&] {
        rg_graph *g = L->g;
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        unsigned long long t[NTICK * NPHASE];
        HC(hipMemcpyAsync(t, L->ticks, sizeof(t), hipMemcpyDeviceToHost, g->st));
        HC(hipStreamSynchronize(g->st));
        int khz = 0;
        int dev = 0;
        HC(hipGetDevice(&dev));
        HC(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        double per_us = khz > 0 ? khz / 1000.0 : 100.0;
        for (int p = 0; p < NPHASE; p++) {
            for (int k = 0; k < 4; k++)
                if (stage_us) stage_us[k * NPHASE + p] = (double)t[(2 + k) * NPHASE + p] / per_us;
            if (work) work[p] = t[6 * NPHASE + p];
            if (chunks) chunks[p] = t[7 * NPHASE + p];
        }
    });
}

int rg_degrees(rg_graph *g, uint32_t *out_deg, uint32_t *in_deg)
{
    if (!g) return RG_EINVAL;
    return guard([&] {
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        if (out_deg) copy_degrees(g->lists[0], out_deg, g->V, g->st);
        if (in_deg) copy_degrees(g->symmetric ? g->lists[0] : g->lists[1], in_deg, g->V, g->st);
        HC(hipStreamSynchronize(g->st));
    });
}

extern "C++" {

// Latency percentiles of a serving round: for each q, the k-th smallest
// latency (k = ceil(q m) - 1) among the m updates u < n whose want(u) holds,
// the value nth_element would put at rank k.  Rounds hold tens of millions of
// updates, so nothing is copied whole: a histogram of 1-us bins finds the bin
// that holds rank k, and only that bin's values are ordered.
template <class Want>
static void percentiles_of(const std::vector<double> &lat, uint64_t n, Want want, const double *q, double *out,
                           int nq, uint64_t *count)
{
    constexpr uint32_t NB = 1u << 21;   // up to ~2.1 s; the last bin takes everything above
    auto bin = [](double x) -> uint32_t {
        const double b = x * 1000.0;
        return b >= (double)(NB - 1) ? NB - 1 : (b > 0 ? (uint32_t)b : 0u);
    };
    std::vector<uint32_t> h(NB, 0u);
    uint64_t m = 0;
    for (uint64_t u = 0; u < n; u++)
        if (want(u)) {
            h[bin(lat[u])]++;
            m++;
        }
    if (count) *count = m;
    std::vector<double> in;
    for (int i = 0; i < nq; i++) {
        if (!m) {
            out[i] = 0;
            continue;
        }
        const uint64_t k = (uint64_t)std::min<double>((double)m - 1, std::ceil(q[i] * (double)m) - 1);
        uint64_t below = 0;
        uint32_t b = 0;
        while (below + h[b] <= k) below += h[b++];
        in.clear();
        for (uint64_t u = 0; u < n; u++)
            if (want(u) && bin(lat[u]) == b) in.push_back(lat[u]);
        std::nth_element(in.begin(), in.begin() + (k - below), in.end());
        out[i] = in[k - below];
    }
}

// Persistent helper threads of a serving loop.  post() hands out jobs
// 0..n-1 and returns at once; wait() returns when every job of the last
// post has run (and rethrows the first exception a job threw).
class HostPool {
public:
    explicit HostPool(unsigned n)
    {
        for (unsigned i = 0; i < n; i++) th_.emplace_back([this] { loop(); });
    }
    ~HostPool()
    {
        {
            std::lock_guard<std::mutex> l(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    void post(unsigned n, std::function<void(unsigned)> f)
    {
        wait();
        {
            std::lock_guard<std::mutex> l(m_);
            f_ = std::move(f);
            n_ = n;
            next_ = 0;
            left_ = n;
        }
        cv_.notify_all();
    }
    void wait()
    {
        std::unique_lock<std::mutex> l(m_);
        done_.wait(l, [this] { return left_ == 0; });
        if (err_) {
            std::exception_ptr e = err_;
            err_ = nullptr;
            std::rethrow_exception(e);
        }
    }

private:
    void loop()
    {
        std::unique_lock<std::mutex> l(m_);
        for (;;) {
            cv_.wait(l, [this] { return stop_ || next_ < n_; });
            if (stop_) return;
            const unsigned j = next_++;
            l.unlock();
            std::exception_ptr e;
            try {
                f_(j);
            } catch (...) {
                e = std::current_exception();
            }
            l.lock();
            if (e && !err_) err_ = e;
            if (--left_ == 0) done_.notify_all();
        }
    }
    std::vector<std::thread> th_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    std::function<void(unsigned)> f_;
    unsigned n_ = 0, next_ = 0, left_ = 0;
    bool stop_ = false;
    std::exception_ptr err_;
};

// Poisson arrival times of a round: exponential gaps from a counter-based
// uniform draw per update, summed per chunk on several host threads, then
// offset by the chunks before (the round's clock starts at 0)
static void poisson_arrivals(std::vector<double> &arr, uint64_t seed, double rate)
{
    const uint64_t n = arr.size();
    const uint64_t key = seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    const unsigned nt = n < (1u << 20) ? 1u : std::min(8u, std::max(1u, std::thread::hardware_concurrency()));
    const uint64_t chunk = (n + nt - 1) / nt;
    std::vector<double> sum(nt, 0.0);
    auto fill = [&](unsigned c) {
        double t = 0;
        for (uint64_t u = c * chunk; u < std::min(n, (c + 1) * chunk); u++) {
            uint64_t z = key + (u + 1) * 0x9E3779B97F4A7C15ull;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            const double uni = ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
            t += -std::log(uni) / rate;
            arr[u] = t;
        }
        sum[c] = t;
    };
    auto shift = [&](unsigned c, double off) {
        for (uint64_t u = c * chunk; u < std::min(n, (c + 1) * chunk); u++) arr[u] += off;
    };
    std::vector<std::thread> th;
    for (unsigned c = 1; c < nt; c++) th.emplace_back(fill, c);
    fill(0);
    for (auto &x : th) x.join();
    th.clear();
    double off = 0;
    for (unsigned c = 1; c < nt; c++) {
        off += sum[c - 1];
        th.emplace_back(shift, c, off);
    }
    for (auto &x : th) x.join();
}

}  // extern "C++"

int rg_serve(rg_graph *g, rg_labels **L, int nL, rg_stream *s, const rg_serve_params *p, rg_serve_result *r,
             double *lat_ms)
{
    if (!g || !s || !p || !r || (nL && !L) || nL < 0 || p->rate <= 0 || p->batch_cap == 0) return RG_EINVAL;
    return guard([&] {
        std::lock_guard<std::recursive_mutex> lk(g->mu);
        HC(hipStreamSynchronize(s->st));
        const uint64_t n = p->n_updates;
        // Poisson arrivals: exponential gaps from a counter-based uniform draw
        std::vector<double> arr(n), lat(n);
        poisson_arrivals(arr, p->seed, p->rate);
        rg_batch_stats st{};
        // cap: updates per batch, or unsafe updates per batch with cap_unsafe
        // (a batch then admits cap / frac updates, frac = the running unsafe
        // fraction of split batches)
        const bool cap_u = p->cap_unsafe && p->split;
        double cap = p->batch_cap;
        double frac = 0.25;
        const double cap_max = std::max<double>(p->cap_max ? p->cap_max : p->batch_cap, 1.0);
        auto limit = [&] { return std::min(cap_max, std::max(1.0, cap_u ? cap / frac : cap)); };
        const double max_wait = std::max(10.0 * p->target_ms, 200.0) * 1e-3;
        uint64_t next = 0, batches = 0, max_batch = 0;
        std::vector<uint8_t> waited(n, 1);   // 1: the update waited for its batch's repair
        bool overloaded = false;
        double gpu_ms = 0;
        auto t0 = std::chrono::steady_clock::now();
        auto now = [&] {
            return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        };
        SplitOut so;
        so.flags = p->split != 0;
        // The safe-answer step of a batch runs on helper threads while this
        // thread goes on driving the repair (a tree that leaves the team
        // launch needs the host for its grid-wide steps), split into SAFE_PARTS
        // ranges; at a million updates one thread took as long as the repair.
        // Each range keeps the updates that wait for the repair; those of the
        // previous batch, completed at t_prev, get their latencies in the
        // same step (one more job), so the device does not idle between
        // batches for it.
        constexpr unsigned SAFE_PARTS = 4;
        HostPool pool(SAFE_PARTS);
        std::vector<std::vector<uint64_t>> unsafe_u(SAFE_PARTS), unsafe_prev(SAFE_PARTS);
        double t_prev = 0;
        auto settle_prev = [&] {
            for (auto &part : unsafe_prev) {
                for (uint64_t u : part) lat[u] = (t_prev - arr[u]) * 1e3;
                part.clear();
            }
        };
        bool timed = true;                // latencies of the current batch are recorded
        double t_repair = 0;              // the current batch's repair completed
        auto run = [&](uint64_t b, uint64_t e) {
            // updates [b, e): adds are the even indices, deletes the odd ones
            uint64_t a0 = (b + 1) / 2, a1 = (e + 1) / 2;
            uint64_t d0 = b / 2, d1 = e / 2;
            so.on_safe = [&, b, e, a0, a1, d0](const uint8_t *flag) {
                const double ts = std::chrono::duration<double>(so.t_safe - t0).count();
                pool.post(SAFE_PARTS + 1, [&, b, e, a0, a1, d0, flag, ts](unsigned j) {
                    if (j == SAFE_PARTS) {
                        settle_prev();
                        return;
                    }
                    auto &mine = unsafe_u[j];
                    mine.clear();
                    const uint64_t pb = b + (e - b) * j / SAFE_PARTS, pe = b + (e - b) * (j + 1) / SAFE_PARTS;
                    for (uint64_t u = pb; u < pe; u++) {
                        // op index of update u: inserts first, then deletes
                        const uint64_t op = (u & 1) ? (a1 - a0) + (u / 2 - d0) : u / 2 - a0;
                        const uint8_t w = flag[op];
                        waited[u] = w;
                        if (w) mine.push_back(u);
                        else if (timed) lat[u] = (ts - arr[u]) * 1e3;
                    }
                });
            };
            try {
                apply_batch(g, L, nL, s, (p->ins_first + a0) % s->n, a1 - a0, (p->del_first + d0) % s->n,
                            d1 - d0, p->directed, &st, p->split ? &so : nullptr);
            } catch (...) {
                // the helpers may still read the batch's flags
                try {
                    pool.wait();
                } catch (...) {
                }
                throw;
            }
            t_repair = now();
            pool.wait();
            gpu_ms += g->last_ms;
        };
        while (next < n) {
            double tn = now();
            while (arr[next] > tn) tn = now();   // idle until the next request arrives
            if (tn - arr[next] > max_wait) {
                overloaded = true;
                break;
            }
            uint64_t arrived = (uint64_t)(std::upper_bound(arr.begin() + next, arr.end(), tn) - arr.begin());
            uint64_t end = std::min<uint64_t>(arrived, next + (uint64_t)std::floor(limit()));
            run(next, end);
            const double tc = t_repair;
            settle_prev();   // a batch without the split had no safe-answer step
            if (so.done) {
                std::swap(unsafe_u, unsafe_prev);
                t_prev = tc;
            } else {
                for (uint64_t u = next; u < end; u++) lat[u] = (tc - arr[u]) * 1e3;
            }
            if (so.done) frac = std::max(0.01, 0.8 * frac + 0.2 * (double)so.unsafe / (double)(end - next));
            if (p->target_ms > 0) {
                double service_ms = (tc - tn) * 1e3;
                uint64_t backlog = (uint64_t)(std::upper_bound(arr.begin() + end, arr.end(), tc) - arr.begin()) - end;
                // a backlog grows the cap first (larger batches drain it in
                // fewer fixed costs); without one, a slow batch shrinks it
                const double cap_top = cap_u ? cap_max * frac : cap_max;
                if ((double)backlog >= limit()) cap = std::min(cap_top, cap * 1.5 + 16);
                else if (service_ms > 0.25 * p->target_ms) cap = std::max(1.0, cap * 0.7);
            }
            batches++;
            max_batch = std::max<uint64_t>(max_batch, end - next);
            next = end;
        }
        const uint64_t served = next;
        double wall = now();
        settle_prev();
        // an overloaded run still applies the rest (untimed) so the stream
        // window stays consistent for the next run
        timed = false;
        while (next < n) {
            uint64_t end = std::min<uint64_t>(n, next + (uint64_t)cap_max);
            run(next, end);
            next = end;
        }
        {
            const double q50 = 0.50, qu[2] = {0.50, 0.999};
            double vu[2];
            uint64_t n_safe = 0, n_unsafe = 0;
            percentiles_of(lat, served, [&](uint64_t u) { return waited[u] == 0; }, &q50, &r->p50_safe_ms, 1,
                           &n_safe);
            percentiles_of(lat, served, [&](uint64_t u) { return waited[u] != 0; }, qu, vu, 2, &n_unsafe);
            r->unsafe = n_unsafe;
            r->safe_frac = served ? (double)n_safe / (double)served : 0;
            r->p50_unsafe_ms = vu[0];
            r->p999_unsafe_ms = vu[1];
        }
        r->overloaded = overloaded ? 1 : 0;
        r->served = served;
        r->gpu_ms = gpu_ms;
        r->wall_s = wall;
        r->updates_per_s = wall > 0 ? (double)served / wall : 0;
        r->batches = batches;
        r->max_batch = max_batch;
        r->mean_batch = batches ? (double)served / (double)batches : 0;
        r->grid_steps = st.grid_steps;
        r->rebuilds = st.rebuilds;
        double sum = 0, mx = 0;
        for (uint64_t u = 0; u < served; u++) {
            sum += lat[u];
            mx = std::max(mx, lat[u]);
        }
        r->mean_ms = served ? sum / (double)served : 0;
        r->max_ms = mx;
        {
            const double q[3] = {0.50, 0.99, 0.999};
            double v[3];
            percentiles_of(lat, served, [](uint64_t) { return true; }, q, v, 3, nullptr);
            r->p50_ms = v[0];
            r->p99_ms = v[1];
            r->p999_ms = v[2];
        }
        for (uint64_t u = served; u < n; u++) lat[u] = std::numeric_limits<double>::infinity();
        if (lat_ms) std::copy(lat.begin(), lat.end(), lat_ms);
    });
}

}  // extern "C"
