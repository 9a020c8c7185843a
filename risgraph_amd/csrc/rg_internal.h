// rg_internal.h -- device data layout shared by the store and engine kernels.
//
// HBM layout (one graph):
//   slot pool   nbr[S] (u32: neighbour | DEAD bit), wt[S] (u32, weighted
//               graphs only), cnt[S] (u32 multiplicity).  Each vertex list is
//               one contiguous block [off, off+cap) of the pool; `len` entries
//               are used, tombstones (cnt == 0) keep their place, as
//               IndexedEdgeStorage does (storage.hpp:369-394).
//   hash pool   hb[H]: 16-byte buckets {key (u64 nbr | w << 32, ~0 = empty),
//               local slot, copy of the entry's count} -- per-list open-
//               addressing index built once a list reaches INDEX_THRESHOLD
//               entries (the reference's is 512, storage.hpp:304, 358-367).
//   headers     per direction (out, in): one 64-byte ListHdr per vertex with
//               off, len, cap, hoff, hcap, hcnt, deg, and nops / seg (the
//               current batch's ops on the list and their segment in the
//               batch's op-index buffer).
// One labels instance (one algorithm's dependency tree):
//   vr[V]       16-byte records {val, inv, stamp}: val is the u64 packed
//               word value << 32 | parent (parent 0xFFFFFFFF = empty), the
//               device form of VertexTree (graph.hpp:54-59), updated by
//               64-bit CAS with a strict-improvement check; inv the
//               invalidation mark (graph.hpp:1729); stamp the frontier
//               dedupe epoch.
//   init[V]     init labels; mstamp[V] modified-set dedupe epochs;
//               q[3][V] vertex queues (cur / next / tree); cq chunk queue of
//               long lists; mlist[V] modified set; ctl control block.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rg {

enum Algo : int { BFS = 0, SSSP = 1, SSWP = 2, WCC = 3 };

constexpr uint32_t DEAD = 0x80000000u;
constexpr uint32_t NBR_MASK = 0x7FFFFFFFu;
constexpr uint32_t NILP = 0xFFFFFFFFu;
// A list gets its open-addressing index at this many entries.  The
// reference indexes at 512 (storage.hpp:304); on the device a 2-lane scan of
// even a hundred entries is a chain of dependent loads, while a 16-byte
// bucket brings key, slot and count in one line, so lists are indexed from
// 32 entries (indexes start at 64 buckets).  Returns and counts do not
// depend on it.
constexpr uint32_t INDEX_THRESHOLD = 32;
constexpr uint32_t INDEX_MIN_CAP = 64;
constexpr uint64_t HEMPTY = ~0ull;
constexpr uint32_t LONG_LIST = 32;          // longer lists go to the chunk queue
constexpr uint32_t CHUNK = 128;             // entries per chunk item (two wave steps)
constexpr int GROUP = 16;                   // lanes per vertex in vertex passes
constexpr int WAVE = 64;

// One bucket of a list's hash index: the key, the entry's local index and a
// copy of the entry's count.  cnt[] stays the count of record (traversal,
// copies, dumps and host reads use it); every count write on an indexed list
// updates the copy too, so a probe that hits reads the count from the line it
// already loaded instead of two more dependent lines (index, then count).
struct alignas(16) HBucket {
    unsigned long long key;
    uint32_t idx;
    uint32_t cnt;
};
static_assert(sizeof(HBucket) == 16, "hash bucket is 16 bytes");

struct Slots {
    uint32_t *nbr;
    uint32_t *wt;
    uint32_t *cnt;
    HBucket *hb;
};

// One 64-byte record per (direction, vertex): a list op of the store touches
// one cache line of header state, and traversal reads the first 16 bytes.
struct alignas(64) ListHdr {
    uint64_t off;     // block offset in the slot pool
    uint32_t len;     // entries used (live and tombstones)
    uint32_t cap;     // block capacity (power of two, 0 = no block)
    uint64_t hoff;    // index offset in the hash pool (HEMPTY: not indexed)
    uint32_t hcap;    // index capacity
    uint32_t hcnt;    // keys in the index
    uint32_t deg;     // live multiplicity (storage.hpp:396)
    uint32_t nops;    // list ops of the current batch (0 between batches)
    uint32_t seg;     // their segment start in ListOps::segbuf
    uint32_t ncap;    // block reserved for this batch's growth (0: none) ...
    uint64_t noff;    // ... and its offset in the slot pool
    uint32_t pad[2];
};
static_assert(sizeof(ListHdr) == 64, "list header is one 64-byte line");

struct List {
    ListHdr *h;
};

// block offset and length of v's list: one 16-byte load
__device__ __forceinline__ void list_span(const List &l, uint32_t v, uint64_t &off, uint32_t &len)
{
    const uint4 x = *reinterpret_cast<const uint4 *>(&l.h[v]);
    off = (uint64_t)x.x | ((uint64_t)x.y << 32);
    len = x.z;
}


// Slab allocator over the slot pool (0) and the hash pool (1): blocks are
// powers of two; a freed block goes to its class's pending list and becomes
// reusable after the batch (pending -> free between kernels), so a pop and a
// push never touch the same list inside one kernel.  Pops take from the end
// of the free list; the bump pointer serves what the free lists cannot.
constexpr int NCLS = 32;
// Entries per class list (StoreArgs::fcap, sized from the slot pool at
// graph creation); frees past it are counted (Pool::dropped) and leak.

struct Pool {
    unsigned long long slot_top, slot_cap;
    unsigned long long hash_top, hash_cap;
    unsigned int oom;
    unsigned int n_heavy;   // current batch: heavy lists (a workgroup each)
    // current batch: segment entries reserved (low word) and light lists (high word)
    unsigned long long seg_pack;
    unsigned int free_n[2][NCLS];
    unsigned int free_take[2][NCLS];
    unsigned int pend_n[2][NCLS];
    unsigned long long recycled[2];
    // frees that found no room (pending list past fcap, or free list full):
    // blocks and entries that no later batch can reuse
    unsigned long long dropped[2], dropped_entries[2];
};

// edge-op kinds (bitmask)
constexpr uint8_t K_INS = 1;      // store insert (+1), seed if old==0 and live
constexpr uint8_t K_DEL = 2;      // store delete (-1), root if old==1
constexpr uint8_t K_FORCE = 4;    // engine-only edge from update_tree_add/del
constexpr uint8_t K_UNDIR = 8;    // undirected op (both directions)
constexpr uint8_t K_NOSTORE = 16; // do not touch the store

// What the store reports for one edge op; the list owner that writes the
// return value usually writes the liveness too, so both share a word pair.
struct OpResult {
    uint32_t old;       // reference add/del return value
    uint32_t live;      // the (src->dst) entry of src's out-list is live after the batch
};

struct EdgeOps {
    uint32_t *src, *dst, *w;
    uint8_t *kind;
    OpResult *res;
    uint32_t n;
};

// list-op flags
constexpr uint8_t L_IN = 1;       // in-list (else out-list)
constexpr uint8_t L_RET = 2;      // its old count is the edge op's return value
constexpr uint8_t L_SLOT = 4;     // its entry's final count gives the edge op's liveness
constexpr uint8_t L_SELF = 8;     // undirected self-loop folded into one +-2 op

// One list op as a 32-byte record: the apply reaches a list's ops in
// segment order, which is not index order, so each op's fields come from one
// memory line instead of one line per field array.
struct alignas(16) LOp {
    uint32_t v;         // list vertex (NILP: unused slot)
    uint32_t nbr, w;    // the entry's key
    uint32_t edge;      // edge op it came from
    uint32_t rank;      // position among the batch's ops on the same list
    uint32_t slot;      // local index the op resolved to (NILP: none)
    int8_t delta;
    uint8_t flags;
    uint16_t pad0;
    uint32_t pad1;
};
static_assert(sizeof(LOp) == 32, "list op record is 32 bytes");

struct ListOps {
    LOp *op;
    uint32_t *segbuf;   // list-op indices grouped by list (segments)
    // the light lists (dense): a copy of each list's rank-0 op record with
    // v = vertex | in-list << 31 and rank = that op's index, written by the
    // segment kernel in index order and read in order by the apply, so a
    // single-op list's fields cost no random line
    LOp *owner;
    uint32_t *heavy;    // rank-0 list ops of the heavy lists (dense)
    uint32_t n;
};

enum Phase : uint32_t {
    PH_DONE = 0,
    PH_SEED = 1,
    PH_PUSH = 2,
    PH_DEL_ROOTS = 3,
    PH_DEL_RESCUE = 4,
    PH_DEL_EXPAND = 5,
    PH_DEL_RESET = 6,
    PH_DEL_PULL = 7,
    PH_BUILD_INIT = 8,
    PH_REBUILD = 9,
};

constexpr int NPHASE = 10;
// batches of at most this many edge ops run the one-workgroup store kernel
constexpr uint32_t STORE_SMALL_OPS = 64;
// per-phase team statistics, NPHASE words each: wall-clock ticks, passes,
// then workgroup 0's vertex stage, first barrier, chunk stage and second
// barrier ticks, summed pass work and summed chunk items
constexpr int NTICK = 8;

constexpr uint32_t F_NEED_REBUILD = 1;
constexpr uint32_t F_QUEUE_OVERFLOW = 2;
constexpr uint32_t F_ALL_MODIFIED = 4;
constexpr uint32_t F_ABORT = 8;        // a team barrier timed out
constexpr uint32_t F_BUILD = 16;       // a full build: its push passes may run dense (pull)
constexpr uint32_t F_DENSE = 32;       // the current push pass pulls over every vertex

struct Ctl {
    uint32_t phase, stage, level, seed;   // seed: inserts still to seed (mixed batch)
    uint32_t n_cur, n_next, n_tree, n_chunk;
    uint32_t n_mod, n_ops;
    uint32_t sel_cur, sel_next, sel_tree;
    uint32_t epoch, mepoch;
    int32_t idx;
    uint32_t flags;
    uint32_t bar_count;    // team barrier: arrivals (monotonic within a launch)
    uint32_t bar_gen;      // team barrier: last released episode
    uint32_t bar_exit;     // team workgroups that left the launch
    uint32_t rbound;       // smallest deletion-root value (largest for SSWP)
    uint32_t work_max;     // largest pass the team runs (larger: grid-wide)
    // builds: live edges the next frontier pushes over (sum of its push-list
    // degrees), and every live edge of the push lists (from the init pass)
    unsigned long long e_next, e_live, levels;
    uint32_t seq;          // passes closed (every phase transition)
    uint32_t dense;        // dense passes of the run
};

// One launch repairs up to MAXT trees; each tree gets W workgroups that
// synchronise through a barrier in the tree's control block.
constexpr int MAXT = 4;
// The team barrier counts arrivals in up to BAR_SUBS sub-counters, one
// 128-byte line apart, whose last arrivals count on the control block.
constexpr uint32_t BAR_SUBS = 8;
constexpr uint32_t BAR_STRIDE = 32;

constexpr uint32_t RESCUE_NEVER = 0xFFFFFFFFu;   // rescue_level: no level qualifies

// One 16-byte record per vertex of a tree: a pass that visits a vertex reads
// or updates its label word, its invalidation mark and its frontier stamp
// together, one memory line instead of three.
struct alignas(16) VRec {
    uint64_t val;     // value << 32 | parent vertex (parent 0xFFFFFFFF: empty)
    int32_t inv;      // invalidation mark: run index, -index while rescued, else stale
    uint32_t stamp;   // frontier dedupe epoch
};
static_assert(sizeof(VRec) == 16, "vertex record is 16 bytes");

struct Tuning {
    uint32_t rescue_level;        // rescue when level <= this (RESCUE_NEVER: none) ...
    uint32_t rescue_frontier;     // ... or frontier < this (graph.hpp:1324)
    uint32_t recompute_level;     // fallback checked up to this level (graph.hpp:1252)
    uint32_t recompute_abs;       // ... when subtree+frontier > this ...
    float recompute_frac;         // ... and > frac * V
    uint32_t walk_limit;          // has_parents step cap (graph.hpp:1900)
    uint32_t dense;               // builds switch to pull passes by the reference's rule (0: push only)
};

// One chunk of a long list for a pass's chunk stage.  It carries where its
// entries start, so the wave that takes it goes straight to them instead of
// first loading the list's header.
struct alignas(16) ChunkItem {
    uint64_t off;    // first slot of the chunk
    uint32_t v;      // the list's vertex
    uint32_t n_op;   // entries (1..CHUNK) | op << 16
};

struct EngineArgs {
    uint32_t V;
    int algo;
    int trace;                    // maintain the modified set
    int undirected;               // algorithm sees edges in both directions
    Slots slots;
    List lists[2];
    int npush, push_l[2];         // lists scanned by push / expand
    int npull, pull_l[2];         // lists scanned by pull / rescue
    VRec *vr;                     // label word, mark and frontier stamp per vertex
    uint32_t *init;
    uint32_t *init_active;        // nullptr: only vertex `root` starts active
    uint32_t root;
    uint32_t *mstamp;
    uint32_t *q[3];
    uint32_t *bnd;                // SSWP: lower bound of an invalidated vertex's new value
    ChunkItem *cq;
    uint32_t cq_cap;
    uint32_t *mlist;
    Ctl *ctl;
    Ctl *hctl;                    // pinned host mirror, written at the end of each launch
    unsigned long long *ticks;    // [NTICK*NPHASE] per-phase pass statistics (team mode)
    uint32_t *bar;                // [BAR_SUBS*BAR_STRIDE] team barrier sub-counters (zero between launches)
    EdgeOps ops;
    Tuning tune;
};

struct StoreArgs {
    uint32_t V;
    int symmetric;
    int weighted;
    Slots slots;
    List lists[2];
    Pool *pool;
    uint64_t *free_off;   // [2][NCLS][fcap] reusable block offsets
    uint64_t *pend_off;   // [2][NCLS][fcap] blocks freed in the current batch
    EdgeOps ops;
    ListOps lops;
    uint32_t *err;   // bad vertex ids etc.
    uint32_t *hstatus;  // pinned host words: err, oom (written after each batch)
    uint32_t *cnt_tmp[2];  // per-vertex insert counts for bulk reservation
    // captured batch pipelines: the batch's edge-op count comes from this
    // device word instead of ops.n (the launch grids are sized for the
    // bucket's largest batch); null = ops.n
    const uint32_t *dyn_n;
    uint32_t fcap;   // entries per size-class free list
    // list-op slots per edge op: the most list ops one edge op of the batch
    // makes (1 directed on a symmetric store, 2 directed with in-lists or
    // undirected on a symmetric store, 4 undirected with in-lists)
    uint32_t lstride;
};

// The window of one captured stream batch: written by the host into pinned
// memory and copied to the device by the graph's first node.
struct BatchDesc {
    uint64_t ins_first, del_first;
    uint32_t n_ins, n_del, n, pad;
};

// host-side launchers (store.hip / engine.hip)
// ev_a / ev_b (may be null) are recorded around the apply kernels.  Returns
// true when the batch's pool recycling is already done (small batches take a
// one-workgroup path that includes it); otherwise launch_pool_recycle must
// follow.
bool launch_store_batch(const StoreArgs &a, hipStream_t s, hipEvent_t ev_a = nullptr,
                        hipEvent_t ev_b = nullptr);
void launch_store_count(const StoreArgs &a, hipStream_t s);
void launch_store_reserve(const StoreArgs &a, hipStream_t s);
void launch_pool_recycle(const StoreArgs &a, hipStream_t s);
// empty lists (no block, no index) for vertices [0, V)
void launch_list_init(List l, uint32_t V, hipStream_t s);
// out[v] = live degree of v's list (device array of V words)
void launch_list_degrees(List l, uint32_t V, uint32_t *out, hipStream_t s);
// an edge record as the C ABI's rg_edge (src, dst, data)
struct rg_edge_rec {
    uint64_t src, dst, data;
};
// live multiset of lists [v0, v1) as records: off[k] = first record of
// vertex v0 + k (relative to off[0]), off has v1 - v0 + 1 entries
void launch_list_dump(List l, Slots sl, uint32_t v0, uint32_t v1, const uint64_t *off, rg_edge_rec *out,
                      hipStream_t s);
// Persistent team launch of the trees whose passes fit it.  The launch has
// XCD_SLOTS * rounds workgroups, dealt to the XCDs round-robin (workgroup b
// runs on XCD b % 8): tree t owns k[t] of the eight XCD slots, so its
// W = rounds * k[t] workgroups share those XCDs' L2s, and a tree with more
// work gets more slots.  k[t] = 0: tree t is not in this launch.  A tree
// leaves the launch when its next pass exceeds work_max[t] (grid mode takes
// over) or its phase is done.  Bit t of init_mask: tree t's control block is
// taken from its pinned mirror (first launch of a run).
constexpr uint32_t XCD_SLOTS = 8;
struct TeamLaunch {
    const EngineArgs *a[MAXT];   // device copies of the trees' arguments
    uint32_t k[MAXT];
    uint32_t work_max[MAXT];
    uint32_t chunk_max[MAXT];    // a pass with more chunk items hands its chunk stage to grid mode
    uint32_t rounds;
    uint32_t init_mask;
};
void launch_engine_team(const TeamLaunch &tl, hipStream_t s);
// workgroups a team launch may have (all co-resident, as its barrier needs)
uint32_t team_resident_max();
// `steps` grid-wide passes enqueued back to back, sized for the pass the host
// last saw; each is a no-op unless the tree's current pass is grid-sized
// (checked on the device control block)
void launch_engine_grid_steps(const EngineArgs &a, const Ctl &host_ctl, int steps, hipStream_t s);
void launch_engine_fill_init(const EngineArgs &a, hipStream_t s);
// inv = 0 for every vertex record (the mark index wrapped)
void launch_clear_marks(VRec *vr, uint32_t V, hipStream_t s);
// out[i] = vr[i].val for i < n
void launch_read_labels(const VRec *vr, uint32_t n, uint64_t *out, hipStream_t s);
// vr[i].val = words[i], marks cleared, for i < n
void launch_write_labels(VRec *vr, uint32_t n, const uint64_t *words, hipStream_t s);
// out[i] = init[i] for i < n
void launch_read_init(const uint32_t *init, uint32_t n, uint64_t *out, hipStream_t s);
void launch_gen_rmat(uint64_t *src, uint64_t *dst, uint64_t n, uint64_t start, int scale,
                     uint64_t seed, double a, double b, double c, hipStream_t s);
void launch_gen_er(uint64_t *src, uint64_t *dst, uint64_t n, uint64_t start, uint64_t V,
                   uint64_t seed, hipStream_t s);
void launch_gen_weights(uint32_t *w, uint64_t n, uint64_t start, uint64_t seed, hipStream_t s);
void launch_classify(const EngineArgs &a, uint32_t n, uint8_t *unsafe, hipStream_t s);

// Safe/unsafe split of a store batch for up to MAXT trees (the classify
// step of the rt drivers' safe phase, bfs_inc_rt.cpp:211-326): an edge op is
// unsafe when it reached the engine's candidates (an insert that created a
// live entry, a delete that removed the last copy) and need_update_tree_add
// / need_update_tree_del holds for it in some tree.  The unsafe ops are
// compacted into `out` (order not kept), their count goes to *count (zeroed
// by the caller's earlier kernel) and to each tree's pinned control-block
// mirror as the engine run's n_ops; unsafe[i] (may be null) flags op i.
struct SplitParams {
    const VRec *vr[MAXT];
    int algo[MAXT];
    int undirected[MAXT];
    Ctl *hctl[MAXT];
    int ntrees;
    EdgeOps in, out;
    uint8_t *unsafe;
    uint32_t *count;
    const uint32_t *dyn_n;   // captured pipelines: the op count (else the launch's n)
};
void launch_split(const SplitParams &p, uint32_t n, hipStream_t s);
void launch_max_degree(const uint64_t *src, uint64_t n, uint64_t V, uint32_t *deg_tmp,
                       unsigned long long *best, hipStream_t s);
void launch_ops_from_stream(const uint64_t *src, const uint64_t *dst, const uint32_t *w,
                            uint64_t first, uint32_t n, uint8_t kind, EdgeOps ops,
                            uint32_t base, hipStream_t s);
// the same window read from a device BatchDesc, with a grid for up to
// n_max edge ops (captured pipelines)
void launch_ops_from_window_dyn(const uint64_t *src, const uint64_t *dst, const uint32_t *w, uint64_t len,
                                const BatchDesc *desc, uint32_t n_max, uint8_t kind_ins, uint8_t kind_del,
                                EdgeOps ops, uint32_t *zero, hipStream_t s);
// store batch launches for a captured pipeline: every kernel takes its size
// from a.dyn_n, grids sized for n_max edge ops (n_max > STORE_SMALL_OPS)
void launch_store_batch_dyn(const StoreArgs &a, uint32_t n_max, hipStream_t s);
// `zero` (may be null) is set to 0 by the launch (the split's counter)
void launch_ops_from_window(const uint64_t *src, const uint64_t *dst, const uint32_t *w, uint64_t len,
                            uint64_t ins_first, uint32_t n_ins, uint64_t del_first, uint32_t n_del,
                            uint8_t kind_ins, uint8_t kind_del, EdgeOps ops, uint32_t *zero, hipStream_t s);

// Device history of one labels instance (history.hip): a ring of records
// addressed by logical index, per-vertex chains through `prev`.
constexpr uint64_t HNIL = ~0ull;
struct HRec {
    uint64_t prev;      // older record of the same vertex (logical index), HNIL = none
    uint32_t version;
    uint32_t value;
    uint32_t vertex;
    uint32_t pad;
};
static_assert(sizeof(HRec) == 24, "history record is 24 bytes");
struct HistArgs {
    HRec *rec;
    uint64_t *head;      // [V] newest record of each vertex
    uint32_t *base;      // [V] value at the base version
    uint64_t cap;        // ring slots
    uint64_t tail;       // records below this logical index are folded
};
void launch_hist_init(const HistArgs &h, const VRec *vr, uint32_t V, hipStream_t s);
// list == nullptr: vertices [0, n)
void launch_hist_append(const HistArgs &h, const VRec *vr, const uint32_t *list, uint32_t n, uint32_t version,
                        uint64_t top, uint32_t *count, hipStream_t s);
void launch_hist_fold(const HistArgs &h, uint64_t new_tail, hipStream_t s);
void launch_hist_read(const HistArgs &h, uint32_t first, uint32_t n, uint32_t version, uint64_t *out, hipStream_t s);

}  // namespace rg
