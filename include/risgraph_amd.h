/*
 * risgraph_amd.h -- C ABI of the MI355X (gfx950) incremental graph hot path.
 *
 * This is the drop-in boundary under the Graph<> interface of the reference
 * (core/graph.hpp).  Plain pointers and sizes only; every call returns an
 * int status (RG_OK on success) and rg_last_error() holds a per-thread
 * message.  Handles own their device memory.  Calls on one graph handle are
 * serialised by the library (the reference engine is not reentrant either:
 * graph.hpp:1728-1730 shares active sets between calls).
 *
 * Vertex ids are u64 at the boundary and must be < 2^31 (the device keeps
 * u32 ids; the reference allows 48-bit ids, type.hpp:23).  Labels are u64 at
 * the boundary and must be < 2^32 (every reference algorithm's labels are:
 * 65536, 2^27 and vertex ids).  Empty parent = V (graph.hpp:90).
 */
#ifndef RISGRAPH_AMD_H
#define RISGRAPH_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define RG_OK 0
#define RG_EINVAL_VERTEX 1   /* graph.hpp:205, 231 "VertexId error." */
#define RG_ENOMEM 2
#define RG_EALGO 3           /* unrecognised algorithm */
#define RG_EDEVICE 4         /* HIP error or device-side failure */
#define RG_EINVAL 5          /* bad argument */

/* algorithms (driver lambdas: bfs_inc.cpp:77-92, sssp_inc.cpp:82-97,
 * sswp_inc.cpp:82-97, wcc_inc.cpp:82-97) */
#define RG_BFS 0
#define RG_SSSP 1
#define RG_SSWP 2
#define RG_WCC 3

/* Edge<u64> (type.hpp:55-64): src, dst, data (weight; 0 when unweighted). */
typedef struct rg_edge {
    uint64_t src, dst, data;
} rg_edge;

typedef struct rg_graph rg_graph;
typedef struct rg_labels rg_labels;
typedef struct rg_stream rg_stream;

/* Per-thread description of the last error. */
const char *rg_last_error(void);

/* Graph(vertices, hint_edges, symmetric, dual) -- graph.hpp:61-100.
 * `weighted` selects Graph<u64> (weights part of edge identity, type.hpp:32-35)
 * versus Graph<void>.  The slot pool is sized from hint_edges. */
int rg_graph_create(uint64_t vertices, uint64_t hint_edges, int symmetric, int dual,
                    int weighted, rg_graph **out);
int rg_graph_destroy(rg_graph *g);
/* vertices, slot-pool use/capacity (entries), hash-pool use/capacity */
int rg_graph_info(rg_graph *g, uint64_t *vertices, uint64_t *slots_used, uint64_t *slots_cap,
                  uint64_t *hash_used, uint64_t *hash_cap);
/* Slab-pool recycling, six counters: blocks recycled (slot pool, hash pool),
 * blocks dropped (freed past the 262,144 pending frees per size class and
 * batch, or into a full free list: never reused), and the entries those
 * dropped blocks held.  No counterpart in the reference (its per-vertex
 * vectors reallocate through the host allocator). */
int rg_graph_pool_info(rg_graph *g, uint64_t out[6]);

/* add_edge / del_edge -- graph.hpp:203-253 over IndexedEdgeStorage::update_edge
 * (storage.hpp:345-398).  delta = +1 add, -1 del.  *old = the reference return
 * value (old multiplicity; undirected: old(d->s) - (s == d) on add). */
int rg_edge_update(rg_graph *g, uint64_t src, uint64_t dst, uint64_t data, int delta, int directed,
                   uint64_t *old);
/* n edge updates with the same delta, applied as one device batch (the
 * reference batch drivers' parallel add_edge / del_edge loops,
 * bfs_inc_batch.cpp:142-173).  old_out[i] receives each return value. */
int rg_edge_update_batch(rg_graph *g, const rg_edge *edges, uint64_t n, int delta, int directed,
                         uint64_t *old_out);
/* rg_edge_update_batch for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  It
 * has no counterpart in the reference.  d_edges (n rg_edge records) and
 * d_old (n u64 return values, may be NULL) are DEVICE pointers on the
 * graph's device; the caller's writes to d_edges must be complete before the
 * call.  The batch's working arrays live in a static device array of
 * RG_STATIC_MAX_OPS edge ops' worth (about 5 MB per device, held until the
 * process ends); nothing is allocated and no hipMalloc / hipFree is called.
 * n > RG_STATIC_MAX_OPS returns RG_EINVAL.  A batch holding an id >= V or a
 * weight over 32 bits is rejected whole (RG_EINVAL_VERTEX / RG_EINVAL).
 * Calls are synchronous; every call on a device shares that device's array
 * and leaves its contents behind, so calls that overlapped on one device
 * would race: the library serialises them with a process-wide lock. */
#define RG_STATIC_MAX_OPS 16384
int rg_edge_update_batch_static(rg_graph *g, const rg_edge *d_edges, uint64_t n, int delta, int directed,
                                uint64_t *d_old);
/* get_outgoing_degree / get_incoming_degree (graph.hpp:155-164) and
 * get_edge_num (graph.hpp:198-201). */
int rg_degree(rg_graph *g, uint64_t v, int incoming, uint64_t *out);
int rg_edge_num(rg_graph *g, uint64_t src, uint64_t dst, uint64_t data, uint64_t *out);

/* alloc_vertex_tree_array (graph.hpp:255-262) bound to one algorithm.
 * root is used by BFS/SSSP/SSWP; init_values (length V, may be NULL) gives
 * WCC labels (all_inc.cpp:172-178 uses a permutation; NULL = identity).
 * trace_modified keeps the set of vertices whose label changed
 * (graph.hpp:117-130).  Labels are all "empty" until rg_build. */
int rg_labels_create(rg_graph *g, int algo, uint64_t root, const uint64_t *init_values,
                     int trace_modified, rg_labels **out);
int rg_labels_destroy(rg_labels *L);
/* Strategy knobs of do_update_tree_del (graph.hpp:1252, 1324, 1900) plus the
 * largest pass (queued vertices, or edge ops / 16) that runs inside the
 * persistent team launch; larger passes run grid-wide.  A level is rescued
 * when level <= rescue_level (0xFFFFFFFF: never) or its frontier is below
 * rescue_frontier.  Defaults: rescue on the first level only (0, 0) for
 * BFS / SSSP and on every level (0, 0xFFFFFFFF) for SSWP and WCC; the
 * reference's policy is (5, 16384).  Full recompute when level <= 10 and
 * subtree+frontier > max(0.05*V, 2000000); ancestor walks capped at
 * walk_limit steps (0 = min(V/10, 65536)); block_max = RG_AUTO (1024 per
 * team workgroup; the team width grows with the batch, 8 to 64 workgroups
 * per tree).  block_max = 0 runs every pass grid-wide.  Values never depend
 * on these; only the work done does. */
#define RG_AUTO 0xFFFFFFFFu
int rg_labels_set_tuning(rg_labels *L, uint32_t rescue_level, uint32_t rescue_frontier,
                         uint32_t recompute_level, uint32_t recompute_abs, double recompute_frac,
                         uint32_t walk_limit, uint32_t block_max);
/* Direction of a build's passes (stream_edges, graph.hpp:323-394): with
 * mode 1 a push pass whose frontier's live out-edges reach both V and a
 * twentieth of the live edges runs dense instead, every vertex pulling its
 * best candidate over its in-edges (graph.hpp:380-394); mode 0 (the
 * default, also RG_DENSE=0; RG_DENSE=1 makes 1 the default) keeps builds
 * push-only, which measured faster on MI355X; -1 leaves the mode.  Returns in dense_passes the
 * dense passes this instance's builds and rebuilds ran so far.  Values never
 * depend on the mode. */
int rg_labels_dense(rg_labels *L, int mode, uint64_t *dense_passes);
/* The tree's share of a batch's persistent repair launch: the launch's
 * eight XCD slots (each a 1/8 of its workgroups) are split over its trees by
 * weight, one slot each at least.  weight 0 (default) = by algorithm (BFS 1,
 * SSSP 3, SSWP 2, WCC 2), else 1-64.  No counterpart in the reference (its
 * trees share the host's OpenMP threads); values never depend on it. */
int rg_labels_team_weight(rg_labels *L, uint32_t weight);

/* build_tree -- graph.hpp:873-985: full fixpoint from the init labels. */
int rg_build(rg_labels *L, uint64_t *result);
/* update_tree_add (single and batch) -- graph.hpp:1124-1218.  The edges are
 * already in the store (the caller ran rg_edge_update and saw old == 0). */
int rg_inc_add(rg_labels *L, const rg_edge *edges, uint64_t n, int directed, uint64_t *result);
/* update_tree_del (single and batch) -- graph.hpp:1622-1718.  The edges are
 * already removed from the store (old == 1). */
int rg_inc_del(rg_labels *L, const rg_edge *edges, uint64_t n, int directed, uint64_t *result);
/* need_update_tree_add / need_update_tree_del (graph.hpp:987-1000,
 * 1220-1233) for n edges at once: unsafe[i] = 1 when edge i would change a
 * label (kind[i] = +1 add, -1 del).  Reads the current labels. */
int rg_classify(rg_labels *L, const rg_edge *edges, const int8_t *kind, uint64_t n, int directed,
                uint8_t *unsafe);

/* labels[first .. first+count): value and parent vertex (V = empty). */
int rg_labels_read(rg_labels *L, uint64_t first, uint64_t count, uint64_t *values, uint64_t *parents);
/* get_modified() (graph.hpp:117-125): vertices whose label changed since the
 * last clear.  *n = count written (<= cap); *all = 1 after a full rebuild
 * (every vertex may have changed). */
int rg_modified_read(rg_labels *L, uint64_t *vids, uint64_t cap, uint64_t *n, int *all);
/* clear_modified() (graph.hpp:127-130). */
int rg_modified_clear(rg_labels *L);

/* A batch of transactions from host buffers (the transaction server of
 * all_inc_rpc_txn.cpp; ClientTxnUpdateRequest, rpc.hpp:40-61): ntxn
 * transactions of txn_len[t] updates each (1..RG_TXN_MAX), the updates laid
 * out one transaction after another in edges / kind (+1 add, -1 del).  All
 * of them go to the store as one device batch (inserts before deletes, the
 * batch drivers' order, bfs_inc_batch.cpp:142-178), then the updates are
 * classified against the labels of every tree in L[0..nL) (1 <= nL <= 4):
 * a transaction is unsafe when any of its updates is (all_inc_rpc_txn.cpp:
 * is_safe is the AND over its parts), and only unsafe updates seed the
 * repair.  txn_unsafe[t] (may be NULL) receives that flag and txn_version[t]
 * (may be NULL) the version at which transaction t takes effect: the
 * version before the batch for a safe transaction, the one the batch ends
 * at for an unsafe one; one version per transaction either way. */
#define RG_TXN_MAX 16
int rg_txn_apply(rg_graph *g, rg_labels **L, int nL, const rg_edge *edges, const int8_t *kind, uint64_t n,
                 const uint32_t *txn_len, uint64_t ntxn, int directed, uint8_t *txn_unsafe, uint64_t *txn_version);

/* Durability (the reference's WAL, all_inc_rpc.cpp:305-336, 572-590, plus
 * the snapshot and replay it lacks).
 * rg_wal_open: every later rg_txn_apply batch is appended to the file at
 * `path` (created when absent; an existing log of the same graph shape is
 * appended to) as one record of 32-byte entries before it is applied, and
 * with sync = 1 fdatasync'ed first (group commit: the whole batch's
 * transactions are durable before any of them is acknowledged).
 * rg_wal_replay applies, in order, the logged batches whose version is
 * above the graph's (the ones a snapshot does not hold) and stops at the
 * first record that is cut short or fails its checksum; *batches (may be
 * NULL) counts the batches applied.  Only rg_txn_apply batches are logged. */
int rg_wal_open(rg_graph *g, const char *path, int sync);
int rg_wal_close(rg_graph *g);
int rg_wal_replay(rg_graph *g, rg_labels **L, int nL, const char *path, uint64_t *batches);
/* rg_snapshot_save writes the graph's shape and version, the live edge
 * multiset of its out-lists and, for each labels instance in L[0..nL), its
 * algorithm, root, WCC init labels, values and parents (written to
 * path.tmp, fsync'ed, renamed).  rg_snapshot_load makes a new graph and
 * labels from such a file (at most `cap` labels into L; *nL = how many),
 * at the snapshot's version; the graph's edge lists are rebuilt from the
 * multiset, so their entry order (and dropped tombstones) may differ. */
int rg_snapshot_save(rg_graph *g, rg_labels **L, int nL, const char *path);
int rg_snapshot_load(const char *path, rg_graph **g, rg_labels **L, int cap, int *nL);
/* A device-resident stream made from host edges (an edge file read with
 * mmap_binary, io.hpp:26-36, or a snapshot's edges); weights from data. */
int rg_stream_from_edges(const rg_edge *edges, uint64_t n, int weighted, rg_stream **out);
/* Stream batches of more than 64 edge ops with trees to repair run their
 * window, store and split launches as one captured hipGraph per batch-size
 * bucket (powers of two from 128), replayed with the batch's size in a
 * device descriptor; on = 0 launches them one by one.  The default is on
 * unless the environment sets RG_GRAPH=0.  No counterpart in the reference
 * (its batches are host loops); results are the same either way. */
int rg_graph_capture(rg_graph *g, int on);
/* Whether capture is on (it turns off by itself where the runtime cannot
 * capture), the pipelines captured so far and the batches replayed. */
int rg_graph_capture_info(rg_graph *g, int *on, uint32_t *pipelines, uint64_t *replays);
/* Autotuner of the batch engine (no counterpart in the reference, whose
 * knobs are compile-time constants).  For four batch sizes (16,384, 65,536,
 * 262,144 and 1,048,576 edge ops, half inserts of stream window
 * [ins_first, ...) and half deletes of window [del_first, ...)) it times
 * forward batches and their inverses (the deleted window inserted back, the
 * inserted one deleted: the graph and the labels' values return to where
 * they were) under candidate settings and keeps, per batch-size bucket
 * (<= 32k, <= 128k, <= 512k, more edge ops), the fastest: the team cap
 * (passes larger than cap_x4/4 x 1024 vertices, or cap_x4/4 x 1024 chunk
 * items, per team workgroup go grid-wide) and the team weights of the four
 * algorithms.  Later batches of the graph run with the table.  With
 * target_ms > 0 it also returns batch_cap, the largest serving batch whose
 * measured device time, doubled (an update of a capped batch also waits for
 * the batch before it), fits a quarter of the target (the serving loop's
 * own shrink rule), for
 * rg_serve_params.cap_max.  Every edge the deleted window holds must be
 * live (it is when the windows are the stream's current window, as in the
 * serving loop); the degrees are checked afterwards and a graph that did not
 * return is reported as RG_EINVAL.  The graph's version advances by
 * `batches`, and the labels' modified sets hold what the batches touched. */
#define RG_TUNE_BUCKETS 4
typedef struct rg_autotune_params {
    uint64_t ins_first, del_first;
    int directed;
    uint32_t pairs;      /* timed forward/inverse pairs per setting (0: 3) */
    double target_ms;    /* serving target for batch_cap (0: not chosen) */
} rg_autotune_params;
typedef struct rg_autotune_result {
    uint32_t ops[RG_TUNE_BUCKETS];          /* edge ops of the timed batches */
    uint32_t cap_x4[RG_TUNE_BUCKETS];       /* chosen team cap, x 1/4 of the default */
    uint32_t weight[RG_TUNE_BUCKETS][4];    /* chosen team weights (BFS, SSSP, SSWP, WCC) */
    double default_ms[RG_TUNE_BUCKETS];     /* device ms of a pair, default settings */
    double best_ms[RG_TUNE_BUCKETS];        /* device ms of a pair, chosen settings */
    uint32_t batch_cap;
    uint64_t batches;
} rg_autotune_result;
int rg_autotune(rg_graph *g, rg_labels **L, int nL, rg_stream *s, const rg_autotune_params *p,
                rg_autotune_result *r);
/* The shape a graph was created with, and a labels instance's algorithm and
 * root (for handles that rg_snapshot_load made). */
int rg_graph_shape(rg_graph *g, int *symmetric, int *dual, int *weighted);
int rg_labels_algo(rg_labels *L, int *algo, uint64_t *root);
/* Overwrite a labels instance with given values and parent vertices
 * (parent V = empty); the modified set then reports every vertex. */
int rg_labels_assign(rg_labels *L, const uint64_t *values, const uint64_t *parents);

/* Request server: the network front-end of the reference's RPC servers
 * (all_inc_rpc.cpp:397-542, all_inc_rpc_txn.cpp; UCX there, TCP on
 * 127.0.0.1 here) feeding a device request loop.  Requests are the
 * ClientUpdateRequest / ClientTxnUpdateRequest fields of rpc.hpp:19-61 in a
 * packed little-endian layout:
 *   update (txn = 0), 40 bytes: u32 client_id, u32 type (0 add, 1 del,
 *     2 end), u64 src, u64 dst, u64 data, i64 request_time;
 *   transaction (txn = 1), 464 bytes: u32 client_id, u32 num_updates (1..16),
 *     u32 types[16], u64 srcs[16], u64 dsts[16], u64 datas[16],
 *     i64 request_time;
 * replies: update, 24 bytes: u32 client_id, u32 status (0 or an RG_* code),
 *   u64 version_id, i64 request_time (echoed); transaction, 144 bytes: u32
 *   client_id, u32 status, u64 version_ids[16], i64 request_time.
 * Every request that arrived while the device was busy is admitted as one
 * batch (up to cap_max updates; 0 = 2^20): logged when the graph has a WAL
 * open (rg_wal_open), stored, split into safe and unsafe requests against
 * the trees in L[0..nL) (1 <= nL <= 4); safe requests are answered as soon
 * as the split is on the host, with the version before the batch, unsafe
 * ones after the repair, with the version after it.  A type-2 request closes
 * its connection.  port 0 picks a free port (rg_server_port).  While the
 * server runs, other calls on the graph wait for the batch in progress. */
typedef struct rg_server rg_server;
typedef struct rg_server_params {
    uint16_t port;
    int txn;                  /* 1: transaction requests, else single updates */
    int directed;
    uint32_t cap_max;
} rg_server_params;
int rg_server_start(rg_graph *g, rg_labels **L, int nL, const rg_server_params *p, rg_server **out);
int rg_server_port(rg_server *s, uint32_t *port);
/* requests and updates served, device batches, unsafe requests; RG_EDEVICE
 * when a batch failed (its requests got the error code as status) */
int rg_server_stats(rg_server *s, uint64_t *requests, uint64_t *updates, uint64_t *batches, uint64_t *unsafe);
/* stops accepting, answers what is queued, closes every connection */
int rg_server_stop(rg_server *s);

/* Version clock of the device batches: every rg_stream_apply and
 * rg_txn_apply batch (and every batch of rg_serve) that applies at least one
 * update advances it by one.  The safe updates of a split batch take effect at the version before
 * it, its unsafe ones at the version it ends (the rt drivers' label_version,
 * bfs_inc_rt.cpp:340-390). */
int rg_graph_version(rg_graph *g, uint64_t *version);

/* Device-resident value history (MVVec, mvvec.hpp:23-162, filled per batch
 * as bfs_inc_rt.cpp:353-358 fills it from the modified set).  enable takes a
 * snapshot of the current values as the base at the current version and
 * keeps a ring of `records` (version, value) records (24 bytes each); after
 * every batch each vertex whose value changed gets one record of the
 * batch's version.  The labels must have been created with trace_modified;
 * the modified set then covers the last batch only.  When the ring is full
 * the oldest batches are folded into the base (MVVec::gc); a batch changing
 * more vertices than the ring holds fails with RG_ENOMEM.
 * read: each vertex's value at `version` (MVVec::get), for versions from
 * the oldest kept (rg_history_info) to the current one; others RG_EINVAL.
 * gc: fold every batch of version <= `version` (reads at `version` and
 * later stay exact).  Only stream / serve batches are recorded; the
 * host-buffer calls (rg_edge_update*, rg_inc_*, rg_build) are not. */
int rg_history_enable(rg_labels *L, uint64_t records);
int rg_history_read(rg_labels *L, uint64_t version, uint64_t first, uint64_t count, uint64_t *values);
int rg_history_gc(rg_labels *L, uint64_t version);
int rg_history_info(rg_labels *L, uint64_t *oldest_version, uint64_t *records, uint64_t *capacity);

/* Device-resident synthetic streams (edge arrays that never leave HBM).
 * kind 0 = R-MAT (scale_or_v = scale, a,b,c probabilities), 1 = Erdos-Renyi
 * (scale_or_v = V).  Weights are generated when wseed != 0. */
int rg_stream_create(int kind, uint64_t scale_or_v, uint64_t n_edges, uint64_t seed, double a,
                     double b, double c, uint64_t wseed, rg_stream **out);
int rg_stream_destroy(rg_stream *s);
/* copy stream edges [first, first+n) to host arrays (any may be NULL) */
int rg_stream_read(rg_stream *s, uint64_t first, uint64_t n, uint64_t *src, uint64_t *dst, uint32_t *w);
/* vertex with the largest out-degree among stream edges [0, n) (ties: lowest id) */
int rg_stream_max_degree_vertex(rg_stream *s, uint64_t n, uint64_t V, uint64_t *out);

/* Bulk add_edge of stream edges [first, first+n) (the drivers' load loop,
 * bfs_inc.cpp:59-69).  No labels are touched. */
int rg_stream_load(rg_graph *g, rg_stream *s, uint64_t first, uint64_t n, int directed);

/* One sliding-window batch on device-resident data, the fused form of a
 * batch driver round (bfs_inc_batch.cpp:142-178): add stream edges
 * [ins_first, ins_first+n_ins), delete [del_first, del_first+n_del), then
 * repair every labels instance in L[0..nL).  Ranges wrap around the end of
 * the stream (positions are taken mod the stream length), so a sliding
 * window can run past it: update k inserts e[(n0+k) mod n] and deletes
 * e[k mod n].  n_ins = n_del = 1 is one plain update (bfs_inc.cpp:160-189).
 * The repair of a mixed batch handles the deletions first and seeds the
 * inserts in the same pass as the deletion pull; values are the unique
 * fixpoint either way. */
typedef struct rg_batch_stats {
    uint64_t seeds;          /* insert edges that reached the engine */
    uint64_t roots;          /* delete edges whose multiplicity hit 0 */
    uint64_t levels;         /* engine levels run (all labels) */
    uint64_t grid_steps;     /* passes that ran grid-wide */
    uint64_t rebuilds;       /* full recomputes taken by the deletion fallback */
} rg_batch_stats;
int rg_stream_apply(rg_graph *g, rg_labels **L, int nL, rg_stream *s, uint64_t ins_first,
                    uint64_t n_ins, uint64_t del_first, uint64_t n_del, int directed,
                    rg_batch_stats *stats);

/* Run `n_steps` plain updates back to back (step k: add ins_first+k, then
 * delete del_first+k, positions mod the stream length, each followed by its
 * repair), the per-update GPU path
 * of bfs_inc.cpp:160-189.  lat_ns (may be NULL) receives each step's wall
 * time in ns. */
int rg_stream_plain(rg_graph *g, rg_labels **L, int nL, rg_stream *s, uint64_t ins_first,
                    uint64_t del_first, uint64_t n_steps, int directed, double *lat_ns,
                    rg_batch_stats *stats);

/* Device timing of the most recent rg_stream_apply batch: GPU milliseconds
 * between events recorded before its store kernels and after its last
 * engine kernel. */
int rg_last_gpu_ms(rg_graph *g, double *ms);
/* Edge ops of the last stream batch that the split found unsafe (all of
 * them for a batch that was not split). */
int rg_last_unsafe(rg_graph *g, uint64_t *n);

/* Kernel timing with HIP events on the streams the kernels run on.  Bit k of
 * `mask` times kernels of kind k: the store apply kernels (kind 0), the
 * persistent team engine launches (kind 1) and grid-wide engine passes
 * (kind 2).  rg_profile_read returns the launch count, summed device
 * milliseconds and summed work units (edge ops for kind 0, trees for kind 1,
 * pass work for kind 2) accumulated since the graph was made; kind 3 + p
 * reads the grid-wide passes of engine phase p (the phase numbers of
 * rg_labels_phase_time) alone.  Kind 0 counts batches of more than 64 edge
 * ops (the apply kernel k_apply_all); batches of at most 64 run the whole store
 * step in one kernel, timed under mask bit 0 and read as kind 13. */
int rg_profile_enable(rg_graph *g, int mask);
int rg_profile_read(rg_graph *g, int kind, uint64_t *launches, double *ms, uint64_t *units);

/* Per-phase device time of the persistent team's engine passes on one
 * labels instance since its creation: us[p] microseconds and passes[p] for
 * phase p (0 done, 1 seed, 2 push, 3 delete roots, 4 rescue + expand level,
 * 5 expand-only level, 6 unused, 7 reset + pull (+ seeds of a mixed batch),
 * 8 build init, 9 rebuild); arrays of 10. */
int rg_labels_phase_time(rg_labels *L, double *us, uint64_t *passes);

/* Where the persistent team's passes spend their time, per phase p (arrays
 * as above): stage_us[k*10 + p] for k = 0 vertex stage, 1 first barrier wait,
 * 2 chunk stage, 3 second barrier wait (as seen by the tree's first
 * workgroup); work[p] summed pass work (queued vertices, or ops/16);
 * chunks[p] summed chunk items.  Any pointer may be NULL. */
int rg_labels_phase_detail(rg_labels *L, double *stage_us, uint64_t *work, uint64_t *chunks);

/* Live multiplicity degree of every vertex (out, and in for directed
 * graphs; either pointer may be NULL), the `degree` of IndexedEdgeStorage
 * (storage.hpp:396). */
int rg_degrees(rg_graph *g, uint32_t *out_deg, uint32_t *in_deg);

/* Serving loop: the device-side replacement of the latency-aware scheduler
 * of the *_inc_rt drivers (bfs_inc_rt.cpp:154-431).  Updates u = 0..n-1
 * alternate add (stream edge ins_first + u/2) and delete (del_first + u/2),
 * positions mod the stream length,
 * the all_inc stream (all_inc.cpp:254-300), and arrive by a Poisson process
 * of `rate` updates/s from `clients` simulated clients (client u % clients).
 * Every loop iteration admits all arrived updates up to the current batch
 * cap (the admission budget), applies them as one device batch to the store
 * and every labels instance in L, and completes them together.  With
 * target_ms > 0 the cap adapts in the spirit of the reference's
 * num_changes_threshold (bfs_inc_rt.cpp:409-427), but on device service
 * time: x1.5+16 while the backlog exceeds the cap, otherwise x0.7 when a
 * batch took longer than target/4, clamped to [1, cap_max].  batch_cap = 1 with no
 * target is the plain one-update-at-a-time path.  When the oldest waiting
 * request is older than max(10 * target, 200 ms) the run is overloaded: the
 * rest is applied untimed in cap_max batches and r->overloaded is set.
 * Latency = completion time - arrival time (no coordinated omission).
 * The per-update bookkeeping of a split batch (flags, latencies) runs on
 * four host threads of the call's own while the calling thread drives the
 * batch's repair. */
typedef struct rg_serve_params {
    uint64_t ins_first, del_first, n_updates;
    double rate;
    uint64_t seed;
    uint32_t clients;
    uint32_t batch_cap;       /* initial cap (1 = plain path) */
    uint32_t cap_max;
    double target_ms;         /* 0 = fixed cap */
    int directed;
    /* 1: safe/unsafe split (the safe phase of bfs_inc_rt.cpp:211-326): a
     * batch of more than 64 updates is classified against the labels before
     * its repair (need_update_tree_add / _del in every tree, store returns
     * applied); its safe updates complete as soon as the store batch and the
     * split are done, only the unsafe ones seed the repair and wait for it */
    int split;
    /* 1: the admission cap counts unsafe updates (the rt drivers'
     * num_changes_threshold, bfs_inc_rt.cpp:409-427): a batch admits
     * cap / f updates, f = the running unsafe fraction (needs split) */
    int cap_unsafe;
} rg_serve_params;
typedef struct rg_serve_result {
    double wall_s, updates_per_s;
    double p50_ms, p99_ms, p999_ms, max_ms, mean_ms;
    double mean_batch;
    uint64_t batches, max_batch, grid_steps, rebuilds;
    uint64_t overloaded;      /* 1 when the run fell behind (see above) */
    uint64_t served;          /* updates completed with a measured latency */
    double gpu_ms;            /* summed device time of the batches */
    uint64_t unsafe;          /* served updates that waited for a repair */
    double safe_frac;         /* served updates that completed at the split */
    double p50_safe_ms, p50_unsafe_ms, p999_unsafe_ms;
} rg_serve_result;
int rg_serve(rg_graph *g, rg_labels **L, int nL, rg_stream *s, const rg_serve_params *p,
             rg_serve_result *r, double *lat_ms);

#ifdef __cplusplus
}
#endif
#endif
