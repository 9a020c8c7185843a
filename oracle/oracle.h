/*
 * oracle.h -- CPU oracle for the RisGraph incremental hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker / the timed CPU port.
 *
 * Three parts:
 *   1. Static closed forms on a live multigraph (BFS levels, min-plus
 *      distances, max-min widths, min-label components) with the reference
 *      sentinels.  These are the "full recompute" check: every monotone
 *      relaxation in the reference has a unique fixpoint, so the incremental
 *      engine, its build_tree and these textbook algorithms agree on values.
 *   2. A dependency-tree validity checker.
 *   3. "cpu_ref": a sequential C restatement of the reference incremental
 *      engine (store + build_tree + update_tree_add/del with rescue), used as
 *      the timed CPU baseline and as a second value oracle per update.
 *
 * Parity status of each function is listed in the header of oracle.c.
 */
#ifndef RG_ORACLE_H
#define RG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_BFS = 0, OR_SSSP = 1, OR_SSWP = 2, OR_WCC = 3 };

/* Sentinels of the reference drivers: bfs_inc.cpp:72 (65536),
 * sssp_inc.cpp:72 / sswp_inc.cpp:72 (134217728). */
uint64_t or_maxl(int algo);

/* Init label of the reference drivers (bfs_inc.cpp:89-92, sssp_inc.cpp:94-97,
 * sswp_inc.cpp:94-97, wcc_inc.cpp:94-97).  For WCC, `init` (length V) may
 * override the identity labelling (all_inc.cpp:172-178); NULL = identity. */
uint64_t or_init_value(int algo, uint64_t v, uint64_t root, const uint64_t *init);

/* 1. Static closed form on the directed multiset of live edges
 * (src[i] -> dst[i], weight w[i]; w may be NULL for BFS/WCC).  WCC treats
 * every edge as undirected.  Writes V values.  Returns 0 or -1 on bad input. */
int or_static_values(int algo, uint64_t V, uint64_t m, const uint64_t *src,
                     const uint64_t *dst, const uint64_t *w, uint64_t root,
                     const uint64_t *init, uint64_t *values);

/* 2. Tree validity: for every v, value == init(v) <=> parent == V (empty);
 * otherwise a live edge (parent -> v, w) (either direction when
 * `undirected`) with f(value[parent], w) == value[v] exists, and the parent
 * chain ends at an empty parent without a cycle.  Returns 0 when valid, else
 * 1 + the first offending vertex (as int64). */
int64_t or_check_tree(int algo, int undirected, uint64_t V, uint64_t m,
                      const uint64_t *src, const uint64_t *dst,
                      const uint64_t *w, uint64_t root, const uint64_t *init,
                      const uint64_t *values, const uint64_t *parents);

/* 3. cpu_ref: sequential restatement of core/graph.hpp's Graph<> engine. */
typedef struct or_graph or_graph;
typedef struct or_labels or_labels;

or_graph *or_graph_create(uint64_t V, uint64_t hint_edges, int symmetric,
                          int dual, int weighted);
void or_graph_destroy(or_graph *g);
/* add/del with the reference return value (old multiplicity). */
uint64_t or_add_edge(or_graph *g, uint64_t s, uint64_t d, uint64_t w, int directed);
uint64_t or_del_edge(or_graph *g, uint64_t s, uint64_t d, uint64_t w, int directed);
/* Bulk import of n edges, same result as n sequential or_add_edge calls. */
void or_add_edges(or_graph *g, uint64_t n, const uint64_t *src,
                  const uint64_t *dst, const uint64_t *w, int directed);
/* Bulk import into an EMPTY graph: same store as n sequential adds (list
 * entries in (nbr, data) order).  Returns -1 if the graph is not empty or an
 * id is out of range. */
int or_bulk_load(or_graph *g, uint64_t n, const uint64_t *src, const uint64_t *dst,
                 const uint64_t *w, int directed);
uint64_t or_degree(or_graph *g, uint64_t v, int incoming);
uint64_t or_edge_num(or_graph *g, uint64_t s, uint64_t d, uint64_t w);
/* Number of (src,dst,w) entries with num>0 (out store); fills arrays when non-NULL. */
uint64_t or_live_edges(or_graph *g, uint64_t *src, uint64_t *dst, uint64_t *w);

or_labels *or_labels_create(or_graph *g, int algo, uint64_t root, const uint64_t *init);
void or_labels_destroy(or_labels *L);
uint64_t or_build_tree(or_graph *g, or_labels *L);
int or_need_update_add(or_graph *g, or_labels *L, uint64_t s, uint64_t d, uint64_t w, int directed);
int or_need_update_del(or_graph *g, or_labels *L, uint64_t s, uint64_t d, uint64_t w, int directed);
uint64_t or_update_add(or_graph *g, or_labels *L, uint64_t s, uint64_t d, uint64_t w, int directed);
uint64_t or_update_del(or_graph *g, or_labels *L, uint64_t s, uint64_t d, uint64_t w, int directed);
uint64_t or_update_add_batch(or_graph *g, or_labels *L, uint64_t n, const uint64_t *s,
                             const uint64_t *d, const uint64_t *w, int directed);
uint64_t or_update_del_batch(or_graph *g, or_labels *L, uint64_t n, const uint64_t *s,
                             const uint64_t *d, const uint64_t *w, int directed);
/* values[V]; parents[V] = parent vertex (V = empty); pweights may be NULL. */
void or_labels_read(or_labels *L, uint64_t *values, uint64_t *parents, uint64_t *pweights);
/* Number of full recomputes the deletion fallback took (graph.hpp:1252). */
uint64_t or_recompute_count(or_labels *L);
/* Start from a known fixpoint instead of build_tree: values and parent
 * vertices (V = empty); parent weights are resolved from the store.  Returns
 * 0, or 1 + the first vertex whose parent has no consistent live edge. */
int64_t or_labels_assign(or_graph *g, or_labels *L, const uint64_t *values, const uint64_t *parents);

/* One plain sliding-window step of the reference drivers (bfs_inc.cpp:160-189):
 * add (s1,d1,w1); if old==0 update_tree_add; del (s2,d2,w2); if old==1
 * update_tree_del.  Runs `n` steps over the given arrays; returns the number
 * of steps whose insert or delete ran incremental work. */
uint64_t or_stream_plain(or_graph *g, or_labels *L, uint64_t n,
                         const uint64_t *as, const uint64_t *ad, const uint64_t *aw,
                         const uint64_t *ds, const uint64_t *dd, const uint64_t *dw,
                         int directed, double *lat_ns);

/* The all_inc alternation over four trees: BFS/SSSP/SSWP labels (Ld[0..nLd))
 * on directed graph gd and WCC (Lu) on symmetric graph gu.  Update u adds
 * edge u/2 (even u) or deletes edge u/2 (odd u).  lat_ns[u] = wall ns. */
uint64_t or_stream_mixed(or_graph *gd, or_labels **Ld, int nLd, or_graph *gu, or_labels *Lu,
                         uint64_t n_updates, const uint64_t *as, const uint64_t *ad, const uint64_t *aw,
                         const uint64_t *ds, const uint64_t *dd, const uint64_t *dw, double *lat_ns);

/* The rt driver's serving loop (bfs_inc_rt.cpp:154-431) on host threads
 * (OpenMP, omp_get_max_threads()) over the all_inc stream's four trees
 * (BFS/SSSP/SSWP labels Ld on gd, WCC Lu on gu; update u adds edge u/2 for
 * even u, deletes it for odd u).  Update u arrives at arrival[u] seconds
 * after the call starts.  Rounds: a parallel safe phase classifies arrived
 * updates (need_update_tree_add / _del in every tree), applies the safe
 * ones to the store under per-list locks and completes them, queues the
 * unsafe ones until their count passes the threshold or the oldest has
 * waited 0.8 x timeout; then a serial phase applies the queued ones in
 * arrival order with the incremental engine.  The threshold starts at the
 * thread count and moves by the reference's rule every 3 rounds.
 * lat_ms[u] = completion - arrival.  stats[7]: served, rounds, unsafe,
 * final threshold, overloaded (a request waited 10 x timeout), threads,
 * wall ns.  Returns the number served. */
uint64_t or_serve_rt(or_graph *gd, or_labels **Ld, int nLd, or_graph *gu, or_labels *Lu, uint64_t n_updates,
                     const uint64_t *as, const uint64_t *ad, const uint64_t *aw, const uint64_t *ds,
                     const uint64_t *dd, const uint64_t *dw, const double *arrival, double timeout_ms,
                     double target_timeout_rate, double *lat_ms, uint64_t *stats);

#ifdef __cplusplus
}
#endif
#endif
