/*
 * oracle.c -- CPU oracle for the RisGraph incremental hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the shipped library never links or calls this
 * file.  tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg are
 * the only users, as the checker and as the timed CPU port.
 *
 * Parity pinning (what checks each function, see tests/test_oracle.py):
 *   or_static_values  BFS / SSSP / WCC against scipy.sparse.csgraph
 *                     (breadth_first_order, dijkstra, connected_components)
 *                     and hand-derived golden vectors in tests/golden/;
 *                     SSWP against the threshold characterisation
 *                     (width(v) >= t  <=>  v reachable from the root over
 *                     edges of weight >= t, checked with scipy BFS for every
 *                     distinct t) and brute-force path enumeration on tiny
 *                     graphs.
 *   or_check_tree     planted faults (wrong parent, missing edge, cycle,
 *                     wrong value, non-empty parent at init) must be caught.
 *   or_graph_* store  brute-force Python multiset model (old-count return,
 *                     clamping at zero, undirected self-loop rule,
 *                     weighted identity (nbr, w)), plus hub lists >= 512.
 *   or_need_update_add / or_need_update_del
 *                     an add is unsafe exactly when the static fixpoint of
 *                     G + e differs from G's; a delete of a single-copy edge
 *                     is unsafe exactly when it is some vertex's parent edge
 *                     (the tree's edges stop being live on G - e), and a
 *                     safe delete leaves the values and a valid tree; all
 *                     four algorithms, directed and undirected; planted
 *                     one-direction and reversed-comparison mistakes fail.
 *   or_serve_rt       the rt loop's final values equal the closed form on the
 *                     final window for all four trees (valid trees too),
 *                     under 1, 2, 4 and 8 threads; every update served.
 *   or_build_tree /
 *   or_update_*       equality with or_static_values after every update of
 *                     seeded sliding-window streams (all four algorithms,
 *                     single and batched), and tree validity each time.
 *
 * The restated engine follows core/graph.hpp of the reference:
 *   add_edge/del_edge           graph.hpp:203-253
 *   IndexedEdgeStorage          storage.hpp:282-398
 *   build_tree                  graph.hpp:873-985 (+ stream_edges 323-396)
 *   update_tree_add (+batch)    graph.hpp:1124-1218, do_update_tree_add 1002-1116
 *   update_tree_del (+batch)    graph.hpp:1622-1718, do_update_tree_del 1235-1612
 *   update_label (CAS form)     graph.hpp:1861-1887
 *   has_parents                 graph.hpp:1889-1907
 * It is sequential: the reference's parallel loops become plain loops,
 * which is the reference's own behaviour below OPENMP_THRESHOLD (type.hpp:105).
 */
#define _POSIX_C_SOURCE 199309L
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* algorithm lambdas (bfs_inc.cpp:77-92, sssp_inc.cpp:82-97,            */
/* sswp_inc.cpp:82-97, wcc_inc.cpp:82-97)                               */
/* ------------------------------------------------------------------ */

uint64_t or_maxl(int algo)
{
    switch (algo) {
    case OR_BFS: return 65536ull;
    case OR_SSSP:
    case OR_SSWP: return 134217728ull;
    default: return 0;
    }
}

uint64_t or_init_value(int algo, uint64_t v, uint64_t root, const uint64_t *init)
{
    switch (algo) {
    case OR_BFS:
    case OR_SSSP: return v == root ? 0 : or_maxl(algo);
    case OR_SSWP: return v == root ? or_maxl(algo) : 0;
    default: return init ? init[v] : v;
    }
}

static int init_active(int algo, uint64_t v, uint64_t root)
{
    return algo == OR_WCC ? 1 : v == root;
}

/* candidate value pushed along an edge of weight w */
static inline uint64_t cand(int algo, uint64_t s, uint64_t w)
{
    switch (algo) {
    case OR_BFS: return s + 1;
    case OR_SSSP: return s + w;
    case OR_SSWP: return s < w ? s : w;
    default: return s;
    }
}

/* update_func(...).first: strict improvement */
static inline int improves(int algo, uint64_t c, uint64_t d)
{
    return algo == OR_SSWP ? c > d : c < d;
}

/* ------------------------------------------------------------------ */
/* 1. static closed forms                                              */
/* ------------------------------------------------------------------ */

typedef struct { uint64_t *off; uint64_t *nbr; uint64_t *w; } csr_t;

static void csr_build(csr_t *c, uint64_t V, uint64_t m, const uint64_t *src,
                      const uint64_t *dst, const uint64_t *w, int both)
{
    uint64_t mm = both ? 2 * m : m;
    c->off = calloc(V + 1, sizeof(uint64_t));
    c->nbr = malloc((mm ? mm : 1) * sizeof(uint64_t));
    c->w = malloc((mm ? mm : 1) * sizeof(uint64_t));
    for (uint64_t i = 0; i < m; i++) {
        c->off[src[i] + 1]++;
        if (both) c->off[dst[i] + 1]++;
    }
    for (uint64_t v = 0; v < V; v++) c->off[v + 1] += c->off[v];
    uint64_t *pos = malloc((V + 1) * sizeof(uint64_t));
    memcpy(pos, c->off, (V + 1) * sizeof(uint64_t));
    for (uint64_t i = 0; i < m; i++) {
        uint64_t wi = w ? w[i] : 0;
        uint64_t p = pos[src[i]]++;
        c->nbr[p] = dst[i];
        c->w[p] = wi;
        if (both) {
            p = pos[dst[i]]++;
            c->nbr[p] = src[i];
            c->w[p] = wi;
        }
    }
    free(pos);
}

static void csr_free(csr_t *c) { free(c->off); free(c->nbr); free(c->w); }

/* binary heap of (key, vertex); max-heap when `maxh` */
typedef struct { uint64_t *k; uint64_t *v; uint64_t n, cap; int maxh; } heap_t;

static int hless(heap_t *h, uint64_t a, uint64_t b)
{
    return h->maxh ? h->k[a] > h->k[b] : h->k[a] < h->k[b];
}

static void hswap(heap_t *h, uint64_t a, uint64_t b)
{
    uint64_t t = h->k[a]; h->k[a] = h->k[b]; h->k[b] = t;
    t = h->v[a]; h->v[a] = h->v[b]; h->v[b] = t;
}

static void hpush(heap_t *h, uint64_t k, uint64_t v)
{
    if (h->n == h->cap) {
        h->cap = h->cap ? 2 * h->cap : 1024;
        h->k = realloc(h->k, h->cap * sizeof(uint64_t));
        h->v = realloc(h->v, h->cap * sizeof(uint64_t));
    }
    uint64_t i = h->n++;
    h->k[i] = k; h->v[i] = v;
    while (i && hless(h, i, (i - 1) / 2)) { hswap(h, i, (i - 1) / 2); i = (i - 1) / 2; }
}

static void hpop(heap_t *h, uint64_t *k, uint64_t *v)
{
    *k = h->k[0]; *v = h->v[0];
    h->n--;
    h->k[0] = h->k[h->n]; h->v[0] = h->v[h->n];
    uint64_t i = 0;
    for (;;) {
        uint64_t l = 2 * i + 1, r = l + 1, b = i;
        if (l < h->n && hless(h, l, b)) b = l;
        if (r < h->n && hless(h, r, b)) b = r;
        if (b == i) break;
        hswap(h, i, b); i = b;
    }
}

static uint64_t uf_find(uint64_t *p, uint64_t x)
{
    while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
    return x;
}

int or_static_values(int algo, uint64_t V, uint64_t m, const uint64_t *src,
                     const uint64_t *dst, const uint64_t *w, uint64_t root,
                     const uint64_t *init, uint64_t *values)
{
    if (algo < OR_BFS || algo > OR_WCC) return -1;
    for (uint64_t i = 0; i < m; i++)
        if (src[i] >= V || dst[i] >= V) return -1;
    if (algo != OR_WCC && root >= V) return -1;
    if ((algo == OR_SSSP || algo == OR_SSWP) && !w) return -1;

    if (algo == OR_WCC) {
        uint64_t *p = malloc(V * sizeof(uint64_t));
        for (uint64_t v = 0; v < V; v++) p[v] = v;
        for (uint64_t i = 0; i < m; i++) {
            uint64_t a = uf_find(p, src[i]), b = uf_find(p, dst[i]);
            if (a != b) p[a] = b;
        }
        uint64_t *best = malloc(V * sizeof(uint64_t));
        for (uint64_t v = 0; v < V; v++) best[v] = UINT64_MAX;
        for (uint64_t v = 0; v < V; v++) {
            uint64_t r = uf_find(p, v), l = or_init_value(algo, v, root, init);
            if (l < best[r]) best[r] = l;
        }
        for (uint64_t v = 0; v < V; v++) values[v] = best[uf_find(p, v)];
        free(p); free(best);
        return 0;
    }

    csr_t c;
    csr_build(&c, V, m, src, dst, w, 0);
    for (uint64_t v = 0; v < V; v++) values[v] = or_init_value(algo, v, root, init);
    uint64_t maxl = or_maxl(algo);
    if (algo == OR_BFS) {
        uint64_t *q = malloc(V * sizeof(uint64_t));
        uint64_t qh = 0, qt = 0;
        q[qt++] = root;
        while (qh < qt) {
            uint64_t u = q[qh++];
            uint64_t nd = values[u] + 1;
            if (nd >= maxl) continue;              /* s+1 < MAXL never holds */
            for (uint64_t e = c.off[u]; e < c.off[u + 1]; e++) {
                uint64_t x = c.nbr[e];
                if (nd < values[x]) { values[x] = nd; q[qt++] = x; }
            }
        }
        free(q);
    } else {
        heap_t h = {0};
        h.maxh = (algo == OR_SSWP);
        hpush(&h, values[root], root);
        while (h.n) {
            uint64_t k, u;
            hpop(&h, &k, &u);
            if (k != values[u]) continue;
            for (uint64_t e = c.off[u]; e < c.off[u + 1]; e++) {
                uint64_t x = c.nbr[e];
                uint64_t nv = cand(algo, k, c.w[e]);
                if (improves(algo, nv, values[x])) { values[x] = nv; hpush(&h, nv, x); }
            }
        }
        free(h.k); free(h.v);
    }
    csr_free(&c);
    return 0;
}

/* ------------------------------------------------------------------ */
/* 2. tree validity                                                    */
/* ------------------------------------------------------------------ */

int64_t or_check_tree(int algo, int undirected, uint64_t V, uint64_t m,
                      const uint64_t *src, const uint64_t *dst,
                      const uint64_t *w, uint64_t root, const uint64_t *init,
                      const uint64_t *values, const uint64_t *parents)
{
    /* incoming CSR (both directions when undirected): in[v] holds (p, w) */
    csr_t in;
    csr_build(&in, V, m, dst, src, w, undirected);
    int64_t bad = 0;
    for (uint64_t v = 0; v < V && !bad; v++) {
        uint64_t iv = or_init_value(algo, v, root, init);
        uint64_t p = parents[v];
        if (values[v] == iv) {
            if (p != V) bad = (int64_t)v + 1;
            continue;
        }
        if (p >= V) { bad = (int64_t)v + 1; break; }
        int ok = 0;
        for (uint64_t e = in.off[v]; e < in.off[v + 1] && !ok; e++)
            if (in.nbr[e] == p && cand(algo, values[p], in.w[e]) == values[v]) ok = 1;
        if (!ok) bad = (int64_t)v + 1;
    }
    if (!bad) {
        /* acyclic: colour walk along parent pointers */
        uint8_t *col = calloc(V, 1);
        for (uint64_t v = 0; v < V && !bad; v++) {
            uint64_t x = v;
            while (x < V && col[x] == 0) { col[x] = 1; x = parents[x]; }
            if (x < V && col[x] == 1) bad = (int64_t)x + 1;  /* closed a cycle */
            x = v;
            while (x < V && col[x] == 1) { col[x] = 2; x = parents[x]; }
        }
        free(col);
    }
    csr_free(&in);
    return bad;
}

/* ------------------------------------------------------------------ */
/* 3. cpu_ref store: IndexedEdgeStorage (storage.hpp:282-398)           */
/* ------------------------------------------------------------------ */

#define OR_INDEX_THRESHOLD 512   /* storage.hpp:304 */

typedef struct { uint64_t nbr; uint64_t data; uint16_t num; } or_ent;

typedef struct {
    or_ent *a;
    uint64_t len, cap, degree;
    uint64_t *hidx;              /* hub index: slot+1, 0 = empty */
    uint64_t hcap, hcount;
} or_adj;

typedef struct { or_adj *v; uint64_t n; } or_store;

static inline uint64_t mix64(uint64_t x)
{
    x ^= x >> 33; x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33; return x;
}

static inline uint64_t key_hash(uint64_t nbr, uint64_t data)
{
    return mix64(nbr * 17 + data);
}

static void hidx_insert(or_adj *a, uint64_t slot)
{
    uint64_t mask = a->hcap - 1;
    uint64_t h = key_hash(a->a[slot].nbr, a->a[slot].data) & mask;
    while (a->hidx[h]) h = (h + 1) & mask;
    a->hidx[h] = slot + 1;
    a->hcount++;
}

static void hidx_rebuild(or_adj *a, uint64_t cap)
{
    free(a->hidx);
    a->hcap = cap;
    a->hcount = 0;
    a->hidx = calloc(cap, sizeof(uint64_t));
    for (uint64_t i = 0; i < a->len; i++) hidx_insert(a, i);
}

/* find (nbr,data) or append it with num 0; returns slot */
static uint64_t adj_find_or_append(or_adj *a, uint64_t nbr, uint64_t data)
{
    if (a->len >= OR_INDEX_THRESHOLD && !a->hidx) {
        uint64_t cap = 1024;
        while (cap < 2 * a->len) cap <<= 1;
        hidx_rebuild(a, cap);
    }
    if (a->hidx) {
        uint64_t mask = a->hcap - 1;
        uint64_t h = key_hash(nbr, data) & mask;
        while (a->hidx[h]) {
            or_ent *e = &a->a[a->hidx[h] - 1];
            if (e->nbr == nbr && e->data == data) return a->hidx[h] - 1;
            h = (h + 1) & mask;
        }
    } else {
        for (uint64_t i = 0; i < a->len; i++)
            if (a->a[i].nbr == nbr && a->a[i].data == data) return i;
    }
    if (a->len == a->cap) {
        a->cap = a->cap ? 2 * a->cap : 4;
        a->a = realloc(a->a, a->cap * sizeof(or_ent));
    }
    uint64_t s = a->len++;
    a->a[s].nbr = nbr; a->a[s].data = data; a->a[s].num = 0;
    if (a->hidx) {
        if (2 * (a->hcount + 1) > a->hcap) hidx_rebuild(a, a->hcap * 2);
        else hidx_insert(a, s);
    }
    return s;
}

/* update_edge (storage.hpp:345-398): returns old num; never below zero;
 * degree += update when update>0 or old >= -update. */
static uint64_t store_update(or_store *st, uint64_t vid, uint64_t nbr, uint64_t data, int update)
{
    or_adj *a = &st->v[vid];
    uint64_t s = adj_find_or_append(a, nbr, data);
    uint64_t cur = a->a[s].num;
    if ((int)a->a[s].num + update >= 0) a->a[s].num = (uint16_t)(a->a[s].num + update);
    if (update > 0 || cur >= (uint64_t)(-update)) a->degree += (uint64_t)(int64_t)update;
    return cur;
}

/* ------------------------------------------------------------------ */
/* graph                                                               */
/* ------------------------------------------------------------------ */

typedef struct { uint64_t *a; uint64_t n, cap; } vec_t;

static inline void vpush(vec_t *v, uint64_t x)
{
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 256;
        v->a = realloc(v->a, v->cap * sizeof(uint64_t));
    }
    v->a[v->n++] = x;
}

struct or_graph {
    uint64_t V, hint_edges;
    int symmetric, dual, weighted;
    or_store out, in;
    /* engine scratch shared by all labels (graph.hpp:1728-1730) */
    vec_t act_in, act_out, act_tree;
    int64_t *invalidated;
    int64_t invalidated_idx;
    uint8_t *bm_in, *bm_out;
};

struct or_labels {
    uint64_t V;
    int algo;
    uint64_t root;
    uint64_t *init;        /* WCC custom labels or NULL */
    uint64_t *data;        /* value */
    uint64_t *pnbr;        /* parent vertex, V = empty (graph.hpp:90) */
    uint64_t *pdata;       /* parent edge weight */
    uint64_t recomputes;
};

or_graph *or_graph_create(uint64_t V, uint64_t hint_edges, int symmetric, int dual, int weighted)
{
    or_graph *g = calloc(1, sizeof(or_graph));
    g->V = V;
    g->hint_edges = hint_edges ? hint_edges : V * 16;   /* graph.hpp:64 */
    g->symmetric = symmetric; g->dual = dual; g->weighted = weighted;
    g->out.n = V; g->out.v = calloc(V ? V : 1, sizeof(or_adj));
    if (!symmetric) { g->in.n = V; g->in.v = calloc(V ? V : 1, sizeof(or_adj)); }
    g->invalidated = calloc(V ? V : 1, sizeof(int64_t));
    g->bm_in = calloc(V ? V : 1, 1);
    g->bm_out = calloc(V ? V : 1, 1);
    return g;
}

static void store_free(or_store *s)
{
    for (uint64_t i = 0; i < s->n; i++) { free(s->v[i].a); free(s->v[i].hidx); }
    free(s->v);
}

void or_graph_destroy(or_graph *g)
{
    if (!g) return;
    store_free(&g->out);
    if (!g->symmetric) store_free(&g->in);
    free(g->act_in.a); free(g->act_out.a); free(g->act_tree.a);
    free(g->invalidated); free(g->bm_in); free(g->bm_out);
    free(g);
}

static inline or_store *outgoing(or_graph *g) { return &g->out; }
static inline or_store *incoming(or_graph *g) { return g->symmetric ? &g->out : &g->in; }

static uint64_t add_directed(or_graph *g, uint64_t s, uint64_t d, uint64_t w, int upd)
{
    uint64_t data = g->weighted ? w : 0;
    uint64_t cur = store_update(&g->out, s, d, data, upd);
    if (!g->symmetric) store_update(&g->in, d, s, data, upd);
    return cur;
}

/* graph.hpp:203-227 */
uint64_t or_add_edge(or_graph *g, uint64_t s, uint64_t d, uint64_t w, int directed)
{
    if (s >= g->V || d >= g->V) return UINT64_MAX;
    if (!directed) {
        add_directed(g, s, d, w, 1);
        return add_directed(g, d, s, w, 1) - (s == d);
    }
    return add_directed(g, s, d, w, 1);
}

/* graph.hpp:229-253 */
uint64_t or_del_edge(or_graph *g, uint64_t s, uint64_t d, uint64_t w, int directed)
{
    if (s >= g->V || d >= g->V) return UINT64_MAX;
    if (!directed) {
        add_directed(g, s, d, w, -1);
        return add_directed(g, d, s, w, -1);
    }
    return add_directed(g, s, d, w, -1);
}

void or_add_edges(or_graph *g, uint64_t n, const uint64_t *src, const uint64_t *dst,
                  const uint64_t *w, int directed)
{
    for (uint64_t i = 0; i < n; i++) or_add_edge(g, src[i], dst[i], w ? w[i] : 0, directed);
}

/* Bulk load into an empty graph: the same multiset store (entries, counts,
 * degrees, hub indexes) that n sequential or_add_edge calls build, made by
 * counting sort + per-list sort and run-length merge.  Entries sit in
 * (nbr, data) order instead of insertion order, which changes no value. */
typedef struct { uint64_t nbr, data; } key_t_;

static int key_cmp(const void *a, const void *b)
{
    const key_t_ *x = a, *y = b;
    if (x->nbr != y->nbr) return x->nbr < y->nbr ? -1 : 1;
    if (x->data != y->data) return x->data < y->data ? -1 : 1;
    return 0;
}

/* list ops of add_edge (graph.hpp:203-227) for one direction: pass 0 fills
 * out-lists (src gets (dst, w)), pass 1 in-lists (dst gets (src, w)); an
 * undirected add contributes both orientations. */
static inline void bulk_op(uint64_t i, int k, int pass, const uint64_t *src, const uint64_t *dst,
                           uint64_t *v, uint64_t *nbr)
{
    uint64_t a = src[i], b = dst[i];
    if (k) { uint64_t t = a; a = b; b = t; }
    if (pass) { uint64_t t = a; a = b; b = t; }
    *v = a;
    *nbr = b;
}

static void bulk_store(or_store *st, uint64_t V, uint64_t n, const uint64_t *src, const uint64_t *dst,
                       const uint64_t *w, int weighted, int per, int pass)
{
    uint64_t m = n * per;
    uint64_t *off = calloc(V + 1, sizeof(uint64_t));
    for (uint64_t i = 0; i < n; i++)
        for (int k = 0; k < per; k++) {
            uint64_t v, nb;
            bulk_op(i, k, pass, src, dst, &v, &nb);
            off[v + 1]++;
        }
    for (uint64_t v = 0; v < V; v++) off[v + 1] += off[v];
    uint64_t *pos = malloc((V + 1) * sizeof(uint64_t));
    memcpy(pos, off, (V + 1) * sizeof(uint64_t));
    key_t_ *kk = malloc((m ? m : 1) * sizeof(key_t_));
    for (uint64_t i = 0; i < n; i++)
        for (int k = 0; k < per; k++) {
            uint64_t v, nb;
            bulk_op(i, k, pass, src, dst, &v, &nb);
            uint64_t p = pos[v]++;
            kk[p].nbr = nb;
            kk[p].data = weighted && w ? w[i] : 0;
        }
    free(pos);
#pragma omp parallel for schedule(dynamic, 256)
    for (uint64_t v = 0; v < V; v++) {
        uint64_t b = off[v], e = off[v + 1];
        if (b == e) continue;
        qsort(kk + b, e - b, sizeof(key_t_), key_cmp);
        or_adj *a = &st->v[v];
        uint64_t uniq = 1;
        for (uint64_t i = b + 1; i < e; i++) uniq += key_cmp(&kk[i - 1], &kk[i]) != 0;
        a->a = malloc(uniq * sizeof(or_ent));
        a->cap = uniq;
        a->len = 0;
        for (uint64_t i = b; i < e; i++) {
            if (i == b || key_cmp(&kk[i - 1], &kk[i]) != 0) {
                a->a[a->len].nbr = kk[i].nbr;
                a->a[a->len].data = kk[i].data;
                a->a[a->len].num = 0;
                a->len++;
            }
            a->a[a->len - 1].num = (uint16_t)(a->a[a->len - 1].num + 1);
        }
        a->degree = e - b;
        if (a->len >= OR_INDEX_THRESHOLD) {
            uint64_t cap = 1024;
            while (cap < 2 * a->len) cap <<= 1;
            hidx_rebuild(a, cap);
        }
    }
    free(kk);
    free(off);
}

int or_bulk_load(or_graph *g, uint64_t n, const uint64_t *src, const uint64_t *dst,
                 const uint64_t *w, int directed)
{
    for (uint64_t v = 0; v < g->V; v++)
        if (g->out.v[v].len || (!g->symmetric && g->in.v[v].len)) return -1;
    for (uint64_t i = 0; i < n; i++)
        if (src[i] >= g->V || dst[i] >= g->V) return -1;
    int per = directed ? 1 : 2;
    bulk_store(&g->out, g->V, n, src, dst, w, g->weighted, per, 0);
    if (!g->symmetric) bulk_store(&g->in, g->V, n, src, dst, w, g->weighted, per, 1);
    return 0;
}

uint64_t or_degree(or_graph *g, uint64_t v, int inc)
{
    return (inc ? incoming(g) : outgoing(g))->v[v].degree;
}

uint64_t or_edge_num(or_graph *g, uint64_t s, uint64_t d, uint64_t w)
{
    or_adj *a = &g->out.v[s];
    uint64_t data = g->weighted ? w : 0;
    for (uint64_t i = 0; i < a->len; i++)
        if (a->a[i].nbr == d && a->a[i].data == data) return a->a[i].num;
    return 0;
}

uint64_t or_live_edges(or_graph *g, uint64_t *src, uint64_t *dst, uint64_t *w)
{
    uint64_t n = 0;
    for (uint64_t v = 0; v < g->V; v++) {
        or_adj *a = &g->out.v[v];
        for (uint64_t i = 0; i < a->len; i++) {
            if (!a->a[i].num) continue;
            if (src) { src[n] = v; dst[n] = a->a[i].nbr; if (w) w[n] = a->a[i].data; }
            n++;
        }
    }
    return n;
}

/* ------------------------------------------------------------------ */
/* labels + engine                                                     */
/* ------------------------------------------------------------------ */

or_labels *or_labels_create(or_graph *g, int algo, uint64_t root, const uint64_t *init)
{
    or_labels *L = calloc(1, sizeof(or_labels));
    L->algo = algo; L->root = root; L->V = g->V;
    uint64_t V = g->V ? g->V : 1;
    if (init) { L->init = malloc(V * sizeof(uint64_t)); memcpy(L->init, init, g->V * sizeof(uint64_t)); }
    L->data = calloc(V, sizeof(uint64_t));
    L->pnbr = malloc(V * sizeof(uint64_t));
    L->pdata = calloc(V, sizeof(uint64_t));
    for (uint64_t v = 0; v < g->V; v++) L->pnbr[v] = g->V;   /* alloc_vertex_tree_array */
    return L;
}

void or_labels_destroy(or_labels *L)
{
    if (!L) return;
    free(L->init); free(L->data); free(L->pnbr); free(L->pdata); free(L);
}

void or_labels_read(or_labels *L, uint64_t *values, uint64_t *parents, uint64_t *pweights)
{
    uint64_t n = L->V;
    if (values) memcpy(values, L->data, n * sizeof(uint64_t));
    if (parents) memcpy(parents, L->pnbr, n * sizeof(uint64_t));
    if (pweights) memcpy(pweights, L->pdata, n * sizeof(uint64_t));
}

uint64_t or_recompute_count(or_labels *L) { return L->recomputes; }

/* Set the labels to a given fixpoint (values + parent vertices, V = empty)
 * instead of running build_tree.  Each parent edge's weight is resolved
 * from the store as a live in-edge (p -> v, w) with f(value[p], w) ==
 * value[v].  Returns 0, or 1 + the first vertex without such an edge. */
int64_t or_labels_assign(or_graph *g, or_labels *L, const uint64_t *values, const uint64_t *parents)
{
    or_store *is = g->symmetric ? &g->out : &g->in;
    int64_t bad = 0;
    for (uint64_t v = 0; v < g->V; v++) {
        L->data[v] = values[v];
        L->pnbr[v] = parents[v] >= g->V ? g->V : parents[v];
        L->pdata[v] = 0;
    }
#pragma omp parallel for schedule(dynamic, 4096)
    for (uint64_t v = 0; v < g->V; v++) {
        uint64_t p = L->pnbr[v];
        if (p == g->V) continue;
        or_adj *a = &is->v[v];
        int ok = 0;
        for (uint64_t i = 0; i < a->len && !ok; i++) {
            or_ent e = a->a[i];
            if (e.num && e.nbr == p && cand(L->algo, values[p], e.data) == values[v]) {
                L->pdata[v] = e.data;
                ok = 1;
            }
        }
        if (!ok) {
#pragma omp critical
            if (!bad || (int64_t)v + 1 < bad) bad = (int64_t)v + 1;
        }
    }
    return bad;
}

static inline uint64_t init_of(or_graph *g, or_labels *L, uint64_t v)
{
    (void)g;
    return or_init_value(L->algo, v, L->root, L->init);
}

/* update_label, CAS form (graph.hpp:1861-1887), sequential: write value and
 * parent edge when update_func strictly improves. */
static inline int update_label(or_graph *g, or_labels *L, uint64_t src, uint64_t dst,
                               uint64_t w, uint64_t src_data)
{
    (void)g;
    uint64_t c = cand(L->algo, src_data, w);
    if (!improves(L->algo, c, L->data[dst])) return 0;
    L->data[dst] = c;
    L->pnbr[dst] = src;
    L->pdata[dst] = w;
    return 1;
}

/* one push level over a sparse frontier (do_update_tree_add body,
 * graph.hpp:1013-1045) */
static uint64_t push_level(or_graph *g, or_labels *L, vec_t *fin, vec_t *fout)
{
    uint64_t res = 0;
    or_store *os = outgoing(g);
    for (uint64_t k = 0; k < fin->n; k++) {
        uint64_t src = fin->a[k];
        or_adj *a = &os->v[src];
        if (!a->degree) continue;
        for (uint64_t i = 0; i < a->len; i++) {
            or_ent e = a->a[i];
            if (!e.num) continue;
            uint64_t dst = e.nbr;
            if (improves(L->algo, cand(L->algo, L->data[src], e.data), L->data[dst])) {
                if (update_label(g, L, src, dst, e.data, L->data[src])) {
                    vpush(fout, dst);
                    res++;
                }
            }
        }
    }
    return res;
}

static uint64_t do_update_tree_add(or_graph *g, or_labels *L)
{
    uint64_t total = 0;
    for (;;) {
        g->act_out.n = 0;
        uint64_t local = push_level(g, L, &g->act_in, &g->act_out);
        vec_t t = g->act_in; g->act_in = g->act_out; g->act_out = t;
        total += local;
        if (!local) break;
    }
    return total;
}

/* build_tree (graph.hpp:873-985) with the push/pull switch of
 * stream_edges(Bitmap) (graph.hpp:323-396). */
uint64_t or_build_tree(or_graph *g, or_labels *L)
{
    uint64_t V = g->V;
    memset(g->bm_in, 0, V);
    for (uint64_t v = 0; v < V; v++) {
        L->data[v] = init_of(g, L, v);
        L->pnbr[v] = V;
        L->pdata[v] = 0;
        if (init_active(L->algo, v, L->root)) g->bm_in[v] = 1;
    }
    uint64_t total = 0;
    or_store *os = outgoing(g), *is = incoming(g);
    for (;;) {
        memset(g->bm_out, 0, V);
        uint64_t active_edges = 0;
        for (uint64_t v = 0; v < V; v++) if (g->bm_in[v]) active_edges += os->v[v].degree;
        int sparse = active_edges < V || active_edges < g->hint_edges / 20 || !g->dual;
        uint64_t local = 0;
        if (sparse) {
            for (uint64_t src = 0; src < V; src++) {
                if (!g->bm_in[src] || !os->v[src].degree) continue;
                or_adj *a = &os->v[src];
                for (uint64_t i = 0; i < a->len; i++) {
                    or_ent e = a->a[i];
                    if (!e.num) continue;
                    uint64_t sd = L->data[src];
                    if (improves(L->algo, cand(L->algo, sd, e.data), L->data[e.nbr]) &&
                        update_label(g, L, src, e.nbr, e.data, sd)) {
                        g->bm_out[e.nbr] = 1;
                        local++;
                    }
                }
            }
        } else {
            for (uint64_t dst = 0; dst < V; dst++) {
                or_adj *a = &is->v[dst];
                if (!a->degree) continue;
                uint64_t best = L->data[dst], bsrc = UINT64_MAX, bw = 0;
                for (uint64_t i = 0; i < a->len; i++) {
                    or_ent e = a->a[i];
                    if (!e.num) continue;
                    uint64_t c = cand(L->algo, L->data[e.nbr], e.data);
                    if (improves(L->algo, c, best)) { best = c; bsrc = e.nbr; bw = e.data; }
                }
                if (bsrc == UINT64_MAX) continue;
                uint64_t sd = L->data[bsrc];
                if (improves(L->algo, cand(L->algo, sd, bw), L->data[dst]) &&
                    update_label(g, L, bsrc, dst, bw, sd)) {
                    g->bm_out[dst] = 1;
                    local++;
                }
            }
        }
        uint8_t *t = g->bm_in; g->bm_in = g->bm_out; g->bm_out = t;
        total += local;
        if (!local) break;
    }
    return total;
}

/* need_update_tree_add (graph.hpp:987-1000) */
int or_need_update_add(or_graph *g, or_labels *L, uint64_t s, uint64_t d, uint64_t w, int directed)
{
    uint64_t data = g->weighted ? w : 0;
    if (!directed)
        return or_need_update_add(g, L, s, d, w, 1) || or_need_update_add(g, L, d, s, w, 1);
    return improves(L->algo, cand(L->algo, L->data[s], data), L->data[d]);
}

/* need_update_tree_del (graph.hpp:1220-1233): parent edge equality
 * (nbr, data) as AdjEdge::operator== (type.hpp:32-35, 49-52). */
int or_need_update_del(or_graph *g, or_labels *L, uint64_t s, uint64_t d, uint64_t w, int directed)
{
    uint64_t data = g->weighted ? w : 0;
    if (!directed)
        return or_need_update_del(g, L, s, d, w, 1) || or_need_update_del(g, L, d, s, w, 1);
    return L->pnbr[d] == s && L->pdata[d] == data;
}

static uint64_t seed_add(or_graph *g, or_labels *L, uint64_t s, uint64_t d, uint64_t data)
{
    if (!improves(L->algo, cand(L->algo, L->data[s], data), L->data[d])) return 0;
    if (update_label(g, L, s, d, data, L->data[s])) { vpush(&g->act_in, d); return 1; }
    return 0;
}

/* update_tree_add single (graph.hpp:1124-1160) */
uint64_t or_update_add(or_graph *g, or_labels *L, uint64_t s, uint64_t d, uint64_t w, int directed)
{
    uint64_t data = g->weighted ? w : 0;
    if (!directed)
        return or_update_add(g, L, s, d, w, 1) + or_update_add(g, L, d, s, w, 1);
    if (!improves(L->algo, cand(L->algo, L->data[s], data), L->data[d])) return 0;
    g->act_in.n = 0;
    seed_add(g, L, s, d, data);
    return do_update_tree_add(g, L);
}

/* update_tree_add batch (graph.hpp:1162-1218) */
uint64_t or_update_add_batch(or_graph *g, or_labels *L, uint64_t n, const uint64_t *s,
                             const uint64_t *d, const uint64_t *w, int directed)
{
    if (n == 1) return or_update_add(g, L, s[0], d[0], w ? w[0] : 0, directed);
    g->act_in.n = 0;
    uint64_t total = 0;
    for (uint64_t i = 0; i < n; i++) {
        uint64_t data = g->weighted && w ? w[i] : 0;
        total += seed_add(g, L, s[i], d[i], data);
        if (!directed) total += seed_add(g, L, d[i], s[i], data);
    }
    if (!g->act_in.n) return total;
    return do_update_tree_add(g, L);
}

/* has_parents (graph.hpp:1889-1907) */
static int has_parents(or_graph *g, or_labels *L, uint64_t x, int *need_recompute)
{
    uint64_t num = 0;
    int64_t idx = g->invalidated_idx;
    for (;;) {
        if (g->invalidated[x] == idx || g->invalidated[x] == -idx) return 1;
        if (L->pnbr[x] != g->V) x = L->pnbr[x];
        else break;
        if ((double)(++num) > (double)g->V * 0.1) { *need_recompute = 1; return 1; }
    }
    return 0;
}

/* do_update_tree_del (graph.hpp:1235-1612) */
static uint64_t do_update_tree_del(or_graph *g, or_labels *L)
{
    uint64_t V = g->V;
    int64_t idx = g->invalidated_idx;
    uint64_t size_tree = 0;
    int need_recompute = 0;
    or_store *os = outgoing(g), *is = incoming(g);
    uint64_t active_vertices = g->act_in.n;
    for (uint64_t i = 0; i < V && active_vertices; i++) {
        if (need_recompute ||
            (i <= 10 && (double)(size_tree + active_vertices) > 0.05 * (double)V &&
             size_tree + active_vertices > 2000000)) {
            memset(g->invalidated, 0, V * sizeof(int64_t));
            L->recomputes++;
            return or_build_tree(g, L);
        }
        /* rescue over in-edges (graph.hpp:1323-1361) */
        if (i <= 5 || active_vertices < 8192 * 2) {
            for (uint64_t k = 0; k < g->act_in.n; k++) {
                uint64_t dst = g->act_in.a[k];
                or_adj *a = &is->v[dst];
                if (!a->degree) continue;
                for (uint64_t j = 0; j < a->len; j++) {
                    or_ent e = a->a[j];
                    if (!e.num) continue;
                    uint64_t src = e.nbr;
                    if (cand(L->algo, L->data[src], e.data) == L->data[dst] &&
                        g->invalidated[dst] == idx && !has_parents(g, L, src, &need_recompute)) {
                        g->invalidated[dst] = -idx;
                        L->pnbr[dst] = src;
                        L->pdata[dst] = e.data;
                    }
                }
            }
        }
        if (need_recompute) continue;
        g->act_out.n = 0;
        for (uint64_t k = 0; k < g->act_in.n; k++) {
            uint64_t vid = g->act_in.a[k];
            if (g->invalidated[vid] == idx) {
                vpush(&g->act_out, vid);
                vpush(&g->act_tree, vid);
                size_tree++;
            } else if (g->invalidated[vid] == -idx) {
                g->invalidated[vid] = 0;
            }
        }
        g->act_in.n = 0;
        active_vertices = 0;
        for (uint64_t k = 0; k < g->act_out.n; k++) {
            uint64_t src = g->act_out.a[k];
            or_adj *a = &os->v[src];
            if (!a->degree) continue;
            for (uint64_t j = 0; j < a->len; j++) {
                or_ent e = a->a[j];
                if (!e.num) continue;
                uint64_t dst = e.nbr;
                if (L->pnbr[dst] == src && L->pdata[dst] == e.data && g->invalidated[dst] == 0) {
                    g->invalidated[dst] = idx;
                    vpush(&g->act_in, dst);
                    active_vertices++;
                }
            }
        }
    }
    /* reset the invalidated subtree (graph.hpp:1420-1430) */
    for (uint64_t k = 0; k < g->act_tree.n; k++) {
        uint64_t vid = g->act_tree.a[k];
        L->data[vid] = init_of(g, L, vid);
        L->pnbr[vid] = V;
        L->pdata[vid] = 0;
        g->invalidated[vid] = 0;
    }
    vec_t t = g->act_in; g->act_in = g->act_tree; g->act_tree = t;
    g->act_tree.n = 0;
    /* pull from in-neighbours (graph.hpp:1433-1510) */
    uint64_t total = 0;
    for (uint64_t k = 0; k < g->act_in.n; k++) {
        uint64_t dst = g->act_in.a[k];
        or_adj *a = &is->v[dst];
        if (!a->degree) continue;
        uint64_t best = L->data[dst], bsrc = UINT64_MAX, bw = 0;
        for (uint64_t j = 0; j < a->len; j++) {
            or_ent e = a->a[j];
            if (!e.num) continue;
            uint64_t c = cand(L->algo, L->data[e.nbr], e.data);
            if (improves(L->algo, c, best)) { best = c; bsrc = e.nbr; bw = e.data; }
        }
        if (bsrc == UINT64_MAX) continue;
        if (improves(L->algo, cand(L->algo, L->data[bsrc], bw), L->data[dst]) &&
            update_label(g, L, bsrc, dst, bw, L->data[bsrc]))
            total++;
    }
    /* push to the fixpoint (graph.hpp:1513-1608) */
    return total + do_update_tree_add(g, L);
}

/* update_tree_del single (graph.hpp:1622-1656) */
uint64_t or_update_del(or_graph *g, or_labels *L, uint64_t s, uint64_t d, uint64_t w, int directed)
{
    uint64_t data = g->weighted ? w : 0;
    if (!directed)
        return or_update_del(g, L, s, d, w, 1) + or_update_del(g, L, d, s, w, 1);
    if (!(L->pnbr[d] == s && L->pdata[d] == data)) return 0;
    g->act_in.n = 0;
    g->act_tree.n = 0;
    vpush(&g->act_in, d);
    g->invalidated_idx++;
    g->invalidated[d] = g->invalidated_idx;
    return do_update_tree_del(g, L);
}

/* update_tree_del batch (graph.hpp:1658-1718) */
uint64_t or_update_del_batch(or_graph *g, or_labels *L, uint64_t n, const uint64_t *s,
                             const uint64_t *d, const uint64_t *w, int directed)
{
    if (n == 1) return or_update_del(g, L, s[0], d[0], w ? w[0] : 0, directed);
    g->act_in.n = 0;
    g->act_tree.n = 0;
    g->invalidated_idx++;
    int64_t idx = g->invalidated_idx;
    for (uint64_t i = 0; i < n; i++) {
        uint64_t data = g->weighted && w ? w[i] : 0;
        for (int dir = 0; dir < (directed ? 1 : 2); dir++) {
            uint64_t a = dir ? d[i] : s[i], b = dir ? s[i] : d[i];
            if (L->pnbr[b] == a && L->pdata[b] == data && g->invalidated[b] != idx) {
                g->invalidated[b] = idx;
                vpush(&g->act_in, b);
            }
        }
    }
    if (!g->act_in.n) return 0;
    return do_update_tree_del(g, L);
}

uint64_t or_stream_plain(or_graph *g, or_labels *L, uint64_t n,
                         const uint64_t *as, const uint64_t *ad, const uint64_t *aw,
                         const uint64_t *ds, const uint64_t *dd, const uint64_t *dw,
                         int directed, double *lat_ns)
{
    uint64_t unsafe = 0;
    for (uint64_t i = 0; i < n; i++) {
        struct timespec t0, t1;
        if (lat_ns) clock_gettime(CLOCK_MONOTONIC, &t0);
        int any = 0;
        uint64_t w1 = aw ? aw[i] : 0, w2 = dw ? dw[i] : 0;
        if (or_add_edge(g, as[i], ad[i], w1, directed) == 0) {
            or_update_add(g, L, as[i], ad[i], w1, directed);
            any = 1;
        }
        if (or_del_edge(g, ds[i], dd[i], w2, directed) == 1) {
            or_update_del(g, L, ds[i], dd[i], w2, directed);
            any = 1;
        }
        unsafe += any;
        if (lat_ns) {
            clock_gettime(CLOCK_MONOTONIC, &t1);
            lat_ns[i] = (double)(t1.tv_sec - t0.tv_sec) * 1e9 + (double)(t1.tv_nsec - t0.tv_nsec);
        }
    }
    return unsafe;
}

/* The all_inc alternation (all_inc.cpp:254-300) over four trees, as the
 * reference runs them: BFS/SSSP/SSWP labels on one directed weighted graph
 * (nLd labels) and WCC on a symmetric graph (Lu may be NULL).  Update u is an
 * add of stream edge u/2 (even u) or a delete of edge u/2 (odd u); each
 * follows the plain drivers' rule (bfs_inc.cpp:160-189): incremental work
 * when the add returned 0 or the delete returned 1. */
uint64_t or_stream_mixed(or_graph *gd, or_labels **Ld, int nLd, or_graph *gu, or_labels *Lu,
                         uint64_t n_updates, const uint64_t *as, const uint64_t *ad, const uint64_t *aw,
                         const uint64_t *ds, const uint64_t *dd, const uint64_t *dw, double *lat_ns)
{
    uint64_t unsafe = 0;
    for (uint64_t u = 0; u < n_updates; u++) {
        struct timespec t0, t1;
        if (lat_ns) clock_gettime(CLOCK_MONOTONIC, &t0);
        uint64_t k = u / 2;
        int any = 0;
        if ((u & 1) == 0) {
            uint64_t s = as[k], d = ad[k], w = aw ? aw[k] : 0;
            if (gd && or_add_edge(gd, s, d, w, 1) == 0) {
                for (int i = 0; i < nLd; i++) or_update_add(gd, Ld[i], s, d, w, 1);
                any = 1;
            }
            if (gu && or_add_edge(gu, s, d, 0, 0) == 0 && Lu) {
                or_update_add(gu, Lu, s, d, 0, 0);
                any = 1;
            }
        } else {
            uint64_t s = ds[k], d = dd[k], w = dw ? dw[k] : 0;
            if (gd && or_del_edge(gd, s, d, w, 1) == 1) {
                for (int i = 0; i < nLd; i++) or_update_del(gd, Ld[i], s, d, w, 1);
                any = 1;
            }
            if (gu && or_del_edge(gu, s, d, 0, 0) == 1 && Lu) {
                or_update_del(gu, Lu, s, d, 0, 0);
                any = 1;
            }
        }
        unsafe += any;
        if (lat_ns) {
            clock_gettime(CLOCK_MONOTONIC, &t1);
            lat_ns[u] = (double)(t1.tv_sec - t0.tv_sec) * 1e9 + (double)(t1.tv_nsec - t0.tv_nsec);
        }
    }
    return unsafe;
}


/* ------------------------------------------------------------------ */
/* The rt driver's serving loop on host threads (bfs_inc_rt.cpp:154-431), */
/* over the four trees of the all_inc stream: a parallel safe phase and a */
/* serial unsafe phase per round, with the num_changes_threshold cap.   */
/* ------------------------------------------------------------------ */

static double now_s(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static int u64_cmp(const void *a, const void *b)
{
    const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}

/* per-list spin locks: the safe phase's concurrent add_edge / del_edge
 * (IndexedEdgeStorage locks each list, storage.hpp:348) */
static void lock_list(uint8_t *l)
{
    while (__atomic_test_and_set(l, __ATOMIC_ACQUIRE))
        while (__atomic_load_n(l, __ATOMIC_RELAXED)) { }
}
static void unlock_list(uint8_t *l) { __atomic_clear(l, __ATOMIC_RELEASE); }

static uint64_t locked_directed(or_graph *g, uint8_t *lk, uint64_t s, uint64_t d, uint64_t w, int upd)
{
    uint64_t data = g->weighted ? w : 0;
    lock_list(&lk[s]);
    uint64_t cur = store_update(&g->out, s, d, data, upd);
    unlock_list(&lk[s]);
    if (!g->symmetric) {
        lock_list(&lk[g->V + d]);
        store_update(&g->in, d, s, data, upd);
        unlock_list(&lk[g->V + d]);
    }
    return cur;
}

static void locked_update(or_graph *g, uint8_t *lk, uint64_t s, uint64_t d, uint64_t w, int directed, int upd)
{
    locked_directed(g, lk, s, d, w, upd);
    if (!directed) locked_directed(g, lk, d, s, w, upd);
}

/* update u of the all_inc alternation: add edge u/2 (even u) or delete it (odd u) */
static int rt_unsafe(or_graph *gd, or_labels **Ld, int nLd, or_graph *gu, or_labels *Lu, int add, uint64_t s,
                     uint64_t d, uint64_t w)
{
    for (int i = 0; i < nLd; i++)
        if (add ? or_need_update_add(gd, Ld[i], s, d, w, 1) : or_need_update_del(gd, Ld[i], s, d, w, 1)) return 1;
    if (gu && Lu && (add ? or_need_update_add(gu, Lu, s, d, 0, 0) : or_need_update_del(gu, Lu, s, d, 0, 0)))
        return 1;
    return 0;
}

uint64_t or_serve_rt(or_graph *gd, or_labels **Ld, int nLd, or_graph *gu, or_labels *Lu, uint64_t n_updates,
                     const uint64_t *as, const uint64_t *ad, const uint64_t *aw, const uint64_t *ds,
                     const uint64_t *dd, const uint64_t *dw, const double *arrival, double timeout_ms,
                     double target_timeout_rate, double *lat_ms, uint64_t *stats)
{
    const uint64_t V = gd ? gd->V : gu->V;
    uint8_t *lkd = gd ? calloc(2 * V, 1) : NULL;
    uint8_t *lku = gu ? calloc(2 * V, 1) : NULL;
    uint64_t *unsafe = malloc((n_updates ? n_updates : 1) * sizeof(uint64_t));
    int threads = 1;
#ifdef _OPENMP
    threads = omp_get_max_threads();
#endif
    uint64_t threshold = (uint64_t)threads, rounds = 0, n_unsafe_total = 0, timeouts = 0;
    uint64_t last_timeouts = 0, counted = 0, last_counted = 0, overloaded = 0;
    uint8_t *state = calloc(n_updates ? n_updates : 1, 1);   /* 1: done (or queued this round) */
    uint64_t lo = 0;             /* every update below lo is done */
    const double t0 = now_s();
    const double to = timeout_ms * 1e-3;
    while (lo < n_updates) {
        double tn = now_s() - t0;
        while (arrival[lo] > tn) tn = now_s() - t0;   /* idle until a request arrives */
        if (tn - arrival[lo] > 10.0 * to + 0.2) { overloaded = 1; break; }
        /* the updates that have arrived */
        uint64_t l = lo, h = n_updates;
        while (l < h) {
            const uint64_t m = l + (h - l) / 2;
            if (arrival[m] <= tn) l = m + 1; else h = m;
        }
        const uint64_t hi = l;
        /* safe phase (bfs_inc_rt.cpp:211-326): threads classify arrived
         * updates against the labels; safe ones are applied to the store at
         * once and complete, unsafe ones wait for the serial phase.  The
         * phase stops taking updates once the unsafe ones exceed the
         * threshold or the oldest of them has waited 0.8 x timeout
         * (need_break_func, bfs_inc_rt.cpp:156-164); the rest stay pending */
        uint64_t n_uns = 0;
        uint64_t first_uns = UINT64_MAX;   /* smallest unsafe update (arrivals grow with u) */
        int stop = 0;
#pragma omp parallel for schedule(dynamic, 64)
        for (uint64_t u = lo; u < hi; u++) {
            int st;
#pragma omp atomic read
            st = stop;
            if (st || state[u]) continue;
            uint64_t nu;
#pragma omp atomic read
            nu = n_uns;
            const uint64_t fu = __atomic_load_n(&first_uns, __ATOMIC_RELAXED);
            const double t = now_s() - t0;
            if (nu > threshold || (fu != UINT64_MAX && t - arrival[fu] > 0.8 * to)) {
#pragma omp atomic write
                stop = 1;
                continue;
            }
            const uint64_t k = u / 2;
            const int add = (u & 1) == 0;
            const uint64_t s = add ? as[k] : ds[k], d = add ? ad[k] : dd[k];
            const uint64_t w = add ? (aw ? aw[k] : 0) : (dw ? dw[k] : 0);
            if (rt_unsafe(gd, Ld, nLd, gu, Lu, add, s, d, w)) {
                uint64_t pos;
#pragma omp atomic capture
                pos = n_uns++;
                unsafe[pos] = u;
                uint64_t cur = __atomic_load_n(&first_uns, __ATOMIC_RELAXED);
                while (u < cur && !__atomic_compare_exchange_n(&first_uns, &cur, u, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
                }
            } else {
                if (gd) locked_update(gd, lkd, s, d, w, 1, add ? 1 : -1);
                if (gu) locked_update(gu, lku, s, d, 0, 0, add ? 1 : -1);
                lat_ms[u] = (now_s() - t0 - arrival[u]) * 1e3;
            }
            state[u] = 1;
        }
        /* serial phase, in arrival order (bfs_inc_rt.cpp:329-405) */
        qsort(unsafe, n_uns, sizeof(uint64_t), u64_cmp);
        for (uint64_t i = 0; i < n_uns; i++) {
            const uint64_t u = unsafe[i], k = u / 2;
            const int add = (u & 1) == 0;
            const uint64_t s = add ? as[k] : ds[k], d = add ? ad[k] : dd[k], w = add ? (aw ? aw[k] : 0) : (dw ? dw[k] : 0);
            if (add) {
                if (gd && or_add_edge(gd, s, d, w, 1) == 0)
                    for (int t = 0; t < nLd; t++) or_update_add(gd, Ld[t], s, d, w, 1);
                if (gu && or_add_edge(gu, s, d, 0, 0) == 0 && Lu) or_update_add(gu, Lu, s, d, 0, 0);
            } else {
                if (gd && or_del_edge(gd, s, d, w, 1) == 1)
                    for (int t = 0; t < nLd; t++) or_update_del(gd, Ld[t], s, d, w, 1);
                if (gu && or_del_edge(gu, s, d, 0, 0) == 1 && Lu) or_update_del(gu, Lu, s, d, 0, 0);
            }
            lat_ms[u] = (now_s() - t0 - arrival[u]) * 1e3;
        }
        n_unsafe_total += n_uns;
        rounds++;
        while (lo < n_updates && state[lo]) lo++;
        /* AIMD on the unsafe-change threshold every 3 rounds (bfs_inc_rt.cpp:409-427) */
        for (; counted < lo; counted++)
            if (lat_ms[counted] > timeout_ms) timeouts++;
        if (rounds % 3 == 0 && counted > last_counted) {
            const double local = (double)(timeouts - last_timeouts) / (double)(counted - last_counted);
            const double global = (double)timeouts / (double)counted;
            if (local > 0.8 * target_timeout_rate) {
                threshold = (uint64_t)(threshold * 0.9);
                if (threshold < (uint64_t)threads) threshold = (uint64_t)threads;
            }
            if (local < 0.8 * target_timeout_rate && global < target_timeout_rate) {
                if (threshold < 1000) threshold += 10;
                else threshold = (uint64_t)(threshold * 1.01);
                if (threshold > 1000000) threshold = 1000000;
            }
            last_timeouts = timeouts;
            last_counted = counted;
        }
    }
    const double wall = now_s() - t0;
    if (stats) {
        stats[0] = lo;                         /* served (every update below lo) */
        stats[1] = rounds;
        stats[2] = n_unsafe_total;
        stats[3] = threshold;
        stats[4] = overloaded;
        stats[5] = (uint64_t)threads;
        stats[6] = (uint64_t)(wall * 1e9);
    }
    free(lkd);
    free(lku);
    free(unsafe);
    free(state);
    return lo;
}
