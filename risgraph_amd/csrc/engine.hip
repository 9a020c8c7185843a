// engine.hip -- incremental dependency-tree kernels (gfx950).
//
// Device restatement of the reference engine in core/graph.hpp:
//   seeding          update_tree_add single/batch      graph.hpp:1124-1218
//   push relax       do_update_tree_add                graph.hpp:1002-1116
//   delete roots     update_tree_del single/batch      graph.hpp:1622-1718
//   rescue           do_update_tree_del + has_parents  graph.hpp:1323-1361, 1889-1907
//   subtree expand   do_update_tree_del                graph.hpp:1364-1417
//   reset/pull/push  do_update_tree_del                graph.hpp:1420-1608
//   full build       build_tree                        graph.hpp:873-985
//   label update     update_label (16-B CAS form)      graph.hpp:1861-1887
//
// The label word is 64 bits: value << 32 | parent vertex.  A relax writes
// only when update_func strictly improves, as the reference's CAS form does:
// a CAS loop for SSWP and WCC, one 64-bit atomicMin for BFS and SSSP (whose
// parents always hold strictly smaller values), so parent pointers stay
// acyclic either way (relax() below).  Parent identity is the vertex only; a
// delete of any (p -> v) edge whose multiplicity reached 0 while v's parent is
// p marks v, and the rescue step re-parents v onto a surviving parallel edge.
//
// Work is level-synchronous over vertex queues.  A vertex pass gives each
// queued vertex a 16-lane group (4 lanes in grid-wide passes); lists longer
// than 32 entries are split into 128-entry chunk items handled a wave each in
// the pass's chunk stage, so hub vertices spread over many CUs.  All trees
// of a batch repair in one persistent launch (the team kernel): each tree
// gets W workgroups that separate passes with a device barrier, and a pass
// too large for the team leaves it for grid-wide launches (one per pass);
// the phase state machine lives in the device control block so the two
// modes hand over at any pass boundary.
#include "rg_internal.h"

#include <algorithm>

namespace rg {

// ---------------------------------------------------------------------
// algorithm lambdas (bfs_inc.cpp:77-92, sssp_inc.cpp:82-97,
// sswp_inc.cpp:82-97, wcc_inc.cpp:82-97)
// ---------------------------------------------------------------------
template <int A>
__device__ __forceinline__ uint64_t cand(uint32_t s, uint32_t w)
{
    if constexpr (A == BFS) return (uint64_t)s + 1;
    else if constexpr (A == SSSP) return (uint64_t)s + w;
    else if constexpr (A == SSWP) return s < w ? s : w;
    else return s;
}

template <int A>
__device__ __forceinline__ bool better(uint64_t c, uint32_t d)
{
    if constexpr (A == SSWP) return c > (uint64_t)d;
    else return c < (uint64_t)d;
}

__device__ __forceinline__ uint64_t pack(uint32_t v, uint32_t p) { return ((uint64_t)v << 32) | p; }
__device__ __forceinline__ uint32_t vval(uint64_t x) { return (uint32_t)(x >> 32); }
__device__ __forceinline__ uint32_t vpar(uint64_t x) { return (uint32_t)x; }

__device__ __forceinline__ uint32_t ald(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int32_t ald(const int32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void ast(uint32_t *p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void ast(int32_t *p, int32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Per-workgroup staging of queue appends in LDS.  Frontier, subtree and
// chunk-queue appends of a pass go to the workgroup's buffers with LDS
// atomics; at the end of the stage the workgroup reserves its range of each
// global queue with one counter atomic and copies the buffer out.  A full
// buffer spills straight to the global queue.  (With thousands of waves
// appending, single-counter device atomics were the pass's critical path.)
constexpr uint32_t SQ_NEXT = 6144, SQ_TREE = 1536, SQ_CHUNK = 512;
struct Stage {
    unsigned long long e_next, e_live;   // builds: degree sums (frontier, whole graph)
    uint32_t n_next, n_tree, n_chunk;
    uint32_t rb_min, rb_max;   // value band of the deletion roots staged here
    uint32_t base[3];
    uint32_t next[SQ_NEXT];
    uint32_t tree[SQ_TREE];
    ChunkItem chunk[SQ_CHUNK];
};

// snapshot of the control block taken at a pass boundary
struct Snap {
    uint32_t phase, stage, level, n_cur, n_tree, n_chunk, n_ops;
    uint32_t sel_cur, sel_next, sel_tree, epoch, mepoch, seed, rbound, flags, work_max, seq;
    int32_t idx;
    Ctl *c;       // live control block: LDS copy for a one-workgroup team, else HBM
    Stage *stg;   // this workgroup's append buffers (LDS)
};

// Read in 8-byte pairs (11 loads instead of 18): every workgroup of a team
// takes the snapshot at once after a barrier, all on the block's one line.
__device__ __forceinline__ uint64_t ald2(const Ctl *c, uint32_t word)
{
    return __hip_atomic_load((const uint64_t *)((const uint32_t *)c + word), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void take_snap(Ctl *c, Snap &s)
{
    static_assert(offsetof(Ctl, phase) == 0 && offsetof(Ctl, level) == 8 && offsetof(Ctl, n_cur) == 16 &&
                      offsetof(Ctl, n_tree) == 24 && offsetof(Ctl, n_mod) == 32 && offsetof(Ctl, sel_cur) == 40 &&
                      offsetof(Ctl, sel_tree) == 48 && offsetof(Ctl, mepoch) == 56 && offsetof(Ctl, flags) == 64 &&
                      offsetof(Ctl, rbound) == 80 && offsetof(Ctl, seq) == 112,
                  "pairs below follow the control block's layout");
    const uint64_t p0 = ald2(c, 0), p1 = ald2(c, 2), p2 = ald2(c, 4), p3 = ald2(c, 6), p4 = ald2(c, 8);
    const uint64_t p5 = ald2(c, 10), p6 = ald2(c, 12), p7 = ald2(c, 14), p8 = ald2(c, 16), p10 = ald2(c, 20);
    const uint64_t p14 = ald2(c, 28);
    auto lo = [](uint64_t x) { return (uint32_t)x; };
    auto hi = [](uint64_t x) { return (uint32_t)(x >> 32); };
    s.c = c;
    s.phase = lo(p0);
    s.stage = hi(p0);
    s.level = lo(p1);
    s.seed = hi(p1);
    s.n_cur = lo(p2);
    s.n_tree = lo(p3);
    s.n_chunk = hi(p3);
    s.n_ops = hi(p4);
    s.sel_cur = lo(p5);
    s.sel_next = hi(p5);
    s.sel_tree = lo(p6);
    s.epoch = hi(p6);
    s.mepoch = lo(p7);
    s.idx = (int32_t)hi(p7);
    s.flags = lo(p8);
    s.rbound = lo(p10);
    s.work_max = hi(p10);
    s.seq = lo(p14);
}

// The same snapshot at the start of a grid-wide kernel, where the block only
// changes between launches: a launch begins with its caches invalidated, so
// plain loads (a few wide scalar loads per block) see the previous kernels'
// writes, and a grid of up to 1024 blocks does not queue 17 single-word
// reads each on the control block's line.
__device__ __forceinline__ void take_snap_launch(Ctl *c, Snap &s)
{
    const Ctl l = *c;
    s.c = c;
    s.phase = l.phase;
    s.stage = l.stage;
    s.level = l.level;
    s.n_cur = l.n_cur;
    s.n_tree = l.n_tree;
    s.n_chunk = l.n_chunk;
    s.n_ops = l.n_ops;
    s.sel_cur = l.sel_cur;
    s.sel_next = l.sel_next;
    s.sel_tree = l.sel_tree;
    s.epoch = l.epoch;
    s.mepoch = l.mepoch;
    s.seed = l.seed;
    s.rbound = l.rbound;
    s.flags = l.flags;
    s.work_max = l.work_max;
    s.idx = l.idx;
    s.seq = l.seq;
}

// Strict-improvement update of the packed label word (update_label, CAS
// form, graph.hpp:1861-1887); true when this call lowered (SSWP: raised) the
// value.  The callers have already seen the candidate beat a plain load.
//
// BFS, and SSSP over an edge of weight >= 1, take one 64-bit atomicMin on
// value << 32 | parent instead of a CAS loop (no retries when many edges
// relax one hub at once).  A candidate equal to the stored value can then
// replace the parent with a smaller id, but such a parent's value is
// strictly below the child's, so parent chains still strictly decrease and
// stay acyclic.  SSWP (min(s, w) can equal s), WCC (a child's value equals
// its parent's) and zero-weight SSSP edges keep the CAS loop, which never
// re-parents on a tie.
template <int A>
__device__ __forceinline__ bool relax(uint64_t *word, uint64_t c, uint32_t parent, bool tie_acyclic)
{
    unsigned long long *p = (unsigned long long *)word;
    if constexpr (A == BFS || A == SSSP) {
        if (tie_acyclic) {
            const unsigned long long prev = atomicMin(p, (unsigned long long)pack((uint32_t)c, parent));
            return c < (uint64_t)vval(prev);
        }
    }
    unsigned long long cur = *p;
    while (better<A>(c, vval(cur))) {
        unsigned long long nw = pack((uint32_t)c, parent);
        unsigned long long prev = atomicCAS(p, cur, nw);
        if (prev == cur) return true;
        cur = prev;
    }
    return false;
}

// one counter atomic per wave: the active lanes (all at this call site, so all
// appending to the same queue) take consecutive positions
__device__ __forceinline__ uint32_t wave_reserve(uint32_t *count)
{
    const uint64_t m = __ballot(1);
    const uint32_t lane = __lane_id();
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1;
    const uint32_t rank = (uint32_t)__popcll(m & ((1ull << lane) - 1));
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
    base = __shfl(base, (int)leader);
    return base + rank;
}

// bounded append to a vertex queue (capacity V; dedupe keeps it in range)
__device__ __forceinline__ void q_append(const EngineArgs &a, Ctl *c, uint32_t *q, uint32_t *count, uint32_t v)
{
    uint32_t pos = wave_reserve(count);
    if (pos < a.V) q[pos] = v;
    else atomicOr(&c->flags, F_QUEUE_OVERFLOW);
}

// append to the next frontier / the subtree list through the LDS stage
__device__ __forceinline__ void stage_next(const EngineArgs &a, const Snap &s, uint32_t v)
{
    const uint32_t pos = wave_reserve(&s.stg->n_next);
    if (pos < SQ_NEXT) s.stg->next[pos] = v;
    else q_append(a, s.c, a.q[s.sel_next], &s.c->n_next, v);
}

__device__ __forceinline__ void stage_tree(const EngineArgs &a, const Snap &s, uint32_t v)
{
    const uint32_t pos = wave_reserve(&s.stg->n_tree);
    if (pos < SQ_TREE) s.stg->tree[pos] = v;
    else q_append(a, s.c, a.q[s.sel_tree], &s.c->n_tree, v);
}

__device__ void stage_init(Stage *st)
{
    if (threadIdx.x == 0) {
        st->n_next = st->n_tree = st->n_chunk = 0;
        st->e_next = st->e_live = 0;
        st->rb_min = 0xFFFFFFFFu;
        st->rb_max = 0;
    }
    __syncthreads();
}

// end of a stage: every thread of the workgroup calls this
__device__ void stage_flush(const EngineArgs &a, const Snap &s)
{
    Stage *st = s.stg;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t nn = min(st->n_next, SQ_NEXT), nt = min(st->n_tree, SQ_TREE);
        const uint32_t nc = min(st->n_chunk, SQ_CHUNK);
        st->base[0] = nn ? atomicAdd(&s.c->n_next, nn) : 0;
        st->base[1] = nt ? atomicAdd(&s.c->n_tree, nt) : 0;
        st->base[2] = nc ? atomicAdd(&s.c->n_chunk, nc) : 0;
        // one band update per workgroup instead of one per deletion root
        if (st->rb_min != 0xFFFFFFFFu) atomicMin(&s.c->rbound, st->rb_min);
        if (st->rb_max != 0) atomicMax(&s.c->rbound, st->rb_max);
        if (st->e_next) atomicAdd(&s.c->e_next, st->e_next);
        if (st->e_live) atomicAdd(&s.c->e_live, st->e_live);
        st->e_next = st->e_live = 0;
        st->rb_min = 0xFFFFFFFFu;
        st->rb_max = 0;
        st->n_next = nn;
        st->n_tree = nt;
        st->n_chunk = nc;
    }
    __syncthreads();
    bool over = false;
    uint32_t *qn = a.q[s.sel_next], *qt = a.q[s.sel_tree];
    for (uint32_t i = threadIdx.x; i < st->n_next; i += blockDim.x) {
        const uint32_t p = st->base[0] + i;
        if (p < a.V) qn[p] = st->next[i];
        else over = true;
    }
    for (uint32_t i = threadIdx.x; i < st->n_tree; i += blockDim.x) {
        const uint32_t p = st->base[1] + i;
        if (p < a.V) qt[p] = st->tree[i];
        else over = true;
    }
    for (uint32_t i = threadIdx.x; i < st->n_chunk; i += blockDim.x) {
        const uint32_t p = st->base[2] + i;
        if (p < a.cq_cap) a.cq[p] = st->chunk[i];
        else over = true;
    }
    if (over) atomicOr(&s.c->flags, F_QUEUE_OVERFLOW);
    __syncthreads();
    if (threadIdx.x == 0) st->n_next = st->n_tree = st->n_chunk = 0;
    __syncthreads();
}

// live edges a push from v goes over (graph.hpp:327-331 sums the out-degrees
// of the active vertices)
__device__ __forceinline__ uint32_t push_degree(const EngineArgs &a, uint32_t v)
{
    uint32_t d = 0;
    for (int li = 0; li < a.npush; li++) d += a.lists[a.push_l[li]].h[v].deg;
    return d;
}

__device__ __forceinline__ void push_next(const EngineArgs &a, const Snap &s, uint32_t v)
{
    if (atomicExch(&a.vr[v].stamp, s.epoch) != s.epoch) {
        stage_next(a, s, v);
        // a build sizes its next pass by the frontier's edges
        if (s.flags & F_BUILD) atomicAdd(&s.stg->e_next, (unsigned long long)push_degree(a, v));
    }
}

__device__ __forceinline__ void mark_mod(const EngineArgs &a, const Snap &s, uint32_t v)
{
    if (a.trace && atomicExch(&a.mstamp[v], s.mepoch) != s.mepoch) {
        uint32_t pos = wave_reserve(&s.c->n_mod);
        if (pos < a.V) a.mlist[pos] = v;
    }
}

enum ChunkOp : uint32_t { CO_PUSH = 0, CO_RESCUE = 1, CO_EXPAND = 2, CO_PULL = 3 };

__device__ __forceinline__ uint32_t scan_rot(uint32_t d, int32_t idx);

// the first `len` entries of v's list at slot `off`, as chunk items; a
// rescue scan's items start at a rotated chunk (scan_rot), since the first
// items of the queue are the first to run
__device__ __forceinline__ void defer_chunks(const EngineArgs &a, const Snap &s, uint32_t v, uint64_t off,
                                             uint32_t len, ChunkOp op)
{
    const uint32_t nch = (len + CHUNK - 1) / CHUNK;
    const uint32_t c0 = op == CO_RESCUE ? scan_rot(v, s.idx) % nch : 0u;
    auto item = [&](uint32_t k) {
        const uint32_t ck = k + c0 < nch ? k + c0 : k + c0 - nch;
        ChunkItem it;
        it.off = off + (uint64_t)ck * CHUNK;
        it.v = v;
        it.n_op = min(CHUNK, len - ck * CHUNK) | ((uint32_t)op << 16);
        return it;
    };
    // the items that fit go to the stage, the rest straight to the queue
    const uint32_t sp = atomicAdd(&s.stg->n_chunk, nch);
    const uint32_t fit = sp < SQ_CHUNK ? min(nch, SQ_CHUNK - sp) : 0;
    for (uint32_t k = 0; k < fit; k++) s.stg->chunk[sp + k] = item(k);
    if (fit == nch) return;
    const uint32_t rest = nch - fit;
    const uint32_t pos = atomicAdd(&s.c->n_chunk, rest);
    if (pos + rest > a.cq_cap) {
        atomicOr(&s.c->flags, F_QUEUE_OVERFLOW);
        return;
    }
    for (uint32_t k = fit; k < nch; k++) a.cq[pos + k - fit] = item(k);
}

// Invalidation marks (graph.hpp:1729) carry the deletion run's index: idx =
// invalid in this run, -idx = rescued in this run.  Marks of earlier runs are
// left in place and read as "not marked" (runs get fresh indices; the host
// clears the array when the index wraps).
__device__ __forceinline__ bool marked(int32_t iv, int32_t idx) { return iv == idx || iv == -idx; }

// mark x invalid unless this run already marked it: run indices only grow
// (until the host clears the marks on wrap), so one atomicMax replaces any
// older mark, 0, or an older rescue mark.  A vertex rescued in this run
// (-idx, then 0) is never a child of a vertex being expanded: its parent is
// its new parent outside the subtree.
__device__ __forceinline__ bool mark_inv(const EngineArgs &a, const Snap &s, uint32_t x)
{
    return atomicMax(&a.vr[x].inv, s.idx) != s.idx;
}

// has_parents (graph.hpp:1889-1907): true when an ancestor of x (x
// included) carries a mark of this run.  Values are monotone along tree
// paths (BFS/SSSP non-decreasing, SSWP non-increasing away from the root),
// and every marked vertex descends from a deletion root, so an ancestor whose
// value is below the smallest root value (above the largest, for SSWP) has
// no marked ancestor either and the walk stops there.  A walk is given up
// after RESCUE_WALK steps and the candidate treated as invalid (the vertex
// then stays in the subtree and is re-pulled: more work, same values); the
// tuning's walk_limit, when smaller, keeps the reference's 0.1*V guard and
// asks for a full rebuild instead.
//
// Walks are remembered.  Every deletion root is marked before the first
// rescue, and parents change during the deletion levels only for marked
// vertices (rescue), so a vertex whose ancestor chain ends at the tree root,
// at a vertex outside the value band or at a vertex already proven valid,
// with no mark on the way, can never be marked in this run.  Such a walk
// stamps the vertices it passed with valid_mark(idx), and later walks stop
// at the first stamp.  A rescued vertex takes the stamp too (its new parent
// was proven valid).  The stamp is negative: it reads as "not marked" and
// loses to any later run's index in mark_inv's atomicMax.
constexpr uint32_t RESCUE_WALK = 64;
constexpr int32_t VALID_BIT = 0x40000000;   // run indices stay below it

__device__ __forceinline__ int32_t valid_mark(int32_t idx) { return -(idx | VALID_BIT); }

template <int A>
__device__ bool has_parents(const EngineArgs &a, const Snap &s, uint32_t x, int32_t idx)
{
    const int32_t vm = valid_mark(idx);
    const uint32_t x0 = x;
    uint32_t steps = 0;
    for (;;) {
        const uint64_t w = a.vr[x].val;
        const int32_t iv = a.vr[x].inv;
        if (iv == vm) break;
        if constexpr (A == SSWP) {
            if (vval(w) > s.rbound) break;
        } else if constexpr (A != WCC) {
            if (vval(w) < s.rbound) break;
        }
        if (marked(iv, idx)) return true;
        uint32_t p = vpar(w);
        if (p == NILP) break;
        x = p;
        if (++steps > a.tune.walk_limit) {
            atomicOr(&s.c->flags, F_NEED_REBUILD);
            return true;
        }
        if (steps > RESCUE_WALK) return true;
    }
    // x proved the chain valid: stamp the vertices below it (their parents
    // are fixed for the rest of the run, so this retraces the same path)
    uint32_t y = x0;
    for (uint32_t k = 0; k < steps && y != x; k++) {
        const uint32_t p = vpar(a.vr[y].val);
        a.vr[y].inv = vm;
        y = p;
    }
    return false;
}

// ---------------------------------------------------------------------
// per-entry work shared by vertex groups (16 or 4 lanes) and chunk waves (64)
// ---------------------------------------------------------------------
template <int A>
__device__ __forceinline__ void entry_w(const EngineArgs &a, uint64_t slot, uint32_t &w)
{
    if constexpr (A == SSSP || A == SSWP) w = a.slots.wt[slot];
    else w = 0;
}

template <int A>
__device__ void push_range(const EngineArgs &a, const Snap &s, uint32_t u, uint32_t uv, uint64_t off,
                           uint32_t b, uint32_t e, int lane, int stride)
{
    for (uint32_t i = b + lane; i < e; i += stride) {
        uint32_t n = a.slots.nbr[off + i];
        if (n & DEAD) continue;
        uint32_t w;
        entry_w<A>(a, off + i, w);
        uint64_t c = cand<A>(uv, w);
        if (!better<A>(c, vval(a.vr[n].val))) continue;
        if (relax<A>(&a.vr[n].val, c, u, A == BFS || w != 0)) {
            push_next(a, s, n);
            mark_mod(a, s, n);
        }
    }
}

// rescue (graph.hpp:1323-1361): re-parent d onto a live in-edge whose source
// gives d's current value and lies outside the invalidated subtree.  The
// winning lane returns true.  The mark goes idx -> -idx -> valid_mark(idx):
// once d has its new parent it is a valid ancestor again for later rescues
// (a reader that still sees -idx, or the stamp with the old parent, backs
// off or walks the old chain to a mark, which is safe).
//
// SSWP also collects in `tb` the best value a valid in-neighbour offers below
// d's value: d's new value is at least that (the neighbour's tree path
// survives the run), which lets the expand leave alone the children whose
// value does not exceed it.
//
// A scan starts at a per-run hash of d and wraps around (the reference scans
// from the front, graph.hpp:1330-1359).  List positions follow the order in
// which keys first entered the list, so a front-first scan makes rescues
// adopt the oldest entries; a sliding window over a circular stream deletes
// those together when it comes round to them, and the adopted tree edges
// then fall in one burst (on R-MAT scale 24, every circle of the stream
// brought a stretch with two to three times the usual share of unsafe
// updates).  Every candidate the test accepts gives d the same value.
__device__ __forceinline__ uint32_t scan_rot(uint32_t d, int32_t idx)
{
    uint32_t h = d * 0x9E3779B1u ^ (uint32_t)idx * 0x85EBCA77u;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    h *= 0x297A2D39u;
    h ^= h >> 15;
    return h;
}

template <int A>
__device__ bool rescue_range(const EngineArgs &a, const Snap &s, uint32_t d, uint32_t dv, uint64_t off,
                             uint32_t b, uint32_t e, int lane, int stride, uint32_t &tb)
{
    const uint32_t len = e - b;
    const uint32_t rot = len ? scan_rot(d, s.idx) % len : 0u;
    for (uint32_t j = (uint32_t)lane; j < len; j += stride) {
        const uint32_t i = b + (j + rot < len ? j + rot : j + rot - len);
        if (a.vr[d].inv != s.idx) break;
        uint32_t n = a.slots.nbr[off + i];
        if (n & DEAD) continue;
        uint32_t w;
        entry_w<A>(a, off + i, w);
        // the weight alone rules most entries out, before their label load:
        // SSSP needs val[n] + w == dv, so w <= dv; SSWP min(val[n], w) is at
        // most w, which must reach dv, or pass tb for the bound; BFS needs
        // val[n] == dv - 1
        if constexpr (A == SSSP) {
            if (w > dv) continue;
        } else if constexpr (A == SSWP) {
            if (w < dv && w <= tb) continue;
        } else if constexpr (A == BFS) {
            // value 0 is the root's alone: a vertex at level 1 has no other parent
            if (dv == 1 && n != a.root) continue;
        }
        const uint64_t nv = a.vr[n].val;
        // a child of d hangs below d's mark: its walk would end at d at once
        if (vpar(nv) == d) continue;
        const uint64_t c = cand<A>(vval(nv), w);
        if (c != (uint64_t)dv) {
            if constexpr (A == SSWP) {
                if (c > tb && c < dv && !has_parents<A>(a, s, n, s.idx)) tb = (uint32_t)c;
            }
            continue;
        }
        if (has_parents<A>(a, s, n, s.idx)) continue;
        if (atomicCAS(&a.vr[d].inv, s.idx, -s.idx) == s.idx) {
            a.vr[d].val = pack(dv, n);
            a.vr[d].inv = valid_mark(s.idx);
            return true;
        }
    }
    return false;
}

// children of x through live out-edges whose target still names x as parent
// join the next level (graph.hpp:1382-1417).
//
// SSWP: x's new value is at least bx (its subtree root's bound).  A child c
// with val[c] <= bx < val[x] has val[c] = w(x, c), so min(new val[x], w(x, c))
// still equals val[c]: c keeps its value and its parent, and so does its whole
// subtree (values only shrink along a path).  Such children are not marked;
// the others inherit bx.
// SSWP: a child reached over an entry of weight w <= bx has val[c] =
// min(val[x], w) <= bx, so that entry is passed over before the child's
// label load (another entry to the same child, the tree edge, still counts)
template <int A>
__device__ __forceinline__ bool expand_skip(const EngineArgs &a, uint64_t slot, uint32_t bx)
{
    if constexpr (A == SSWP) return a.slots.wt[slot] <= bx;
    else return false;
}

template <int A>
__device__ __forceinline__ bool expand_child(const EngineArgs &a, const Snap &s, uint32_t x, uint32_t c,
                                             uint32_t bx)
{
    const uint64_t cv = a.vr[c].val;
    if (vpar(cv) != x) return false;
    if constexpr (A == SSWP) {
        if (vval(cv) <= bx) return false;
    }
    if (!mark_inv(a, s, c)) return false;
    if constexpr (A == SSWP) a.bnd[c] = bx;
    return true;
}

template <int A>
__device__ void expand_range(const EngineArgs &a, const Snap &s, uint32_t x, uint64_t off, uint32_t b,
                             uint32_t e, int lane, int stride, uint32_t bx)
{
    for (uint32_t i = b + lane; i < e; i += stride) {
        uint32_t c = a.slots.nbr[off + i];
        if ((c & DEAD) || expand_skip<A>(a, off + i, bx)) continue;
        if (expand_child<A>(a, s, x, c, bx)) stage_next(a, s, c);
    }
}

// best candidate word over a range (min or max of value<<32|parent); in-
// neighbours inside the invalidated subtree are skipped, their values are
// being reset in the same pass (a build's dense passes have no subtree: its
// init pass cleared every mark)
template <int A>
__device__ __forceinline__ uint64_t reduce_best(uint64_t x, int width)
{
    for (int m = width / 2; m > 0; m >>= 1) {
        uint64_t y = __shfl_xor((unsigned long long)x, m, width);
        if (A == SSWP ? y > x : y < x) x = y;
    }
    return x;
}

// Steps are uniform across the G-lane group (the bounds depend on the list
// only).  Between steps the group shares its best candidate, and an entry
// whose weight alone shows it cannot beat that is skipped before its label
// load: an SSSP candidate val + w is at least w, an SSWP min(val, w) at most
// w (strictly, so ties still compare parents).
template <int A>
__device__ uint64_t pull_range(const EngineArgs &a, const Snap &s, uint64_t off, uint32_t b, uint32_t e,
                               int lane, int stride)
{
    uint64_t best = (A == SSWP) ? 0ull : ~0ull;
    for (uint32_t base = b; base < e; base += stride) {
        const uint32_t i = base + (uint32_t)lane;
        if (i < e) {
            const uint32_t n = a.slots.nbr[off + i];
            uint32_t w;
            entry_w<A>(a, off + i, w);
            bool skip = (n & DEAD) != 0;
            if constexpr (A == SSSP) skip |= w > vval(best);
            else if constexpr (A == SSWP) skip |= w < vval(best);
            if (!skip && !((s.flags & F_BUILD) == 0 && a.vr[n].inv == s.idx)) {
                const uint64_t c = cand<A>(vval(a.vr[n].val), w);
                if (A == SSWP || c <= 0xFFFFFFFFull) {
                    const uint64_t word = pack((uint32_t)c, n);
                    if (A == SSWP ? word > best : word < best) best = word;
                }
            }
        }
        if constexpr (A == SSSP || A == SSWP)
            if (base + stride < e) best = reduce_best<A>(best, stride);
    }
    return best;
}

template <int A>
__device__ __forceinline__ void pull_apply(const EngineArgs &a, const Snap &s, uint32_t x, uint64_t best)
{
    if (A == SSWP ? best == 0ull : best == ~0ull) return;
    if (relax<A>(&a.vr[x].val, vval(best), vpar(best), A == BFS)) {
        push_next(a, s, x);
        mark_mod(a, s, x);
    }
}

// any lane of the calling lane's G-lane group
template <int G>
__device__ __forceinline__ bool group_any(bool p)
{
    const uint64_t b = __ballot(p);
    const uint32_t base = __lane_id() & ~(uint32_t)(G - 1);
    return ((b >> base) & ((1ull << G) - 1)) != 0;
}

// ---------------------------------------------------------------------
// passes
// ---------------------------------------------------------------------
template <int A>
__device__ void seed_edge(const EngineArgs &a, const Snap &s, uint32_t src, uint32_t dst, uint32_t w)
{
    uint64_t c = cand<A>(vval(a.vr[src].val), w);
    if (!better<A>(c, vval(a.vr[dst].val))) return;
    if (relax<A>(&a.vr[dst].val, c, src, A == BFS || w != 0)) {
        push_next(a, s, dst);
        mark_mod(a, s, dst);
    }
}

// a root also widens the value band of the invalidated subtree (has_parents)
template <int A>
__device__ __forceinline__ void mark_root(const EngineArgs &a, const Snap &s, uint32_t src, uint32_t dst)
{
    const uint64_t w = a.vr[dst].val;
    if (vpar(w) != src) return;
    if (mark_inv(a, s, dst)) {
        if constexpr (A == SSWP) a.bnd[dst] = 0;   // no bound known yet
        if constexpr (A == SSWP) atomicMax(&s.stg->rb_max, vval(w));
        else if constexpr (A != WCC) atomicMin(&s.stg->rb_min, vval(w));
        stage_next(a, s, dst);
    }
}

// One thread per edge op.  SEED: relax the inserted edges that created a live
// entry (update_tree_add batch seeding, graph.hpp:1162-1218).  DEL_ROOTS: mark
// targets of deleted tree edges (update_tree_del batch, graph.hpp:1658-1714).
// DEL_PULL: seed the inserted edges of a mixed batch whose endpoints both lie
// outside the invalidated subtree (an edge touching the subtree is covered by
// the pull into its target or the push from its source).
template <int A>
__device__ void pass_ops(const EngineArgs &a, const Snap &s, uint32_t tid, uint32_t nthr)
{
    const bool pull = s.phase == PH_DEL_PULL;
    for (uint32_t i = tid; i < s.n_ops; i += nthr) {
        uint8_t k = a.ops.kind[i];
        uint32_t src = a.ops.src[i], dst = a.ops.dst[i];
        uint32_t w = (A == SSSP || A == SSWP) ? a.ops.w[i] : 0;
        bool both = a.undirected || (k & K_UNDIR);
        if (s.phase == PH_SEED || pull) {
            if (!(k & K_INS)) continue;
            // an insert reaches the engine when it created a live entry
            // (bfs_inc.cpp:162-165: add_edge returned 0 and the batch left it)
            const OpResult r = a.ops.res[i];
            const bool seed = (k & K_FORCE) || (r.old == 0 && r.live);
            if (!seed) continue;
            if (pull && (a.vr[src].inv == s.idx || a.vr[dst].inv == s.idx)) continue;
            seed_edge<A>(a, s, src, dst, w);
            if (both) seed_edge<A>(a, s, dst, src, w);
        } else {
            if (!(k & K_DEL)) continue;
            if (!((k & K_FORCE) || a.ops.res[i].old == 1)) continue;
            mark_root<A>(a, s, src, dst);
            if (both) mark_root<A>(a, s, dst, src);
        }
    }
}

// queue entries of the deletion levels: bit 31 = rescue already attempted
// (its in-list scan was deferred to chunk items), only expand
constexpr uint32_t Q_EXPAND_ONLY = 0x80000000u;

// Local continuation of the deletion levels: the children a group marks are
// processed by that group in the same pass, depth first, up to CONT_DEPTH
// levels below the queued vertex with at most CONT_STACK pending per group;
// the rest join the next level's queue.  Marking order does not matter for
// the rescue test (has_parents), since every invalidated vertex already has
// a marked deletion root above it once the roots pass is done; levels only
// bound the work per pass.
constexpr uint32_t CONT_DEPTH = 7;   // the most a stack word's 3 depth bits hold
constexpr uint32_t CONT_STACK = 32;
constexpr uint32_t MAX_GROUPS = 1024 / GROUP;   // per workgroup
constexpr uint32_t STACK_WORDS = MAX_GROUPS * 2 * CONT_STACK;
// Grid-wide passes use few large blocks (two per CU) with grid-stride loops:
// each block reserves its queue ranges with one counter atomic per stage,
// and same-address device atomics are the scarce resource here.  Their
// vertex stage gives each vertex 4 lanes instead of 16: a large pass has
// vertices to spare, and most lists are short, so narrower groups keep four
// times as many vertices' loads in flight (grid-wide pull, push and expand
// steps ran 10-25% faster than with 16 lanes on R-MAT scale 24).
constexpr uint32_t GRID_BLOCK = 1024;
constexpr uint32_t GRID_MAX_BLOCKS = 1024;
constexpr int GRID_GROUP = 4;
constexpr uint32_t GRID_CONT_STACK = 16;
constexpr uint32_t GRID_STACK_WORDS = (GRID_BLOCK / GRID_GROUP) * 2 * GRID_CONT_STACK;

// a group's LDS stack: vertices in v[0, CS), depths in d[...]; `top` is the
// same in all lanes of the group
struct GStack {
    uint32_t *v, *d;
    uint32_t top;
};

template <int G>
__device__ __forceinline__ uint32_t group_ballot(bool p)
{
    const uint64_t b = __ballot(p);
    return (uint32_t)((b >> (__lane_id() & ~(uint32_t)(G - 1))) & ((1ull << G) - 1));
}

// expand x's short list with the group's G lanes stepping through it
// together; a stacked child keeps its depth in the low 3 bits of its stack
// word and its SSWP bound above them
static_assert(CONT_DEPTH < 8, "depth fits 3 bits");
template <int A, int G, uint32_t CS>
__device__ void expand_cont(const EngineArgs &a, const Snap &s, uint32_t x, uint64_t off, uint32_t len,
                            int glane, GStack &st, uint32_t depth, uint32_t bx)
{
    for (uint32_t b = 0; b < len; b += G) {
        const uint32_t i = b + glane;
        bool mk = false;
        uint32_t c = 0;
        if (i < len) {
            c = a.slots.nbr[off + i];
            if (!(c & DEAD) && !expand_skip<A>(a, off + i, bx)) mk = expand_child<A>(a, s, x, c, bx);
        }
        const bool keep = mk && depth < CONT_DEPTH;
        const uint32_t kb = group_ballot<G>(keep);
        const uint32_t rank = (uint32_t)__popc(kb & ((1u << glane) - 1));
        const bool fits = keep && st.top + rank < CS;
        if (fits) {
            st.v[st.top + rank] = c;
            st.d[st.top + rank] = (depth + 1) | (A == SSWP ? bx << 3 : 0u);
        }
        if (mk && !fits) stage_next(a, s, c);
        st.top = min(CS, st.top + (uint32_t)__popc(kb));
    }
}

// one vertex of a deletion level: try to rescue it, else keep it invalid,
// record it for the pull, and mark its children
// `bv` is v's SSWP bound when the caller knows it (a stacked child), else
// NO_BOUND: read it from the bound array.
constexpr uint32_t NO_BOUND = 0xFFFFFFFFu;

template <int A, int G, uint32_t CS>
__device__ void del_vertex(const EngineArgs &a, const Snap &s, uint32_t raw, uint32_t depth, int glane,
                           GStack &st, uint32_t bv)
{
    const uint32_t v = raw & ~Q_EXPAND_ONLY;
    if (a.vr[v].inv != s.idx) return;   // rescued
    if constexpr (A == SSWP) {
        if (bv == NO_BOUND) bv = a.bnd[v];
    } else {
        bv = 0;
    }
    if (s.phase == PH_DEL_RESCUE && !(raw & Q_EXPAND_ONLY)) {
        const uint32_t dv = vval(a.vr[v].val);
        bool won = false, deferred = false;
        uint32_t tb = bv;
        for (int li = 0; li < a.npull; li++) {
            int L = a.pull_l[li];
            uint64_t off;
            uint32_t len;
            list_span(a.lists[L], v, off, len);
            if (len > LONG_LIST) {
                // a long in-list is searched whole, spread over chunk items
                // (items after a success return at once).  A hub that stays
                // invalid costs its subtree and its own pull: scanning only a
                // prefix of hub in-lists (1024 entries) made the mean batch
                // twice as long on R-MAT scale 24
                if (glane == 0) defer_chunks(a, s, v, off, len, CO_RESCUE);
                deferred = true;
                continue;
            }
            won |= rescue_range<A>(a, s, v, dv, off, 0, len, glane, G, tb);
        }
        if (group_any<G>(won)) return;
        if constexpr (A == SSWP) {
            // the group's best valid offer is a bound on v's new value; the
            // deferred scan of a long in-list may still raise it
            for (int m = G / 2; m > 0; m >>= 1) tb = max(tb, (uint32_t)__shfl_xor((int)tb, m, G));
            bv = tb;
            if (glane == 0) a.bnd[v] = bv;
        }
        if (deferred) {
            // the long in-list is scanned in this pass's chunk stage; v is
            // expanded at the next level unless that rescues it
            if (glane == 0) stage_next(a, s, v | Q_EXPAND_ONLY);
            return;
        }
    }
    if (glane == 0) stage_tree(a, s, v);
    for (int li = 0; li < a.npush; li++) {
        int L = a.push_l[li];
        uint64_t off;
        uint32_t len;
        list_span(a.lists[L], v, off, len);
        if (len > LONG_LIST) {
            if (glane == 0) defer_chunks(a, s, v, off, len, CO_EXPAND);
            continue;
        }
        expand_cont<A, G, CS>(a, s, v, off, len, glane, st, depth, bv);
    }
}

template <int A, int G, uint32_t CS>
__device__ void pass_vertices(const EngineArgs &a, const Snap &s, uint32_t tid, uint32_t nthr, uint32_t *stk)
{
    const uint32_t ph = s.phase;
    if (ph == PH_BUILD_INIT) {
        // the init pass also counts the live edges of the push lists: the
        // dense-pass rule compares the frontier's edges with a twentieth of them
        unsigned long long e = 0;
        for (uint32_t v = tid; v < a.V; v += nthr) {
            a.vr[v].val = pack(a.init[v], NILP);
            a.vr[v].inv = 0;
            bool act = a.init_active ? a.init_active[v] != 0 : v == a.root;
            if (act) push_next(a, s, v);
            if constexpr (G == GRID_GROUP) {
                if (a.tune.dense) e += push_degree(a, v);
            }
        }
        for (int m = WAVE / 2; m > 0; m >>= 1) e += __shfl_xor(e, m, WAVE);
        if (__lane_id() == 0 && e) atomicAdd(&s.stg->e_live, e);
        return;
    }
    const int glane = tid % G;
    const uint32_t gid = tid / G, ngr = nthr / G;
    // (build init and dense passes always run grid-wide)
    if constexpr (G == GRID_GROUP) if (ph == PH_PUSH && (s.flags & F_DENSE)) {
        // dense pass of a build (graph.hpp:380-394): every vertex takes the
        // best candidate over its live in-edges; long in-lists finish in
        // pull chunks
        for (uint32_t v = gid; v < a.V; v += ngr) {
            uint64_t best = (A == SSWP) ? 0ull : ~0ull;
            for (int li = 0; li < a.npull; li++) {
                int L = a.pull_l[li];
                uint64_t off;
                uint32_t len;
                list_span(a.lists[L], v, off, len);
                if (len > LONG_LIST) {
                    if (glane == 0) defer_chunks(a, s, v, off, len, CO_PULL);
                    continue;
                }
                uint64_t b = pull_range<A>(a, s, off, 0, len, glane, G);
                if (A == SSWP ? b > best : b < best) best = b;
            }
            best = reduce_best<A>(best, G);
            if (glane == 0) pull_apply<A>(a, s, v, best);
        }
        return;
    }
    const uint32_t n = (ph == PH_DEL_PULL) ? s.n_tree : s.n_cur;
    const uint32_t *qv = (ph == PH_DEL_PULL) ? a.q[s.sel_tree] : a.q[s.sel_cur];
    for (uint32_t gi = gid; gi < n; gi += ngr) {
        const uint32_t raw = qv[gi];
        const uint32_t v = raw & ~Q_EXPAND_ONLY;
        if (ph == PH_PUSH) {
            uint32_t uv = vval(a.vr[v].val);
            for (int li = 0; li < a.npush; li++) {
                int L = a.push_l[li];
                uint64_t off;
                uint32_t len;
                list_span(a.lists[L], v, off, len);
                if (len > LONG_LIST) {
                    if (glane == 0) defer_chunks(a, s, v, off, len, CO_PUSH);
                    continue;
                }
                push_range<A>(a, s, v, uv, off, 0, len, glane, G);
            }
        } else if (ph == PH_DEL_RESCUE || ph == PH_DEL_EXPAND) {
            // one deletion level (rescue then expand per vertex), continued
            // depth first through the children this group marks
            GStack st;
            st.v = stk + (threadIdx.x / G) * 2 * CS;
            st.d = st.v + CS;
            st.top = 0;
            del_vertex<A, G, CS>(a, s, raw, 0, glane, st, NO_BOUND);
            while (st.top) {
                st.top--;
                const uint32_t c = st.v[st.top], d = st.d[st.top];
                del_vertex<A, G, CS>(a, s, c, d & 7u, glane, st, d >> 3);
            }
        } else if (ph == PH_DEL_PULL) {
            // reset v and pull its best candidate from in-neighbours outside
            // the subtree (graph.hpp:1420-1510); long lists finish in chunks
            uint64_t best = (A == SSWP) ? 0ull : ~0ull;
            for (int li = 0; li < a.npull; li++) {
                int L = a.pull_l[li];
                uint64_t off;
                uint32_t len;
                list_span(a.lists[L], v, off, len);
                if (len > LONG_LIST) {
                    if (glane == 0) defer_chunks(a, s, v, off, len, CO_PULL);
                    continue;
                }
                uint64_t b = pull_range<A>(a, s, off, 0, len, glane, G);
                if (A == SSWP ? b > best : b < best) best = b;
            }
            best = reduce_best<A>(best, G);
            if (glane == 0) {
                const uint32_t iv = a.init[v];
                const bool take = (A == SSWP) ? (best != 0ull && vval(best) > iv)
                                              : (best != ~0ull && vval(best) < iv);
                a.vr[v].val = take ? best : pack(iv, NILP);
                // a vertex left at its init label improves nobody (except
                // under WCC, where every init label is a candidate): it is
                // pushed once a pull chunk or a push gives it a value
                if (A == WCC || take) push_next(a, s, v);
                mark_mod(a, s, v);
            }
        }
    }
}

// The vertex stage of a pass: op passes, vertex passes, and the mixed
// pull pass (tree vertices plus the seeds of the batch's inserts).
template <int A, int G, uint32_t CS>
__device__ __forceinline__ void pass_main(const EngineArgs &a, const Snap &s, uint32_t tid, uint32_t nthr,
                                          uint32_t *stk)
{
    if (s.phase == PH_SEED || s.phase == PH_DEL_ROOTS) {
        pass_ops<A>(a, s, tid, nthr);
        return;
    }
    pass_vertices<A, G, CS>(a, s, tid, nthr, stk);
    if (s.phase == PH_DEL_PULL && s.seed) pass_ops<A>(a, s, tid, nthr);
}

template <int A>
__device__ void pass_chunks(const EngineArgs &a, const Snap &s, uint32_t tid, uint32_t nthr, uint32_t n_chunk)
{
    const uint32_t ph = s.phase;
    if (ph != PH_PUSH && ph != PH_DEL_RESCUE && ph != PH_DEL_EXPAND && ph != PH_DEL_PULL) return;
    const int lane = tid % WAVE;
    const uint32_t wid = tid / WAVE, nw = nthr / WAVE;
    for (uint32_t ci = wid; ci < n_chunk; ci += nw) {
        const ChunkItem it = a.cq[ci];
        const uint32_t v = it.v, op = it.n_op >> 16;
        const uint64_t off = it.off;
        const uint32_t b = 0, e = it.n_op & 0xFFFFu;
        if (op == CO_PUSH) {
            push_range<A>(a, s, v, vval(a.vr[v].val), off, b, e, lane, WAVE);
        } else if (op == CO_RESCUE) {
            if (a.vr[v].inv != s.idx) continue;
            // SSWP: offers at or under v's bound so far add nothing
            uint32_t tb = A == SSWP ? a.bnd[v] : 0u;
            rescue_range<A>(a, s, v, vval(a.vr[v].val), off, b, e, lane, WAVE, tb);
            if constexpr (A == SSWP) {
                for (int m = WAVE / 2; m > 0; m >>= 1) tb = max(tb, (uint32_t)__shfl_xor((int)tb, m, WAVE));
                if (lane == 0 && tb) atomicMax(&a.bnd[v], tb);
            }
        } else if (op == CO_EXPAND) {
            expand_range<A>(a, s, v, off, b, e, lane, WAVE, A == SSWP ? a.bnd[v] : 0u);
        } else {
            uint64_t best = pull_range<A>(a, s, off, b, e, lane, WAVE);
            best = reduce_best<A>(best, WAVE);
            if (lane == 0) pull_apply<A>(a, s, v, best);
        }
    }
}

// phase transitions, run by one thread after a pass completed.  A batch with
// deletions runs them first (roots, levels, pull), and its inserts are seeded
// in the pull pass, so one push phase settles both; values are the unique
// fixpoint either way, as in the reference's add-then-delete batch order.
__device__ __forceinline__ void advance_state(const EngineArgs &a, Ctl &c)
{
    const uint32_t ph = c.phase;
    auto swap_queues = [&]() {
        const uint32_t cur = c.sel_cur;
        c.sel_cur = c.sel_next;
        c.sel_next = cur;
        c.n_cur = c.n_next;
        c.n_next = 0;
        c.epoch += 1;
        c.levels += 1;
    };
    auto rescue_on = [&](uint32_t level, uint32_t n) {
        return (a.tune.rescue_level != RESCUE_NEVER && level <= a.tune.rescue_level) || n < a.tune.rescue_frontier;
    };
    auto recompute = [&](uint32_t level, uint32_t n) {
        uint64_t tot = (uint64_t)c.n_tree + n;
        return level <= a.tune.recompute_level && (double)tot > (double)a.tune.recompute_frac * a.V &&
               tot > a.tune.recompute_abs;
    };
    auto level_next = [&](uint32_t level) {
        const uint32_t n = c.n_cur;
        if (c.flags & F_NEED_REBUILD) c.phase = PH_REBUILD;
        else if (n == 0) c.phase = PH_DEL_PULL;
        else if (recompute(level, n)) c.phase = PH_REBUILD;
        else c.phase = rescue_on(level, n) ? PH_DEL_RESCUE : PH_DEL_EXPAND;
    };
    c.n_chunk = 0;
    c.stage = 0;
    switch (ph) {
    case PH_SEED:
    case PH_BUILD_INIT:
    case PH_PUSH:
    case PH_DEL_PULL:
        swap_queues();
        c.seed = 0;
        c.phase = c.n_cur ? PH_PUSH : PH_DONE;
        if (c.flags & F_BUILD) {
            // graph.hpp:334: a pass goes dense when its frontier's edges
            // reach both the vertex count and a twentieth of the live edges
            const unsigned long long e = c.e_next;
            c.e_next = 0;
            const bool dense = a.tune.dense && c.phase == PH_PUSH && a.npull > 0 && e >= a.V && e >= c.e_live / 20;
            c.flags = dense ? (c.flags | F_DENSE) : (c.flags & ~F_DENSE);
            if (dense) c.dense += 1;
        }
        break;
    case PH_DEL_ROOTS:
        swap_queues();
        c.level = 0;
        c.n_tree = 0;
        if (c.n_cur == 0) c.phase = c.seed ? PH_SEED : PH_DONE;
        else level_next(0);
        break;
    case PH_DEL_RESCUE:
    case PH_DEL_EXPAND:
        swap_queues();
        c.level += 1;
        level_next(c.level);
        break;
    default:
        // unknown phase: stop instead of spinning
        c.flags |= F_QUEUE_OVERFLOW;
        c.phase = PH_DONE;
        break;
    }
}

// The control block is read once (independent loads, one round trip), the
// transition runs on a private copy in registers, and the words that changed
// go back.  Nothing else touches the block meanwhile: the caller is the last
// arriver of a team barrier, or the single thread of the grid-mode advance
// kernel.
__device__ void advance(const EngineArgs &a, Ctl *c)
{
    Ctl l;
    l.phase = ald(&c->phase);
    l.stage = ald(&c->stage);
    l.level = ald(&c->level);
    l.seed = ald(&c->seed);
    l.n_cur = ald(&c->n_cur);
    l.n_next = ald(&c->n_next);
    l.n_tree = ald(&c->n_tree);
    l.n_chunk = ald(&c->n_chunk);
    l.sel_cur = ald(&c->sel_cur);
    l.sel_next = ald(&c->sel_next);
    l.epoch = ald(&c->epoch);
    l.flags = ald(&c->flags);
    l.levels = c->levels;
    l.e_next = __hip_atomic_load(&c->e_next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    l.e_live = __hip_atomic_load(&c->e_live, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    l.dense = c->dense;
    const Ctl o = l;
    advance_state(a, l);
    ast(&c->seq, ald(&c->seq) + 1u);
    if (l.phase != o.phase) ast(&c->phase, l.phase);
    if (l.stage != o.stage) ast(&c->stage, l.stage);
    if (l.level != o.level) ast(&c->level, l.level);
    if (l.seed != o.seed) ast(&c->seed, l.seed);
    if (l.n_cur != o.n_cur) ast(&c->n_cur, l.n_cur);
    if (l.n_next != o.n_next) ast(&c->n_next, l.n_next);
    if (l.n_tree != o.n_tree) ast(&c->n_tree, l.n_tree);
    if (l.n_chunk != o.n_chunk) ast(&c->n_chunk, l.n_chunk);
    if (l.sel_cur != o.sel_cur) ast(&c->sel_cur, l.sel_cur);
    if (l.sel_next != o.sel_next) ast(&c->sel_next, l.sel_next);
    if (l.epoch != o.epoch) ast(&c->epoch, l.epoch);
    if (l.flags != o.flags) ast(&c->flags, l.flags);
    c->levels = l.levels;
    if (l.e_next != o.e_next) __hip_atomic_store(&c->e_next, l.e_next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    c->dense = l.dense;
}

// work of the current pass, in units comparable with the team cap; an op
// pass (one thread per op, no list scans) counts 1/16 per op
__device__ __forceinline__ uint32_t phase_work(const EngineArgs &a, const Snap &s)
{
    switch (s.phase) {
    case PH_SEED:
    case PH_DEL_ROOTS: return s.n_ops / 16;
    case PH_DEL_PULL: return s.n_tree + (s.seed ? s.n_ops / 16 : 0);
    case PH_BUILD_INIT: return a.V;
    case PH_PUSH: return (s.flags & F_DENSE) ? a.V : s.n_cur;
    default: return s.n_cur;
    }
}

// the tree's current pass runs grid-wide: a full rebuild, a chunk stage
// handed over by the team, or more work than the team takes
__device__ __forceinline__ bool grid_needed(const EngineArgs &a, const Snap &s)
{
    if (s.phase == PH_DONE || s.phase == PH_REBUILD) return false;
    if (s.flags & (F_QUEUE_OVERFLOW | F_ABORT)) return false;   // the host reports these
    return s.phase == PH_BUILD_INIT || (s.phase == PH_PUSH && (s.flags & F_DENSE)) || s.stage != 0 ||
           phase_work(a, s) > s.work_max;
}

// copy the control block to the pinned host mirror (one thread), so the
// host reads the state after a stream sync without a copy launch
__device__ void publish(const EngineArgs &a, const Ctl *c)
{
    if (!a.hctl) return;
    const uint32_t *src = (const uint32_t *)c;
    uint32_t *dst = (uint32_t *)a.hctl;
    for (uint32_t i = 0; i < sizeof(Ctl) / 4; i++) dst[i] = ald(src + i);
    __threadfence_system();
}

// ---------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------
// Team barrier over the W workgroups of one tree (thread 0 of each
// workgroup arrives).  Workgroup wg arrives on sub-counter wg % ns (ns =
// min(W, BAR_SUBS), each on its own cache line); the last arrival of a
// sub-group arrives on the control block's counter, so a barrier costs about
// W/ns + ns serialised same-address atomics instead of W.  Counters grow
// monotonically within a launch, so episode e of a sub-group of size m is
// complete when its count reaches e*m; the last arriver overall runs `last`
// (the pass's serial step) and then releases the episode by writing it into
// a release word on every sub-counter's line.  Each waiter spins on its own
// sub-group's line, so the spinning reads stay off the control block, whose
// line the serial step and the queue counters use.  Waiting is bounded in
// wall-clock time: a timeout sets F_ABORT, which every waiter of the tree
// checks every 64 spins, so all waves leave the launch and the host reports
// an error.
constexpr uint64_t TEAM_SPIN_TICKS = 200000000ull;   // 2 s at the 100 MHz wall clock

template <class F>
__device__ __forceinline__ void team_sync(Ctl *c, uint32_t *bar, uint32_t W, uint32_t wg, F &&last)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        if (W == 1) {
            last();
        } else {
            const uint32_t ns = W < BAR_SUBS ? W : BAR_SUBS;
            const uint32_t sub = wg % ns;
            const uint32_t m = W / ns + (sub < W % ns ? 1u : 0u);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            const uint32_t old = atomicAdd(&bar[sub * BAR_STRIDE], 1u);
            const uint32_t ep = old / m + 1;
            bool release = false;
            if (old % m == m - 1) {
                // the sub-group is complete: pass its writes on to the top count
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
                const uint32_t top = atomicAdd(&c->bar_count, 1u);
                release = top % ns == ns - 1;
            }
            if (release) {
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                last();
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                for (uint32_t k = 0; k < ns; k++)
                    __hip_atomic_store(&bar[k * BAR_STRIDE + 1], ep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                const uint32_t *gen = &bar[sub * BAR_STRIDE + 1];
                const uint64_t t0 = wall_clock64();
                uint32_t spins = 0;
                while ((int32_t)(ald(gen) - ep) < 0) {
                    __builtin_amdgcn_s_sleep(1);
                    if ((++spins & 63u) != 0) continue;
                    if (ald(&c->flags) & F_ABORT) break;
                    if (wall_clock64() - t0 > TEAM_SPIN_TICKS) {
                        atomicOr(&c->flags, F_ABORT);
                        break;
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
    }
    __syncthreads();
}

struct TeamParams {
    const EngineArgs *a[MAXT];   // device copies of the trees' arguments
    uint32_t k[MAXT];            // XCD slots of each tree
    uint32_t work_max[MAXT];     // a larger pass leaves the team for grid mode
    uint32_t chunk_max[MAXT];    // more chunk items than this leave for grid mode
    uint32_t rounds;             // workgroups per XCD slot
    uint32_t init_mask;
    // per slot s, bits 4s..4s+3: its tree (0xF: none) and its rank among
    // the tree's slots (packed: a byte array indexed by the slot would be
    // copied to private memory)
    uint32_t slot_tree, slot_rank;
};

// The level-synchronous loop of one tree on its W workgroups.  With W == 1
// the control block lives in LDS for the whole launch (workgroup-local
// atomics, no barrier traffic) and goes back to HBM at exit; with W > 1 it
// stays in HBM and passes are separated by the team barrier.
template <int A>
__device__ __forceinline__ void team_run(const EngineArgs &a, uint32_t W, uint32_t work_max, uint32_t chunk_max,
                                         uint32_t wg, bool init, Ctl *lds_ctl, Snap &snap, uint32_t *sh,
                                         uint32_t *stk, Stage *stg)
{
    stage_init(stg);
    const uint32_t tid = wg * blockDim.x + threadIdx.x, nthr = W * blockDim.x;
    Ctl *c = (W == 1) ? lds_ctl : a.ctl;
    if (threadIdx.x == 0 && (W == 1 || (init && wg == 0))) {
        // a run's first launch takes its initial control block from the
        // pinned host mirror (no separate copy command); the barrier words
        // are left alone, other workgroups may already be arriving
        const uint32_t *src = (const uint32_t *)(init ? a.hctl : a.ctl);
        uint32_t *dst = (uint32_t *)c;
        const uint32_t b0 = offsetof(Ctl, bar_count) / 4, b2 = offsetof(Ctl, bar_exit) / 4;
        for (uint32_t i = 0; i < sizeof(Ctl) / 4; i++) {
            if (W == 1) dst[i] = ald(src + i);
            else if (i < b0 || i > b2) ast(dst + i, src[i]);
        }
    }
    if (W > 1 && init) team_sync(c, a.bar, W, wg, [] {});
    // every pass either finishes the phase or hands it to grid mode; the cap
    // only bounds the launch if the state machine were ever corrupted
    for (uint32_t iter = 0; iter < (1u << 24); iter++) {
        __syncthreads();
        if (threadIdx.x == 0) {
            take_snap(c, snap);
            snap.stg = stg;
        }
        __syncthreads();
        const Snap s = snap;
        if (s.phase == PH_DONE || s.phase == PH_REBUILD || s.phase == PH_BUILD_INIT) break;
        if (s.flags & F_ABORT) break;
        if (phase_work(a, s) > work_max) break;
        if (s.phase == PH_PUSH && (s.flags & F_DENSE)) break;   // dense passes run grid-wide
        // stage clocks of this workgroup: start, vertex stage done, first
        // barrier passed, chunk stage done
        uint64_t tk[4];
        tk[0] = wall_clock64();
        tk[1] = tk[2] = tk[0];
        // the pass's serial step: statistics, then the phase transition
        auto close_pass = [&] {
            if (a.ticks && s.phase < NPHASE) {
                a.ticks[s.phase] += wall_clock64() - tk[0];
                a.ticks[NPHASE + s.phase] += 1;
            }
            advance(a, c);
        };
        if (s.stage == 0) {
            pass_main<A, GROUP, CONT_STACK>(a, s, tid, nthr, stk);
            stage_flush(a, s);
            tk[1] = wall_clock64();
            // the chunk queue is complete once every workgroup finished; a
            // pass without chunk items closes right here
            team_sync(c, a.bar, W, wg, [&] {
                const uint32_t nch = ald(&c->n_chunk);
                if (nch > chunk_max) ast(&c->stage, 1u);
                else if (nch == 0) close_pass();
            });
            tk[2] = wall_clock64();
            if (threadIdx.x == 0) {
                sh[0] = ald(&c->n_chunk);
                sh[1] = ald(&c->stage);
                sh[2] = ald(&c->seq) != s.seq ? 1u : 0u;
            }
            __syncthreads();
            if (sh[2]) {   // uniform: closed at the first barrier
                if (a.ticks && wg == 0 && threadIdx.x == 0 && s.phase < NPHASE) {
                    unsigned long long *d = a.ticks + 2 * NPHASE + s.phase;
                    d[0] += tk[1] - tk[0];
                    d[NPHASE] += tk[2] - tk[1];
                    d[4 * NPHASE] += phase_work(a, s);
                }
                continue;
            }
            if (sh[1] != 0) break;   // uniform: every workgroup read the same block
        } else {
            if (threadIdx.x == 0) sh[0] = ald(&c->n_chunk);
            __syncthreads();
        }
        pass_chunks<A>(a, s, tid, nthr, sh[0]);
        stage_flush(a, s);
        tk[3] = wall_clock64();
        team_sync(c, a.bar, W, wg, close_pass);
        if (a.ticks && wg == 0 && threadIdx.x == 0 && s.phase < NPHASE) {
            // workgroup 0's view: vertex stage, first barrier wait, chunk
            // stage, second barrier wait; pass work and chunk items
            unsigned long long *d = a.ticks + 2 * NPHASE + s.phase;
            d[0] += tk[1] - tk[0];
            d[NPHASE] += tk[2] - tk[1];
            d[2 * NPHASE] += tk[3] - tk[2];
            d[3 * NPHASE] += wall_clock64() - tk[3];
            d[4 * NPHASE] += phase_work(a, s);
            d[5 * NPHASE] += sh[0];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (W == 1) {
            const uint32_t *src = (const uint32_t *)c;
            uint32_t *dst = (uint32_t *)a.ctl;
            for (uint32_t i = 0; i < sizeof(Ctl) / 4; i++) ast(dst + i, src[i]);
            publish(a, a.ctl);
        } else {
            // the last workgroup out restarts the barrier for the next launch
            // (nobody is waiting any more) and publishes the final block
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            if (atomicAdd(&c->bar_exit, 1u) == W - 1) {
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                ast(&c->bar_count, 0u);
                ast(&c->bar_gen, 0u);
                ast(&c->bar_exit, 0u);
                for (uint32_t k = 0; k < BAR_SUBS; k++) {
                    ast(&a.bar[k * BAR_STRIDE], 0u);
                    ast(&a.bar[k * BAR_STRIDE + 1], 0u);
                }
                publish(a, c);
            }
        }
    }
}

template <class T>
__device__ __forceinline__ T pick(const T (&x)[MAXT], uint32_t t)
{
    static_assert(MAXT == 4, "selection below");
    return t == 0 ? x[0] : t == 1 ? x[1] : t == 2 ? x[2] : x[3];
}

// Workgroup b runs on XCD b % 8 (workgroups are dealt to the XCDs round-
// robin) and serves the tree that owns that XCD slot, as workgroup
// (b / 8) * k + rank of the tree's W = rounds * k.
__global__ __launch_bounds__(1024) void k_team(TeamParams P)
{
    __shared__ EngineArgs sa;
    __shared__ Ctl lds_ctl;
    __shared__ Snap snap;
    __shared__ uint32_t sh[3];
    __shared__ uint32_t stk[STACK_WORDS];
    __shared__ Stage stg;
    const uint32_t slot = blockIdx.x % XCD_SLOTS, round = blockIdx.x / XCD_SLOTS;
    const uint32_t t = (P.slot_tree >> (4 * slot)) & 0xFu;
    if (t >= (uint32_t)MAXT) return;   // a slot of a tree not in this launch
    const uint32_t k = pick(P.k, t), W = P.rounds * k, wg = round * k + ((P.slot_rank >> (4 * slot)) & 0xFu);
    const uint32_t work_max = pick(P.work_max, t), chunk_max = pick(P.chunk_max, t);
    static_assert(sizeof(EngineArgs) % 4 == 0, "word copy below");
    // the tree's arguments are staged in LDS: every pass reads them often
    const uint32_t *src = (const uint32_t *)pick(P.a, t);
    for (uint32_t i = threadIdx.x; i < sizeof(EngineArgs) / 4; i += blockDim.x) ((uint32_t *)&sa)[i] = src[i];
    __syncthreads();
    const bool init = (P.init_mask >> t) & 1u;
    switch (sa.algo) {
    case BFS: team_run<BFS>(sa, W, work_max, chunk_max, wg, init, &lds_ctl, snap, sh, stk, &stg); break;
    case SSSP: team_run<SSSP>(sa, W, work_max, chunk_max, wg, init, &lds_ctl, snap, sh, stk, &stg); break;
    case SSWP: team_run<SSWP>(sa, W, work_max, chunk_max, wg, init, &lds_ctl, snap, sh, stk, &stg); break;
    default: team_run<WCC>(sa, W, work_max, chunk_max, wg, init, &lds_ctl, snap, sh, stk, &stg); break;
    }
}

template <int A>
__global__ __launch_bounds__(GRID_BLOCK) void k_grid_main(EngineArgs a)
{
    __shared__ Snap snap;
    __shared__ uint32_t stk[GRID_STACK_WORDS];
    __shared__ Stage stg;
    stage_init(&stg);
    if (threadIdx.x == 0) {
        take_snap_launch(a.ctl, snap);
        snap.stg = &stg;
    }
    __syncthreads();
    const Snap s = snap;
    uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x, nthr = gridDim.x * blockDim.x;
    if (!grid_needed(a, s) || s.stage != 0) return;
    pass_main<A, GRID_GROUP, GRID_CONT_STACK>(a, s, tid, nthr, stk);
    stage_flush(a, s);
}

template <int A>
__global__ __launch_bounds__(GRID_BLOCK) void k_grid_chunks(EngineArgs a)
{
    __shared__ Snap snap;
    __shared__ Stage stg;
    stage_init(&stg);
    if (threadIdx.x == 0) {
        take_snap_launch(a.ctl, snap);
        snap.stg = &stg;
    }
    __syncthreads();
    const Snap s = snap;
    uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x, nthr = gridDim.x * blockDim.x;
    if (!grid_needed(a, s)) return;
    pass_chunks<A>(a, s, tid, nthr, s.n_chunk);
    stage_flush(a, s);
}

// closes a grid step (the same test as its two pass kernels, on the same
// unchanged fields) and publishes the control block either way
__global__ void k_advance(EngineArgs a)
{
    Snap s;
    take_snap(a.ctl, s);
    if (grid_needed(a, s)) advance(a, a.ctl);
    publish(a, a.ctl);
}

// Team workgroups that can be resident at once on the current device: the
// team barrier needs every workgroup of a launch resident, so wider requests
// are narrowed to fit.
uint32_t team_resident_max()
{
    static uint32_t cached[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 1;
    if (!cached[dev]) {
        hipDeviceProp_t p{};
        int occ = 0;
        if (hipGetDeviceProperties(&p, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_team, 1024, 0) != hipSuccess)
            return 1;
        cached[dev] = (uint32_t)std::max(1, occ) * (uint32_t)std::max(1, p.multiProcessorCount);
    }
    return cached[dev];
}

void launch_engine_team(const TeamLaunch &tl, hipStream_t st)
{
    TeamParams P{};
    uint32_t s = 0;
    for (int t = 0; t < MAXT; t++) {
        for (uint32_t r = 0; r < tl.k[t] && s < XCD_SLOTS; r++, s++) {
            P.slot_tree |= (uint32_t)t << (4 * s);
            P.slot_rank |= r << (4 * s);
        }
    }
    if (s == 0 || tl.rounds == 0) return;
    for (; s < XCD_SLOTS; s++) P.slot_tree |= 0xFu << (4 * s);
    // every workgroup of the launch must be resident at once
    const uint32_t rounds = std::max<uint32_t>(1, std::min(tl.rounds, team_resident_max() / XCD_SLOTS));
    const EngineArgs *any = nullptr;
    for (int t = 0; t < MAXT; t++)
        if (tl.k[t]) any = tl.a[t];
    for (int t = 0; t < MAXT; t++) {
        P.a[t] = tl.k[t] ? tl.a[t] : any;
        P.k[t] = tl.k[t];
        P.work_max[t] = tl.work_max[t];
        P.chunk_max[t] = tl.chunk_max[t];
    }
    P.rounds = rounds;
    P.init_mask = tl.init_mask;
    hipLaunchKernelGGL(k_team, dim3(XCD_SLOTS * rounds), dim3(1024), 0, st, P);
}

template <int A>
static void grid_launch(const EngineArgs &a, const Ctl &h, int steps, hipStream_t st)
{
    uint64_t threads;
    switch (h.phase) {
    case PH_SEED:
    case PH_DEL_ROOTS: threads = h.n_ops; break;
    case PH_BUILD_INIT: threads = a.V; break;
    case PH_DEL_PULL: threads = std::max<uint64_t>((uint64_t)h.n_tree * GRID_GROUP, h.seed ? h.n_ops : 0); break;
    default: threads = (uint64_t)h.n_cur * GRID_GROUP; break;
    }
    if (h.phase == PH_PUSH && (h.flags & F_DENSE)) threads = (uint64_t)a.V * GRID_GROUP;
    uint64_t blocks = (threads + GRID_BLOCK - 1) / GRID_BLOCK;
    if (blocks < 1) blocks = 1;
    static const uint32_t max_blocks = [] {
        const char *e = getenv("RG_GRID_BLOCKS");
        const long v = e ? atol(e) : 0;
        return v > 0 && v <= 4096 ? (uint32_t)v : GRID_MAX_BLOCKS;
    }();
    if (blocks > max_blocks) blocks = max_blocks;
    for (int k = 0; k < steps; k++) {
        // later steps do not know their pass's size: full width
        if (k == 1) blocks = max_blocks;
        hipLaunchKernelGGL(k_grid_main<A>, dim3((uint32_t)blocks), dim3(GRID_BLOCK), 0, st, a);
        hipLaunchKernelGGL(k_grid_chunks<A>, dim3(std::max<uint32_t>(1, max_blocks / 2)), dim3(GRID_BLOCK), 0, st, a);
        hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, st, a);
    }
}

void launch_engine_grid_steps(const EngineArgs &a, const Ctl &h, int steps, hipStream_t st)
{
    switch (a.algo) {
    case BFS: grid_launch<BFS>(a, h, steps, st); break;
    case SSSP: grid_launch<SSSP>(a, h, steps, st); break;
    case SSWP: grid_launch<SSWP>(a, h, steps, st); break;
    default: grid_launch<WCC>(a, h, steps, st); break;
    }
}

// need_update_tree_add / need_update_tree_del (graph.hpp:987-1000,
// 1220-1233) for a batch of edges: an add is unsafe when it strictly improves
// the target label, a del when the target's parent is the source (either
// direction for undirected ops or an undirected algorithm).
template <int A>
__global__ void k_classify(EngineArgs a, uint32_t n, uint8_t *unsafe)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint8_t k = a.ops.kind[i];
    uint32_t s = a.ops.src[i], d = a.ops.dst[i];
    uint32_t w = (A == SSSP || A == SSWP) ? a.ops.w[i] : 0;
    bool both = a.undirected || (k & K_UNDIR);
    bool u;
    if (k & K_INS) {
        u = better<A>(cand<A>(vval(a.vr[s].val), w), vval(a.vr[d].val));
        if (both) u = u || better<A>(cand<A>(vval(a.vr[d].val), w), vval(a.vr[s].val));
    } else {
        u = vpar(a.vr[d].val) == s;
        if (both) u = u || vpar(a.vr[s].val) == d;
    }
    unsafe[i] = u;
}

void launch_classify(const EngineArgs &a, uint32_t n, uint8_t *unsafe, hipStream_t st)
{
    if (!n) return;
    dim3 g((n + 255) / 256), b(256);
    switch (a.algo) {
    case BFS: hipLaunchKernelGGL(k_classify<BFS>, g, b, 0, st, a, n, unsafe); break;
    case SSSP: hipLaunchKernelGGL(k_classify<SSSP>, g, b, 0, st, a, n, unsafe); break;
    case SSWP: hipLaunchKernelGGL(k_classify<SSWP>, g, b, 0, st, a, n, unsafe); break;
    default: hipLaunchKernelGGL(k_classify<WCC>, g, b, 0, st, a, n, unsafe); break;
    }
}

// need_update_tree_add / need_update_tree_del (graph.hpp:987-1000,
// 1220-1233) from the label words of the op's source (ws) and target (wd)
template <int A>
__device__ __forceinline__ bool op_unsafe(uint64_t ws, uint64_t wd, bool ins, bool both, uint32_t s, uint32_t d,
                                          uint32_t w)
{
    if (ins) {
        bool u = better<A>(cand<A>(vval(ws), w), vval(wd));
        if (both) u = u || better<A>(cand<A>(vval(wd), w), vval(ws));
        return u;
    }
    bool u = vpar(wd) == s;
    if (both) u = u || vpar(ws) == d;
    return u;
}

// One thread per edge op of a store batch, before any tree's repair starts:
// the labels are still the fixpoint of the graph before the batch, so an op
// that is safe here changes no label (an insert that improves nothing, a
// delete of no tree edge), and the batch's unsafe ops alone seed the repair
// (a safe insert that a deletion later makes useful is found by the pull or
// push over the store, where it already is).
__global__ void k_split(SplitParams P, uint32_t n)
{
    if (P.dyn_n) n = *P.dyn_n;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool u = false;
    uint32_t s = 0, d = 0, w = 0;
    uint8_t k = 0;
    OpResult r{};
    if (i < n) {
        k = P.in.kind[i];
        r = P.in.res[i];
        s = P.in.src[i];
        d = P.in.dst[i];
        w = P.in.w[i];
        const bool ins = (k & K_INS) != 0;
        const bool cand = (k & K_FORCE) || (ins ? (r.old == 0 && r.live != 0) : ((k & K_DEL) && r.old == 1));
        if (cand) {
            // every tree's label words are loaded before any is tested
            // (most ops are safe in every tree and need them all)
            uint64_t ws[MAXT], wd[MAXT];
#pragma unroll
            for (int t = 0; t < MAXT; t++) {
                ws[t] = wd[t] = 0;
                if (t < P.ntrees) {
                    const bool both = P.undirected[t] || (k & K_UNDIR);
                    wd[t] = P.vr[t][d].val;
                    if (ins || both) ws[t] = P.vr[t][s].val;
                }
            }
#pragma unroll
            for (int t = 0; t < MAXT; t++) {
                if (t >= P.ntrees) continue;
                const bool both = P.undirected[t] || (k & K_UNDIR);
                switch (P.algo[t]) {
                case BFS: u |= op_unsafe<BFS>(ws[t], wd[t], ins, both, s, d, w); break;
                case SSSP: u |= op_unsafe<SSSP>(ws[t], wd[t], ins, both, s, d, w); break;
                case SSWP: u |= op_unsafe<SSWP>(ws[t], wd[t], ins, both, s, d, w); break;
                default: u |= op_unsafe<WCC>(ws[t], wd[t], ins, both, s, d, w); break;
                }
            }
        }
        if (P.unsafe) P.unsafe[i] = u ? 1 : 0;
    }
    const uint32_t pos = u ? wave_reserve(P.count) : 0u;
    if (u) {
        P.out.src[pos] = s;
        P.out.dst[pos] = d;
        P.out.w[pos] = w;
        P.out.kind[pos] = k;
        P.out.res[pos] = r;
    }
}

// the compacted count becomes each tree's n_ops in its pinned control-block
// mirror, from which the run's first launch takes its control block
__global__ void k_split_publish(SplitParams P)
{
    const uint32_t c = __hip_atomic_load(P.count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int t = 0; t < P.ntrees; t++)
        if (P.hctl[t]) P.hctl[t]->n_ops = c;
    __threadfence_system();
}

void launch_split(const SplitParams &p, uint32_t n, hipStream_t st)
{
    if (n) hipLaunchKernelGGL(k_split, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
    hipLaunchKernelGGL(k_split_publish, dim3(1), dim3(1), 0, st, p);
}

// init labels of the reference drivers: BFS/SSSP root 0 else MAXL, SSWP
// root MAXL else 0 (MAXL = 65536 for BFS, 2^27 otherwise); WCC takes the
// host-provided labels (identity or a permutation) and is not filled here.
__global__ void k_fill_init(uint32_t *init, uint32_t V, int algo, uint32_t root)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        uint32_t x;
        if (algo == BFS) x = v == root ? 0u : 65536u;
        else if (algo == SSSP) x = v == root ? 0u : 134217728u;
        else if (algo == SSWP) x = v == root ? 134217728u : 0u;
        else x = v;
        init[v] = x;
    }
}

__global__ void k_clear_marks(VRec *vr, uint32_t V)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) vr[v].inv = 0;
}

void launch_clear_marks(VRec *vr, uint32_t V, hipStream_t st)
{
    if (V) hipLaunchKernelGGL(k_clear_marks, dim3(std::min<uint32_t>((V + 255) / 256, 16384)), dim3(256), 0, st, vr, V);
}

// vr[i].val = words[i] with the marks cleared (labels restored from a snapshot)
__global__ void k_write_labels(VRec *vr, uint32_t n, const uint64_t *words)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        vr[i].val = words[i];
        vr[i].inv = 0;
    }
}

void launch_write_labels(VRec *vr, uint32_t n, const uint64_t *words, hipStream_t st)
{
    if (n) hipLaunchKernelGGL(k_write_labels, dim3(std::min<uint32_t>((n + 255) / 256, 16384)), dim3(256), 0, st, vr, n, words);
}

__global__ void k_read_init(const uint32_t *init, uint32_t n, uint64_t *out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = init[i];
}

void launch_read_init(const uint32_t *init, uint32_t n, uint64_t *out, hipStream_t st)
{
    if (n) hipLaunchKernelGGL(k_read_init, dim3(std::min<uint32_t>((n + 255) / 256, 16384)), dim3(256), 0, st, init, n, out);
}

__global__ void k_read_labels(const VRec *vr, uint32_t n, uint64_t *out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = vr[i].val;
}

void launch_read_labels(const VRec *vr, uint32_t n, uint64_t *out, hipStream_t st)
{
    if (n) hipLaunchKernelGGL(k_read_labels, dim3(std::min<uint32_t>((n + 255) / 256, 16384)), dim3(256), 0, st, vr, n, out);
}

void launch_engine_fill_init(const EngineArgs &a, hipStream_t st)
{
    uint64_t blocks = (a.V + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    if (!blocks) blocks = 1;
    hipLaunchKernelGGL(k_fill_init, dim3((uint32_t)blocks), dim3(256), 0, st, a.init, a.V, a.algo, a.root);
}

}  // namespace rg
