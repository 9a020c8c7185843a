// store.hip -- HBM slab adjacency store: batched edge insert/delete.
//
// Replaces IndexedEdgeStorage::update_edge (reference core/storage.hpp:345-398)
// called from Graph::add_edge / del_edge (core/graph.hpp:203-253).
//
// A batch of edge ops is expanded into list ops (out-list of src and in-list
// of dst, or both endpoint lists of a symmetric graph).  Each list op takes
// a rank on its list with one atomicAdd; the first op of every list reserves
// the list's segment of an op-index buffer and every op drops its index into
// place, so each list has exactly one owner: no locks, no cross-owner races,
// and the block of a list may be re-allocated (doubling, from the slab pool)
// before the owner appends to it.  The owner applies the list's inserts
// first, then its deletes, so a batch behaves like the reference batch driver
// (all adds, then all dels: bfs_inc_batch.cpp:142-178).
// k_apply_all runs both kinds of list side by side in one launch:
//   light lists    one 2-lane group per list; ops one at a time (lookups
//                  scan the list 8 entries per lane per step, or probe its
//                  open-addressing index once it holds >= INDEX_THRESHOLD
//                  entries).
//   heavy lists    one 256-thread workgroup per list for hub segments (>= 8
//                  ops on an indexed list), index builds and rebuilds, and
//                  block copies longer than 256 entries: 256 ops per round,
//                  same-key ops of a round folded into one count update
//                  through an LDS key table.
#include "rg_internal.h"
#include "rg_gen.h"

namespace rg {

__device__ __forceinline__ uint64_t mix64(uint64_t x)
{
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x;
}

__device__ __forceinline__ uint64_t hkey_of(uint32_t nbr, uint32_t w)
{
    return (uint64_t)(nbr & NBR_MASK) | ((uint64_t)w << 32);
}

__device__ __forceinline__ uint32_t next_pow2(uint32_t x)
{
    uint32_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

// ---------------------------------------------------------------------
// expand edge ops into list ops and link them into per-list chains
// ---------------------------------------------------------------------
// Edge op i -> list ops S*i..S*i+S-1, S = the batch's lstride (unused slots
// get v = NILP; an op needing more slots than S flags the batch).  With `rank`,
// each list op takes its rank on its list with one atomicAdd on the header's
// nops (the segment path); without, the caller groups the ops itself.
__device__ void expand_op(const StoreArgs &a, uint32_t i, bool rank, uint32_t n_ops)
{
    uint32_t s = a.ops.src[i], d = a.ops.dst[i], w = a.weighted ? a.ops.w[i] : 0;
    uint8_t kind = a.ops.kind[i];
    a.ops.res[i] = OpResult{0u, 0u};
    const uint32_t S = a.lstride;
    uint32_t base = S * i;
    for (uint32_t k = 0; k < S; k++) a.lops.op[base + k].v = NILP;
    if (kind & K_NOSTORE) return;
    if (s >= a.V || d >= a.V) {
        atomicOr(a.err, 1u);
        return;
    }
    int8_t sign = (kind & K_INS) ? 1 : -1;
    bool undir = kind & K_UNDIR;
    uint32_t nl = 0;
    auto emit = [&](uint32_t v, uint32_t nbr, int8_t delta, uint8_t flags) {
        if (nl >= S) {
            atomicOr(a.err, 2u);   // the host chose too small a stride
            return;
        }
        uint32_t j = base + nl++;
        LOp o{};
        o.v = v;
        o.nbr = nbr;
        o.w = w;
        o.edge = i;
        o.delta = delta;
        o.flags = flags;
        o.slot = NILP;
        int L = (flags & L_IN) ? 1 : 0;
        if (rank) o.rank = atomicAdd(&a.lists[L].h[v].nops, 1u);
        a.lops.op[j] = o;
    };
    // graph.hpp:203-253: a directed op touches out[src] (+ in[dst] unless
    // symmetric); an undirected op is the directed op both ways and returns
    // the old count of the (d->s) direction, minus 1 for a self-loop on add.
    // In a batch, undirected updates keep both orientations at the same
    // count, so the return value is taken from the list of min(s, d): ops on
    // one edge in either orientation then share one list owner and see one
    // serial order.  A single op returns old(d->s) exactly as the reference does
    // (that differs only if directed ops were mixed into the same edge).
    // A self-loop's two same-key updates fold into one +-2 op.
    if (!undir) {
        emit(s, d, sign, L_RET | L_SLOT);
        if (!a.symmetric) emit(d, s, sign, L_IN);
    } else if (s != d) {
        bool ret_sd = n_ops > 1 && s < d;
        emit(s, d, sign, (uint8_t)(L_SLOT | (ret_sd ? L_RET : 0)));
        if (!a.symmetric) emit(d, s, sign, L_IN);
        emit(d, s, sign, (uint8_t)(ret_sd ? 0 : L_RET));
        if (!a.symmetric) emit(s, d, sign, L_IN);
    } else {
        emit(s, s, (int8_t)(2 * sign), L_RET | L_SLOT | L_SELF);
        if (!a.symmetric) emit(s, s, (int8_t)(2 * sign), L_IN | L_SELF);
    }
}

// The batch's edge-op count: captured pipelines take it from device memory.
// (Kernels read it into a local and never write their StoreArgs argument: a
// written argument becomes a per-lane private copy of the whole struct.)
__device__ __forceinline__ uint32_t op_count(const StoreArgs &a)
{
    return a.dyn_n ? *a.dyn_n : a.ops.n;
}

__global__ void k_expand_link(StoreArgs a)
{
    const uint32_t n = op_count(a);
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) expand_op(a, i, true, n);
}

// ---------------------------------------------------------------------
// slab pool
// ---------------------------------------------------------------------
__device__ __forceinline__ int log2_pow2(uint64_t n) { return 63 - __clzll((long long)n); }

// allocate a power-of-two block: reuse a block freed by an earlier batch,
// else bump the pool pointer
__device__ uint64_t pool_alloc(const StoreArgs &a, int h, uint64_t n)
{
    Pool *p = a.pool;
    int c = log2_pow2(n);
    if (c < NCLS) {
        unsigned int avail = p->free_n[h][c];
        if (avail) {
            unsigned int k = atomicAdd(&p->free_take[h][c], 1u);
            if (k < avail) return a.free_off[((uint64_t)h * NCLS + c) * a.fcap + (avail - 1 - k)];
        }
    }
    unsigned long long *top = h ? &p->hash_top : &p->slot_top;
    unsigned long long cap = h ? p->hash_cap : p->slot_cap;
    unsigned long long o = atomicAdd(top, (unsigned long long)n);
    if (o + n > cap) {
        atomicOr(&p->oom, 1u);
        return HEMPTY;
    }
    return o;
}

// return a block; it becomes allocatable after this batch
__device__ void pool_free(const StoreArgs &a, int h, uint64_t off, uint64_t n)
{
    if (off == HEMPTY || n == 0) return;
    int c = log2_pow2(n);
    if (c >= NCLS) return;
    unsigned int k = atomicAdd(&a.pool->pend_n[h][c], 1u);
    if (k < a.fcap) a.pend_off[((uint64_t)h * NCLS + c) * a.fcap + k] = off;
}

// ---------------------------------------------------------------------
// segments: group the batch's list ops by list
// ---------------------------------------------------------------------
// The first op of every list reserves the list's segment and an entry in one
// of two dense owner lists (rank-0 ops): light lists, which k_apply_all
// walks with a 2-lane group each, and heavy lists, which it gives a
// workgroup each.  A list is heavy when its segment is a hub segment, or when
// its inserts (at most its op count) could need an index build or rebuild or
// a block copy longer than LIGHT_COPY entries.  A block sums its segments and
// lists with prefix scans and reserves them with one 64-bit atomic (segment
// entries low, light lists high) plus one for its heavy lists.
constexpr int SEG_BLOCK = 1024;
constexpr uint32_t PAR_MIN = 8;        // hub segments: at least this many ops on an indexed list
constexpr uint32_t LIGHT_COPY = 256;   // longer lists grow in the heavy kernel

__device__ __forceinline__ bool heavy_list(const ListHdr &h, uint32_t n)
{
    const uint32_t need = h.len + n;
    const bool indexed = h.hoff != HEMPTY;
    return (n >= PAR_MIN && (indexed || need >= INDEX_THRESHOLD)) || (!indexed && need >= INDEX_THRESHOLD) ||
           (indexed && 2ull * (h.hcnt + n) > h.hcap) || (need > h.cap && h.len > LIGHT_COPY);
}

// Growth blocks of the light lists are reserved here, per block: the lists
// that may outgrow their block (op count as the bound on inserts) count per
// size class in LDS, one thread per class pops the class's share of the free
// list with one atomic, and the rest comes from one bump of the pool pointer
// for the whole block (a per-list bump is a same-address atomic each).
__device__ void reserve_growth(const StoreArgs &a, uint32_t cls, bool want, uint64_t &out)
{
    __shared__ uint32_t dem[NCLS], take[NCLS], nfree[NCLS], avail[NCLS];
    __shared__ unsigned long long bump_off[NCLS];
    __shared__ uint32_t any;
    if (threadIdx.x < NCLS) dem[threadIdx.x] = 0;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    uint32_t r = 0;
    if (want) {
        r = atomicAdd(&dem[cls], 1u);
        any = 1;
    }
    __syncthreads();
    if (!any) return;
    Pool *p = a.pool;
    if (threadIdx.x < NCLS) {
        const uint32_t c = threadIdx.x, d = dem[c];
        uint32_t t = 0, f = 0, av = 0;
        if (d) {
            av = p->free_n[0][c];
            if (av) {
                t = atomicAdd(&p->free_take[0][c], d);
                f = t < av ? min(d, av - t) : 0;
            }
        }
        take[c] = t;
        nfree[c] = f;
        avail[c] = av;
        bump_off[c] = (unsigned long long)(d - f) << c;   // slots this class needs from the bump pointer
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        for (int c = 0; c < NCLS; c++) {
            const unsigned long long n = bump_off[c];
            bump_off[c] = tot;
            tot += n;
        }
        unsigned long long base = tot ? atomicAdd(&p->slot_top, tot) : 0ull;
        if (tot && base + tot > p->slot_cap) {
            atomicOr(&p->oom, 1u);
            base = HEMPTY;
        }
        for (int c = 0; c < NCLS; c++) bump_off[c] = base == HEMPTY ? HEMPTY : bump_off[c] + base;
    }
    __syncthreads();
    if (want) {
        if (r < nfree[cls])
            out = a.free_off[(uint64_t)cls * a.fcap + (avail[cls] - 1 - (take[cls] + r))];
        else if (bump_off[cls] == HEMPTY)
            out = HEMPTY;
        else
            out = bump_off[cls] + ((uint64_t)(r - nfree[cls]) << cls);
    }
}

__global__ void __launch_bounds__(SEG_BLOCK) k_segment(StoreArgs a)
{
    __shared__ uint32_t wsum[SEG_BLOCK / WAVE][3];
    __shared__ unsigned long long bbase;
    __shared__ uint32_t hbase;
    const uint32_t nlops = a.lstride * op_count(a);
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t cnt = 0, lt = 0, hv = 0, v = NILP, gcls = 0;
    bool grow = false;
    int L = 0;
    if (j < nlops) {
        v = a.lops.op[j].v;
        if (v != NILP && a.lops.op[j].rank == 0) {
            L = (a.lops.op[j].flags & L_IN) ? 1 : 0;
            const ListHdr &h = a.lists[L].h[v];
            cnt = h.nops;
            if (heavy_list(h, cnt)) {
                hv = 1;
            } else {
                lt = 1;
                // a single delete never grows its list
                const bool ins = cnt > 1 || a.lops.op[j].delta > 0;
                if (ins && h.len + cnt > h.cap) {
                    grow = true;
                    gcls = (uint32_t)log2_pow2(max(4u, next_pow2(h.len + cnt)));
                }
            }
        }
    }
    uint64_t noff = HEMPTY;
    reserve_growth(a, gcls, grow, noff);
    if (grow) {
        ListHdr &h = a.lists[L].h[v];
        h.noff = noff;
        h.ncap = noff == HEMPTY ? 0u : 1u << gcls;
    }
    const uint32_t lane = __lane_id(), wid = threadIdx.x / WAVE;
    uint32_t ic = cnt, il = lt, ih = hv;
    for (int d = 1; d < WAVE; d <<= 1) {
        const uint32_t yc = __shfl_up(ic, d), yl = __shfl_up(il, d), yh = __shfl_up(ih, d);
        if (lane >= (uint32_t)d) {
            ic += yc;
            il += yl;
            ih += yh;
        }
    }
    if (lane == WAVE - 1) {
        wsum[wid][0] = ic;
        wsum[wid][1] = il;
        wsum[wid][2] = ih;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tc = 0, tl = 0, th = 0;
        for (int w = 0; w < SEG_BLOCK / WAVE; w++) {
            const uint32_t c = wsum[w][0], l = wsum[w][1], h = wsum[w][2];
            wsum[w][0] = tc;
            wsum[w][1] = tl;
            wsum[w][2] = th;
            tc += c;
            tl += l;
            th += h;
        }
        bbase = (tc | tl) ? atomicAdd(&a.pool->seg_pack, ((unsigned long long)tl << 32) | tc) : 0ull;
        hbase = th ? atomicAdd(&a.pool->n_heavy, th) : 0u;
    }
    __syncthreads();
    if (lt | hv) {
        const uint32_t seg = (uint32_t)bbase + wsum[wid][0] + ic - cnt;
        a.lists[L].h[v].seg = seg;
        // the rank-0 op takes the first place of its segment here (segments
        // of neighbouring ops lie side by side, so these writes coalesce);
        // a light list with this one op never reads its segment
        if (hv || cnt > 1) a.lops.segbuf[seg] = j;
    }
    if (lt) {
        LOp o = a.lops.op[j];
        o.v = v | ((uint32_t)L << 31);
        o.rank = j;
        a.lops.owner[(uint32_t)(bbase >> 32) + wsum[wid][1] + il - 1] = o;
    }
    if (hv) a.lops.heavy[hbase + wsum[wid][2] + ih - 1] = j;
}

__global__ void k_scatter(StoreArgs a)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.lstride * op_count(a)) return;
    const uint32_t v = a.lops.op[j].v;
    // rank-0 ops were placed by k_segment: only lists with several ops of
    // the batch reach the header read below
    if (v == NILP || a.lops.op[j].rank == 0) return;
    const int L = (a.lops.op[j].flags & L_IN) ? 1 : 0;
    a.lops.segbuf[a.lists[L].h[v].seg + a.lops.op[j].rank] = j;
}

// ---------------------------------------------------------------------
// apply: one group owns one list (its rank-0 op): AG lanes for a light
// list, a whole workgroup for a heavy one
// ---------------------------------------------------------------------
struct ListState {
    uint64_t off, hoff;
    uint32_t len, cap, deg, hcap, hcnt;
};

__device__ __forceinline__ ListState load_state(const ListHdr &h)
{
    ListState st;
    st.off = h.off;
    st.len = h.len;
    st.cap = h.cap;
    st.hoff = h.hoff;
    st.hcap = h.hcap;
    st.hcnt = h.hcnt;
    st.deg = h.deg;
    return st;
}

__device__ __forceinline__ void store_state(ListHdr &h, const ListState &st)
{
    h.off = st.off;
    h.len = st.len;
    h.cap = st.cap;
    h.hoff = st.hoff;
    h.hcap = st.hcap;
    h.hcnt = st.hcnt;
    h.deg = st.deg;
}

// A list is owned by a group of G lanes (G = 16 in the apply and reserve
// kernels, so one wave serves four lists).  Group control flow is uniform;
// group ballots are the group's bits of the wave ballot.
struct Grp {
    int lane;   // lane within the group
    int base;   // first wave lane of the group
};

template <int G>
__device__ __forceinline__ uint32_t gballot(bool p, const Grp &g)
{
    uint64_t b = __ballot(p);
    if constexpr (G == 64) return (uint32_t)b;   // only used with G <= 32
    else return (uint32_t)((b >> g.base) & ((1ull << G) - 1));
}

template <int G>
__device__ __forceinline__ uint64_t gbcast(uint64_t x, const Grp &g)
{
    return __shfl((unsigned long long)x, 0, G);
}

// entries per lane kept in flight by the copy and index-build loops (a
// group owns whole hub lists, so these loops are latency-bound otherwise)
constexpr int UNR = 8;

// group-parallel (re)build of a list's hash index at capacity `hcap`
template <int G>
__device__ bool build_index(const StoreArgs &a, ListState &st, uint32_t hcap, const Grp &g)
{
    uint64_t h0 = 0;
    if (g.lane == 0) {
        h0 = pool_alloc(a, 1, hcap);
        if (h0 != HEMPTY && st.hoff != HEMPTY) pool_free(a, 1, st.hoff, st.hcap);
    }
    h0 = gbcast<G>(h0, g);
    if (h0 == HEMPTY) return false;
    for (uint32_t b = g.lane; b < hcap; b += G) a.slots.hb[h0 + b].key = HEMPTY;
    __threadfence_block();
    const uint32_t mask = hcap - 1;
    for (uint32_t i0 = g.lane; i0 < st.len; i0 += G * UNR) {
        uint64_t key[UNR];
        uint32_t bk[UNR], cn[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const uint32_t i = i0 + u * G;
            key[u] = HEMPTY;
            cn[u] = 0;
            if (i < st.len) {
                const uint32_t nb = a.slots.nbr[st.off + i];
                const uint32_t w = a.slots.wt ? a.slots.wt[st.off + i] : 0;
                key[u] = hkey_of(nb, w);
                cn[u] = a.slots.cnt[st.off + i];
            }
            bk[u] = (uint32_t)mix64(key[u]) & mask;
        }
        unsigned long long prev[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++)
            prev[u] = key[u] == HEMPTY ? 0ull
                                       : atomicCAS(&a.slots.hb[h0 + bk[u]].key, (unsigned long long)HEMPTY,
                                                   (unsigned long long)key[u]);
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            if (key[u] == HEMPTY) continue;
            uint32_t b = bk[u], probes = 0;
            while (prev[u] != HEMPTY) {   // taken: linear probing (cannot fill below 50% load)
                b = (b + 1) & mask;
                if (++probes >= hcap) break;
                prev[u] = atomicCAS(&a.slots.hb[h0 + b].key, (unsigned long long)HEMPTY,
                                    (unsigned long long)key[u]);
            }
            if (probes >= hcap) {
                atomicOr(&a.pool->oom, 2u);
            } else {
                a.slots.hb[h0 + b].idx = i0 + u * G;
                a.slots.hb[h0 + b].cnt = cn[u];
            }
        }
    }
    __threadfence_block();
    st.hoff = h0;
    st.hcap = hcap;
    st.hcnt = st.len;
    return true;
}

// move a list into a block of capacity next_pow2(need) (>= 4); the old block
// is returned to the slab free lists
template <int G, int U = UNR>
__device__ void move_block(const StoreArgs &a, ListState &st, uint64_t o, uint32_t ncap, const Grp &g);

template <int G, int U = UNR>
__device__ bool grow_block(const StoreArgs &a, ListState &st, uint32_t need, const Grp &g)
{
    uint32_t ncap = next_pow2(need);
    if (ncap < 4) ncap = 4;
    uint64_t o = 0;
    if (g.lane == 0) o = pool_alloc(a, 0, ncap);
    o = gbcast<G>(o, g);
    if (o == HEMPTY) return false;
    move_block<G, U>(a, st, o, ncap, g);
    return true;
}

// copy a list into block [o, o + ncap) and free its old block
template <int G, int U>
__device__ void move_block(const StoreArgs &a, ListState &st, uint64_t o, uint32_t ncap, const Grp &g)
{
    for (uint32_t i0 = g.lane; i0 < st.len; i0 += G * U) {
        uint32_t nb[U], cn[U], wt[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint32_t i = i0 + u * G;
            if (i < st.len) {
                nb[u] = a.slots.nbr[st.off + i];
                cn[u] = a.slots.cnt[st.off + i];
                wt[u] = a.slots.wt ? a.slots.wt[st.off + i] : 0;
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint32_t i = i0 + u * G;
            if (i < st.len) {
                a.slots.nbr[o + i] = nb[u];
                a.slots.cnt[o + i] = cn[u];
                if (a.slots.wt) a.slots.wt[o + i] = wt[u];
            }
        }
    }
    __threadfence_block();
    if (g.lane == 0 && st.cap) pool_free(a, 0, st.off, st.cap);
    st.off = o;
    st.cap = ncap;
}

// find key in the list: local index or NILP.  Hashed lists also return in
// *bucket the bucket that holds the key, or the empty bucket where it would
// go (NILP: index full), and on a hit the count copy of that bucket in
// *hcnt (the probe's one 16-byte load per lane brings key, index and count).
template <int G>
__device__ uint32_t list_find(const StoreArgs &a, const ListState &st, uint32_t key_nbr, uint32_t key_w,
                              const Grp &g, uint32_t *bucket, uint32_t *hcnt)
{
    if (st.hoff != HEMPTY) {
        uint64_t key = hkey_of(key_nbr, key_w);
        uint32_t mask = st.hcap - 1;
        uint32_t h = (uint32_t)mix64(key) & mask;
        for (uint32_t probe = 0; probe < st.hcap; probe += G) {
            uint32_t b = (h + probe + g.lane) & mask;
            const HBucket e = a.slots.hb[st.hoff + b];
            uint32_t hit = gballot<G>(e.key == key, g);
            uint32_t emp = gballot<G>(e.key == HEMPTY, g);
            uint32_t any = hit | emp;
            if (any) {
                int first = __ffs(any) - 1;
                *bucket = (h + probe + first) & mask;
                const int src = g.base + first;
                const uint32_t idx = (uint32_t)__shfl((int)e.idx, src);
                const uint32_t c = (uint32_t)__shfl((int)e.cnt, src);
                if ((hit >> first) & 1u) {
                    *hcnt = c;
                    return idx;
                }
                return NILP;
            }
        }
        *bucket = NILP;
        return NILP;
    }
    // Linear scan, 8 entries per lane per step (two 16-byte loads in flight):
    // blocks hold a power of two >= 4 entries at a multiple-of-4 offset, so
    // a quad that starts below len lies inside the block.  Keys are unique in
    // a list, so any match is the match.
    const uint32_t *nb = a.slots.nbr + st.off;
    for (uint32_t b = 0; b < st.len; b += 8 * G) {
        uint32_t found = NILP;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const uint32_t i0 = b + (uint32_t)h * 4 * G + 4 * (uint32_t)g.lane;
            uint4 q = make_uint4(NILP, NILP, NILP, NILP);
            if (i0 < st.len) q = *reinterpret_cast<const uint4 *>(nb + i0);
            const uint32_t e[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const uint32_t i = i0 + (uint32_t)c;
                if (found == NILP && i < st.len && (e[c] & NBR_MASK) == key_nbr &&
                    (!a.slots.wt || a.slots.wt[st.off + i] == key_w))
                    found = i;
            }
        }
        const uint32_t hit = gballot<G>(found != NILP, g);
        if (hit) return (uint32_t)__shfl((int)found, g.base + __ffs(hit) - 1);
    }
    return NILP;
}

// lanes per list in the light apply: 2 measured best at 300k-op batches
// (0.50 ms store batch against 0.54 with 4 and 0.59 with 8; 1 was best at
// million-op batches by 3%)
constexpr int AG = 2;

template <int G>
__device__ __forceinline__ int32_t gsum(int32_t x)
{
    for (int m = G / 2; m > 0; m >>= 1) x += __shfl_xor(x, m, G);
    return x;
}

// One op at a time, in segment order.  The caller guarantees the list needs
// no index build or rebuild (a new key of an indexed list finds an empty
// bucket at load <= 1/2) and has room for every insert.
__device__ void apply_serial(const StoreArgs &a, ListState &st, const uint32_t *ops, uint32_t n, const Grp &g)
{
    for (int pass = 0; pass < 2; pass++) {
        for (uint32_t i = 0; i < n; i++) {
            const uint32_t k = ops[i];
            int8_t delta = a.lops.op[k].delta;
            if ((pass == 0) != (delta > 0)) continue;
            uint32_t key_nbr = a.lops.op[k].nbr;
            uint32_t key_w = a.slots.wt ? a.lops.op[k].w : 0;
            const bool hashed = st.hoff != HEMPTY;
            uint32_t bucket = NILP, hc = 0;
            uint32_t idx = list_find<AG>(a, st, key_nbr, key_w, g, &bucket, &hc);
            uint32_t c = 0;
            if (idx == NILP && delta > 0) {
                // append a new entry with count 0 (storage.hpp:371-374, 389-392)
                idx = st.len;
                if (g.lane == 0) {
                    a.slots.nbr[st.off + idx] = key_nbr | DEAD;
                    a.slots.cnt[st.off + idx] = 0;
                    if (a.slots.wt) a.slots.wt[st.off + idx] = key_w;
                    if (hashed && bucket != NILP)
                        a.slots.hb[st.hoff + bucket] = HBucket{hkey_of(key_nbr, key_w), idx, 0u};
                }
                st.len++;
                if (hashed) st.hcnt++;
                __threadfence_block();
            } else if (idx != NILP) {
                c = hashed ? hc : a.slots.cnt[st.off + idx];
            }
            uint32_t nc, ret;
            int32_t dchange;
            switch (delta) {
            case 1: nc = c + 1; ret = c; dchange = 1; break;
            case -1: nc = c ? c - 1 : 0; ret = c; dchange = c ? -1 : 0; break;
            case 2: nc = c + 2; ret = c; dchange = 2; break;
            default: /* -2: undirected self-loop delete, two -1 steps */
                ret = c ? c - 1 : 0;
                nc = c >= 2 ? c - 2 : 0;
                dchange = -(int32_t)(c >= 2 ? 2 : c);
                break;
            }
            if (idx != NILP && g.lane == 0) {
                a.slots.cnt[st.off + idx] = nc;
                if (hashed && bucket != NILP) a.slots.hb[st.hoff + bucket].cnt = nc;
                if ((c == 0) != (nc == 0))
                    a.slots.nbr[st.off + idx] = key_nbr | (nc ? 0u : DEAD);
            }
            st.deg = (uint32_t)((int32_t)st.deg + dchange);
            if (g.lane == 0) {
                if (a.lops.op[k].flags & L_RET) a.ops.res[a.lops.op[k].edge].old = ret;
                a.lops.op[k].slot = idx;
            }
            __threadfence_block();
        }
    }
}

// A light list with exactly one op (the common case): the op's fields come
// from the caller, the return value and the liveness follow from the one
// count update.
// Room for `need` entries in a light list: move into the block k_segment
// reserved for it (or return that block unused).
__device__ __forceinline__ bool light_room(const StoreArgs &a, ListState &st, uint32_t need, uint32_t rcap,
                                           uint64_t roff, const Grp &g)
{
    if (need > st.cap) {
        if (rcap >= need) {
            move_block<AG, 1>(a, st, roff, rcap, g);
            return true;
        }
        return grow_block<AG, 1>(a, st, need, g);   // no reservation (pool exhausted)
    }
    if (rcap && g.lane == 0) pool_free(a, 0, roff, rcap);
    return true;
}

__device__ __forceinline__ void apply_single(const StoreArgs &a, ListState &st, uint32_t k, int8_t delta,
                                             uint8_t flags, uint32_t key_nbr, uint32_t key_w, uint32_t edge,
                                             uint32_t rcap, uint64_t roff, const Grp &g)
{
    if (!light_room(a, st, st.len + (delta > 0 ? 1u : 0u), rcap, roff, g)) return;
    const bool hashed = st.hoff != HEMPTY;
    uint32_t bucket = NILP, c = 0, hc = 0;
    bool appended = false;
    uint32_t idx = list_find<AG>(a, st, key_nbr, key_w, g, &bucket, &hc);
    if (idx == NILP) {
        if (delta > 0) {
            // append a new entry (storage.hpp:371-374, 389-392)
            idx = st.len;
            appended = true;
            if (g.lane == 0) {
                a.slots.nbr[st.off + idx] = key_nbr | DEAD;
                if (a.slots.wt) a.slots.wt[st.off + idx] = key_w;
            }
            st.len++;
            if (hashed) st.hcnt++;
        }
    } else {
        c = hashed ? hc : a.slots.cnt[st.off + idx];
    }
    uint32_t nc, ret;
    int32_t dchange;
    switch (delta) {
    case 1: nc = c + 1; ret = c; dchange = 1; break;
    case -1: nc = c ? c - 1 : 0; ret = c; dchange = c ? -1 : 0; break;
    case 2: nc = c + 2; ret = c; dchange = 2; break;
    default: /* -2: undirected self-loop delete, two -1 steps */
        ret = c ? c - 1 : 0;
        nc = c >= 2 ? c - 2 : 0;
        dchange = -(int32_t)(c >= 2 ? 2 : c);
        break;
    }
    st.deg = (uint32_t)((int32_t)st.deg + dchange);
    if (g.lane == 0) {
        if (idx != NILP) {
            a.slots.cnt[st.off + idx] = nc;
            if (hashed && bucket != NILP) {
                if (appended) a.slots.hb[st.hoff + bucket] = HBucket{hkey_of(key_nbr, key_w), idx, nc};
                else a.slots.hb[st.hoff + bucket].cnt = nc;
            }
            if ((c == 0) != (nc == 0)) a.slots.nbr[st.off + idx] = key_nbr | (nc ? 0u : DEAD);
        }
        const uint32_t live = (idx != NILP && nc > 0) ? 1u : 0u;
        if ((flags & (L_RET | L_SLOT)) == (L_RET | L_SLOT)) {
            a.ops.res[edge] = OpResult{ret, live};
        } else {
            if (flags & L_RET) a.ops.res[edge].old = ret;
            if (flags & L_SLOT) a.ops.res[edge].live = live;
        }
        // (no slot write-back: only multi-op lists read it, in write_liveness)
    }
}

// liveness of the edge ops whose (src -> dst) entry is on this list, from the
// count after the whole batch
__device__ __forceinline__ void write_liveness(const StoreArgs &a, const ListState &st, const uint32_t *ops,
                                               uint32_t n, uint32_t t0, uint32_t nt)
{
    for (uint32_t i = t0; i < n; i += nt) {
        const uint32_t k = ops[i];
        if (!(a.lops.op[k].flags & L_SLOT)) continue;
        const uint32_t idx = a.lops.op[k].slot;
        a.ops.res[a.lops.op[k].edge].live = (idx != NILP && a.slots.cnt[st.off + idx] > 0) ? 1u : 0u;
    }
}

// Light lists: one AG-lane group per list (its rank-0 op, from the light
// owner list); k_segment already sent every list that could need an index
// build, a rebuild, a long copy or hub rounds to the heavy list.  `blk` and
// `nblk` are this block's index among the light blocks and their count.
__device__ void light_lists_apply(const StoreArgs &a, uint32_t blk, uint32_t nblk)
{
    const uint32_t t = blk * blockDim.x + threadIdx.x;
    Grp g;
    g.lane = threadIdx.x % AG;
    g.base = (threadIdx.x % WAVE) - g.lane;
    const uint32_t gid = t / AG;
    const uint32_t ng = (nblk * blockDim.x) / AG;
    const uint32_t nown = (uint32_t)(a.pool->seg_pack >> 32);
    for (uint32_t oi = gid; oi < nown; oi += ng) {
        // the owner entry carries the rank-0 op's record (most lists have
        // this one op only), so only the header line is a random access
        const LOp o = a.lops.owner[oi];
        const uint32_t j = o.rank;
        const uint32_t v = o.v & 0x7FFFFFFFu;
        const int L = (int)(o.v >> 31);
        const uint8_t fl = o.flags;
        const int8_t dj = o.delta;
        const uint32_t key_nbr = o.nbr;
        const uint32_t key_w = a.slots.wt ? o.w : 0;
        const uint32_t ej = o.edge;
        ListHdr &hd = a.lists[L].h[v];
        const uint32_t n = hd.nops;
        const uint32_t rcap = hd.ncap;
        const uint64_t roff = hd.noff;
        ListState st = load_state(hd);
        if (n == 1) {
            apply_single(a, st, j, dj, fl, key_nbr, key_w, ej, rcap, roff, g);
            if (g.lane == 0) {
                store_state(hd, st);
                hd.nops = 0;
                hd.ncap = 0;
            }
            continue;
        }
        const uint32_t *ops = a.lops.segbuf + hd.seg;

        int32_t n_ins = 0;
        for (uint32_t i = g.lane; i < n; i += AG) n_ins += a.lops.op[ops[i]].delta > 0 ? 1 : 0;
        const uint32_t need = st.len + (uint32_t)gsum<AG>(n_ins);
        if (light_room(a, st, need, rcap, roff, g)) {
            apply_serial(a, st, ops, n, g);
            __threadfence_block();
            write_liveness(a, st, ops, n, g.lane, AG);
        }
        if (g.lane == 0) {
            store_state(hd, st);
            hd.nops = 0;
            hd.ncap = 0;
        }
    }
}

// ---------------------------------------------------------------------
// heavy lists: one 256-thread workgroup per list
// ---------------------------------------------------------------------
constexpr int HB = 256;           // threads per heavy list
constexpr uint32_t HT = 2 * HB;   // LDS key table of one round (load <= 1/2)
// entries per thread in flight in copies and index builds: 4, not 8, keeps
// k_apply_all at 74 VGPRs (6 waves per SIMD instead of 4), which its light
// lists' dependent loads need more (a 512k-update batch's apply 386 -> 314 us)
constexpr int HUNR = 4;

struct HeavyShared {
    unsigned long long key[HT];
    uint32_t cnt[HT];
    uint32_t idx[HT];
    uint32_t c0[HT];
    unsigned long long bcast;
    uint32_t n_app;
    int32_t sum;
};

__device__ bool wg_grow(const StoreArgs &a, ListState &st, uint32_t need, HeavyShared &sh)
{
    uint32_t ncap = next_pow2(need);
    if (ncap < 4) ncap = 4;
    if (threadIdx.x == 0) sh.bcast = pool_alloc(a, 0, ncap);
    __syncthreads();
    const uint64_t o = sh.bcast;
    __syncthreads();
    if (o == HEMPTY) return false;
    for (uint32_t i0 = threadIdx.x; i0 < st.len; i0 += HB * HUNR) {
        uint32_t nb[HUNR], cn[HUNR], wt[HUNR];
#pragma unroll
        for (int u = 0; u < HUNR; u++) {
            const uint32_t i = i0 + u * HB;
            if (i < st.len) {
                nb[u] = a.slots.nbr[st.off + i];
                cn[u] = a.slots.cnt[st.off + i];
                wt[u] = a.slots.wt ? a.slots.wt[st.off + i] : 0;
            }
        }
#pragma unroll
        for (int u = 0; u < HUNR; u++) {
            const uint32_t i = i0 + u * HB;
            if (i < st.len) {
                a.slots.nbr[o + i] = nb[u];
                a.slots.cnt[o + i] = cn[u];
                if (a.slots.wt) a.slots.wt[o + i] = wt[u];
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && st.cap) pool_free(a, 0, st.off, st.cap);
    st.off = o;
    st.cap = ncap;
    return true;
}

__device__ bool wg_build_index(const StoreArgs &a, ListState &st, uint32_t hcap, HeavyShared &sh)
{
    if (threadIdx.x == 0) {
        const uint64_t h0 = pool_alloc(a, 1, hcap);
        if (h0 != HEMPTY && st.hoff != HEMPTY) pool_free(a, 1, st.hoff, st.hcap);
        sh.bcast = h0;
    }
    __syncthreads();
    const uint64_t h0 = sh.bcast;
    __syncthreads();
    if (h0 == HEMPTY) return false;
    for (uint32_t b = threadIdx.x; b < hcap; b += HB) a.slots.hb[h0 + b].key = HEMPTY;
    __syncthreads();
    const uint32_t mask = hcap - 1;
    for (uint32_t i0 = threadIdx.x; i0 < st.len; i0 += HB * HUNR) {
        uint64_t key[HUNR];
        uint32_t bk[HUNR], cn[HUNR];
        unsigned long long prev[HUNR];
#pragma unroll
        for (int u = 0; u < HUNR; u++) {
            const uint32_t i = i0 + u * HB;
            key[u] = HEMPTY;
            cn[u] = 0;
            if (i < st.len) {
                key[u] = hkey_of(a.slots.nbr[st.off + i], a.slots.wt ? a.slots.wt[st.off + i] : 0);
                cn[u] = a.slots.cnt[st.off + i];
            }
            bk[u] = (uint32_t)mix64(key[u]) & mask;
        }
#pragma unroll
        for (int u = 0; u < HUNR; u++)
            prev[u] = key[u] == HEMPTY ? 0ull
                                       : atomicCAS(&a.slots.hb[h0 + bk[u]].key, (unsigned long long)HEMPTY,
                                                   (unsigned long long)key[u]);
#pragma unroll
        for (int u = 0; u < HUNR; u++) {
            if (key[u] == HEMPTY) continue;
            uint32_t b = bk[u], probes = 0;
            while (prev[u] != HEMPTY) {   // linear probing (cannot fill below 50% load)
                b = (b + 1) & mask;
                if (++probes >= hcap) break;
                prev[u] = atomicCAS(&a.slots.hb[h0 + b].key, (unsigned long long)HEMPTY,
                                    (unsigned long long)key[u]);
            }
            if (probes >= hcap) {
                atomicOr(&a.pool->oom, 2u);
            } else {
                a.slots.hb[h0 + b].idx = i0 + u * HB;
                a.slots.hb[h0 + b].cnt = cn[u];
            }
        }
    }
    __syncthreads();
    st.hoff = h0;
    st.hcap = hcap;
    st.hcnt = st.len;
    return true;
}

// Apply an indexed list's segment HB ops per round, one op per thread.  Ops
// of a round with the same key meet in an LDS table: each adds its |delta|
// to the key's total and gets the sum of the ops before it (`pre`), in the
// order the LDS atomics took them.  The first op of each key (pre == 0)
// looks the key up, appends a new entry if needed, and applies the total to
// the count; every op then reports the count it would have seen had the
// key's ops run one by one in that order (clamped at 0 for deletes, as
// storage.hpp:378/393 do), so the returns of a key are the reference's
// multiset.
__device__ void wg_apply_rounds(const StoreArgs &a, ListState &st, const uint32_t *ops, uint32_t n, HeavyShared &sh)
{
    const uint32_t tid = threadIdx.x;
    const uint32_t mask = st.hcap - 1;
    for (int pass = 0; pass < 2; pass++) {
        for (uint32_t r0 = 0; r0 < n; r0 += HB) {
            for (uint32_t b = tid; b < HT; b += HB) {
                sh.key[b] = HEMPTY;
                sh.cnt[b] = 0;
            }
            if (tid == 0) {
                sh.n_app = 0;
                sh.sum = 0;
            }
            __syncthreads();
            const uint32_t i = r0 + tid;
            const uint32_t k = i < n ? ops[i] : 0;
            const int8_t delta = i < n ? a.lops.op[k].delta : (int8_t)0;
            const bool act = i < n && ((pass == 0) == (delta > 0));
            uint32_t key_nbr = 0, key_w = 0, ls = 0, pre = 0;
            uint64_t key = HEMPTY;
            if (act) {
                key_nbr = a.lops.op[k].nbr;
                key_w = a.slots.wt ? a.lops.op[k].w : 0;
                key = hkey_of(key_nbr, key_w);
                ls = (uint32_t)mix64(key) & (HT - 1);
                for (uint32_t p = 0; p < HT; p++) {   // at most HB keys in HT buckets: always ends
                    const unsigned long long prev =
                        atomicCAS(&sh.key[ls], (unsigned long long)HEMPTY, (unsigned long long)key);
                    if (prev == HEMPTY || prev == key) break;
                    ls = (ls + 1) & (HT - 1);
                }
                pre = atomicAdd(&sh.cnt[ls], (uint32_t)(delta < 0 ? -delta : delta));
            }
            __syncthreads();
            if (act && pre == 0) {
                uint32_t idx = NILP, bucket = NILP, hb_at = NILP, c = 0;
                uint32_t b = (uint32_t)mix64(key) & mask;
                for (uint32_t p = 0; p < st.hcap; p++, b = (b + 1) & mask) {
                    const HBucket e = a.slots.hb[st.hoff + b];
                    if (e.key == key) {
                        idx = e.idx;
                        c = e.cnt;
                        hb_at = b;
                        break;
                    }
                    if (e.key == HEMPTY) {
                        bucket = b;
                        break;
                    }
                }
                if (idx == NILP && pass == 0 && bucket != NILP) {
                    // a key new to the list: append an entry with count 0
                    idx = st.len + atomicAdd(&sh.n_app, 1u);
                    a.slots.nbr[st.off + idx] = key_nbr | DEAD;
                    a.slots.cnt[st.off + idx] = 0;
                    if (a.slots.wt) a.slots.wt[st.off + idx] = key_w;
                    uint32_t bb = bucket;
                    for (uint32_t p = 0; p < st.hcap; p++, bb = (bb + 1) & mask) {
                        if (atomicCAS(&a.slots.hb[st.hoff + bb].key, (unsigned long long)HEMPTY,
                                      (unsigned long long)key) == HEMPTY) {
                            a.slots.hb[st.hoff + bb].idx = idx;
                            hb_at = bb;
                            break;
                        }
                    }
                } else if (idx == NILP && pass == 0) {
                    atomicOr(&a.pool->oom, 2u);   // index full (cannot happen at load <= 1/2)
                }
                if (idx != NILP) {
                    const uint32_t tot = sh.cnt[ls];
                    const uint32_t cf = pass == 0 ? c + tot : (c > tot ? c - tot : 0);
                    a.slots.cnt[st.off + idx] = cf;
                    if (hb_at != NILP) a.slots.hb[st.hoff + hb_at].cnt = cf;
                    if ((c == 0) != (cf == 0)) a.slots.nbr[st.off + idx] = key_nbr | (cf ? 0u : DEAD);
                    atomicAdd(&sh.sum, (int32_t)cf - (int32_t)c);
                }
                sh.idx[ls] = idx;
                sh.c0[ls] = c;
            }
            __syncthreads();
            if (act) {
                const uint32_t idx = sh.idx[ls], c = sh.c0[ls];
                const uint32_t cb = pass == 0 ? c + pre : (c > pre ? c - pre : 0);
                const uint32_t ret = delta == -2 ? (cb ? cb - 1 : 0) : cb;
                if (a.lops.op[k].flags & L_RET) a.ops.res[a.lops.op[k].edge].old = ret;
                a.lops.op[k].slot = idx;
            }
            st.len += sh.n_app;
            st.hcnt += sh.n_app;
            st.deg = (uint32_t)((int32_t)st.deg + sh.sum);
            __syncthreads();
        }
    }
}

// One heavy list with the whole HB-thread workgroup: room for its inserts,
// index build or rebuild, then rounds (or, for a short unindexed list that
// only needed a long copy, the serial path on the first AG lanes); liveness
// and the header last.
__device__ void heavy_list_apply(const StoreArgs &a, ListHdr &hd, const uint32_t *ops, uint32_t n, HeavyShared &sh)
{
    ListState st = load_state(hd);
    if (threadIdx.x == 0) sh.sum = 0;
    __syncthreads();
    int32_t my = 0;
    for (uint32_t i = threadIdx.x; i < n; i += HB) my += a.lops.op[ops[i]].delta > 0 ? 1 : 0;
    if (my) atomicAdd(&sh.sum, my);
    __syncthreads();
    const uint32_t n_ins = (uint32_t)sh.sum;
    __syncthreads();
    const uint32_t need = st.len + n_ins;
    bool ok = true;
    if (need > st.cap) ok = wg_grow(a, st, need, sh);
    if (ok) {
        if (st.hoff != HEMPTY || need >= INDEX_THRESHOLD) {
            // indexed: room for every new key at load <= 1/2
            uint32_t hc = next_pow2(2 * need);
            if (hc < INDEX_MIN_CAP) hc = INDEX_MIN_CAP;
            if (st.hoff == HEMPTY || 2ull * (st.hcnt + n_ins) > st.hcap) ok = wg_build_index(a, st, hc, sh);
            if (ok) wg_apply_rounds(a, st, ops, n, sh);
        } else if (threadIdx.x < AG) {
            // a grown short list: the first AG lanes apply its few ops
            Grp g;
            g.lane = threadIdx.x;
            g.base = 0;
            apply_serial(a, st, ops, n, g);
        }
    }
    // every thread takes the list state of thread 0 (the serial path
    // advanced only the first group's copy)
    __syncthreads();
    if (threadIdx.x == 0) {
        store_state(hd, st);
        hd.nops = 0;
    }
    __syncthreads();
    if (ok) {
        st.off = hd.off;
        write_liveness(a, st, ops, n, threadIdx.x, HB);
    }
    __syncthreads();
}

// The apply step of a batch in one launch: blocks [0, hblocks) take the
// heavy lists, one list per block at a time; the rest walk the light owner
// list.  The two sets of lists are disjoint, light lists already hold their
// growth blocks, and freed blocks only go to the pending lists, so both kinds
// run side by side (the heavy blocks come first and start first).
static_assert(HB == 256, "light and heavy blocks share the launch's block size");
__global__ void __launch_bounds__(HB) k_apply_all(StoreArgs a, uint32_t hblocks)
{
    __shared__ HeavyShared sh;
    if (blockIdx.x >= hblocks) {
        light_lists_apply(a, blockIdx.x - hblocks, gridDim.x - hblocks);
        return;
    }
    const uint32_t nh = a.pool->n_heavy;
    for (uint32_t hi = blockIdx.x; hi < nh; hi += hblocks) {
        const uint32_t j = a.lops.heavy[hi];   // the list's rank-0 op
        const uint32_t v = a.lops.op[j].v;
        const int L = (a.lops.op[j].flags & L_IN) ? 1 : 0;
        ListHdr &hd = a.lists[L].h[v];
        heavy_list_apply(a, hd, a.lops.segbuf + hd.seg, hd.nops, sh);
    }
}

constexpr uint32_t SMALL_OPS = STORE_SMALL_OPS;
__global__ void __launch_bounds__(HB) k_store_small(StoreArgs a);

bool launch_store_batch(const StoreArgs &a, hipStream_t s, hipEvent_t ev_a, hipEvent_t ev_b)
{
    if (a.ops.n == 0) return false;
    const uint32_t n = a.ops.n;
    const uint32_t nl = a.lstride * n;
    if (n <= SMALL_OPS) {
        // one launch for the whole batch, pool recycling included
        if (ev_a) (void)hipEventRecord(ev_a, s);
        hipLaunchKernelGGL(k_store_small, dim3(1), dim3(HB), 0, s, a);
        if (ev_b) (void)hipEventRecord(ev_b, s);
        return true;
    }
    hipLaunchKernelGGL(k_expand_link, dim3((n + 255) / 256), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_segment, dim3((nl + SEG_BLOCK - 1) / SEG_BLOCK), dim3(SEG_BLOCK), 0, s, a);
    hipLaunchKernelGGL(k_scatter, dim3((nl + 255) / 256), dim3(256), 0, s, a);
    uint64_t groups = (uint64_t)nl;
    uint64_t blocks = (groups * AG + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    // heavy lists: at most one per list op; blocks beyond the queue exit at once
    const uint32_t hblocks = nl < 2048 ? nl : 2048;
    if (ev_a) (void)hipEventRecord(ev_a, s);
    hipLaunchKernelGGL(k_apply_all, dim3(hblocks + (uint32_t)blocks), dim3(HB), 0, s, a, hblocks);
    if (ev_b) (void)hipEventRecord(ev_b, s);
    return false;
}

void launch_store_batch_dyn(const StoreArgs &a, uint32_t n_max, hipStream_t s)
{
    const uint32_t nl = a.lstride * n_max;
    hipLaunchKernelGGL(k_expand_link, dim3((n_max + 255) / 256), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_segment, dim3((nl + SEG_BLOCK - 1) / SEG_BLOCK), dim3(SEG_BLOCK), 0, s, a);
    hipLaunchKernelGGL(k_scatter, dim3((nl + 255) / 256), dim3(256), 0, s, a);
    uint64_t blocks = ((uint64_t)nl * AG + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    const uint32_t hblocks = nl < 2048 ? nl : 2048;
    hipLaunchKernelGGL(k_apply_all, dim3(hblocks + (uint32_t)blocks), dim3(HB), 0, s, a, hblocks);
}

// ---------------------------------------------------------------------
// bulk reservation: count the entries a whole load will append per list,
// then size every block (and hub index) once, so the load batches never
// re-allocate.  The count is an upper bound (duplicates count each time).
// ---------------------------------------------------------------------
__global__ void k_count(StoreArgs a)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.ops.n) return;
    uint8_t kind = a.ops.kind[i];
    if (!(kind & K_INS)) return;
    uint32_t s = a.ops.src[i], d = a.ops.dst[i];
    if (s >= a.V || d >= a.V) return;
    atomicAdd(&a.cnt_tmp[0][s], 1u);
    if (!a.symmetric) atomicAdd(&a.cnt_tmp[1][d], 1u);
    if (kind & K_UNDIR) {
        atomicAdd(&a.cnt_tmp[0][d], 1u);
        if (!a.symmetric) atomicAdd(&a.cnt_tmp[1][s], 1u);
    }
}

__global__ void k_reserve(StoreArgs a)
{
    Grp g;
    g.lane = threadIdx.x % AG;
    g.base = (threadIdx.x % WAVE) - g.lane;
    const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) / AG;
    const uint32_t nw = (gridDim.x * blockDim.x) / AG;
    const int lane = g.lane;
    const int nlists = a.symmetric ? 1 : 2;
    for (uint32_t v = wid; v < a.V; v += nw) {
        for (int L = 0; L < nlists; L++) {
            uint32_t add = a.cnt_tmp[L][v];
            if (!add) continue;
            ListHdr &hd = a.lists[L].h[v];
            ListState st = load_state(hd);
            uint32_t need = st.len + add;
            bool ok = true;
            if (need > st.cap) ok = grow_block<AG>(a, st, need, g);
            if (ok && need >= INDEX_THRESHOLD && (st.hoff == HEMPTY || st.hcap < 2 * need)) {
                uint32_t hc = next_pow2(2 * need);
                if (hc < INDEX_MIN_CAP) hc = INDEX_MIN_CAP;
                ok = build_index<AG>(a, st, hc, g);
            }
            if (lane == 0) {
                store_state(hd, st);
                a.cnt_tmp[L][v] = 0;
            }
        }
    }
}

void launch_store_count(const StoreArgs &a, hipStream_t s)
{
    if (a.ops.n == 0) return;
    hipLaunchKernelGGL(k_count, dim3((a.ops.n + 255) / 256), dim3(256), 0, s, a);
}

void launch_store_reserve(const StoreArgs &a, hipStream_t s)
{
    uint64_t blocks = ((uint64_t)a.V * AG + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_reserve, dim3((uint32_t)blocks), dim3(256), 0, s, a);
}

__global__ void k_list_init(List l, uint32_t V)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    ListHdr h{};
    h.hoff = HEMPTY;
    l.h[v] = h;
}

void launch_list_init(List l, uint32_t V, hipStream_t s)
{
    if (V) hipLaunchKernelGGL(k_list_init, dim3((V + 255) / 256), dim3(256), 0, s, l, V);
}

__global__ void k_list_degrees(List l, uint32_t V, uint32_t *out)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) out[v] = l.h[v].deg;
}

void launch_list_degrees(List l, uint32_t V, uint32_t *out, hipStream_t s)
{
    if (V) hipLaunchKernelGGL(k_list_degrees, dim3((V + 255) / 256), dim3(256), 0, s, l, V, out);
}

// The live multiset of vertices [v0, v1)'s lists as edge records, `deg`
// copies of each (v, nbr, w) entry with its count (snapshots): vertex v's
// records start at out[off[v - v0] - off[0]].  One thread per vertex; a
// list writes at most its header's live degree (the sum of its counts).
__global__ void k_list_dump(List l, Slots sl, uint32_t v0, uint32_t v1, const uint64_t *off, rg_edge_rec *out)
{
    const uint32_t v = v0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= v1) return;
    const ListHdr h = l.h[v];
    uint64_t o = off[v - v0] - off[0];
    const uint64_t end = o + h.deg;
    for (uint32_t i = 0; i < h.len && o < end; i++) {
        const uint64_t slot = h.off + i;
        uint32_t c = sl.cnt[slot];
        const uint32_t nb = sl.nbr[slot] & NBR_MASK;
        const uint64_t w = sl.wt ? sl.wt[slot] : 0;
        for (; c > 0 && o < end; c--, o++) {
            out[o].src = v;
            out[o].dst = nb;
            out[o].data = w;
        }
    }
}

void launch_list_dump(List l, Slots sl, uint32_t v0, uint32_t v1, const uint64_t *off, rg_edge_rec *out,
                      hipStream_t s)
{
    if (v1 > v0)
        hipLaunchKernelGGL(k_list_dump, dim3((v1 - v0 + 255) / 256), dim3(256), 0, s, l, sl, v0, v1, off, out);
}

// pending frees of size class c of pool h -> its free list (threads t0 of
// nt move the offsets); one caller per class, after every pop of the batch
__device__ void recycle_class(const StoreArgs &a, int h, int c, uint32_t t0, uint32_t nt)
{
    Pool *p = a.pool;
    unsigned int avail = p->free_n[h][c];
    unsigned int take = p->free_take[h][c];
    unsigned int rem = take >= avail ? 0 : avail - take;   // pops took from the end
    const unsigned int pn_all = p->pend_n[h][c];
    const unsigned int pn = pn_all > a.fcap ? a.fcap : pn_all;
    unsigned int room = a.fcap - rem;
    unsigned int mv = pn < room ? pn : room;
    uint64_t base = ((uint64_t)h * NCLS + c) * a.fcap;
    for (unsigned int k = t0; k < mv; k += nt) a.free_off[base + rem + k] = a.pend_off[base + k];
    __syncthreads();
    if (t0 == 0) {
        p->free_n[h][c] = rem + mv;
        p->free_take[h][c] = 0;
        p->pend_n[h][c] = 0;
        if (mv) atomicAdd(&p->recycled[h], (unsigned long long)mv);
        if (pn_all > mv) {
            atomicAdd(&p->dropped[h], (unsigned long long)(pn_all - mv));
            atomicAdd(&p->dropped_entries[h], (unsigned long long)(pn_all - mv) << c);
        }
    }
}

// the batch is over: reset the per-batch counters and publish the batch
// status (bad ids, pool exhaustion) to pinned host memory
__device__ void end_batch(const StoreArgs &a)
{
    Pool *p = a.pool;
    p->seg_pack = 0;   // segments and owners of the next batch
    p->n_heavy = 0;
    if (a.hstatus) {
        a.hstatus[0] = __hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        a.hstatus[1] = __hip_atomic_load(&p->oom, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence_system();
    }
}

// pending frees -> free lists, once per batch (one block per pool/class)
__global__ void k_pool_recycle(StoreArgs a)
{
    recycle_class(a, blockIdx.x / NCLS, blockIdx.x % NCLS, threadIdx.x, blockDim.x);
    if (blockIdx.x == 0 && threadIdx.x == 0) end_batch(a);
}

// A batch of at most SMALL_OPS edge ops in one workgroup (the per-update
// path): expand, group the list ops by list with a stable rank sort in LDS,
// apply the light lists with AG-lane groups and the heavy ones with the
// whole workgroup (the same routines as the batch kernels), then recycle the
// pool's pending frees and publish the batch status.  One launch instead of
// six.
__global__ void __launch_bounds__(HB) k_store_small(StoreArgs a)
{
    constexpr uint32_t SL = 4 * SMALL_OPS;
    static_assert(SL <= HB, "one thread per list op");
    __shared__ HeavyShared sh;
    __shared__ uint32_t key[SL], order[SL], sstart[SL + 1], hseg[SL];
    __shared__ uint32_t wcnt[HB / WAVE];
    __shared__ uint32_t nvalid, nseg, nheavy;
    const uint32_t tid = threadIdx.x, lane = __lane_id(), wv = tid / WAVE;
    const uint32_t n = a.ops.n, nl = a.lstride * n;
    if (tid == 0) {
        nvalid = 0;
        nheavy = 0;
    }
    if (tid < n) expand_op(a, tid, false, n);
    __syncthreads();
    // list key (L << 31 | v) of every list op; NILP for unused slots
    uint32_t kt = NILP;
    if (tid < nl) {
        const uint32_t v = a.lops.op[tid].v;
        if (v != NILP) kt = v | ((a.lops.op[tid].flags & L_IN) ? 0x80000000u : 0u);
    }
    if (tid < SL) key[tid] = kt;
    if (kt != NILP) atomicAdd(&nvalid, 1u);
    __syncthreads();
    // stable rank sort: ops of one list become a contiguous run in op order
    if (kt != NILP) {
        uint32_t r = 0;
        for (uint32_t u = 0; u < nl; u++) {
            const uint32_t ku = key[u];
            r += (ku != NILP && (ku < kt || (ku == kt && u < tid))) ? 1u : 0u;
        }
        order[r] = tid;
    }
    __syncthreads();
    const uint32_t nv = nvalid;
    if (tid < nv) a.lops.segbuf[tid] = order[tid];
    // run starts, compacted in order (wave ballots plus wave offsets)
    const bool head = tid < nv && (tid == 0 || key[order[tid]] != key[order[tid - 1]]);
    const uint64_t bal = __ballot(head);
    if (lane == 0) wcnt[wv] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t off = 0;
    for (uint32_t q = 0; q < wv; q++) off += wcnt[q];
    if (head) sstart[off + (uint32_t)__popcll(bal & ((1ull << lane) - 1))] = tid;
    if (tid == 0) {
        uint32_t t = 0;
        for (uint32_t q = 0; q < HB / WAVE; q++) t += wcnt[q];
        nseg = t;
        sstart[t] = nv;
    }
    __syncthreads();
    // light lists: one AG-lane group each; heavy ones are queued
    {
        Grp g;
        g.lane = (int)(tid % AG);
        g.base = (int)(tid % WAVE) - g.lane;
        for (uint32_t sg = tid / AG; sg < nseg; sg += HB / AG) {
            const uint32_t b = sstart[sg], cnt = sstart[sg + 1] - b;
            const uint32_t j0 = order[b];
            const uint32_t v = a.lops.op[j0].v;
            const int L = (a.lops.op[j0].flags & L_IN) ? 1 : 0;
            ListHdr &hd = a.lists[L].h[v];
            if (heavy_list(hd, cnt)) {
                if (g.lane == 0) hseg[atomicAdd(&nheavy, 1u)] = sg;
                continue;
            }
            ListState st = load_state(hd);
            const uint32_t *ops = a.lops.segbuf + b;
            int32_t n_ins = 0;
            for (uint32_t i = g.lane; i < cnt; i += AG) n_ins += a.lops.op[ops[i]].delta > 0 ? 1 : 0;
            const uint32_t need = st.len + (uint32_t)gsum<AG>(n_ins);
            if (light_room(a, st, need, 0u, 0ull, g)) {
                apply_serial(a, st, ops, cnt, g);
                __threadfence_block();
                write_liveness(a, st, ops, cnt, g.lane, AG);
            }
            if (g.lane == 0) store_state(hd, st);
        }
    }
    __syncthreads();
    // heavy lists, one at a time with the whole workgroup
    const uint32_t nh = nheavy;
    for (uint32_t q = 0; q < nh; q++) {
        const uint32_t sg = hseg[q];
        const uint32_t b = sstart[sg], cnt = sstart[sg + 1] - b;
        const uint32_t j0 = order[b];
        ListHdr &hd = a.lists[(a.lops.op[j0].flags & L_IN) ? 1 : 0].h[a.lops.op[j0].v];
        heavy_list_apply(a, hd, a.lops.segbuf + b, cnt, sh);
    }
    __syncthreads();
    __threadfence();
    // every pop of the batch is done: pending frees -> free lists, all
    // classes at once (HB / (2 NCLS) threads per class; the counters are
    // read again only by the next batch's kernels)
    {
        constexpr uint32_t TPC = HB / (2 * NCLS);
        static_assert(TPC >= 1, "a thread per class at least");
        const uint32_t hcls = tid / TPC, t0 = tid % TPC;
        if (hcls < 2 * NCLS) {
            const int h = (int)(hcls / NCLS), c = (int)(hcls % NCLS);
            Pool *p = a.pool;
            const unsigned int avail = p->free_n[h][c], take = p->free_take[h][c];
            const unsigned int rem = take >= avail ? 0 : avail - take;
            const unsigned int pn_all = p->pend_n[h][c];
            const unsigned int pn = pn_all > a.fcap ? a.fcap : pn_all;
            const unsigned int mv = pn < a.fcap - rem ? pn : a.fcap - rem;
            const uint64_t base = ((uint64_t)h * NCLS + c) * a.fcap;
            for (unsigned int k = t0; k < mv; k += TPC) a.free_off[base + rem + k] = a.pend_off[base + k];
            if (t0 == 0) {
                p->free_n[h][c] = rem + mv;
                p->free_take[h][c] = 0;
                p->pend_n[h][c] = 0;
                if (mv) atomicAdd(&p->recycled[h], (unsigned long long)mv);
                if (pn_all > mv) {
                    atomicAdd(&p->dropped[h], (unsigned long long)(pn_all - mv));
                    atomicAdd(&p->dropped_entries[h], (unsigned long long)(pn_all - mv) << c);
                }
            }
        }
    }
    __syncthreads();
    if (tid == 0) end_batch(a);
}

void launch_pool_recycle(const StoreArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(k_pool_recycle, dim3(2 * NCLS), dim3(256), 0, s, a);
}

// ---------------------------------------------------------------------
// synthetic streams (rg_gen.h) and stream -> edge-op conversion
// ---------------------------------------------------------------------
__global__ void k_gen_rmat(uint64_t *src, uint64_t *dst, uint64_t n, uint64_t start, int scale,
                           uint64_t key, RmatThresholds t)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        rmat_edge(key, start + i, scale, t, &src[i], &dst[i]);
}

__global__ void k_gen_er(uint64_t *src, uint64_t *dst, uint64_t n, uint64_t start, uint64_t V, uint64_t key)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        er_edge(key, start + i, V, &src[i], &dst[i]);
}

__global__ void k_gen_w(uint32_t *w, uint64_t n, uint64_t start, uint64_t key)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        w[i] = stream_weight(key, start + i);
}

static dim3 gen_grid(uint64_t n)
{
    uint64_t b = (n + 255) / 256;
    if (b > 16384) b = 16384;
    if (b == 0) b = 1;
    return dim3((uint32_t)b);
}

void launch_gen_rmat(uint64_t *src, uint64_t *dst, uint64_t n, uint64_t start, int scale,
                     uint64_t seed, double a, double b, double c, hipStream_t s)
{
    hipLaunchKernelGGL(k_gen_rmat, gen_grid(n), dim3(256), 0, s, src, dst, n, start, scale,
                       draw_key(seed), rmat_thresholds(a, b, c));
}

void launch_gen_er(uint64_t *src, uint64_t *dst, uint64_t n, uint64_t start, uint64_t V,
                   uint64_t seed, hipStream_t s)
{
    hipLaunchKernelGGL(k_gen_er, gen_grid(n), dim3(256), 0, s, src, dst, n, start, V, draw_key(seed));
}

void launch_gen_weights(uint32_t *w, uint64_t n, uint64_t start, uint64_t seed, hipStream_t s)
{
    hipLaunchKernelGGL(k_gen_w, gen_grid(n), dim3(256), 0, s, w, n, start, draw_key(seed));
}

// out-degree histogram of a stream prefix and its arg-max (ties -> lowest id),
// the drivers' root choice in these workloads
__global__ void k_hist(const uint64_t *src, uint64_t n, uint64_t V, uint32_t *deg)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        if (src[i] < V) atomicAdd(&deg[src[i]], 1u);
}

__global__ void k_argmax(const uint32_t *deg, uint64_t V, unsigned long long *best)
{
    unsigned long long b = 0;
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V;
         v += (uint64_t)gridDim.x * blockDim.x) {
        unsigned long long x = ((unsigned long long)deg[v] << 32) | (0xFFFFFFFFull - v);
        if (x > b) b = x;
    }
    atomicMax(best, b);
}

void launch_max_degree(const uint64_t *src, uint64_t n, uint64_t V, uint32_t *deg_tmp,
                       unsigned long long *best, hipStream_t s)
{
    (void)hipMemsetAsync(deg_tmp, 0, V * sizeof(uint32_t), s);
    (void)hipMemsetAsync(best, 0, sizeof(unsigned long long), s);
    hipLaunchKernelGGL(k_hist, gen_grid(n), dim3(256), 0, s, src, n, V, deg_tmp);
    hipLaunchKernelGGL(k_argmax, gen_grid(V), dim3(256), 0, s, deg_tmp, V, best);
}

__global__ void k_ops_from_stream(const uint64_t *src, const uint64_t *dst, const uint32_t *w,
                                  uint64_t first, uint32_t n, uint8_t kind, EdgeOps ops, uint32_t base)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ops.src[base + i] = (uint32_t)src[first + i];
    ops.dst[base + i] = (uint32_t)dst[first + i];
    ops.w[base + i] = w ? w[first + i] : 0;
    ops.kind[base + i] = kind;
}

void launch_ops_from_stream(const uint64_t *src, const uint64_t *dst, const uint32_t *w,
                            uint64_t first, uint32_t n, uint8_t kind, EdgeOps ops,
                            uint32_t base, hipStream_t s)
{
    if (!n) return;
    hipLaunchKernelGGL(k_ops_from_stream, dim3((n + 255) / 256), dim3(256), 0, s, src, dst, w, first, n,
                       kind, ops, base);
}

// one batch of a sliding window over a circular stream of `len` edges: ops
// [0, n_ins) insert stream edges ins_first.., ops [n_ins, n_ins + n_del)
// delete stream edges del_first.., both positions taken modulo len
__global__ void k_ops_from_window(const uint64_t *src, const uint64_t *dst, const uint32_t *w, uint64_t len,
                                  uint64_t ins_first, uint32_t n_ins, uint64_t del_first, uint32_t n_del,
                                  uint8_t kind_ins, uint8_t kind_del, EdgeOps ops, uint32_t *zero)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && zero) *zero = 0;
    if (i >= n_ins + n_del) return;
    const bool ins = i < n_ins;
    uint64_t e = ins ? ins_first + i : del_first + (i - n_ins);
    e %= len;
    ops.src[i] = (uint32_t)src[e];
    ops.dst[i] = (uint32_t)dst[e];
    ops.w[i] = w ? w[e] : 0;
    ops.kind[i] = ins ? kind_ins : kind_del;
}

__global__ void k_ops_from_window_dyn(const uint64_t *src, const uint64_t *dst, const uint32_t *w, uint64_t len,
                                      const BatchDesc *desc, uint8_t kind_ins, uint8_t kind_del, EdgeOps ops,
                                      uint32_t *zero)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && zero) *zero = 0;
    const BatchDesc d = *desc;
    if (i >= d.n_ins + d.n_del) return;
    const bool ins = i < d.n_ins;
    uint64_t e = ins ? d.ins_first + i : d.del_first + (i - d.n_ins);
    e %= len;
    ops.src[i] = (uint32_t)src[e];
    ops.dst[i] = (uint32_t)dst[e];
    ops.w[i] = w ? w[e] : 0;
    ops.kind[i] = ins ? kind_ins : kind_del;
}

void launch_ops_from_window_dyn(const uint64_t *src, const uint64_t *dst, const uint32_t *w, uint64_t len,
                                const BatchDesc *desc, uint32_t n_max, uint8_t kind_ins, uint8_t kind_del,
                                EdgeOps ops, uint32_t *zero, hipStream_t s)
{
    hipLaunchKernelGGL(k_ops_from_window_dyn, dim3((n_max + 255) / 256), dim3(256), 0, s, src, dst, w, len, desc,
                       kind_ins, kind_del, ops, zero);
}

void launch_ops_from_window(const uint64_t *src, const uint64_t *dst, const uint32_t *w, uint64_t len,
                            uint64_t ins_first, uint32_t n_ins, uint64_t del_first, uint32_t n_del,
                            uint8_t kind_ins, uint8_t kind_del, EdgeOps ops, uint32_t *zero, hipStream_t s)
{
    const uint32_t n = n_ins + n_del;
    if (!n) return;
    hipLaunchKernelGGL(k_ops_from_window, dim3((n + 255) / 256), dim3(256), 0, s, src, dst, w, len, ins_first,
                       n_ins, del_first, n_del, kind_ins, kind_del, ops, zero);
}

}  // namespace rg
