// history.hip -- device-resident multi-version vertex values.
//
// The device form of the reference's MVVec (core/mvvec.hpp:23-162) as the
// rt / rpc drivers fill it: after every batch, each vertex whose value
// changed gets a record (version, value) (bfs_inc_rt.cpp:353-358,
// all_inc_rpc.cpp:656-670), and a read at version t returns each vertex's
// newest value at or before t (MVVec::get).
//
// Layout (one labels instance):
//   rec   ring of `cap` 24-byte records {prev, version, value, vertex},
//         addressed by logical index i (slot i % cap); records are appended
//         in version order, one batch after another;
//   head  per vertex, the logical index of its newest record (NIL = none);
//         `prev` links a vertex's records newest to oldest;
//   base  per vertex, its value at `base_version` (the snapshot taken when
//         history was enabled, advanced by folds).
// Records with a logical index below `tail` are gone: a chain that reaches
// one ends there and the read falls back to base.  A fold (MVVec::gc) moves
// the tail past the oldest batches, writing each vertex's newest folded
// value into base, so the ring never overwrites a live record.
#include "rg_internal.h"

namespace rg {

__device__ __forceinline__ uint32_t hist_value(const HistArgs &h, uint32_t v)
{
    const uint64_t i = h.head[v];
    if (i != HNIL && i >= h.tail) return h.rec[i % h.cap].value;
    return h.base[v];
}

// base = current values, no records
__global__ void k_hist_init(HistArgs h, const VRec *vr, uint32_t V)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        h.base[v] = (uint32_t)(vr[v].val >> 32);
        h.head[v] = HNIL;
    }
}

// One thread per candidate vertex (the batch's modified set, or every
// vertex after a full rebuild): a vertex whose value differs from its newest
// record gets a record of `version`.  Records of one batch are distinct
// vertices (the modified set has no duplicates), so head updates never
// race.  *count (zeroed by the host before) receives the number appended;
// the host has made room for n records.
__global__ void k_hist_append(HistArgs h, const VRec *vr, const uint32_t *list, uint32_t n, uint32_t version,
                              uint64_t top, uint32_t *count)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool add = false;
    uint32_t v = 0, val = 0;
    if (i < n) {
        v = list ? list[i] : i;
        val = (uint32_t)(vr[v].val >> 32);
        add = val != hist_value(h, v);
    }
    const uint64_t m = __ballot(add);
    if (!m) return;
    const uint32_t lane = __lane_id();
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
    base = __shfl(base, (int)leader);
    if (!add) return;
    const uint64_t idx = top + base + (uint32_t)__popcll(m & ((1ull << lane) - 1));
    HRec r;
    r.prev = h.head[v];
    r.version = version;
    r.value = val;
    r.vertex = v;
    r.pad = 0;
    h.rec[idx % h.cap] = r;
    h.head[v] = idx;
}

// Fold records [h.tail, new_tail): each vertex's newest record in that range
// becomes its base value (MVVec::gc keeps, per vertex, the newest record at
// or before the gc version).  A record is its vertex's newest in the range
// when walking the vertex's chain from its head, skipping records at or
// above new_tail, reaches it first.
__global__ void k_hist_fold(HistArgs h, uint64_t new_tail)
{
    const uint64_t n = new_tail - h.tail;
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t idx = h.tail + k;
        const HRec r = h.rec[idx % h.cap];
        uint64_t j = h.head[r.vertex];
        while (j != HNIL && j >= new_tail) j = h.rec[j % h.cap].prev;
        if (j == idx) h.base[r.vertex] = r.value;
    }
}

// values at `version` of vertices [first, first + n)
__global__ void k_hist_read(HistArgs h, uint32_t first, uint32_t n, uint32_t version, uint64_t *out)
{
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t v = first + k;
        uint64_t j = h.head[v];
        uint32_t val = h.base[v];
        while (j != HNIL && j >= h.tail) {
            const HRec r = h.rec[j % h.cap];
            if (r.version <= version) {
                val = r.value;
                break;
            }
            j = r.prev;
        }
        out[k] = val;
    }
}

static dim3 hgrid(uint64_t n)
{
    uint64_t b = (n + 255) / 256;
    if (b > 16384) b = 16384;
    if (b < 1) b = 1;
    return dim3((uint32_t)b);
}

void launch_hist_init(const HistArgs &h, const VRec *vr, uint32_t V, hipStream_t s)
{
    hipLaunchKernelGGL(k_hist_init, hgrid(V), dim3(256), 0, s, h, vr, V);
}

void launch_hist_append(const HistArgs &h, const VRec *vr, const uint32_t *list, uint32_t n, uint32_t version,
                        uint64_t top, uint32_t *count, hipStream_t s)
{
    if (!n) return;
    hipLaunchKernelGGL(k_hist_append, dim3((n + 255) / 256), dim3(256), 0, s, h, vr, list, n, version, top, count);
}

void launch_hist_fold(const HistArgs &h, uint64_t new_tail, hipStream_t s)
{
    if (new_tail > h.tail) hipLaunchKernelGGL(k_hist_fold, hgrid(new_tail - h.tail), dim3(256), 0, s, h, new_tail);
}

void launch_hist_read(const HistArgs &h, uint32_t first, uint32_t n, uint32_t version, uint64_t *out, hipStream_t s)
{
    if (n) hipLaunchKernelGGL(k_hist_read, hgrid(n), dim3(256), 0, s, h, first, n, version, out);
}

}  // namespace rg
