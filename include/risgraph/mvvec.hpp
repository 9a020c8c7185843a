// risgraph/mvvec.hpp -- multi-version per-vertex values on the host, the
// interface of the reference's core/mvvec.hpp:23-162 that the rt / rpc
// drivers fill from the modified set (bfs_inc_rt.cpp:353-358,
// all_inc_rpc.cpp:656-670):
//
//   set(id, version, value[, gc_version])  record `value` as id's value from
//                                          `version` on (a same-version set
//                                          overwrites; an unchanged value
//                                          adds nothing)
//   get(id, version)                       id's value at `version`: the
//                                          newest record at or before it
//   gc(id, version) / gc(version)          drop records that no read at or
//                                          after `version` can reach
//
// Layout: one newest-first chain of records per vertex.  Records live in a
// pool of fixed-size chunks that never move (a chunk table reserved up
// front), so set() on different vertices can run concurrently: a new
// record takes its slot with one atomic increment and is published with a
// compare-and-swap on the vertex's head.  Freed records go to a free list
// and are reused.  Concurrent set() on the same vertex is safe; get() and
// gc() of a vertex must not overlap set() of that vertex, as in the
// reference.
//
// The device keeps its own history store for the serving loop
// (rg_history_* in risgraph_amd.h); this class is for drivers that write
// the history themselves from get_modified().
#pragma once
#include <atomic>
#include <cstdint>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <vector>

template <typename DataType>
class MVVec {
    struct Rec {
        uint64_t version;
        uint64_t prev;   // older record of the same vertex, NIL = none
        DataType data;
    };
    static constexpr uint64_t NIL = ~0ull;
    static constexpr uint64_t CHUNK_BITS = 16;
    static constexpr uint64_t CHUNK = 1ull << CHUNK_BITS;
    static constexpr uint64_t MAX_CHUNKS = 1ull << 20;   // 2^36 records at most

public:
    explicit MVVec(uint64_t length)
        : n_(length), head_(new std::atomic<uint64_t>[length ? length : 1]), chunks_(MAX_CHUNKS)
    {
        for (uint64_t i = 0; i < n_; i++) head_[i].store(NIL, std::memory_order_relaxed);
        for (auto &c : chunks_) c.store(nullptr, std::memory_order_relaxed);
    }
    ~MVVec()
    {
        for (auto &c : chunks_) delete[] c.load(std::memory_order_relaxed);
    }
    MVVec(MVVec &&o) noexcept
        : n_(o.n_), head_(std::move(o.head_)), chunks_(MAX_CHUNKS), top_(o.top_.load()), free_(std::move(o.free_))
    {
        for (uint64_t i = 0; i < MAX_CHUNKS; i++) {
            chunks_[i].store(o.chunks_[i].load(std::memory_order_relaxed), std::memory_order_relaxed);
            o.chunks_[i].store(nullptr, std::memory_order_relaxed);
        }
        o.n_ = 0;
    }
    MVVec(const MVVec &) = delete;
    MVVec &operator=(const MVVec &) = delete;

    uint64_t size() const { return n_; }

    // Returns true when a new record was added.
    bool set(uint64_t id, uint64_t version, const DataType &data, uint64_t gc_version = NIL)
    {
        check(id);
        bool added = false;
        uint64_t h = head_[id].load(std::memory_order_acquire);
        if (h != NIL && rec(h).version == version) {
            rec(h).data = data;   // same version: overwrite
        } else if (h == NIL || (rec(h).version < version && !(rec(h).data == data))) {
            const uint64_t r = alloc();
            rec(r).version = version;
            rec(r).data = data;
            for (;;) {
                rec(r).prev = h;
                if (head_[id].compare_exchange_weak(h, r, std::memory_order_acq_rel)) {
                    added = true;
                    break;
                }
                // another writer published first: keep ours only if still newer
                if (h != NIL && !(rec(h).version < version)) {
                    if (rec(h).version == version) rec(h).data = data;
                    release(r);
                    break;
                }
            }
        }
        if (gc_version != NIL) gc(id, gc_version);
        return added;
    }

    DataType get(uint64_t id, uint64_t version) const
    {
        check(id);
        for (uint64_t r = head_[id].load(std::memory_order_acquire); r != NIL; r = rec(r).prev)
            if (rec(r).version <= version) return rec(r).data;
        throw std::out_of_range("MVVec::get: no value at or before this version");
    }

    // Keep the newest record at or before `version` (reads at `version` and
    // later need nothing older) and recycle the records behind it.
    void gc(uint64_t id, uint64_t version)
    {
        check(id);
        uint64_t r = head_[id].load(std::memory_order_acquire);
        while (r != NIL && rec(r).version > version) r = rec(r).prev;
        if (r == NIL) return;
        uint64_t old = rec(r).prev;
        rec(r).prev = NIL;
        while (old != NIL) {
            const uint64_t nx = rec(old).prev;
            release(old);
            old = nx;
        }
    }
    void gc(uint64_t version)
    {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4096)
#endif
        for (uint64_t i = 0; i < n_; i++) gc(i, version);
    }

private:
    uint64_t n_;
    std::unique_ptr<std::atomic<uint64_t>[]> head_;
    std::vector<std::atomic<Rec *>> chunks_;
    std::atomic<uint64_t> top_{0};
    std::mutex free_mu_;
    std::vector<uint64_t> free_;

    void check(uint64_t id) const
    {
        if (id >= n_) throw std::out_of_range("MVVec: vertex id out of range");
    }
    Rec &rec(uint64_t r) const { return chunks_[r >> CHUNK_BITS].load(std::memory_order_acquire)[r & (CHUNK - 1)]; }

    uint64_t alloc()
    {
        {
            std::lock_guard<std::mutex> lk(free_mu_);
            if (!free_.empty()) {
                const uint64_t r = free_.back();
                free_.pop_back();
                return r;
            }
        }
        const uint64_t r = top_.fetch_add(1, std::memory_order_relaxed);
        const uint64_t c = r >> CHUNK_BITS;
        if (c >= MAX_CHUNKS) throw std::length_error("MVVec: record pool exhausted");
        if (!chunks_[c].load(std::memory_order_acquire)) {
            Rec *fresh = new Rec[CHUNK];
            Rec *expect = nullptr;
            if (!chunks_[c].compare_exchange_strong(expect, fresh, std::memory_order_acq_rel)) delete[] fresh;
        }
        return r;
    }
    void release(uint64_t r)
    {
        std::lock_guard<std::mutex> lk(free_mu_);
        free_.push_back(r);
    }
};
