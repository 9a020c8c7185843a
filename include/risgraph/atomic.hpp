// risgraph/atomic.hpp -- the host atomics the drivers call (reference
// core/atomic.hpp:47-108, 205-223): cas, write_min, write_max and
// atomic_append.
#pragma once
#include <atomic>
#include <cstdint>
#include <mutex>
#include <type_traits>
#include <vector>

template <typename T>
inline bool cas(T *ptr, T oldv, T newv)
{
    static_assert(std::is_integral<T>::value, "cas on integers");
    return __atomic_compare_exchange_n(ptr, &oldv, newv, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
}

template <typename T>
inline bool write_min(T *a, T b)
{
    T c = __atomic_load_n(a, __ATOMIC_RELAXED);
    while (c > b) {
        if (__atomic_compare_exchange_n(a, &c, b, false, __ATOMIC_SEQ_CST, __ATOMIC_RELAXED)) return true;
    }
    return false;
}

template <typename T>
inline bool write_max(T *a, T b)
{
    T c = __atomic_load_n(a, __ATOMIC_RELAXED);
    while (c < b) {
        if (__atomic_compare_exchange_n(a, &c, b, false, __ATOMIC_SEQ_CST, __ATOMIC_RELAXED)) return true;
    }
    return false;
}

// Append `data` to a vector shared by many threads whose logical length is
// `length`; returns its index.  The slot is taken with one atomic
// increment; only a slot past the vector's size takes the mutex to grow it
// (doubling), and a writer never stores into storage a grow may move: it
// waits until the vector covers its index, then stores under the mutex.
template <typename T>
inline uint64_t atomic_append(std::vector<T> &array, std::atomic_uint64_t &length, const T &data, std::mutex &mutex)
{
    const uint64_t idx = length.fetch_add(1, std::memory_order_relaxed);
    std::lock_guard<std::mutex> lk(mutex);
    if (idx >= array.size()) array.resize(std::max<uint64_t>(idx + 1, 2 * array.size()));
    array[idx] = data;
    return idx;
}
