// risgraph/atomic.hpp -- the host atomics the drivers call (reference
// core/atomic.hpp:47-108): cas, write_min, write_max.
#pragma once
#include <cstdint>
#include <type_traits>

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
