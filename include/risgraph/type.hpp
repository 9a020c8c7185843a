// risgraph/type.hpp -- edge types of the Graph<> interface (drop-in for the
// reference core/type.hpp:21-81).  Field names and meaning are kept because
// drivers use them directly: AdjEdge{nbr, num, data} with equality on
// (nbr, data) only, Edge{src, dst, data}.
#pragma once
#include <cstdint>
#include <functional>

template <typename Data>
struct AdjEdge {
    uint64_t nbr : 48;
    uint16_t num : 16;
    Data data;
    bool operator==(const AdjEdge &o) const { return nbr == o.nbr && data == o.data; }
    bool operator<(const AdjEdge &o) const { return nbr != o.nbr ? nbr < o.nbr : data < o.data; }
    using datatype = Data;
};

template <>
struct AdjEdge<void> {
    uint64_t nbr : 48;
    uint16_t num : 16;
    static const uint64_t data = 0;
    bool operator==(const AdjEdge &o) const { return nbr == o.nbr; }
    bool operator<(const AdjEdge &o) const { return nbr < o.nbr; }
    using datatype = uint64_t;
};

template <typename Data>
struct Edge {
    uint64_t src, dst;
    Data data;
    operator AdjEdge<Data>() const { return AdjEdge<Data>{dst, 0, data}; }
    using datatype = Data;
};

template <>
struct Edge<void> {
    uint64_t src, dst;
    static const uint64_t data = 0;
    operator AdjEdge<void>() const { return AdjEdge<void>{dst, 0}; }
    using datatype = uint64_t;
};

static_assert(sizeof(AdjEdge<void>) == 8, "AdjEdge<void> is one word");
static_assert(sizeof(Edge<void>) == 16, "Edge<void> is two words");

// sequential cut of the reference's parallel loops (type.hpp:105)
const uint64_t OPENMP_THRESHOLD = 8192;
