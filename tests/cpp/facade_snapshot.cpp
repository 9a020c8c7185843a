// Copy semantics of the facade's labels (include/risgraph/graph.hpp): a copy
// of the label array is a snapshot, as a copy of the reference's
// std::vector<VertexTree> is (`auto labels_pre = labels;`, bfs_inc.cpp:110).
// Build BFS on a path 0 -> 1 -> ... -> 9 plus a shortcut, snapshot, delete
// the shortcut: the live labels change, the snapshot keeps the old levels.
#include <cstdio>

#include "risgraph/graph.hpp"

int main()
{
    const uint64_t V = 10, MAXL = 65536;
    Graph<void> graph(V, 64, false, true);
    for (uint64_t v = 0; v + 1 < V; v++) graph.add_edge({v, v + 1}, true);
    graph.add_edge({0, 7}, true);
    auto labels = graph.alloc_vertex_tree_array<uint64_t>();
    auto upd = [](uint64_t, uint64_t, uint64_t s, uint64_t d, decltype(graph)::adjedge_type) {
        return std::make_pair(s + 1 < d, s + 1);
    };
    auto eq = [](uint64_t, uint64_t, uint64_t s, uint64_t d, decltype(graph)::adjedge_type) { return s + 1 == d; };
    auto init = [](uint64_t v) -> std::pair<uint64_t, bool> { return {v == 0 ? 0 : MAXL, v == 0}; };
    auto cont = [](uint64_t, uint64_t t, uint64_t l) { return std::make_pair(l > 0, t + l); };
    auto act = [](uint64_t r, uint64_t, uint64_t, uint64_t, uint64_t, uint64_t) { return r + 1; };
    graph.build_tree<uint64_t, uint64_t>(init, cont, upd, act, labels);
    int bad = 0;
    if (labels[7].data != 1 || labels[9].data != 3) bad |= 1;
    auto snap = labels;                      // snapshot
    if (snap.bound() || !labels.bound()) bad |= 2;
    if (graph.del_edge({0, 7}, true) == 1)
        graph.update_tree_del<uint64_t, uint64_t>(init, cont, upd, act, eq, labels, {0, 7}, true);
    if (labels[7].data != 7 || labels[9].data != 9) bad |= 4;     // live labels moved
    if (snap[7].data != 1 || snap[9].data != 3 || snap[8].parent.nbr != 7) bad |= 8;   // snapshot did not
    auto moved = std::move(labels);          // a move keeps the device binding
    if (!moved.bound() || moved[9].data != 9) bad |= 16;
    printf("facade snapshot check: %s (%d)\n", bad ? "FAIL" : "ok", bad);
    return bad;
}
