// inc.cpp -- sliding-window incremental driver over the Graph<> facade.
//
// Built once per algorithm (-DALGO_BFS / -DALGO_SSSP / -DALGO_SSWP / -DALGO_WCC)
// and per mode (plain, or -DBATCH), it runs the workload of the reference
// drivers bfs_inc / sssp_inc / sswp_inc / wcc_inc (plain: add edge i, then
// delete edge i - imported, each followed by its incremental repair;
// bfs_inc.cpp:113-192) and *_inc_batch (batched adds, then batched deletes;
// bfs_inc_batch.cpp:136-186) through the same Graph<> calls.
//
// usage:  <prog> graph_file root imported_rate [batch]      (BFS/SSSP/SSWP)
//         <prog> graph_file imported_rate [batch]           (WCC)
// Weights follow the reference drivers: (src + dst) % 128 (sssp_inc.cpp:65).
// At the end the incremental labels are compared with a fresh build_tree on
// the final graph (the reference drivers' own check, bfs_inc.cpp:220-293);
// RG_DUMP=<file> writes the final values as u64.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "risgraph/graph.hpp"

#if defined(ALGO_BFS)
using ED = void;
static const char *NAME = "bfs";
static const bool SYMMETRIC = false;
#elif defined(ALGO_SSSP)
using ED = uint64_t;
static const char *NAME = "sssp";
static const bool SYMMETRIC = false;
#elif defined(ALGO_SSWP)
using ED = uint64_t;
static const char *NAME = "sswp";
static const bool SYMMETRIC = false;
#elif defined(ALGO_WCC)
using ED = void;
static const char *NAME = "wcc";
static const bool SYMMETRIC = true;
#else
#error "define ALGO_BFS, ALGO_SSSP, ALGO_SSWP or ALGO_WCC"
#endif

using G = Graph<ED>;
using EdgeT = G::edge_type;

static EdgeT make_edge(uint64_t s, uint64_t d)
{
    EdgeT e;
    e.src = s;
    e.dst = d;
#if defined(ALGO_SSSP) || defined(ALGO_SSWP)
    e.data = (s + d) % 128;
#endif
    return e;
}

static double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char **argv)
{
    const int need = SYMMETRIC ? 3 : 4;
    if (argc < need) {
        fprintf(stderr, SYMMETRIC ? "usage: %s graph imported_rate [batch]\n" : "usage: %s graph root imported_rate [batch]\n",
                argv[0]);
        return 1;
    }
    auto file = mmap_binary(argv[1]);
    const std::pair<uint64_t, uint64_t> *raw = file.first;
    const uint64_t len = file.second;
    uint64_t root = SYMMETRIC ? 0 : std::stoull(argv[2]);
    double rate = std::stod(argv[SYMMETRIC ? 2 : 3]);
    uint64_t batch = argc > need ? std::stoull(argv[need]) : 0;
#ifdef BATCH
    if (!batch) batch = 1024;
#endif
    uint64_t V = 0;
    for (uint64_t i = 0; i < len; i++) {
        V = std::max(V, raw[i].first + 1);
        V = std::max(V, raw[i].second + 1);
    }
    const bool directed = !SYMMETRIC;
    G graph(V, len, SYMMETRIC, !SYMMETRIC);
    const uint64_t imported = (uint64_t)(len * rate);
    {
        std::vector<EdgeT> e(imported);
        for (uint64_t i = 0; i < imported; i++) e[i] = make_edge(raw[i].first, raw[i].second);
        graph.add_edges(e.data(), imported, directed);
    }
    auto labels = graph.alloc_vertex_tree_array<uint64_t>();

    // the reference driver lambdas (bfs_inc.cpp:72-92 and siblings)
#if defined(ALGO_BFS)
    const uint64_t MAXL = 65536;
    auto update_func = [](uint64_t, uint64_t, uint64_t s, uint64_t d, G::adjedge_type) {
        return std::make_pair(s + 1 < d, s + 1);
    };
    auto equal_func = [](uint64_t, uint64_t, uint64_t s, uint64_t d, G::adjedge_type) { return s + 1 == d; };
    auto init_label_func = [=](uint64_t v) { return std::make_pair(v == root ? 0 : MAXL, v == root); };
#elif defined(ALGO_SSSP)
    const uint64_t MAXL = 134217728;
    auto update_func = [](uint64_t, uint64_t, uint64_t s, uint64_t d, G::adjedge_type e) {
        return std::make_pair(s + e.data < d, s + e.data);
    };
    auto equal_func = [](uint64_t, uint64_t, uint64_t s, uint64_t d, G::adjedge_type e) { return s + e.data == d; };
    auto init_label_func = [=](uint64_t v) { return std::make_pair(v == root ? 0 : MAXL, v == root); };
#elif defined(ALGO_SSWP)
    const uint64_t MAXL = 134217728;
    auto update_func = [](uint64_t, uint64_t, uint64_t s, uint64_t d, G::adjedge_type e) {
        uint64_t c = std::min<uint64_t>(s, e.data);
        return std::make_pair(c > d, c);
    };
    auto equal_func = [](uint64_t, uint64_t, uint64_t s, uint64_t d, G::adjedge_type e) {
        return std::min<uint64_t>(s, e.data) == d;
    };
    auto init_label_func = [=](uint64_t v) { return std::make_pair(v == root ? MAXL : 0, v == root); };
#else
    auto update_func = [](uint64_t, uint64_t, uint64_t s, uint64_t d, G::adjedge_type) {
        return std::make_pair(s < d, s);
    };
    auto equal_func = [](uint64_t, uint64_t, uint64_t s, uint64_t d, G::adjedge_type) { return s == d; };
    auto init_label_func = [](uint64_t v) { return std::make_pair(v, true); };
#endif
    auto continue_reduce_func = [](uint64_t, uint64_t total, uint64_t local) {
        return std::make_pair(local > 0, total + local);
    };
    auto active_result_func = [](uint64_t old, uint64_t, uint64_t, uint64_t, uint64_t, uint64_t) { return old + 1; };

    double t0 = now_s();
    graph.build_tree<uint64_t, uint64_t>(init_label_func, continue_reduce_func, update_func, active_result_func,
                                         labels);
    fprintf(stderr, "%s: |V|=%lu |E|=%lu imported=%lu build %.3fs\n", NAME, V, len, imported, now_s() - t0);

    uint64_t updates = 0;
    t0 = now_s();
    if (!batch) {
        for (uint64_t i = imported; i < len; i++) {
            EdgeT a = make_edge(raw[i].first, raw[i].second);
            if (graph.add_edge(a, directed) == 0)
                graph.update_tree_add<uint64_t, uint64_t>(continue_reduce_func, update_func, active_result_func,
                                                          labels, a, directed);
            EdgeT d = make_edge(raw[i - imported].first, raw[i - imported].second);
            if (graph.del_edge(d, directed) == 1)
                graph.update_tree_del<uint64_t, uint64_t>(init_label_func, continue_reduce_func, update_func,
                                                          active_result_func, equal_func, labels, d, directed);
            updates += 2;
        }
    } else {
        std::vector<EdgeT> adds, dels, fresh;
        std::vector<uint64_t> old;
        for (uint64_t b = imported; b < len; b += batch) {
            uint64_t e = std::min(b + batch, len);
            adds.clear();
            dels.clear();
            for (uint64_t i = b; i < e; i++) {
                adds.push_back(make_edge(raw[i].first, raw[i].second));
                dels.push_back(make_edge(raw[i - imported].first, raw[i - imported].second));
            }
            old.resize(adds.size());
            graph.add_edges(adds.data(), adds.size(), directed, old.data());
            fresh.clear();
            for (size_t k = 0; k < adds.size(); k++)
                if (old[k] == 0) fresh.push_back(adds[k]);
            if (!fresh.empty())
                graph.update_tree_add<uint64_t, uint64_t>(continue_reduce_func, update_func, active_result_func,
                                                          labels, fresh, fresh.size(), directed);
            graph.del_edges(dels.data(), dels.size(), directed, old.data());
            fresh.clear();
            for (size_t k = 0; k < dels.size(); k++)
                if (old[k] == 1) fresh.push_back(dels[k]);
            if (!fresh.empty())
                graph.update_tree_del<uint64_t, uint64_t>(init_label_func, continue_reduce_func, update_func,
                                                          active_result_func, equal_func, labels, fresh,
                                                          fresh.size(), directed);
            updates += 2 * (e - b);
        }
    }
    double el = now_s() - t0;
    fprintf(stderr, "%s: %lu updates in %.3fs (%.0f updates/s)\n", NAME, updates, el, updates / std::max(el, 1e-9));

    std::vector<uint64_t> inc(V);
    uint64_t sum = 0, reached = 0;
    for (uint64_t v = 0; v < V; v++) {
        inc[v] = labels[v].data;
        sum += inc[v];
        reached += labels[v].parent.nbr != V;
    }
    auto fresh = graph.alloc_vertex_tree_array<uint64_t>();
    graph.build_tree<uint64_t, uint64_t>(init_label_func, continue_reduce_func, update_func, active_result_func,
                                         fresh);
    uint64_t diff = 0;
    for (uint64_t v = 0; v < V; v++) diff += fresh[v].data != inc[v];
    printf("%s values: sum=%lu with_parent=%lu rebuild_mismatch=%lu\n", NAME, sum, reached, diff);
    if (const char *dump = getenv("RG_DUMP")) {
        FILE *f = fopen(dump, "wb");
        if (f) {
            fwrite(inc.data(), 8, V, f);
            fclose(f);
        }
    }
    return diff == 0 ? 0 : 2;
}
