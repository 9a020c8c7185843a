// Host MVVec (include/risgraph/mvvec.hpp) against a brute-force model: a
// per-vertex map version -> value.  Sets come from many OpenMP threads on
// distinct vertices per round, as the rt drivers' stream_vertices_hybrid
// over the modified set does (bfs_inc_rt.cpp:353-358); reads at every
// earlier version must return the value that was current then.  gc(v)
// must keep every read at or after v intact.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <vector>

#include "risgraph/mvvec.hpp"

int main()
{
    const uint64_t V = 5000, ROUNDS = 300;
    MVVec<uint64_t> h(V);
    std::vector<std::map<uint64_t, uint64_t>> model(V);
    std::mt19937_64 rng(7);
    for (uint64_t v = 0; v < V; v++) {
        h.set(v, 0, v);
        model[v][0] = v;
    }
    for (uint64_t ver = 1; ver <= ROUNDS; ver++) {
        std::vector<uint64_t> ids, vals;
        for (uint64_t k = 0; k < 400; k++) {
            uint64_t v = rng() % V;
            ids.push_back(v);
            vals.push_back(rng() % 8);   // small range: unchanged values happen
        }
        // distinct vertices per parallel pass; a vertex touched twice in a
        // round is set twice in the same version (the second overwrites)
        std::vector<char> seen(V, 0);
        std::vector<uint64_t> a, b, a2, b2;
        for (size_t i = 0; i < ids.size(); i++) {
            if (!seen[ids[i]]) { seen[ids[i]] = 1; a.push_back(ids[i]); b.push_back(vals[i]); }
            else { a2.push_back(ids[i]); b2.push_back(vals[i]); }
        }
#pragma omp parallel for
        for (size_t i = 0; i < a.size(); i++) h.set(a[i], ver, b[i]);
        for (size_t i = 0; i < a2.size(); i++) h.set(a2[i], ver, b2[i]);
        for (size_t i = 0; i < a.size(); i++) model[a[i]][ver] = b[i];
        for (size_t i = 0; i < a2.size(); i++) model[a2[i]][ver] = b2[i];
    }
    auto at = [&](uint64_t v, uint64_t ver) { return std::prev(model[v].upper_bound(ver))->second; };
    long bad = 0;
    for (uint64_t v = 0; v < V; v++)
        for (uint64_t ver = 0; ver <= ROUNDS; ver += 7)
            if (h.get(v, ver) != at(v, ver)) bad++;
    const uint64_t g = 150;
    h.gc(g);
    for (uint64_t v = 0; v < V; v++)
        for (uint64_t ver = g; ver <= ROUNDS; ver += 3)
            if (h.get(v, ver) != at(v, ver)) bad++;
    // records freed by gc are reused by later sets
    for (uint64_t v = 0; v < V; v++) h.set(v, ROUNDS + 1, 1000 + v);
    for (uint64_t v = 0; v < V; v++)
        if (h.get(v, ROUNDS + 1) != 1000 + v || h.get(v, ROUNDS) != at(v, ROUNDS)) bad++;
    printf("mismatches %ld\n", bad);
    return bad ? 1 : 0;
}
