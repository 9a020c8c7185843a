"""Device value history (rg_history_*, the MVVec of mvvec.hpp:23-162 filled
per batch from the modified set as bfs_inc_rt.cpp:353-358 does) against
cpu_ref snapshots.

BFS and SSSP trees on one directed weighted store get a history ring small
enough that the oldest batches are folded while the stream runs.  After
every device batch cpu_ref runs the same batch (reference batch order: adds,
then deletes) and its values are kept under the graph's version.  At the
end, every version the history still keeps must read back exactly cpu_ref's
values of that version; versions already folded, or not reached yet, are
refused.  An explicit gc then folds up to a chosen version and the same
holds for what remains.
"""
import numpy as np
import pytest

import oracle
import risgraph_amd as rg

pytestmark = pytest.mark.gpu


def test_history_reads_past_versions_exactly():
    V, E = 4096, 1 << 16
    s = rg.Stream("rmat", 12, E, seed=21, wseed=22)
    src, dst, w = s.read()
    w = w.astype(np.uint64)
    n0 = E // 2   # the window stays inside the stream (no wrap on the host side)
    g = rg.Graph(V, E, symmetric=False, dual=True, weighted=True)
    g.load_stream(s, 0, n0, True)
    root = s.max_degree_vertex(n0, V)
    algos = (rg.BFS, rg.SSSP)
    Ls = [g.alloc_vertex_tree_array(a, root, trace_modified=True) for a in algos]
    for L in Ls:
        L.build_tree()
    ref = oracle.RefGraph(V, E, symmetric=False, dual=True, weighted=True)
    ref.bulk_load(src[:n0], dst[:n0], w[:n0], True)
    Rs = [oracle.RefLabels(ref, a, root) for a in algos]
    for R in Rs:
        R.build()
    cap = 6000
    for L in Ls:
        L.enable_history(cap)
    snaps = {g.version(): [R.read()[0] for R in Rs]}
    rng = np.random.default_rng(5)
    pos = 0
    sizes = [1, 3, 40, 200, 1000, 2500] + [int(x) for x in rng.integers(1, 1500, 14)]
    assert n0 + sum(sizes) + 5 <= E
    for B in sizes:
        g.apply_stream(Ls, s, n0 + pos, B, pos, B, True)
        idx = np.arange(n0 + pos, n0 + pos + B)
        olds = np.array([ref.add_edge(int(src[i]), int(dst[i]), int(w[i])) for i in idx])
        new = idx[olds == 0]
        didx = np.arange(pos, pos + B)
        for R in Rs:
            if len(new):
                R.inc_add_batch(src[new], dst[new], w[new], True)
        olds = np.array([ref.del_edge(int(src[i]), int(dst[i]), int(w[i])) for i in didx])
        gone = didx[olds == 1]
        for R in Rs:
            if len(gone):
                R.inc_del_batch(src[gone], dst[gone], w[gone], True)
        pos += B
        snaps[g.version()] = [R.read()[0] for R in Rs]
        for L, R in zip(Ls, Rs):
            assert np.array_equal(L.read()[0], R.read()[0])
    # plain one-update batches advance the clock too
    for k in range(5):
        g.stream_plain(Ls, s, n0 + pos, pos, 1, True, latency=False)
        i, j = n0 + pos, pos
        if ref.add_edge(int(src[i]), int(dst[i]), int(w[i])) == 0:
            for R in Rs:
                R.inc_add(int(src[i]), int(dst[i]), int(w[i]))
        if ref.del_edge(int(src[j]), int(dst[j]), int(w[j])) == 1:
            for R in Rs:
                R.inc_del(int(src[j]), int(dst[j]), int(w[j]))
        pos += 1
        snaps[g.version()] = [R.read()[0] for R in Rs]
    cur = g.version()
    assert cur == len(sizes) + 5
    info = [L.history_info() for L in Ls]
    assert all(i["capacity"] == cap and 0 < i["records"] <= cap for i in info)
    assert any(i["oldest_version"] > 0 for i in info), "the ring must have folded some batches"
    checked = 0
    for ver, vals in snaps.items():
        for t, L in enumerate(Ls):
            if ver >= info[t]["oldest_version"]:
                assert np.array_equal(L.history_read(ver), vals[t]), (ver, L.algo)
                checked += 1
            else:
                with pytest.raises(rg.RgError):
                    L.history_read(ver)
    assert checked >= len(snaps)
    with pytest.raises(rg.RgError):
        Ls[0].history_read(cur + 1)
    # explicit gc: reads at the gc version and later stay exact
    keep = sorted(snaps)[-4]
    for L in Ls:
        L.history_gc(keep)
    for t, L in enumerate(Ls):
        assert L.history_info()["oldest_version"] <= keep
        for ver in sorted(snaps)[-4:]:
            if ver in snaps:
                assert np.array_equal(L.history_read(ver), snaps[ver][t]), (ver, L.algo)
