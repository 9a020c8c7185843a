"""Durability: write-ahead log + snapshot + replay (the reference's WAL,
all_inc_rpc.cpp:305-336 and 572-590, made recoverable).

Three trees (BFS, SSSP and WCC on the undirected view) run transaction
batches with a WAL open; a snapshot is taken after the third batch; four
more batches follow; then the process "crashes" (the graph is destroyed and
a partial record is appended to the log, as a write cut short would leave
it).  Recovery loads the snapshot and replays the log: it must skip the
batches the snapshot holds, apply the four after it, stop at the torn tail,
and end at the pre-crash version with the pre-crash values, degrees and
edge counts, values equal to the closed form and valid trees.
"""
import os
import struct

import numpy as np
import pytest

import oracle
import risgraph_amd as rg

pytestmark = pytest.mark.gpu


def _random_txns(rng, live, src, dst, w, nxt, V):
    lens = rng.integers(1, 17, int(rng.integers(30, 120))).astype(np.uint32)
    ops, used = [], set()
    while len(ops) < int(lens.sum()):
        if rng.random() < 0.45 and live:
            j = int(rng.integers(0, len(live)))
            live[j], live[-1] = live[-1], live[j]
            e = live.pop()
            if e in used:
                live.insert(0, e)
                continue
            ops.append(e + (-1,))
        else:
            e = (int(src[nxt[0]]), int(dst[nxt[0]]), int(w[nxt[0]]))
            nxt[0] += 1
            if e in used:
                continue
            ops.append(e + (1,))
        used.add(ops[-1][:3])
    return ops, lens


def test_snapshot_and_wal_recover_the_graph(tmp_path):
    rng = np.random.default_rng(11)
    V, E = 4096, 1 << 16
    s = rg.Stream("rmat", 12, E, seed=31, wseed=32)
    src, dst, w = s.read()
    w = w.astype(np.uint64)
    n0 = E // 2
    g = rg.Graph(V, E, symmetric=False, dual=True, weighted=True)
    g.load_stream(s, 0, n0, True)
    root = s.max_degree_vertex(n0, V)
    Ls = [g.alloc_vertex_tree_array(a, root) for a in (rg.BFS, rg.SSSP, rg.WCC)]
    for L in Ls:
        L.build_tree()
    wal = str(tmp_path / "graph.wal")
    snap = str(tmp_path / "graph.snap")
    g.wal_open(wal, sync=True)
    live = [(int(a), int(b), int(c)) for a, b, c in zip(src[:n0], dst[:n0], w[:n0])]
    nxt = [n0]
    for b in range(7):
        if b == 3:
            g.snapshot_save(Ls, snap)
            assert os.path.exists(snap)
        ops, lens = _random_txns(rng, live, src, dst, w, nxt, V)
        a = np.array([o[0] for o in ops], np.uint64)
        d = np.array([o[1] for o in ops], np.uint64)
        ww = np.array([o[2] for o in ops], np.uint64)
        k = np.array([o[3] for o in ops], np.int8)
        g.txn_apply(Ls, a, d, ww, k, lens, True)
        live.extend(o[:3] for o in ops if o[3] > 0)
    assert g.version() == 7
    before = [L.read()[0] for L in Ls]
    deg = g.degrees()
    sample = live[:: max(1, len(live) // 300)]
    counts = [g.get_edge_num(a, b, c) for a, b, c in sample]
    g.close()
    # a record cut short at the end of the log
    with open(wal, "ab") as f:
        f.write(struct.pack("<IIQQ", 0x31424752, 5, 8, 0) + b"\x08" * 40)   # header + 40 of 160 entry bytes
    g2, L2 = rg.snapshot_load(snap)
    assert g2.version() == 3 and [L.algo for L in L2] == [rg.BFS, rg.SSSP, rg.WCC]
    assert g2.wal_replay(L2, wal) == 4
    assert g2.version() == 7
    for L, v in zip(L2, before):
        assert np.array_equal(L.read()[0], v), L.algo
    d2 = g2.degrees()
    assert np.array_equal(deg[0], d2[0]) and np.array_equal(deg[1], d2[1])
    assert counts == [g2.get_edge_num(a, b, c) for a, b, c in sample]
    # the closed form on the recovered store, and valid trees
    m = {}
    for e in live:
        m[e] = m.get(e, 0) + 1
    keys = sorted(m)
    ls = np.array([e[0] for e in keys for _ in range(m[e])], np.uint64)
    ld = np.array([e[1] for e in keys for _ in range(m[e])], np.uint64)
    lw = np.array([e[2] for e in keys for _ in range(m[e])], np.uint64)
    assert len(ls) == int(d2[0].sum())
    assert all(g2.get_edge_num(*e) == m[e] for e in keys[:: max(1, len(keys) // 200)])
    for L in L2:
        vals, par = L.read()
        assert np.array_equal(vals, oracle.static_values(L.algo, V, ls, ld, lw, root)), L.algo
        assert oracle.check_tree(L.algo, V, ls, ld, lw, root, vals, par, undirected=L.algo == rg.WCC) == 0
    # replaying again changes nothing (every logged batch is at or below the version)
    assert g2.wal_replay(L2, wal) == 0
