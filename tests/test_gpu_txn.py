"""Transactions of up to 16 updates (rg_txn_apply; ClientTxnUpdateRequest,
rpc.hpp:40-61, and the transaction server all_inc_rpc_txn.cpp) against
cpu_ref.

BFS on a directed store and WCC on its undirected view.  Each round builds
random transactions of 1..16 updates mixing deletes of live edges with
inserts of new ones and applies them as one device batch.  cpu_ref starts
each round from the device's tree (values and parents) and runs the same
batch in the reference batch order (inserts, then deletes).  Checked:

* after every round the device values equal cpu_ref's and the closed form;
* a transaction is unsafe exactly when one of its updates reached the
  repair (an insert that created a live entry, a delete that removed the
  last copy) and need_update_tree_add / need_update_tree_del of cpu_ref
  (graph.hpp:987-1000, 1220-1233) holds for it in either tree, evaluated on
  the labels before the round;
* every transaction gets one version: the one before the batch when safe,
  the one after it when unsafe.
"""
import numpy as np
import pytest

import oracle
import risgraph_amd as rg

pytestmark = pytest.mark.gpu


def test_transactions_match_cpu_ref():
    rng = np.random.default_rng(3)
    V = 2048
    s = rg.Stream("rmat", 11, 30000, seed=9)
    src, dst, _ = s.read()
    n0 = 20000
    g = rg.Graph(V, 60000, symmetric=False, dual=True, weighted=False)
    g.load_stream(s, 0, n0, True)
    root = s.max_degree_vertex(n0, V)
    Lb = g.alloc_vertex_tree_array(rg.BFS, root)
    Lw = g.alloc_vertex_tree_array(rg.WCC, root)
    for L in (Lb, Lw):
        L.build_tree()
    gd = oracle.RefGraph(V, 60000, symmetric=False, dual=True, weighted=False)
    gd.add_edges(src[:n0], dst[:n0], None, True)
    gu = oracle.RefGraph(V, 60000, symmetric=True, dual=False, weighted=False)
    gu.add_edges(src[:n0], dst[:n0], None, False)
    live = [(int(a), int(b)) for a, b in zip(src[:n0], dst[:n0])]
    nxt = n0
    n_safe = n_unsafe = 0
    for rnd in range(8):
        ntx = int(rng.integers(20, 300))
        lens = rng.integers(1, 17, ntx).astype(np.uint32)
        # one update per edge key and round: inside one batch the returned
        # old counts of a key's updates are a multiset, not an order
        ops, used = [], set()
        while len(ops) < int(lens.sum()):
            if rng.random() < 0.5 and live:
                j = int(rng.integers(0, len(live)))
                live[j], live[-1] = live[-1], live[j]
                a, b = live.pop()
                if (a, b) in used:
                    live.insert(0, (a, b))
                    continue
                ops.append((a, b, -1))
            else:
                a, b = int(src[nxt % 30000]), int(dst[nxt % 30000])
                if rng.random() < 0.3:
                    a, b = int(rng.integers(0, V)), int(rng.integers(0, V))
                nxt += 1
                if (a, b) in used:
                    continue
                ops.append((a, b, 1))
            used.add((a, b))
        # cpu_ref starts from the device's tree
        vb, pb = Lb.read()
        vw, pw = Lw.read()
        Rb = oracle.RefLabels(gd, rg.BFS, root)
        Rb.assign(vb, pb)
        Rw = oracle.RefLabels(gu, rg.WCC, root)
        Rw.assign(vw, pw)
        need = [Rb.need_add(a, b) or Rw.need_add(a, b, 0, False) if k > 0 else
                Rb.need_del(a, b) or Rw.need_del(a, b, 0, False) for a, b, k in ops]
        v0 = g.version()
        a_s = np.array([o[0] for o in ops], np.uint64)
        a_d = np.array([o[1] for o in ops], np.uint64)
        kind = np.array([o[2] for o in ops], np.int8)
        unsafe, ver = g.txn_apply([Lb, Lw], a_s, a_d, np.zeros(len(ops), np.uint64), kind, lens, True)
        assert g.version() == v0 + 1
        # cpu_ref, batch order: inserts then deletes
        old = [0] * len(ops)
        for i, (a, b, k) in enumerate(ops):
            if k > 0:
                old[i] = gd.add_edge(a, b)
                gu.add_edge(a, b, 0, False)
        for i, (a, b, k) in enumerate(ops):
            if k < 0:
                old[i] = gd.del_edge(a, b)
                gu.del_edge(a, b, 0, False)
        ins = [i for i, o in enumerate(ops) if o[2] > 0 and old[i] == 0]
        dels = [i for i, o in enumerate(ops) if o[2] < 0 and old[i] == 1]
        if ins:
            Rb.inc_add_batch(a_s[ins], a_d[ins], None, True)
            Rw.inc_add_batch(a_s[ins], a_d[ins], None, False)
        if dels:
            Rb.inc_del_batch(a_s[dels], a_d[dels], None, True)
            Rw.inc_del_batch(a_s[dels], a_d[dels], None, False)
        assert np.array_equal(Lb.read()[0], Rb.read()[0]), rnd
        assert np.array_equal(Lw.read()[0], Rw.read()[0]), rnd
        ls, ld, lw = gd.live_edges()
        assert np.array_equal(Lb.read()[0], oracle.static_values(rg.BFS, V, ls, ld, lw, root))
        # expected flags: the update reached the repair and is unsafe in a tree
        cand = []
        for i, (a, b, k) in enumerate(ops):
            if k > 0:
                cand.append(old[i] == 0 and gd.edge_num(a, b) > 0)
            else:
                cand.append(old[i] == 1)
        k0 = 0
        for t, m in enumerate(lens):
            want = any(cand[k0 + j] and need[k0 + j] for j in range(int(m)))
            assert bool(unsafe[t]) == want, (rnd, t)
            assert int(ver[t]) == (v0 + 1 if want else v0), (rnd, t)
            k0 += int(m)
        n_safe += int((~unsafe).sum())
        n_unsafe += int(unsafe.sum())
        for i, (a, b, k) in enumerate(ops):
            if k > 0:
                live.append((a, b))
    assert n_safe > 0 and n_unsafe > 0
