"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

* static closed forms against hand-derived golden vectors and scipy's
  textbook routines (BFS, Dijkstra, connected components), SSWP against the
  threshold characterisation and brute-force path enumeration;
* the tree checker against planted faults;
* the cpu_ref store against the brute-force multiset model and golden ops;
* the cpu_ref incremental engine against the static closed form after every
  update of seeded sliding-window streams (all four algorithms, single and
  batched), with tree validity each time.
"""
import itertools
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse import csgraph

import oracle
from oracle import BFS, SSSP, SSWP, WCC
from oracle.model import MultigraphModel
import streams

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")
MAXL_BFS, MAXL_W = 65536, 134217728


def _golden():
    with open(GOLDEN) as f:
        return json.load(f)


@pytest.mark.parametrize("g", _golden()["graphs"], ids=lambda g: g["name"])
def test_static_known_answers(g):
    e = np.array(g["edges"], dtype=np.uint64)
    for name, algo in oracle.ALGOS.items():
        if name not in g:
            continue
        got = oracle.static_values(algo, g["V"], e[:, 0], e[:, 1], e[:, 2], g["root"], g.get("init"))
        assert got.tolist() == g[name], (g["name"], name)


def _random_graph(V, m, seed, wmax=20):
    rng = np.random.default_rng(seed)
    s = rng.integers(0, V, m).astype(np.uint64)
    d = rng.integers(0, V, m).astype(np.uint64)
    w = rng.integers(1, wmax + 1, m).astype(np.uint64)
    return s, d, w


def _min_weight_csr(V, s, d, w):
    """Directed CSR keeping the minimum weight of parallel edges (scipy sums duplicates)."""
    best = {}
    for a, b, c in zip(s.tolist(), d.tolist(), w.tolist()):
        if a == b:
            continue
        best[(a, b)] = min(best.get((a, b), c), c)
    if not best:
        return sp.csr_matrix((V, V))
    k = np.array(list(best.keys()))
    return sp.csr_matrix((np.array(list(best.values()), dtype=np.float64), (k[:, 0], k[:, 1])), shape=(V, V))


@pytest.mark.parametrize("seed", range(6))
def test_static_vs_scipy(seed):
    V, m = 300, 900 + 150 * seed
    s, d, w = _random_graph(V, m, seed)
    root = int(s[0])
    A = _min_weight_csr(V, s, d, w)
    # BFS levels: unweighted shortest paths
    U = A.copy()
    U.data[:] = 1.0
    lev = csgraph.shortest_path(U, method="D", directed=True, indices=root, unweighted=True)
    exp = np.where(np.isinf(lev), MAXL_BFS, lev).astype(np.uint64)
    assert np.array_equal(oracle.static_values(BFS, V, s, d, w, root), exp)
    # SSSP: Dijkstra
    dist = csgraph.dijkstra(A, directed=True, indices=root)
    exp = np.where(np.isinf(dist), MAXL_W, dist).astype(np.uint64)
    assert np.array_equal(oracle.static_values(SSSP, V, s, d, w, root), exp)
    # WCC: min vertex id per undirected component
    n, lab = csgraph.connected_components(sp.csr_matrix((np.ones(m), (s.astype(np.int64), d.astype(np.int64))),
                                                        shape=(V, V)), directed=False)
    first = np.full(n, V)
    for v in range(V):
        first[lab[v]] = min(first[lab[v]], v)
    assert np.array_equal(oracle.static_values(WCC, V, s, d), first[lab].astype(np.uint64))


@pytest.mark.parametrize("seed", range(4))
def test_sswp_threshold_characterisation(seed):
    """width(v) >= t  <=>  v reachable from root over edges with weight >= t."""
    V, m = 200, 700
    s, d, w = _random_graph(V, m, 100 + seed, wmax=50)
    root = int(s[0])
    width = oracle.static_values(SSWP, V, s, d, w, root)
    assert width[root] == MAXL_W
    for t in sorted(set(w.tolist())):
        keep = w >= t
        A = sp.csr_matrix((np.ones(int(keep.sum())), (s[keep].astype(np.int64), d[keep].astype(np.int64))),
                          shape=(V, V))
        reach = np.zeros(V, bool)
        reach[csgraph.breadth_first_order(A, root, directed=True, return_predecessors=False)] = True
        assert np.array_equal(width >= t, reach), t
    # unreached vertices have width 0, never a positive value
    assert np.all((width > 0) == np.isin(np.arange(V), csgraph.breadth_first_order(
        sp.csr_matrix((np.ones(m), (s.astype(np.int64), d.astype(np.int64))), shape=(V, V)), root,
        return_predecessors=False)))


def _brute_paths(V, edges, root):
    """All simple paths from root (tiny graphs): BFS/SSSP minima and SSWP maxima."""
    adj = {}
    for a, b, c in edges:
        adj.setdefault(a, []).append((b, c))
    best = {"bfs": [MAXL_BFS] * V, "sssp": [MAXL_W] * V, "sswp": [0] * V}
    best["bfs"][root] = 0
    best["sssp"][root] = 0
    best["sswp"][root] = MAXL_W

    def rec(u, seen, hops, dist, width):
        for v, c in adj.get(u, []):
            if v in seen:
                continue
            h, dd, ww = hops + 1, dist + c, min(width, c)
            best["bfs"][v] = min(best["bfs"][v], h)
            best["sssp"][v] = min(best["sssp"][v], dd)
            best["sswp"][v] = max(best["sswp"][v], ww)
            rec(v, seen | {v}, h, dd, ww)

    rec(root, {root}, 0, 0, MAXL_W)
    return best


@pytest.mark.parametrize("seed", range(10))
def test_static_vs_brute_force(seed):
    V, m = 7, 14
    s, d, w = _random_graph(V, m, 500 + seed, wmax=9)
    root = int(s[0])
    exp = _brute_paths(V, list(zip(s.tolist(), d.tolist(), w.tolist())), root)
    for name in ("bfs", "sssp", "sswp"):
        got = oracle.static_values(oracle.ALGOS[name], V, s, d, w, root)
        assert got.tolist() == exp[name], name


def test_bfs_sentinel_caps_long_paths():
    """s+1 < 65536 never holds at depth 65535, so deeper vertices stay 65536."""
    V = 66000
    s = np.arange(V - 1, dtype=np.uint64)
    d = s + np.uint64(1)
    got = oracle.static_values(BFS, V, s, d, None, 0)
    exp = np.minimum(np.arange(V), MAXL_BFS).astype(np.uint64)
    assert np.array_equal(got, exp)
    g = oracle.RefGraph(V, V, symmetric=False, dual=True)
    g.add_edges(s, d)
    L = oracle.RefLabels(g, BFS, 0)
    L.build()
    assert np.array_equal(L.read()[0], exp)


def test_tree_checker_catches_planted_faults():
    g = [x for x in _golden()["graphs"] if x["name"] == "diamond"][0]
    e = np.array(g["edges"], dtype=np.uint64)
    s, d, w = e[:, 0], e[:, 1], e[:, 2]
    V = g["V"]
    vals = np.array(g["sssp"], dtype=np.uint64)
    par = np.array([V, 0, 0, 1, 3], dtype=np.uint64)     # 3 via 1 (2+4) is valid
    assert oracle.check_tree(SSSP, V, s, d, w, 0, vals, par) == 0
    par2 = par.copy(); par2[3] = 2                         # 5+1 == 6 also valid (tie)
    assert oracle.check_tree(SSSP, V, s, d, w, 0, vals, par2) == 0
    bad = par.copy(); bad[4] = 1                           # no edge 1->4
    assert oracle.check_tree(SSSP, V, s, d, w, 0, vals, bad) == 5
    bad = par.copy(); bad[0] = 1                           # root must have no parent
    assert oracle.check_tree(SSSP, V, s, d, w, 0, vals, bad) == 1
    badv = vals.copy(); badv[4] = 12                       # value inconsistent with edge
    assert oracle.check_tree(SSSP, V, s, d, w, 0, badv, par) == 5
    # cycle with consistent values: WCC labels, 0<->1 pointing at each other
    cs = np.array([0, 1], np.uint64); cd = np.array([1, 0], np.uint64)
    cv = np.array([5, 5, 2], np.uint64)
    init = np.array([7, 8, 2], np.uint64)
    cp = np.array([1, 0, 3], np.uint64)
    assert oracle.check_tree(WCC, 3, cs, cd, None, 0, cv, cp, init=init) != 0


@pytest.mark.parametrize("case", _golden()["store_ops"], ids=lambda c: c["name"])
def test_store_golden_ops(case):
    g = oracle.RefGraph(case["V"], 16, symmetric=case["symmetric"], dual=True, weighted=case["weighted"])
    m = MultigraphModel(case["V"], symmetric=case["symmetric"], weighted=case["weighted"])
    for op, a, b, w, directed, expect in case["ops"]:
        f = g.add_edge if op == "add" else g.del_edge
        fm = m.add_edge if op == "add" else m.del_edge
        assert f(a, b, w, directed) == expect
        assert fm(a, b, w, directed) == expect
    for a, b, w, c in case["final_counts"]:
        assert g.edge_num(a, b, w) == c
        assert m.count(a, b, w) == c


@pytest.mark.parametrize("symmetric,weighted", [(False, False), (False, True), (True, False)])
def test_store_matches_model_with_hubs(symmetric, weighted):
    """Random add/del mix incl. a hub list above the 512-entry index threshold."""
    rng = np.random.default_rng(7)
    V = 1500
    g = oracle.RefGraph(V, 0, symmetric=symmetric, dual=True, weighted=weighted)
    m = MultigraphModel(V, symmetric=symmetric, weighted=weighted)
    for k in range(6000):
        a = 0 if rng.random() < 0.5 else int(rng.integers(0, V))
        b = int(rng.integers(0, V))
        w = int(rng.integers(1, 4))
        directed = True if not symmetric else bool(rng.random() < 0.2)
        if rng.random() < 0.65:
            assert g.add_edge(a, b, w, directed) == m.add_edge(a, b, w, directed), k
        else:
            assert g.del_edge(a, b, w, directed) == m.del_edge(a, b, w, directed), k
    assert g.degree(0) == m.degree(0)
    assert g.degree(0) > 512
    s1, d1, w1 = g.live_edges()
    s2, d2, w2 = m.live_edges()
    assert sorted(zip(s1.tolist(), d1.tolist(), w1.tolist())) == sorted(zip(s2.tolist(), d2.tolist(), w2.tolist()))


@pytest.mark.parametrize("symmetric,weighted,directed", [(False, False, True), (False, True, True),
                                                          (True, False, False), (False, True, False)])
def test_bulk_load_equals_sequential_adds(symmetric, weighted, directed):
    V, E = 700, 9000
    s, d = streams.rmat(10, E, 5)
    s, d = s % V, d % V
    w = streams.weights(E, 6) % np.uint64(3) + np.uint64(1)
    s[:800] = 0                                   # a hub above the 512-entry index threshold
    a = oracle.RefGraph(V, E, symmetric, True, weighted)
    b = oracle.RefGraph(V, E, symmetric, True, weighted)
    a.add_edges(s, d, w, directed)
    b.bulk_load(s, d, w, directed)
    la, lb = a.live_edges(), b.live_edges()
    ka = sorted(zip(*[x.tolist() for x in la]))
    kb = sorted(zip(*[x.tolist() for x in lb]))
    assert ka == kb
    for v in range(V):
        assert a.degree(v) == b.degree(v)
        assert a.degree(v, True) == b.degree(v, True)
    for i in range(0, E, 37):
        assert a.edge_num(int(s[i]), int(d[i]), int(w[i])) == b.edge_num(int(s[i]), int(d[i]), int(w[i]))
    with pytest.raises(ValueError):
        b.bulk_load(s, d, w, directed)
    # incremental updates behave the same afterwards
    for i in range(300):
        x, y, z = int(d[i]), int(s[i]), int(w[i])
        assert a.add_edge(x, y, z, directed) == b.add_edge(x, y, z, directed)
        assert a.del_edge(int(s[i]), int(d[i]), z, directed) == b.del_edge(int(s[i]), int(d[i]), z, directed)


def test_labels_assign_matches_build():
    V, E = 1024, 12000
    s, d = streams.rmat(10, E, 8)
    w = streams.weights(E, 9)
    for algo in (BFS, SSSP, SSWP, WCC):
        sym = algo == WCC
        g = oracle.RefGraph(V, E, sym, not sym, algo in (SSSP, SSWP))
        g.bulk_load(s, d, w, not sym)
        root = streams.max_out_degree_root(s, V)
        L1 = oracle.RefLabels(g, algo, root)
        L1.build()
        v1, p1, w1 = L1.read()
        L2 = oracle.RefLabels(g, algo, root)
        L2.assign(v1, p1)
        v2, p2, _ = L2.read()
        assert np.array_equal(v1, v2) and np.array_equal(p1, p2)
        # a parent with no live edge into the vertex must be rejected
        ls, ld, _ = g.live_edges()
        edges = set(zip(ls.tolist(), ld.tolist()))
        v = int(np.nonzero(p1 < V)[0][0])
        p = next(x for x in range(V) if (x, v) not in edges and (v, x) not in edges)
        bad = p1.copy()
        bad[v] = p
        L3 = oracle.RefLabels(g, algo, root)
        with pytest.raises(ValueError):
            L3.assign(v1, bad)


def _window_stream(kind, V, E, load_frac, seed=1, wseed=2):
    if kind == "rmat":
        s, d = streams.rmat(int(np.log2(V)), E, seed)
    else:
        s, d = streams.erdos_renyi(V, E, seed)
    w = streams.weights(E, wseed)
    return s, d, w, int(E * load_frac)


CASES = [(BFS, "rmat", False), (SSSP, "rmat", True), (SSWP, "rmat", True), (WCC, "rmat", False),
         (WCC, "er", False), (BFS, "er", False)]


@pytest.mark.parametrize("algo,kind,weighted", CASES)
def test_cpu_ref_plain_stream_matches_static(algo, kind, weighted):
    V, E = 512, 4096
    s, d, w, n0 = _window_stream(kind, V, E, 0.9)
    sym = algo == WCC
    directed = not sym
    g = oracle.RefGraph(V, E, symmetric=sym, dual=not sym, weighted=weighted)
    m = MultigraphModel(V, symmetric=sym, weighted=weighted)
    g.add_edges(s[:n0], d[:n0], w[:n0], directed)
    for i in range(n0):
        m.add_edge(int(s[i]), int(d[i]), int(w[i]), directed)
    root = streams.max_out_degree_root(s[:n0], V)
    L = oracle.RefLabels(g, algo, root)
    L.build()
    for k in range(E - n0):
        i, j = n0 + k, k
        ww = int(w[i]) if weighted else 0
        old = g.add_edge(int(s[i]), int(d[i]), ww, directed)
        assert old == m.add_edge(int(s[i]), int(d[i]), ww, directed)
        if old == 0:
            L.inc_add(int(s[i]), int(d[i]), ww, directed)
        wj = int(w[j]) if weighted else 0
        old = g.del_edge(int(s[j]), int(d[j]), wj, directed)
        assert old == m.del_edge(int(s[j]), int(d[j]), wj, directed)
        if old == 1:
            L.inc_del(int(s[j]), int(d[j]), wj, directed)
        ls, ld, lw = m.live_edges()
        exp = oracle.static_values(algo, V, ls, ld, lw, root)
        vals, par, _ = L.read()
        assert np.array_equal(vals, exp), k
        assert oracle.check_tree(algo, V, ls, ld, lw, root, vals, par) == 0, k


@pytest.mark.parametrize("algo,kind,weighted", CASES)
@pytest.mark.parametrize("batch", [7, 64])
def test_cpu_ref_batched_stream_matches_static(algo, kind, weighted, batch):
    V, E = 512, 4096
    s, d, w, n0 = _window_stream(kind, V, E, 0.8, seed=3)
    sym = algo == WCC
    directed = not sym
    g = oracle.RefGraph(V, E, symmetric=sym, dual=not sym, weighted=weighted)
    g.add_edges(s[:n0], d[:n0], w[:n0], directed)
    root = streams.max_out_degree_root(s[:n0], V)
    L = oracle.RefLabels(g, algo, root)
    L.build()
    wz = w if weighted else np.zeros_like(w)
    for b0 in range(n0, E, batch):
        b1 = min(b0 + batch, E)
        idx = np.arange(b0, b1)
        olds = [g.add_edge(int(s[i]), int(d[i]), int(wz[i]), directed) for i in idx]
        new = idx[np.array(olds) == 0]
        if len(new):
            L.inc_add_batch(s[new], d[new], wz[new], directed)
        didx = idx - n0
        olds = [g.del_edge(int(s[i]), int(d[i]), int(wz[i]), directed) for i in didx]
        gone = didx[np.array(olds) == 1]
        if len(gone):
            L.inc_del_batch(s[gone], d[gone], wz[gone], directed)
        ls, ld, lw = g.live_edges()
        exp = oracle.static_values(algo, V, ls, ld, lw, root)
        vals, par, _ = L.read()
        assert np.array_equal(vals, exp), b0
        assert oracle.check_tree(algo, V, ls, ld, lw, root, vals, par) == 0, b0


def test_cpu_ref_recompute_fallback_keeps_values():
    """The deletion fallback (graph.hpp:1252) is a full build: values unchanged."""
    V = 5000
    # a long chain 0->1->...->V-1: deleting 0->1 invalidates > 0.05V but the
    # 2,000,000 floor keeps the reference from recomputing; check values anyway
    s = np.arange(V - 1, dtype=np.uint64)
    d = s + np.uint64(1)
    g = oracle.RefGraph(V, V, False, True, False)
    g.add_edges(s, d)
    g.add_edge(0, V // 2)
    L = oracle.RefLabels(g, BFS, 0)
    L.build()
    assert g.del_edge(0, 1) == 1
    L.inc_del(0, 1)
    ls, ld, lw = g.live_edges()
    assert np.array_equal(L.read()[0], oracle.static_values(BFS, V, ls, ld, lw, 0))


# ---------------------------------------------------------------------------
# the safe/unsafe classifier (need_update_tree_add / need_update_tree_del,
# graph.hpp:987-1000, 1220-1233) against what the mathematics fixes:
#  * an add is unsafe exactly when the fixpoint of G + e differs from that of
#    G (labels are the unique fixpoint of a monotone relaxation, so an edge
#    that does not strictly improve its target leaves every value alone);
#  * a delete of an edge held once is unsafe exactly when the edge is some
#    vertex's parent edge, i.e. when the tree stops being a set of live
#    edges on G - e; a safe delete leaves a valid tree with the same values.
# ---------------------------------------------------------------------------
def _classifier_case(algo, directed, seed):
    V = 48
    m = 40 if algo == WCC else 150   # sparse enough for several components
    rng = np.random.default_rng(seed)
    weighted = algo in (SSSP, SSWP)
    s = rng.integers(0, V, m).astype(np.uint64)
    d = rng.integers(0, V, m).astype(np.uint64)
    w = rng.integers(1, 6, m).astype(np.uint64) if weighted else np.zeros(m, np.uint64)
    sym = not directed
    g = oracle.RefGraph(V, m, symmetric=sym, dual=not sym, weighted=weighted)
    g.add_edges(s, d, w if weighted else None, directed)
    root = int(s[0])
    L = oracle.RefLabels(g, algo, root)
    L.build()
    return V, g, L, root, rng, weighted


def _add_changes_fixpoint(algo, V, g, root, a, b, c, directed):
    ls, ld, lw = g.live_edges()
    before = oracle.static_values(algo, V, ls, ld, lw, root)
    ns, nd, nw = [a], [b], [c]
    if not directed:
        ns, nd, nw = [a, b], [b, a], [c, c]
    after = oracle.static_values(algo, V, np.concatenate([ls, np.array(ns, np.uint64)]),
                                 np.concatenate([ld, np.array(nd, np.uint64)]),
                                 np.concatenate([lw, np.array(nw, np.uint64)]), root)
    return not np.array_equal(before, after)


def _parent_edges_live_without(g, L, V, a, b, c, directed):
    """Is every vertex's parent edge (parent vertex, parent weight) still a
    live edge once the single copy of (a -> b, c) is gone?"""
    ls, ld, lw = g.live_edges()
    live = {}
    for x, y, z in zip(ls.tolist(), ld.tolist(), lw.tolist()):
        live[(x, y, z)] = live.get((x, y, z), 0) + 1
    for key in ([(a, b, c)] if directed else [(a, b, c), (b, a, c)]):
        live[key] -= 1
    _, par, pw = L.read()
    for v in range(V):
        p = int(par[v])
        if p == V:
            continue
        if live.get((p, v, int(pw[v])), 0) <= 0:
            return False
    return True


@pytest.mark.parametrize("directed", [True, False], ids=["directed", "undirected"])
@pytest.mark.parametrize("algo", [BFS, SSSP, SSWP, WCC], ids=["bfs", "sssp", "sswp", "wcc"])
def test_classifier_add_iff_fixpoint_changes(algo, directed):
    if algo == WCC and directed:
        pytest.skip("WCC runs on symmetric graphs only (wcc_inc.cpp:56)")
    for seed in range(3):
        V, g, L, root, rng, weighted = _classifier_case(algo, directed, 100 + seed)
        n_unsafe = 0
        for _ in range(120):
            a, b = int(rng.integers(0, V)), int(rng.integers(0, V))
            c = int(rng.integers(1, 6)) if weighted else 0
            got = L.need_add(a, b, c, directed)
            assert got == _add_changes_fixpoint(algo, V, g, root, a, b, c, directed), (seed, a, b, c)
            n_unsafe += got
        assert 0 < n_unsafe < 120, "the sample must hold both safe and unsafe adds"


@pytest.mark.parametrize("directed", [True, False], ids=["directed", "undirected"])
@pytest.mark.parametrize("algo", [BFS, SSSP, SSWP, WCC], ids=["bfs", "sssp", "sswp", "wcc"])
def test_classifier_del_iff_parent_edge(algo, directed):
    if algo == WCC and directed:
        pytest.skip("WCC runs on symmetric graphs only (wcc_inc.cpp:56)")
    for seed in range(3):
        V, g, L, root, rng, weighted = _classifier_case(algo, directed, 200 + seed)
        ls, ld, lw = g.live_edges()
        vals = L.read()[0]
        n_unsafe = 0
        seen = set()
        for x, y, z in zip(ls.tolist(), ld.tolist(), lw.tolist()):
            key = (x, y, z) if directed else (min(x, y), max(x, y), z)
            if key in seen or g.edge_num(x, y, z) != 1:
                continue
            seen.add(key)
            got = L.need_del(x, y, z, directed)
            assert got == (not _parent_edges_live_without(g, L, V, x, y, z, directed)), (seed, x, y, z)
            n_unsafe += got
            if not got:
                # a safe delete keeps the same values and a valid tree
                keep = np.ones(len(ls), bool)
                for i in range(len(ls)):
                    if (ls[i], ld[i], lw[i]) == (x, y, z) or (not directed and (ls[i], ld[i], lw[i]) == (y, x, z)):
                        keep[i] = False
                exp = oracle.static_values(algo, V, ls[keep], ld[keep], lw[keep], root)
                assert np.array_equal(vals, exp), (seed, x, y, z)
                par = L.read()[1]
                assert oracle.check_tree(algo, V, ls[keep], ld[keep], lw[keep], root, vals, par,
                                         undirected=not directed or algo == WCC) == 0
        assert n_unsafe > 0 and len(seen) > n_unsafe


def test_classifier_checks_catch_planted_mistakes():
    """The two properties above reject plausible classifier mistakes: an
    undirected add or delete that looks at one direction only, and an SSWP
    add test with the comparison the wrong way round."""
    # one-direction undirected add: misses an edge that improves src from dst
    caught = False
    for seed in range(3):
        V, g, L, root, rng, _ = _classifier_case(BFS, False, 300 + seed)
        vals = L.read()[0]
        for _ in range(200):
            a, b = int(rng.integers(0, V)), int(rng.integers(0, V))
            one_way = vals[a] + 1 < vals[b] and vals[a] < 65536
            if one_way != _add_changes_fixpoint(BFS, V, g, root, a, b, 0, False):
                caught = True
    assert caught
    # SSWP add with the comparison reversed
    V, g, L, root, rng, _ = _classifier_case(SSWP, True, 310)
    vals = L.read()[0]
    caught = False
    for _ in range(200):
        a, b, c = int(rng.integers(0, V)), int(rng.integers(0, V)), int(rng.integers(1, 6))
        wrong = min(int(vals[a]), c) < int(vals[b])
        if wrong != _add_changes_fixpoint(SSWP, V, g, root, a, b, c, True):
            caught = True
    assert caught
    # one-direction undirected delete: misses a tree edge held as (b -> a)
    V, g, L, root, rng, _ = _classifier_case(WCC, False, 320)
    _, par, _ = L.read()
    ls, ld, lw = g.live_edges()
    caught = False
    for x, y, z in zip(ls.tolist(), ld.tolist(), lw.tolist()):
        if g.edge_num(x, y, z) != 1:
            continue
        one_way = int(par[y]) == x
        if one_way != (not _parent_edges_live_without(g, L, V, x, y, z, False)):
            caught = True
    assert caught


def test_rt_serving_loop_matches_static():
    """or_serve_rt (the rt driver's safe / unsafe rounds on host threads)
    ends at the same values as the closed form on the final window, for all
    four trees, with every update served and its latency recorded."""
    V, E = 1024, 12000
    s, d = streams.rmat(10, E, 3)
    w = streams.weights(E, 4)
    n0 = 9000
    gd = oracle.RefGraph(V, E, False, True, True)
    gd.add_edges(s[:n0], d[:n0], w[:n0], True)
    gu = oracle.RefGraph(V, E, True, False, False)
    gu.add_edges(s[:n0], d[:n0], None, False)
    root = streams.max_out_degree_root(s[:n0], V)
    Ld = [oracle.RefLabels(gd, a, root) for a in (BFS, SSSP, SSWP)]
    Lu = oracle.RefLabels(gu, WCC, root)
    for L in Ld + [Lu]:
        L.build()
    n = 4000
    arrival = np.arange(n) / 2e5          # 200k updates/s offered
    st, lat = oracle.serve_rt(gd, Ld, gu, Lu, n, s[n0:], d[n0:], w[n0:], s, d, w, arrival)
    assert st["served"] == n and not st["overloaded"] and st["threads"] >= 1
    assert 0 < st["unsafe"] < n and np.all(np.isfinite(lat)) and np.all(lat >= 0)
    k = n // 2
    ls, ld, lw = s[k:n0 + k], d[k:n0 + k], w[k:n0 + k]
    for L in Ld + [Lu]:
        vals, par, _ = L.read()
        assert np.array_equal(vals, oracle.static_values(L.algo, V, ls, ld, lw, root)), L.algo
        assert oracle.check_tree(L.algo, V, ls, ld, lw, root, vals, par, undirected=L.algo == WCC) == 0
