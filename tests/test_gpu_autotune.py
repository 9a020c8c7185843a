"""The batch-engine autotuner (rg_autotune): forward batches of the stream
window and their inverses under candidate team settings, per batch-size
bucket.  The graph and the labels' values must come back unchanged (degrees
and values equal to the closed form on the loaded window, valid trees), the
table must be filled for the sizes the stream holds, and later batches run
with the tuned table must stay exact.  Windows whose deleted edges are not
live are reported as an error.
"""
import numpy as np
import pytest

import oracle
import risgraph_amd as rg

pytestmark = pytest.mark.gpu


def _setup(seed):
    V, E = 4096, 1 << 16
    s = rg.Stream("rmat", 12, E, seed=seed, wseed=seed + 1)
    src, dst, w = s.read()
    w = w.astype(np.uint64)
    n0 = E // 2
    g = rg.Graph(V, E, symmetric=False, dual=True, weighted=True)
    g.load_stream(s, 0, n0, True)
    root = s.max_degree_vertex(n0, V)
    Ls = [g.alloc_vertex_tree_array(a, root) for a in (rg.BFS, rg.SSSP, rg.SSWP, rg.WCC)]
    for L in Ls:
        L.build_tree()
    return V, E, s, src, dst, w, n0, g, root, Ls


def _check(Ls, V, src, dst, w, lo, hi, root):
    idx = np.arange(lo, hi) % len(src)
    ls, ld, lw = src[idx], dst[idx], w[idx]
    for L in Ls:
        vals, par = L.read()
        assert np.array_equal(vals, oracle.static_values(L.algo, V, ls, ld, lw, root)), L.algo
        assert oracle.check_tree(L.algo, V, ls, ld, lw, root, vals, par, undirected=L.algo == rg.WCC) == 0


def test_autotune_restores_the_graph_and_tunes_later_batches():
    V, E, s, src, dst, w, n0, g, root, Ls = _setup(91)
    deg = g.degrees()
    vals = [L.read()[0] for L in Ls]
    v0 = g.version()
    res = g.autotune(Ls, s, n0, 0, target_ms=4.0, pairs=2)
    # a stream of 65,536 edges holds the 16,384- and 65,536-op batches
    assert [b["ops"] for b in res["buckets"]] == [16384, 65536], res
    for b in res["buckets"]:
        assert b["cap_x4"] in (2, 4, 8, 16) and b["best_ms"] <= b["default_ms"] + 1e-9, b
        assert all(x >= 1 for x in b["weight"]), b
    assert res["batch_cap"] >= 1024 and res["batches"] == g.version() - v0
    d2 = g.degrees()
    assert np.array_equal(deg[0], d2[0]) and np.array_equal(deg[1], d2[1])
    for L, v in zip(Ls, vals):
        assert np.array_equal(L.read()[0], v), L.algo
    _check(Ls, V, src, dst, w, 0, n0, root)
    # batches of both tuned buckets afterwards
    pos = 0
    for half in (3000, 20000, 5000, 1):
        g.apply_stream(Ls, s, n0 + pos, half, pos, half, True)
        pos += half
        _check(Ls, V, src, dst, w, pos, n0 + pos, root)


def test_autotune_rejects_windows_that_do_not_return():
    V, E, s, src, dst, w, n0, g, root, Ls = _setup(93)
    # the deleted window [n0 + 20000, ...) was never loaded, nor inserted by
    # the 8,192-edge insert window [n0, n0 + 8192)
    with pytest.raises(rg.RgError):
        g.autotune(Ls, s, n0, n0 + 20000, pairs=1)
