"""Captured batch pipelines (rg_graph_capture): a stream batch of more than
64 edge ops with trees to repair runs its window, store and split launches
as one hipGraph per batch-size bucket, replayed with the batch's size read
from a device descriptor.  The replays must give what the launches one by
one give: a captured graph and an eager twin run the same batches (sizes
that hit the same bucket again, sizes on bucket edges, a bucket's first
capture after eager batches) and must hold the same labels, parents and
degrees after every batch, equal to the closed form on the live window
with valid trees.  Trees freed and allocated again, and a second stream,
must not replay a pipeline captured for other buffers.
"""
import numpy as np
import pytest

import oracle
import risgraph_amd as rg

pytestmark = pytest.mark.gpu

ALGOS = (rg.BFS, rg.SSSP, rg.SSWP, rg.WCC)


def _graph(s, V, E, n0, root, capture):
    g = rg.Graph(V, E, symmetric=False, dual=True, weighted=True)
    g.capture(capture)
    g.load_stream(s, 0, n0, True)
    Ls = [g.alloc_vertex_tree_array(a, root) for a in ALGOS]
    for L in Ls:
        L.build_tree()
    return g, Ls


def _check(g, Ls, V, src, dst, w, lo, hi, root):
    """The live window is stream edges [lo, hi) (mod the stream length)."""
    idx = np.arange(lo, hi) % len(src)
    ls, ld, lw = src[idx], dst[idx], w[idx]
    out = []
    for L in Ls:
        vals, par = L.read()
        assert np.array_equal(vals, oracle.static_values(L.algo, V, ls, ld, lw, root)), L.algo
        assert oracle.check_tree(L.algo, V, ls, ld, lw, root, vals, par, undirected=L.algo == rg.WCC) == 0
        out.append((vals, par))
    return out


def test_captured_batches_match_eager_launches():
    V, E = 4096, 1 << 16
    s = rg.Stream("rmat", 12, E, seed=51, wseed=52)
    src, dst, w = s.read()
    w = w.astype(np.uint64)
    n0 = E // 2
    root = s.max_degree_vertex(n0, V)
    gc, Lc = _graph(s, V, E, n0, root, True)
    ge, Le = _graph(s, V, E, n0, root, False)
    pos = 0
    sizes = (33, 50, 64, 64, 100, 37, 250, 1000, 600, 64, 2048, 1500, 33, 2049, 5000, 60, 128, 7000)
    for b, half in enumerate(sizes):
        # a batch of `half` inserts and `half` deletes (2 * half edge ops)
        if b == 9:
            gc.capture(False)          # eager batches in between ...
        if b == 11:
            gc.capture(True)           # ... and replays of earlier buckets after them
        for g, Ls in ((gc, Lc), (ge, Le)):
            g.apply_stream(Ls, s, (n0 + pos) % E, half, pos % E, half, True)
        pos += half
        assert gc.version() == ge.version() == b + 1
        got = _check(gc, Lc, V, src, dst, w, pos, n0 + pos, root)
        for (vc, pc), L in zip(got, Le):
            ve, pe = L.read()
            assert np.array_equal(vc, ve), (b, L.algo)
        dc, de = gc.degrees(), ge.degrees()
        assert np.array_equal(dc[0], de[0]) and np.array_equal(dc[1], de[1]), b
    ic, ie = gc.capture_info(), ge.capture_info()
    # batches of more than 64 edge ops, less the two run eagerly
    assert ic["on"] and ic["replays"] == sum(2 * h > 64 for h in sizes) - 2, ic
    # buckets 128, 256, 512, 2048, 4096, 8192, 16384
    assert ic["pipelines"] == 7, ic
    assert not ie["on"] and ie["replays"] == 0 and ie["pipelines"] == 0
    # every block freed by these batches went back to a free list
    for g in (gc, ge):
        pi = g.pool_info()
        assert pi["dropped"] == [0, 0] and pi["dropped_entries"] == [0, 0] and pi["recycled"][0] > 0, pi


def test_captured_batches_after_new_trees_and_stream():
    V, E = 4096, 1 << 16
    s = rg.Stream("rmat", 12, E, seed=61, wseed=62)
    src, dst, w = s.read()
    w = w.astype(np.uint64)
    n0 = E // 2
    root = s.max_degree_vertex(n0, V)
    g, Ls = _graph(s, V, E, n0, root, True)
    pos = 0
    for half in (300, 200):
        g.apply_stream(Ls, s, n0 + pos, half, pos, half, True)
        pos += half
    _check(g, Ls, V, src, dst, w, pos, n0 + pos, root)
    # new trees in place of two old ones: the bucket of 512 must capture again
    for L in Ls[:2]:
        L.close()
    Ls = [g.alloc_vertex_tree_array(a, root) for a in (rg.SSSP, rg.BFS)] + Ls[2:]
    for L in Ls[:2]:
        L.build_tree()
    g.apply_stream(Ls, s, n0 + pos, 250, pos, 250, True)
    pos += 250
    _check(g, Ls, V, src, dst, w, pos, n0 + pos, root)
    # a second stream holding the same edges rotated: a batch from it inserts
    # stream edges [n0 + pos, n0 + pos + 240) and deletes [pos, pos + 240)
    rot = n0 + pos
    s2 = rg.Stream.from_edges(np.roll(src, -rot), np.roll(dst, -rot), np.roll(w, -rot))
    g.apply_stream(Ls, s2, 0, 240, (E - n0) % E, 240, True)
    pos += 240
    _check(g, Ls, V, src, dst, w, pos, n0 + pos, root)
    g.apply_stream(Ls, s, n0 + pos, 200, pos, 200, True)
    pos += 200
    _check(g, Ls, V, src, dst, w, pos, n0 + pos, root)
    info = g.capture_info()
    assert info["replays"] == 5 and 2 <= info["pipelines"] <= 4, info


@pytest.mark.parametrize("capture", [True, False])
def test_serving_with_and_without_capture(capture):
    """The serving loop's batches (sizes set by the arrival rate) through
    captured or eager pipelines end on the closed form of the final window."""
    V, E = 4096, 1 << 16
    s = rg.Stream("rmat", 12, E, seed=71, wseed=72)
    src, dst, w = s.read()
    w = w.astype(np.uint64)
    n0 = int(0.9 * E)
    root = s.max_degree_vertex(n0, V)
    g, Ls = _graph(s, V, E, n0, root, capture)
    n = 20000
    res, _ = g.serve(Ls, s, n0, 0, n, 8e6, batch_cap=64, cap_max=8192, target_ms=4.0, split=True)
    assert res["served"] == n and not res["overloaded"]
    assert (g.capture_info()["replays"] > 0) == capture
    _check(g, Ls, V, src, dst, w, n // 2, n0 + n // 2, root)
