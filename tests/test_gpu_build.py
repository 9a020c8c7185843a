"""Direction-optimizing builds (stream_edges, graph.hpp:323-394): with the
switch on, a build's push pass whose frontier's live out-edges reach both V
and a twentieth of the live edges runs dense, every vertex pulling its best
candidate over its in-edges.  Builds with the switch and push-only builds
(the default) of the same graph must give the same values, equal to the closed form, with valid
trees; the switch must actually run dense passes on these graphs, and the
push-only mode none.
"""
import numpy as np
import pytest

import oracle
import risgraph_amd as rg

pytestmark = pytest.mark.gpu


def _graph(kind, scale, ef, symmetric, seed):
    V, E = 1 << scale, ef << scale
    s = rg.Stream(kind, scale if kind == "rmat" else V, E, seed=seed, wseed=seed + 1)
    src, dst, w = s.read()
    w = w.astype(np.uint64)
    g = rg.Graph(V, E, symmetric=symmetric, dual=not symmetric, weighted=True)
    # a symmetric store takes undirected ops (both directions)
    g.load_stream(s, 0, E, not symmetric)
    root = s.max_degree_vertex(E, V)
    if symmetric:
        src, dst, w = np.concatenate([src, dst]), np.concatenate([dst, src]), np.concatenate([w, w])
    return g, V, src, dst, w, root


@pytest.mark.parametrize("kind,symmetric", [("rmat", False), ("rmat", True), ("er", True)],
                         ids=["rmat-directed", "rmat-symmetric", "er-symmetric"])
def test_dense_build_matches_push_build(kind, symmetric):
    g, V, src, dst, w, root = _graph(kind, 13, 16, symmetric, 81)
    for algo in (rg.BFS, rg.SSSP, rg.SSWP, rg.WCC):
        La = g.alloc_vertex_tree_array(algo, root)
        Lb = g.alloc_vertex_tree_array(algo, root)
        La.dense(1)
        Lb.dense(0)
        La.build_tree()
        Lb.build_tree()
        va, pa = La.read()
        vb, _ = Lb.read()
        exp = oracle.static_values(algo, V, src, dst, w, root)
        assert np.array_equal(va, vb), algo
        assert np.array_equal(va, exp), algo
        assert oracle.check_tree(algo, V, src, dst, w, root, va, pa) == 0, algo
        assert La.dense() > 0, algo
        assert Lb.dense() == 0, algo
        La.close()
        Lb.close()
