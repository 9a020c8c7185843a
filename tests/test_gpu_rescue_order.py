"""Rescue choices carry no stream-position bias.

The reference's rescue scans an in-list from its front (graph.hpp:1330-1359).
List positions follow the order in which keys first entered the list, so a
front-first scan keeps adopting the oldest keys as tree edges; a sliding
window over a circular stream deletes those together when it comes round to
them, and every circle of the stream then brings a stretch of extra unsafe
updates (with front-first scans this setup measured 1.52x the median unsafe
count in the tenth of a circle after the window wraps, against 1.07x with
the rotated scans the library uses).  The test slides a window of batches
over 2.5 circles of an R-MAT scale-16 stream with all four trees, requires
the unsafe count per tenth of a circle to stay within 25% of the median from
the second circle on, and checks the final values against the closed form
and the trees for validity.
"""
import numpy as np
import pytest

import oracle
import risgraph_amd as rg

pytestmark = pytest.mark.gpu


def test_unsafe_share_is_flat_along_the_circular_stream():
    scale = 16
    V, E = 1 << scale, 16 << scale
    n0 = int(0.9 * E)
    B = 4096                      # insert + delete pairs per batch
    s = rg.Stream("rmat", scale, E, seed=1, wseed=2)
    g = rg.Graph(V, E, symmetric=False, dual=True, weighted=True)
    g.load_stream(s, 0, n0, True)
    root = s.max_degree_vertex(n0, V)
    algos = (rg.BFS, rg.SSSP, rg.SSWP, rg.WCC)
    Ls = [g.alloc_vertex_tree_array(a, root, trace_modified=False) for a in algos]
    for L in Ls:
        L.build_tree()
    pos, unsafe = 0, []
    while pos < 2.5 * E:
        g.apply_stream(Ls, s, (n0 + pos) % E, B, pos % E, B, True)
        unsafe.append(g.last_unsafe())
        pos += B
    unsafe = np.asarray(unsafe, np.float64)
    per = E // B                  # batches per circle
    w = per // 10
    later = unsafe[per:]
    windows = np.array([later[i:i + w].mean() for i in range(0, len(later) - w + 1, w)])
    med = float(np.median(later))
    assert med > 0
    assert windows.max() <= 1.25 * med, (windows.round(0).tolist(), med)

    # the window after pos pairs: stream edges [pos, pos + n0) modulo E
    src, dst, wt = s.read()
    idx = (np.arange(n0, dtype=np.int64) + pos) % E
    ls, ld, lw = src[idx], dst[idx], wt[idx].astype(np.uint64)
    for a, L in zip(algos, Ls):
        vals, par = L.read()
        assert np.array_equal(vals, oracle.static_values(a, V, ls, ld, lw, root)), a
        assert oracle.check_tree(a, V, ls, ld, lw, root, vals, par, undirected=a == rg.WCC) == 0, a
