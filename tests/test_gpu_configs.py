"""GPU parity on the shapes of the reference workloads (configs 2-4).

Each test runs the HIP path through the C ABI on a sliding window over a
seeded device-generated stream.  The live multigraph after a step is exactly
the window of stream edges [pos, n0 + pos), so the static closed form is
computed straight from the stream; cpu_ref (the oracle's restatement of the
reference engine) is run beside the device where its cost allows.

* config 4, scaled to 2^16 vertices / 2^20 edges: BFS and WCC on a
  symmetric Erdos-Renyi graph, batch sizes floor(2^(16 U)) (log-uniform in
  [1, 65536]), every batch checked against cpu_ref and the closed form;
* config 4 at full size (2^24 vertices, 2^28 edges, 2^27 loaded): the same
  batches, closed form and tree validity at the end;
* config 2 at full size (R-MAT scale 20, 2^24 edges, 90% loaded): plain
  SSSP and SSWP updates, cpu_ref after every update, closed form every
  PLAIN_CHECK updates;
* config 3 at full size (R-MAT scale 22, 2^26 edges, 80% loaded): SSSP and
  WCC batches of 1, 10, 100, 1k, 10k and 100k insert+delete pairs, closed
  form after every batch (bfs_inc_batch.cpp:136-186).
"""
import numpy as np
import pytest

import oracle
import risgraph_amd as rg
import streams

pytestmark = pytest.mark.gpu

C2_UPDATES = 2000      # plain updates per algorithm at scale 20
PLAIN_CHECK = 500      # closed-form check interval of the plain runs


def _log_uniform_sizes(n, seed):
    rng = np.random.default_rng(seed)
    return [max(1, int(np.floor(2.0 ** (16.0 * u)))) for u in rng.random(n)]


def _window_static(algo, V, src, dst, w, lo, hi, root, undirected):
    s, d = src[lo:hi], dst[lo:hi]
    ww = None if w is None else w[lo:hi]
    if undirected and algo != rg.WCC:
        s, d = np.concatenate([s, d]), np.concatenate([d, s])
        ww = None if ww is None else np.concatenate([ww, ww])
    return s, d, ww, oracle.static_values(algo, V, s, d, ww, root)


def _check_window(L, algo, V, src, dst, w, lo, hi, root, undirected, tag):
    s, d, ww, exp = _window_static(algo, V, src, dst, w, lo, hi, root, undirected)
    vals, par = L.read()
    bad = np.nonzero(vals != exp)[0]
    assert len(bad) == 0, f"{tag}: {len(bad)} vertices differ, first {bad[:5]}"
    # undirected windows are already doubled for BFS/SSSP/SSWP; WCC's checker
    # looks both ways by itself
    t = oracle.check_tree(algo, V, s, d, ww, root, vals, par, undirected=algo == rg.WCC)
    assert t == 0, f"{tag}: invalid tree at vertex {t - 1}"


@pytest.mark.parametrize("name", ["bfs", "wcc"])
def test_config4_scaled_undirected_er_log_uniform_batches(name):
    algo = rg.ALGOS[name]
    V, E = 1 << 16, 1 << 20
    s = rg.Stream("er", V, E, seed=1)
    src, dst, _ = s.read()
    n0 = E // 2
    g = rg.Graph(V, E, symmetric=True, dual=False, weighted=False)
    g.load_stream(s, 0, n0, False)
    root = s.max_degree_vertex(n0, V)
    L = g.alloc_vertex_tree_array(algo, root)
    L.build_tree()
    ref = oracle.RefGraph(V, E, symmetric=True, dual=False, weighted=False)
    ref.bulk_load(src[:n0], dst[:n0], None, False)
    RL = oracle.RefLabels(ref, algo, root)
    RL.build()
    assert np.array_equal(L.read()[0], RL.read()[0])
    pos = 0
    sizes = _log_uniform_sizes(14, 4)
    assert max(sizes) > 8192 and min(sizes) < 64
    for B in sizes:
        st = g.apply_stream([L], s, n0 + pos, B, pos, B, False)
        # reference batch driver order: all adds, then all dels (bfs_inc_batch.cpp:142-178)
        idx = np.arange(n0 + pos, n0 + pos + B)
        olds = np.array([ref.add_edge(int(src[i]), int(dst[i]), 0, False) for i in idx])
        new = idx[olds == 0]
        if len(new):
            RL.inc_add_batch(src[new], dst[new], None, False)
        didx = np.arange(pos, pos + B)
        olds = np.array([ref.del_edge(int(src[i]), int(dst[i]), 0, False) for i in didx])
        gone = didx[olds == 1]
        if len(gone):
            RL.inc_del_batch(src[gone], dst[gone], None, False)
        pos += B
        ref_vals = RL.read()[0]
        vals = L.read()[0]
        assert np.array_equal(vals, ref_vals), (B, st)
        _check_window(L, algo, V, src, dst, None, pos, n0 + pos, root, True, f"batch {B}")


@pytest.mark.slow
def test_config4_full_size_undirected_er():
    """2^24 vertices, 2^28 stream edges, 2^27 loaded, symmetric store; BFS and
    WCC maintained together over log-uniform batches; closed form and tree
    validity for both at the end."""
    V, E = 1 << 24, 1 << 28
    s = rg.Stream("er", V, E, seed=1)
    n0 = E // 2
    g = rg.Graph(V, E, symmetric=True, dual=False, weighted=False)
    g.load_stream(s, 0, n0, False)
    root = s.max_degree_vertex(n0, V)
    Ls = [g.alloc_vertex_tree_array(a, root, trace_modified=False) for a in (rg.BFS, rg.WCC)]
    for L in Ls:
        L.build_tree()
    pos = 0
    for B in _log_uniform_sizes(8, 5):
        g.apply_stream(Ls, s, n0 + pos, B, pos, B, False)
        pos += B
    src, dst, _ = s.read(0, n0 + pos)
    for L in Ls:
        _check_window(L, L.algo, V, src, dst, None, pos, n0 + pos, root, True, f"algo {L.algo}")


@pytest.mark.slow
@pytest.mark.parametrize("name", ["sssp", "sswp"])
def test_config2_full_size_plain_stream(name):
    """R-MAT scale 20 (2^20 vertices, 2^24 edges, weights 1..1000), 90% loaded,
    root = max out-degree; plain sliding-window updates (sssp_inc.cpp /
    sswp_inc.cpp main loops), values equal to cpu_ref after every update."""
    algo = rg.ALGOS[name]
    V, E = 1 << 20, 1 << 24
    s = rg.Stream("rmat", 20, E, seed=1, wseed=2)
    src, dst, w = s.read()
    w = w.astype(np.uint64)
    n0 = int(0.9 * E)
    g = rg.Graph(V, E, symmetric=False, dual=True, weighted=True)
    g.load_stream(s, 0, n0, True)
    root = s.max_degree_vertex(n0, V)
    assert root == streams.max_out_degree_root(src[:n0], V)
    L = g.alloc_vertex_tree_array(algo, root, trace_modified=False)
    L.build_tree()
    ref = oracle.RefGraph(V, E, symmetric=False, dual=True, weighted=True)
    ref.bulk_load(src[:n0], dst[:n0], w[:n0], True)
    RL = oracle.RefLabels(ref, algo, root)
    RL.build()
    assert np.array_equal(L.read()[0], RL.read()[0])
    for k in range(C2_UPDATES):
        i, j = n0 + k, k
        g.stream_plain([L], s, i, j, 1, True, latency=False)
        if ref.add_edge(int(src[i]), int(dst[i]), int(w[i])) == 0:
            RL.inc_add(int(src[i]), int(dst[i]), int(w[i]))
        if ref.del_edge(int(src[j]), int(dst[j]), int(w[j])) == 1:
            RL.inc_del(int(src[j]), int(dst[j]), int(w[j]))
        vals = L.read()[0]
        assert np.array_equal(vals, RL.read()[0]), k
        if k % PLAIN_CHECK == PLAIN_CHECK - 1:
            _check_window(L, algo, V, src, dst, w, k + 1, n0 + k + 1, root, False, f"update {k}")


@pytest.mark.slow
@pytest.mark.parametrize("name", ["sssp", "wcc"])
def test_config3_full_size_batches(name):
    """R-MAT scale 22 (2^22 vertices, 2^26 edges, weights 1..1000), 80%
    loaded; SSSP on the directed store, WCC on a symmetric one (undirected
    edges, wcc_inc_batch); sliding-window batches of 1 .. 100k insert+delete
    pairs, closed form after every batch."""
    algo = rg.ALGOS[name]
    V, E = 1 << 22, 1 << 26
    sym = algo == rg.WCC
    s = rg.Stream("rmat", 22, E, seed=1, wseed=2)
    n0 = int(0.8 * E)
    g = rg.Graph(V, E, symmetric=sym, dual=not sym, weighted=not sym)
    g.load_stream(s, 0, n0, not sym)
    root = s.max_degree_vertex(n0, V)
    L = g.alloc_vertex_tree_array(algo, root, trace_modified=False)
    L.build_tree()
    pos = 0
    batches = (1, 10, 100, 1000, 10000, 100000)
    total = sum(batches)
    src, dst, w = s.read(0, n0 + total)
    w = None if sym else w.astype(np.uint64)
    for B in batches:
        g.apply_stream([L], s, n0 + pos, B, pos, B, not sym)
        pos += B
        _check_window(L, algo, V, src, dst, w, pos, n0 + pos, root, sym, f"batch {B}")
