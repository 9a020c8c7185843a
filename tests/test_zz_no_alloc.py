"""The store entry point that allocates nothing (rg_edge_update_batch_static)
against the host-buffer entry point (rg_edge_update_batch), through the C ABI.

Two graphs get the same load; then the same three batches go through one
entry point each: inserts of 12,000 edges, deletes of 3,000 (a smaller
second call), and 40 inserts (the one-workgroup store path).  After every
call the returned old multiplicities agree per key as multisets (the
order among a key's ops inside one batch is not fixed, storage.hpp:378/393
semantics hold for each key), and the degrees of every vertex and a sample of
edge counts agree exactly.  Oversized and invalid batches are rejected with
the graph left unchanged.

Device buffers come from the HIP runtime the library itself loaded
(libamdhip64.so.7 through ctypes), so the pointers belong to the same runtime.
"""
import ctypes
from collections import defaultdict

import numpy as np
import pytest

import risgraph_amd as rg

pytestmark = pytest.mark.gpu


class _Hip:
    def __init__(self):
        rg.load_library()
        self.lib = ctypes.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    def h2d(self, dst, arr):
        assert self.lib.hipMemcpy(dst, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes, 1) == 0

    def d2h(self, arr, src):
        assert self.lib.hipMemcpy(arr.ctypes.data_as(ctypes.c_void_p), src, arr.nbytes, 2) == 0

    def free(self, p):
        self.lib.hipFree(p)


def _per_key(src, dst, w, olds):
    out = defaultdict(list)
    for a, b, c, o in zip(src.tolist(), dst.tolist(), w.tolist(), olds.tolist()):
        out[(a, b, c)].append(o)
    return {k: sorted(v) for k, v in out.items()}


def test_static_entry_point_matches_batch_entry_point():
    hip = _Hip()
    V, E = 4096, 1 << 16
    s = rg.Stream("rmat", 12, E, seed=13, wseed=14)
    src, dst, w = s.read()
    w = w.astype(np.uint64)
    n0 = E // 2
    ga = rg.Graph(V, E, symmetric=False, dual=True, weighted=True)
    gb = rg.Graph(V, E, symmetric=False, dual=True, weighted=True)
    for g in (ga, gb):
        g.load_stream(s, 0, n0, True)
    cap = 12000
    d_edges = hip.alloc(cap * 24)
    d_old = hip.alloc(cap * 8)
    try:
        calls = [(np.arange(n0, n0 + 12000), 1), (np.arange(0, 3000), -1), (np.arange(n0 + 12000, n0 + 12040), 1)]
        for k, (idx, delta) in enumerate(calls):
            n = len(idx)
            want = ga.update_edges(src[idx], dst[idx], w[idx], delta, True)
            rec = np.ascontiguousarray(np.stack([src[idx], dst[idx], w[idx]], axis=1).astype(np.uint64))
            hip.h2d(d_edges, rec)
            gb.update_edges_static(d_edges.value, n, delta, True, d_old.value)
            got = np.empty(n, np.uint64)
            hip.d2h(got, d_old)
            assert _per_key(src[idx], dst[idx], w[idx], got) == _per_key(src[idx], dst[idx], w[idx], want), k
            oa, ia = ga.degrees()
            ob, ib = gb.degrees()
            assert np.array_equal(oa, ob) and np.array_equal(ia, ib), k
            for i in idx[:: max(1, n // 200)]:
                a, b, c = int(src[i]), int(dst[i]), int(w[i])
                assert ga.get_edge_num(a, b, c) == gb.get_edge_num(a, b, c), (k, a, b, c)
        # without an output buffer the batch still applies
        idx = np.arange(n0 + 12040, n0 + 12100)
        ga.update_edges(src[idx], dst[idx], w[idx], 1, True)
        rec = np.ascontiguousarray(np.stack([src[idx], dst[idx], w[idx]], axis=1).astype(np.uint64))
        hip.h2d(d_edges, rec)
        gb.update_edges_static(d_edges.value, len(idx), 1, True, None)
        assert np.array_equal(ga.degrees()[0], gb.degrees()[0])
        # too large: rejected before any device work
        with pytest.raises(rg.RgError) as e:
            gb.update_edges_static(d_edges.value, rg.STATIC_MAX_OPS + 1, 1, True, d_old.value)
        assert e.value.code == rg.RG_EINVAL
        # a bad vertex id rejects the whole batch; the store is unchanged
        before = gb.degrees()[0]
        bad = rec.copy()
        bad[5, 1] = V
        hip.h2d(d_edges, bad)
        with pytest.raises(rg.RgError) as e:
            gb.update_edges_static(d_edges.value, len(bad), 1, True, d_old.value)
        assert e.value.code == rg.RG_EINVAL_VERTEX
        assert np.array_equal(before, gb.degrees()[0])
        # so does a weight over 32 bits
        bad = rec.copy()
        bad[7, 2] = 1 << 33
        hip.h2d(d_edges, bad)
        with pytest.raises(rg.RgError) as e:
            gb.update_edges_static(d_edges.value, len(bad), 1, True, d_old.value)
        assert e.value.code == rg.RG_EINVAL
        assert np.array_equal(before, gb.degrees()[0])
        # a valid call afterwards works (the flags were cleared)
        hip.h2d(d_edges, rec)
        gb.update_edges_static(d_edges.value, len(rec), -1, True, d_old.value)
        ga.update_edges(src[idx], dst[idx], w[idx], -1, True)
        assert np.array_equal(ga.degrees()[0], gb.degrees()[0])
    finally:
        hip.free(d_edges)
        hip.free(d_old)
