"""risgraph_amd -- MI355X-native incremental graph hot path (Python side).

Thin ctypes mirror of the C ABI in ``include/risgraph_amd.h``.  The compute
path is the HIP library ``risgraph_amd/librisgraph_amd.so`` (built by
``risgraph_amd/build.py``); there is no CPU fallback: importing the bindings
without the library, or calling them without a GPU, raises.

Class and method names follow the reference ``Graph<>`` interface
(core/graph.hpp): ``add_edge`` / ``del_edge`` return the old multiplicity,
``build_tree``, ``update_tree_add`` / ``update_tree_del`` (single edge or
batch), ``need_update_tree_add`` / ``need_update_tree_del``,
``get_modified`` / ``clear_modified``.
"""
import ctypes
import os

import numpy as np

BFS, SSSP, SSWP, WCC = 0, 1, 2, 3
ALGOS = {"bfs": BFS, "sssp": SSSP, "sswp": SSWP, "wcc": WCC}
RG_OK, RG_EINVAL_VERTEX, RG_ENOMEM, RG_EALGO, RG_EDEVICE, RG_EINVAL = range(6)
STATIC_MAX_OPS = 16384   # RG_STATIC_MAX_OPS: largest batch of rg_edge_update_batch_static

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librisgraph_amd.so")

u64 = ctypes.c_uint64
u64p = ctypes.POINTER(ctypes.c_uint64)
vp = ctypes.c_void_p
ci = ctypes.c_int


class RgEdge(ctypes.Structure):
    _fields_ = [("src", u64), ("dst", u64), ("data", u64)]


class BatchStats(ctypes.Structure):
    _fields_ = [("seeds", u64), ("roots", u64), ("levels", u64), ("grid_steps", u64),
                ("rebuilds", u64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


TUNE_BUCKETS = 4


class AutotuneParams(ctypes.Structure):
    _fields_ = [("ins_first", ctypes.c_uint64), ("del_first", ctypes.c_uint64), ("directed", ctypes.c_int),
                ("pairs", ctypes.c_uint32), ("target_ms", ctypes.c_double)]


class AutotuneResult(ctypes.Structure):
    _fields_ = [("ops", ctypes.c_uint32 * TUNE_BUCKETS), ("cap_x4", ctypes.c_uint32 * TUNE_BUCKETS),
                ("weight", (ctypes.c_uint32 * 4) * TUNE_BUCKETS), ("default_ms", ctypes.c_double * TUNE_BUCKETS),
                ("best_ms", ctypes.c_double * TUNE_BUCKETS), ("batch_cap", ctypes.c_uint32),
                ("batches", ctypes.c_uint64)]

    def as_dict(self):
        buckets = []
        for b in range(TUNE_BUCKETS):
            if self.ops[b]:
                buckets.append({"ops": int(self.ops[b]), "cap_x4": int(self.cap_x4[b]),
                                "weight": [int(x) for x in self.weight[b]],
                                "default_ms": round(self.default_ms[b], 4), "best_ms": round(self.best_ms[b], 4)})
        return {"buckets": buckets, "batch_cap": int(self.batch_cap), "batches": int(self.batches)}


class ServeParams(ctypes.Structure):
    _fields_ = [("ins_first", u64), ("del_first", u64), ("n_updates", u64), ("rate", ctypes.c_double),
                ("seed", u64), ("clients", ctypes.c_uint32), ("batch_cap", ctypes.c_uint32),
                ("cap_max", ctypes.c_uint32), ("target_ms", ctypes.c_double), ("directed", ci),
                ("split", ci), ("cap_unsafe", ci)]


class ServeResult(ctypes.Structure):
    _fields_ = [("wall_s", ctypes.c_double), ("updates_per_s", ctypes.c_double),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double), ("p999_ms", ctypes.c_double),
                ("max_ms", ctypes.c_double), ("mean_ms", ctypes.c_double), ("mean_batch", ctypes.c_double),
                ("batches", u64), ("max_batch", u64), ("grid_steps", u64), ("rebuilds", u64),
                ("overloaded", u64), ("served", u64), ("gpu_ms", ctypes.c_double), ("unsafe", u64),
                ("safe_frac", ctypes.c_double), ("p50_safe_ms", ctypes.c_double),
                ("p50_unsafe_ms", ctypes.c_double), ("p999_unsafe_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: (float(getattr(self, k)) if t is ctypes.c_double else int(getattr(self, k)))
                for k, t in self._fields_}


class RgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rg error {code}: {msg}")
        self.code = code


# exported symbols and signatures (must match include/risgraph_amd.h)
SIGNATURES = {
    "rg_last_error": (ctypes.c_char_p, []),
    "rg_graph_create": (ci, [u64, u64, ci, ci, ci, ctypes.POINTER(vp)]),
    "rg_graph_destroy": (ci, [vp]),
    "rg_graph_info": (ci, [vp, u64p, u64p, u64p, u64p, u64p]),
    "rg_txn_apply": (ci, [vp, vp, ci, vp, vp, u64, vp, u64, ci, vp, vp]),
    "rg_wal_open": (ci, [vp, ctypes.c_char_p, ci]),
    "rg_wal_close": (ci, [vp]),
    "rg_wal_replay": (ci, [vp, vp, ci, ctypes.c_char_p, u64p]),
    "rg_snapshot_save": (ci, [vp, vp, ci, ctypes.c_char_p]),
    "rg_snapshot_load": (ci, [ctypes.c_char_p, ctypes.POINTER(vp), vp, ci, ctypes.POINTER(ci)]),
    "rg_stream_from_edges": (ci, [vp, u64, ci, ctypes.POINTER(vp)]),
    "rg_labels_assign": (ci, [vp, vp, vp]),
    "rg_graph_shape": (ci, [vp, ctypes.POINTER(ci), ctypes.POINTER(ci), ctypes.POINTER(ci)]),
    "rg_graph_capture": (ci, [vp, ci]),
    "rg_last_unsafe": (ci, [vp, u64p]),
    "rg_autotune": (ci, [vp, vp, ci, vp, vp, vp]),
    "rg_graph_pool_info": (ci, [vp, vp]),
    "rg_graph_capture_info": (ci, [vp, ctypes.POINTER(ci), ctypes.POINTER(ctypes.c_uint32), u64p]),
    "rg_labels_algo": (ci, [vp, ctypes.POINTER(ci), u64p]),
    "rg_server_start": (ci, [vp, vp, ci, vp, ctypes.POINTER(vp)]),
    "rg_server_port": (ci, [vp, ctypes.POINTER(ctypes.c_uint32)]),
    "rg_server_stats": (ci, [vp, u64p, u64p, u64p, u64p]),
    "rg_server_stop": (ci, [vp]),
    "rg_graph_version": (ci, [vp, u64p]),
    "rg_history_enable": (ci, [vp, u64]),
    "rg_history_read": (ci, [vp, u64, u64, u64, vp]),
    "rg_history_gc": (ci, [vp, u64]),
    "rg_history_info": (ci, [vp, u64p, u64p, u64p]),
    "rg_edge_update_batch_static": (ci, [vp, vp, u64, ci, ci, vp]),
    "rg_edge_update": (ci, [vp, u64, u64, u64, ci, ci, u64p]),
    "rg_edge_update_batch": (ci, [vp, vp, u64, ci, ci, vp]),
    "rg_degree": (ci, [vp, u64, ci, u64p]),
    "rg_edge_num": (ci, [vp, u64, u64, u64, u64p]),
    "rg_labels_create": (ci, [vp, ci, u64, vp, ci, ctypes.POINTER(vp)]),
    "rg_labels_destroy": (ci, [vp]),
    "rg_labels_set_tuning": (ci, [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                  ctypes.c_uint32, ctypes.c_double, ctypes.c_uint32,
                                  ctypes.c_uint32]),
    "rg_build": (ci, [vp, u64p]),
    "rg_inc_add": (ci, [vp, vp, u64, ci, u64p]),
    "rg_inc_del": (ci, [vp, vp, u64, ci, u64p]),
    "rg_classify": (ci, [vp, vp, vp, u64, ci, vp]),
    "rg_labels_read": (ci, [vp, u64, u64, vp, vp]),
    "rg_modified_read": (ci, [vp, vp, u64, u64p, ctypes.POINTER(ci)]),
    "rg_modified_clear": (ci, [vp]),
    "rg_stream_create": (ci, [ci, u64, u64, u64, ctypes.c_double, ctypes.c_double,
                              ctypes.c_double, u64, ctypes.POINTER(vp)]),
    "rg_stream_destroy": (ci, [vp]),
    "rg_stream_read": (ci, [vp, u64, u64, vp, vp, vp]),
    "rg_stream_max_degree_vertex": (ci, [vp, u64, u64, u64p]),
    "rg_stream_load": (ci, [vp, vp, u64, u64, ci]),
    "rg_stream_apply": (ci, [vp, vp, ci, vp, u64, u64, u64, u64, ci, vp]),
    "rg_stream_plain": (ci, [vp, vp, ci, vp, u64, u64, u64, ci, vp, vp]),
    "rg_last_gpu_ms": (ci, [vp, ctypes.POINTER(ctypes.c_double)]),
    "rg_serve": (ci, [vp, vp, ci, vp, vp, vp, vp]),
    "rg_profile_enable": (ci, [vp, ci]),
    "rg_profile_read": (ci, [vp, ci, u64p, ctypes.POINTER(ctypes.c_double), u64p]),
    "rg_degrees": (ci, [vp, vp, vp]),
    "rg_labels_phase_time": (ci, [vp, vp, vp]),
    "rg_labels_dense": (ci, [vp, ci, u64p]),
    "rg_labels_team_weight": (ci, [vp, ctypes.c_uint32]),
    "rg_labels_phase_detail": (ci, [vp, vp, vp, vp]),
}

PHASES = ["done", "seed", "push", "del_roots", "rescue", "expand", "reset", "pull", "build_init", "rebuild"]

_lib = None


def load_library(path=LIB_PATH):
    """Load the HIP library; raises if it is missing (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: run risgraph_amd/build.py (hipcc, gfx950) first")
        L = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != RG_OK:
        raise RgError(rc, load_library().rg_last_error().decode(errors="replace"))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(vp)


def _edges(src, dst, w=None):
    src = np.atleast_1d(np.asarray(src, dtype=np.uint64))
    dst = np.atleast_1d(np.asarray(dst, dtype=np.uint64))
    e = np.zeros((len(src), 3), dtype=np.uint64)
    e[:, 0] = src
    e[:, 1] = dst
    if w is not None:
        e[:, 2] = np.atleast_1d(np.asarray(w, dtype=np.uint64))
    return np.ascontiguousarray(e)


class Stream:
    """Device-resident synthetic edge stream (R-MAT or Erdos-Renyi)."""

    def __init__(self, kind, scale_or_v, n_edges, seed=1, a=0.57, b=0.19, c=0.19, wseed=0):
        h = vp()
        _check(load_library().rg_stream_create(0 if kind == "rmat" else 1, scale_or_v, n_edges, seed,
                                               a, b, c, wseed, ctypes.byref(h)))
        self.h = h
        self.n = n_edges
        self.weighted = bool(wseed)

    def read(self, first=0, n=None):
        n = self.n - first if n is None else n
        s = np.empty(n, np.uint64)
        d = np.empty(n, np.uint64)
        w = np.empty(n, np.uint32) if self.weighted else None
        _check(load_library().rg_stream_read(self.h, first, n, _ptr(s), _ptr(d), _ptr(w)))
        return s, d, w

    @classmethod
    def from_edges(cls, src, dst, w=None):
        """rg_stream_from_edges: a device stream of the given host edges."""
        e = _edges(src, dst, w)
        h = vp()
        _check(load_library().rg_stream_from_edges(_ptr(e), len(e), int(w is not None), ctypes.byref(h)))
        obj = cls.__new__(cls)
        obj.h, obj.n, obj.weighted = h, len(e), w is not None
        return obj

    def max_degree_vertex(self, n, V):
        out = u64()
        _check(load_library().rg_stream_max_degree_vertex(self.h, n, V, ctypes.byref(out)))
        return int(out.value)

    def close(self):
        if getattr(self, "h", None):
            load_library().rg_stream_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """Graph<EdgeData> (core/graph.hpp:41-100) backed by the HBM slab store."""

    def __init__(self, vertices, hint_edges=0, symmetric=False, dual=True, weighted=False):
        h = vp()
        _check(load_library().rg_graph_create(vertices, hint_edges, int(symmetric), int(dual),
                                              int(weighted), ctypes.byref(h)))
        self.h = h
        self.V = vertices
        self.symmetric = symmetric
        self.weighted = weighted
        self._labels = []

    def close(self):
        for L in self._labels:
            L.close()
        self._labels = []
        if getattr(self, "h", None):
            load_library().rg_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- edge store (graph.hpp:203-253) ---
    def add_edge(self, src, dst, data=0, directed=True):
        old = u64()
        _check(load_library().rg_edge_update(self.h, src, dst, data, 1, int(directed), ctypes.byref(old)))
        return int(old.value)

    def del_edge(self, src, dst, data=0, directed=True):
        old = u64()
        _check(load_library().rg_edge_update(self.h, src, dst, data, -1, int(directed), ctypes.byref(old)))
        return int(old.value)

    def update_edges(self, src, dst, w=None, delta=1, directed=True):
        e = _edges(src, dst, w)
        old = np.empty(len(e), np.uint64)
        _check(load_library().rg_edge_update_batch(self.h, _ptr(e), len(e), delta, int(directed), _ptr(old)))
        return old

    def update_edges_static(self, d_edges, n, delta=1, directed=True, d_old=None):
        """rg_edge_update_batch_static: device pointers (ints) to n rg_edge
        records and n u64 return values; nothing is allocated."""
        _check(load_library().rg_edge_update_batch_static(self.h, d_edges, n, delta, int(directed), d_old))

    def get_degree(self, v, incoming=False):
        out = u64()
        _check(load_library().rg_degree(self.h, v, int(incoming), ctypes.byref(out)))
        return int(out.value)

    def get_edge_num(self, src, dst, data=0):
        out = u64()
        _check(load_library().rg_edge_num(self.h, src, dst, data, ctypes.byref(out)))
        return int(out.value)

    def info(self):
        vals = [u64() for _ in range(5)]
        _check(load_library().rg_graph_info(self.h, *[ctypes.byref(v) for v in vals]))
        keys = ("vertices", "slots_used", "slots_cap", "hash_used", "hash_cap")
        return {k: int(v.value) for k, v in zip(keys, vals)}

    def pool_info(self):
        """rg_graph_pool_info: recycled / dropped blocks and dropped entries
        of the slot pool and the hash pool."""
        out = np.zeros(6, np.uint64)
        _check(load_library().rg_graph_pool_info(self.h, _ptr(out)))
        o = [int(x) for x in out]
        return {"recycled": o[0:2], "dropped": o[2:4], "dropped_entries": o[4:6]}

    def load_stream(self, stream, first, n, directed=True):
        _check(load_library().rg_stream_load(self.h, stream.h, first, n, int(directed)))

    # --- labels (graph.hpp:255-262) ---
    def alloc_vertex_tree_array(self, algo, root=0, init=None, trace_modified=True):
        L = Labels(self, algo, root, init, trace_modified)
        self._labels.append(L)
        return L

    def apply_stream(self, labels, stream, ins_first, n_ins, del_first, n_del, directed=True):
        arr = (vp * max(1, len(labels)))(*[L.h for L in labels])
        st = BatchStats()
        _check(load_library().rg_stream_apply(self.h, arr, len(labels), stream.h, ins_first, n_ins,
                                              del_first, n_del, int(directed), ctypes.byref(st)))
        return st.as_dict()

    def stream_plain(self, labels, stream, ins_first, del_first, n_steps, directed=True, latency=True):
        arr = (vp * max(1, len(labels)))(*[L.h for L in labels])
        st = BatchStats()
        lat = np.empty(n_steps, np.float64) if latency else None
        _check(load_library().rg_stream_plain(self.h, arr, len(labels), stream.h, ins_first, del_first,
                                              n_steps, int(directed), _ptr(lat), ctypes.byref(st)))
        return lat, st.as_dict()

    def serve(self, labels, stream, ins_first, del_first, n_updates, rate, batch_cap=1, cap_max=0,
              target_ms=0.0, seed=4, clients=256, directed=True, latencies=False, split=True, cap_unsafe=False):
        """Poisson-arrival serving loop (rg_serve); returns (result dict, latencies_ms or None)."""
        arr = (vp * max(1, len(labels)))(*[L.h for L in labels])
        p = ServeParams(ins_first, del_first, n_updates, rate, seed, clients, batch_cap,
                        cap_max or batch_cap, target_ms, int(directed), int(split), int(cap_unsafe))
        r = ServeResult()
        lat = np.empty(n_updates, np.float64) if latencies else None
        _check(load_library().rg_serve(self.h, arr, len(labels), stream.h, ctypes.byref(p),
                                       ctypes.byref(r), _ptr(lat)))
        return r.as_dict(), lat

    def profile(self, mask=3):
        """Time kernels with HIP events: bit 0 store apply, bit 1 team engine
        launches, bit 2 grid-wide engine passes."""
        _check(load_library().rg_profile_enable(self.h, int(mask)))

    def profile_read(self, kind):
        """kind 0 = store apply kernels (batches over 64 edge ops),
        1 = persistent team engine launches, 2 = grid-wide engine passes,
        3 + p = grid-wide passes of phase p (PHASES[p]), 13 = the
        one-kernel store step of batches of at most 64 edge ops."""
        n, units, ms = u64(), u64(), ctypes.c_double()
        _check(load_library().rg_profile_read(self.h, kind, ctypes.byref(n), ctypes.byref(ms),
                                              ctypes.byref(units)))
        return {"launches": int(n.value), "ms": ms.value, "units": int(units.value)}

    def degrees(self):
        out = np.empty(self.V, np.uint32)
        inc = np.empty(self.V, np.uint32)
        _check(load_library().rg_degrees(self.h, _ptr(out), _ptr(inc)))
        return out, inc

    @classmethod
    def _wrap(cls, h):
        obj = cls.__new__(cls)
        obj.h = h
        V = u64()
        _check(load_library().rg_graph_info(h, ctypes.byref(V), None, None, None, None))
        sym, dual, wt = ci(), ci(), ci()
        _check(load_library().rg_graph_shape(h, ctypes.byref(sym), ctypes.byref(dual), ctypes.byref(wt)))
        obj.V, obj.symmetric, obj.weighted, obj._labels = int(V.value), bool(sym.value), bool(wt.value), []
        return obj

    def autotune(self, labels, stream, ins_first, del_first, target_ms=0.0, pairs=3, directed=True):
        """rg_autotune: time forward/inverse batch pairs per batch-size bucket
        and keep the fastest engine settings; returns the table and, with a
        target, the serving batch cap."""
        arr = (vp * max(1, len(labels)))(*[L.h for L in labels])
        p = AutotuneParams(ins_first, del_first, int(directed), pairs, target_ms)
        r = AutotuneResult()
        _check(load_library().rg_autotune(self.h, arr, len(labels), stream.h, ctypes.byref(p), ctypes.byref(r)))
        return r.as_dict()

    def capture(self, on=True):
        """rg_graph_capture: replay stream batches' window/store/split
        launches from captured hipGraphs (on) or launch them one by one."""
        _check(load_library().rg_graph_capture(self.h, int(on)))

    def capture_info(self):
        """rg_graph_capture_info: {"on", "pipelines", "replays"}."""
        on, n, r = ci(), ctypes.c_uint32(), u64()
        _check(load_library().rg_graph_capture_info(self.h, ctypes.byref(on), ctypes.byref(n), ctypes.byref(r)))
        return {"on": bool(on.value), "pipelines": int(n.value), "replays": int(r.value)}

    def wal_open(self, path, sync=True):
        """rg_wal_open: log every later transaction batch to `path`."""
        _check(load_library().rg_wal_open(self.h, os.fsencode(path), int(sync)))

    def wal_close(self):
        _check(load_library().rg_wal_close(self.h))

    def wal_replay(self, labels, path):
        """rg_wal_replay: apply the logged batches above the graph's version; returns their count."""
        arr = (vp * max(1, len(labels)))(*[L.h for L in labels])
        n = u64()
        _check(load_library().rg_wal_replay(self.h, arr, len(labels), os.fsencode(path), ctypes.byref(n)))
        return int(n.value)

    def snapshot_save(self, labels, path):
        arr = (vp * max(1, len(labels)))(*[L.h for L in labels])
        _check(load_library().rg_snapshot_save(self.h, arr, len(labels), os.fsencode(path)))

    def txn_apply(self, labels, src, dst, w, kind, txn_len, directed=True):
        """rg_txn_apply: a batch of transactions; returns (unsafe flags, versions) per transaction."""
        arr = (vp * max(1, len(labels)))(*[L.h for L in labels])
        e = _edges(src, dst, w)
        k = np.ascontiguousarray(kind, dtype=np.int8)
        tl = np.ascontiguousarray(txn_len, dtype=np.uint32)
        unsafe = np.empty(len(tl), np.uint8)
        ver = np.empty(len(tl), np.uint64)
        _check(load_library().rg_txn_apply(self.h, arr, len(labels), _ptr(e), _ptr(k), len(e), _ptr(tl), len(tl),
                                           int(directed), _ptr(unsafe), _ptr(ver)))
        return unsafe.astype(bool), ver

    def version(self):
        """rg_graph_version: device batches applied so far."""
        v = u64()
        _check(load_library().rg_graph_version(self.h, ctypes.byref(v)))
        return int(v.value)

    def last_gpu_ms(self):
        ms = ctypes.c_double()
        _check(load_library().rg_last_gpu_ms(self.h, ctypes.byref(ms)))
        return ms.value

    def last_unsafe(self):
        """rg_last_unsafe: unsafe edge ops of the last stream batch."""
        n = u64()
        _check(load_library().rg_last_unsafe(self.h, ctypes.byref(n)))
        return int(n.value)


class Labels:
    """One algorithm's dependency tree: VertexTree array (graph.hpp:54-59)."""

    def __init__(self, g, algo, root=0, init=None, trace_modified=True):
        self.g = g
        self.algo = algo
        self.root = root
        init_arr = None if init is None else np.ascontiguousarray(init, dtype=np.uint64)
        h = vp()
        _check(load_library().rg_labels_create(g.h, algo, root, _ptr(init_arr), int(trace_modified),
                                               ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            load_library().rg_labels_destroy(self.h)
            self.h = None

    def set_tuning(self, rescue_level=0, rescue_frontier=None, recompute_level=10,
                   recompute_abs=2000000, recompute_frac=0.05, walk_limit=0, block_max=None):
        """Deletion strategy knobs (rg_labels_set_tuning); block_max None =
        automatic; rescue_frontier None = the library default for the
        algorithm (0, or 0xFFFFFFFF for SSWP and WCC: rescue on every level)."""
        if rescue_frontier is None:
            rescue_frontier = 0xFFFFFFFF if self.algo in (WCC, SSWP) else 0
        bm = 0xFFFFFFFF if block_max is None else int(block_max)
        _check(load_library().rg_labels_set_tuning(self.h, rescue_level, rescue_frontier, recompute_level,
                                                   recompute_abs, recompute_frac, walk_limit, bm))

    def dense(self, mode=-1):
        """rg_labels_dense: set the build direction rule (1 reference
        push/pull switch, 0 push only, -1 unchanged); returns the dense
        passes run so far."""
        n = u64()
        _check(load_library().rg_labels_dense(self.h, int(mode), ctypes.byref(n)))
        return int(n.value)

    def team_weight(self, weight=0):
        """rg_labels_team_weight: share of the batch repair launch (0 = by algorithm)."""
        _check(load_library().rg_labels_team_weight(self.h, int(weight)))

    def build_tree(self):
        r = u64()
        _check(load_library().rg_build(self.h, ctypes.byref(r)))
        return int(r.value)

    def update_tree_add(self, src, dst, w=None, directed=True):
        e = _edges(src, dst, w)
        r = u64()
        _check(load_library().rg_inc_add(self.h, _ptr(e), len(e), int(directed), ctypes.byref(r)))
        return int(r.value)

    def update_tree_del(self, src, dst, w=None, directed=True):
        e = _edges(src, dst, w)
        r = u64()
        _check(load_library().rg_inc_del(self.h, _ptr(e), len(e), int(directed), ctypes.byref(r)))
        return int(r.value)

    def classify(self, src, dst, w, kind, directed=True):
        e = _edges(src, dst, w)
        k = np.ascontiguousarray(kind, dtype=np.int8)
        out = np.empty(len(e), np.uint8)
        _check(load_library().rg_classify(self.h, _ptr(e), _ptr(k), len(e), int(directed), _ptr(out)))
        return out.astype(bool)

    def need_update_tree_add(self, src, dst, w=0, directed=True):
        return bool(self.classify([src], [dst], [w], [1], directed)[0])

    def need_update_tree_del(self, src, dst, w=0, directed=True):
        return bool(self.classify([src], [dst], [w], [-1], directed)[0])

    def read(self, first=0, count=None):
        count = self.g.V - first if count is None else count
        v = np.empty(count, np.uint64)
        p = np.empty(count, np.uint64)
        _check(load_library().rg_labels_read(self.h, first, count, _ptr(v), _ptr(p)))
        return v, p

    def phase_time(self):
        """Block-mode engine time per phase: {phase: (microseconds, passes)}."""
        us = np.zeros(10, np.float64)
        n = np.zeros(10, np.uint64)
        _check(load_library().rg_labels_phase_time(self.h, _ptr(us), _ptr(n)))
        return {PHASES[p]: (float(us[p]), int(n[p])) for p in range(10) if n[p]}

    def phase_detail(self):
        """Team pass breakdown per phase: {phase: dict(vertex_us, barrier1_us,
        chunk_us, barrier2_us, work, chunks)} (first workgroup's clocks)."""
        st = np.zeros(40, np.float64)
        work = np.zeros(10, np.uint64)
        ch = np.zeros(10, np.uint64)
        _check(load_library().rg_labels_phase_detail(self.h, _ptr(st), _ptr(work), _ptr(ch)))
        keys = ("vertex_us", "barrier1_us", "chunk_us", "barrier2_us")
        out = {}
        for p in range(10):
            d = {k: float(st[i * 10 + p]) for i, k in enumerate(keys)}
            d["work"] = int(work[p])
            d["chunks"] = int(ch[p])
            if any(d.values()):
                out[PHASES[p]] = d
        return out

    def get_modified(self):
        n = u64()
        allf = ci()
        _check(load_library().rg_modified_read(self.h, None, 0, ctypes.byref(n), ctypes.byref(allf)))
        out = np.empty(int(n.value), np.uint64)
        _check(load_library().rg_modified_read(self.h, _ptr(out), len(out), ctypes.byref(n), ctypes.byref(allf)))
        return out[: int(n.value)], bool(allf.value)

    @classmethod
    def _wrap(cls, g, h):
        obj = cls.__new__(cls)
        a, r = ci(), u64()
        _check(load_library().rg_labels_algo(h, ctypes.byref(a), ctypes.byref(r)))
        obj.g, obj.h, obj.algo, obj.root = g, h, int(a.value), int(r.value)
        g._labels.append(obj)
        return obj

    def assign(self, values, parents):
        """rg_labels_assign: set values and parent vertices (V = empty)."""
        v = np.ascontiguousarray(values, dtype=np.uint64)
        p = np.ascontiguousarray(parents, dtype=np.uint64)
        _check(load_library().rg_labels_assign(self.h, _ptr(v), _ptr(p)))

    def enable_history(self, records):
        """rg_history_enable: keep a device ring of `records` value records."""
        _check(load_library().rg_history_enable(self.h, int(records)))

    def history_read(self, version, first=0, count=None):
        """Values at a past version (rg_history_read)."""
        count = self.g.V - first if count is None else count
        out = np.empty(count, np.uint64)
        _check(load_library().rg_history_read(self.h, int(version), first, count, _ptr(out)))
        return out

    def history_gc(self, version):
        _check(load_library().rg_history_gc(self.h, int(version)))

    def history_info(self):
        a, b, c = u64(), u64(), u64()
        _check(load_library().rg_history_info(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return {"oldest_version": int(a.value), "records": int(b.value), "capacity": int(c.value)}

    def clear_modified(self):
        _check(load_library().rg_modified_clear(self.h))


def snapshot_load(path, cap=8):
    """rg_snapshot_load: (Graph, [Labels]) rebuilt from a snapshot file."""
    g = vp()
    arr = (vp * cap)()
    n = ci()
    _check(load_library().rg_snapshot_load(os.fsencode(path), ctypes.byref(g), arr, cap, ctypes.byref(n)))
    G = Graph._wrap(g)
    return G, [Labels._wrap(G, vp(arr[i])) for i in range(n.value)]


class ServerParams(ctypes.Structure):
    _fields_ = [("port", ctypes.c_uint16), ("txn", ci), ("directed", ci), ("cap_max", ctypes.c_uint32)]


class Server:
    """rg_server_*: TCP request front-end (127.0.0.1) feeding the device request loop."""

    def __init__(self, graph, labels, port=0, txn=False, directed=True, cap_max=0):
        self.graph, self.labels = graph, list(labels)
        arr = (vp * len(self.labels))(*[L.h for L in self.labels])
        p = ServerParams(port, int(txn), int(directed), cap_max)
        h = vp()
        _check(load_library().rg_server_start(graph.h, arr, len(self.labels), ctypes.byref(p), ctypes.byref(h)))
        self.h = h
        pt = ctypes.c_uint32()
        _check(load_library().rg_server_port(h, ctypes.byref(pt)))
        self.port = int(pt.value)

    def stats(self):
        v = [u64() for _ in range(4)]
        rc = load_library().rg_server_stats(self.h, *[ctypes.byref(x) for x in v])
        d = dict(zip(("requests", "updates", "batches", "unsafe"), (int(x.value) for x in v)))
        d["ok"] = rc == RG_OK
        return d

    def stop(self):
        if getattr(self, "h", None):
            rc = load_library().rg_server_stop(self.h)
            self.h = None
            _check(rc)
