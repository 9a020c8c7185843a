"""CPU oracle for the RisGraph incremental hot path (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this package, and only as the checker or the timed CPU port.
The shipped library (``risgraph_amd``) never imports it.

``liboracle.so`` is built from ``oracle/oracle.c`` by ``oracle/Makefile``.
Which check pins each function is listed at the top of ``oracle.c``.
The reference engine itself cannot be compiled in this image (core/graph.hpp
includes tbb/tbb.h; TBB is absent).  ``oracle/_ref`` holds the reference's
plain drivers compiled against this repository's Graph<> facade (the
drop-in check, ``build_ref_drivers``).
"""
import ctypes
import os
import subprocess

import numpy as np

BFS, SSSP, SSWP, WCC = 0, 1, 2, 3
ALGOS = {"bfs": BFS, "sssp": SSSP, "sswp": SSWP, "wcc": WCC}

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")
_lib = None

u64 = ctypes.c_uint64
u64p = ctypes.POINTER(ctypes.c_uint64)
vp = ctypes.c_void_p


def build():
    """Compile liboracle.so with gcc (via oracle/Makefile)."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "liboracle.so"], check=True)


REF = "/root/reference"


def build_ref_drivers():
    """Compile the reference's plain drivers against the Graph<> facade into
    oracle/_ref/ (needs the reference sources and the built library); does
    nothing where the reference sources are absent."""
    if not os.path.isdir(os.path.join(REF, "src")):
        return False
    subprocess.run(["make", "-s", "-B", "-C", _HERE, f"REF={REF}", "ref_drivers"], check=True)
    return True


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        sig = {
            "or_maxl": (u64, [ctypes.c_int]),
            "or_static_values": (ctypes.c_int, [ctypes.c_int, u64, u64, vp, vp, vp, u64, vp, vp]),
            "or_check_tree": (ctypes.c_int64, [ctypes.c_int, ctypes.c_int, u64, u64, vp, vp, vp,
                                               u64, vp, vp, vp]),
            "or_graph_create": (vp, [u64, u64, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "or_graph_destroy": (None, [vp]),
            "or_add_edge": (u64, [vp, u64, u64, u64, ctypes.c_int]),
            "or_del_edge": (u64, [vp, u64, u64, u64, ctypes.c_int]),
            "or_add_edges": (None, [vp, u64, vp, vp, vp, ctypes.c_int]),
            "or_bulk_load": (ctypes.c_int, [vp, u64, vp, vp, vp, ctypes.c_int]),
            "or_labels_assign": (ctypes.c_int64, [vp, vp, vp, vp]),
            "or_degree": (u64, [vp, u64, ctypes.c_int]),
            "or_edge_num": (u64, [vp, u64, u64, u64]),
            "or_live_edges": (u64, [vp, vp, vp, vp]),
            "or_labels_create": (vp, [vp, ctypes.c_int, u64, vp]),
            "or_labels_destroy": (None, [vp]),
            "or_build_tree": (u64, [vp, vp]),
            "or_need_update_add": (ctypes.c_int, [vp, vp, u64, u64, u64, ctypes.c_int]),
            "or_need_update_del": (ctypes.c_int, [vp, vp, u64, u64, u64, ctypes.c_int]),
            "or_update_add": (u64, [vp, vp, u64, u64, u64, ctypes.c_int]),
            "or_update_del": (u64, [vp, vp, u64, u64, u64, ctypes.c_int]),
            "or_update_add_batch": (u64, [vp, vp, u64, vp, vp, vp, ctypes.c_int]),
            "or_update_del_batch": (u64, [vp, vp, u64, vp, vp, vp, ctypes.c_int]),
            "or_labels_read": (None, [vp, vp, vp, vp]),
            "or_recompute_count": (u64, [vp]),
            "or_stream_plain": (u64, [vp, vp, u64, vp, vp, vp, vp, vp, vp, ctypes.c_int, vp]),
            "or_stream_mixed": (u64, [vp, vp, ctypes.c_int, vp, vp, u64, vp, vp, vp, vp, vp, vp, vp]),
            "or_serve_rt": (u64, [vp, vp, ctypes.c_int, vp, vp, u64, vp, vp, vp, vp, vp, vp, vp,
                                  ctypes.c_double, ctypes.c_double, vp, vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(vp)


def _u64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.uint64)


def maxl(algo):
    return int(lib().or_maxl(algo))


def static_values(algo, V, src, dst, w=None, root=0, init=None):
    """Closed-form values on the live multigraph (edges src->dst, weight w)."""
    src, dst, w, init = _u64(src), _u64(dst), _u64(w), _u64(init)
    out = np.empty(V, dtype=np.uint64)
    rc = lib().or_static_values(algo, V, len(src), _p(src), _p(dst), _p(w), root, _p(init), _p(out))
    if rc != 0:
        raise ValueError("or_static_values: bad input")
    return out


def check_tree(algo, V, src, dst, w, root, values, parents, init=None, undirected=None):
    """0 if the dependency tree is valid, else 1 + first bad vertex."""
    if undirected is None:
        undirected = algo == WCC
    src, dst, w, init = _u64(src), _u64(dst), _u64(w), _u64(init)
    values, parents = _u64(values), _u64(parents)
    return int(lib().or_check_tree(algo, int(undirected), V, len(src), _p(src), _p(dst), _p(w),
                                   root, _p(init), _p(values), _p(parents)))


class RefGraph:
    """cpu_ref: sequential restatement of core/graph.hpp's Graph<> engine."""

    def __init__(self, V, hint_edges=0, symmetric=False, dual=True, weighted=False):
        self.V = V
        self.weighted = weighted
        self.symmetric = symmetric
        self.h = lib().or_graph_create(V, hint_edges, int(symmetric), int(dual), int(weighted))

    def close(self):
        if self.h:
            lib().or_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_edge(self, s, d, w=0, directed=True):
        return int(lib().or_add_edge(self.h, s, d, w, int(directed)))

    def del_edge(self, s, d, w=0, directed=True):
        return int(lib().or_del_edge(self.h, s, d, w, int(directed)))

    def add_edges(self, src, dst, w=None, directed=True):
        src, dst, w = _u64(src), _u64(dst), _u64(w)
        lib().or_add_edges(self.h, len(src), _p(src), _p(dst), _p(w), int(directed))

    def bulk_load(self, src, dst, w=None, directed=True):
        """Same store as add_edges, built by sorting (graph must be empty)."""
        src, dst, w = _u64(src), _u64(dst), _u64(w)
        rc = lib().or_bulk_load(self.h, len(src), _p(src), _p(dst), _p(w), int(directed))
        if rc != 0:
            raise ValueError("or_bulk_load: graph not empty or id out of range")

    def degree(self, v, incoming=False):
        return int(lib().or_degree(self.h, v, int(incoming)))

    def edge_num(self, s, d, w=0):
        return int(lib().or_edge_num(self.h, s, d, w))

    def live_edges(self):
        n = int(lib().or_live_edges(self.h, None, None, None))
        s = np.empty(n, np.uint64)
        d = np.empty(n, np.uint64)
        w = np.empty(n, np.uint64)
        lib().or_live_edges(self.h, _p(s), _p(d), _p(w))
        return s, d, w


class RefLabels:
    def __init__(self, g, algo, root=0, init=None):
        self.g = g
        self.algo = algo
        self.root = root
        self.init = None if init is None else _u64(init)
        self.h = lib().or_labels_create(g.h, algo, root, _p(self.init))

    def close(self):
        if self.h:
            lib().or_labels_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build(self):
        return int(lib().or_build_tree(self.g.h, self.h))

    def need_add(self, s, d, w=0, directed=True):
        return bool(lib().or_need_update_add(self.g.h, self.h, s, d, w, int(directed)))

    def need_del(self, s, d, w=0, directed=True):
        return bool(lib().or_need_update_del(self.g.h, self.h, s, d, w, int(directed)))

    def inc_add(self, s, d, w=0, directed=True):
        return int(lib().or_update_add(self.g.h, self.h, s, d, w, int(directed)))

    def inc_del(self, s, d, w=0, directed=True):
        return int(lib().or_update_del(self.g.h, self.h, s, d, w, int(directed)))

    def inc_add_batch(self, s, d, w=None, directed=True):
        s, d, w = _u64(s), _u64(d), _u64(w)
        return int(lib().or_update_add_batch(self.g.h, self.h, len(s), _p(s), _p(d), _p(w),
                                             int(directed)))

    def inc_del_batch(self, s, d, w=None, directed=True):
        s, d, w = _u64(s), _u64(d), _u64(w)
        return int(lib().or_update_del_batch(self.g.h, self.h, len(s), _p(s), _p(d), _p(w),
                                             int(directed)))

    def read(self):
        V = self.g.V
        v = np.empty(V, np.uint64)
        p = np.empty(V, np.uint64)
        pw = np.empty(V, np.uint64)
        lib().or_labels_read(self.h, _p(v), _p(p), _p(pw))
        return v, p, pw

    def assign(self, values, parents):
        """Start from a given fixpoint (values, parent vertices) instead of build()."""
        values, parents = _u64(values), _u64(parents)
        bad = int(lib().or_labels_assign(self.g.h, self.h, _p(values), _p(parents)))
        if bad:
            raise ValueError(f"or_labels_assign: no consistent parent edge at vertex {bad - 1}")

    def recomputes(self):
        return int(lib().or_recompute_count(self.h))

    def stream_plain(self, add_s, add_d, add_w, del_s, del_d, del_w, directed=True, latency=True):
        """Plain sliding-window steps (bfs_inc.cpp:160-189); returns (#unsafe, lat_ns)."""
        a = [_u64(x) for x in (add_s, add_d, add_w, del_s, del_d, del_w)]
        n = len(a[0])
        lat = np.empty(n, np.float64) if latency else None
        unsafe = lib().or_stream_plain(self.g.h, self.h, n, *[_p(x) for x in a], int(directed),
                                       _p(lat))
        return int(unsafe), lat


def stream_mixed(gd, Ld, gu, Lu, n_updates, add_s, add_d, add_w, del_s, del_d, del_w, latency=True):
    """all_inc-style alternation over BFS/SSSP/SSWP (Ld on gd) and WCC (Lu on gu).

    Returns (#updates that ran incremental work, per-update wall ns or None)."""
    arr = (vp * max(1, len(Ld)))(*[L.h for L in Ld])
    a = [_u64(x) for x in (add_s, add_d, add_w, del_s, del_d, del_w)]
    lat = np.empty(n_updates, np.float64) if latency else None
    unsafe = lib().or_stream_mixed(gd.h if gd else None, arr, len(Ld), gu.h if gu else None,
                                   Lu.h if Lu else None, n_updates, *[_p(x) for x in a], _p(lat))
    return int(unsafe), lat


def serve_rt(gd, Ld, gu, Lu, n_updates, add_s, add_d, add_w, del_s, del_d, del_w, arrival,
             timeout_ms=20.0, target_timeout_rate=0.001):
    """The rt driver's serving loop on host threads (or_serve_rt); returns (stats, latencies_ms)."""
    arr = (vp * max(1, len(Ld)))(*[L.h for L in Ld])
    a = [_u64(x) for x in (add_s, add_d, add_w, del_s, del_d, del_w)]
    t = np.ascontiguousarray(arrival, dtype=np.float64)
    lat = np.full(n_updates, np.inf, np.float64)
    st = np.zeros(7, np.uint64)
    lib().or_serve_rt(gd.h if gd else None, arr, len(Ld), gu.h if gu else None, Lu.h if Lu else None,
                      n_updates, *[_p(x) for x in a], _p(t), timeout_ms, target_timeout_rate, _p(lat), _p(st))
    keys = ("served", "rounds", "unsafe", "threshold", "overloaded", "threads", "wall_ns")
    return dict(zip(keys, (int(x) for x in st))), lat
