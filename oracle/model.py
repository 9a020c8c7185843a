"""Brute-force multigraph model of the reference edge store (test infrastructure only).

Restates the observable behaviour of ``Graph::add_edge`` / ``del_edge``
(reference core/graph.hpp:203-253) over ``IndexedEdgeStorage::update_edge``
(core/storage.hpp:345-398) with a plain dict keyed by (vertex, nbr, weight):

* an entry is identified by (nbr, data) -- AdjEdge::operator== ignores ``num``
  (core/type.hpp:32-35), so parallel edges with different weights are distinct;
* ``update_edge`` returns the multiplicity before the change and never lets it
  go below zero (storage.hpp:378, 393);
* directed graphs update the out-list of src and the in-list of dst; a
  symmetric graph has only out-lists;
* ``add_edge(e, directed=False)`` adds both directions and returns
  ``old(d->s) - (s == d)``; ``del_edge(e, directed=False)`` returns
  ``old(d->s)`` (graph.hpp:206-211, 232-237).
"""
from collections import defaultdict

import numpy as np


class MultigraphModel:
    def __init__(self, V, symmetric=False, weighted=False):
        self.V = V
        self.symmetric = symmetric
        self.weighted = weighted
        self.out = defaultdict(int)   # (v, nbr, w) -> count
        self.inc = defaultdict(int)

    def _upd(self, s, d, w, delta):
        w = w if self.weighted else 0
        key = (s, d, w)
        old = self.out[key]
        if old + delta >= 0:
            self.out[key] = old + delta
        if not self.symmetric:
            ik = (d, s, w)
            o2 = self.inc[ik]
            if o2 + delta >= 0:
                self.inc[ik] = o2 + delta
        return old

    def add_edge(self, s, d, w=0, directed=True):
        if s >= self.V or d >= self.V:
            raise IndexError("VertexId error.")
        if not directed:
            self._upd(s, d, w, 1)
            return self._upd(d, s, w, 1) - (s == d)
        return self._upd(s, d, w, 1)

    def del_edge(self, s, d, w=0, directed=True):
        if s >= self.V or d >= self.V:
            raise IndexError("VertexId error.")
        if not directed:
            self._upd(s, d, w, -1)
            return self._upd(d, s, w, -1)
        return self._upd(s, d, w, -1)

    def count(self, s, d, w=0):
        return self.out.get((s, d, w if self.weighted else 0), 0)

    def degree(self, v):
        return sum(c for (a, _, _), c in self.out.items() if a == v)

    def live_edges(self):
        items = [(s, d, w) for (s, d, w), c in self.out.items() if c > 0]
        if not items:
            z = np.zeros(0, np.uint64)
            return z, z.copy(), z.copy()
        a = np.array(items, dtype=np.uint64)
        return a[:, 0].copy(), a[:, 1].copy(), a[:, 2].copy()
