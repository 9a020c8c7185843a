"""The request server (rg_server_*): the TCP front-end on 127.0.0.1 in place
of the reference's UCX RPC servers (all_inc_rpc.cpp:397-542,
all_inc_rpc_txn.cpp), feeding the device request loop.

Eight client threads, one connection each, send the all_inc alternation
(add stream edge n0 + k, delete stream edge k; all_inc.cpp:254-300) closed
loop, as rpc_client.cpp does: one request, then its reply.  Every reply
must carry the client's id, status 0, the echoed time stamp and a version
no older than the client's previous one; the newest version any client
saw is the graph's final one, or the one before it when the final batch held
safe updates only (those take effect at the version before their batch).
After the server stops, the four trees (BFS / SSSP / SSWP on the directed
store, WCC on its undirected view) must equal the closed form on the final
window, with valid trees.  The same runs with transactions of 1
to 16 updates (every update of a transaction gets the same version).
"""
import socket
import struct
import threading

import numpy as np
import pytest

import oracle
import risgraph_amd as rg

pytestmark = pytest.mark.gpu

REQ = struct.Struct("<IIQQQq")                   # 40 bytes
TXN = struct.Struct("<II16I16Q16Q16Qq")          # 464 bytes
REP = struct.Struct("<IIQq")                     # 24 bytes
TXN_REP = struct.Struct("<II16Qq")               # 144 bytes


def _recv(sock, n):
    buf = b""
    while len(buf) < n:
        chunk = sock.recv(n - len(buf))
        if not chunk:
            raise ConnectionError("server closed the connection")
        buf += chunk
    return buf


def _updates(src, dst, w, n0, n):
    """Update u of the alternation: (type, src, dst, w); 0 add, 1 del."""
    out = []
    for u in range(n):
        k = u // 2
        i = n0 + k if u % 2 == 0 else k
        out.append((u % 2, int(src[i]), int(dst[i]), int(w[i])))
    return out


def _setup():
    V, E = 4096, 1 << 16
    s = rg.Stream("rmat", 12, E, seed=41, wseed=42)
    src, dst, w = s.read()
    w = w.astype(np.uint64)
    n0 = E // 2
    g = rg.Graph(V, E, symmetric=False, dual=True, weighted=True)
    g.load_stream(s, 0, n0, True)
    root = s.max_degree_vertex(n0, V)
    Ls = [g.alloc_vertex_tree_array(a, root) for a in (rg.BFS, rg.SSSP, rg.SSWP, rg.WCC)]
    for L in Ls:
        L.build_tree()
    return V, src, dst, w, n0, g, root, Ls


def _check_final(V, src, dst, w, n0, k, g, root, Ls):
    ls, ld, lw = src[k:n0 + k], dst[k:n0 + k], w[k:n0 + k]
    for L in Ls:
        vals, par = L.read()
        assert np.array_equal(vals, oracle.static_values(L.algo, V, ls, ld, lw, root)), L.algo
        assert oracle.check_tree(L.algo, V, ls, ld, lw, root, vals, par, undirected=L.algo == rg.WCC) == 0


def test_server_single_updates():
    V, src, dst, w, n0, g, root, Ls = _setup()
    n, clients = 4000, 8
    ups = _updates(src, dst, w, n0, n)
    srv = rg.Server(g, Ls, port=0)
    errors, last = [], [0] * clients

    def client(c):
        try:
            sock = socket.create_connection(("127.0.0.1", srv.port))
            sock.setsockopt(socket.IPPROTO_TCP, socket.TCP_NODELAY, 1)
            prev = 0
            for u in range(c, n, clients):
                t, a, b, x = ups[u]
                sock.sendall(REQ.pack(c, t, a, b, x, 1000 * u + c))
                cid, status, ver, echo = REP.unpack(_recv(sock, REP.size))
                assert (cid, status, echo) == (c, 0, 1000 * u + c), (cid, status, echo)
                assert ver >= prev
                prev = ver
            last[c] = prev
            sock.sendall(REQ.pack(c, 2, 0, 0, 0, 0))
            sock.close()
        except Exception as e:  # reported by the main thread
            errors.append(repr(e))

    th = [threading.Thread(target=client, args=(c,)) for c in range(clients)]
    for x in th:
        x.start()
    for x in th:
        x.join(timeout=120)
    st = srv.stats()
    srv.stop()
    assert not errors, errors[:3]
    assert st["ok"] and st["requests"] == n and st["updates"] == n and 1 <= st["batches"] <= n
    # the final batch's replies carry its version when it held an unsafe
    # update, else the version before it
    assert g.version() - 1 <= max(last) <= g.version()
    _check_final(V, src, dst, w, n0, n // 2, g, root, Ls)


def test_server_transactions():
    V, src, dst, w, n0, g, root, Ls = _setup()
    n, clients = 4000, 6
    ups = _updates(src, dst, w, n0, n)
    # client c takes the contiguous slice [c*n/clients, (c+1)*n/clients) in transactions of 1..16
    rng = np.random.default_rng(8)
    plans = []
    per = n // clients
    for c in range(clients):
        lo, hi = c * per, (c + 1) * per
        txns, u = [], lo
        while u < hi:
            m = int(min(hi - u, rng.integers(1, 17)))
            txns.append(list(range(u, u + m)))
            u += m
        plans.append(txns)
    srv = rg.Server(g, Ls, port=0, txn=True)
    errors = []

    def client(c):
        try:
            sock = socket.create_connection(("127.0.0.1", srv.port))
            prev = 0
            for j, tx in enumerate(plans[c]):
                types = [ups[u][0] for u in tx] + [0] * (16 - len(tx))
                srcs = [ups[u][1] for u in tx] + [0] * (16 - len(tx))
                dsts = [ups[u][2] for u in tx] + [0] * (16 - len(tx))
                datas = [ups[u][3] for u in tx] + [0] * (16 - len(tx))
                sock.sendall(TXN.pack(c, len(tx), *types, *srcs, *dsts, *datas, j))
                r = TXN_REP.unpack(_recv(sock, TXN_REP.size))
                cid, status, vers, echo = r[0], r[1], r[2:18], r[18]
                assert (cid, status, echo) == (c, 0, j)
                assert len(set(vers[:len(tx)])) == 1, "one version per transaction"
                assert vers[0] >= prev
                prev = vers[0]
            sock.close()
        except Exception as e:  # reported by the main thread
            errors.append(repr(e))

    th = [threading.Thread(target=client, args=(c,)) for c in range(clients)]
    for x in th:
        x.start()
    for x in th:
        x.join(timeout=120)
    st = srv.stats()
    srv.stop()
    assert not errors, errors[:3]
    assert st["ok"] and st["updates"] == per * clients
    # the window: adds n0 .. n0 + n/2 and deletes 0 .. n/2 (each slice holds
    # whole add/del pairs because per is even)
    assert per % 2 == 0
    _check_final(V, src, dst, w, n0, per * clients // 2, g, root, Ls)
