// server.hip -- the network request front-end and device request loop.
//
// Replaces the reference's UCX RPC servers (all_inc_rpc.cpp:397-542 for
// single updates, all_inc_rpc_txn.cpp for transactions) with TCP on the
// loopback interface and one batching loop that feeds the device:
//
//   acceptor      accepts connections on 127.0.0.1:port;
//   readers       one thread per connection reads fixed-size requests
//                 (the ClientUpdateRequest / ClientTxnUpdateRequest fields
//                 of rpc.hpp:19-61 in a packed little-endian layout) into
//                 the shared request queue;
//   batcher       takes every queued request (up to cap_max updates) as one
//                 device batch: the WAL record when the graph has a log
//                 open, the store, the safe/unsafe split, the repair;
//                 responses of safe requests are sent as soon as the split
//                 is on the host (version before the batch), those of unsafe
//                 ones after the repair (version after it).  One version per
//                 request, every update of a transaction included.
//
// A request of type End (2) ends its connection after the responses already
// owed on it.  The requests' time stamps are echoed back for the client's
// latency bookkeeping.
#include <arpa/inet.h>
#include <netinet/in.h>
#include <netinet/tcp.h>
#include <poll.h>
#include <sys/socket.h>
#include <unistd.h>

#include <atomic>
#include <cerrno>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/risgraph_amd.h"
#include "host_batch.h"

namespace {

// wire layout (packed, little-endian)
#pragma pack(push, 1)
struct WireUpdate {          // ClientUpdateRequest (rpc.hpp:19-32)
    uint32_t client_id;
    uint32_t type;           // 0 add, 1 del, 2 end
    uint64_t src, dst, data;
    int64_t request_time;
};
struct WireTxn {             // ClientTxnUpdateRequest (rpc.hpp:40-55)
    uint32_t client_id;
    uint32_t num_updates;
    uint32_t types[RG_TXN_MAX];
    uint64_t srcs[RG_TXN_MAX], dsts[RG_TXN_MAX], datas[RG_TXN_MAX];
    int64_t request_time;
};
struct WireReply {           // ClientUpdateResponse (rpc.hpp:34-38)
    uint32_t client_id;
    uint32_t status;         // 0 ok, else an RG_* code
    uint64_t version_id;
    int64_t request_time;
};
struct WireTxnReply {        // ClientTxnUpdateResponse (rpc.hpp:57-61)
    uint32_t client_id;
    uint32_t status;
    uint64_t version_ids[RG_TXN_MAX];
    int64_t request_time;
};
#pragma pack(pop)
static_assert(sizeof(WireUpdate) == 40 && sizeof(WireTxn) == 464, "request layout");
static_assert(sizeof(WireReply) == 24 && sizeof(WireTxnReply) == 144, "reply layout");

struct Conn {
    int fd = -1;
    std::mutex wmu;                  // reply writes
    std::atomic<bool> open{true};
    std::thread reader;
};

struct Req {
    std::shared_ptr<Conn> c;
    uint32_t client_id = 0, n = 0;
    int64_t t = 0;
    rg_edge e[RG_TXN_MAX];
    int8_t kind[RG_TXN_MAX];
};

bool read_full(int fd, void *p, size_t n)
{
    char *b = (char *)p;
    while (n) {
        ssize_t k = recv(fd, b, n, 0);
        if (k < 0 && errno == EINTR) continue;
        if (k <= 0) return false;
        b += k;
        n -= (size_t)k;
    }
    return true;
}

bool write_full(int fd, const void *p, size_t n)
{
    const char *b = (const char *)p;
    while (n) {
        ssize_t k = send(fd, b, n, MSG_NOSIGNAL);
        if (k < 0 && errno == EINTR) continue;
        if (k <= 0) return false;
        b += k;
        n -= (size_t)k;
    }
    return true;
}

}  // namespace

struct rg_server {
    rg_graph *g = nullptr;
    std::vector<rg_labels *> L;
    rg_server_params p{};
    int lfd = -1;
    uint16_t port = 0;
    std::atomic<bool> stop{false};
    std::thread acceptor, batcher;
    std::mutex mu;                   // queue and connection list
    std::condition_variable cv;
    std::deque<Req> q;
    std::vector<std::shared_ptr<Conn>> conns;
    std::atomic<uint64_t> n_req{0}, n_upd{0}, n_batches{0}, n_unsafe{0}, n_errors{0};

    void reply(const Req &r, uint64_t version, uint32_t status)
    {
        std::lock_guard<std::mutex> lk(r.c->wmu);
        if (!r.c->open) return;
        bool ok;
        if (p.txn) {
            WireTxnReply m{};
            m.client_id = r.client_id;
            m.status = status;
            for (uint32_t i = 0; i < r.n; i++) m.version_ids[i] = version;
            m.request_time = r.t;
            ok = write_full(r.c->fd, &m, sizeof(m));
        } else {
            WireReply m{r.client_id, status, version, r.t};
            ok = write_full(r.c->fd, &m, sizeof(m));
        }
        if (!ok) r.c->open = false;
    }

    void read_loop(std::shared_ptr<Conn> c)
    {
        for (;;) {
            Req r;
            r.c = c;
            if (p.txn) {
                WireTxn m;
                if (!read_full(c->fd, &m, sizeof(m))) break;
                if (m.num_updates == 0 || m.num_updates > RG_TXN_MAX) break;
                bool end = false;
                r.client_id = m.client_id;
                r.t = m.request_time;
                for (uint32_t i = 0; i < m.num_updates; i++) {
                    if (m.types[i] == 2) end = true;
                    r.e[r.n] = rg_edge{m.srcs[i], m.dsts[i], m.datas[i]};
                    r.kind[r.n++] = m.types[i] == 1 ? -1 : 1;
                }
                if (end) break;
            } else {
                WireUpdate m;
                if (!read_full(c->fd, &m, sizeof(m)) || m.type == 2) break;
                r.client_id = m.client_id;
                r.t = m.request_time;
                r.e[0] = rg_edge{m.src, m.dst, m.data};
                r.kind[0] = m.type == 1 ? -1 : 1;
                r.n = 1;
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                q.push_back(std::move(r));
            }
            cv.notify_one();
        }
    }

    void accept_loop()
    {
        while (!stop) {
            pollfd pf{lfd, POLLIN, 0};
            if (poll(&pf, 1, 50) <= 0) continue;
            int fd = accept(lfd, nullptr, nullptr);
            if (fd < 0) continue;
            int one = 1;
            (void)setsockopt(fd, IPPROTO_TCP, TCP_NODELAY, &one, sizeof(one));
            auto c = std::make_shared<Conn>();
            c->fd = fd;
            std::lock_guard<std::mutex> lk(mu);
            conns.push_back(c);
            c->reader = std::thread([this, c] { read_loop(c); });
        }
    }

    void batch_loop()
    {
        std::vector<Req> b;
        std::vector<rg_edge> e;
        std::vector<int8_t> kind;
        std::vector<uint32_t> txn_of;
        std::vector<uint8_t> flags;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return stop || !q.empty(); });
                if (q.empty() && stop) return;
                // admission: every arrived request, up to cap_max updates
                b.clear();
                uint64_t m = 0;
                while (!q.empty() && (b.empty() || m + q.front().n <= p.cap_max)) {
                    m += q.front().n;
                    b.push_back(std::move(q.front()));
                    q.pop_front();
                }
            }
            e.clear();
            kind.clear();
            txn_of.clear();
            for (uint32_t t = 0; t < b.size(); t++)
                for (uint32_t i = 0; i < b[t].n; i++) {
                    e.push_back(b[t].e[i]);
                    kind.push_back(b[t].kind[i]);
                    txn_of.push_back(t);
                }
            flags.assign(e.size(), 1);
            std::vector<uint8_t> sent(b.size(), 0);
            uint64_t v0 = 0, v1 = 0;
            // safe requests are answered as soon as the split reaches the host
            auto on_safe = [&](const uint8_t *f) {
                size_t k = 0;
                for (size_t t = 0; t < b.size(); t++) {
                    bool u = false;
                    for (uint32_t i = 0; i < b[t].n; i++, k++) u = u || f[k];
                    if (!u) {
                        reply(b[t], v0, 0);   // host_batch set v0 before the batch started
                        sent[t] = 1;
                    }
                }
            };
            std::string err;
            const int rc = rg::host_batch(g, L.data(), (int)L.size(), e.data(), kind.data(), txn_of.data(), e.size(),
                                          p.directed, flags.data(), on_safe, &v0, &v1, &err);
            n_batches++;
            n_req += b.size();
            n_upd += e.size();
            size_t k = 0;
            for (size_t t = 0; t < b.size(); t++) {
                bool u = false;
                for (uint32_t i = 0; i < b[t].n; i++, k++) u = u || flags[k];
                if (u) n_unsafe++;
                if (rc != RG_OK) {
                    n_errors++;
                    reply(b[t], v0, (uint32_t)rc);
                } else if (!sent[t]) {
                    reply(b[t], u ? v1 : v0, 0);
                }
            }
        }
    }

};

extern "C" {

int rg_server_start(rg_graph *g, rg_labels **L, int nL, const rg_server_params *p, rg_server **out)
{
    if (!g || !p || !out || nL < 1 || nL > 4 || !L) return RG_EINVAL;
    *out = nullptr;
    auto *s = new rg_server();
    s->g = g;
    s->L.assign(L, L + nL);
    s->p = *p;
    if (!s->p.cap_max) s->p.cap_max = 1u << 20;
    s->lfd = socket(AF_INET, SOCK_STREAM, 0);
    int one = 1;
    sockaddr_in a{};
    a.sin_family = AF_INET;
    a.sin_port = htons(p->port);
    a.sin_addr.s_addr = htonl(INADDR_LOOPBACK);
    socklen_t al = sizeof(a);
    if (s->lfd < 0 || setsockopt(s->lfd, SOL_SOCKET, SO_REUSEADDR, &one, sizeof(one)) != 0 ||
        bind(s->lfd, (sockaddr *)&a, sizeof(a)) != 0 || listen(s->lfd, 256) != 0 ||
        getsockname(s->lfd, (sockaddr *)&a, &al) != 0) {
        if (s->lfd >= 0) close(s->lfd);
        delete s;
        return RG_EINVAL;
    }
    s->port = ntohs(a.sin_port);
    s->acceptor = std::thread([s] { s->accept_loop(); });
    s->batcher = std::thread([s] { s->batch_loop(); });
    *out = s;
    return RG_OK;
}

int rg_server_port(rg_server *s, uint32_t *port)
{
    if (!s || !port) return RG_EINVAL;
    *port = s->port;
    return RG_OK;
}

int rg_server_stats(rg_server *s, uint64_t *requests, uint64_t *updates, uint64_t *batches, uint64_t *unsafe)
{
    if (!s) return RG_EINVAL;
    if (requests) *requests = s->n_req;
    if (updates) *updates = s->n_upd;
    if (batches) *batches = s->n_batches;
    if (unsafe) *unsafe = s->n_unsafe;
    return s->n_errors ? RG_EDEVICE : RG_OK;
}

int rg_server_stop(rg_server *s)
{
    if (!s) return RG_OK;
    s->stop = true;
    s->cv.notify_all();
    if (s->acceptor.joinable()) s->acceptor.join();
    // the batcher drains what is queued, then returns
    if (s->batcher.joinable()) s->batcher.join();
    std::vector<std::shared_ptr<Conn>> cs;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        cs = s->conns;
    }
    for (auto &c : cs) {
        {
            std::lock_guard<std::mutex> lk(c->wmu);
            c->open = false;
        }
        shutdown(c->fd, SHUT_RDWR);
    }
    for (auto &c : cs) {
        if (c->reader.joinable()) c->reader.join();
        close(c->fd);
    }
    close(s->lfd);
    const int rc = s->n_errors ? RG_EDEVICE : RG_OK;
    delete s;
    return rc;
}

}  // extern "C"
