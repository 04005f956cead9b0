// order.cpp -- a locality order of a graph's ids (host side of libdlg.so).
//
// TC in the bitmap layout sets, per round and source x, the bits of x's
// newly reached ends in row x of F.  On road-like and mesh-like graphs those
// ends form a ring around x; in input order a ring's ids are scattered over
// the whole row, so nearly every end touches a different word and a
// different cache line of F.  Renumbered so that nodes close in the graph
// get close ids, a ring crosses far fewer words and lines (on the 400 x 400
// grid: 42 % fewer words, 63 % fewer 128-byte lines, 31 % fewer table
// entries per round), and the engine maps the result back at read-out.
//
// Recursive BFS bisection (no coordinates needed): a set of nodes is split
// into its connected components (undirected); a component larger than a leaf
// is ordered by a BFS from a peripheral node and cut at the median of that
// order into two halves, each ordered recursively; components of up to LEAF
// nodes are kept whole, after the large ones.  Most sets are connected and
// come with a peripheral node from their parent's order, so one BFS per set
// and level does.  Halves are ordered in place on up to 16 host threads.
// Deterministic: every choice follows the order of the ids and of the
// adjacency lists.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <future>
#include <thread>
#include <vector>

namespace dlg {

namespace {

constexpr uint32_t LEAF = 32;           // ids per 32-bit word of a bitmap row
constexpr size_t PAR_MIN = 1u << 13;    // sets at least this large may go to another thread

struct Ctx {
    const uint64_t *off;
    const uint32_t *adj;
    std::atomic<uint32_t> *mark;   // the set each node is in (sets ordered concurrently are disjoint)
    std::atomic<uint32_t> *vis;    // the last BFS that reached each node
    std::atomic<uint32_t> next_stamp{1};
    std::atomic<int> threads{1};
    int max_threads = 1;
};

// BFS over the nodes of `set` from s, appending to q (which also serves as
// the queue); returns the new end.
size_t bfs(Ctx &c, uint32_t set, uint32_t stamp, uint32_t s, uint32_t *q, size_t end) {
    size_t h = end;
    q[end++] = s;
    c.vis[s].store(stamp, std::memory_order_relaxed);
    for (; h < end; h++) {
        const uint32_t v = q[h];
        for (uint64_t e = c.off[v]; e < c.off[v + 1]; e++) {
            const uint32_t w = c.adj[e];
            if (c.mark[w].load(std::memory_order_relaxed) != set) continue;
            if (c.vis[w].load(std::memory_order_relaxed) == stamp) continue;
            c.vis[w].store(stamp, std::memory_order_relaxed);
            q[end++] = w;
        }
    }
    return end;
}

// Orders the set nodes[0 .. n) in place; tmp[0 .. n) is scratch.  `hint`,
// when not NONE, is a node of the set at its periphery: a BFS from it orders
// a connected set directly.  Halves of a BFS order from a peripheral node
// come with their own peripheral node: the first half (a ball around the
// start) its last node, the second half the order's last node.
constexpr uint32_t NONE = 0xffffffffu;

void order_span(Ctx &c, uint32_t *nodes, size_t n, uint32_t *tmp, uint32_t hint) {
    if (n <= LEAF) return;
    const uint32_t set = c.next_stamp.fetch_add(1);
    for (size_t i = 0; i < n; i++) c.mark[nodes[i]].store(set, std::memory_order_relaxed);
    struct Half {
        size_t at, n;
        uint32_t hint;
    };
    std::vector<Half> halves;
    const uint32_t first = c.next_stamp.fetch_add(1);
    size_t end = bfs(c, set, first, hint != NONE ? hint : nodes[0], tmp, 0);
    if (end == n && hint != NONE) {   // connected, ordered from the periphery
        std::copy(tmp, tmp + n, nodes);
        halves.push_back({0, n / 2, nodes[n / 2 - 1]});
        halves.push_back({n / 2, n - n / 2, nodes[n - 1]});
    } else {
        // connected components (each in BFS order from its first node) into tmp
        std::vector<std::pair<size_t, size_t>> comps{{0, end}};
        for (size_t i = 0; i < n; i++) {
            if (c.vis[nodes[i]].load(std::memory_order_relaxed) == first) continue;
            const size_t b = end;
            end = bfs(c, set, first, nodes[i], tmp, end);
            comps.push_back({b, end});
        }
        // large components: BFS from their far end (the last node the first
        // BFS reached), halved at the median; small ones after them
        size_t at = 0;
        for (auto &cp : comps) {
            const size_t sz = cp.second - cp.first;
            if (sz <= LEAF) continue;
            const uint32_t far = tmp[cp.second - 1], far_stamp = c.next_stamp.fetch_add(1);
            bfs(c, set, far_stamp, far, nodes + at, 0);
            halves.push_back({at, sz / 2, nodes[at + sz / 2 - 1]});
            halves.push_back({at + sz / 2, sz - sz / 2, nodes[at + sz - 1]});
            at += sz;
        }
        for (auto &cp : comps)
            if (cp.second - cp.first <= LEAF)
                for (size_t k = cp.first; k < cp.second; k++) nodes[at++] = tmp[k];
    }
    std::vector<std::future<void>> jobs;
    for (const Half &h : halves) {
        if (h.n >= PAR_MIN && c.threads.fetch_add(1) < c.max_threads) {
            jobs.push_back(std::async(std::launch::async, [&c, h, nodes, tmp]() {
                order_span(c, nodes + h.at, h.n, tmp + h.at, h.hint);
                c.threads.fetch_sub(1);
            }));
        } else {
            if (h.n >= PAR_MIN) c.threads.fetch_sub(1);
            order_span(c, nodes + h.at, h.n, tmp + h.at, h.hint);
        }
    }
    for (auto &j : jobs) j.get();
}

}  // namespace

// perm[new id] = old id over the local ids 0 .. W-1 of the edges (src, dst)
// given as local ids; returns the inverse, sigma[old] = new.
std::vector<uint32_t> locality_order(const std::vector<uint32_t> &src, const std::vector<uint32_t> &dst, uint32_t W) {
    std::vector<uint64_t> off((size_t)W + 1, 0);
    for (size_t i = 0; i < src.size(); i++) {
        if (src[i] == dst[i]) continue;
        off[src[i] + 1]++;
        off[dst[i] + 1]++;
    }
    for (uint32_t v = 0; v < W; v++) off[v + 1] += off[v];
    std::vector<uint32_t> adj(off[W]);
    {
        std::vector<uint64_t> pos(off.begin(), off.end() - 1);
        for (size_t i = 0; i < src.size(); i++) {
            if (src[i] == dst[i]) continue;
            adj[pos[src[i]]++] = dst[i];
            adj[pos[dst[i]]++] = src[i];
        }
        // each neighbour once (a symmetric input lists every edge twice here)
        uint64_t k = 0;
        for (uint32_t v = 0; v < W; v++) {
            const uint64_t b = off[v], e = off[v + 1];
            std::sort(adj.begin() + (ptrdiff_t)b, adj.begin() + (ptrdiff_t)e);
            off[v] = k;
            for (uint64_t j = b; j < e; j++)
                if (j == b || adj[j] != adj[j - 1]) adj[k++] = adj[j];
        }
        off[W] = k;
        adj.resize(k);
    }
    std::vector<std::atomic<uint32_t>> mark(W), vis(W);
    for (uint32_t v = 0; v < W; v++) {
        mark[v].store(0, std::memory_order_relaxed);
        vis[v].store(0, std::memory_order_relaxed);
    }
    Ctx c;
    c.off = off.data();
    c.adj = adj.data();
    c.mark = mark.data();
    c.vis = vis.data();
    unsigned hw = std::thread::hardware_concurrency();
    c.max_threads = (int)std::min<unsigned>(hw ? hw : 1, 16);
    std::vector<uint32_t> perm(W), tmp(W);
    for (uint32_t v = 0; v < W; v++) perm[v] = v;
    order_span(c, perm.data(), W, tmp.data(), NONE);
    std::vector<uint32_t> sigma(W);
    for (uint32_t k = 0; k < W; k++) sigma[perm[k]] = k;
    return sigma;
}

}  // namespace dlg
