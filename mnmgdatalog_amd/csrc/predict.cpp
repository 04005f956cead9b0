// predict.cpp -- size of a transitive closure before running it (host side of
// libdlg.so), for the served form's admission and the batched workloads'
// output caps.
//
// SCC condensation: within a strongly connected component every node reaches
// the same nodes, so with R*(C) = C plus every node reachable from C,
//     |TC| = sum over components C of |C| * (|R*(C)| - (C reaches itself ? 0 : |C|))
// where C reaches itself when it has two or more nodes or a self-loop (tc.cu's
// closure holds (x, x) exactly for the x on a cycle).  Components come out of
// Tarjan's algorithm sinks first, so R*(C) = C u R*(successors), each a
// finished bitset over components when it is needed.  Above BITSET_MAX_COMPS
// components the bitsets would not fit; then |R*(C)| is found by a search of
// the condensation from components drawn by size, and the sum is estimated.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/dlg.h"

namespace {

constexpr uint32_t BITSET_MAX_COMPS = 16384;   // k^2 bits = 32 MiB of reach sets
constexpr int SAMPLES = 4096;
constexpr uint64_t SAMPLE_WORK = 200000000ull;   // condensation steps over all samples

struct Graph {
    uint32_t V = 0;
    std::vector<uint64_t> off;   // CSR over compacted ids
    std::vector<uint32_t> dst;
    std::vector<uint8_t> self_loop;
};

Graph build(const int32_t *rows, int64_t n) {
    Graph g;
    std::vector<int32_t> ids(rows, rows + 2 * n);
    std::sort(ids.begin(), ids.end());
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    g.V = (uint32_t)ids.size();
    auto id = [&](int32_t x) { return (uint32_t)(std::lower_bound(ids.begin(), ids.end(), x) - ids.begin()); };
    std::vector<uint32_t> src(n);
    g.off.assign((size_t)g.V + 1, 0);
    g.self_loop.assign(g.V, 0);
    for (int64_t i = 0; i < n; i++) {
        src[i] = id(rows[2 * i]);
        g.off[src[i] + 1]++;
    }
    for (uint32_t v = 0; v < g.V; v++) g.off[v + 1] += g.off[v];
    g.dst.resize(n);
    std::vector<uint64_t> pos(g.off.begin(), g.off.end() - 1);
    for (int64_t i = 0; i < n; i++) {
        const uint32_t d = id(rows[2 * i + 1]);
        g.dst[pos[src[i]]++] = d;
        if (d == src[i]) g.self_loop[d] = 1;
    }
    return g;
}

// Tarjan's algorithm without recursion; comp[v] numbers components in order
// of completion (every component completes after those it reaches).
uint32_t tarjan(const Graph &g, std::vector<uint32_t> &comp) {
    const uint32_t NONE = 0xffffffffu;
    std::vector<uint32_t> index(g.V, NONE), low(g.V, 0), stack;
    std::vector<uint8_t> on(g.V, 0);
    std::vector<std::pair<uint32_t, uint64_t>> call;   // node, next edge
    comp.assign(g.V, NONE);
    uint32_t next = 0, k = 0;
    for (uint32_t s = 0; s < g.V; s++) {
        if (index[s] != NONE) continue;
        call.push_back({s, g.off[s]});
        index[s] = low[s] = next++;
        stack.push_back(s);
        on[s] = 1;
        while (!call.empty()) {
            const uint32_t v = call.back().first;
            uint64_t &e = call.back().second;
            if (e < g.off[v + 1]) {
                const uint32_t w = g.dst[e++];
                if (index[w] == NONE) {
                    index[w] = low[w] = next++;
                    stack.push_back(w);
                    on[w] = 1;
                    call.push_back({w, g.off[w]});
                } else if (on[w]) {
                    low[v] = std::min(low[v], index[w]);
                }
                continue;
            }
            if (low[v] == index[v]) {
                uint32_t w;
                do {
                    w = stack.back();
                    stack.pop_back();
                    on[w] = 0;
                    comp[w] = k;
                } while (w != v);
                k++;
            }
            call.pop_back();
            if (!call.empty()) {
                const uint32_t u = call.back().first;
                low[u] = std::min(low[u], low[v]);
            }
        }
    }
    return k;
}

}  // namespace

extern "C" dlg_status dlg_predict_tc_size(const int32_t *h_rows, int64_t n, int64_t *tc_size, int32_t *exact) {
    if (!tc_size || n < 0 || (n && !h_rows) || n >= 0x7fffffffLL) return DLG_INVALID_INPUT;
    *tc_size = 0;
    if (exact) *exact = 1;
    if (n == 0) return DLG_OK;
    try {
        const Graph g = build(h_rows, n);
        std::vector<uint32_t> comp;
        const uint32_t k = tarjan(g, comp);
        std::vector<uint64_t> size(k, 0);
        std::vector<uint8_t> self(k, 0);
        for (uint32_t v = 0; v < g.V; v++) {
            size[comp[v]]++;
            if (g.self_loop[v]) self[comp[v]] = 1;
        }
        for (uint32_t c = 0; c < k; c++)
            if (size[c] > 1) self[c] = 1;
        // condensation edges, grouped by source component, deduplicated
        std::vector<uint64_t> coff((size_t)k + 1, 0);
        std::vector<std::pair<uint32_t, uint32_t>> ce;
        ce.reserve(g.dst.size());
        for (uint32_t v = 0; v < g.V; v++)
            for (uint64_t e = g.off[v]; e < g.off[v + 1]; e++)
                if (comp[g.dst[e]] != comp[v]) ce.push_back({comp[v], comp[g.dst[e]]});
        std::sort(ce.begin(), ce.end());
        ce.erase(std::unique(ce.begin(), ce.end()), ce.end());
        for (auto &p : ce) coff[p.first + 1]++;
        for (uint32_t c = 0; c < k; c++) coff[c + 1] += coff[c];
        auto reached = [&](uint32_t c, uint64_t rstar) { return rstar - (self[c] ? 0 : size[c]); };
        long double total = 0.0L;
        if (k <= BITSET_MAX_COMPS) {
            const uint32_t W = (k + 63) / 64;
            std::vector<uint64_t> bits((size_t)k * W, 0);
            std::vector<uint32_t> big;   // components of more than one node: counted by size
            for (uint32_t c = 0; c < k; c++)
                if (size[c] > 1) big.push_back(c);
            for (uint32_t c = 0; c < k; c++) {   // successors have smaller numbers: already done
                uint64_t *b = &bits[(size_t)c * W];
                b[c >> 6] |= 1ull << (c & 63);
                for (uint64_t e = coff[c]; e < coff[c + 1]; e++) {
                    const uint64_t *s = &bits[(size_t)ce[e].second * W];
                    for (uint32_t w = 0; w < W; w++) b[w] |= s[w];
                }
                uint64_t rstar = 0;
                for (uint32_t w = 0; w < W; w++) rstar += (uint64_t)__builtin_popcountll(b[w]);
                for (uint32_t x : big)
                    if ((b[x >> 6] >> (x & 63)) & 1ull) rstar += size[x] - 1;
                total += (long double)size[c] * (long double)reached(c, rstar);
            }
        } else {
            if (exact) *exact = 0;
            // components drawn by size (node ids by a fixed multiplicative stride),
            // each searched once; the sum is V times the mean over the draws
            std::vector<uint64_t> memo(k, ~0ull);
            std::vector<uint32_t> stamp(k, 0), queue;
            uint64_t work = 0;
            long double acc = 0.0L;
            int drawn = 0;
            for (int s = 0; s < SAMPLES && (work < SAMPLE_WORK || drawn < 32); s++) {
                const uint32_t v = (uint32_t)(((uint64_t)s * 0x9E3779B97F4A7C15ull >> 17) % g.V);
                const uint32_t c = comp[v];
                if (memo[c] == ~0ull) {
                    uint64_t rstar = 0;
                    queue.assign(1, c);
                    stamp[c] = (uint32_t)s + 1;
                    for (size_t h = 0; h < queue.size(); h++) {
                        const uint32_t x = queue[h];
                        rstar += size[x];
                        for (uint64_t e = coff[x]; e < coff[x + 1]; e++) {
                            const uint32_t y = ce[e].second;
                            if (stamp[y] != (uint32_t)s + 1) {
                                stamp[y] = (uint32_t)s + 1;
                                queue.push_back(y);
                            }
                        }
                        work += 1 + (coff[x + 1] - coff[x]);
                    }
                    memo[c] = reached(c, rstar);
                }
                acc += (long double)memo[c];
                drawn++;
            }
            total = acc / (long double)drawn * (long double)g.V;
        }
        *tc_size = total > 9.0e18L ? (int64_t)9000000000000000000LL : (int64_t)(total + 0.5L);
    } catch (const std::bad_alloc &) {
        return DLG_OUT_OF_MEMORY;
    }
    return DLG_OK;
}
