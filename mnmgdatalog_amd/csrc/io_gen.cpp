// io_gen.cpp -- input formats and synthetic workloads (host side of libdlg.so).
//
// File format (common/parallel_io.cu:5-41, binary_file_utils.py:5-33): raw
// native-endian int32 pairs, no header, rows = filesize / 8.
// Random relation of single_join (common/parallel_io.cu:79-108): glibc
// srand(rank + 1) / rand() % range + 1, rank 0 here.
// Synthetic graphs: SplitMix64 seeds a PCG32 (XSH-RR) stream; draws never go
// through std::*_distribution, so a seed gives the same graph everywhere.
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/dlg.h"

namespace {

struct Rng {
    uint64_t state, inc;
    static uint64_t splitmix(uint64_t &x) {
        uint64_t z = (x += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed) {
        uint64_t s = seed;
        state = splitmix(s);
        inc = splitmix(s) | 1ull;
        next();
    }
    uint32_t next() {  // PCG32 XSH-RR (O'Neill 2014)
        uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
        uint32_t rot = (uint32_t)(old >> 59);
        return (xs >> rot) | (xs << ((32 - rot) & 31));
    }
    uint32_t below(uint32_t bound) {  // unbiased, bound > 0
        uint32_t threshold = (uint32_t)(-bound) % bound;
        while (true) {
            uint32_t r = next();
            if (r >= threshold) return r % bound;
        }
    }
    double unit() {  // [0, 1)
        uint64_t hi = next(), lo = next();
        return (double)(((hi << 32) | lo) >> 11) * (1.0 / 9007199254740992.0);
    }
};

int32_t *to_rows(const std::vector<uint64_t> &e, int64_t *n_rows) {
    int32_t *r = (int32_t *)std::malloc(sizeof(int32_t) * 2 * (e.size() ? e.size() : 1));
    if (!r) return nullptr;
    for (size_t i = 0; i < e.size(); i++) {
        r[2 * i] = (int32_t)(uint32_t)(e[i] >> 32);
        r[2 * i + 1] = (int32_t)(uint32_t)e[i];
    }
    *n_rows = (int64_t)e.size();
    return r;
}

inline uint64_t ekey(uint32_t u, uint32_t v) { return ((uint64_t)u << 32) | v; }

inline uint64_t mix64(uint64_t x) {
    uint64_t s = x;
    return Rng::splitmix(s);
}

int host_threads(size_t n) {
    unsigned hw = std::thread::hardware_concurrency();
    size_t t = std::min<size_t>(hw ? hw : 1, 16);
    return (int)std::max<size_t>(1, std::min(t, n / (1u << 20)));
}

// Runs f(t, lo, hi) over T equal slices of [0, n) on T threads.
template <typename F>
void parallel_slices(size_t n, int T, F f) {
    if (T <= 1) {
        f(0, (size_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < T; t++) th.emplace_back([&, t] { f(t, n * t / T, n * (t + 1) / T); });
    for (auto &x : th) x.join();
}

// Sort of u64 keys: LSD radix with 8-bit digits on up to 16 host threads
// (per-thread digit counts, a prefix over (digit, thread), stable scatter);
// passes whose digit is the same for every key are skipped.
void sort_u64(std::vector<uint64_t> &e) {
    const size_t n = e.size();
    const int T = host_threads(n);
    if (T <= 1) {
        std::sort(e.begin(), e.end());
        return;
    }
    uint64_t vary = 0;
    {
        std::vector<uint64_t> o(T, 0), a(T, ~0ull);
        parallel_slices(n, T, [&](int t, size_t lo, size_t hi) {
            uint64_t x = 0, y = ~0ull;
            for (size_t i = lo; i < hi; i++) {
                x |= e[i];
                y &= e[i];
            }
            o[t] = x;
            a[t] = y;
        });
        uint64_t x = 0, y = ~0ull;
        for (int t = 0; t < T; t++) {
            x |= o[t];
            y &= a[t];
        }
        vary = x ^ y;   // bits that differ somewhere
    }
    std::vector<uint64_t> tmp(n);
    std::vector<size_t> cnt((size_t)T * 256);
    for (int shift = 0; shift < 64; shift += 8) {
        if (((vary >> shift) & 0xff) == 0) continue;
        std::fill(cnt.begin(), cnt.end(), 0);
        parallel_slices(n, T, [&](int t, size_t lo, size_t hi) {
            size_t *c = &cnt[(size_t)t * 256];
            for (size_t i = lo; i < hi; i++) c[(e[i] >> shift) & 0xff]++;
        });
        size_t sum = 0;
        for (int d = 0; d < 256; d++)
            for (int t = 0; t < T; t++) {
                size_t c = cnt[(size_t)t * 256 + d];
                cnt[(size_t)t * 256 + d] = sum;
                sum += c;
            }
        parallel_slices(n, T, [&](int t, size_t lo, size_t hi) {
            size_t *c = &cnt[(size_t)t * 256];
            for (size_t i = lo; i < hi; i++) tmp[c[(e[i] >> shift) & 0xff]++] = e[i];
        });
        e.swap(tmp);
    }
}

// Keep exactly m distinct edges out of a larger sorted-unique set: the m with
// the smallest seeded hash (a bijection of the edge, so no ties), kept in
// their sorted order.
void trim_to(std::vector<uint64_t> &e, size_t m, uint64_t seed) {
    if (e.size() <= m) return;
    const uint64_t salt = seed * 0x2545f4914f6cdd1dull;
    std::vector<uint64_t> h(e.size());
    parallel_slices(e.size(), host_threads(e.size()), [&](int, size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; i++) h[i] = mix64(e[i] ^ salt);
    });
    std::vector<uint64_t> sel(h);
    std::nth_element(sel.begin(), sel.begin() + (ptrdiff_t)(m - 1), sel.end());
    const uint64_t cut = sel[m - 1];   // the m-th smallest hash
    size_t k = 0;
    for (size_t i = 0; i < e.size(); i++)
        if (h[i] <= cut) e[k++] = e[i];
    e.resize(k);
}

void sort_unique(std::vector<uint64_t> &e) {
    sort_u64(e);
    e.erase(std::unique(e.begin(), e.end()), e.end());
}

}  // namespace

extern "C" {


dlg_status dlg_read_rows(const char *path, int32_t **h_rows, int64_t *n) {
    if (!path || !h_rows || !n) return DLG_INVALID_INPUT;
    *h_rows = nullptr;
    *n = 0;
    struct stat st;
    if (stat(path, &st) != 0) return DLG_IO_ERROR;
    FILE *f = std::fopen(path, "rb");
    if (!f) return DLG_IO_ERROR;
    int64_t rows = (int64_t)st.st_size / 8;
    int32_t *buf = (int32_t *)std::malloc(sizeof(int32_t) * 2 * (size_t)(rows ? rows : 1));
    if (!buf) {
        std::fclose(f);
        return DLG_OUT_OF_MEMORY;
    }
    size_t got = rows ? std::fread(buf, 8, (size_t)rows, f) : 0;
    std::fclose(f);
    if ((int64_t)got != rows) {
        std::free(buf);
        return DLG_IO_ERROR;
    }
    *h_rows = buf;
    *n = rows;
    return DLG_OK;
}

dlg_status dlg_write_rows(const char *path, const int32_t *h_rows, int64_t n) {
    if (!path || n < 0 || (n && !h_rows)) return DLG_INVALID_INPUT;
    FILE *f = std::fopen(path, "wb");
    if (!f) return DLG_IO_ERROR;
    size_t put = n ? std::fwrite(h_rows, 8, (size_t)n, f) : 0;
    int rc = std::fclose(f);
    return ((int64_t)put == n && rc == 0) ? DLG_OK : DLG_IO_ERROR;
}

dlg_status dlg_generate_rand_rows(int64_t rows, int32_t range, int32_t **h_rows) {
    if (rows < 0 || range <= 0 || !h_rows) return DLG_INVALID_INPUT;
    int32_t *v = (int32_t *)std::malloc(sizeof(int32_t) * 2 * (size_t)(rows ? rows : 1));
    if (!v) return DLG_OUT_OF_MEMORY;
    srand(1u);  // seed = rank + 1, rank 0
    for (int64_t i = 0; i < 2 * rows; i++) v[i] = (rand() % range) + 1;
    *h_rows = v;
    return DLG_OK;
}

// Power-law out-degrees: every vertex draws a degree from P(d) ~ d^-alpha on
// [1, n-1]; single vertices are then redrawn until the degrees sum to exactly
// m; destinations are drawn without replacement per source (uniform over the
// other n-1 vertices), so all m edges are distinct.
dlg_status dlg_gen_powerlaw(int32_t n, int64_t m, double alpha, uint64_t seed, int32_t **h_rows,
                            int64_t *n_rows) {
    if (n < 2 || m < 0 || m > (int64_t)n * (n - 1) || alpha <= 1.0 || !h_rows || !n_rows)
        return DLG_INVALID_INPUT;
    Rng rng(seed);
    int dmax = n - 1;
    std::vector<double> cdf(dmax);
    double acc = 0;
    for (int d = 1; d <= dmax; d++) {
        acc += std::pow((double)d, -alpha);
        cdf[d - 1] = acc;
    }
    for (auto &x : cdf) x /= acc;
    auto draw = [&]() {
        double u = rng.unit();
        int d = (int)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) + 1;
        return std::min(d, dmax);
    };
    std::vector<int> deg(n);
    int64_t s = 0;
    for (int u = 0; u < n; u++) {
        deg[u] = draw();
        s += deg[u];
    }
    // Resample single vertices' degrees (fresh power-law draws), keeping a
    // redraw only when the degree sum does not move away from m.
    for (int64_t step = 0; s != m && step < 100000000; step++) {
        int u = (int)rng.below((uint32_t)n);
        int d = draw();
        int64_t ns = s - deg[u] + d;
        if (std::llabs(ns - m) <= std::llabs(s - m)) {
            deg[u] = d;
            s = ns;
        }
    }
    while (s > m) {  // only reached when m < n (every vertex draws at least 1)
        int u = (int)rng.below((uint32_t)n);
        if (deg[u] > 0) { deg[u]--; s--; }
    }
    while (s < m) {
        int u = (int)rng.below((uint32_t)n);
        if (deg[u] < dmax) { deg[u]++; s++; }
    }
    std::vector<uint64_t> e;
    e.reserve((size_t)m);
    std::vector<uint32_t> pick;
    std::vector<uint8_t> used((size_t)n, 0);
    for (int u = 0; u < n; u++) {
        pick.clear();
        // sample deg[u] distinct destinations from [0, n) \ {u}
        while ((int)pick.size() < deg[u]) {
            uint32_t v = rng.below((uint32_t)(n - 1));
            if (v >= (uint32_t)u) v++;
            if (!used[v]) {
                used[v] = 1;
                pick.push_back(v);
            }
        }
        for (uint32_t v : pick) {
            used[v] = 0;
            e.push_back(ekey((uint32_t)u, v));
        }
    }
    sort_unique(e);
    *h_rows = to_rows(e, n_rows);
    return *h_rows ? DLG_OK : DLG_OUT_OF_MEMORY;
}

// Road-network-like grid: id = r * side + c.
dlg_status dlg_gen_grid(int32_t side, double p, double frac, int32_t maxd, uint64_t seed,
                        int32_t **h_rows, int64_t *n_rows) {
    if (side < 2 || p < 0 || p > 1 || frac < 0 || maxd < 1 || !h_rows || !n_rows)
        return DLG_INVALID_INPUT;
    Rng rng(seed);
    std::vector<uint64_t> e;
    int64_t kept = 0;
    for (int r = 0; r < side; r++)
        for (int c = 0; c < side; c++) {
            uint32_t u = (uint32_t)(r * side + c);
            if (c + 1 < side && rng.unit() < p) {
                e.push_back(ekey(u, u + 1));
                e.push_back(ekey(u + 1, u));
                kept++;
            }
            if (r + 1 < side && rng.unit() < p) {
                e.push_back(ekey(u, u + side));
                e.push_back(ekey(u + side, u));
                kept++;
            }
        }
    int64_t nshort = (int64_t)std::llround(frac * (double)kept);
    for (int64_t s = 0; s < nshort;) {
        int r = (int)rng.below((uint32_t)side), c = (int)rng.below((uint32_t)side);
        int dr = (int)rng.below((uint32_t)(2 * maxd + 1)) - maxd;
        int dc = (int)rng.below((uint32_t)(2 * maxd + 1)) - maxd;
        int dist = std::abs(dr) + std::abs(dc);
        if (dist < 1 || dist > maxd) continue;
        int r2 = r + dr, c2 = c + dc;
        if (r2 < 0 || r2 >= side || c2 < 0 || c2 >= side) continue;
        uint32_t u = (uint32_t)(r * side + c), v = (uint32_t)(r2 * side + c2);
        e.push_back(ekey(u, v));
        e.push_back(ekey(v, u));
        s++;
    }
    sort_unique(e);
    *h_rows = to_rows(e, n_rows);
    return *h_rows ? DLG_OK : DLG_OUT_OF_MEMORY;
}

// 5-tetrahedron split of each hex cell; corner b = (bx, by, bz) bits (1, 2, 4).
// Even cells use the tetrahedra {1,2,4,7}, {0,1,2,4}, {3,1,2,7}, {5,1,4,7},
// {6,2,4,7}; odd cells the mirror image (b ^ 1), so face diagonals match.
dlg_status dlg_gen_tetmesh(int32_t cells, int32_t **h_rows, int64_t *n_rows) {
    if (cells < 1 || cells > 1000 || !h_rows || !n_rows) return DLG_INVALID_INPUT;
    static const int tets[5][4] = {{1, 2, 4, 7}, {0, 1, 2, 4}, {3, 1, 2, 7}, {5, 1, 4, 7}, {6, 2, 4, 7}};
    const int64_t nv = cells + 1;
    std::vector<uint64_t> e;
    e.reserve((size_t)cells * cells * cells * 30);
    for (int i = 0; i < cells; i++)
        for (int j = 0; j < cells; j++)
            for (int k = 0; k < cells; k++) {
                int par = (i + j + k) & 1;
                uint32_t id[8];
                for (int b = 0; b < 8; b++) {
                    int bb = par ? (b ^ 1) : b;
                    int64_t x = i + (bb & 1), y = j + ((bb >> 1) & 1), z = k + ((bb >> 2) & 1);
                    id[b] = (uint32_t)((x * nv + y) * nv + z);
                }
                for (int t = 0; t < 5; t++)
                    for (int a = 0; a < 4; a++)
                        for (int b = a + 1; b < 4; b++) {
                            uint32_t u = id[tets[t][a]], v = id[tets[t][b]];
                            if (u > v) std::swap(u, v);
                            e.push_back(ekey(u, v));   // low -> high
                        }
            }
    sort_unique(e);
    *h_rows = to_rows(e, n_rows);
    return *h_rows ? DLG_OK : DLG_OUT_OF_MEMORY;
}

dlg_status dlg_gen_erdos_renyi(int32_t n, int64_t m, uint64_t seed, int32_t **h_rows,
                               int64_t *n_rows) {
    if (n < 2 || m < 0 || m > (int64_t)n * (n - 1) / 2 || !h_rows || !n_rows) return DLG_INVALID_INPUT;
    Rng rng(seed);
    std::vector<uint64_t> e;
    e.reserve((size_t)(m * 1.1) + 16);
    while ((int64_t)e.size() < m) {
        int64_t need = m - (int64_t)e.size();
        int64_t batch = need + need / 8 + 16;
        for (int64_t b = 0; b < batch; b++) {
            uint32_t u = rng.below((uint32_t)n), v = rng.below((uint32_t)n);
            if (u != v) e.push_back(ekey(u, v));
        }
        sort_unique(e);
    }
    trim_to(e, (size_t)m, seed);
    *h_rows = to_rows(e, n_rows);
    return *h_rows ? DLG_OK : DLG_OUT_OF_MEMORY;
}

// R-MAT (Chakrabarti, Zhan, Faloutsos 2004): each edge descends `scale`
// quadrant choices with probabilities a, b, c, 1-a-b-c.  Batches are drawn on
// up to 16 host threads, each with its own seeded stream.
dlg_status dlg_gen_rmat(int32_t scale, int64_t m, double a, double b, double c, uint64_t seed,
                        int undirected, int32_t **h_rows, int64_t *n_rows) {
    if (scale < 1 || scale > 30 || m < 0 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0 || !h_rows ||
        !n_rows)
        return DLG_INVALID_INPUT;
    double nn = std::ldexp(1.0, scale);
    if ((double)m > nn * (nn - 1) / 4) return DLG_INVALID_INPUT;
    const double ab = a + b, abc = a + b + c;
    std::vector<uint64_t> e;
    int round = 0;
    while ((int64_t)e.size() < m) {
        int64_t need = m - (int64_t)e.size();
        int64_t batch = need + need / 8 + 64;
        int T = (int)std::min<int64_t>(16, std::max<int64_t>(1, batch / 200000));
        std::vector<std::vector<uint64_t>> part(T);
        std::vector<std::thread> th;
        for (int t = 0; t < T; t++)
            th.emplace_back([&, t] {
                Rng rng(seed ^ mix64(((uint64_t)round << 20) + (uint64_t)t + 1));
                int64_t lo = batch * t / T, hi = batch * (t + 1) / T;
                auto &out = part[t];
                out.reserve((size_t)(hi - lo));
                for (int64_t q = lo; q < hi; q++) {
                    uint32_t u = 0, v = 0;
                    for (int l = 0; l < scale; l++) {
                        double r = rng.unit();
                        uint32_t bu = 0, bv = 0;
                        if (r < a) {
                        } else if (r < ab) {
                            bv = 1;
                        } else if (r < abc) {
                            bu = 1;
                        } else {
                            bu = 1;
                            bv = 1;
                        }
                        u = (u << 1) | bu;
                        v = (v << 1) | bv;
                    }
                    if (u == v) continue;
                    if (undirected && u > v) std::swap(u, v);
                    out.push_back(ekey(u, v));
                }
            });
        for (auto &x : th) x.join();
        for (auto &p : part) e.insert(e.end(), p.begin(), p.end());
        sort_unique(e);
        round++;
    }
    trim_to(e, (size_t)m, seed);
    if (undirected) {
        size_t k = e.size();
        e.resize(2 * k);
        for (size_t i = 0; i < k; i++) e[k + i] = ekey((uint32_t)e[i], (uint32_t)(e[i] >> 32));
        sort_unique(e);
    }
    *h_rows = to_rows(e, n_rows);
    return *h_rows ? DLG_OK : DLG_OUT_OF_MEMORY;
}

}  // extern "C"
