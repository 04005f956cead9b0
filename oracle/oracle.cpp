/*
 * oracle.cpp -- CPU restatement of the MNMGDatalog semi-naive fixpoint path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the thing measured or shipped):
 * loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * It restates the reference ALGORITHM step by step with std::sort / unique /
 * set_difference / merge on packed 64-bit tuples, the same relational steps the
 * reference runs with Thrust on the GPU:
 *   - tuple order: signed lexicographic (key, value)   common/utils.cu:115-124 (set_cmp)
 *   - join:        (ht.value, probe.value) for every ht slot with ht.key == probe.key,
 *                  a multiset (duplicate relation rows give duplicate join rows)
 *                                                     common/join.cu:1-48, common/hash_table.cu:1-55
 *   - TC loop:     tc.cu:132-314      - SG loop: sg.cu:117-390
 *   - WCC loop:    wcc.cu:140-485     - single join: single_join.cu:100-283
 *   - generator:   common/parallel_io.cu:79-108 (glibc srand/rand)
 *
 * Pinning (each function is checked by tests/test_oracle.py against something
 * the reference produced or the mathematics fixes, not against a retyping):
 *   orc_tc           README.md:37-45 counts (hipc, data_10, OL, TG), data/data_5_TC.txt,
 *                    data/dummy_TC.txt tuple sets, chain closed form (data_22: 253 in 22),
 *                    SF.cedge 287 / 80,498,014 (logs/small-merged-job-output.txt:108),
 *                    brute-force closure on random small graphs
 *   orc_sg           README.md:56-58 (hipc, data_10, OL), ego-Facebook 13 / 15,018,986
 *                    (logs/small-merged-job-output.txt:269), brute-force definition on small graphs
 *   orc_wcc          data/cc_process.txt (paper: 4 iterations, 2 CC, largest 3),
 *                    README.md:67 loc-Brightkite 12 / 547 / 56,739,
 *                    rough/README.md:446-447 SF.cedge 306 / 17 / 115,921, union-find on
 *                    graphs where the termination rule does not cut the run short
 *   orc_single_join  logs/single-join-small-job-output.txt:34 (10M rows, range 1M -> 99,998,429),
 *                    brute force on small relations
 *   orc_generate_rand_rows  pinned through the single-join count above (it is glibc rand()).
 * The reference itself cannot be built here (CUDA, Thrust/CUDA and MPI are absent).
 */
#include "oracle.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

namespace {

typedef std::vector<uint64_t> Rel;

inline uint64_t pack(int32_t k, int32_t v) {
    return (uint64_t((uint32_t)k ^ 0x80000000u) << 32) | uint64_t((uint32_t)v ^ 0x80000000u);
}
inline int32_t key_of(uint64_t t) { return (int32_t)((uint32_t)(t >> 32) ^ 0x80000000u); }
inline int32_t val_of(uint64_t t) { return (int32_t)((uint32_t)t ^ 0x80000000u); }

/* Sort with up to `threads` workers: chunked std::sort then pairwise merges. */
void psort(Rel &a, int threads) {
    size_t n = a.size();
    if (threads <= 1 || n < (1u << 16)) {
        std::sort(a.begin(), a.end());
        return;
    }
    int t = threads;
    std::vector<size_t> cut(t + 1);
    for (int i = 0; i <= t; i++) cut[i] = n * (size_t)i / (size_t)t;
    {
        std::vector<std::thread> w;
        for (int i = 0; i < t; i++)
            w.emplace_back([&, i] { std::sort(a.begin() + cut[i], a.begin() + cut[i + 1]); });
        for (auto &x : w) x.join();
    }
    for (int width = 1; width < t; width *= 2) {
        std::vector<std::thread> w;
        for (int i = 0; i + width < t; i += 2 * width) {
            size_t lo = cut[i], mid = cut[i + width], hi = cut[std::min(i + 2 * width, t)];
            w.emplace_back([&a, lo, mid, hi] {
                std::inplace_merge(a.begin() + lo, a.begin() + mid, a.begin() + hi);
            });
        }
        for (auto &x : w) x.join();
    }
}

void sort_unique(Rel &a, int threads) {
    psort(a, threads);
    a.erase(std::unique(a.begin(), a.end()), a.end());
}

/* Index of a relation for the join: all rows sorted, so equal keys are
 * contiguous.  Emits the same multiset as the reference's open-addressed
 * multimap (one output per slot whose key equals the probe key). */
struct Index {
    Rel rows;
    explicit Index(const Rel &r) : rows(r) { std::sort(rows.begin(), rows.end()); }
    std::pair<size_t, size_t> range(int32_t k) const {
        uint64_t lo = pack(k, INT32_MIN), hi = pack(k, INT32_MAX);
        size_t a = std::lower_bound(rows.begin(), rows.end(), lo) - rows.begin();
        size_t b = std::upper_bound(rows.begin(), rows.end(), hi) - rows.begin();
        return {a, b};
    }
};

/* J = {(ht.value, probe.value)}  (common/join.cu:38-41) */
Rel join(const Index &ht, const Rel &probe, int threads, int64_t *raw) {
    size_t n = probe.size();
    int t = (threads > 1 && n > 4096) ? threads : 1;
    std::vector<Rel> part(t);
    auto work = [&](int w) {
        size_t lo = n * (size_t)w / (size_t)t, hi = n * (size_t)(w + 1) / (size_t)t;
        Rel &o = part[w];
        for (size_t i = lo; i < hi; i++) {
            auto r = ht.range(key_of(probe[i]));
            int32_t pv = val_of(probe[i]);
            for (size_t j = r.first; j < r.second; j++) o.push_back(pack(val_of(ht.rows[j]), pv));
        }
    };
    if (t == 1) {
        work(0);
    } else {
        std::vector<std::thread> w;
        for (int i = 0; i < t; i++) w.emplace_back(work, i);
        for (auto &x : w) x.join();
    }
    Rel out;
    size_t total = 0;
    for (auto &p : part) total += p.size();
    out.reserve(total);
    for (auto &p : part) out.insert(out.end(), p.begin(), p.end());
    if (raw) *raw += (int64_t)total;
    return out;
}

Rel swap_all(const Rel &a) {
    Rel o(a.size());
    for (size_t i = 0; i < a.size(); i++) o[i] = pack(val_of(a[i]), key_of(a[i]));
    return o;
}

Rel from_rows(const int32_t *rows, int64_t n, bool reversed) {
    Rel r((size_t)n);
    for (int64_t i = 0; i < n; i++)
        r[i] = reversed ? pack(rows[2 * i + 1], rows[2 * i]) : pack(rows[2 * i], rows[2 * i + 1]);
    return r;
}

int32_t *emit(const Rel &f, bool swap_out) {
    int32_t *o = (int32_t *)malloc(sizeof(int32_t) * 2 * (f.size() ? f.size() : 1));
    for (size_t i = 0; i < f.size(); i++) {
        int32_t k = key_of(f[i]), v = val_of(f[i]);
        o[2 * i] = swap_out ? v : k;
        o[2 * i + 1] = swap_out ? k : v;
    }
    return o;
}

/* F <- F u N for disjoint sorted F, N; N <- U \ F.  (tc.cu:257-283) */
void diff_and_merge(Rel &U, Rel &F) {
    Rel N;
    N.reserve(U.size());
    std::set_difference(U.begin(), U.end(), F.begin(), F.end(), std::back_inserter(N));
    Rel M(F.size() + N.size());
    std::merge(F.begin(), F.end(), N.begin(), N.end(), M.begin());
    F.swap(M);
    U.swap(N);
}

int tc_impl(const int32_t *rows, int64_t n, int threads, int32_t max_it, orc_stats *st,
            int32_t **out) {
    if (n < 0 || (n > 0 && !rows)) return 1;
    orc_stats s{};
    Rel edge = from_rows(rows, n, false);   /* keyed by src, NOT deduplicated (tc.cu:210) */
    Rel delta = from_rows(rows, n, true);   /* (dst, src)  (tc.cu:132-133)                */
    sort_unique(delta, threads);            /* tc.cu:177-181                               */
    Rel full = delta;                       /* tc.cu:186-189                               */
    Index ht(edge);
    while (true) {
        Rel j = edge.empty() ? Rel() : join(ht, delta, threads, &s.raw_join_rows);
        sort_unique(j, threads);
        size_t before = full.size();
        diff_and_merge(j, full);
        delta.swap(j);
        s.iterations++;
        if (full.size() == before) break;                /* tc.cu:294-300 */
        if (max_it > 0 && s.iterations >= max_it) break;
    }
    s.output_size = (int64_t)full.size();
    if (st) *st = s;
    if (out) *out = emit(full, true);       /* reverse_t_full: (src,dst) rows */
    return 0;
}

int sg_impl(const int32_t *rows, int64_t n, int threads, int32_t max_it, orc_stats *st,
            int32_t **out) {
    if (n < 0 || (n > 0 && !rows)) return 1;
    orc_stats s{};
    Rel edge = from_rows(rows, n, false);
    sort_unique(edge, threads);             /* sg.cu:141-156, also the hash-table relation */
    Index ht(edge);
    Rel base = edge.empty() ? Rel() : join(ht, edge, threads, &s.raw_join_rows);  /* (y, x) */
    base.erase(std::remove_if(base.begin(), base.end(),
                              [](uint64_t t) { return key_of(t) == val_of(t); }),
               base.end());                 /* sg.cu:175-182 */
    sort_unique(base, threads);
    Rel full = base, delta = base;
    while (true) {
        Rel tmp = edge.empty() ? Rel() : swap_all(join(ht, delta, threads, &s.raw_join_rows));
        sort_unique(tmp, threads);          /* tmp(b, x), sg.cu:248-290 */
        Rel j = edge.empty() ? Rel() : swap_all(join(ht, tmp, threads, &s.raw_join_rows));
        sort_unique(j, threads);            /* (x, y), sg.cu:295-338 */
        size_t before = full.size();
        diff_and_merge(j, full);
        delta.swap(j);
        s.iterations++;
        if (full.size() == before) break;   /* sg.cu:376-389 */
        if (max_it > 0 && s.iterations >= max_it) break;
    }
    s.output_size = (int64_t)full.size();
    if (st) *st = s;
    if (out) *out = emit(full, false);
    return 0;
}

/* Symmetric deduplicated edge relation and the initial cc = {(n, n)}. */
void wcc_setup(const int32_t *rows, int64_t n, Rel &edge, Rel &cc) {
    edge.resize(2 * (size_t)n);
    for (int64_t i = 0; i < n; i++) {
        edge[i] = pack(rows[2 * i], rows[2 * i + 1]);
        edge[n + i] = pack(rows[2 * i + 1], rows[2 * i]);
    }
    sort_unique(edge, 1);                   /* wcc.cu:140-170 */
    cc.clear();
    for (size_t i = 0; i < edge.size(); i++) {
        int32_t k = key_of(edge[i]);
        if (cc.empty() || key_of(cc.back()) != k) cc.push_back(pack(k, k));   /* wcc.cu:203-224 */
    }
}

/* One WCC round: returns the new delta; cc is replaced by the per-node minimum. */
Rel wcc_round(const Index &ht, const Rel &delta, Rel &cc, int threads, int64_t *raw) {
    Rel j = ht.rows.empty() ? Rel() : join(ht, delta, threads, raw);   /* (m, l)       */
    sort_unique(j, threads);                                           /* wcc.cu:285-294 */
    Rel merged(cc.size() + j.size());
    std::merge(cc.begin(), cc.end(), j.begin(), j.end(), merged.begin());   /* wcc.cu:302-312 */
    Rel next;
    next.reserve(cc.size());
    for (size_t i = 0; i < merged.size(); i++)   /* unique by key keeps the first = min label */
        if (next.empty() || key_of(next.back()) != key_of(merged[i])) next.push_back(merged[i]);
    Rel d;
    std::set_difference(next.begin(), next.end(), cc.begin(), cc.end(), std::back_inserter(d));
    cc.swap(next);
    return d;
}

}  // namespace

extern "C" {

int orc_tc(const int32_t *rows, int64_t n, int threads, orc_stats *st, int32_t **out) {
    return tc_impl(rows, n, threads, 0, st, out);
}
int orc_tc_bounded(const int32_t *rows, int64_t n, int threads, int32_t max_it, orc_stats *st,
                   int32_t **out) {
    return tc_impl(rows, n, threads, max_it, st, out);
}
int orc_sg(const int32_t *rows, int64_t n, int threads, orc_stats *st, int32_t **out) {
    return sg_impl(rows, n, threads, 0, st, out);
}
int orc_sg_bounded(const int32_t *rows, int64_t n, int threads, int32_t max_it, orc_stats *st,
                   int32_t **out) {
    return sg_impl(rows, n, threads, max_it, st, out);
}

int orc_wcc(const int32_t *rows, int64_t n, int threads, orc_stats *st, int32_t **out) {
    if (n < 0 || (n > 0 && !rows)) return 1;
    orc_stats s{};
    Rel edge, cc;
    wcc_setup(rows, n, edge, cc);
    s.nodes = (int64_t)cc.size();
    Index ht(edge);
    Rel delta = cc;
    int64_t prev = (int64_t)cc.size();      /* wcc.cu:236-242 */
    while (true) {
        delta = wcc_round(ht, delta, cc, threads, &s.raw_join_rows);
        int64_t cur = (int64_t)delta.size();
        s.iterations++;
        if (cur == prev) break;             /* wcc.cu:366-383: equal consecutive delta sizes */
        prev = cur;
    }
    /* labels -> (label, 0), sorted; component sizes and distinct labels (wcc.cu:389-472) */
    std::vector<int32_t> labels(cc.size());
    for (size_t i = 0; i < cc.size(); i++) labels[i] = val_of(cc[i]);
    std::sort(labels.begin(), labels.end());
    Rel rowsout;
    int64_t largest = -1;                   /* thrust::reduce(..., -1, maximum) on no components */
    for (size_t i = 0; i < labels.size();) {
        size_t j = i;
        while (j < labels.size() && labels[j] == labels[i]) j++;
        largest = std::max<int64_t>(largest, (int64_t)(j - i));
        rowsout.push_back(pack(labels[i], 0));
        i = j;
    }
    s.output_size = (int64_t)rowsout.size();
    s.secondary = largest;
    if (st) *st = s;
    if (out) *out = emit(rowsout, false);
    return 0;
}

int orc_wcc_fixpoint(const int32_t *rows, int64_t n, orc_stats *st, int32_t **labels_out) {
    if (n < 0 || (n > 0 && !rows)) return 1;
    orc_stats s{};
    Rel edge, cc;
    wcc_setup(rows, n, edge, cc);
    s.nodes = (int64_t)cc.size();
    Index ht(edge);
    Rel delta = cc;
    while (true) {
        delta = wcc_round(ht, delta, cc, 1, &s.raw_join_rows);
        s.iterations++;
        if (delta.empty()) break;
    }
    std::vector<int32_t> labels(cc.size());
    for (size_t i = 0; i < cc.size(); i++) labels[i] = val_of(cc[i]);
    std::vector<int32_t> sorted = labels;
    std::sort(sorted.begin(), sorted.end());
    int64_t comps = 0, largest = -1;
    for (size_t i = 0; i < sorted.size();) {
        size_t j = i;
        while (j < sorted.size() && sorted[j] == sorted[i]) j++;
        comps++;
        largest = std::max<int64_t>(largest, (int64_t)(j - i));
        i = j;
    }
    s.output_size = comps;
    s.secondary = largest;
    if (st) *st = s;
    if (labels_out) {
        /* rows (node, label) in node order */
        int32_t *o = (int32_t *)malloc(sizeof(int32_t) * 2 * (cc.size() ? cc.size() : 1));
        for (size_t i = 0; i < cc.size(); i++) {
            o[2 * i] = key_of(cc[i]);
            o[2 * i + 1] = val_of(cc[i]);
        }
        *labels_out = o;
    }
    return 0;
}

int orc_single_join(const int32_t *rows, int64_t n, int threads, orc_stats *st, int32_t **out) {
    if (n < 0 || (n > 0 && !rows)) return 1;
    orc_stats s{};
    Rel r = from_rows(rows, n, false);
    sort_unique(r, threads);                /* single_join.cu:164-168 */
    Rel rr = from_rows(rows, n, true);
    sort_unique(rr, threads);               /* single_join.cu:200-203 */
    Index ht(r);
    Rel j = r.empty() ? Rel() : join(ht, rr, threads, &s.raw_join_rows);   /* (c, a) */
    sort_unique(j, threads);                /* single_join.cu:259-264 */
    s.iterations = 1;
    s.output_size = (int64_t)j.size();
    if (st) *st = s;
    if (out) *out = emit(j, true);          /* get_reverse_int_ar_from_entity_ar: (a, c) */
    return 0;
}

int orc_generate_rand_rows(int64_t rows, int32_t range, int32_t *out) {
    if (rows < 0 || range <= 0 || (rows > 0 && !out)) return 1;
    srand(1u);                               /* seed = rank + 1 on rank 0 */
    for (int64_t i = 0; i < 2 * rows; i++) out[i] = (rand() % range) + 1;
    return 0;
}

void orc_free(void *p) { free(p); }

}  // extern "C"
