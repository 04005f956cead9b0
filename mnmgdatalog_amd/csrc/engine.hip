// engine.hip -- semi-naive fixpoint loops (TC, SG, WCC, single join) and the
// operator-level C ABI of libdlg.so.
//
// Every relation is a device array of packed u64 tuples (see dlg.h); all
// work runs on the context's stream; buffers come from the context's pool.
// The loops keep the reference's relational steps and iteration counting:
//   TC  tc.cu:132-314    Δ0 = unique(E reversed); F = Δ0; repeat
//                        { J = Δ ⋈ E; N = unique(J) \ F; F = F ∪ N; it++ } until N = ∅
//   SG  sg.cu:117-390    base = unique(E ⋈ E) minus (x,x); repeat
//                        { T = unique(Δ ⋈ E reversed); J = T ⋈ E reversed; N = unique(J) \ F; ... }
//   WCC wcc.cu:140-485   min-label propagation, stop when two consecutive |Δ| are equal
//   SJ  single_join.cu:100-283  unique((R⁻¹) ⋈ R) over deduplicated R, R⁻¹
// Full relation F: sorted array; N is found by one fused compaction
// ("first of an equal run AND not found in F"), then F' = merge(F, N).
#include "engine.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>

namespace dlg {

PhaseClock::~PhaseClock() {
    for (auto &m : marks) ctx->event_free.push_back(m.second);
}

void PhaseClock::mark(int ph) {
    // the same phase again (every round of a loop) adds nothing to the totals: no event
    if (!marks.empty() && marks.back().first == ph && ph >= 0) return;
    if (ctx->event_free.empty()) ctx->event_free.push_back(timing_event());
    hipEvent_t e = ctx->event_free.back();
    ctx->event_free.pop_back();
    HIPCHK(hipEventRecord(e, ctx->stream));
    marks.push_back({ph, e});
}

void PhaseClock::finish(double *out) {
    mark(-1);
    HIPCHK(hipEventSynchronize(marks.back().second));
    for (size_t i = 1; i < marks.size(); i++) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, marks[i - 1].second, marks[i].second));
        if (marks[i - 1].first >= 0) out[marks[i - 1].first] += ms * 1e-3;
    }
}

Opts read_opts(dlg_ctx *ctx, const dlg_opts *o) {
    Opts r{ctx->load_factor, ctx->radix_bits, 0, true, 0, 0, false, 0, 0, 0, 0};
    if (o) {
        if (o->load_factor < 0 || o->load_factor > 0.95)
            fail(DLG_INVALID_INPUT, "load_factor out of (0, 0.95]");
        if (o->radix_bits != 0 && (o->radix_bits < 4 || o->radix_bits > 8))
            fail(DLG_INVALID_INPUT, "radix_bits out of 4..8");
        if (o->strategy < 0 || o->strategy > 4) fail(DLG_INVALID_INPUT, "strategy out of 0..4");
        if (o->max_iterations < 0) fail(DLG_INVALID_INPUT, "negative max_iterations");
        if (o->block_size != 0 && (o->block_size < 64 || o->block_size > 1024 || o->block_size % 64))
            fail(DLG_INVALID_INPUT, "block_size must be 64..1024 and a multiple of 64");
        r.block_size = o->block_size;
        if (o->load_factor > 0) r.lf = o->load_factor;
        if (o->radix_bits > 0) r.bits = o->radix_bits;
        r.max_it = o->max_iterations;
        r.keep = o->keep_output != 0;
        r.check = o->check_output != 0;
        if (o->join_method < 0 || o->join_method > 1) fail(DLG_INVALID_INPUT, "join_method out of 0..1");
        r.join = o->join_method;
        r.strategy = o->strategy;
        if (o->reorder < 0 || o->reorder > 2) fail(DLG_INVALID_INPUT, "reorder out of 0..2");
        r.reorder = o->reorder;
        if (o->relax < 0 || o->relax > 2) fail(DLG_INVALID_INPUT, "relax out of 0..2");
        r.relax = o->relax;
        if (o->queue < 0 || o->queue > dlg_ctx::MAPPED_SLOTS) fail(DLG_INVALID_INPUT, "queue out of 0..8");
        r.queue = o->queue;
    }
    return r;
}

// sort + unique of buf[0..n) into out; returns distinct count.  buf is clobbered.
static u64 sort_unique_into(dlg_ctx *ctx, DBuf<u64> &buf, u64 n, DBuf<u64> &tmp, DBuf<u64> &out,
                            int mode, const u64 *other, u64 n_other, int bits,
                            const u32 *bounds = nullptr) {
    if (n == 0) return 0;
    radix_sort(ctx, buf, n, tmp, bits, bounds);
    out.ensure(n);
    return select(ctx, mode, buf.p, n, other, n_other, out.p);
}

// J = probe ⋈ index into `out` (resized); returns |J|.
static u64 join_into(dlg_ctx *ctx, const Index &ix, const u64 *probe, u64 n, int proj,
                     JoinScratch &js, DBuf<u64> &out) {
    u64 total = join_count(ctx, ix, probe, n, js);
    if (total) {
        out.ensure(total);
        join_expand(ctx, ix, probe, n, js, total, proj, out.p);
    }
    return total;
}

void download_rows(dlg_ctx *ctx, const u64 *d, u64 n, int orient, int32_t *h_rows) {
    if (!n) return;
    DBuf<int32_t> r(ctx, 2 * n);
    unpack_rows(ctx, d, n, orient, r.p);
    HIPCHK(hipMemcpyAsync(h_rows, r.p, sizeof(int32_t) * 2 * n, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
}

static void result_rows(dlg_ctx *ctx, const u64 *d, u64 n, int orient, dlg_result *res) {
    res->n_rows = (int64_t)n;
    res->rows = (int32_t *)host_block_alloc(sizeof(int32_t) * 2 * (n ? n : 1));
    if (!res->rows) fail(DLG_OUT_OF_MEMORY, "host allocation of result rows");
    download_rows(ctx, d, n, orient, res->rows);
}

static void result_digest(dlg_ctx *ctx, const u64 *d, u64 n, int orient, dlg_result *res) {
    u64 sum = 0, breaks = 0;
    rows_digest(ctx, d, n, orient, &sum, &breaks);
    res->output_sum = sum;
    res->output_order_breaks = (int64_t)breaks;
}

void check_rows_arg(int64_t n, const void *rows) {
    if (n < 0) fail(DLG_INVALID_INPUT, "negative row count");
    if (n > 0 && !rows) fail(DLG_INVALID_INPUT, "null rows with n > 0");
    if ((u64)n >= 0x7fffffffull) fail(DLG_OVERFLOW, "more than 2^31-2 input rows");
}

// The full relation F, three layouts with identical results:
//  1  one sorted array; every round's new tuples are merged into it (the
//     reference's scheme, tc.cu:268-286): each round rewrites all of F.
//  2  sorted, disjoint runs whose sizes fall geometrically from the oldest to
//     the newest; a round's new tuples become the newest run and runs of
//     similar size are merged (a binary counter), so each tuple is rewritten
//     O(log rounds) times; N = unique(J) \ F searches every run.
//  3  an append-only array plus a device hash set of F: membership costs one
//     or two slot reads however large F is, and F is sorted once at the end.
// collapse() leaves the whole relation sorted in one array.
struct FullRel {
    dlg_ctx *ctx;
    int mode;                       // 1 one array, 2 tiered runs, 3 append + hash set
    std::vector<DBuf<u64>> runs;
    std::vector<u64> sizes;
    u64 total = 0;
    // mode 3: every tuple of F in arrival order, and a hash set of them
    DBuf<u64> all;
    HashSet hs;
    double lf;
    int bits;
    const u32 *bounds;
    FullRel(dlg_ctx *c, const Opts &o, const u32 *b)
        : ctx(c), mode(o.strategy == 1 ? 1 : (o.strategy == 2 ? 2 : 3)), all(c), hs(c), lf(o.lf),
          bits(o.bits), bounds(b) {}

    void merge_last_two() {
        size_t k = runs.size();
        DBuf<u64> m(ctx, sizes[k - 2] + sizes[k - 1]);
        merge(ctx, runs[k - 2].p, sizes[k - 2], runs[k - 1].p, sizes[k - 1], m.p);
        u64 s = sizes[k - 2] + sizes[k - 1];
        runs.pop_back();
        sizes.pop_back();
        runs.back() = std::move(m);
        sizes.back() = s;
    }
    // Takes buf[0..n) (sorted, disjoint from F); buf is left empty.
    void add(DBuf<u64> &buf, u64 n) {
        if (n == 0) return;
        if (mode == 3) {
            if (all.cap < total + n) {                       // grow x2, keep the contents
                DBuf<u64> g(ctx, std::max<u64>(2 * (total + n), 1024));
                copy_u64(ctx, g.p, all.p, total);
                all = std::move(g);
            }
            copy_u64(ctx, all.p + total, buf.p, n);
            hashset_reserve(ctx, hs, total + n, lf, all.p, total);
            hashset_insert(ctx, hs, buf.p, n);
            total += n;
            buf.release();
            return;
        }
        total += n;
        runs.push_back(std::move(buf));
        sizes.push_back(n);
        buf = DBuf<u64>(ctx);
        if (mode == 1) {
            if (runs.size() == 2) merge_last_two();
            return;
        }
        while (runs.size() >= 2 &&
               (sizes[runs.size() - 2] <= 2 * sizes[runs.size() - 1] || runs.size() >= (size_t)MAX_RUNS))
            merge_last_two();
    }
    // N = first of each run of equal tuples in sorted J, minus F.
    u64 new_tuples(const u64 *J, u64 nj, u64 *N) {
        if (mode == 3) return select_unique_not_in_hash(ctx, J, nj, hs, N);
        if (runs.size() == 1) return select(ctx, SEL_UNIQUE_NOT_IN, J, nj, runs[0].p, sizes[0], N);
        RunSet rs;
        rs.k = (int)runs.size();
        for (int j = 0; j < rs.k; j++) {
            rs.p[j] = runs[j].p;
            rs.n[j] = sizes[j];
        }
        return select_unique_not_in_runs(ctx, J, nj, rs, N);
    }
    // The whole relation, sorted, moved into `out`.
    void collapse(DBuf<u64> &out) {
        if (mode == 3) {
            hs.slots.release();
            DBuf<u64> tmp(ctx);
            if (total) radix_sort(ctx, all, total, tmp, bits, bounds);
            if (!all.p) all.ensure(1);
            out = std::move(all);
            return;
        }
        while (runs.size() >= 2) merge_last_two();
        if (runs.empty()) {
            out.ensure(1);
        } else {
            out = std::move(runs[0]);
            runs.clear();
            sizes.clear();
        }
    }
};

// Every tuple of a TC / SG fixpoint pairs two node ids of the input, so the
// id range of the edge relation bounds both halves of every tuple; the radix
// sorts of the loop use it instead of a min/max pass over each join output.
static void id_bounds(dlg_ctx *ctx, const u64 *edge, u64 n, u32 *b) {
    u32 h[4];
    tuple_bounds(ctx, edge, n, h);
    u32 lo = std::min(h[0], h[2]), hi = std::max(h[1], h[3]);
    if (n == 0) lo = hi = 0;
    b[0] = lo; b[1] = hi; b[2] = lo; b[3] = hi;
}

// The bitmap layout of F (strategy 4, or automatic when the id range is
// narrow enough for a W*W-bit map).
static bool use_bitmap(const Opts &o, u64 n, const u32 *ib) {
    if (n == 0) return false;
    u64 W = (u64)(ib[1] - ib[0]) + 1;
    if (o.strategy == 4) {
        if (W > BITMAP_FORCED_MAX_W) fail(DLG_INVALID_INPUT, "bitmap layout: id range too wide");
        return true;
    }
    return o.strategy == 0 && W <= BITMAP_MAX_W;
}

// Sort J and keep its tuples that are new: N = unique(J) \ F.
static u64 new_tuples(dlg_ctx *ctx, DBuf<u64> &J, u64 nj, DBuf<u64> &tmp, DBuf<u64> &N, FullRel &F,
                      const Opts &o, PhaseClock &pc, const u32 *bounds) {
    pc.mark(DLG_PH_DEDUP);
    if (nj == 0) return 0;
    radix_sort(ctx, J, nj, tmp, o.bits, bounds);
    pc.mark(DLG_PH_MERGE);
    N.ensure(nj);
    return F.new_tuples(J.p, nj, N.p);
}

// ------------------------------------------------------------------ TC
// Round k: J = Δ ⋈ E, then Δ (the previous round's N) joins F, then
// N = unique(J) \ F.  Adding Δ to F after its join is the reference's
// "F = F ∪ N" of the previous round (tc.cu:268-286) done one step later,
// so every set difference sees the same F as the reference's.
// TC's bitmap layout: one transposed map per id range (one for a plain run).
static bool tc_use_rows(const Opts &o, u64 n, const u32 *ib, const std::vector<u32> *blocks) {
    if (n == 0 || (o.strategy != 0 && o.strategy != 4)) return false;
    u64 bits = rowmap_bits(ib[0], ib[1], blocks), wmax = (u64)(ib[1] - ib[0]) + 1;
    if (blocks) {
        wmax = 0;
        for (size_t i = 0; i + 1 < blocks->size(); i++)
            wmax = std::max<u64>(wmax, (u64)(*blocks)[i + 1] - (*blocks)[i]);
    }
    if (o.strategy == 4) {
        if (wmax > BITMAP_FORCED_MAX_W) fail(DLG_INVALID_INPUT, "bitmap layout: id range too wide");
        return true;
    }
    return wmax <= BITMAP_MAX_W && bits <= BITMAP_MAX_W * BITMAP_MAX_W;
}

// The locality order of a plain TC run in the bitmap layout (order.cpp):
// sigma (device, sigma[old - lo] = new - lo) from the rows, which are
// downloaded for it, and the renumbered rows.
static void tc_order(dlg_ctx *ctx, const int32_t *d_rows, u64 n, u32 lo, u64 W, DBuf<u32> &sigma,
                     DBuf<int32_t> &rows) {
    std::vector<int32_t> h(2 * n);
    HIPCHK(hipMemcpyAsync(h.data(), d_rows, sizeof(int32_t) * 2 * n, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    std::vector<u32> src(n), dst(n);
    for (u64 i = 0; i < n; i++) {
        src[i] = ((u32)h[2 * i] ^ 0x80000000u) - lo;
        dst[i] = ((u32)h[2 * i + 1] ^ 0x80000000u) - lo;
    }
    std::vector<u32> sg = locality_order(src, dst, (u32)W);
    sigma.ensure(W);
    HIPCHK(hipMemcpyAsync(sigma.p, sg.data(), 4 * W, hipMemcpyHostToDevice, ctx->stream));
    rows.ensure(2 * n);
    relabel_rows(ctx, d_rows, n, sigma.p, lo, rows.p);
    ctx->sync();   // the host vectors go out of scope
}

constexpr u64 ORDER_MIN_W = 1ull << 14;      // automatic locality order from this id range on,
constexpr u64 ORDER_MAX_ROWS = 1ull << 24;   // up to this many edges (host time),
constexpr double ORDER_MIN_RECIPROCAL = 0.9; // for (nearly) symmetric graphs: their closures are
                                             // complete per component, so each source's rounds
                                             // sweep rings through all of its component

void tc_fixpoint(dlg_ctx *ctx, const int32_t *d_rows, u64 n, const Opts &o, PhaseClock &pc,
                 RoundHook *hook, FixpointOut &out, const std::vector<u32> *blocks) {
    pc.mark(DLG_PH_INIT);
    DBuf<u64> edge(ctx, n), delta(ctx, n), J(ctx, n), N(ctx), tmp(ctx);
    pack_rows(ctx, d_rows, n, DLG_ORIENT_KV, edge.p);   // e(src, dst) keyed by src
    u32 ib[4];
    id_bounds(ctx, edge.p, n, ib);
    pc.mark(DLG_PH_HASHTABLE);
    Index ix(ctx);
    auto build_index = [&]() {
        if (o.join == 1) {
            // tc_nl.cu:182-185 sorts and deduplicates the edge relation it probes
            DBuf<u64> ue(ctx);
            u64 ne = sort_unique_into(ctx, edge, n, tmp, ue, SEL_UNIQUE, nullptr, 0, o.bits, ib);
            ix.rows.swap(ue);   // the index keeps the sorted rows in place
            index_build(ctx, ix, ix.rows.p, ne, o.lf, true, true);
        } else {
            index_build(ctx, ix, edge.p, n, o.lf, false);  // duplicates kept, as tc.cu:210
        }
    };
    build_index();
    // renumber into a locality order (order.cpp): plain runs in the bitmap layout
    DBuf<u32> sigma(ctx);
    DBuf<int32_t> reordered(ctx);
    const u64 W = (u64)(ib[1] - ib[0]) + 1;
    bool reorder = !blocks && !hook && o.reorder != 1 && tc_use_rows(o, n, ib, blocks);
    if (reorder && o.reorder != 2)
        reorder = W >= ORDER_MIN_W && n <= ORDER_MAX_ROWS && reciprocal_fraction(ctx, ix) >= ORDER_MIN_RECIPROCAL;
    if (reorder) {
        tc_order(ctx, d_rows, n, ib[0], W, sigma, reordered);
        d_rows = reordered.p;
        pack_rows(ctx, d_rows, n, DLG_ORIENT_KV, edge.p);
        build_index();
    }
    pack_rows(ctx, d_rows, n, DLG_ORIENT_VK, J.p);      // Δ0 rows (dst, src)
    edge.release();
    JoinScratch js(ctx);
    int it = 0;
    if (tc_use_rows(o, n, ib, blocks)) {
        // F as transposed bitmaps (row = source; one per id range) and Δ grouped
        // by source: one wave per source row and round (k_tcrows.hip).
        RowMap F(ctx);
        rowmap_init(ctx, F, ib[0], ib[1], blocks);
        J.release();
        TcRows tr(ctx);
        pc.mark(DLG_PH_DEDUP);
        tc_rows_setup(ctx, ix, F, tr);                  // Δ0 = F = unique(E reversed)
        u64 total = tr.nd, raw = 0;
        pc.mark(DLG_PH_JOIN);
        if (!hook) {
            it = tc_rows_run(ctx, F, tr, o.max_it, o.queue, &total, &raw);   // to |F| unchanged (tc.cu:294-300)
        } else {
            while (true) {
                u64 nn = tc_rows_round(ctx, F, tr, &raw);
                total += nn;
                it++;
                if (hook->by_source()) {   // the new ends by source, no tuples built
                    hook->on_round_sources(ctx, tr.ax.p, tr.alen.p, nn ? tr.nact : 0, nn, it);
                } else {
                    tc_rows_tuples(ctx, tr, N);
                    hook->on_round(ctx, N.p, nn, it, false);
                }
                if (nn == 0) break;                     // |F| unchanged (tc.cu:294-300)
                if (o.max_it > 0 && it >= o.max_it) break;
            }
        }
        pc.mark(DLG_PH_FINAL);
        out.raw += (i64)raw;
        out.nfull = total;
        if (reorder) {   // back to the input's ids
            tr = TcRows(ctx);
            ix = Index(ctx);
            rowmap_emit_renumbered(ctx, F, out.full, total, sigma.p);
        } else {
            rowmap_emit(ctx, F, out.full, total);
        }
        out.iterations = it;
        return;
    }
    pc.mark(DLG_PH_DEDUP);
    u64 nd = sort_unique_into(ctx, J, n, tmp, delta, SEL_UNIQUE, nullptr, 0, o.bits, ib);
    FullRel F(ctx, o, ib);                              // F = Δ0, added after the first join
    while (true) {
        pc.mark(DLG_PH_JOIN);
        u64 nj = join_into(ctx, ix, delta.p, nd, DLG_PROJ_HV_PV, js, J);   // (z, x)
        out.raw += (i64)nj;
        pc.mark(DLG_PH_MERGE);
        F.add(delta, nd);
        u64 nn = new_tuples(ctx, J, nj, tmp, N, F, o, pc, ib);
        it++;
        if (hook) hook->on_round(ctx, N.p, nn, it, true);
        if (nn == 0) break;                         // |F| unchanged (tc.cu:294-300)
        delta.swap(N);
        nd = nn;
        if (o.max_it > 0 && it >= o.max_it) {
            F.add(delta, nd);
            break;
        }
    }
    pc.mark(DLG_PH_MERGE);
    out.nfull = F.total;
    F.collapse(out.full);
    out.iterations = it;
}

// ------------------------------------------------------------------ SG
void sg_fixpoint(dlg_ctx *ctx, const int32_t *d_rows, u64 n, const Opts &o, PhaseClock &pc,
                 RoundHook *hook, FixpointOut &out) {
    pc.mark(DLG_PH_INIT);
    DBuf<u64> edge(ctx), J(ctx, n), T(ctx), delta(ctx), N(ctx), tmp(ctx);
    pack_rows(ctx, d_rows, n, DLG_ORIENT_KV, J.p);
    u32 ib[4];
    id_bounds(ctx, J.p, n, ib);
    pc.mark(DLG_PH_DEDUP);
    u64 ne = sort_unique_into(ctx, J, n, tmp, edge, SEL_UNIQUE, nullptr, 0, o.bits, ib);
    pc.mark(DLG_PH_HASHTABLE);
    Index ix(ctx);
    index_build(ctx, ix, edge.p, ne, o.lf, true, o.join == 1);
    JoinScratch js(ctx);
    int it = 0;
    if (use_bitmap(o, ne, ib)) {
        // F, the round's T and the round's N as bitmaps; both joins of a
        // round are row operations on them (k_sgrows.hip), Δ never leaves its
        // map D.  The base rule's rows go straight into F and D (k_bitmap.hip).
        // SG is symmetric (both rules are), so F equals its transpose and is
        // emitted without a sort.
        Bitmap F(ctx), R(ctx), D(ctx);
        bitmap_init(ctx, F, ib[0], ib[1], true);
        bitmap_init(ctx, R, ib[0], ib[1], false);
        bitmap_init(ctx, D, ib[0], ib[1], true);
        SgRows sr(ctx);
        sg_rows_setup(ctx, ix, F, sr);
        if (!ix.sorted) index_dense(ctx, ix, F.lo, F.W);   // the joins' count pass reads offsets
        u64 raw = 0;
        pc.mark(DLG_PH_JOIN);
        bitmap_join_mark(ctx, ix, edge.p, ne, js, DLG_PROJ_HV_PV, true, F, &D, &raw);   // base (y, x)
        edge.release();
        u64 total = bitmap_count(ctx, D);
        while (true) {
            pc.mark(DLG_PH_JOIN);
            const u64 nn = sg_rows_round(ctx, sr, R, F, D, &raw);   // Δ in D -> T in R -> N in D (and F)
            total += nn;
            it++;
            if (hook) {
                bitmap_drain(ctx, D, N, true, false);   // N as (x, y) tuples, D kept for the next round
                hook->on_round(ctx, N.p, nn, it, false);
            }
            if (nn == 0) break;                     // sg.cu:376-389
            if (o.max_it > 0 && it >= o.max_it) break;
        }
        pc.mark(DLG_PH_FINAL);
        out.raw += (i64)raw;
        out.nfull = total;
        bitmap_emit(ctx, F, out.full, total, true);
        out.iterations = it;
        return;
    }
    pc.mark(DLG_PH_JOIN);
    u64 nb = join_into(ctx, ix, edge.p, ne, DLG_PROJ_HV_PV, js, J);   // (y, x)
    out.raw += (i64)nb;
    edge.release();
    pc.mark(DLG_PH_DEDUP);
    u64 nd = sort_unique_into(ctx, J, nb, tmp, delta, SEL_UNIQUE_NODIAG, nullptr, 0, o.bits, ib);
    FullRel F(ctx, o, ib);                              // F = base, added after the first join
    while (true) {
        pc.mark(DLG_PH_JOIN);
        u64 n1 = join_into(ctx, ix, delta.p, nd, DLG_PROJ_PV_HV, js, J);  // (b, x)
        out.raw += (i64)n1;
        pc.mark(DLG_PH_MERGE);
        F.add(delta, nd);
        pc.mark(DLG_PH_DEDUP);
        u64 nt = sort_unique_into(ctx, J, n1, tmp, T, SEL_UNIQUE, nullptr, 0, o.bits, ib);
        pc.mark(DLG_PH_JOIN);
        u64 n2 = join_into(ctx, ix, T.p, nt, DLG_PROJ_PV_HV, js, J);      // (x, y)
        out.raw += (i64)n2;
        u64 nn = new_tuples(ctx, J, n2, tmp, N, F, o, pc, ib);
        it++;
        if (hook) hook->on_round(ctx, N.p, nn, it, true);
        if (nn == 0) break;                         // sg.cu:376-389
        delta.swap(N);
        nd = nn;
        if (o.max_it > 0 && it >= o.max_it) {
            F.add(delta, nd);
            break;
        }
    }
    pc.mark(DLG_PH_MERGE);
    out.nfull = F.total;
    F.collapse(out.full);
    out.iterations = it;
}

// ------------------------------------------------------------------ single join
// The map form (strategy 4, or automatic when the id range allows the bitmap
// layout and the W*W-bit map is no larger than the raw join's tuples): every
// output row built in LDS and read out of the map in tuple order, so the raw
// join is never materialised or sorted.
// The row form (strategy 2, or automatic when the map form is not taken):
// each output row sorted and deduplicated in LDS, for id ranges up to
// SJ_ROWS_MAX_W wide with dense offsets no more than a few per row, and rows
// of at most SJW_CAP candidates (else the sorted raw join).
constexpr u64 SJ_ROWS_MAX_W = 1ull << 27;
constexpr u64 SJ_ROWS_W_PER_ROW = 4, SJ_ROWS_W_SLACK = 1ull << 20;

static bool sj_use_map(const Opts &o, u64 W, u64 raw) {
    const u64 wpr = (W + 31) / 32;
    if (o.strategy == 4) {
        if (W > BITMAP_FORCED_MAX_W || 4 * wpr > SJ_ROW_LDS_MAX)
            fail(DLG_INVALID_INPUT, "bitmap layout: id range too wide");
        return true;
    }
    return o.strategy == 0 && o.join == 0 && W <= BITMAP_MAX_W && 4 * W * wpr <= 8 * raw;
}

void sj_fixpoint(dlg_ctx *ctx, const int32_t *d_rows, u64 n, const Opts &o, PhaseClock &pc,
                 FixpointOut &out) {
    pc.mark(DLG_PH_INIT);
    DBuf<u64> a(ctx, n), R(ctx), Rr(ctx), J(ctx), tmp(ctx);
    pack_rows(ctx, d_rows, n, DLG_ORIENT_VK, a.p);   // (b, a) for each (a, b)
    u32 ib[4];
    id_bounds(ctx, a.p, n, ib);
    pc.mark(DLG_PH_DEDUP);
    u64 nrr = sort_unique_into(ctx, a, n, tmp, Rr, SEL_UNIQUE, nullptr, 0, o.bits, ib);
    out.iterations = 1;
    if (nrr == 0) return;
    pc.mark(DLG_PH_HASHTABLE);
    const u64 W = (u64)(ib[1] - ib[0]) + 1;
    const bool try_map = o.strategy == 4 || (o.strategy == 0 && o.join == 0 && W <= BITMAP_MAX_W);
    const bool try_rows = (o.strategy == 2 || (o.strategy == 0 && o.join == 0)) && W <= SJ_ROWS_MAX_W &&
                          W <= SJ_ROWS_W_PER_ROW * nrr + SJ_ROWS_W_SLACK;
    if (try_map || try_rows) {
        Index rx(ctx);   // the (b, a) rows with dense offsets: b's run of a's
        rx.rows.swap(Rr);
        rx.n = nrr;
        if (try_rows || W <= BITMAP_FORCED_MAX_W) index_dense(ctx, rx, ib[0], W);
        const u64 raw = rx.dense_w && try_map ? self_join_raw(ctx, rx.rows.p, nrr, rx.dense.p, ib[0]) : 0;
        if (try_map && sj_use_map(o, W, raw)) {
            pc.mark(DLG_PH_JOIN);
            Bitmap F(ctx);
            bitmap_init(ctx, F, ib[0], ib[1], false, false);
            bitmap_self_join(ctx, rx.rows.p, nrr, rx.dense.p, F);   // row c: {a : (a, b), (b, c)}
            out.raw += (i64)raw;
            pc.mark(DLG_PH_DEDUP);
            out.nfull = bitmap_drain(ctx, F, out.full, false, false);   // (c, a) in order
            return;
        }
        if (o.strategy == 4) fail(DLG_INVALID_INPUT, "bitmap layout: id range too wide");
        if (try_rows && rx.dense_w) {
            pc.mark(DLG_PH_JOIN);
            u64 nout = 0, rraw = 0;
            if (self_join_rows(ctx, rx.rows.p, nrr, rx.dense.p, ib[0], W, out.full, &nout, &rraw)) {   // (c, a)
                out.raw += (i64)rraw;
                out.nfull = nout;
                return;
            }
        }
        Rr.swap(rx.rows);
    }
    pc.mark(DLG_PH_INIT);
    pack_rows(ctx, d_rows, n, DLG_ORIENT_KV, a.p);
    pc.mark(DLG_PH_DEDUP);
    u64 nr = sort_unique_into(ctx, a, n, tmp, R, SEL_UNIQUE, nullptr, 0, o.bits, ib);
    a.release();
    pc.mark(DLG_PH_HASHTABLE);
    Index ix(ctx);
    ix.rows.swap(R);
    index_build(ctx, ix, ix.rows.p, nr, o.lf, true, o.join == 1);
    pc.mark(DLG_PH_JOIN);
    JoinScratch js(ctx);
    u64 nj = join_into(ctx, ix, Rr.p, nrr, DLG_PROJ_HV_PV, js, J);   // (c, a)
    out.raw += (i64)nj;
    pc.mark(DLG_PH_DEDUP);
    out.nfull = sort_unique_into(ctx, J, nj, tmp, out.full, SEL_UNIQUE, nullptr, 0, o.bits);
}

// ------------------------------------------------------------------ WCC
// Edge lists of at least this many rows are checked for arriving sorted (one read).
constexpr u64 WCC_PRESORTED_MIN = 1ull << 20;

void wcc_graph_build(dlg_ctx *ctx, const int32_t *d_rows, u64 n, const Opts &o, PhaseClock &pc,
                     WccGraph &g) {
    pc.mark(DLG_PH_INIT);
    // E then E⁻¹ in one buffer (wcc.cu:140-160), with the input's order breaks
    // and column bounds taken in the same read
    DBuf<u64> sym(ctx, 2 * n), edge(ctx), tmp(ctx);
    u32 cb[4];
    const u64 breaks = pack_sym_rows_scan(ctx, d_rows, n, sym.p, sym.p + n, cb);
    const bool presorted = n >= WCC_PRESORTED_MIN && breaks == 0;
    pc.mark(DLG_PH_DEDUP);
    if (presorted) {
        // E arrives sorted by (src, dst): E⁻¹ in that order lists each dst's sources in
        // ascending order, so a stable sort by its key alone sorts it by (dst, src) --
        // the key's passes only -- and one merge keeping one copy of each edge gives
        // E ∪ E⁻¹ sorted and deduplicated, as the full sort + unique would
        const u32 rb[4] = {cb[2], cb[3], cb[0], cb[1]};   // E⁻¹: key b, value a
        DBuf<u64> view(ctx);   // E⁻¹ in place, borrowed from sym
        view.p = sym.p + n;
        view.cap = n;
        radix_sort(ctx, view, n, tmp, o.bits, rb, true);
        const u64 *rev = view.p;                  // in place, or in tmp's block
        if (view.p != sym.p + n) tmp.swap(view);  // tmp owns its block again
        view.p = nullptr;
        view.cap = 0;
        edge.ensure(2 * n);
        g.ne = merge_unique(ctx, sym.p, n, rev, n, edge.p);
    } else {
        const u32 lo = std::min(cb[0], cb[2]), hi = std::max(cb[1], cb[3]);
        const u32 ub[4] = {lo, hi, lo, hi};
        g.ne = sort_unique_into(ctx, sym, 2 * n, tmp, edge, SEL_UNIQUE, nullptr, 0, o.bits, ub);
    }
    sym.release();
    tmp.release();
    pc.mark(DLG_PH_HASHTABLE);
    g.ix.rows.swap(edge);   // the index keeps the sorted edges in place
    index_build(ctx, g.ix, g.ix.rows.p, g.ne, o.lf, true, true);
    // the rounds read the groups' runs; the neighbour lookup takes the rank
    // directory when it applies, else hash probes (slots built only then)
    if (o.join != 1) {
        if (index_rank_lookup_applies(ctx, g.ix)) g.ix.sorted = false;
        else index_hash(ctx, g.ix, o.lf);
    }
    g.V = g.ix.groups;
    g.nbr.ensure(g.ne ? g.ne : 1);
    index_lookup_values(ctx, g.ix, g.nbr.p);               // neighbour dense ids
}

static void run_tc(dlg_ctx *ctx, const int32_t *d_rows, u64 n, const Opts &o, dlg_result *res) {
    PhaseClock pc(ctx);
    auto t0 = Clock::now();
    FixpointOut f(ctx);
    tc_fixpoint(ctx, d_rows, n, o, pc, nullptr, f);
    pc.mark(DLG_PH_FINAL);
    ctx->sync();
    res->device_seconds = secs(t0, Clock::now());
    res->iterations = f.iterations;
    res->output_size = (int64_t)f.nfull;
    res->raw_join_rows = f.raw;
    if (o.check) result_digest(ctx, f.full.p, f.nfull, DLG_ORIENT_VK, res);
    if (o.keep) result_rows(ctx, f.full.p, f.nfull, DLG_ORIENT_VK, res);   // (src, dst)
    pc.finish(res->phase_seconds);
}

static void run_sg(dlg_ctx *ctx, const int32_t *d_rows, u64 n, const Opts &o, dlg_result *res) {
    PhaseClock pc(ctx);
    auto t0 = Clock::now();
    FixpointOut f(ctx);
    sg_fixpoint(ctx, d_rows, n, o, pc, nullptr, f);
    pc.mark(DLG_PH_FINAL);
    ctx->sync();
    res->device_seconds = secs(t0, Clock::now());
    res->iterations = f.iterations;
    res->output_size = (int64_t)f.nfull;
    res->raw_join_rows = f.raw;
    if (o.check) result_digest(ctx, f.full.p, f.nfull, DLG_ORIENT_KV, res);
    if (o.keep) result_rows(ctx, f.full.p, f.nfull, DLG_ORIENT_KV, res);   // (x, y)
    pc.finish(res->phase_seconds);
}

static void run_sj(dlg_ctx *ctx, const int32_t *d_rows, u64 n, const Opts &o, dlg_result *res) {
    PhaseClock pc(ctx);
    auto t0 = Clock::now();
    FixpointOut f(ctx);
    sj_fixpoint(ctx, d_rows, n, o, pc, f);
    pc.mark(DLG_PH_FINAL);
    ctx->sync();
    res->device_seconds = secs(t0, Clock::now());
    res->iterations = 1;
    res->output_size = (int64_t)f.nfull;
    res->raw_join_rows = f.raw;
    if (o.check) result_digest(ctx, f.full.p, f.nfull, DLG_ORIENT_VK, res);
    if (o.keep) result_rows(ctx, f.full.p, f.nfull, DLG_ORIENT_VK, res);   // (a, c)
    pc.finish(res->phase_seconds);
}

// Min-label propagation over a built graph; fills iterations, sizes and
// (with o.keep) the (label, 0) rows.  The graph is only read, so one
// resident graph can serve queries from several contexts at once.
void wcc_solve(dlg_ctx *ctx, const WccGraph &g, const Opts &o, PhaseClock &pc, dlg_result *res) {
    pc.mark(DLG_PH_INIT);
    WccState s(ctx);
    wcc_setup(ctx, s, g.V);
    int it = 0;
    const bool graph = o.strategy == 2 || (o.strategy == 0 && o.relax == WCC_AUTO && g.ne <= (1ull << 25));
    pc.mark(DLG_PH_JOIN);
    if (graph) {
        it = wcc_rounds_graph(ctx, s, g.ix, g.nbr.p, g.ne, o.max_it, 16);
    } else {
        u64 nf = g.V;
        u64 prev = g.V;                                    // wcc.cu:236-242
        while (true) {
            pc.mark(DLG_PH_JOIN);
            u64 cur = 0;
            if (nf && it == 0 && o.relax == WCC_AUTO)
                cur = wcc_first_round(ctx, s, g.ix, g.nbr.p);      // labels are still the node ids
            else if (nf)
                cur = wcc_round(ctx, s, g.ix, g.nbr.p, nf, (u32)(it + 1), o.relax);
            it++;
            nf = cur;
            if (cur == prev) break;                        // wcc.cu:366-383
            prev = cur;
            if (o.max_it > 0 && it >= o.max_it) break;
        }
    }
    pc.mark(DLG_PH_FINAL);
    DBuf<u64> out(ctx);
    i64 largest = -1;
    u64 ncomp = wcc_finalize(ctx, s, g.ix, out, &largest);
    ctx->sync();
    res->iterations = it;
    res->output_size = (int64_t)ncomp;
    res->secondary = largest;
    res->nodes = (int64_t)g.V;
    if (o.check) result_digest(ctx, out.p, ncomp, DLG_ORIENT_KV, res);
    if (o.keep) result_rows(ctx, out.p, ncomp, DLG_ORIENT_KV, res);   // (label, 0)
}

static void run_wcc(dlg_ctx *ctx, const int32_t *d_rows, u64 n, const Opts &o, dlg_result *res) {
    PhaseClock pc(ctx);
    auto t0 = Clock::now();
    WccGraph g(ctx);
    wcc_graph_build(ctx, d_rows, n, o, pc, g);
    wcc_solve(ctx, g, o, pc, res);
    res->device_seconds = secs(t0, Clock::now());
    pc.finish(res->phase_seconds);
}

}  // namespace dlg

// ====================================================================== C ABI
using namespace dlg;

typedef void (*runner_t)(dlg_ctx *, const int32_t *, u64, const Opts &, dlg_result *);

static dlg_status run_device(dlg_ctx *ctx, runner_t fn, const int32_t *d_rows, int64_t n,
                             const dlg_opts *o, dlg_result *res) {
    if (!ctx || !res) return DLG_INVALID_INPUT;
    std::memset(res, 0, sizeof(*res));
    API_BEGIN(ctx)
    check_rows_arg(n, d_rows);
    Opts op = read_opts(ctx, o);
    BlockSizeScope bss(ctx, op);
    res->input_rows = n;
    fn(ctx, d_rows, (u64)n, op, res);
    API_END(ctx)
}

static dlg_status run_host(dlg_ctx *ctx, runner_t fn, const int32_t *h_rows, int64_t n,
                           const dlg_opts *o, dlg_result *res) {
    if (!ctx || !res) return DLG_INVALID_INPUT;
    std::memset(res, 0, sizeof(*res));
    API_BEGIN(ctx)
    check_rows_arg(n, h_rows);
    Opts op = read_opts(ctx, o);
    BlockSizeScope bss(ctx, op);
    res->input_rows = n;
    auto t0 = Clock::now();
    DBuf<int32_t> d(ctx, 2 * (size_t)(n ? n : 1));
    if (n)
        HIPCHK(hipMemcpyAsync(d.p, h_rows, sizeof(int32_t) * 2 * (size_t)n, hipMemcpyHostToDevice,
                              ctx->stream));
    ctx->sync();
    double up = secs(t0, Clock::now());
    fn(ctx, d.p, (u64)n, op, res);
    res->phase_seconds[DLG_PH_INIT] += up;
    API_END(ctx)
}

extern "C" {

dlg_status dlg_run_tc(dlg_ctx *c, const int32_t *h, int64_t n, const dlg_opts *o, dlg_result *r) {
    return run_host(c, run_tc, h, n, o, r);
}
dlg_status dlg_run_sg(dlg_ctx *c, const int32_t *h, int64_t n, const dlg_opts *o, dlg_result *r) {
    return run_host(c, run_sg, h, n, o, r);
}
dlg_status dlg_run_wcc(dlg_ctx *c, const int32_t *h, int64_t n, const dlg_opts *o, dlg_result *r) {
    return run_host(c, run_wcc, h, n, o, r);
}
dlg_status dlg_run_single_join(dlg_ctx *c, const int32_t *h, int64_t n, const dlg_opts *o,
                               dlg_result *r) {
    return run_host(c, run_sj, h, n, o, r);
}
dlg_status dlg_run_tc_device(dlg_ctx *c, const int32_t *d, int64_t n, const dlg_opts *o,
                             dlg_result *r) {
    return run_device(c, run_tc, d, n, o, r);
}
dlg_status dlg_run_sg_device(dlg_ctx *c, const int32_t *d, int64_t n, const dlg_opts *o,
                             dlg_result *r) {
    return run_device(c, run_sg, d, n, o, r);
}
dlg_status dlg_run_wcc_device(dlg_ctx *c, const int32_t *d, int64_t n, const dlg_opts *o,
                              dlg_result *r) {
    return run_device(c, run_wcc, d, n, o, r);
}
dlg_status dlg_run_single_join_device(dlg_ctx *c, const int32_t *d, int64_t n, const dlg_opts *o,
                                      dlg_result *r) {
    return run_device(c, run_sj, d, n, o, r);
}

void dlg_result_free(dlg_result *r) {
    if (!r) return;
    host_block_free(r->rows);
    r->rows = nullptr;
    r->n_rows = 0;
}

dlg_status dlg_pack_rows(dlg_ctx *c, const int32_t *d_rows, int64_t n, int orient, uint64_t *d_t) {
    if (!c || n < 0 || (n && (!d_rows || !d_t)) || (orient != 0 && orient != 1))
        return DLG_INVALID_INPUT;
    API_BEGIN(c)
    pack_rows(c, d_rows, (u64)n, orient, d_t);
    c->sync();
    API_END(c)
}

dlg_status dlg_unpack_rows(dlg_ctx *c, const uint64_t *d_t, int64_t n, int orient, int32_t *d_rows) {
    if (!c || n < 0 || (n && (!d_rows || !d_t)) || (orient != 0 && orient != 1))
        return DLG_INVALID_INPUT;
    API_BEGIN(c)
    unpack_rows(c, d_t, (u64)n, orient, d_rows);
    c->sync();
    API_END(c)
}

dlg_status dlg_hash_build(dlg_ctx *c, const uint64_t *d_rel, int64_t n, double lf, dlg_hash **out) {
    if (!c || !out || n < 0 || (n && !d_rel) || lf > 0.95) return DLG_INVALID_INPUT;
    *out = nullptr;
    dlg_hash *h = nullptr;
    API_BEGIN(c)
    h = new dlg_hash(c);
    index_build(c, h->ix, d_rel, (u64)n, lf, false);
    c->sync();
    *out = h;
    h = nullptr;
    API_END_CLEANUP(c, delete h)
}

dlg_status dlg_hash_info(const dlg_hash *h, int64_t *slots, int64_t *keys, int64_t *rows) {
    if (!h) return DLG_INVALID_INPUT;
    if (slots) *slots = (int64_t)h->ix.nslots;
    if (keys) *keys = (int64_t)h->ix.groups;
    if (rows) *rows = (int64_t)h->ix.n;
    return DLG_OK;
}

void dlg_hash_free(dlg_ctx *c, dlg_hash *h) {
    if (!h) return;
    if (c) {
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
    }
    delete h;
}

dlg_status dlg_join(dlg_ctx *c, const dlg_hash *h, const uint64_t *d_probe, int64_t n, int proj,
                    uint64_t **d_out, int64_t *out_n) {
    if (!c || !h || !d_out || !out_n || n < 0 || (n && !d_probe) || proj < DLG_PROJ_HV_PV ||
        proj > DLG_PROJ_PV_HV_DROP_EQ)
        return DLG_INVALID_INPUT;
    *d_out = nullptr;
    *out_n = 0;
    API_BEGIN(c)
    JoinScratch js(c);
    u64 total = join_count(c, h->ix, d_probe, (u64)n, js);
    const bool drop_eq = proj >= DLG_PROJ_HV_PV_DROP_EQ;
    const int base = drop_eq ? proj - DLG_PROJ_HV_PV_DROP_EQ : proj;
    if (total && !drop_eq) {
        u64 *o = (u64 *)c->pool->get(total * sizeof(u64));
        join_expand(c, h->ix, d_probe, (u64)n, js, total, base, o);
        c->sync();
        *d_out = o;
    } else if (total) {
        // get_join, then remove_if(is_key_equal_value): a stable compaction
        DBuf<u64> all(c, total), kept(c, total);
        join_expand(c, h->ix, d_probe, (u64)n, js, total, base, all.p);
        total = select(c, SEL_NODIAG, all.p, total, nullptr, 0, kept.p);
        c->sync();
        if (total) {   // the kept rows' buffer goes to the caller as it is
            *d_out = kept.p;
            kept.p = nullptr;
            kept.cap = 0;
        }
    }
    *out_n = (int64_t)total;
    API_END(c)
}

dlg_status dlg_wcc_relax(dlg_ctx *c, const dlg_hash *h, const uint64_t *d_delta, int64_t n_delta, uint64_t *d_cc,
                         int64_t n_cc, uint64_t *d_next, int64_t *n_next) {
    if (!c || !h || !n_next || n_delta < 0 || n_cc < 0 || (n_delta && !d_delta) || (n_cc && (!d_cc || !d_next)))
        return DLG_INVALID_INPUT;
    *n_next = 0;
    API_BEGIN(c)
    if ((u64)n_cc != h->ix.groups) fail(DLG_INVALID_INPUT, "dlg_wcc_relax: n_cc differs from the index's key count");
    if (n_delta && n_cc) {
        if (!h->nbr_built) {   // the index's group of every row's value, kept with the index
            h->nbr.ensure(h->ix.n);
            index_lookup_values(c, h->ix, h->nbr.p);
            h->nbr_built = true;
        }
        *n_next = (int64_t)wcc_relax_tuples(c, h->ix, h->nbr.p, d_delta, (u64)n_delta, d_cc, (u64)n_cc, d_next);
    }
    API_END(c)
}

dlg_status dlg_sort(dlg_ctx *c, uint64_t *d, int64_t n) {
    if (!c || n < 0 || (n && !d)) return DLG_INVALID_INPUT;
    API_BEGIN(c)
    DBuf<u64> tmp(c);
    radix_sort(c, d, (u64)n, tmp);
    c->sync();
    API_END(c)
}

dlg_status dlg_sort_unique(dlg_ctx *c, uint64_t *d, int64_t n, int64_t *out_n) {
    if (!c || !out_n || n < 0 || (n && !d)) return DLG_INVALID_INPUT;
    *out_n = 0;
    API_BEGIN(c)
    if (n) {
        DBuf<u64> tmp(c), out(c, (size_t)n);
        radix_sort(c, d, (u64)n, tmp);
        u64 k = select(c, SEL_UNIQUE, d, (u64)n, nullptr, 0, out.p);
        copy_u64(c, d, out.p, k);
        c->sync();
        *out_n = (int64_t)k;
    }
    API_END(c)
}

dlg_status dlg_set_difference(dlg_ctx *c, const uint64_t *a, int64_t na, const uint64_t *b,
                              int64_t nb, uint64_t *out, int64_t *out_n) {
    if (!c || !out_n || na < 0 || nb < 0 || (na && (!a || !out)) || (nb && !b))
        return DLG_INVALID_INPUT;
    *out_n = 0;
    API_BEGIN(c)
    if (na) {
        DBuf<u64> o(c, (size_t)na);
        u64 k = select(c, SEL_NOT_IN, a, (u64)na, b, (u64)nb, o.p);
        copy_u64(c, out, o.p, k);
        c->sync();
        *out_n = (int64_t)k;
    }
    API_END(c)
}

dlg_status dlg_merge(dlg_ctx *c, const uint64_t *a, int64_t na, const uint64_t *b, int64_t nb,
                     uint64_t *out) {
    if (!c || na < 0 || nb < 0 || (na && !a) || (nb && !b) || ((na || nb) && !out))
        return DLG_INVALID_INPUT;
    API_BEGIN(c)
    merge(c, a, (u64)na, b, (u64)nb, out);
    c->sync();
    API_END(c)
}

dlg_status dlg_merge_unique(dlg_ctx *c, const uint64_t *a, int64_t na, const uint64_t *b, int64_t nb,
                            uint64_t *out, int64_t *out_n) {
    if (!c || !out_n || na < 0 || nb < 0 || (na && !a) || (nb && !b) || ((na || nb) && !out))
        return DLG_INVALID_INPUT;
    API_BEGIN(c)
    *out_n = (int64_t)merge_unique(c, a, (u64)na, b, (u64)nb, out);
    API_END(c)
}

dlg_status dlg_exclusive_scan(dlg_ctx *c, const uint32_t *in, int64_t n, uint64_t *out,
                              uint64_t *total) {
    if (!c || n < 0 || (n && (!in || !out))) return DLG_INVALID_INPUT;
    API_BEGIN(c)
    u64 t = n ? scan_u32_u64(c, in, (u64)n, out) : 0;
    if (total) *total = t;
    API_END(c)
}

}  // extern "C"

namespace dlg {
// Plain run of one program on device rows (used by the served form).
void run_program(dlg_ctx *ctx, int prog, const int32_t *d_rows, u64 n, const Opts &o, dlg_result *res) {
    switch (prog) {
        case DLG_PROG_TC: run_tc(ctx, d_rows, n, o, res); break;
        case DLG_PROG_SG: run_sg(ctx, d_rows, n, o, res); break;
        case DLG_PROG_WCC: run_wcc(ctx, d_rows, n, o, res); break;
        case DLG_PROG_SJ: run_sj(ctx, d_rows, n, o, res); break;
        default: fail(DLG_INVALID_INPUT, "unknown program");
    }
}
}  // namespace dlg
