// engine.h -- fixpoint drivers shared by the plain (engine.hip), batched
// (batch.hip) and served (serve.hip) forms.
#pragma once

#include <chrono>
#include <utility>
#include <vector>

#include "dlg_internal.h"

namespace dlg {

struct Opts {
    double lf;
    int bits;
    int max_it;
    bool keep;
    int strategy;
    int block_size;  // 0 = the context's
    bool check;      // output digest
    int join;        // 0 hash index, 1 sorted index (binary search)
    int reorder;     // TC bitmap layout: 0 auto, 1 never, 2 always renumber into a locality order
    int relax;       // WCC eager rounds: 0 auto, 1 push, 2 pull
    int queue;       // TC bitmap rounds queued ahead of the host: 0 auto, 1..8
};
Opts read_opts(dlg_ctx *ctx, const dlg_opts *o);

// Applies a call's block size to the context for the duration of the call.
struct BlockSizeScope {
    dlg_ctx *ctx;
    int saved;
    BlockSizeScope(dlg_ctx *c, const Opts &o) : ctx(c), saved(c->block_size) {
        if (o.block_size) ctx->block_size = o.block_size;
    }
    ~BlockSizeScope() { ctx->block_size = saved; }
};

// Phase timing with events on the context stream: mark(p) starts phase p.
struct PhaseClock {
    dlg_ctx *ctx;
    std::vector<std::pair<int, hipEvent_t>> marks;
    explicit PhaseClock(dlg_ctx *c) : ctx(c) {}
    ~PhaseClock();
    void mark(int ph);
    void finish(double *out);  // adds seconds per phase into out[DLG_PH_COUNT]
};

typedef std::chrono::steady_clock Clock;
inline double secs(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double>(b - a).count();
}

// Called once per fixpoint round with the round's new tuples N (unique,
// disjoint from the full relation; `sorted` tells whether N is in tuple
// order) -- the batched form counts them per instance.
struct RoundHook {
    virtual ~RoundHook() {}
    virtual void on_round(dlg_ctx *ctx, const u64 *N, u64 nn, int round, bool sorted) = 0;
    // TC in the bitmap layout keeps a round's new tuples grouped by source:
    // a hook that only counts them may take the sources instead (nact of
    // them, ids in the tuples' packed form, source i with len[i] new tuples;
    // nn in all), so the tuples are never built.
    virtual bool by_source() const { return false; }
    virtual void on_round_sources(dlg_ctx *, const u32 *, const u32 *, u64, u64, int) {}
};

// Id-range width up to which TC / SG keep the full relation as a bitmap
// (W*W bits <= 2^36 bits = 8 GiB) when the strategy is automatic, and the
// widest range the bitmap layout takes when asked for (32-bit word indexes:
// W * ceil(W / 32) < 2^32 - 1; 16 GiB).
constexpr u64 BITMAP_MAX_W = 1ull << 18;
constexpr u64 BITMAP_FORCED_MAX_W = 370720;

struct FixpointOut {
    DBuf<u64> full;   // final relation, sorted
    u64 nfull = 0;
    int iterations = 0;
    i64 raw = 0;
    explicit FixpointOut(dlg_ctx *c) : full(c) {}
};

// Relations from int32 device rows (n rows).  `hook` may be null.
// `blocks` (batched form): packed id-range starts of the instances, plus the
// end; tuples never pair ids of two ranges.
void tc_fixpoint(dlg_ctx *ctx, const int32_t *d_rows, u64 n, const Opts &o, PhaseClock &pc,
                 RoundHook *hook, FixpointOut &out, const std::vector<u32> *blocks = nullptr);
void sg_fixpoint(dlg_ctx *ctx, const int32_t *d_rows, u64 n, const Opts &o, PhaseClock &pc,
                 RoundHook *hook, FixpointOut &out);
void sj_fixpoint(dlg_ctx *ctx, const int32_t *d_rows, u64 n, const Opts &o, PhaseClock &pc,
                 FixpointOut &out);

// WCC set-up shared by the plain and batched forms: symmetric deduplicated
// edges, index over them (groups = nodes) and dense neighbour ids.
struct WccGraph {
    Index ix;
    DBuf<u32> nbr;
    u64 V = 0, ne = 0;
    explicit WccGraph(dlg_ctx *c) : ix(c), nbr(c) {}
};
void wcc_graph_build(dlg_ctx *ctx, const int32_t *d_rows, u64 n, const Opts &o, PhaseClock &pc,
                     WccGraph &g);

void wcc_solve(dlg_ctx *ctx, const WccGraph &g, const Opts &o, PhaseClock &pc, dlg_result *res);

// Plain run of one program on device rows (engine.hip).
void run_program(dlg_ctx *ctx, int prog, const int32_t *d_rows, u64 n, const Opts &o, dlg_result *res);

// Batched form (batch.hip): B instances of one program as one fixpoint.
void batch_run(dlg_ctx *ctx, int prog, int B, const int32_t *const *rows, const int64_t *n,
               const Opts &o, dlg_result *res, dlg_batch_stats *st);

void download_rows(dlg_ctx *ctx, const u64 *d, u64 n, int orient, int32_t *h_rows);
void check_rows_arg(int64_t n, const void *rows);

}  // namespace dlg

// The index handed across the C ABI (dlg_hash_build).  `nbr` (the index's
// group of every row's value, ~0 when the value is no key) is built on the
// first dlg_wcc_relax call that needs it.
struct dlg_hash {
    dlg::Index ix;
    mutable dlg::DBuf<dlg::u32> nbr;
    mutable bool nbr_built = false;
    explicit dlg_hash(dlg_ctx *c) : ix(c), nbr(c) {}
};
