// serve.hip -- served form: a request loop with HBM-budget admission and
// continuous batching (no counterpart in the reference, which runs one
// fixpoint per process).
//
// Slots.  `workers` host threads, each owning a dlg_ctx (its own HIP stream
// and memory pool); kernels of requests in different slots overlap on the
// device.  Streams keep the default priority.
// Admission.  A request reserves its predicted peak footprint.  Requests are
// taken in arrival order, skipping ahead past one that does not fit yet;
// while a request waits to run alone (see below) nothing new is admitted, so
// it cannot starve.  A request whose prediction exceeds the whole budget is
// refused at submission.
// Batching.  A slot that frees up takes the oldest admissible request; if it
// is small, every other waiting small request of the same program that fits
// joins it, and they run as one disjoint-union batch (batch_run: results are
// exact per request).
// Reservations are enforced: a slot's pool may hold at most its reservation
// plus `retain_bytes` of cache.  A request that outgrows it is re-queued at
// the front, first to run without batching, then alone with the whole budget.
#include <cmath>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "engine.h"

namespace dlg {

enum { MODE_BATCHABLE = 0, MODE_SINGLE = 1, MODE_SOLO = 2 };

struct Request {
    u64 ticket = 0;
    int prog = 0;
    int graph = -1;
    std::vector<int32_t> rows;
    int64_t n = 0;
    size_t reserve = 0;
    int mode = MODE_BATCHABLE;
    bool done = false;
    dlg_status status = DLG_PENDING;
    std::string err;
    dlg_result res;
    dlg_request_info info;
};

struct ResidentGraph {
    std::unique_ptr<WccGraph> g;
    size_t bytes = 0;
};

}  // namespace dlg

using namespace dlg;

struct dlg_server {
    dlg_server_opts o;
    Opts opts;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> pending;
    std::unordered_map<u64, std::unique_ptr<Request>> reqs;
    std::vector<ResidentGraph> graphs;
    dlg_ctx *loader = nullptr;
    std::vector<dlg_ctx *> ctxs;
    std::vector<std::thread> threads;
    size_t budget = 0, in_use = 0;
    bool stop = false;
    u64 next_ticket = 1;
    Clock::time_point t0;
    std::string error;
};

namespace {

double now_s(const dlg_server *s) { return secs(s->t0, Clock::now()); }

// Predicted peak device bytes of one request, by the layout its run will
// use.  TC and SG keep the full relation as a bitmap when the id range W is
// at most BITMAP_MAX_W wide: F and its transposed copy for the read-out,
// W * ceil(W / 32) words each (SG: F, the round's maps D and R), plus the
// sorted output and the round buffers, about 32 B per output tuple, plus
// the edge index.  Wider ranges keep F as sorted tuples with a hash set and
// a final sort: about 48 B per output tuple.  The output is |TC| by SCC
// condensation (dlg_predict_tc_size, estimates padded by a quarter), or the
// bound nodes^2 (SG; TC when even the bound fits or the request has more than
// PREDICT_MAX_ROWS rows), with nodes bounded by the id range and twice the rows.
// A single join reserves 384 B per row for its relations, plus (when the id
// range allows degree counts) its map and output in the map form, or 24 B
// per raw join row sorted, the raw join bounded by in-degree x out-degree.
constexpr int64_t PREDICT_MAX_ROWS = 4000000;

size_t predict(const dlg_server *s, int prog, const int32_t *rows, int64_t n, int graph) {
    const size_t slack = 32ull << 20;
    if (graph >= 0) return 64 * (size_t)s->graphs[graph].g->V + slack;
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    for (int64_t i = 0; i < 2 * n; i++) {
        lo = std::min<int64_t>(lo, rows[i]);
        hi = std::max<int64_t>(hi, rows[i]);
    }
    const double W = n ? (double)(hi - lo + 1) : 0.0;
    const double nodes = std::min(2.0 * (double)n, W);
    double b;
    switch (prog) {
        case DLG_PROG_TC:
        case DLG_PROG_SG: {
            const bool bitmap = (s->opts.strategy == 0 && W <= (double)BITMAP_MAX_W) ||
                                (s->opts.strategy == 4 && W <= (double)BITMAP_FORCED_MAX_W);
            const double maps = bitmap ? (prog == DLG_PROG_SG ? 3.0 : 2.0) * W * std::ceil(W / 32.0) * 4.0 : 0.0;
            const double per_tuple = bitmap ? 32.0 : 48.0;
            double out = nodes * nodes;
            // the condensation costs O(rows log rows) host time: up to PREDICT_MAX_ROWS rows
            if (prog == DLG_PROG_TC && n <= PREDICT_MAX_ROWS &&
                maps + per_tuple * out + 160.0 * (double)n > (double)(256ull << 20)) {
                int64_t k = 0;
                int32_t exact = 0;
                if (dlg_predict_tc_size(rows, n, &k, &exact) == DLG_OK) out = (double)k * (exact ? 1.0 : 1.25);
            }
            b = maps + per_tuple * out + 160.0 * (double)n;
            break;
        }
        case DLG_PROG_WCC:
            b = 200.0 * (double)n + 64.0 * nodes;
            break;
        default: {          // single join
            b = 48.0 * (double)n * 8.0;   // R, R^-1 and their index
            const int st = s->opts.strategy;
            if (n && W <= (double)BITMAP_FORCED_MAX_W && (st == 4 || (st == 0 && s->opts.join == 0))) {
                // the raw join's size with duplicates (a bound of the engine's count over
                // distinct rows): the sum over ids b of in-degree x out-degree
                std::vector<uint32_t> in((size_t)W, 0u), out((size_t)W, 0u);
                for (int64_t i = 0; i < n; i++) {
                    out[(size_t)(rows[2 * i] - lo)]++;
                    in[(size_t)(rows[2 * i + 1] - lo)]++;
                }
                double raw = 0.0;
                for (size_t k = 0; k < (size_t)W; k++) raw += (double)in[k] * (double)out[k];
                const double mapb = W * std::ceil(W / 32.0) * 4.0;
                if (st == 4 || (W <= (double)BITMAP_MAX_W && mapb <= 8.0 * raw))
                    b += mapb + 8.0 * std::min(raw, W * W);   // the map form: map + output
                else
                    b += 24.0 * raw;                          // J, its sort scratch, the output
            } else {
                b += 24.0 * 8.0 * (double)n;   // no degree counts over a wide id range
            }
            break;
        }
    }
    if (b > 1e18) b = 1e18;
    return (size_t)b + slack;
}

bool batchable(const dlg_server *s, const Request *r) {
    return r->mode == MODE_BATCHABLE && r->graph < 0 && r->n <= s->o.small_rows;
}

// Under the lock: choose the next work unit, or return false.
bool pick(dlg_server *s, std::vector<Request *> &batch, size_t &reserved) {
    batch.clear();
    reserved = 0;
    if (s->pending.empty()) return false;
    size_t avail = s->budget - s->in_use;
    // a request waiting to run alone drains the device first
    for (auto it = s->pending.begin(); it != s->pending.end(); ++it) {
        if ((*it)->mode == MODE_SOLO) {
            if (s->in_use != 0) return false;
            Request *r = *it;
            s->pending.erase(it);
            batch.push_back(r);
            reserved = s->budget;
            s->in_use += reserved;   // nothing else is admitted until it ends (the worker gives it back)
            return true;
        }
    }
    auto first = s->pending.end();
    for (auto it = s->pending.begin(); it != s->pending.end(); ++it)
        if ((*it)->reserve <= avail) {
            first = it;
            break;
        }
    if (first == s->pending.end()) return false;
    Request *r = *first;
    s->pending.erase(first);
    batch.push_back(r);
    reserved = r->reserve;
    if (batchable(s, r)) {
        for (auto it = s->pending.begin(); it != s->pending.end() && (int)batch.size() < s->o.max_batch;) {
            Request *q = *it;
            if (q->prog == r->prog && batchable(s, q) && reserved + q->reserve <= avail) {
                reserved += q->reserve;
                batch.push_back(q);
                it = s->pending.erase(it);
            } else {
                ++it;
            }
        }
    }
    s->in_use += reserved;
    return true;
}

void run_unit(dlg_server *s, dlg_ctx *ctx, std::vector<Request *> &batch, dlg_status &st, std::string &err) {
    st = DLG_OK;
    try {
        (void)hipSetDevice(ctx->device);
        BlockSizeScope bss(ctx, s->opts);
        Request *r0 = batch[0];
        if (r0->graph >= 0) {
            PhaseClock pc(ctx);
            auto t = Clock::now();
            wcc_solve(ctx, *s->graphs[r0->graph].g, s->opts, pc, &r0->res);
            r0->res.device_seconds = secs(t, Clock::now());
            pc.finish(r0->res.phase_seconds);
        } else if (batch.size() == 1) {
            DBuf<int32_t> d(ctx, 2 * (size_t)(r0->n ? r0->n : 1));
            if (r0->n)
                HIPCHK(hipMemcpyAsync(d.p, r0->rows.data(), sizeof(int32_t) * 2 * (size_t)r0->n,
                                      hipMemcpyHostToDevice, ctx->stream));
            r0->res.input_rows = r0->n;
            run_program(ctx, r0->prog, d.p, (u64)r0->n, s->opts, &r0->res);
        } else {
            int B = (int)batch.size();
            std::vector<const int32_t *> ptrs(B);
            std::vector<int64_t> ns(B);
            std::vector<dlg_result> out(B);
            for (int i = 0; i < B; i++) {
                ptrs[i] = batch[i]->rows.data();
                ns[i] = batch[i]->n;
            }
            dlg_batch_stats bst;
            std::memset(&bst, 0, sizeof(bst));
            batch_run(ctx, r0->prog, B, ptrs.data(), ns.data(), s->opts, out.data(), &bst);
            for (int i = 0; i < B; i++) {
                batch[i]->res = out[i];
                batch[i]->res.device_seconds = bst.device_seconds;
            }
        }
    } catch (const Error &e) {
        st = e.status;
        err = e.msg;
    } catch (const std::bad_alloc &) {
        st = DLG_OUT_OF_MEMORY;
        err = "host allocation failed";
    }
}

void worker(dlg_server *s, int slot) {
    dlg_ctx *ctx = s->ctxs[slot];
    std::vector<Request *> batch;
    while (true) {
        size_t reserved = 0;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return pick(s, batch, reserved) || (s->stop && s->pending.empty()); });
            if (batch.empty()) return;
            double t = now_s(s);
            for (Request *r : batch) {
                r->info.start_s = t;
                r->info.slot = slot;
                r->info.batch_size = (int)batch.size();
                r->info.reserved_bytes = (int64_t)r->reserve;
                r->info.attempts++;
            }
        }
        ctx->pool->set_budget(reserved + (size_t)s->o.retain_bytes);
        dlg_status st;
        std::string err;
        run_unit(s, ctx, batch, st, err);
        try {
            ctx->sync();
        } catch (const Error &e) {
            if (st == DLG_OK) {
                st = e.status;
                err = e.msg;
            }
        }
        ctx->pool->set_budget(0);
        ctx->pool->trim_to((size_t)s->o.retain_bytes);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->in_use -= reserved;
            double t = now_s(s);
            for (auto it = batch.rbegin(); it != batch.rend(); ++it) {
                Request *r = *it;
                if (st == DLG_OUT_OF_MEMORY && r->mode < MODE_SOLO) {
                    dlg_result_free(&r->res);
                    std::memset(&r->res, 0, sizeof(r->res));
                    r->mode = (batch.size() > 1 && r->mode == MODE_BATCHABLE) ? MODE_SINGLE : MODE_SOLO;
                    if (r->mode == MODE_SOLO) r->reserve = s->budget;
                    s->pending.push_front(r);
                    continue;
                }
                r->status = st;
                r->err = err;
                r->done = true;
                r->info.status = st;
                r->info.end_s = t;
                if (st != DLG_OK) {
                    dlg_result_free(&r->res);
                    std::memset(&r->res, 0, sizeof(r->res));
                }
            }
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
}

}  // namespace

extern "C" {

dlg_status dlg_server_create(const dlg_server_opts *o, dlg_server **out) {
    if (!out) return DLG_INVALID_INPUT;
    *out = nullptr;
    dlg_server *s = new dlg_server();
    if (o) s->o = *o;
    else std::memset(&s->o, 0, sizeof(s->o));
    if (s->o.workers <= 0) s->o.workers = 4;
    if (s->o.workers > 16) s->o.workers = 16;
    if (s->o.max_batch <= 0) s->o.max_batch = 64;
    if (s->o.small_rows <= 0) s->o.small_rows = 20000;
    if (s->o.retain_bytes <= 0) s->o.retain_bytes = 256ll << 20;
    s->t0 = Clock::now();
    dlg_status st = dlg_ctx_create(s->o.device, &s->loader);
    if (st != DLG_OK) {
        delete s;
        return st;
    }
    try {
        s->opts = read_opts(s->loader, &s->o.tune);
    } catch (const Error &e) {
        dlg_ctx_destroy(s->loader);
        delete s;
        return e.status;
    }
    s->opts.keep = true;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = (size_t)s->loader->hbm_bytes;
    size_t slots_cache = (size_t)s->o.workers * (size_t)s->o.retain_bytes;
    size_t budget = s->o.hbm_budget > 0 ? (size_t)s->o.hbm_budget : (size_t)(0.85 * (double)free_b);
    if (budget <= slots_cache + (64ull << 20)) {
        dlg_ctx_destroy(s->loader);
        delete s;
        return DLG_INVALID_INPUT;
    }
    s->budget = budget - slots_cache;
    for (int i = 0; i < s->o.workers; i++) {
        dlg_ctx *c = nullptr;
        st = dlg_ctx_create(s->o.device, &c);
        if (st != DLG_OK) {
            for (auto x : s->ctxs) dlg_ctx_destroy(x);
            dlg_ctx_destroy(s->loader);
            delete s;
            return st;
        }
        s->ctxs.push_back(c);
    }
    for (int i = 0; i < s->o.workers; i++) s->threads.emplace_back(worker, s, i);
    *out = s;
    return DLG_OK;
}

const char *dlg_server_error(const dlg_server *s) { return s ? s->error.c_str() : "null server"; }

dlg_status dlg_server_add_graph(dlg_server *s, const int32_t *h_rows, int64_t n, int32_t *graph_id) {
    if (!s || !graph_id || n < 0 || (n && !h_rows)) return DLG_INVALID_INPUT;
    dlg_ctx *ctx = s->loader;
    try {
        (void)hipSetDevice(ctx->device);
        check_rows_arg(n, h_rows);
        size_t before = ctx->pool->live_bytes();
        ResidentGraph rg;
        rg.g.reset(new WccGraph(ctx));
        {
            DBuf<int32_t> d(ctx, 2 * (size_t)(n ? n : 1));
            if (n)
                HIPCHK(hipMemcpyAsync(d.p, h_rows, sizeof(int32_t) * 2 * (size_t)n, hipMemcpyHostToDevice,
                                      ctx->stream));
            PhaseClock pc(ctx);
            wcc_graph_build(ctx, d.p, (u64)n, s->opts, pc, *rg.g);
            ctx->sync();
        }
        ctx->pool->trim();
        rg.bytes = ctx->pool->live_bytes() - before;
        std::lock_guard<std::mutex> lk(s->mu);
        if (rg.bytes + s->in_use > s->budget) fail(DLG_OUT_OF_MEMORY, "resident graph does not fit the budget");
        s->budget -= rg.bytes;
        s->graphs.push_back(std::move(rg));
        *graph_id = (int32_t)s->graphs.size() - 1;
    } catch (const Error &e) {
        s->error = e.msg;
        return e.status;
    }
    return DLG_OK;
}

dlg_status dlg_server_submit(dlg_server *s, int prog, const int32_t *h_rows, int64_t n, int32_t graph_id,
                             uint64_t *ticket) {
    if (!s || !ticket || prog < 0 || prog > 3) return DLG_INVALID_INPUT;
    if (graph_id >= 0 && (prog != DLG_PROG_WCC || graph_id >= (int32_t)s->graphs.size())) return DLG_INVALID_INPUT;
    if (graph_id < 0 && (n < 0 || (n && !h_rows) || n >= 0x7fffffffLL)) return DLG_INVALID_INPUT;
    std::unique_ptr<Request> r(new Request());
    std::memset(&r->res, 0, sizeof(r->res));
    std::memset(&r->info, 0, sizeof(r->info));
    r->prog = prog;
    r->graph = graph_id;
    if (graph_id < 0) {
        r->n = n;
        r->rows.assign(h_rows, h_rows + 2 * n);
    }
    r->reserve = predict(s, prog, r->rows.data(), r->n, graph_id);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return DLG_INVALID_INPUT;
    size_t cap = s->o.max_request_bytes > 0 ? std::min(s->budget, (size_t)s->o.max_request_bytes) : s->budget;
    if (r->reserve > cap) {
        s->error = "request refused: predicted device memory exceeds the request limit";
        return DLG_REFUSED;
    }
    r->ticket = s->next_ticket++;
    r->info.status = DLG_PENDING;
    r->info.submit_s = now_s(s);
    *ticket = r->ticket;
    s->pending.push_back(r.get());
    s->reqs[r->ticket] = std::move(r);
    s->cv_work.notify_one();
    return DLG_OK;
}

dlg_status dlg_server_wait(dlg_server *s, uint64_t ticket, double timeout_s, dlg_result *res,
                           dlg_request_info *info) {
    if (!s || !res) return DLG_INVALID_INPUT;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return DLG_INVALID_INPUT;
    Request *r = it->second.get();
    auto ready = [&] { return r->done; };
    if (timeout_s < 0) {
        s->cv_done.wait(lk, ready);
    } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) {
        return DLG_PENDING;
    }
    *res = r->res;
    if (info) *info = r->info;
    if (r->status != DLG_OK) s->error = r->err;
    s->reqs.erase(it);
    return DLG_OK;
}

void dlg_server_destroy(dlg_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->threads) t.join();
    for (auto &kv : s->reqs) dlg_result_free(&kv.second->res);
    s->reqs.clear();
    s->graphs.clear();
    for (auto c : s->ctxs) dlg_ctx_destroy(c);
    dlg_ctx_destroy(s->loader);
    delete s;
}

}  // extern "C"
