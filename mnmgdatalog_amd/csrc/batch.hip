// batch.hip -- batched form: many independent fixpoint instances as one run.
//
// Instance i's ids [lo_i, hi_i] are shifted into a private range
// [base_i, base_i + width_i) of the int32 id space (base_0 = INT32_MIN), and
// the union of all instances runs through the plain fixpoint drivers.  No
// tuple ever relates two instances (every rule joins on a shared id), so the
// union's relations are disjoint unions of the instances' relations, and in
// every sorted relation an instance occupies one contiguous key range.
// Per-round accounting:
//   TC / SG   2 binary searches per instance boundary in the round's new
//             tuples N give |N_i| when N is sorted, else a histogram of N's
//             keys over the instance ranges (LDS-privatised per workgroup);
//             an instance's iteration count is the first round with
//             |N_i| = 0 (its plain run stops there).  TC's bitmap layout
//             keeps one map per instance range (block-diagonal).
//   WCC       frontier nodes are counted per instance (dense node ids are
//             sorted by id, so instances are contiguous node ranges); each
//             instance applies the reference's stop rule (two consecutive
//             equal |Δ_i|) on its own and is frozen -- its frontier entries
//             are dropped -- when it fires.
#include "engine.h"

#include <algorithm>
#include <climits>
#include <cstdlib>

namespace dlg {

// pos[i] = lower_bound(a, (bounds[i] << 32)) for i in [0, nb)
__global__ void bounds_k(const u64 *a, u64 n, const u32 *bounds, u32 nb, u64 *pos) {
    u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nb) pos[i] = lower_bound_u64(a, 0, n, (u64)bounds[i] << 32);
}

// gkey[g] = key of node group g
__global__ void group_keys_k(const u64 *rows, const u32 *gstart, u64 G, u32 *gkey) {
    u64 g = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; g < G; g += stride) gkey[g] = tkey(rows[gstart[g]]);
}

__global__ void lower_bound_u32_k(const u32 *a, u64 n, const u32 *x, u32 nx, u32 *out) {
    u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nx) return;
    u64 lo = 0, hi = n;
    u32 v = x[i];
    while (lo < hi) {
        u64 mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1; else hi = mid;
    }
    out[i] = (u32)lo;
}

__device__ __forceinline__ u32 inst_of(const u32 *vstart, u32 B, u32 v) {
    // last i with vstart[i] <= v  (vstart[0] = 0, vstart[B] = V)
    u32 lo = 0, hi = B;
    while (hi - lo > 1) {
        u32 mid = (lo + hi) >> 1;
        if (vstart[mid] <= v) lo = mid; else hi = mid;
    }
    return lo;
}

constexpr int MAX_LDS_INST = 4096;

// Per-instance counts of frontier nodes: an LDS histogram per workgroup when
// the instances fit, flushed with one device atomic per nonzero bin.
__global__ void frontier_hist_k(const u32 *node, u64 n, const u32 *vstart, u32 B, u32 *cnt) {
    __shared__ u32 vs[MAX_LDS_INST + 1];
    __shared__ u32 s_cnt[MAX_LDS_INST];
    const bool lds = B <= MAX_LDS_INST;
    if (lds)
        for (u32 i = threadIdx.x; i <= B; i += blockDim.x) {
            vs[i] = vstart[i];
            if (i < B) s_cnt[i] = 0u;
        }
    __syncthreads();
    const u32 *tab = lds ? vs : vstart;
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        u32 b = inst_of(tab, B, node[i]);
        if (lds) atomicAdd(&s_cnt[b], 1u);
        else atomicAdd(&cnt[b], 1u);
    }
    if (!lds) return;
    __syncthreads();
    for (u32 j = threadIdx.x; j < B; j += blockDim.x)
        if (s_cnt[j]) atomicAdd(&cnt[j], s_cnt[j]);
}

// Keep frontier entries of live instances (order is irrelevant: Δ is a set).
__global__ void frontier_filter_k(const u32 *node, const u32 *label, u64 n, const u32 *vstart, u32 B,
                                  const u32 *live, u32 *onode, u32 *olabel, u32 *counter) {
    const int lane = threadIdx.x & (WAVE - 1);
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    u64 lim = (n + WAVE - 1) / WAVE * WAVE;   // whole waves iterate together
    for (; i < lim; i += stride) {
        bool keep = false;
        u32 v = 0, l = 0;
        if (i < n) {
            v = node[i];
            l = label[i];
            keep = live[inst_of(vstart, B, v)] != 0;
        }
        u64 bm = __ballot(keep);
        if (bm) {
            int leader = __ffsll((unsigned long long)bm) - 1;
            u32 b = 0;
            if (lane == leader) b = atomicAdd(counter, (u32)__popcll(bm));
            b = __shfl(b, leader, WAVE);
            if (keep) {
                u32 p = b + (u32)__popcll(bm & lanemask_lt());
                onode[p] = v;
                olabel[p] = l;
            }
        }
    }
}

__global__ void comp_largest_k(const u32 *ccount, u64 V, const u32 *vstart, u32 B, u32 *largest) {
    u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; v < V; v += stride) {
        u32 c = ccount[v];
        if (c) atomicMax(&largest[inst_of(vstart, B, (u32)v)], c);
    }
}

__global__ void comp_count_k(const u32 *L, u64 V, u32 *ccount) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    const u64 lim = (V + WAVE - 1) / WAVE * WAVE;   // whole waves iterate together
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < lim; i += stride)
        wave_count_by_key(ccount, i < V ? L[i] : 0u, i < V);
}

__global__ void comp_flags_k(const u32 *ccount, u64 V, u32 *flags) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < V; i += stride) flags[i] = ccount[i] ? 1u : 0u;
}

__global__ void comp_emit_k(const u32 *flags, const u64 *offs, u64 V, const u32 *gkey, u64 *out) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < V; i += stride)
        if (flags[i]) out[offs[i]] = tmake(gkey[i], 0x80000000u);
}

struct Layout {
    int B = 0;
    std::vector<int64_t> lo, shift;   // id' = id + shift[i]
    std::vector<u32> ub;              // flipped key of range starts, B + 1 entries
    int64_t total_rows = 0;
    int exhausted = 0;                // first instance that starts past the last id (B: none)
};

static Layout make_layout(int B, const int32_t *const *rows, const int64_t *n) {
    Layout L;
    L.B = B;
    L.lo.resize(B);
    L.shift.resize(B);
    L.ub.resize(B + 1);
    int64_t base = INT32_MIN;
    L.exhausted = B;
    for (int i = 0; i < B; i++) {
        if (n[i] < 0 || (n[i] > 0 && !rows[i])) fail(DLG_INVALID_INPUT, "bad instance rows");
        if (base > (int64_t)INT32_MAX && L.exhausted == B) L.exhausted = i;
        int64_t lo = INT64_MAX, hi = INT64_MIN;
        for (int64_t r = 0; r < 2 * n[i]; r++) {
            lo = std::min<int64_t>(lo, rows[i][r]);
            hi = std::max<int64_t>(hi, rows[i][r]);
        }
        int64_t width = n[i] ? hi - lo + 1 : 0;
        L.lo[i] = n[i] ? lo : 0;
        L.shift[i] = base - L.lo[i];
        L.ub[i] = base - 1 >= (int64_t)INT32_MAX ? 0xffffffffu : (u32)((uint64_t)(base + 2147483648LL));
        base += width;
        if (base - 1 > (int64_t)INT32_MAX)
            fail(DLG_OVERFLOW, "batch id ranges exceed 2^32 ids; split the batch");
        L.total_rows += n[i];
    }
    L.ub[B] = base - 1 >= INT32_MAX ? 0xffffffffu : (u32)((uint64_t)(base + 2147483648LL));
    return L;
}

// Boundary positions of the instance key ranges in a sorted relation.
static std::vector<u64> boundaries(dlg_ctx *ctx, const u64 *a, u64 n, const DBuf<u32> &ub, int B,
                                   DBuf<u64> &pos) {
    std::vector<u64> h(B + 1, n);
    if (n) {
        DLG_LAUNCH(ctx, DLG_K_MISC, 8.0 * (B + 1), bounds_k, grid_for(B + 1, 256), 256, 0, a, n,
                   (const u32 *)ub.p, (u32)(B + 1), pos.p);
        HIPCHK(hipMemcpyAsync(h.data(), pos.p, 8 * (B + 1), hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
    } else {
        std::fill(h.begin(), h.end(), 0);
    }
    // the last range is open-ended when it reaches the top of the id space
    h[B] = n;
    return h;
}

// Per-instance count of unordered tuples: instance = last i with ub[i] <= key.
// Per-instance counts of the keys of a: the instance by binary search over
// the range starts; counted in an LDS histogram per workgroup (when the
// instances fit) and flushed with one device atomic per nonzero bin.
__global__ void key_hist_k(const u64 *a, u64 n, const u32 *ub, u32 B, u32 *cnt) {
    __shared__ u32 s_ub[MAX_LDS_INST + 1];
    __shared__ u32 s_cnt[MAX_LDS_INST];
    const bool lds = B <= MAX_LDS_INST;
    if (lds)
        for (u32 i = threadIdx.x; i <= B; i += blockDim.x) {
            s_ub[i] = ub[i];
            if (i < B) s_cnt[i] = 0u;
        }
    __syncthreads();
    const u32 *tab = lds ? s_ub : ub;
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        u32 k = tkey(a[i]);
        u32 lo = 0, hi = B;   // tab[0] <= every key; the last range is open-ended
        while (hi - lo > 1) {
            u32 mid = (lo + hi) >> 1;
            if (tab[mid] <= k) lo = mid; else hi = mid;
        }
        if (lds) atomicAdd(&s_cnt[lo], 1u);
        else atomicAdd(&cnt[lo], 1u);
    }
    if (!lds) return;
    __syncthreads();
    for (u32 j = threadIdx.x; j < B; j += blockDim.x)
        if (s_cnt[j]) atomicAdd(&cnt[j], s_cnt[j]);
}

// The same counts from a round's sources: source ax[i] adds len[i].
__global__ void source_hist_k(const u32 *ax, const u32 *len, u64 n, const u32 *ub, u32 B, u32 *cnt) {
    __shared__ u32 s_ub[MAX_LDS_INST + 1];
    __shared__ u32 s_cnt[MAX_LDS_INST];
    const bool lds = B <= MAX_LDS_INST;
    if (lds)
        for (u32 i = threadIdx.x; i <= B; i += blockDim.x) {
            s_ub[i] = ub[i];
            if (i < B) s_cnt[i] = 0u;
        }
    __syncthreads();
    const u32 *tab = lds ? s_ub : ub;
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const u32 k = ax[i], w = len[i];
        if (!w) continue;
        u32 lo = 0, hi = B;   // tab[0] <= every key; the last range is open-ended
        while (hi - lo > 1) {
            const u32 mid = (lo + hi) >> 1;
            if (tab[mid] <= k) lo = mid; else hi = mid;
        }
        if (lds) atomicAdd(&s_cnt[lo], w);
        else atomicAdd(&cnt[lo], w);
    }
    if (!lds) return;
    __syncthreads();
    for (u32 j = threadIdx.x; j < B; j += blockDim.x)
        if (s_cnt[j]) atomicAdd(&cnt[j], s_cnt[j]);
}

struct BatchHook : RoundHook {
    const DBuf<u32> &ub;
    int B;
    DBuf<u64> pos;
    DBuf<u32> cnt;
    std::vector<int> iters;
    std::vector<char> done;
    BatchHook(dlg_ctx *ctx, const DBuf<u32> &u, int b)
        : ub(u), B(b), pos(ctx, b + 1), cnt(ctx, b), iters(b, 0), done(b, 0) {}
    void on_round(dlg_ctx *ctx, const u64 *N, u64 nn, int round, bool sorted) override {
        std::vector<u64> c(B, 0);
        if (sorted) {
            std::vector<u64> p = boundaries(ctx, N, nn, ub, B, pos);
            for (int i = 0; i < B; i++) c[i] = p[i + 1] - p[i];
        } else if (nn) {
            fill_u32(ctx, cnt.p, B, 0u);
            DLG_LAUNCH(ctx, DLG_K_MISC, 8.0 * nn, key_hist_k, sgrid(ctx, nn), ctx->block_size, 0, N, nn,
                       (const u32 *)ub.p, (u32)B, cnt.p);
            std::vector<u32> h(B);
            HIPCHK(hipMemcpyAsync(h.data(), cnt.p, 4 * B, hipMemcpyDeviceToHost, ctx->stream));
            ctx->sync();
            for (int i = 0; i < B; i++) c[i] = h[i];
        }
        record(c, round);
    }
    bool by_source() const override { return true; }
    void on_round_sources(dlg_ctx *ctx, const u32 *ax, const u32 *len, u64 nact, u64 nn, int round) override {
        std::vector<u64> c(B, 0);
        if (nn && nact) {
            fill_u32(ctx, cnt.p, B, 0u);
            DLG_LAUNCH(ctx, DLG_K_MISC, 8.0 * nact, source_hist_k, sgrid(ctx, nact), ctx->block_size, 0, ax, len,
                       nact, (const u32 *)ub.p, (u32)B, cnt.p);
            std::vector<u32> h(B);
            HIPCHK(hipMemcpyAsync(h.data(), cnt.p, 4 * B, hipMemcpyDeviceToHost, ctx->stream));
            ctx->sync();
            for (int i = 0; i < B; i++) c[i] = h[i];
        }
        record(c, round);
    }
    // an instance's iteration count is the first round that gave it nothing new
    void record(const std::vector<u64> &c, int round) {
        for (int i = 0; i < B; i++) {
            if (done[i]) continue;
            iters[i] = round;
            if (c[i] == 0) done[i] = 1;
        }
    }
};

static void instance_digest(dlg_ctx *ctx, const u64 *d, u64 n, int orient, u32 sub0, u32 sub1, dlg_result *r) {
    u64 sum = 0, breaks = 0;
    rows_digest(ctx, d, n, orient, &sum, &breaks, sub0, sub1);
    r->output_sum = sum;
    r->output_order_breaks = (int64_t)breaks;
}

// Rows of every instance in its own ids: tuple i of the sorted output (of
// the instance j with bnd[j] <= i < bnd[j + 1], found by a binary search)
// unpacked in `orient` with shift[j] taken off the first column (and off the
// second too when both hold ids).
__global__ void unpack_instances_k(const u64 *t, u64 n, const u64 *bnd, const u32 *shift, u32 B, int orient,
                                   int both, int32_t *rows) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        u32 lo = 0, hi = B;   // last j with bnd[j] <= i
        while (hi - lo > 1) {
            const u32 mid = (lo + hi) >> 1;
            if (bnd[mid] <= i) lo = mid; else hi = mid;
        }
        const u32 k = tkey(t[i]) ^ 0x80000000u, v = tval(t[i]) ^ 0x80000000u, s = shift[lo];
        const u32 r0 = orient == DLG_ORIENT_KV ? k : v, r1 = orient == DLG_ORIENT_KV ? v : k;
        rows[2 * i] = (int32_t)(r0 - s);
        rows[2 * i + 1] = (int32_t)(both ? r1 - s : r1);
    }
}

// Every instance's size and, with `keep`, its host rows: unpacked and shifted
// back on the device, then one copy into each instance's own array.
static void split_rows(dlg_ctx *ctx, const u64 *d, const std::vector<u64> &b, const Layout &L, int orient, bool both,
                       bool keep, dlg_result *res) {
    for (int i = 0; i < L.B; i++) res[i].output_size = (int64_t)(b[i + 1] - b[i]);
    if (!keep) return;
    const u64 n = b[L.B];
    DBuf<int32_t> rows(ctx, 2 * (n ? n : 1));
    if (n) {
        std::vector<u32> sh(L.B);
        for (int i = 0; i < L.B; i++) sh[i] = (u32)L.shift[i];
        DBuf<u64> bnd(ctx, L.B + 1);
        DBuf<u32> dsh(ctx, L.B);
        HIPCHK(hipMemcpyAsync(bnd.p, b.data(), 8 * (L.B + 1), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(dsh.p, sh.data(), 4 * L.B, hipMemcpyHostToDevice, ctx->stream));
        DLG_LAUNCH(ctx, DLG_K_PACK, 16.0 * n, unpack_instances_k, sgrid(ctx, n), ctx->block_size, 0, d, n,
                   (const u64 *)bnd.p, (const u32 *)dsh.p, (u32)L.B, orient, both ? 1 : 0, rows.p);
        ctx->sync();   // the host vectors go out of scope
    }
    for (int i = 0; i < L.B; i++) {
        const u64 cnt = b[i + 1] - b[i];
        res[i].n_rows = (int64_t)cnt;
        res[i].rows = (int32_t *)host_block_alloc(sizeof(int32_t) * 2 * (cnt ? cnt : 1));
        if (!res[i].rows) fail(DLG_OUT_OF_MEMORY, "host allocation of batch rows");
        if (cnt)
            HIPCHK(hipMemcpyAsync(res[i].rows, rows.p + 2 * b[i], sizeof(int32_t) * 2 * cnt, hipMemcpyDeviceToHost,
                                  ctx->stream));
    }
    ctx->sync();
}

void batch_run(dlg_ctx *ctx, int prog, int B, const int32_t *const *rows, const int64_t *n,
                      const Opts &o, dlg_result *res, dlg_batch_stats *st) {
    Layout L = make_layout(B, rows, n);
    if (L.exhausted < B) {
        // The instances from L.exhausted on start where the ids are used up, so
        // they are empty (a wider one overflows above); their range start would
        // be the last id of the instance before them, whose tuples the range
        // searches would then credit to them.  They run as a batch of their own.
        const int E = L.exhausted;
        dlg_batch_stats s2;
        std::memset(&s2, 0, sizeof(s2));
        batch_run(ctx, prog, E, rows, n, o, res, st);
        batch_run(ctx, prog, B - E, rows + E, n + E, o, res + E, &s2);
        st->device_seconds += s2.device_seconds;
        st->rounds = std::max(st->rounds, s2.rounds);
        st->total_output += s2.total_output;
        st->union_rows += s2.union_rows;
        return;
    }
    if (L.total_rows >= 0x7fffffffLL) fail(DLG_OVERFLOW, "batch has more than 2^31-2 rows");
    std::vector<int32_t> un((size_t)(2 * L.total_rows + 2));
    {
        size_t k = 0;
        for (int i = 0; i < B; i++)
            for (int64_t r = 0; r < 2 * n[i]; r++) un[k++] = (int32_t)((int64_t)rows[i][r] + L.shift[i]);
    }
    for (int i = 0; i < B; i++) {
        std::memset(&res[i], 0, sizeof(dlg_result));
        res[i].input_rows = n[i];
        res[i].raw_join_rows = -1;
    }
    auto t0 = Clock::now();
    u64 N = (u64)L.total_rows;
    DBuf<int32_t> d_rows(ctx, 2 * (N ? N : 1));
    if (N)
        HIPCHK(hipMemcpyAsync(d_rows.p, un.data(), sizeof(int32_t) * 2 * N, hipMemcpyHostToDevice,
                              ctx->stream));
    DBuf<u32> ub(ctx, B + 1);
    HIPCHK(hipMemcpyAsync(ub.p, L.ub.data(), 4 * (B + 1), hipMemcpyHostToDevice, ctx->stream));
    PhaseClock pc(ctx);
    int rounds = 0;
    if (prog == DLG_PROG_TC || prog == DLG_PROG_SG || prog == DLG_PROG_SJ) {
        FixpointOut f(ctx);
        BatchHook hook(ctx, ub, B);
        if (prog == DLG_PROG_TC) tc_fixpoint(ctx, d_rows.p, N, o, pc, &hook, f, &L.ub);
        else if (prog == DLG_PROG_SG) sg_fixpoint(ctx, d_rows.p, N, o, pc, &hook, f);
        else sj_fixpoint(ctx, d_rows.p, N, o, pc, f);
        rounds = f.iterations;
        std::vector<u64> b = boundaries(ctx, f.full.p, f.nfull, ub, B, hook.pos);
        ctx->sync();
        st->device_seconds = secs(t0, Clock::now());
        split_rows(ctx, f.full.p, b, L, prog == DLG_PROG_SG ? DLG_ORIENT_KV : DLG_ORIENT_VK, true, o.keep, res);
        for (int i = 0; i < B; i++)
            res[i].iterations = prog == DLG_PROG_SJ ? 1 : (hook.iters[i] ? hook.iters[i] : rounds);
        if (o.check) {   // each instance's digest over its own ids
            const int orient = prog == DLG_PROG_SG ? DLG_ORIENT_KV : DLG_ORIENT_VK;
            for (int i = 0; i < B; i++)
                instance_digest(ctx, f.full.p + b[i], b[i + 1] - b[i], orient, (u32)L.shift[i], (u32)L.shift[i],
                                &res[i]);
        }
    } else if (prog == DLG_PROG_WCC) {
        WccGraph g(ctx);
        wcc_graph_build(ctx, d_rows.p, N, o, pc, g);
        u64 V = g.V;
        DBuf<u32> gkey(ctx, V ? V : 1), vstart(ctx, B + 1), cnt(ctx, B), live(ctx, B), largest(ctx, B);
        if (V)
            DLG_LAUNCH(ctx, DLG_K_MISC, 12.0 * V, group_keys_k, sgrid(ctx, V), ctx->block_size, 0,
                       (const u64 *)g.ix.rows.p, (const u32 *)g.ix.gstart.p, V, gkey.p);
        DLG_LAUNCH(ctx, DLG_K_MISC, 4.0 * B, lower_bound_u32_k, grid_for(B + 1, 256), 256, 0,
                   (const u32 *)gkey.p, V, (const u32 *)ub.p, (u32)(B + 1), vstart.p);
        std::vector<u32> hv(B + 1);
        HIPCHK(hipMemcpyAsync(hv.data(), vstart.p, 4 * (B + 1), hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
        hv[B] = (u32)V;
        HIPCHK(hipMemcpyAsync(vstart.p, hv.data(), 4 * (B + 1), hipMemcpyHostToDevice, ctx->stream));
        pc.mark(DLG_PH_INIT);
        WccState s(ctx);
        wcc_setup(ctx, s, V);
        std::vector<i64> prev(B), cur(B);
        std::vector<int> iters(B, 0);
        std::vector<char> done(B, 0);
        std::vector<u32> hlive(B, 1);
        for (int i = 0; i < B; i++) prev[i] = (i64)(hv[i + 1] - hv[i]);   // wcc.cu:236-242
        int live_n = B;
        u64 nf = V;
        DBuf<u32> tnode(ctx, V ? V : 1), tlabel(ctx, V ? V : 1);
        std::vector<u32> hc(B);
        while (live_n > 0) {
            pc.mark(DLG_PH_JOIN);
            rounds++;
            u64 nn = nf ? wcc_round(ctx, s, g.ix, g.nbr.p, nf, (u32)rounds) : 0;
            fill_u32(ctx, cnt.p, B, 0u);
            if (nn)
                DLG_LAUNCH(ctx, DLG_K_MISC, 4.0 * nn, frontier_hist_k, sgrid(ctx, nn), ctx->block_size, 0,
                           (const u32 *)s.fnode.p, nn, (const u32 *)vstart.p, (u32)B, cnt.p);
            HIPCHK(hipMemcpyAsync(hc.data(), cnt.p, 4 * B, hipMemcpyDeviceToHost, ctx->stream));
            ctx->sync();
            bool froze = false;
            for (int i = 0; i < B; i++) {
                if (done[i]) continue;
                iters[i]++;
                cur[i] = hc[i];
                if (cur[i] == prev[i]) {                  // wcc.cu:366-383, per instance
                    done[i] = 1;
                    hlive[i] = 0;
                    live_n--;
                    froze = froze || hc[i] > 0;
                } else {
                    prev[i] = cur[i];
                }
            }
            nf = nn;
            if (froze && nf) {
                HIPCHK(hipMemcpyAsync(live.p, hlive.data(), 4 * B, hipMemcpyHostToDevice, ctx->stream));
                fill_u32(ctx, s.counter.p, 1, 0u);
                DLG_LAUNCH(ctx, DLG_K_MISC, 12.0 * nf, frontier_filter_k, sgrid(ctx, nf), ctx->block_size, 0,
                           (const u32 *)s.fnode.p, (const u32 *)s.flabel.p, nf, (const u32 *)vstart.p,
                           (u32)B, (const u32 *)live.p, tnode.p, tlabel.p, s.counter.p);
                HIPCHK(hipMemcpyAsync(ctx->h_scalars, s.counter.p, 4, hipMemcpyDeviceToHost, ctx->stream));
                ctx->sync();
                nf = *(u32 *)ctx->h_scalars;
                s.fnode.swap(tnode);
                s.flabel.swap(tlabel);
            }
            if (o.max_it > 0 && rounds >= o.max_it) break;
        }
        // components per instance
        pc.mark(DLG_PH_FINAL);
        DBuf<u32> ccount(ctx, V ? V : 1), flags(ctx, V ? V : 1);
        DBuf<u64> offs(ctx, V + 1), out(ctx);
        fill_u32(ctx, largest.p, B, 0u);
        u64 ncomp = 0;
        if (V) {
            fill_u32(ctx, ccount.p, V, 0u);
            DLG_LAUNCH(ctx, DLG_K_WCC_FINAL, 12.0 * V, comp_count_k, sgrid(ctx, V), ctx->block_size, 0,
                       (const u32 *)s.L.p, V, ccount.p);
            DLG_LAUNCH(ctx, DLG_K_WCC_FINAL, 8.0 * V, comp_flags_k, sgrid(ctx, V), ctx->block_size, 0,
                       (const u32 *)ccount.p, V, flags.p);
            DLG_LAUNCH(ctx, DLG_K_WCC_FINAL, 8.0 * V, comp_largest_k, sgrid(ctx, V), ctx->block_size, 0,
                       (const u32 *)ccount.p, V, (const u32 *)vstart.p, (u32)B, largest.p);
            ncomp = scan_u32_u64(ctx, flags.p, V, offs.p);
            out.ensure(ncomp ? ncomp : 1);
            DLG_LAUNCH(ctx, DLG_K_WCC_FINAL, 16.0 * V, comp_emit_k, sgrid(ctx, V), ctx->block_size, 0,
                       (const u32 *)flags.p, (const u64 *)offs.p, V, (const u32 *)gkey.p, out.p);
        }
        DBuf<u64> pos(ctx, B + 1);
        std::vector<u64> b = boundaries(ctx, out.p, ncomp, ub, B, pos);
        std::vector<u32> hl(B);
        HIPCHK(hipMemcpyAsync(hl.data(), largest.p, 4 * B, hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
        st->device_seconds = secs(t0, Clock::now());
        split_rows(ctx, out.p, b, L, DLG_ORIENT_KV, false, o.keep, res);   // (label, 0): labels are ids
        if (o.check)   // (label, 0): only the label column is shifted
            for (int i = 0; i < B; i++)
                instance_digest(ctx, out.p + b[i], b[i + 1] - b[i], DLG_ORIENT_KV, (u32)L.shift[i], 0u, &res[i]);
        for (int i = 0; i < B; i++) {
            res[i].iterations = iters[i];
            res[i].secondary = res[i].output_size ? (i64)hl[i] : -1;
            res[i].nodes = (i64)(hv[i + 1] - hv[i]);
        }
    } else {
        fail(DLG_INVALID_INPUT, "unknown program");
    }
    st->rounds = rounds;
    st->union_rows = L.total_rows;
    st->total_output = 0;
    for (int i = 0; i < B; i++) st->total_output += res[i].output_size;
}

}  // namespace dlg

using namespace dlg;

extern "C" dlg_status dlg_batch_run(dlg_ctx *ctx, int prog, int32_t count, const int32_t *const *h_rows,
                                    const int64_t *n, const dlg_opts *o, dlg_result *results,
                                    dlg_batch_stats *stats) {
    if (!ctx || count < 0 || (count && (!h_rows || !n || !results)) || !stats) return DLG_INVALID_INPUT;
    std::memset(stats, 0, sizeof(*stats));
    if (count == 0) return DLG_OK;
    API_BEGIN(ctx)
    Opts op = read_opts(ctx, o);
    BlockSizeScope bss(ctx, op);
    try {
        batch_run(ctx, prog, count, h_rows, n, op, results, stats);
    } catch (...) {
        for (int i = 0; i < count; i++) dlg_result_free(&results[i]);
        throw;
    }
    API_END(ctx)
}

// ------------------------------------------------------------------ autotuner
// Grid search over (load factor, radix digit width, block size): every
// candidate runs the whole program on the same device-resident input, the
// fastest (min over `repeats`) wins, and every candidate must reproduce the
// first candidate's iteration count and output size.
extern "C" dlg_status dlg_autotune(dlg_ctx *ctx, int prog, const int32_t *h_rows, int64_t n,
                                   const double *lfs, const int32_t *bits, const int32_t *bss,
                                   int32_t repeats, dlg_opts *best, double *best_seconds) {
    if (!ctx || !best || n < 0 || (n && !h_rows) || prog < 0 || prog > 3) return DLG_INVALID_INPUT;
    static const double def_lf[] = {0.3, 0.5, 0.7, 0.85, 0};
    static const int32_t def_bits[] = {6, 7, 8, 0};
    static const int32_t def_bs[] = {128, 256, 512, 0};
    if (!lfs) lfs = def_lf;
    if (!bits) bits = def_bits;
    if (!bss) bss = def_bs;
    if (repeats < 1) repeats = 1;
    int saved_bs = ctx->block_size;
    API_BEGIN(ctx)
    check_rows_arg(n, h_rows);
    DBuf<int32_t> d(ctx, 2 * (size_t)(n ? n : 1));
    if (n)
        HIPCHK(hipMemcpyAsync(d.p, h_rows, sizeof(int32_t) * 2 * (size_t)n, hipMemcpyHostToDevice,
                              ctx->stream));
    ctx->sync();
    typedef dlg_status (*fn_t)(dlg_ctx *, const int32_t *, int64_t, const dlg_opts *, dlg_result *);
    fn_t fns[4] = {dlg_run_tc_device, dlg_run_sg_device, dlg_run_wcc_device, dlg_run_single_join_device};
    double best_t = 1e300;
    int ref_it = -1;
    int64_t ref_out = -1;
    for (const int32_t *bs = bss; *bs; bs++) {
        if (*bs < 64 || *bs > 1024 || (*bs % 64)) fail(DLG_INVALID_INPUT, "block size must be 64..1024, multiple of 64");
        for (const double *lf = lfs; *lf > 0; lf++)
            for (const int32_t *b = bits; *b; b++) {
                dlg_opts o;
                std::memset(&o, 0, sizeof(o));
                o.load_factor = *lf;
                o.radix_bits = *b;
                o.keep_output = 0;
                ctx->block_size = *bs;
                double t = 1e300;
                for (int r = 0; r < repeats; r++) {
                    dlg_result res;
                    dlg_status s = fns[prog](ctx, d.p, n, &o, &res);
                    if (s != DLG_OK) fail(s, ctx->last_error);
                    if (ref_it < 0) {
                        ref_it = res.iterations;
                        ref_out = res.output_size;
                    } else if (res.iterations != ref_it || res.output_size != ref_out) {
                        fail(DLG_DEVICE_ERROR, "autotune: a candidate changed the result");
                    }
                    t = std::min(t, res.device_seconds);
                    dlg_result_free(&res);
                }
                if (t < best_t) {
                    best_t = t;
                    *best = o;
                    best->block_size = *bs;
                }
            }
    }
    ctx->block_size = saved_bs;
    if (best_seconds) *best_seconds = best_t;
    API_END_CLEANUP(ctx, ctx->block_size = saved_bs)
}
