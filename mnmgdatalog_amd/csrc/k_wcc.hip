// k_wcc.hip -- weakly connected components by min-label propagation.
//
// Reference: wcc.cu:140-485.  Per round the reference joins the delta
// Δ(n, l) with the symmetric edge relation, merges the result into
// cc(node, label), keeps the minimum label per node (unique_by_key after a
// sorted merge) and takes Δ' = cc' \ cc.  The same relation is computed here
// on dense node ids (the sorted distinct keys of the symmetric edge set):
//   L[v]      current label of node v (a dense id; dense order == id order,
//             so min over dense ids == min over the original ids)
//   frontier  (v, l) pairs of Δ; l is the label v had when it entered Δ, so
//             sources read a snapshot (Jacobi), exactly as the reference's
//             join reads Δ's value column
//   relax     for every edge (v, m) of a frontier node: atomicMin(L[m], l);
//             a node whose label dropped is appended once per round
//             (mark[m] == round), giving Δ' = {(m, L'(m)) : L'(m) < L(m)}
// Edge traversal is load balanced: degree counts -> scan -> one thread per
// edge, its node found from the tile's node starts set as bits in LDS.
//
// Two other forms of the same round (same labels, same Δ'):
//   round one  every node is in Δ with its own id as label, so L'(m) is the
//              least of m and its neighbours; neighbour lists are sorted
//              (the index keeps each key's rows in value order), so that is
//              min(m, first neighbour): one read per node instead of per edge
//   pull       L'(m) = min(L(m), min over ALL neighbours v of L(v)).  Equal to
//              the push round because a neighbour v outside Δ has already
//              pushed its current label: v's label last changed in a round
//              after which v was in Δ, and the next round lowered m to it
//              (round one has every node in Δ).  One thread per edge again,
//              but reading labels instead of taking atomics: a segmented
//              minimum over the wave's lanes of one node, then one atomic per
//              node and wave into a copy Ln of L; Δ' = {m : Ln(m) < L(m)}.
#include "dlg_internal.h"

#include <algorithm>

namespace dlg {

__global__ void wcc_init_k(u32 *L, u32 *fnode, u32 *flabel, u32 *mark, u64 V) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < V; i += stride) {
        L[i] = (u32)i;
        fnode[i] = (u32)i;
        flabel[i] = (u32)i;
        mark[i] = 0;
    }
}

__global__ void wcc_degree_k(const u32 *fnode, u64 nf, const u32 *gstart, u32 *cnt, u32 *pstart) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < nf; i += stride) {
        u32 v = fnode[i];
        u32 s = gstart[v];
        cnt[i] = gstart[v + 1] - s;
        pstart[i] = s;
    }
}

// Appends of a block's tile with ONE counter atomic per block: an atomic per
// wave on one address serialises over the whole grid (~70k waves for a
// round that lowers 4.6M nodes).  Each thread offers I items; the slots
// follow (item, wave, lane) order.  s_cnt: I * T / WAVE + 1 u32 of LDS, at
// most 65.  Every thread of the block must call it.
template <int T, int I>
__device__ __forceinline__ void block_append(const bool (&app)[I], u32 *count, u32 *s_cnt, u32 (&slot)[I]) {
    constexpr int W = T / WAVE, K = I * W;
    static_assert(K <= WAVE, "one wave scans the per-wave counts");
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
    u64 bm[I];
#pragma unroll
    for (int r = 0; r < I; r++) {
        bm[r] = __ballot(app[r]);
        if (lane == 0) s_cnt[r * W + wave] = (u32)__popcll(bm[r]);
    }
    __syncthreads();
    if (wave == 0) {
        const u32 c = lane < K ? s_cnt[lane] : 0u;
        u32 tot = 0;
        const u32 ex = wave_excl_scan32(c, &tot);
        u32 b = 0;
        if (lane == 0 && tot) b = atomicAdd(count, tot);
        b = __shfl(b, 0, WAVE);
        if (lane < K) s_cnt[lane] = b + ex;
    }
    __syncthreads();
    const u64 lt = lanemask_lt();
#pragma unroll
    for (int r = 0; r < I; r++) slot[r] = s_cnt[r * W + wave] + (u32)__popcll(bm[r] & lt);
    __syncthreads();   // s_cnt may be reused
}

// The owners of a tile's TS_TILE positions (edges, or join rows) base ..
// base + TS_TILE - 1 under offsets that rise strictly (every node has an
// edge): offs[lo] <= base and nodes lo+1 .. hi-1 start inside the tile.  Their
// starts are set as bits in LDS; the owner of tile position p is lo plus the
// starts at or before p -- what a binary search of offs gives, without global
// loads per position.
constexpr int TS_TILE = 2048, TS_WORDS = TS_TILE / 32;
static_assert(TS_WORDS == WAVE, "one wave scans the start words");
struct TileStarts {
    u32 bits[TS_WORDS];     // bit p: a node starts at base + p
    u32 before[TS_WORDS];   // starts in the words before
};

// Before the caller's barrier that publishes lo and hi.
__device__ __forceinline__ void tile_starts_clear(TileStarts &ts) {
    if (threadIdx.x < TS_WORDS) ts.bits[threadIdx.x] = 0u;
}

// Every thread of the block calls it (two barriers).
template <int T, typename O>
__device__ __forceinline__ void tile_starts_set(TileStarts &ts, const O *offs, u64 lo, u64 hi, u64 base) {
    for (u64 j = lo + 1 + threadIdx.x; j < hi; j += T) {   // base < offs[j] < base + TS_TILE
        const u32 p = (u32)((u64)offs[j] - base);
        atomicOr(&ts.bits[p >> 5], 1u << (p & 31u));
    }
    __syncthreads();
    if (threadIdx.x < WAVE) {
        u32 tot;
        ts.before[threadIdx.x] = wave_excl_scan32((u32)__popc(ts.bits[threadIdx.x]), &tot);
    }
    __syncthreads();
}

// Owner of tile position p (< TS_TILE), relative to lo.
__device__ __forceinline__ u32 tile_owner(const TileStarts &ts, u32 p) {
    return ts.before[p >> 5] + (u32)__popc(ts.bits[p >> 5] & (0xffffffffu >> (31u - (p & 31u))));
}

constexpr int WR_T = 256, WR_I = 8, WR_TILE = WR_T * WR_I;
static_assert(WR_TILE == TS_TILE, "relax tiles");

__global__ __launch_bounds__(WR_T) void wcc_relax_k(const u32 *flabel, u64 nf, const u32 *pstart,
                                                    const u64 *offs, u64 total, const u32 *nbr,
                                                    u32 *L, u32 *mark, u32 round, u32 *next_node,
                                                    u32 *next_count) {
    __shared__ u64 s_lo, s_hi;
    __shared__ u32 s_cnt[WR_I * WR_T / WAVE + 1];
    __shared__ TileStarts ts;
    const u64 base = (u64)blockIdx.x * WR_TILE;
    tile_starts_clear(ts);
    if (threadIdx.x < WAVE) {   // the first wave finds the tile's frontier nodes
        const u64 last = (base + WR_TILE < total ? base + WR_TILE : total) - 1;
        const u64 lo = wave_upper_bound(offs, 0, nf, base) - 1;
        const u64 hi = wave_upper_bound(offs, lo, nf, last);
        if (threadIdx.x == 0) {
            s_lo = lo;
            s_hi = hi;
        }
    }
    __syncthreads();
    const u64 lo = s_lo, hi = s_hi;
    tile_starts_set<WR_T>(ts, offs, lo, hi, base);   // frontier nodes have edges: offsets rise strictly
    bool app[WR_I];
    u32 m[WR_I], slot[WR_I];
#pragma unroll
    for (int r = 0; r < WR_I; r++) {
        u64 o = base + (u64)r * WR_T + threadIdx.x;
        app[r] = false;
        m[r] = 0;
        if (o < total) {
            const u64 i = lo + tile_owner(ts, (u32)(r * WR_T) + threadIdx.x);
            m[r] = nbr[(u64)pstart[i] + (o - offs[i])];
            u32 l = flabel[i];
            if (L[m[r]] > l) {
                u32 old = atomicMin(&L[m[r]], l);
                if (old > l) app[r] = atomicExch(&mark[m[r]], round) != round;
            }
        }
    }
    block_append<WR_T, WR_I>(app, next_count, s_cnt, slot);
#pragma unroll
    for (int r = 0; r < WR_I; r++)
        if (app[r]) next_node[slot[r]] = m[r];
}

constexpr int WA_T = 256, WA_I = 16, WA_TILE = WA_T * WA_I;   // node-parallel kernels: one tile per block

// Round one: L(m) = min(m, first neighbour of m); lowered nodes appended with
// their label.  Every node has a neighbour (the nodes are the edge keys).
__global__ __launch_bounds__(WA_T) void wcc_first_k(const u32 *gstart, const u32 *nbr, u64 V, u32 *L,
                                                    u32 *next_node, u32 *next_label, u32 *count) {
    __shared__ u32 s_cnt[WA_I * WA_T / WAVE + 1];
    const u64 base = (u64)blockIdx.x * WA_TILE;
    bool app[WA_I];
    u32 l[WA_I], slot[WA_I];
#pragma unroll
    for (int r = 0; r < WA_I; r++) {
        const u64 i = base + (u64)r * WA_T + threadIdx.x;
        app[r] = false;
        l[r] = 0;
        if (i < V) {
            const u32 f = nbr[gstart[i]];
            app[r] = f < (u32)i;
            l[r] = app[r] ? f : (u32)i;
            if (app[r]) L[i] = l[r];
        }
    }
    block_append<WA_T, WA_I>(app, count, s_cnt, slot);
#pragma unroll
    for (int r = 0; r < WA_I; r++)
        if (app[r]) {
            next_node[slot[r]] = (u32)(base + (u64)r * WA_T + threadIdx.x);
            next_label[slot[r]] = l[r];
        }
}

// Pull round over all ne edges (CSR gstart/nbr): Ln(m) = min(Ln(m), L(v)) for
// every edge (m, v), with Ln == L on entry.  Lanes of one node are adjacent
// (edges in node order), so a segmented minimum over the wave leaves one
// candidate per node and wave.  An edge's node comes from the tile's node
// starts (TileStarts).
constexpr int WP_T = 256, WP_I = 8, WP_TILE = WP_T * WP_I;
static_assert(WP_TILE == TS_TILE, "pull tiles");

__global__ __launch_bounds__(WP_T) void wcc_pull_k(const u32 *gstart, u64 V, u64 ne, const u32 *nbr,
                                                   const u32 *L, u32 *Ln) {
    __shared__ u64 s_lo, s_hi;
    __shared__ TileStarts ts;
    const u64 base = (u64)blockIdx.x * WP_TILE;
    const int lane = threadIdx.x & (WAVE - 1);
    tile_starts_clear(ts);
    if (threadIdx.x < WAVE) {   // the first wave finds the tile's nodes
        const u64 last = (base + WP_TILE < ne ? base + WP_TILE : ne) - 1;
        const u64 lo = wave_upper_bound(gstart, 0, V, base) - 1;
        const u64 hi = wave_upper_bound(gstart, lo, V, last);
        if (threadIdx.x == 0) {
            s_lo = lo;
            s_hi = hi;
        }
    }
    __syncthreads();
    const u64 lo = s_lo, hi = s_hi;
    tile_starts_set<WP_T>(ts, gstart, lo, hi, base);
    u32 v[WP_I];
#pragma unroll
    for (int r = 0; r < WP_I; r++) {
        const u64 o = base + (u64)r * WP_T + threadIdx.x;
        v[r] = o < ne ? nbr[o] : 0u;
    }
#pragma unroll
    for (int r = 0; r < WP_I; r++) {
        const u64 o = base + (u64)r * WP_T + threadIdx.x;
        u32 m = 0xffffffffu, val = 0xffffffffu;
        if (o < ne) {
            const u32 p = (u32)(r * WP_T) + threadIdx.x;
            m = (u32)lo + tile_owner(ts, p);
            val = L[v[r]];
        }
#pragma unroll
        for (int d = 1; d < WAVE; d <<= 1) {
            const u32 tv = __shfl_up(val, d, WAVE), tm = __shfl_up(m, d, WAVE);
            if (lane >= d && tm == m) val = min(val, tv);
        }
        const u32 nm = __shfl_down(m, 1, WAVE);
        const bool tail = o < ne && (lane == WAVE - 1 || nm != m);
        if (tail && val < Ln[m]) atomicMin(&Ln[m], val);
    }
}

// Δ' of a pull round: nodes whose label dropped, appended with the new label;
// L takes the new labels, so L == Ln again afterwards.
__global__ __launch_bounds__(WA_T) void wcc_changed_k(u32 *L, const u32 *Ln, u64 V, u32 *next_node,
                                                      u32 *next_label, u32 *count) {
    __shared__ u32 s_cnt[WA_I * WA_T / WAVE + 1];
    const u64 base = (u64)blockIdx.x * WA_TILE;
    bool app[WA_I];
    u32 l[WA_I], slot[WA_I];
#pragma unroll
    for (int r = 0; r < WA_I; r++) {
        const u64 i = base + (u64)r * WA_T + threadIdx.x;
        app[r] = false;
        l[r] = 0;
        if (i < V) {
            l[r] = Ln[i];
            app[r] = l[r] < L[i];
            if (app[r]) L[i] = l[r];
        }
    }
    block_append<WA_T, WA_I>(app, count, s_cnt, slot);
#pragma unroll
    for (int r = 0; r < WA_I; r++)
        if (app[r]) {
            next_node[slot[r]] = (u32)(base + (u64)r * WA_T + threadIdx.x);
            next_label[slot[r]] = l[r];
        }
}

__global__ void wcc_gather_k(const u32 *node, u64 n, const u32 *L, u32 *label) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < n; i += stride) label[i] = L[node[i]];
}

// Nodes per label.  A giant component's label covers most of every tile, and
// one atomic per wave on its counter would serialise over the grid: the
// tile's first label is counted in registers and added once per block; the
// other labels go through per-wave grouping.
__global__ __launch_bounds__(WA_T) void wcc_comp_count_k(const u32 *L, u64 V, u32 *ccount) {
    __shared__ u32 s_part[WA_T / WAVE];
    const u64 base = (u64)blockIdx.x * WA_TILE;
    const u32 hot = L[base];
    u32 mine = 0;
#pragma unroll
    for (int r = 0; r < WA_I; r++) {
        const u64 i = base + (u64)r * WA_T + threadIdx.x;
        const bool valid = i < V;
        const u32 k = valid ? L[i] : 0u;
        mine += (valid && k == hot) ? 1u : 0u;
        wave_count_by_key(ccount, k, valid && k != hot);
    }
#pragma unroll
    for (int d = WAVE / 2; d > 0; d >>= 1) mine += (u32)__shfl_xor((int)mine, d, WAVE);
    if ((threadIdx.x & (WAVE - 1)) == 0) s_part[threadIdx.x / WAVE] = mine;
    __syncthreads();
    if (threadIdx.x == 0) {
        u32 s = 0;
        for (int w = 0; w < WA_T / WAVE; w++) s += s_part[w];
        atomicAdd(&ccount[hot], s);
    }
}

// flags[v] = ccount[v] > 0; largest via atomicMax.
__global__ void wcc_comp_flags_k(const u32 *ccount, u64 V, u32 *flags, u32 *largest) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    u32 mx = 0;
    for (; i < V; i += stride) {
        u32 c = ccount[i];
        flags[i] = c > 0 ? 1u : 0u;
        mx = max(mx, c);
    }
#pragma unroll
    for (int d = WAVE / 2; d > 0; d >>= 1) mx = max(mx, (u32)__shfl_xor((int)mx, d, WAVE));
    if ((threadIdx.x & (WAVE - 1)) == 0 && mx) atomicMax(largest, mx);
}

// Output tuples (node key, 0) for every label in use, in node order.
__global__ void wcc_emit_k(const u32 *flags, const u64 *offs, u64 V, const u64 *rows,
                           const u32 *gstart, u64 *out) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < V; i += stride)
        if (flags[i]) out[offs[i]] = tmake(tkey(rows[gstart[i]]), 0x80000000u);
}

static u64 read_counter(dlg_ctx *ctx, WccState &s) {
    HIPCHK(hipMemcpyAsync(ctx->h_scalars, s.counter.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    return *(u32 *)ctx->h_scalars;
}

u64 wcc_first_round(dlg_ctx *ctx, WccState &s, const Index &ix, const u32 *nbr) {
    const u64 V = s.V;
    if (V == 0) return 0;
    fill_u32(ctx, s.counter.p, 1, 0u);
    // 4 B start + 4 B first neighbour per node; lowered nodes: 4 B label + 8 B appended
    DLG_LAUNCH(ctx, DLG_K_WCC_RELAX, 8.0 * V, wcc_first_k, grid_for(V, WA_TILE), WA_T, 0,
               (const u32 *)ix.gstart.p, nbr, V, s.L.p, s.fnode.p, s.flabel.p, s.counter.p);
    return read_counter(ctx, s);
}

static u64 wcc_pull_round(dlg_ctx *ctx, WccState &s, const Index &ix, const u32 *nbr) {
    const u64 V = s.V, ne = ix.n;
    s.Ln.ensure(V);
    HIPCHK(hipMemcpyAsync(s.Ln.p, s.L.p, V * 4, hipMemcpyDeviceToDevice, ctx->stream));
    fill_u32(ctx, s.counter.p, 1, 0u);
    if (ne)
        // 4 B neighbour id + 4 B label read per edge; per node 4 B start and 4 B Ln
        DLG_LAUNCH(ctx, DLG_K_WCC_RELAX, 8.0 * ne + 8.0 * V, wcc_pull_k, grid_for(ne, WP_TILE), WP_T, 0,
                   (const u32 *)ix.gstart.p, V, ne, nbr, (const u32 *)s.L.p, s.Ln.p);
    DLG_LAUNCH(ctx, DLG_K_WCC_RELAX, 8.0 * V, wcc_changed_k, grid_for(V, WA_TILE), WA_T, 0, s.L.p,
               (const u32 *)s.Ln.p, V, s.nnode.p, s.flabel.p, s.counter.p);
    s.fnode.swap(s.nnode);
    return read_counter(ctx, s);
}

// One round.  Returns |Δ'| (the new frontier size); the frontier is replaced.
u64 wcc_round(dlg_ctx *ctx, WccState &s, const Index &ix, const u32 *nbr, u64 nf, u32 round, int form) {
    if (form == WCC_PULL) return wcc_pull_round(ctx, s, ix, nbr);
    s.cnt.ensure(nf);
    s.pstart.ensure(nf);
    s.offs.ensure(nf + 1);
    DLG_LAUNCH(ctx, DLG_K_WCC_RELAX, 16.0 * nf, wcc_degree_k, sgrid(ctx, nf), ctx->block_size, 0,
               (const u32 *)s.fnode.p, nf, (const u32 *)ix.gstart.p, s.cnt.p, s.pstart.p);
    u64 total = scan_u32_u64(ctx, s.cnt.p, nf, s.offs.p);
    // a pull round reads every edge without atomics; measured on R-MAT it costs about
    // what pushing an eighth of the edges does
    if (form == WCC_AUTO && 8 * total >= (u64)ix.n) return wcc_pull_round(ctx, s, ix, nbr);
    fill_u32(ctx, s.counter.p, 1, 0u);
    if (total) {
        // 4 B neighbour id + 4 B label read + 4 B atomicMin per edge, frontier reads amortised
        DLG_LAUNCH(ctx, DLG_K_WCC_RELAX, 12.0 * total + 8.0 * nf, wcc_relax_k, grid_for(total, WR_TILE),
                   WR_T, 0, (const u32 *)s.flabel.p, nf, (const u32 *)s.pstart.p,
                   (const u64 *)s.offs.p, total, nbr, s.L.p, s.mark.p, round, s.nnode.p,
                   s.counter.p);
    }
    u64 h = 0;
    HIPCHK(hipMemcpyAsync(ctx->h_scalars, s.counter.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    h = *(u32 *)ctx->h_scalars;
    if (h) {
        DLG_LAUNCH(ctx, DLG_K_WCC_RELAX, 12.0 * h, wcc_gather_k, sgrid(ctx, h), ctx->block_size, 0,
                   (const u32 *)s.nnode.p, h, (const u32 *)s.L.p, s.flabel.p);
    }
    s.fnode.swap(s.nnode);
    return h;
}

// One thread per join row of Δ ⋈ E (load-balanced like wcc_relax_k): row r
// of probe i gives node m = value of rows[r] at cc position nbr[r] and the
// candidate (m, label of Δ[i]); cc[nbr[r]] takes the lesser.  A value that is
// no key, or a cc tuple of another node, raises *bad.
__global__ __launch_bounds__(WR_T) void wcc_tuple_relax_k(const u64 *delta, u64 nd, const u32 *pstart,
                                                          const u64 *offs, u64 total, const u64 *rows,
                                                          const u32 *nbr, u64 ncc, u64 *cc, u32 *bad) {
    __shared__ u64 s_lo, s_hi;
    const u64 base = (u64)blockIdx.x * WR_TILE;
    if (threadIdx.x < WAVE) {   // the first wave finds the tile's probes
        const u64 last = (base + WR_TILE < total ? base + WR_TILE : total) - 1;
        const u64 lo = wave_upper_bound(offs, 0, nd, base) - 1;
        const u64 hi = wave_upper_bound(offs, lo, nd, last);
        if (threadIdx.x == 0) {
            s_lo = lo;
            s_hi = hi;
        }
    }
    __syncthreads();
    const u64 lo = s_lo, hi = s_hi;
#pragma unroll
    for (int r = 0; r < WR_I; r++) {
        const u64 o = base + (u64)r * WR_T + threadIdx.x;
        if (o < total) {
            const u64 i = upper_bound_u64(offs, lo, hi, o) - 1;
            const u64 row = (u64)pstart[i] + (o - offs[i]);
            const u32 g = nbr[row];
            const u64 t = tmake(tval(rows[row]), tval(delta[i]));
            if (g >= ncc) {
                *bad = 1u;
            } else {
                const u64 cur = cc[g];
                if (tkey(cur) != tkey(t)) *bad = 1u;
                else if (cur > t) atomicMin((unsigned long long *)&cc[g], (unsigned long long)t);
            }
        }
    }
}

u64 wcc_relax_tuples(dlg_ctx *ctx, const Index &ix, const u32 *nbr, const u64 *delta, u64 nd, u64 *cc, u64 ncc,
                     u64 *next) {
    if (nd == 0 || ncc == 0 || ix.n == 0) return 0;
    DBuf<u64> old(ctx, ncc);
    copy_u64(ctx, old.p, cc, ncc);
    JoinScratch js(ctx);
    const u64 total = join_count(ctx, ix, delta, nd, js);
    DBuf<u32> bad(ctx, 1);
    fill_u32(ctx, bad.p, 1, 0u);
    if (total)
        // 4 B group + 8 B row + 8 B cc read per join row (the atomic only when lower), Δ reads amortised
        DLG_LAUNCH(ctx, DLG_K_WCC_RELAX, 20.0 * total + 8.0 * nd, wcc_tuple_relax_k, grid_for(total, WR_TILE), WR_T,
                   0, delta, nd, (const u32 *)js.pstart.p, (const u64 *)js.offs.p, total, (const u64 *)ix.rows.p,
                   nbr, ncc, cc, bad.p);
    u32 hb = 0;
    HIPCHK(hipMemcpyAsync(ctx->h_scalars, bad.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    hb = *(u32 *)ctx->h_scalars;
    if (hb) fail(DLG_INVALID_INPUT, "dlg_wcc_relax: a join row's node is not in cc (edges not symmetric, or cc "
                                    "not one tuple per key in key order)");
    // cc' keeps cc's node order, so both are sorted: Δ' = cc' minus cc
    return select(ctx, SEL_NOT_IN, cc, ncc, old.p, ncc, next);
}

void wcc_setup(dlg_ctx *ctx, WccState &s, u64 V) {
    s.V = V;
    s.L.ensure(V);
    s.mark.ensure(V);
    s.fnode.ensure(V);
    s.flabel.ensure(V);
    s.nnode.ensure(V);
    s.counter.ensure(4);
    if (V)
        DLG_LAUNCH(ctx, DLG_K_MISC, 16.0 * V, wcc_init_k, sgrid(ctx, V), ctx->block_size, 0, s.L.p, s.fnode.p,
                   s.flabel.p, s.mark.p, V);
}

// Components: returns #distinct labels, *largest, and the (label, 0) tuples in `out`.
u64 wcc_finalize(dlg_ctx *ctx, WccState &s, const Index &ix, DBuf<u64> &out, i64 *largest) {
    u64 V = s.V;
    *largest = -1;
    if (V == 0) return 0;
    DBuf<u32> ccount(ctx, V), flags(ctx, V);
    DBuf<u64> offs(ctx, V + 1);
    fill_u32(ctx, ccount.p, V, 0u);
    fill_u32(ctx, s.counter.p, 1, 0u);
    DLG_LAUNCH(ctx, DLG_K_WCC_FINAL, 12.0 * V, wcc_comp_count_k, grid_for(V, WA_TILE), WA_T, 0,
               (const u32 *)s.L.p, V, ccount.p);
    DLG_LAUNCH(ctx, DLG_K_WCC_FINAL, 8.0 * V, wcc_comp_flags_k, sgrid(ctx, V), ctx->block_size, 0,
               (const u32 *)ccount.p, V, flags.p, s.counter.p);
    u64 ncomp = scan_u32_u64(ctx, flags.p, V, offs.p);
    HIPCHK(hipMemcpyAsync(ctx->h_scalars, s.counter.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    *largest = (i64) * (u32 *)ctx->h_scalars;
    out.ensure(ncomp ? ncomp : 1);
    DLG_LAUNCH(ctx, DLG_K_WCC_FINAL, 16.0 * V, wcc_emit_k, sgrid(ctx, V), ctx->block_size, 0,
               (const u32 *)flags.p, (const u64 *)offs.p, V, (const u64 *)ix.rows.p,
               (const u32 *)ix.gstart.p, out.p);
    return ncomp;
}

}  // namespace dlg

// =====================================================================
// Graph-captured rounds.  For long runs with small frontiers the eager
// round costs more in launches and host synchronisation than in work.
// Here every size lives in device memory (WccDev), every kernel has a
// fixed grid sized for the whole graph and exits early past the live size,
// the stop rule runs on the device, and KR rounds (frontier buffers in
// ping-pong) are captured once into a hipGraph that the host replays,
// reading back one flag per replay.  No memset nodes: counters are cleared
// by kernels.  Four kernels a round: the frontier's degrees and label
// snapshot, their scan, the relax, the stop rule.
namespace dlg {

struct WccDev {
    u64 nf;      // live frontier size
    u64 total;   // edges leaving the frontier
    u32 count;   // next frontier size (append counter)
    u32 done;    // the stop rule fired
    u32 it;      // rounds run
    u32 max_it;  // 0 = none
    u64 prev;    // |Δ| of the previous round (wcc.cu:236-242, 366-383)
};

constexpr int GW_T = 256, GW_I = 8, GW_TILE = GW_T * GW_I;

// Degrees of the frontier's nodes and their block sums; each node's label is
// read here too, after the previous round's relax: the snapshot its edges
// carry this round (Jacobi).
__global__ __launch_bounds__(GW_T) void gw_degree_scan_k(const u32 *fnode, const u32 *gstart, const u32 *L,
                                                         u32 *flabel, u32 *cnt, u32 *pstart, u64 *bsum,
                                                         const WccDev *st) {
    __shared__ u64 sc[GW_T / WAVE + 1];
    const u64 nf = st->done ? 0 : st->nf;
    const u64 base = (u64)blockIdx.x * GW_TILE + (u64)threadIdx.x * GW_I;
    u64 s = 0;
#pragma unroll
    for (int i = 0; i < GW_I; i++) {
        u64 j = base + i;
        if (j < nf) {
            u32 v = fnode[j];
            u32 a = gstart[v];
            u32 c = gstart[v + 1] - a;
            cnt[j] = c;
            pstart[j] = a;
            flabel[j] = L[v];
            s += c;
        }
    }
    u64 tot;
    block_exclusive_scan<u64>(s, sc, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Offsets of the frontier's edges: each workgroup adds the block sums before
// its own (read and summed by the whole workgroup) to a scan of its degrees;
// the last workgroup also writes the round's edge total.
__global__ __launch_bounds__(GW_T) void gw_scan_apply_k(const u32 *cnt, const u64 *bsum, u64 *offs, WccDev *st) {
    __shared__ u64 sc[GW_T / WAVE + 1];
    const u64 nf = st->done ? 0 : st->nf;
    const bool last = blockIdx.x == gridDim.x - 1;
    const u64 base = (u64)blockIdx.x * GW_TILE + (u64)threadIdx.x * GW_I;
    if ((u64)blockIdx.x * GW_TILE >= nf && !last) return;   // block-uniform: no frontier nodes here
    u64 p = 0;
    for (u64 i = threadIdx.x; i < blockIdx.x; i += GW_T) p += bsum[i];
    u64 pre;
    block_exclusive_scan<u64>(p, sc, &pre);
    __syncthreads();   // sc is reused
    if (last && threadIdx.x == 0) st->total = pre + bsum[blockIdx.x];   // 0 when the rule has fired
    u32 v[GW_I];
    u64 s = 0;
#pragma unroll
    for (int i = 0; i < GW_I; i++) {
        v[i] = base + i < nf ? cnt[base + i] : 0;
        s += v[i];
    }
    u64 ex = block_exclusive_scan<u64>(s, sc, nullptr) + pre;
#pragma unroll
    for (int i = 0; i < GW_I; i++) {
        if (base + i < nf) offs[base + i] = ex;
        ex += v[i];
    }
}

__global__ __launch_bounds__(WR_T) void gw_relax_k(const u32 *flabel, const u32 *pstart, const u64 *offs,
                                                   const u32 *nbr, u32 *L, u32 *mark, u32 *next_node,
                                                   WccDev *st) {
    __shared__ u64 s_lo, s_hi;
    __shared__ u32 s_cnt[WR_I * WR_T / WAVE + 1];
    __shared__ TileStarts ts;
    const u64 total = st->total, nf = st->nf;
    const u64 base = (u64)blockIdx.x * WR_TILE;
    if (st->done || base >= total) return;   // block-uniform
    const u32 round = st->it + 1;
    tile_starts_clear(ts);
    if (threadIdx.x < WAVE) {   // the first wave finds the tile's frontier nodes
        const u64 last = (base + WR_TILE < total ? base + WR_TILE : total) - 1;
        const u64 lo = wave_upper_bound(offs, 0, nf, base) - 1;
        const u64 hi = wave_upper_bound(offs, lo, nf, last);
        if (threadIdx.x == 0) {
            s_lo = lo;
            s_hi = hi;
        }
    }
    __syncthreads();
    const u64 lo = s_lo, hi = s_hi;
    tile_starts_set<WR_T>(ts, offs, lo, hi, base);   // frontier nodes have edges: offsets rise strictly
    bool app[WR_I];
    u32 m[WR_I], slot[WR_I];
#pragma unroll
    for (int r = 0; r < WR_I; r++) {
        u64 o = base + (u64)r * WR_T + threadIdx.x;
        app[r] = false;
        m[r] = 0;
        if (o < total) {
            const u64 i = lo + tile_owner(ts, (u32)(r * WR_T) + threadIdx.x);
            m[r] = nbr[(u64)pstart[i] + (o - offs[i])];
            u32 l = flabel[i];
            if (L[m[r]] > l) {
                u32 old = atomicMin(&L[m[r]], l);
                if (old > l) app[r] = atomicExch(&mark[m[r]], round) != round;
            }
        }
    }
    block_append<WR_T, WR_I>(app, &st->count, s_cnt, slot);
#pragma unroll
    for (int r = 0; r < WR_I; r++)
        if (app[r]) next_node[slot[r]] = m[r];
}

// The reference's stop rule, on the device: count the round, stop when two
// consecutive delta sizes are equal.  The append counter is cleared for the
// next round.
__global__ void gw_end_k(WccDev *st) {
    if (threadIdx.x != 0 || blockIdx.x != 0 || st->done) return;
    u64 cur = st->count;
    st->count = 0;
    st->it += 1;
    st->nf = cur;
    if (cur == st->prev || (st->max_it && st->it >= st->max_it)) st->done = 1;
    else st->prev = cur;
}

static void gw_capture_round(dlg_ctx *ctx, WccState &s, const Index &ix, const u32 *nbr, u64 V, u64 E,
                             WccDev *st, u32 *fnode, u32 *flabel, u32 *nnode, u64 *bsum) {
    const unsigned gb = grid_for(V, GW_TILE);   // bsum holds gb block sums
    hipLaunchKernelGGL(gw_degree_scan_k, dim3(gb), dim3(GW_T), 0, ctx->stream, (const u32 *)fnode,
                       (const u32 *)ix.gstart.p, (const u32 *)s.L.p, flabel, s.cnt.p, s.pstart.p, bsum,
                       (const WccDev *)st);
    hipLaunchKernelGGL(gw_scan_apply_k, dim3(gb), dim3(GW_T), 0, ctx->stream, (const u32 *)s.cnt.p,
                       (const u64 *)bsum, s.offs.p, st);
    hipLaunchKernelGGL(gw_relax_k, dim3(grid_for(E, WR_TILE)), dim3(WR_T), 0, ctx->stream,
                       (const u32 *)flabel, (const u32 *)s.pstart.p, (const u64 *)s.offs.p, nbr, s.L.p,
                       s.mark.p, nnode, st);
    hipLaunchKernelGGL(gw_end_k, dim3(1), dim3(64), 0, ctx->stream, st);
}

// Runs the rounds of wcc_solve with captured graphs of `kr` rounds (even).
// Returns the iteration count; the labels are left in s.L.
int wcc_rounds_graph(dlg_ctx *ctx, WccState &s, const Index &ix, const u32 *nbr, u64 ne, int max_it,
                     int kr) {
    const u64 V = s.V;
    if (V == 0) return 1;   // empty graph: the first round already sees 0 == 0
    if (kr < 2) kr = 2;
    kr &= ~1;
    s.cnt.ensure(V);
    s.pstart.ensure(V);
    s.offs.ensure(V + 1);
    DBuf<u32> nlabel(ctx, V);
    DBuf<u64> bsum(ctx, (V + GW_TILE - 1) / GW_TILE + 1);
    DBuf<u64> stbuf(ctx, (sizeof(WccDev) + 7) / 8);
    WccDev *st = (WccDev *)stbuf.p;
    WccDev h;
    std::memset(&h, 0, sizeof(h));
    h.nf = V;
    h.prev = V;
    h.max_it = (u32)(max_it > 0 ? max_it : 0);
    HIPCHK(hipMemcpyAsync(st, &h, sizeof(h), hipMemcpyHostToDevice, ctx->stream));
    ctx->sync();
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool prof = ctx->prof;
    ctx->prof = false;
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    for (int r = 0; r < kr; r++) {
        if (r % 2 == 0)
            gw_capture_round(ctx, s, ix, nbr, V, ne, st, s.fnode.p, s.flabel.p, s.nnode.p, bsum.p);
        else
            gw_capture_round(ctx, s, ix, nbr, V, ne, st, s.nnode.p, nlabel.p, s.fnode.p, bsum.p);
    }
    hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
    ctx->prof = prof;
    HIPCHK(ce);
    hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
        (void)hipGraphDestroy(graph);
        HIPCHK(ie);
    }
    int it = 0;
    try {
        while (true) {
            if (ctx->prof) ctx->prof_begin();
            HIPCHK(hipGraphLaunch(exec, ctx->stream));
            if (ctx->prof) ctx->prof_end(DLG_K_WCC_RELAX, 0.0);
            HIPCHK(hipMemcpyAsync(ctx->h_scalars, st, sizeof(WccDev), hipMemcpyDeviceToHost, ctx->stream));
            ctx->sync();
            WccDev cur;
            std::memcpy(&cur, ctx->h_scalars, sizeof(cur));
            it = (int)cur.it;
            if (cur.done) break;
        }
    } catch (...) {
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        throw;
    }
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    return it;
}

}  // namespace dlg
