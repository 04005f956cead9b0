// k_join.hip -- hash index build and the two-pass hash join.
//
// Reference roles: get_hash_table + build_hash_table_entity
// (common/hash_table.cu:1-55) and get_join + get_join_result_size_entity +
// get_join_result_entity (common/join.cu:1-87).
//
// Layout in HBM (one index):
//   rows[n]     the relation sorted by (key, value), duplicates kept (multimap)
//   gstart[G+1] first row of each distinct key ("group"), gstart[G] = n
//   slots[S]    open-addressed table, S = next_pow2(ceil(G / load_factor)),
//               slot = (key << 32) | group, linear probing, EMPTY = ~0
// The reference stores one (key, value) slot per ROW and walks the whole run
// of equal keys in both join passes; here one slot per distinct KEY is found
// once per probe and the matching values are a contiguous range of `rows`,
// so the materialize pass reads them coalesced and the multiset of join rows
// is exactly the reference's (one output per relation row with an equal key).
//
// Join, per probe tuple (k, v):
//   count   : slot lookup -> (row start, count)                 [join_count_k]
//   offsets : exclusive scan of counts (u32 -> u64)             [scan]
//   expand  : one thread per OUTPUT row (load-balanced): binary search of the
//             output index in the offsets, then out = (rows[start+j].value, v)
//             or its reverse                                     [join_expand_k]
#include "dlg_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace dlg {

__device__ __forceinline__ u32 ix_find(const u64 *slots, u64 mask, u32 key) {
    u64 p = fmix32(key) & mask;
    while (true) {
        u64 s = slots[p];
        if (s == EMPTY_SLOT) return 0xffffffffu;
        if (tkey(s) == key) return tval(s);
        p = (p + 1) & mask;
    }
}

__global__ void set_last_k(u32 *gstart, u64 g, u32 n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) gstart[g] = n;
}

__global__ void hash_insert_k(const u64 *rows, const u32 *gstart, u64 G, u64 *slots, u64 mask) {
    u64 g = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; g < G; g += stride) {
        u32 key = tkey(rows[gstart[g]]);
        u64 v = tmake(key, (u32)g);
        u64 p = fmix32(key) & mask;
        while (true) {
            unsigned long long prev =
                atomicCAS((unsigned long long *)&slots[p], (unsigned long long)EMPTY_SLOT,
                          (unsigned long long)v);
            if (prev == EMPTY_SLOT) break;
            p = (p + 1) & mask;
        }
    }
}

void index_build(dlg_ctx *ctx, Index &ix, const u64 *rel, u64 n, double lf, bool presorted,
                 bool sorted_lookup) {
    ix.n = n;
    ix.groups = 0;
    ix.nslots = 0;
    ix.sorted = sorted_lookup;
    ix.dense_w = 0;
    if (n == 0) return;
    if (n >= 0xffffffffull) fail(DLG_OVERFLOW, "index: relation larger than 2^32-2 rows");
    if (!(lf > 0.0)) lf = ctx->load_factor;
    lf = std::min(0.95, std::max(0.05, lf));
    ix.rows.ensure(n);
    if (rel != ix.rows.p) copy_u64(ctx, ix.rows.p, rel, n);
    if (!presorted) {
        DBuf<u64> tmp(ctx);
        radix_sort(ctx, ix.rows, n, tmp);
    }
    ix.gstart.ensure(n + 1);
    u64 G = select(ctx, SEL_KEY_HEAD, ix.rows.p, n, nullptr, 0, ix.gstart.p);
    DLG_LAUNCH(ctx, DLG_K_MISC, 4.0, set_last_k, 1, 64, 0, ix.gstart.p, G, (u32)n);
    ix.groups = G;
    if (sorted_lookup) return;
    index_hash(ctx, ix, lf);
}

void index_hash(dlg_ctx *ctx, Index &ix, double lf) {
    const u64 G = ix.groups;
    if (G == 0) return;
    if (!(lf > 0.0)) lf = ctx->load_factor;
    lf = std::min(0.95, std::max(0.05, lf));
    ix.sorted = false;
    u64 want = (u64)std::ceil((double)G / lf);
    u64 S = 64;
    while (S < want || S <= G) S <<= 1;
    ix.nslots = S;
    ix.slots.ensure(S);
    fill_u64(ctx, ix.slots.p, S, EMPTY_SLOT);
    DLG_LAUNCH(ctx, DLG_K_HASH_BUILD, 8.0 * S + 12.0 * G + 8.0 * G, hash_insert_k, sgrid(ctx, G),
               ctx->block_size, 0,
               (const u64 *)ix.rows.p, (const u32 *)ix.gstart.p, G, ix.slots.p, S - 1);
}

__global__ __launch_bounds__(1024) void join_count_k(const u64 *probe, u64 n, const u64 *slots,
                                                    u64 mask, const u32 *gstart, u32 *cnt,
                                                    u32 *pstart) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        u32 g = ix_find(slots, mask, tkey(probe[i]));
        u32 c = 0, s = 0;
        if (g != 0xffffffffu) {
            s = gstart[g];
            c = gstart[g + 1] - s;
        }
        cnt[i] = c;
        pstart[i] = s;
    }
}

// Dense-offset variant: the key's rows are rows[dense[k - lo] .. dense[k - lo + 1]).
__global__ __launch_bounds__(1024) void join_count_dense_k(const u64 *probe, u64 n, const u64 *dense, u32 lo,
                                                          u64 W, u32 *cnt, u32 *pstart) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        const u32 k = tkey(probe[i]) - lo;
        u32 c = 0, st = 0;
        if (k < W) {
            st = (u32)dense[k];
            c = (u32)(dense[k + 1] - st);
        }
        cnt[i] = c;
        pstart[i] = st;
    }
}
// dense[k] = the first row whose key is at least lo + k, k in [0, W]: one
// binary search per key over the sorted rows.
__global__ __launch_bounds__(256) void dense_bounds_k(const u64 *rows, u64 n, u32 lo, u64 W, u64 *dense) {
    const u64 k = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > W) return;
    const u64 key = (u64)lo + k;
    u64 a = 0, b = n;
    while (a < b) {
        const u64 m = (a + b) >> 1;
        if ((u64)tkey(rows[m]) < key) a = m + 1; else b = m;
    }
    dense[k] = a;
}

void index_dense(dlg_ctx *ctx, Index &ix, u32 lo, u64 W) {
    ix.dense_w = 0;
    if (ix.n == 0 || W == 0) return;
    ix.dense.ensure(W + 1);
    DLG_LAUNCH(ctx, DLG_K_MISC, 8.0 * (W + 1), dense_bounds_k, grid_for(W + 1, 256), 256, 0,
               (const u64 *)ix.rows.p, ix.n, lo, W, ix.dense.p);
    ix.dense_lo = lo;
    ix.dense_w = W;
}

// Sorted-index variant (common/join.cu:89-142): the probe key's run of rows
// by two binary searches over the sorted relation.
__global__ __launch_bounds__(1024) void join_count_sorted_k(const u64 *probe, u64 n, const u64 *rows, u64 nr,
                                                           u32 *cnt, u32 *pstart) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        u64 k = (u64)tkey(probe[i]) << 32;
        u64 s = lower_bound_u64(rows, 0, nr, k);
        u64 e = upper_bound_u64(rows, s, nr, k | 0xffffffffull);
        cnt[i] = (u32)(e - s);
        pstart[i] = (u32)s;
    }
}

// The count kernel of the index's lookup kind: cnt[i], pstart[i] per probe.
static void join_count_launch(dlg_ctx *ctx, const Index &ix, const u64 *probe, u64 n, u32 *cnt, u32 *pstart) {
    if (ix.sorted) {
        // probe read + ~2 log2(n) row reads per probe + count/start writes
        double lg = std::log2((double)ix.n + 1.0);
        DLG_LAUNCH(ctx, DLG_K_JOIN_COUNT, 8.0 * n + 16.0 * lg * n + 8.0 * n, join_count_sorted_k, sgrid(ctx, n),
                   ctx->block_size, 0, probe, n, (const u64 *)ix.rows.p, ix.n, cnt, pstart);
    } else if (ix.dense_w) {
        // probe read + an offset pair per probe + count/start writes
        DLG_LAUNCH(ctx, DLG_K_JOIN_COUNT, 8.0 * n + 16.0 * n + 8.0 * n, join_count_dense_k, sgrid(ctx, n),
                   ctx->block_size, 0, probe, n, (const u64 *)ix.dense.p, ix.dense_lo, ix.dense_w, cnt, pstart);
    } else {
        // probe read + one slot line per probe + gstart pair + count/start writes
        DLG_LAUNCH(ctx, DLG_K_JOIN_COUNT, 8.0 * n + 8.0 * n + 8.0 * n + 8.0 * n, join_count_k, sgrid(ctx, n),
                   ctx->block_size, 0, probe, n, (const u64 *)ix.slots.p, ix.nslots - 1, (const u32 *)ix.gstart.p,
                   cnt, pstart);
    }
}

u64 join_count(dlg_ctx *ctx, const Index &ix, const u64 *probe, u64 n, JoinScratch &js) {
    if (n == 0 || ix.n == 0) return 0;
    js.cnt.ensure(n);
    js.pstart.ensure(n);
    js.offs.ensure(n + 1);
    join_count_launch(ctx, ix, probe, n, js.cnt.p, js.pstart.p);
    return scan_u32_u64(ctx, js.cnt.p, n, js.offs.p);
}

u64 join_count_with(dlg_ctx *ctx, const Index &ix, const u64 *probe, u64 n, u32 *cnt, u32 *pstart, u64 *offs,
                    u64 *bsum, u64 *d_total) {
    if (n == 0 || ix.n == 0) return 0;
    join_count_launch(ctx, ix, probe, n, cnt, pstart);
    scan_u32_u64_with(ctx, cnt, n, offs, bsum, d_total);
    u64 h = 0;
    ctx->read_scalars(d_total, 1, &h);
    return h;
}

constexpr int EX_T = 256, EX_I = 8, EX_TILE = EX_T * EX_I;

// The probes whose output ranges meet this block's EX_TILE outputs are
// [lo, hi); when at most EX_TILE + 1 of them (every probe with a non-empty
// range in the tile adds at least one output), their offsets, row starts and
// values are staged in LDS and each output finds its probe by a binary
// search in LDS.  Tiles spanning more probes (runs of empty probes) search
// global memory.
constexpr int EX_LDS = EX_TILE + 1;

template <int PROJ>
__global__ __launch_bounds__(EX_T) void join_expand_k(const u64 *probe, u64 np, const u32 *pstart,
                                                      const u64 *offs, u64 total, const u64 *rows,
                                                      u64 *out) {
    __shared__ u64 s_lo, s_hi;
    __shared__ u64 l_offs[EX_LDS];
    __shared__ u32 l_start[EX_LDS];
    __shared__ u32 l_pv[EX_LDS];
    const u64 base = (u64)blockIdx.x * EX_TILE;
    if (threadIdx.x < WAVE) {   // the first wave finds the block's probes
        const u64 last = (base + EX_TILE < total ? base + EX_TILE : total) - 1;
        const u64 lo = wave_upper_bound(offs, 0, np, base) - 1;
        const u64 hi = wave_upper_bound(offs, lo, np, last);
        if (threadIdx.x == 0) {
            s_lo = lo;
            s_hi = hi;
        }
    }
    __syncthreads();
    const u64 lo = s_lo, hi = s_hi;
    const u64 span = hi - lo;
    if (span <= (u64)EX_LDS) {
        for (u64 j = threadIdx.x; j < span; j += EX_T) {
            l_offs[j] = offs[lo + j];
            l_start[j] = pstart[lo + j];
            l_pv[j] = tval(probe[lo + j]);
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < EX_I; r++) {
            u64 o = base + (u64)r * EX_T + threadIdx.x;
            if (o < total) {
                u32 a = 0, b = (u32)span;   // last j with l_offs[j] <= o
                while (b - a > 1) {
                    u32 m = (a + b) >> 1;
                    if (l_offs[m] <= o) a = m; else b = m;
                }
                u32 hv = tval(rows[(u64)l_start[a] + (o - l_offs[a])]);
                u32 pv = l_pv[a];
                out[o] = PROJ == DLG_PROJ_HV_PV ? tmake(hv, pv) : tmake(pv, hv);
            }
        }
        return;
    }
#pragma unroll
    for (int r = 0; r < EX_I; r++) {
        u64 o = base + (u64)r * EX_T + threadIdx.x;
        if (o < total) {
            u64 i = upper_bound_u64(offs, lo, hi, o) - 1;
            u32 hv = tval(rows[(u64)pstart[i] + (o - offs[i])]);
            u32 pv = tval(probe[i]);
            out[o] = PROJ == DLG_PROJ_HV_PV ? tmake(hv, pv) : tmake(pv, hv);
        }
    }
}

void join_expand_with(dlg_ctx *ctx, const Index &ix, const u64 *probe, u64 n, const u32 *pstart, const u64 *offs,
                      u64 total, int proj, u64 *out) {
    if (total == 0) return;
    unsigned g = grid_for(total, EX_TILE);
    double bytes = 24.0 * total;  // relation value read, output write, probe/offset reads amortised
    if (proj == DLG_PROJ_PV_HV)
        DLG_LAUNCH(ctx, DLG_K_JOIN_EXPAND, bytes, join_expand_k<DLG_PROJ_PV_HV>, g, EX_T, 0, probe, n, pstart, offs,
                   total, (const u64 *)ix.rows.p, out);
    else
        DLG_LAUNCH(ctx, DLG_K_JOIN_EXPAND, bytes, join_expand_k<DLG_PROJ_HV_PV>, g, EX_T, 0, probe, n, pstart, offs,
                   total, (const u64 *)ix.rows.p, out);
}

void join_expand(dlg_ctx *ctx, const Index &ix, const u64 *probe, u64 n, const JoinScratch &js,
                 u64 total, int proj, u64 *out) {
    join_expand_with(ctx, ix, probe, n, js.pstart.p, js.offs.p, total, proj, out);
}

// Rows whose reverse is a row too (binary search in the sorted rows),
// counted per wave.
__global__ void reciprocal_k(const u64 *rows, u64 n, unsigned long long *count) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    const u64 lim = (n + WAVE - 1) / WAVE * WAVE;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < lim; i += stride) {
        bool f = false;
        if (i < n) {
            const u64 t = rows[i], r = tmake(tval(t), tkey(t));
            const u64 p = lower_bound_u64(rows, 0, n, r);
            f = p < n && rows[p] == r;
        }
        const u64 m = __ballot(f);
        if ((threadIdx.x & (WAVE - 1)) == 0 && m) atomicAdd(count, (unsigned long long)__popcll(m));
    }
}

double reciprocal_fraction(dlg_ctx *ctx, const Index &ix) {
    if (ix.n == 0) return 0.0;
    DBuf<u64> c(ctx, 1);
    fill_u64(ctx, c.p, 1, 0ull);
    DLG_LAUNCH(ctx, DLG_K_MISC, 8.0 * ix.n, reciprocal_k, sgrid(ctx, ix.n), ctx->block_size, 0,
               (const u64 *)ix.rows.p, ix.n, (unsigned long long *)c.p);
    u64 h = 0;
    ctx->read_scalars(c.p, 1, &h);
    return (double)h / (double)ix.n;
}

__global__ void lookup_values_k(const u64 *rows, u64 n, const u64 *slots, u64 mask, u32 *out) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < n; i += stride) out[i] = ix_find(slots, mask, tval(rows[i]));
}

// Group of each row's value by binary search: its first row, then that row's
// group among the group starts.
__global__ void lookup_values_sorted_k(const u64 *rows, u64 n, const u32 *gstart, u64 G, u32 *out) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        u64 k = (u64)tval(rows[i]) << 32;
        u64 r = lower_bound_u64(rows, 0, n, k);
        u32 g = 0xffffffffu;
        if (r < n && (rows[r] >> 32) == (k >> 32)) {
            u64 lo = 0, hi = G;   // the group whose start is r
            while (lo < hi) {
                u64 mid = (lo + hi) >> 1;
                if (gstart[mid] < r) lo = mid + 1; else hi = mid;
            }
            g = (u32)lo;
        }
        out[i] = g;
    }
}

// Rank directory over a dense key range [lo, lo + W): entry j holds the
// number of keys below id lo + 32j (high half) and the presence bits of ids
// lo + 32j .. lo + 32j + 31 (low half).  A key's group is its rank among the
// keys (groups are in key order), so a lookup is one 8-byte read of a table
// of W/4 bytes (2 MiB for 2^23 ids: it stays in L2) and a popcount.
//
// Presence bits, one thread per group: the groups' keys ascend, so lanes of
// one word are adjacent; a segmented OR leaves one atomic per word and wave.
__global__ void rank_bits_k(const u64 *rows, const u32 *gstart, u64 G, u32 lo, u32 *bits) {
    const int lane = threadIdx.x & (WAVE - 1);
    const u64 stride = (u64)gridDim.x * blockDim.x;
    const u64 lim = (G + WAVE - 1) / WAVE * WAVE;   // whole waves iterate together
    for (u64 g = (u64)blockIdx.x * blockDim.x + threadIdx.x; g < lim; g += stride) {
        u32 w = 0xffffffffu, b = 0;
        if (g < G) {
            const u32 d = tkey(rows[gstart[g]]) - lo;
            w = d >> 5;
            b = 1u << (d & 31);
        }
#pragma unroll
        for (int s = 1; s < WAVE; s <<= 1) {
            const u32 tb = __shfl_up(b, s, WAVE), tw = __shfl_up(w, s, WAVE);
            if (lane >= s && tw == w) b |= tb;
        }
        const u32 nw = __shfl_down(w, 1, WAVE);
        if (g < G && (lane == WAVE - 1 || nw != w)) atomicOr(&bits[w], b);
    }
}

__global__ void rank_count_k(const u32 *bits, u64 nw, u32 *cnt) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x; j < nw; j += stride) cnt[j] = (u32)__popc(bits[j]);
}

__global__ void rank_pack_k(const u32 *bits, const u64 *below, u64 nw, u64 *dir) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x; j < nw; j += stride)
        dir[j] = (below[j] << 32) | bits[j];
}

__global__ void lookup_values_rank_k(const u64 *rows, u64 n, const u64 *dir, u32 lo, u64 W, u32 *out) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const u32 d = tval(rows[i]) - lo;
        u32 g = 0xffffffffu;
        if ((u64)d < W) {
            const u64 e = dir[d >> 5];
            const u32 b = (u32)e, bit = d & 31;
            if ((b >> bit) & 1u) g = (u32)(e >> 32) + (u32)__popc(b & ((1u << bit) - 1u));
        }
        out[i] = g;
    }
}

// Rank lookups when the key range is at most 64 ids per key (plus 2^20):
// the directory then takes at most 2 bytes per key over the hash table's slots.
constexpr u64 RANK_IDS_PER_KEY = 64, RANK_SLACK = 1ull << 20;

bool index_rank_lookup_applies(dlg_ctx *ctx, const Index &ix) {
    if (ix.n == 0) return true;
    u64 ends[2];
    HIPCHK(hipMemcpyAsync(ctx->h_scalars, ix.rows.p, 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync((char *)ctx->h_scalars + 8, ix.rows.p + (ix.n - 1), 8, hipMemcpyDeviceToHost,
                          ctx->stream));
    ctx->sync();
    std::memcpy(ends, ctx->h_scalars, 16);
    const u64 W = (u64)(u32)(ends[1] >> 32) - (u32)(ends[0] >> 32) + 1;   // rows are sorted
    return W <= RANK_IDS_PER_KEY * ix.groups + RANK_SLACK;
}

void index_lookup_values(dlg_ctx *ctx, const Index &ix, u32 *out) {
    if (ix.n == 0) return;
    if (ix.sorted) {
        DLG_LAUNCH(ctx, DLG_K_MISC, 20.0 * ix.n, lookup_values_sorted_k, sgrid(ctx, ix.n), ctx->block_size, 0,
                   (const u64 *)ix.rows.p, ix.n, (const u32 *)ix.gstart.p, ix.groups, out);
        return;
    }
    u64 ends[2];
    HIPCHK(hipMemcpyAsync(ctx->h_scalars, ix.rows.p, 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync((char *)ctx->h_scalars + 8, ix.rows.p + (ix.n - 1), 8, hipMemcpyDeviceToHost,
                          ctx->stream));
    ctx->sync();
    std::memcpy(ends, ctx->h_scalars, 16);
    const u32 lo = (u32)(ends[0] >> 32);
    const u64 W = (u64)(u32)(ends[1] >> 32) - lo + 1;   // rows are sorted: first and last key
    if (W <= RANK_IDS_PER_KEY * ix.groups + RANK_SLACK) {
        const u64 nw = (W + 31) / 32;
        DBuf<u32> bits(ctx, nw), cnt(ctx, nw);
        DBuf<u64> below(ctx, nw + 1), dir(ctx, nw);
        fill_u32(ctx, bits.p, nw, 0u);
        DLG_LAUNCH(ctx, DLG_K_MISC, 12.0 * ix.groups, rank_bits_k, sgrid(ctx, ix.groups), ctx->block_size, 0,
                   (const u64 *)ix.rows.p, (const u32 *)ix.gstart.p, ix.groups, lo, bits.p);
        DLG_LAUNCH(ctx, DLG_K_MISC, 8.0 * nw, rank_count_k, sgrid(ctx, nw), ctx->block_size, 0,
                   (const u32 *)bits.p, nw, cnt.p);
        scan_u32_u64(ctx, cnt.p, nw, below.p);
        DLG_LAUNCH(ctx, DLG_K_MISC, 20.0 * nw, rank_pack_k, sgrid(ctx, nw), ctx->block_size, 0,
                   (const u32 *)bits.p, (const u64 *)below.p, nw, dir.p);
        // 8 B row + 4 B group per row; the directory reads hit L2
        DLG_LAUNCH(ctx, DLG_K_MISC, 12.0 * ix.n, lookup_values_rank_k, sgrid(ctx, ix.n), ctx->block_size, 0,
                   (const u64 *)ix.rows.p, ix.n, (const u64 *)dir.p, lo, W, out);
        return;
    }
    if (ix.nslots == 0) fail(DLG_DEVICE_ERROR, "index: value lookup by hash probe without slots");
    DLG_LAUNCH(ctx, DLG_K_MISC, 20.0 * ix.n, lookup_values_k, sgrid(ctx, ix.n), ctx->block_size, 0,
               (const u64 *)ix.rows.p,
               ix.n, (const u64 *)ix.slots.p, ix.nslots - 1, out);
}

}  // namespace dlg
