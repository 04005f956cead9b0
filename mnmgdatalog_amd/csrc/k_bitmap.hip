// k_bitmap.hip -- the full relation as a bitmap over the id range.
//
// Every tuple of a TC / SG fixpoint is a pair of node ids of the input, so
// with W = the width of the id range, tuple (k, v) has the bit
//     (k - lo) * W + (v - lo)            (row-major), or
//     (v - lo) * W + (k - lo)            (transposed)
// of a W*W-bit map (W = 46,656 ids -> 272 MB for the SG mesh).  The
// relational steps of a round collapse into the join's materialize pass:
//   for every join row t:  if bit(t) is clear, atomicOr it; the thread whose
//   atomicOr flipped the bit appends t to N
// which is unique(J) \ F and F = F u N at once (tc.cu:250-283): J is never
// written, sorted or merged.  The bit is read plainly first, so rows already
// in F (most of them once a fixpoint is under way) cost one read, no atomic.
//
// A probe's join rows share the probe's value and differ in the matched
// row's value; the layout puts the probe's half in the bitmap row, so one
// probe's rows land in one bitmap row, usually in a few words.  Lanes of a
// wave that hit the same word OR their bits together and issue one atomic
// (a segmented OR scan over the wave), and the new tuples gather in LDS and
// go out with one counter atomic per ~1K tuples.  N comes out unordered;
// reading the bitmap in bit order emits the final relation (row-major:
// already sorted; transposed: sorted afterwards).
#include "dlg_internal.h"

#include <algorithm>

namespace dlg {

__device__ __forceinline__ u64 bm_bit(u64 t, u32 lo, u64 W, bool tr) {
    u64 a = (u64)(tkey(t) - lo), b = (u64)(tval(t) - lo);
    return tr ? b * W + a : a * W + b;
}

__device__ __forceinline__ bool bm_set(u32 *words, u64 b) {
    u32 mask = 1u << (b & 31);
    u32 *w = words + (b >> 5);
    if (*w & mask) return false;                 // already in F (bits are never cleared mid-round)
    return (atomicOr(w, mask) & mask) == 0;      // this thread set it: a new tuple
}

__device__ __forceinline__ void wave_append(bool app, u64 t, u64 *out, unsigned long long *counter) {
    const int lane = threadIdx.x & (WAVE - 1);
    u64 bm = __ballot(app);
    if (!bm) return;
    int leader = __ffsll((unsigned long long)bm) - 1;
    unsigned long long b = 0;
    if (lane == leader) b = atomicAdd(counter, (unsigned long long)__popcll(bm));
    b = __shfl(b, leader, WAVE);
    if (app) out[b + (u64)__popcll(bm & lanemask_lt())] = t;
}

// Whole-wave insert: lane l offers bit `mask` (0 = nothing) of word `wi`.
// Runs of lanes with the same word combine their bits (segmented inclusive
// OR scan), the run's last lane issues the atomicOr, and a lane's tuple is
// new when its bit was clear before that atomic and no earlier lane of the
// run offered the same bit.  Every lane of the wave must call it.
__device__ __forceinline__ bool wave_insert(u32 *words, u64 wi, u32 mask) {
    const int lane = threadIdx.x & (WAVE - 1);
    u32 wlo = (u32)wi, whi = (u32)(wi >> 32);
    u32 plo = __shfl_up(wlo, 1, WAVE), phi = __shfl_up(whi, 1, WAVE);
    bool head = lane == 0 || plo != wlo || phi != whi;
    u32 v = mask;
    int f = head ? 1 : 0;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        u32 vv = __shfl_up(v, d, WAVE);
        int ff = __shfl_up(f, d, WAVE);
        if (lane >= d && !f) {
            v |= vv;
            f = ff;
        }
    }
    u32 excl = __shfl_up(v, 1, WAVE);
    if (head) excl = 0;
    u64 heads = __ballot(head);
    u64 tails = (heads >> 1) | (1ull << (WAVE - 1));
    int tail = lane + __ffsll((unsigned long long)(tails >> lane)) - 1;
    u32 old = 0;
    if (lane == tail && v) old = atomicOr(words + wi, v);
    old = __shfl(old, tail, WAVE);
    return mask && !(old & mask) && !(excl & mask);
}

constexpr int BJ_T = 256, BJ_I = 8, BJ_TILE = BJ_T * BJ_I, BJ_LDS = BJ_TILE + 1;
constexpr int BJ_BUF = BJ_TILE + 1024;   // new tuples staged per block; flushed above 1024

// Join materialize fused with the bitmap insert.  Persistent blocks walk
// tiles of BJ_TILE join rows (probe location as in join_expand_k: the tile's
// probes staged in LDS when they fit, a binary search per row).
template <int PROJ, bool NODIAG, bool TR>
__global__ __launch_bounds__(BJ_T) void bm_join_insert_k(const u64 *probe, u64 np, const u32 *pstart,
                                                         const u64 *offs, u64 total, const u64 *rows,
                                                         u32 *words, u32 lo, u64 W, u64 *out,
                                                         unsigned long long *counter) {
    __shared__ u64 s_lo, s_hi;
    __shared__ unsigned long long s_base;
    __shared__ u32 s_cnt;
    __shared__ u32 l_offs[BJ_LDS];    // probe offsets relative to the tile base (the first: 0)
    __shared__ u32 l_start[BJ_LDS];   // row of the probe's match at l_offs (the first: adjusted)
    __shared__ u32 l_pv[BJ_LDS];
    __shared__ u64 buf[BJ_BUF];
    const u64 ntiles = (total + BJ_TILE - 1) / BJ_TILE;
    if (threadIdx.x == 0) s_cnt = 0;
    for (u64 tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const u64 base = tile * BJ_TILE;
        __syncthreads();
        if (threadIdx.x == 0) {
            u64 last = (base + BJ_TILE < total ? base + BJ_TILE : total) - 1;
            u64 l = upper_bound_u64(offs, 0, np, base) - 1;
            s_lo = l;
            s_hi = upper_bound_u64(offs, l, np, last);
        }
        __syncthreads();
        const u64 plo = s_lo, phi = s_hi, span = phi - plo;
        const bool lds = span <= (u64)BJ_LDS;
        if (lds) {
            for (u64 j = threadIdx.x; j < span; j += BJ_T) {
                u64 of = offs[plo + j];
                u32 st = pstart[plo + j];
                if (j == 0) {   // the tile may start inside the first probe's rows
                    st += (u32)(base - of);
                    of = base;
                }
                l_offs[j] = (u32)(of - base);
                l_start[j] = st;
                l_pv[j] = tval(probe[plo + j]);
            }
        }
        __syncthreads();
        u64 t[BJ_I], wi[BJ_I];
        u32 mask[BJ_I];
#pragma unroll
        for (int r = 0; r < BJ_I; r++) {
            u64 o = base + (u64)r * BJ_T + threadIdx.x;
            mask[r] = 0;
            wi[r] = ~0ull;
            t[r] = 0;
            if (o < total) {
                u32 hv, pv;
                if (lds) {
                    u32 rel = (u32)(o - base), a = 0, b = (u32)span;
                    while (b - a > 1) {
                        u32 m = (a + b) >> 1;
                        if (l_offs[m] <= rel) a = m; else b = m;
                    }
                    hv = tval(rows[(u64)l_start[a] + (rel - l_offs[a])]);
                    pv = l_pv[a];
                } else {
                    u64 i = upper_bound_u64(offs, plo, phi, o) - 1;
                    hv = tval(rows[(u64)pstart[i] + (o - offs[i])]);
                    pv = tval(probe[i]);
                }
                t[r] = PROJ == DLG_PROJ_HV_PV ? tmake(hv, pv) : tmake(pv, hv);
                if (!NODIAG || hv != pv) {
                    u64 b = bm_bit(t[r], lo, W, TR);
                    wi[r] = b >> 5;
                    mask[r] = 1u << (b & 31);
                }
            }
        }
        u32 cur[BJ_I];
#pragma unroll
        for (int r = 0; r < BJ_I; r++) cur[r] = mask[r] ? words[wi[r]] : 0u;   // all reads in flight
#pragma unroll
        for (int r = 0; r < BJ_I; r++) {
            u32 m = (cur[r] & mask[r]) ? 0u : mask[r];   // bits already in F drop out
            bool app = false;
            if (__ballot(m != 0)) app = wave_insert(words, wi[r], m);
            u64 am = __ballot(app);
            if (am) {
                const int lane = threadIdx.x & (WAVE - 1);
                int leader = __ffsll((unsigned long long)am) - 1;
                u32 p = 0;
                if (lane == leader) p = atomicAdd(&s_cnt, (u32)__popcll(am));
                p = __shfl(p, leader, WAVE);
                if (app) buf[p + (u32)__popcll(am & lanemask_lt())] = t[r];
            }
        }
        __syncthreads();
        const u32 c = s_cnt;
        if (c > (u32)(BJ_BUF - BJ_TILE)) {
            if (threadIdx.x == 0) s_base = atomicAdd(counter, (unsigned long long)c);
            __syncthreads();
            for (u32 j = threadIdx.x; j < c; j += BJ_T) out[s_base + j] = buf[j];
            __syncthreads();
            if (threadIdx.x == 0) s_cnt = 0;
        }
    }
    __syncthreads();
    const u32 c = s_cnt;
    if (c) {
        if (threadIdx.x == 0) s_base = atomicAdd(counter, (unsigned long long)c);
        __syncthreads();
        for (u32 j = threadIdx.x; j < c; j += BJ_T) out[s_base + j] = buf[j];
    }
}

template <bool NODIAG, bool TR>
__global__ void bm_insert_list_k(const u64 *in, u64 n, u32 *words, u32 lo, u64 W, u64 *out,
                                 unsigned long long *counter) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    u64 lim = (n + WAVE - 1) / WAVE * WAVE;   // whole waves iterate together for the ballot
    for (; i < lim; i += stride) {
        bool app = false;
        u64 t = 0;
        if (i < n) {
            t = in[i];
            if (!NODIAG || tkey(t) != tval(t)) app = bm_set(words, bm_bit(t, lo, W, TR));
        }
        wave_append(app, t, out, counter);
    }
}

template <bool TR>
__global__ void bm_clear_list_k(const u64 *in, u64 n, u32 *words, u32 lo, u64 W) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < n; i += stride) words[bm_bit(in[i], lo, W, TR) >> 5] = 0u;   // every set bit is in the list
}

// Emission: per block, 256 threads x 8 words; popcounts -> scan -> write.
constexpr int BE_T = 256, BE_W = 8, BE_TILE = BE_T * BE_W;

__global__ __launch_bounds__(BE_T) void bm_count_k(const u32 *words, u64 nw, u64 *bsum) {
    __shared__ u64 sc[BE_T / WAVE + 1];
    u64 base = (u64)blockIdx.x * BE_TILE + (u64)threadIdx.x * BE_W;
    u64 c = 0;
#pragma unroll
    for (int i = 0; i < BE_W; i++)
        if (base + i < nw) c += __popc(words[base + i]);
    u64 tot;
    block_exclusive_scan<u64>(c, sc, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

template <bool TR>
__global__ __launch_bounds__(BE_T) void bm_emit_k(const u32 *words, u64 nw, const u64 *bpre, u32 lo, u64 W,
                                                  u64 *out) {
    __shared__ u64 sc[BE_T / WAVE + 1];
    u64 base = (u64)blockIdx.x * BE_TILE + (u64)threadIdx.x * BE_W;
    u32 w[BE_W];
    u64 c = 0;
#pragma unroll
    for (int i = 0; i < BE_W; i++) {
        w[i] = base + i < nw ? words[base + i] : 0u;
        c += __popc(w[i]);
    }
    u64 pos = block_exclusive_scan<u64>(c, sc, nullptr) + bpre[blockIdx.x];
#pragma unroll
    for (int i = 0; i < BE_W; i++) {
        u32 x = w[i];
        if (!x) continue;
        u64 b0 = (base + i) * 32;
        u64 k = b0 / W, v = b0 - k * W;   // bit b0 -> (row, column); walk the word's bits in order
        u32 prev = 0;
        while (x) {
            u32 j = __ffs(x) - 1;
            x &= x - 1;
            v += j - prev;
            prev = j;
            while (v >= W) {
                v -= W;
                k++;
            }
            out[pos++] = TR ? tmake(lo + (u32)v, lo + (u32)k) : tmake(lo + (u32)k, lo + (u32)v);
        }
    }
}

void bitmap_init(dlg_ctx *ctx, Bitmap &bm, u32 lo, u32 hi, bool transposed) {
    bm.lo = lo;
    bm.W = (u64)(hi - lo) + 1;
    bm.nbits = bm.W * bm.W;
    bm.nwords = (bm.nbits + 31) / 32;
    bm.tr = transposed;
    bm.words.ensure(bm.nwords);
    fill_u32(ctx, bm.words.p, bm.nwords, 0u);
    bm.count.ensure(1);
}

static u64 read_counter(dlg_ctx *ctx, const Bitmap &bm) {
    u64 h = 0;
    ctx->read_scalars(bm.count.p, 1, &h);
    return h;
}

template <int PROJ, bool NODIAG>
static void launch_join_insert(dlg_ctx *ctx, const Index &ix, const u64 *probe, u64 n, JoinScratch &js, u64 total,
                               Bitmap &bm, DBuf<u64> &out) {
    // per row: the matched row's value + a bitmap word; per new row an atomic and a write
    double bytes = 12.0 * total;
    u64 ntiles = (total + BJ_TILE - 1) / BJ_TILE;
    unsigned g = (unsigned)std::min<u64>(ntiles, (u64)ctx->num_cus * 4);
    unsigned long long *cnt = (unsigned long long *)bm.count.p;
    if (bm.tr)
        DLG_LAUNCH(ctx, DLG_K_BITMAP_INSERT, bytes, (bm_join_insert_k<PROJ, NODIAG, true>), g, BJ_T, 0, probe, n,
                   (const u32 *)js.pstart.p, (const u64 *)js.offs.p, total, (const u64 *)ix.rows.p, bm.words.p,
                   bm.lo, bm.W, out.p, cnt);
    else
        DLG_LAUNCH(ctx, DLG_K_BITMAP_INSERT, bytes, (bm_join_insert_k<PROJ, NODIAG, false>), g, BJ_T, 0, probe, n,
                   (const u32 *)js.pstart.p, (const u64 *)js.offs.p, total, (const u64 *)ix.rows.p, bm.words.p,
                   bm.lo, bm.W, out.p, cnt);
}

u64 bitmap_join_insert(dlg_ctx *ctx, const Index &ix, const u64 *probe, u64 n, JoinScratch &js, int proj,
                       bool nodiag, Bitmap &bm, DBuf<u64> &out, u64 *raw) {
    u64 total = join_count(ctx, ix, probe, n, js);
    if (raw) *raw += total;
    if (total == 0) return 0;
    out.ensure(std::min<u64>(total, bm.nbits));
    fill_u64(ctx, bm.count.p, 1, 0ull);
    if (proj == DLG_PROJ_PV_HV) {
        if (nodiag) launch_join_insert<DLG_PROJ_PV_HV, true>(ctx, ix, probe, n, js, total, bm, out);
        else launch_join_insert<DLG_PROJ_PV_HV, false>(ctx, ix, probe, n, js, total, bm, out);
    } else {
        if (nodiag) launch_join_insert<DLG_PROJ_HV_PV, true>(ctx, ix, probe, n, js, total, bm, out);
        else launch_join_insert<DLG_PROJ_HV_PV, false>(ctx, ix, probe, n, js, total, bm, out);
    }
    return read_counter(ctx, bm);
}

u64 bitmap_insert_list(dlg_ctx *ctx, const u64 *in, u64 n, bool nodiag, Bitmap &bm, u64 *out) {
    if (n == 0) return 0;
    fill_u64(ctx, bm.count.p, 1, 0ull);
    unsigned long long *cnt = (unsigned long long *)bm.count.p;
    unsigned g = sgrid(ctx, n);
    int bs = ctx->block_size;
    if (nodiag) {
        if (bm.tr)
            DLG_LAUNCH(ctx, DLG_K_BITMAP_INSERT, 12.0 * n, (bm_insert_list_k<true, true>), g, bs, 0, in, n,
                       bm.words.p, bm.lo, bm.W, out, cnt);
        else
            DLG_LAUNCH(ctx, DLG_K_BITMAP_INSERT, 12.0 * n, (bm_insert_list_k<true, false>), g, bs, 0, in, n,
                       bm.words.p, bm.lo, bm.W, out, cnt);
    } else {
        if (bm.tr)
            DLG_LAUNCH(ctx, DLG_K_BITMAP_INSERT, 12.0 * n, (bm_insert_list_k<false, true>), g, bs, 0, in, n,
                       bm.words.p, bm.lo, bm.W, out, cnt);
        else
            DLG_LAUNCH(ctx, DLG_K_BITMAP_INSERT, 12.0 * n, (bm_insert_list_k<false, false>), g, bs, 0, in, n,
                       bm.words.p, bm.lo, bm.W, out, cnt);
    }
    return read_counter(ctx, bm);
}

void bitmap_clear_list(dlg_ctx *ctx, const u64 *in, u64 n, Bitmap &bm) {
    if (n == 0) return;
    if (bm.tr)
        DLG_LAUNCH(ctx, DLG_K_MISC, 12.0 * n, bm_clear_list_k<true>, sgrid(ctx, n), ctx->block_size, 0, in, n,
                   bm.words.p, bm.lo, bm.W);
    else
        DLG_LAUNCH(ctx, DLG_K_MISC, 12.0 * n, bm_clear_list_k<false>, sgrid(ctx, n), ctx->block_size, 0, in, n,
                   bm.words.p, bm.lo, bm.W);
}

u64 bitmap_emit(dlg_ctx *ctx, const Bitmap &bm, DBuf<u64> &out, u64 expect, int bits) {
    out.ensure(expect ? expect : 1);
    if (expect == 0) return 0;
    u64 nb = (bm.nwords + BE_TILE - 1) / BE_TILE;
    DBuf<u64> bsum(ctx, nb), tot(ctx, 1);
    DLG_LAUNCH(ctx, DLG_K_MISC, 4.0 * bm.nwords, bm_count_k, (unsigned)nb, BE_T, 0, (const u32 *)bm.words.p,
               bm.nwords, bsum.p);
    scan_blocks(ctx, bsum.p, nb, tot.p);
    if (bm.tr)
        DLG_LAUNCH(ctx, DLG_K_MISC, 4.0 * bm.nwords + 8.0 * expect, bm_emit_k<true>, (unsigned)nb, BE_T, 0,
                   (const u32 *)bm.words.p, bm.nwords, (const u64 *)bsum.p, bm.lo, bm.W, out.p);
    else
        DLG_LAUNCH(ctx, DLG_K_MISC, 4.0 * bm.nwords + 8.0 * expect, bm_emit_k<false>, (unsigned)nb, BE_T, 0,
                   (const u32 *)bm.words.p, bm.nwords, (const u64 *)bsum.p, bm.lo, bm.W, out.p);
    u64 h = 0;
    ctx->read_scalars(tot.p, 1, &h);
    if (h != expect) fail(DLG_DEVICE_ERROR, "bitmap emission count mismatch");
    if (bm.tr) {   // emitted in (value, key) order
        u32 hi = bm.lo + (u32)(bm.W - 1);
        u32 bounds[4] = {bm.lo, hi, bm.lo, hi};
        DBuf<u64> tmp(ctx);
        radix_sort(ctx, out, h, tmp, bits, bounds);
    }
    return h;
}

}  // namespace dlg
