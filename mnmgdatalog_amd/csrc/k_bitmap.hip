// k_bitmap.hip -- the full relation as a bitmap over the id range.
//
// Every tuple of a TC / SG fixpoint is a pair of node ids of the input, so
// with W = the width of the id range and rows padded to whole 32-bit words
// (S = 32 * ceil(W / 32) bits per row), tuple (k, v) has the bit
//     (k - lo) * S + (v - lo)            (row-major), or
//     (v - lo) * S + (k - lo)            (transposed)
// of a W x S-bit map (W = 46,656 ids -> 272 MB for the SG mesh; 160,000 ids
// -> 3.2 GB for the 400 x 400 grid).  The relational steps of a round
// collapse into the join's materialize pass:
//   for every join row t:  if bit(t) is clear, set it; the row that set it
//   is a new tuple
// which is unique(J) \ F and F = F u N at once (tc.cu:250-283): J is never
// written, sorted or merged.
//
// A probe's join rows share the probe's value and differ in the matched
// row's value; the layout puts the probe's half in the bitmap row, so one
// probe's rows land in one bitmap row.  The bits of a tile of join rows are
// ORed together per word in LDS first, so each distinct word costs one plain
// read and at most one atomicOr.  New tuples are appended unordered, or
// marked in a second map that is drained in bit order.  The final relation
// is read out of the map in bit order: a row-major map is sorted as it lies,
// a transposed one is transposed back (32 x 32-bit blocks) first.
#include "dlg_internal.h"

#include <algorithm>

namespace dlg {

__device__ __forceinline__ u64 bm_bit(u64 t, u32 lo, u64 S, bool tr) {
    u64 a = (u64)(tkey(t) - lo), b = (u64)(tval(t) - lo);
    return tr ? b * S + a : a * S + b;
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

constexpr int BJ_T = 256, BJ_I = 8, BJ_TILE = BJ_T * BJ_I, BJ_LDS = BJ_TILE + 1;
constexpr int BH = 2 * BJ_TILE;         // LDS hash slots per tile (load <= 1/2)
constexpr u32 BH_EMPTY = 0xffffffffu;   // word indexes stay below it (nwords < 2^32 - 1)

// What the insert does with the rows that set a clear bit.
enum BmMode {
    BM_APPEND = 0,   // append them to `out` (unordered)
    BM_MARK = 1,     // set the same bits in `mark` (a map of the same geometry)
    BM_SET = 2,      // nothing: only the bits are wanted
};

// LDS layout of the join-insert kernel: the probe staging of a tile
// (offsets, starts, probe values, row owners) and, once the tile's rows are
// located, the tile's word hash table (keys, bits) in the same space.
constexpr int LB_OFFS = 0, LB_START = 2052, LB_PV = 4104, LB_OWNER = 6156, LB_WORDS = LB_OWNER + BJ_TILE;
constexpr int LB_HKEY = 0, LB_HBITS = BH;
static_assert(LB_START >= BJ_LDS && LB_PV - LB_START >= BJ_LDS && LB_OWNER - LB_PV >= BJ_LDS, "staging overlap");
static_assert(LB_OWNER % 4 == 0 && LB_HBITS + BH <= LB_WORDS, "hash table must fit the staging space");

__device__ __forceinline__ u32 bh_hash(u32 w) { return (w * 0x9e3779b1u) >> (32 - 12); }   // BH = 2^12
static_assert(BH == 1 << 12, "bh_hash assumes 4096 slots");

// Join materialize fused with the bitmap insert.  Persistent blocks walk
// tiles of BJ_TILE join rows.  Per tile:
//   1 locate each row's probe: the tile's probes staged in LDS with a
//     row -> probe owner map (a binary search in global memory when the
//     tile spans more probes than fit);
//   2 gather the matched values, form the tuples and their bitmap words;
//   3 OR the tile's bits together per word in an LDS hash table, so a word
//     that many rows of the tile hit costs one global read and at most one
//     atomicOr (device-scope atomics execute at the memory side);
//   4 per distinct word: a plain read; the bits not yet set go in with one
//     atomicOr (APPEND reads its result: the bits this tile set);
//   5 MARK: the same bits are ORed into `mark`; APPEND: each bit this tile
//     set is claimed by the first of the tile's rows that carries it, which
//     appends its tuple through an LDS buffer: one counter atomic per tile.
template <int PROJ, bool NODIAG, bool TR, int MODE>
__global__ __launch_bounds__(BJ_T) void bm_join_insert_k(const u64 *probe, u64 np, const u32 *pstart,
                                                         const u64 *offs, u64 total, const u64 *rows,
                                                         u32 *words, u32 lo, u64 S, u64 *out,
                                                         unsigned long long *counter, u32 *mark) {
    __shared__ u64 s_lo, s_hi;
    __shared__ unsigned long long s_base;
    __shared__ u32 s_cnt, s_nused;
    __shared__ __align__(16) u32 lds[LB_WORDS];
    __shared__ u16 h_used[BJ_TILE];   // the tile's occupied slots
    __shared__ u32 s_wmax[BJ_T / WAVE];
    __shared__ u64 buf[MODE == BM_APPEND ? BJ_TILE : 1];   // the tile's new tuples
    u32 *l_offs = lds + LB_OFFS, *l_start = lds + LB_START, *l_pv = lds + LB_PV, *owner = lds + LB_OWNER;
    u32 *h_key = lds + LB_HKEY, *h_bits = lds + LB_HBITS;
    const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
    const u64 ntiles = (total + BJ_TILE - 1) / BJ_TILE;
    if (threadIdx.x == 0) s_cnt = 0;
    for (u64 tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const u64 base = tile * BJ_TILE;
        __syncthreads();
        if (threadIdx.x < WAVE) {   // the first wave finds the tile's probes
            const u64 last = (base + BJ_TILE < total ? base + BJ_TILE : total) - 1;
            const u64 l = wave_upper_bound(offs, 0, np, base) - 1;
            const u64 h = wave_upper_bound(offs, l, np, last);
            if (threadIdx.x == 0) {
                s_lo = l;
                s_hi = h;
                s_nused = 0;
            }
        }
        __syncthreads();
        const u64 plo = s_lo, phi = s_hi, span = phi - plo;
        const bool staged = span <= (u64)BJ_LDS;
        if (staged) {
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
            __syncthreads();
            // owner[rel] = the last probe starting at or before rel: mark each
            // probe's first row, then a block-wide running max fills forward.
            // Probes without rows share their start with the next probe and
            // are not marked.
            for (int i = threadIdx.x; i < BJ_TILE; i += BJ_T) owner[i] = 0;
            __syncthreads();
            for (u64 j = threadIdx.x; j < span; j += BJ_T)
                if (j + 1 == span || l_offs[j + 1] != l_offs[j]) owner[l_offs[j]] = (u32)j;
            __syncthreads();
            uint4 o0 = *(const uint4 *)&owner[threadIdx.x * BJ_I];
            uint4 o1 = *(const uint4 *)&owner[threadIdx.x * BJ_I + 4];
            u32 loc[BJ_I] = {o0.x, o0.y, o0.z, o0.w, o1.x, o1.y, o1.z, o1.w};
#pragma unroll
            for (int i = 1; i < BJ_I; i++) loc[i] = max(loc[i], loc[i - 1]);
            u32 inc = loc[BJ_I - 1];
#pragma unroll
            for (int d = 1; d < WAVE; d <<= 1) {
                u32 x = __shfl_up(inc, d, WAVE);
                if (lane >= d) inc = max(inc, x);
            }
            if (lane == WAVE - 1) s_wmax[wv] = inc;
            u32 pre = __shfl_up(inc, 1, WAVE);
            if (lane == 0) pre = 0;
            __syncthreads();
            for (int w = 0; w < wv; w++) pre = max(pre, s_wmax[w]);
#pragma unroll
            for (int i = 0; i < BJ_I; i++) loc[i] = max(loc[i], pre);
            *(uint4 *)&owner[threadIdx.x * BJ_I] = make_uint4(loc[0], loc[1], loc[2], loc[3]);
            *(uint4 *)&owner[threadIdx.x * BJ_I + 4] = make_uint4(loc[4], loc[5], loc[6], loc[7]);
            __syncthreads();
        }
        // 1-2: locate, gather, form tuples and words
        u32 ridx[BJ_I], pvs[BJ_I];
#pragma unroll
        for (int r = 0; r < BJ_I; r++) {
            u64 o = base + (u64)r * BJ_T + threadIdx.x;
            ridx[r] = 0;
            pvs[r] = 0;
            if (o < total) {
                if (staged) {
                    u32 rel = (u32)(o - base), a = owner[rel];
                    ridx[r] = l_start[a] + (rel - l_offs[a]);
                    pvs[r] = l_pv[a];
                } else {
                    u64 i = upper_bound_u64(offs, plo, phi, o) - 1;
                    ridx[r] = pstart[i] + (u32)(o - offs[i]);
                    pvs[r] = tval(probe[i]);
                }
            }
        }
        u32 hvs[BJ_I];
#pragma unroll
        for (int r = 0; r < BJ_I; r++) {
            u64 o = base + (u64)r * BJ_T + threadIdx.x;
            hvs[r] = o < total ? tval(rows[ridx[r]]) : 0u;
        }
        u64 t[BJ_I];
        u32 wi[BJ_I], mask[BJ_I];
#pragma unroll
        for (int r = 0; r < BJ_I; r++) {
            u64 o = base + (u64)r * BJ_T + threadIdx.x;
            u32 hv = hvs[r], pv = pvs[r];
            t[r] = PROJ == DLG_PROJ_HV_PV ? tmake(hv, pv) : tmake(pv, hv);
            mask[r] = 0;
            wi[r] = 0;
            if (o < total && (!NODIAG || hv != pv)) {
                u64 b = bm_bit(t[r], lo, S, TR);
                wi[r] = (u32)(b >> 5);
                mask[r] = 1u << (b & 31);
            }
        }
        __syncthreads();   // the staging space becomes the hash table
        for (int i = threadIdx.x * 4; i < BH; i += BJ_T * 4) {
            *(uint4 *)&h_key[i] = make_uint4(BH_EMPTY, BH_EMPTY, BH_EMPTY, BH_EMPTY);
            *(uint4 *)&h_bits[i] = make_uint4(0u, 0u, 0u, 0u);
        }
        __syncthreads();
        // 3: per-word OR in LDS
        u32 slot[BJ_I];
#pragma unroll
        for (int r = 0; r < BJ_I; r++) {
            slot[r] = 0;
            if (!mask[r]) continue;
            u32 h = bh_hash(wi[r]);
            while (true) {
                u32 k = atomicCAS(&h_key[h], BH_EMPTY, wi[r]);
                if (k == BH_EMPTY) {
                    h_used[atomicAdd(&s_nused, 1u)] = (u16)h;
                    break;
                }
                if (k == wi[r]) break;
                h = (h + 1) & (BH - 1);
            }
            atomicOr(&h_bits[h], mask[r]);
            slot[r] = h;
        }
        __syncthreads();
        // 4: one read and at most one atomic per distinct word
        const u32 nused = s_nused;
        u32 sh[BJ_I], sk[BJ_I], sb[BJ_I];
#pragma unroll
        for (int i = 0; i < BJ_I; i++) {
            u32 j = threadIdx.x + i * BJ_T;
            sb[i] = 0;
            sk[i] = 0;
            sh[i] = 0;
            if (j < nused) {
                sh[i] = h_used[j];
                sk[i] = h_key[sh[i]];
                sb[i] = h_bits[sh[i]];
            }
        }
        u32 cur[BJ_I];
#pragma unroll
        for (int i = 0; i < BJ_I; i++) cur[i] = sb[i] ? words[sk[i]] : 0u;
        // MARK needs no atomic results: the plain read sees at least the
        // words as the launch began (every earlier round's bits), so a bit it
        // finds clear is new this round, and marking it twice is harmless;
        // the first tile to set a bit read it clear, so every new bit is
        // marked.  APPEND needs the result: each bit is appended once.
        u32 nb[BJ_I];
#pragma unroll
        for (int i = 0; i < BJ_I; i++) {
            nb[i] = sb[i] & ~cur[i];
            if (nb[i]) {
                if (MODE != BM_APPEND) atomicOr(words + sk[i], nb[i]);
                else cur[i] = atomicOr(words + sk[i], nb[i]);
            }
        }
        if (MODE == BM_SET) continue;
        if (MODE == BM_MARK) {
#pragma unroll
            for (int i = 0; i < BJ_I; i++)
                if (nb[i]) atomicOr(mark + sk[i], nb[i]);
            continue;
        }
#pragma unroll
        for (int i = 0; i < BJ_I; i++) {
            if (nb[i]) nb[i] &= ~cur[i];   // the bits this tile set
            if (sb[i]) h_bits[sh[i]] = nb[i];
        }
        __syncthreads();
        // 5: each set bit goes to the first row of the tile that claims it
#pragma unroll
        for (int r = 0; r < BJ_I; r++) {
            bool app = false;
            if (mask[r] && (h_bits[slot[r]] & mask[r])) app = (atomicAnd(&h_bits[slot[r]], ~mask[r]) & mask[r]) != 0;
            u64 am = __ballot(app);
            if (am) {
                int leader = __ffsll((unsigned long long)am) - 1;
                u32 p = 0;
                if (lane == leader) p = atomicAdd(&s_cnt, (u32)__popcll(am));
                p = __shfl(p, leader, WAVE);
                if (app) buf[p + (u32)__popcll(am & lanemask_lt())] = t[r];
            }
        }
        __syncthreads();
        const u32 c = s_cnt;
        if (c) {
            if (threadIdx.x == 0) s_base = atomicAdd(counter, (unsigned long long)c);
            __syncthreads();
            for (u32 j = threadIdx.x; j < c; j += BJ_T) out[s_base + j] = buf[j];
            __syncthreads();
            if (threadIdx.x == 0) s_cnt = 0;
        }
    }
}

template <bool NODIAG, bool TR>
__global__ void bm_insert_list_k(const u64 *in, u64 n, u32 *words, u32 lo, u64 S, u64 *out,
                                 unsigned long long *counter) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    u64 lim = (n + WAVE - 1) / WAVE * WAVE;   // whole waves iterate together for the ballot
    for (; i < lim; i += stride) {
        bool app = false;
        u64 t = 0;
        if (i < n) {
            t = in[i];
            if (!NODIAG || tkey(t) != tval(t)) app = bm_set(words, bm_bit(t, lo, S, TR));
        }
        wave_append(app, t, out, counter);
    }
}

template <bool TR>
__global__ void bm_clear_list_k(const u64 *in, u64 n, u32 *words, u32 lo, u64 S) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < n; i += stride) words[bm_bit(in[i], lo, S, TR) >> 5] = 0u;   // every set bit is in the list
}

// Emission: per block, 256 threads x 8 words; popcounts -> scan -> write.
// The word buffer is padded to whole blocks (zeros), so the passes read
// without bounds checks, two 16-byte loads per thread.
constexpr int BE_T = 256, BE_W = 8, BE_TILE = BE_T * BE_W;

__device__ __forceinline__ void load8(const u32 *words, u64 base, u32 *w) {
    uint4 a = *(const uint4 *)(words + base), b = *(const uint4 *)(words + base + 4);
    w[0] = a.x; w[1] = a.y; w[2] = a.z; w[3] = a.w;
    w[4] = b.x; w[5] = b.y; w[6] = b.z; w[7] = b.w;
}

__global__ __launch_bounds__(BE_T) void bm_count_k(const u32 *words, u64 *bsum) {
    __shared__ u64 sc[BE_T / WAVE + 1];
    u64 base = (u64)blockIdx.x * BE_TILE + (u64)threadIdx.x * BE_W;
    u32 w[BE_W];
    load8(words, base, w);
    u64 c = 0;
#pragma unroll
    for (int i = 0; i < BE_W; i++) c += __popc(w[i]);
    u64 tot;
    block_exclusive_scan<u64>(c, sc, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// SWAP: write (column, row) instead of (row, column).  CLEAR: zero the words read.
// Emission of a tile: per-thread popcounts -> block scan.  A tile of at most
// 2 BE_STAGE set bits is expanded thread by thread (each thread's words give
// a consecutive run of outputs) into an LDS stage of BE_STAGE tuples at a
// time, which the block then writes out with whole-wave coalesced stores.  A
// denser tile gets a position for every word (in LDS), then each wave walks
// its 512 words two at a time, one lane per bit, so the set bits of a word
// pair are written to consecutive addresses.
// MULTI: the words are several maps, each starting at a whole tile: map m
// covers tiles [mt[m], mt[m + 1]) with ids from mlo[m] and mwpr[m] words per
// row (nm maps); each workgroup finds its tile's map by a binary search.
constexpr int BE_STAGE = 4096;
template <bool SWAP, bool CLEAR, bool MULTI = false>
__global__ __launch_bounds__(BE_T) void bm_emit_k(u32 *words, const u64 *bpre, u32 lo, u64 wpr, u64 *out,
                                                  const u64 *mt = nullptr, const u32 *mlo = nullptr,
                                                  const u64 *mwpr = nullptr, u32 nm = 0) {
    __shared__ u64 sc[BE_T / WAVE + 1];
    __shared__ u64 s_buf[BE_STAGE];   // the stage, or the dense walk's three word tables
    static_assert(3 * BE_TILE * 4 <= BE_STAGE * 8, "word tables must fit the stage");
    u32 *s_w = (u32 *)s_buf, *s_pos = s_w + BE_TILE, *s_row = s_pos + BE_TILE;
    const u64 tile = (u64)blockIdx.x * BE_TILE;
    const u64 base = tile + (u64)threadIdx.x * BE_W;
    u64 mbase = 0;   // the tile's map's first word
    if (MULTI) {
        u32 a = 0, b = nm;   // the last map starting at or before this tile
        while (b - a > 1) {
            const u32 m = (a + b) >> 1;
            if (mt[m] <= (u64)blockIdx.x) a = m; else b = m;
        }
        mbase = mt[a] * BE_TILE;
        lo = mlo[a];
        wpr = mwpr[a];
    }
    const u64 rtile = tile - mbase, rbase = base - mbase;   // word indexes within the map
    u32 w[BE_W];
    load8(words, base, w);
    u64 c = 0;
#pragma unroll
    for (int i = 0; i < BE_W; i++) c += __popc(w[i]);
    u64 tot;
    const u64 rel = block_exclusive_scan<u64>(c, sc, &tot);   // this thread's first output in the tile
    const u64 out0 = bpre[blockIdx.x];
    if (CLEAR && c) {
        *(uint4 *)(words + base) = make_uint4(0u, 0u, 0u, 0u);
        *(uint4 *)(words + base + 4) = make_uint4(0u, 0u, 0u, 0u);
    }
    if (tot == 0) return;
    if (tot <= (u64)(2 * BE_STAGE)) {   // uniform: staged, one or two rounds
        const u64 k0 = rbase / wpr, cw0 = rbase - k0 * wpr;
        for (u64 r0 = 0; r0 < tot; r0 += BE_STAGE) {
            const u64 r1 = r0 + BE_STAGE < tot ? r0 + BE_STAGE : tot;
            if (c && rel < r1 && rel + c > r0) {
                u64 p = rel, k = k0, cw = cw0;
#pragma unroll
                for (int i = 0; i < BE_W; i++) {
                    u32 x = w[i];
                    const u32 c0 = (u32)(cw * 32);
                    while (x) {
                        const u32 v = c0 + (u32)(__ffs(x) - 1);
                        x &= x - 1;
                        if (p >= r0 && p < r1)
                            s_buf[p - r0] = SWAP ? tmake(lo + v, lo + (u32)k) : tmake(lo + (u32)k, lo + v);
                        p++;
                    }
                    if (++cw == wpr) {
                        cw = 0;
                        k++;
                    }
                }
            }
            __syncthreads();
            for (u64 i = threadIdx.x; i < r1 - r0; i += BE_T) out[out0 + r0 + i] = s_buf[i];
            __syncthreads();
        }
        return;
    }
    {
        u64 pos = rel;
        u64 k = rbase / wpr, cw = rbase - k * wpr;   // one division per thread; the words are consecutive
#pragma unroll
        for (int i = 0; i < BE_W; i++) {
            s_w[threadIdx.x * BE_W + i] = w[i];
            s_pos[threadIdx.x * BE_W + i] = (u32)pos;
            s_row[threadIdx.x * BE_W + i] = (u32)k;
            pos += __popc(w[i]);
            if (++cw == wpr) {
                cw = 0;
                k++;
            }
        }
    }
    __syncthreads();
    out += out0;
    const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
    const u32 half = lane >> 5, bit = lane & 31;
    const u32 below = (1u << bit) - 1u;
    for (int j = wv * (BE_TILE / (BE_T / WAVE)); j < (wv + 1) * (BE_TILE / (BE_T / WAVE)); j += 2) {
        const u32 x0 = s_w[j], x1 = s_w[j + 1];
        if ((x0 | x1) == 0u) continue;   // uniform across the wave
        const u32 x = half ? x1 : x0;
        if (!((x >> bit) & 1u)) continue;
        const u32 k = s_row[j + half];                                       // row
        const u32 v = (u32)(rtile + j + half - (u64)k * wpr) * 32u + bit;     // column
        out[s_pos[j + half] + __popc(x & below)] = SWAP ? tmake(lo + v, lo + k) : tmake(lo + k, lo + v);
    }
}

void bitmap_init(dlg_ctx *ctx, Bitmap &bm, u32 lo, u32 hi, bool transposed, bool clear) {
    bm.lo = lo;
    bm.W = (u64)(hi - lo) + 1;
    bm.wpr = (bm.W + 31) / 32;
    bm.S = bm.wpr * 32;
    bm.nwords = bm.W * bm.wpr;
    bm.nbits = bm.nwords * 32;
    if (bm.nwords >= (u64)BH_EMPTY) fail(DLG_INVALID_INPUT, "bitmap layout: id range too wide");
    bm.tr = transposed;
    u64 padded = (bm.nwords + BE_TILE - 1) / BE_TILE * BE_TILE;
    bm.words.ensure(padded);
    if (clear) fill_u32(ctx, bm.words.p, padded, 0u);
    else fill_u32(ctx, bm.words.p + bm.nwords, padded - bm.nwords, 0u);   // the caller writes every row
    bm.count.ensure(1);
}

static u64 read_counter(dlg_ctx *ctx, const Bitmap &bm) {
    u64 h = 0;
    ctx->read_scalars(bm.count.p, 1, &h);
    return h;
}

template <int PROJ, bool NODIAG, int MODE>
static void launch_join_insert(dlg_ctx *ctx, const Index &ix, const u64 *probe, u64 n, JoinScratch &js, u64 total,
                               Bitmap &bm, u64 *out, u32 *mark) {
    // per row: the matched row's value + a bitmap word; per new row an atomic (+ a write)
    double bytes = 12.0 * total;
    u64 ntiles = (total + BJ_TILE - 1) / BJ_TILE;
    unsigned g = (unsigned)std::min<u64>(ntiles, (u64)ctx->num_cus * 4);
    unsigned long long *cnt = (unsigned long long *)bm.count.p;
    if (bm.tr)
        DLG_LAUNCH(ctx, DLG_K_BITMAP_INSERT, bytes, (bm_join_insert_k<PROJ, NODIAG, true, MODE>), g, BJ_T, 0, probe,
                   n, (const u32 *)js.pstart.p, (const u64 *)js.offs.p, total, (const u64 *)ix.rows.p, bm.words.p,
                   bm.lo, bm.S, out, cnt, mark);
    else
        DLG_LAUNCH(ctx, DLG_K_BITMAP_INSERT, bytes, (bm_join_insert_k<PROJ, NODIAG, false, MODE>), g, BJ_T, 0,
                   probe, n, (const u32 *)js.pstart.p, (const u64 *)js.offs.p, total, (const u64 *)ix.rows.p,
                   bm.words.p, bm.lo, bm.S, out, cnt, mark);
}

template <int MODE>
static void join_insert_mode(dlg_ctx *ctx, const Index &ix, const u64 *probe, u64 n, JoinScratch &js, u64 total,
                             int proj, bool nodiag, Bitmap &bm, u64 *out, u32 *mark) {
    if (proj == DLG_PROJ_PV_HV) {
        if (nodiag) launch_join_insert<DLG_PROJ_PV_HV, true, MODE>(ctx, ix, probe, n, js, total, bm, out, mark);
        else launch_join_insert<DLG_PROJ_PV_HV, false, MODE>(ctx, ix, probe, n, js, total, bm, out, mark);
    } else {
        if (nodiag) launch_join_insert<DLG_PROJ_HV_PV, true, MODE>(ctx, ix, probe, n, js, total, bm, out, mark);
        else launch_join_insert<DLG_PROJ_HV_PV, false, MODE>(ctx, ix, probe, n, js, total, bm, out, mark);
    }
}

u64 bitmap_join_insert(dlg_ctx *ctx, const Index &ix, const u64 *probe, u64 n, JoinScratch &js, int proj,
                       bool nodiag, Bitmap &bm, DBuf<u64> &out, u64 *raw) {
    u64 total = join_count(ctx, ix, probe, n, js);
    if (raw) *raw += total;
    if (total == 0) return 0;
    out.ensure(std::min<u64>(total, bm.nbits));
    fill_u64(ctx, bm.count.p, 1, 0ull);
    join_insert_mode<BM_APPEND>(ctx, ix, probe, n, js, total, proj, nodiag, bm, out.p, nullptr);
    return read_counter(ctx, bm);
}

void bitmap_join_mark(dlg_ctx *ctx, const Index &ix, const u64 *probe, u64 n, JoinScratch &js, int proj,
                      bool nodiag, Bitmap &bm, Bitmap *mark, u64 *raw) {
    u64 total = join_count(ctx, ix, probe, n, js);
    if (raw) *raw += total;
    if (total == 0) return;
    if (mark) {
        if (mark->W != bm.W || mark->lo != bm.lo || mark->tr != bm.tr)
            fail(DLG_INVALID_INPUT, "bitmap_join_mark: maps of different geometry");
        join_insert_mode<BM_MARK>(ctx, ix, probe, n, js, total, proj, nodiag, bm, nullptr, mark->words.p);
    } else {
        join_insert_mode<BM_SET>(ctx, ix, probe, n, js, total, proj, nodiag, bm, nullptr, nullptr);
    }
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
                       bm.words.p, bm.lo, bm.S, out, cnt);
        else
            DLG_LAUNCH(ctx, DLG_K_BITMAP_INSERT, 12.0 * n, (bm_insert_list_k<true, false>), g, bs, 0, in, n,
                       bm.words.p, bm.lo, bm.S, out, cnt);
    } else {
        if (bm.tr)
            DLG_LAUNCH(ctx, DLG_K_BITMAP_INSERT, 12.0 * n, (bm_insert_list_k<false, true>), g, bs, 0, in, n,
                       bm.words.p, bm.lo, bm.S, out, cnt);
        else
            DLG_LAUNCH(ctx, DLG_K_BITMAP_INSERT, 12.0 * n, (bm_insert_list_k<false, false>), g, bs, 0, in, n,
                       bm.words.p, bm.lo, bm.S, out, cnt);
    }
    return read_counter(ctx, bm);
}

void bitmap_clear_list(dlg_ctx *ctx, const u64 *in, u64 n, Bitmap &bm) {
    if (n == 0) return;
    if (bm.tr)
        DLG_LAUNCH(ctx, DLG_K_MISC, 12.0 * n, bm_clear_list_k<true>, sgrid(ctx, n), ctx->block_size, 0, in, n,
                   bm.words.p, bm.lo, bm.S);
    else
        DLG_LAUNCH(ctx, DLG_K_MISC, 12.0 * n, bm_clear_list_k<false>, sgrid(ctx, n), ctx->block_size, 0, in, n,
                   bm.words.p, bm.lo, bm.S);
}

// Count pass + block scan over nwords words (padded to whole BE_TILEs);
// returns the number of set bits (synchronises).
static u64 words_count(dlg_ctx *ctx, const u32 *words, u64 nwords, DBuf<u64> &bsum, u64 &nb) {
    nb = (nwords + BE_TILE - 1) / BE_TILE;
    bsum.ensure(nb);
    DBuf<u64> tot(ctx, 1);
    DLG_LAUNCH(ctx, DLG_K_MISC, 4.0 * nb * BE_TILE, bm_count_k, (unsigned)nb, BE_T, 0, words, bsum.p);
    scan_blocks(ctx, bsum.p, nb, tot.p);
    u64 h = 0;
    ctx->read_scalars(tot.p, 1, &h);
    return h;
}

template <bool SWAP, bool CLEAR>
static void emit_pass(dlg_ctx *ctx, u32 lo, u64 wpr, u32 *words, const DBuf<u64> &bsum, u64 nb, u64 h,
                      u64 *out) {
    DLG_LAUNCH(ctx, DLG_K_MISC, (CLEAR ? 8.0 : 4.0) * nb * BE_TILE + 8.0 * h, (bm_emit_k<SWAP, CLEAR>),
               (unsigned)nb, BE_T, 0, words, (const u64 *)bsum.p, lo, wpr, out);
}

u64 bitmap_drain(dlg_ctx *ctx, Bitmap &bm, DBuf<u64> &out, bool swap, bool clear) {
    DBuf<u64> bsum(ctx);
    u64 nb = 0;
    u64 h = words_count(ctx, bm.words.p, bm.nwords, bsum, nb);
    out.ensure(h ? h : 1);
    if (h == 0) return 0;
    if (clear) {
        if (swap) emit_pass<true, true>(ctx, bm.lo, bm.wpr, bm.words.p, bsum, nb, h, out.p);
        else emit_pass<false, true>(ctx, bm.lo, bm.wpr, bm.words.p, bsum, nb, h, out.p);
    } else {
        if (swap) emit_pass<true, false>(ctx, bm.lo, bm.wpr, bm.words.p, bsum, nb, h, out.p);
        else emit_pass<false, false>(ctx, bm.lo, bm.wpr, bm.words.p, bsum, nb, h, out.p);
    }
    return h;
}

u64 bitmap_count(dlg_ctx *ctx, const Bitmap &bm) {
    DBuf<u64> bsum(ctx);
    u64 nb = 0;
    return words_count(ctx, bm.words.p, bm.nwords, bsum, nb);
}

// Bit-matrix transpose of a W x S map in 32 x 32-bit blocks.  A workgroup
// takes a tile of BT x BT blocks (32 BT source rows x BT words): the tile
// comes in with coalesced loads (BT consecutive words per row) into LDS,
// each half wave transposes blocks there -- 32 lanes hold a block's 32 rows,
// five exchange steps swap the off-diagonal halves of 16-, 8-, 4-, 2- and
// 1-bit sub-blocks, and lane i holds the block's column i as a destination
// word -- into a second LDS tile, which goes out with coalesced stores (BT
// consecutive words per destination row).  Rows are padded to BT + 1 words,
// so a block's 32 rows sit in 32 different banks.
// Small maps (a few tiles) go block by block instead: a half wave per block,
// loading the block's 32 rows and storing its 32 columns directly.
__global__ void bm_transpose_small_k(const u32 *src, u32 *dst, u64 W, u64 wpr) {
    const int lane = threadIdx.x & 31;
    const u64 blk = ((u64)blockIdx.x * blockDim.x + threadIdx.x) >> 5;   // one block per half wave
    if (blk >= wpr * wpr) return;   // whole half waves leave together
    const u64 R = blk / wpr, C = blk - R * wpr;   // source block: rows 32R.., word column C
    const u64 r = R * 32 + lane;
    u32 x = r < W ? src[r * wpr + C] : 0u;
    const u32 m[5] = {0x0000ffffu, 0x00ff00ffu, 0x0f0f0f0fu, 0x33333333u, 0x55555555u};
#pragma unroll
    for (int st = 0; st < 5; st++) {
        const int j = 16 >> st;
        const u32 y = __shfl_xor(x, j, 32);
        x = (lane & j) ? ((x & ~m[st]) | ((y & ~m[st]) >> j)) : ((x & m[st]) | ((y & m[st]) << j));
    }
    const u64 c = C * 32 + lane;   // destination row = source column
    if (c < W) dst[c * wpr + R] = x;
}

constexpr int BT = 16;
constexpr u64 BT_SMALL_WPR = 256;   // maps up to this many words per row go block by block
__global__ __launch_bounds__(256) void bm_transpose_k(const u32 *src, u32 *dst, u64 W, u64 wpr) {
    __shared__ u32 tb[32 * BT][BT + 1];   // the tile, transposed in place
    const u64 nt = (wpr + BT - 1) / BT;
    const u64 R0 = (u64)(blockIdx.x / nt) * BT, C0 = (u64)(blockIdx.x % nt) * BT;   // tile: block row, word column
    const int t = threadIdx.x;
    constexpr int NK = 32 * BT * BT / 256;   // words per thread, all loaded before any is stored
    u32 v[NK];
#pragma unroll
    for (int k = 0; k < NK; k++) {
        const int rr = t / BT + (256 / BT) * k, cc = t % BT;
        const u64 r = R0 * 32 + rr, c = C0 + cc;
        v[k] = r < W && c < wpr ? src[r * wpr + c] : 0u;
    }
#pragma unroll
    for (int k = 0; k < NK; k++) tb[t / BT + (256 / BT) * k][t % BT] = v[k];
    __syncthreads();
    // each half wave transposes BT * BT / 8 blocks into registers, then (after
    // every half wave has read its blocks) writes them back transposed
    const int lane = t & 31;
    const u32 m[5] = {0x0000ffffu, 0x00ff00ffu, 0x0f0f0f0fu, 0x33333333u, 0x55555555u};
    constexpr int NB = BT * BT / (256 / 32);
#pragma unroll
    for (int i = 0; i < NB; i++) {
        const int bk = (t >> 5) + (256 / 32) * i, Rl = bk / BT, Cl = bk % BT;
        u32 x = tb[32 * Rl + lane][Cl];
#pragma unroll
        for (int st = 0; st < 5; st++) {
            const int j = 16 >> st;
            const u32 y = __shfl_xor(x, j, 32);
            x = (lane & j) ? ((x & ~m[st]) | ((y & ~m[st]) >> j)) : ((x & m[st]) | ((y & m[st]) << j));
        }
        v[i] = x;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NB; i++) {
        const int bk = (t >> 5) + (256 / 32) * i, Rl = bk / BT, Cl = bk % BT;
        tb[32 * Cl + lane][Rl] = v[i];   // destination row = source column
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NK; k++) {
        const int rr = t / BT + (256 / BT) * k, cc = t % BT;
        const u64 r = C0 * 32 + rr, c = R0 + cc;
        if (r < W && c < wpr) dst[r * wpr + c] = tb[rr][cc];
    }
}

// A W-row block of wpr words per row, transposed (or copied) into a padded
// scratch map and emitted in tuple order at out.  Returns the set bits.
u64 bitmap_emit_block(dlg_ctx *ctx, const u32 *words, u32 lo, u64 W, u64 wpr, bool transposed, DBuf<u32> &scratch,
                      u64 *out) {
    const u64 nwords = W * wpr;
    if (nwords == 0) return 0;
    const u64 padded = (nwords + BE_TILE - 1) / BE_TILE * BE_TILE;
    scratch.ensure(padded);
    fill_u32(ctx, scratch.p + nwords, padded - nwords, 0u);
    if (transposed) {
        const u64 nt = (wpr + BT - 1) / BT;
        if (wpr <= BT_SMALL_WPR)
            DLG_LAUNCH(ctx, DLG_K_MISC, 8.0 * nwords, bm_transpose_small_k, grid_for(wpr * wpr * 32, 256), 256, 0,
                       words, scratch.p, W, wpr);
        else
            DLG_LAUNCH(ctx, DLG_K_MISC, 8.0 * nwords, bm_transpose_k, (unsigned)(nt * nt), 256, 0, words, scratch.p,
                       W, wpr);
    } else {
        HIPCHK(hipMemcpyAsync(scratch.p, words, 4 * nwords, hipMemcpyDeviceToDevice, ctx->stream));
    }
    DBuf<u64> bsum(ctx);
    u64 nb = 0;
    u64 h = words_count(ctx, scratch.p, nwords, bsum, nb);
    if (h) emit_pass<false, false>(ctx, lo, wpr, scratch.p, bsum, nb, h, out);
    return h;
}

// Several maps at once.  Map m: W[m] rows of wpr[m] words at src + sb[m],
// written transposed to dst + db[m]; 32 x 32-bit block q of all the maps'
// blocks (cum[m] of them before map m) is map m's block q - cum[m], one half
// wave each, as bm_transpose_small_k does for one map.
__global__ void bm_transpose_small_multi_k(const u32 *src, u32 *dst, const u64 *cum, const u64 *W, const u64 *wpr,
                                           const u64 *sb, const u64 *db, u32 nm) {
    const int lane = threadIdx.x & 31;
    const u64 q = ((u64)blockIdx.x * blockDim.x + threadIdx.x) >> 5;   // one block per half wave
    if (q >= cum[nm]) return;                                             // whole half waves leave together
    u32 a = 0, b = nm;   // the last map whose blocks start at or before q
    while (b - a > 1) {
        const u32 m = (a + b) >> 1;
        if (cum[m] <= q) a = m; else b = m;
    }
    const u64 blk = q - cum[a], wp = wpr[a], Wm = W[a];
    const u64 R = blk / wp, C = blk - R * wp;   // source block: rows 32R.., word column C
    const u64 r = R * 32 + lane;
    u32 x = r < Wm ? src[sb[a] + r * wp + C] : 0u;
    const u32 m5[5] = {0x0000ffffu, 0x00ff00ffu, 0x0f0f0f0fu, 0x33333333u, 0x55555555u};
#pragma unroll
    for (int st = 0; st < 5; st++) {
        const int j = 16 >> st;
        const u32 y = __shfl_xor(x, j, 32);
        x = (lane & j) ? ((x & ~m5[st]) | ((y & ~m5[st]) >> j)) : ((x & m5[st]) | ((y & m5[st]) << j));
    }
    const u64 c = C * 32 + lane;   // destination row = source column
    if (c < Wm) dst[db[a] + c * wp + R] = x;
}

// Zero map m's words past its own in its last tile: [db[m] + nw[m], db[m + 1]).
__global__ void bm_zero_tails_k(u32 *words, const u64 *db, const u64 *nw, u32 nm) {
    for (u32 m = blockIdx.x; m < nm; m += gridDim.x)
        for (u64 i = db[m] + nw[m] + threadIdx.x; i < db[m + 1]; i += blockDim.x) words[i] = 0u;
}

// dst row r = src row sigma[r] (rows of wpr words): one workgroup per 8 rows, coalesced.
__global__ __launch_bounds__(256) void row_gather_k(const u32 *src, u32 *dst, const u32 *sigma, u64 W, u64 wpr) {
    const u64 r0 = (u64)blockIdx.x * 8;
    for (u64 r = r0; r < r0 + 8 && r < W; r++) {
        const u32 *s = src + (u64)sigma[r] * wpr;
        u32 *d = dst + r * wpr;
        for (u64 j = threadIdx.x; j < wpr; j += 256) d[j] = s[j];
    }
}

u64 bitmap_emit_block_renumbered(dlg_ctx *ctx, u32 *words, u32 lo, u64 W, u64 wpr, const u32 *sigma,
                                 DBuf<u32> &scratch, u64 *out) {
    const u64 nwords = W * wpr;
    if (nwords == 0) return 0;
    const u64 padded = (nwords + BE_TILE - 1) / BE_TILE * BE_TILE;
    scratch.ensure(padded);
    fill_u32(ctx, scratch.p + nwords, padded - nwords, 0u);
    // rows back to input order: row x = the map's row sigma(x) (columns still renumbered)
    DLG_LAUNCH(ctx, DLG_K_MISC, 8.0 * nwords, row_gather_k, grid_for(W, 8), 256, 0, (const u32 *)words, scratch.p,
               sigma, W, wpr);
    // transposed: row = renumbered end, columns = sources in input order
    const u64 nt = (wpr + BT - 1) / BT;
    if (wpr <= BT_SMALL_WPR)
        DLG_LAUNCH(ctx, DLG_K_MISC, 8.0 * nwords, bm_transpose_small_k, grid_for(wpr * wpr * 32, 256), 256, 0,
                   (const u32 *)scratch.p, words, W, wpr);
    else
        DLG_LAUNCH(ctx, DLG_K_MISC, 8.0 * nwords, bm_transpose_k, (unsigned)(nt * nt), 256, 0, (const u32 *)scratch.p,
                   words, W, wpr);
    // rows back to input order: row z = row sigma(z)
    DLG_LAUNCH(ctx, DLG_K_MISC, 8.0 * nwords, row_gather_k, grid_for(W, 8), 256, 0, (const u32 *)words, scratch.p,
               sigma, W, wpr);
    DBuf<u64> bsum(ctx);
    u64 nb = 0;
    u64 h = words_count(ctx, scratch.p, nwords, bsum, nb);
    if (h) emit_pass<false, false>(ctx, lo, wpr, scratch.p, bsum, nb, h, out);
    return h;
}

u64 bitmap_emit(dlg_ctx *ctx, Bitmap &bm, DBuf<u64> &out, u64 expect, bool symmetric) {
    out.ensure(expect ? expect : 1);
    if (expect == 0) return 0;
    DBuf<u64> bsum(ctx);
    u64 h;
    if (bm.tr && !symmetric) {   // back to row-major, so that bit order is tuple order
        DBuf<u32> tw(ctx);
        h = bitmap_emit_block(ctx, bm.words.p, bm.lo, bm.W, bm.wpr, true, tw, out.p);
    } else {
        u64 nb = 0;
        h = words_count(ctx, bm.words.p, bm.nwords, bsum, nb);
        if (h == expect) emit_pass<false, false>(ctx, bm.lo, bm.wpr, bm.words.p, bsum, nb, h, out.p);
    }
    if (h != expect) fail(DLG_DEVICE_ERROR, "bitmap emission count mismatch");
    return h;
}

// ------------------------------------------------- single join into a map
// The relation comes as `rows` = (y, x) for each tuple (x, y): keyed by its
// second column, sorted, with dense[k] = the first row of key lo + k for k in
// [0, W] (dense[W] = the row count); col[i] = row i's first column minus lo.
// Row c of the map gets bit a for every (a, b), (b, c) of the relation: b
// runs over key c's run, a over key b's.  One workgroup per row keeps the
// row's wpr words in LDS.  It takes c's run 256 b's at a time: each thread
// looks up one b's run, a block scan places the runs end to end, and the
// threads then stride over that concatenation (SJR_U independent loads in
// flight each), finding their run by walking the scanned starts forward.
constexpr int SJR_T = 256, SJR_U = 8;
__global__ __launch_bounds__(SJR_T) void sj_rows_k(const u32 *col, const u64 *dense, u64 wpr, u32 *words) {
    extern __shared__ u32 s_row[];
    __shared__ u64 sc[SJR_T / WAVE + 1];
    __shared__ u64 s_pre[SJR_T + 1];
    __shared__ u64 s_q0[SJR_T];
    const u64 c = blockIdx.x;
    for (u64 j = threadIdx.x; j < wpr; j += SJR_T) s_row[j] = 0u;
    const u64 p0 = dense[c], p1 = dense[c + 1];
    for (u64 pb = p0; pb < p1; pb += SJR_T) {
        const u64 p = pb + threadIdx.x;
        u64 q0 = 0, len = 0;
        if (p < p1) {
            const u32 b = col[p];
            q0 = dense[b];
            len = dense[b + 1] - q0;
        }
        u64 tot;
        const u64 ex = block_exclusive_scan<u64>(len, sc, &tot);   // also orders the row's clearing
        s_pre[threadIdx.x] = ex;
        s_q0[threadIdx.x] = q0;
        if (threadIdx.x == 0) s_pre[SJR_T] = tot;
        __syncthreads();
        u32 sg = 0;   // s_pre[sg] <= j < s_pre[sg + 1] for this thread's current j
        for (u64 j0 = threadIdx.x; j0 < tot; j0 += (u64)SJR_U * SJR_T) {
            u64 q[SJR_U];
#pragma unroll
            for (int r = 0; r < SJR_U; r++) {
                const u64 j = j0 + (u64)r * SJR_T;
                q[r] = ~0ull;
                if (j < tot) {
                    while (s_pre[sg + 1] <= j) sg++;
                    q[r] = s_q0[sg] + (j - s_pre[sg]);
                }
            }
            u32 a[SJR_U];
#pragma unroll
            for (int r = 0; r < SJR_U; r++) a[r] = q[r] != ~0ull ? col[q[r]] : 0u;
#pragma unroll
            for (int r = 0; r < SJR_U; r++)
                if (q[r] != ~0ull) atomicOr(&s_row[a[r] >> 5], 1u << (a[r] & 31u));
        }
        __syncthreads();   // s_pre, s_q0 and sc are reused by the next 256 b's
    }
    __syncthreads();
    u32 *dst = words + c * wpr;
    for (u64 j = threadIdx.x; j < wpr; j += SJR_T) dst[j] = s_row[j];
}

__global__ void sj_col_k(const u64 *rows, u64 n, u32 lo, u32 *col) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) col[i] = tval(rows[i]) - lo;
}

// |{(a, b, c)}|: the sum over rows (y, x) of key x's run length.
__global__ __launch_bounds__(256) void sj_raw_k(const u64 *rows, u64 n, const u64 *dense, u32 lo,
                                                unsigned long long *raw) {
    __shared__ u64 s_part[256 / WAVE];
    u64 t = 0;
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const u64 b = (u64)(tval(rows[i]) - lo);
        t += dense[b + 1] - dense[b];
    }
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) t += __shfl_xor(t, o, WAVE);
    if ((threadIdx.x & (WAVE - 1)) == 0) s_part[threadIdx.x / WAVE] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        u64 sum = 0;
        for (int j = 0; j < 256 / WAVE; j++) sum += s_part[j];
        if (sum) atomicAdd(raw, (unsigned long long)sum);
    }
}

u64 self_join_raw(dlg_ctx *ctx, const u64 *rows, u64 n, const u64 *dense, u32 lo) {
    if (n == 0) return 0;
    DBuf<u64> c(ctx, 1);
    fill_u64(ctx, c.p, 1, 0ull);
    DLG_LAUNCH(ctx, DLG_K_MISC, 24.0 * n, sj_raw_k, sgrid(ctx, n), 256, 0, rows, n, dense, lo,
               (unsigned long long *)c.p);
    u64 h = 0;
    ctx->read_scalars(c.p, 1, &h);
    return h;
}

void bitmap_self_join(dlg_ctx *ctx, const u64 *rows, u64 n, const u64 *dense, Bitmap &bm) {
    if (bm.tr) fail(DLG_INVALID_INPUT, "self join map: row-major maps only");
    if (4 * bm.wpr > SJ_ROW_LDS_MAX) fail(DLG_INVALID_INPUT, "self join map: rows wider than LDS");
    DBuf<u32> col(ctx, n);
    DLG_LAUNCH(ctx, DLG_K_MISC, 12.0 * n, sj_col_k, sgrid(ctx, n), ctx->block_size, 0, rows, n, bm.lo, col.p);
    DLG_LAUNCH(ctx, DLG_K_BITMAP_INSERT, 4.0 * bm.nwords + 4.0 * n + 8.0 * bm.W, sj_rows_k, (unsigned)bm.W, SJR_T,
               (unsigned)(4 * bm.wpr), (const u32 *)col.p, dense, bm.wpr, bm.words.p);
}

// ------------------------------------------------- single join by rows
// For id ranges too wide for the map, or outputs too sparse for it: output
// row c is the sorted distinct first columns gathered from c's runs (the
// a's of (a, b) for each b of (b, c)).  One wave per row gathers its
// candidates into LDS (64 b's at a time, their runs laid end to end by a
// wave scan), sorts them there (bitonic), keeps the first of each value and
// writes them at the row's place in a candidate array; a scan of the kept
// counts then places each row in the output.  Rows with more than SJW_CAP
// candidates are left to the sorted raw join (the host checks the maximum).
constexpr int SJW_T = 256, SJW_W = SJW_T / WAVE;

__device__ __forceinline__ void sjw_sync() {   // LDS written by some lanes of a wave, read by others
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// rraw[c] = c's candidates: the sum over c's run of the run lengths of its b's;
// the largest to mx[0], the longest run of b's to mx[1].
__global__ __launch_bounds__(256) void sj_row_raw_k(const u32 *col, const u64 *dense, u64 W, u64 *rraw,
                                                    unsigned long long *mx) {
    __shared__ u64 s_m[2][256 / WAVE];
    u64 m = 0, mb = 0;
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 c = (u64)blockIdx.x * blockDim.x + threadIdx.x; c < W; c += stride) {
        u64 t = 0;
        const u64 p0 = dense[c], p1 = dense[c + 1];
        for (u64 p = p0; p < p1; p++) {
            const u32 b = col[p];
            t += dense[b + 1] - dense[b];
        }
        rraw[c] = t;
        m = t > m ? t : m;
        mb = p1 - p0 > mb ? p1 - p0 : mb;
    }
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        const u64 x = __shfl_xor(m, o, WAVE), y = __shfl_xor(mb, o, WAVE);
        m = x > m ? x : m;
        mb = y > mb ? y : mb;
    }
    if ((threadIdx.x & (WAVE - 1)) == 0) {
        s_m[0][threadIdx.x / WAVE] = m;
        s_m[1][threadIdx.x / WAVE] = mb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 256 / WAVE; w++) {
            m = s_m[0][w] > m ? s_m[0][w] : m;
            mb = s_m[1][w] > mb ? s_m[1][w] : mb;
        }
        // an atomic only where it can still raise the maximum (a stale read only costs an atomic)
        if (m > __hip_atomic_load(mx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(mx, (unsigned long long)m);
        if (mb > __hip_atomic_load(mx + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(mx + 1, (unsigned long long)mb);
    }
}

// Bitonic sort, ascending, of SJW_REG * WAVE values held as v[r] = element
// r WAVE + lane: strides below WAVE exchange with lane ^ j by shuffle, larger
// ones compare registers of the same lane.
constexpr int SJW_REG = 4;
template <int R>
__device__ __forceinline__ void sjw_sort_regs(u32 (&v)[R], int lane) {
    constexpr u32 P = R * WAVE;
#pragma unroll
    for (u32 k = 2; k <= P; k <<= 1) {
#pragma unroll
        for (u32 j = k >> 1; j > 0; j >>= 1) {
            if (j >= (u32)WAVE) {
                const u32 m = j / WAVE;
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int q = r ^ (int)m;
                    if (q > r) {
                        const u32 e = (u32)r * WAVE + (u32)lane;
                        const bool asc = (e & k) == 0;
                        const u32 x = v[r], y = v[q];
                        if ((x > y) == asc) {
                            v[r] = y;
                            v[q] = x;
                        }
                    }
                }
            } else {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const u32 e = (u32)r * WAVE + (u32)lane;
                    const bool asc = (e & k) == 0, low = ((u32)lane & j) == 0;
                    const u32 y = (u32)__shfl_xor((int)v[r], (int)j, WAVE);
                    const u32 mn = v[r] < y ? v[r] : y, mx = v[r] < y ? y : v[r];
                    v[r] = low == asc ? mn : mx;
                }
            }
        }
    }
}

// Rows the register kernel takes: at most SJW_REG * WAVE candidates from at
// most WAVE b's (empty rows too).
__device__ __forceinline__ bool sjw_small(u64 raw, u64 nb) { return raw <= (u64)SJW_REG * WAVE && nb <= (u64)WAVE; }

// A small row's candidates: j = r WAVE + lane found among the lanes' runs
// (q0, pre: this lane's run start and first candidate) by a binary search
// over the lanes' starts, loaded into v[r] (R loads in flight), sorted in
// registers; the first of each value goes to out in order.  Returns their count.
template <int R>
__device__ __forceinline__ u32 sjw_small_row(const u32 *col, u32 q0, u32 pre, u64 nb, u64 raw, int lane, u32 *out) {
    u32 v[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const u32 j = (u32)r * WAVE + (u32)lane;
        int s = 0;   // the last lane whose start is <= j (starts rise with the lane; empty runs share one)
#pragma unroll
        for (int st = WAVE / 2; st > 0; st >>= 1) {
            const u32 ps = __shfl(pre, s + st, WAVE);
            if (s + st < (int)nb && ps <= j) s += st;
        }
        const u32 qs = __shfl(q0, s, WAVE), ps = __shfl(pre, s, WAVE);
        v[r] = (u64)j < raw ? col[(u64)qs + (j - ps)] : 0xffffffffu;
    }
    sjw_sort_regs<R>(v, lane);
    u32 kept = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const u64 e = (u64)r * WAVE + (u64)lane;
        // element e - 1: lane - 1's v[r], or for lane 0 the last lane's v[r - 1]
        const u32 up = __shfl(v[r], (lane + WAVE - 1) & (WAVE - 1), WAVE);
        const u32 wrap = __shfl(v[r > 0 ? r - 1 : 0], WAVE - 1, WAVE);
        const u32 prev = lane > 0 ? up : wrap;
        const bool keep = e < raw && (e == 0 || v[r] != prev);
        const u64 bal = __ballot(keep);
        if (keep) out[kept + __builtin_amdgcn_mbcnt_hi((u32)(bal >> 32), __builtin_amdgcn_mbcnt_lo((u32)bal, 0u))] = v[r];
        kept += (u32)__popcll(bal);
    }
    return kept;
}

// Small rows with no LDS: lane l takes the row's l-th b, a wave scan lays the
// b's runs end to end, and candidate j = r WAVE + lane is found by a binary
// search over the lanes' starts (by shuffle) and loaded straight into v[r]
// (SJW_REG loads in flight), then sorted in registers and deduplicated.
__global__ __launch_bounds__(256) void sj_row_small_k(const u32 *col, const u64 *dense, const u64 *roff, u64 W,
                                                      u32 *cand, u32 *ucnt) {
    const int lane = threadIdx.x & (WAVE - 1);
    for (u64 c = ((u64)blockIdx.x * blockDim.x + threadIdx.x) / WAVE; c < W; c += (u64)gridDim.x * blockDim.x / WAVE) {
        const u64 r0 = roff[c], raw = roff[c + 1] - r0;
        const u64 p0 = dense[c], nb = dense[c + 1] - p0;
        if (!sjw_small(raw, nb)) continue;   // sj_row_sort_k's
        if (raw == 0) {
            if (lane == 0) ucnt[c] = 0u;
            continue;
        }
        u32 q0 = 0, len = 0;   // this lane's b: its run (a row's b's and runs fit 32 bits here)
        if ((u64)lane < nb) {
            const u32 b = col[p0 + (u64)lane];
            const u64 d0 = dense[b];
            q0 = (u32)d0;
            len = (u32)(dense[b + 1] - d0);
        }
        u32 inc = len;
#pragma unroll
        for (int d = 1; d < WAVE; d <<= 1) {
            const u32 t = __shfl_up(inc, d, WAVE);
            if (lane >= d) inc += t;
        }
        const u32 pre = inc - len;   // this lane's first candidate
        u32 kept;
        if (raw <= (u64)WAVE) kept = sjw_small_row<1>(col, q0, pre, nb, raw, lane, cand + r0);
        else if (raw <= 2 * (u64)WAVE) kept = sjw_small_row<2>(col, q0, pre, nb, raw, lane, cand + r0);
        else kept = sjw_small_row<SJW_REG>(col, q0, pre, nb, raw, lane, cand + r0);
        if (lane == 0) ucnt[c] = kept;
    }
}

__global__ __launch_bounds__(SJW_T) void sj_row_sort_k(const u32 *col, const u64 *dense, const u64 *roff, u64 W,
                                                       u32 *cand, u32 *ucnt) {
    __shared__ u32 s_buf[SJW_W][SJW_CAP];
    __shared__ u64 s_pre[SJW_W][WAVE + 1];
    __shared__ u64 s_q0[SJW_W][WAVE];
    const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
    u32 *buf = s_buf[wv];
    u64 *pre = s_pre[wv], *q0s = s_q0[wv];
    for (u64 c = (u64)blockIdx.x * SJW_W + wv; c < W; c += (u64)gridDim.x * SJW_W) {   // wave-uniform
        const u64 r0 = roff[c], raw = roff[c + 1] - r0;
        const u64 p1 = dense[c + 1];
        if (sjw_small(raw, p1 - dense[c]) || raw > SJW_CAP) continue;   // sj_row_small_k's, or a heavy row
        // gather the row's candidates into buf[0, raw)
        u64 filled = 0;
        for (u64 pb = dense[c]; pb < p1; pb += WAVE) {
            u64 q0 = 0, len = 0;
            if (pb + (u64)lane < p1) {
                const u32 b = col[pb + (u64)lane];
                q0 = dense[b];
                len = dense[b + 1] - q0;
            }
            u64 inc = len;   // wave scan of the run lengths
#pragma unroll
            for (int d = 1; d < WAVE; d <<= 1) {
                const u64 t = __shfl_up(inc, d, WAVE);
                if (lane >= d) inc += t;
            }
            const u64 tot = __shfl(inc, WAVE - 1, WAVE);
            pre[lane] = inc - len;
            q0s[lane] = q0;
            if (lane == 0) pre[WAVE] = tot;
            sjw_sync();
            u32 sg = 0;
            for (u64 j = lane; j < tot; j += WAVE) {
                while (pre[sg + 1] <= j) sg++;
                buf[filled + j] = col[q0s[sg] + (j - pre[sg])];
            }
            filled += tot;
            sjw_sync();   // pre and q0s are rewritten for the next 64 b's
        }
        if (raw <= (u64)SJW_REG * WAVE) {   // sorted in registers: element r WAVE + lane in v[r]
            sjw_sync();
            u32 v[SJW_REG];
#pragma unroll
            for (int r = 0; r < SJW_REG; r++) {
                const u64 e = (u64)r * WAVE + (u64)lane;
                v[r] = e < raw ? buf[e] : 0xffffffffu;
            }
            sjw_sort_regs<SJW_REG>(v, lane);
            u32 kept = 0;
#pragma unroll
            for (int r = 0; r < SJW_REG; r++) {
                const u64 e = (u64)r * WAVE + (u64)lane;
                // element e - 1: lane - 1's v[r], or for lane 0 the last lane's v[r - 1]
                const u32 up = __shfl(v[r], (lane + WAVE - 1) & (WAVE - 1), WAVE);
                const u32 wrap = __shfl(v[r > 0 ? r - 1 : 0], WAVE - 1, WAVE);
                const u32 prev = lane > 0 ? up : wrap;
                const bool keep = e < raw && (e == 0 || v[r] != prev);
                const u64 bal = __ballot(keep);
                if (keep)
                    cand[r0 + kept + __builtin_amdgcn_mbcnt_hi((u32)(bal >> 32), __builtin_amdgcn_mbcnt_lo((u32)bal, 0u))] =
                        v[r];
                kept += (u32)__popcll(bal);
            }
            if (lane == 0) ucnt[c] = kept;
            sjw_sync();   // buf is rewritten for the next row
            continue;
        }
        // bitonic sort of buf[0, P), P the next power of two, padded with the largest value
        u32 P = 2;
        while ((u64)P < raw) P <<= 1;
        for (u32 j = (u32)raw + lane; j < P; j += WAVE) buf[j] = 0xffffffffu;
        sjw_sync();
        for (u32 k = 2; k <= P; k <<= 1) {
            for (u32 jj = k >> 1; jj > 0; jj >>= 1) {
                for (u32 i = lane; i < P; i += WAVE) {
                    const u32 l = i ^ jj;
                    if (l > i) {
                        const u32 x = buf[i], y = buf[l];
                        if ((x > y) == ((i & k) == 0)) {
                            buf[i] = y;
                            buf[l] = x;
                        }
                    }
                }
                sjw_sync();
            }
        }
        // the first of each value, in order, at the row's place in cand
        u32 kept = 0;
        for (u64 j0 = 0; j0 < raw; j0 += WAVE) {
            const u64 j = j0 + (u64)lane;
            const u32 v = j < raw ? buf[j] : 0u;
            const bool keep = j < raw && (j == 0 || buf[j - 1] != v);
            const u64 bal = __ballot(keep);
            if (keep)
                cand[r0 + kept + __builtin_amdgcn_mbcnt_hi((u32)(bal >> 32), __builtin_amdgcn_mbcnt_lo((u32)bal, 0u))] = v;
            kept += (u32)__popcll(bal);
        }
        if (lane == 0) ucnt[c] = kept;
        sjw_sync();   // buf is rewritten for the next row
    }
}

// out[uoff[c] ..] = (c, a) for the kept a's of row c: one wave per row.
__global__ __launch_bounds__(256) void sj_row_emit_k(const u32 *cand, const u64 *roff, const u64 *uoff, u64 W, u32 lo,
                                                     u64 *out) {
    const int lane = threadIdx.x & (WAVE - 1);
    for (u64 c = ((u64)blockIdx.x * blockDim.x + threadIdx.x) / WAVE; c < W; c += (u64)gridDim.x * blockDim.x / WAVE) {
        const u64 r0 = roff[c], o0 = uoff[c], m = uoff[c + 1] - o0;
        for (u64 k = lane; k < m; k += WAVE) out[o0 + k] = tmake(lo + (u32)c, lo + cand[r0 + k]);
    }
}

// Heavy rows (more than SJW_CAP candidates) go through a sort: hraw[c] =
// the row's candidates if it is heavy, else 0 (scanned into the heavy rows'
// offsets in J).
__global__ void sj_heavy_raw_k(const u64 *roff, u64 W, u64 *hraw) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 c = (u64)blockIdx.x * blockDim.x + threadIdx.x; c < W; c += stride) {
        const u64 r = roff[c + 1] - roff[c];
        hraw[c] = r > SJW_CAP ? r : 0;
    }
}

// A heavy row's candidates as (c, a) tuples at J[hoff[c] ..]: one wave per
// row, 64 b's at a time, candidate j found among the lanes' runs by a binary
// search over their starts.
__global__ __launch_bounds__(256) void sj_heavy_gather_k(const u32 *col, const u64 *dense, const u64 *hoff, u64 W,
                                                         u32 lo, u64 *J) {
    const int lane = threadIdx.x & (WAVE - 1);
    for (u64 c = ((u64)blockIdx.x * blockDim.x + threadIdx.x) / WAVE; c < W; c += (u64)gridDim.x * blockDim.x / WAVE) {
        const u64 h0 = hoff[c];
        if (hoff[c + 1] == h0) continue;   // a light row
        const u64 p1 = dense[c + 1];
        u64 filled = 0;
        for (u64 pb = dense[c]; pb < p1; pb += WAVE) {
            const u64 nb = p1 - pb < (u64)WAVE ? p1 - pb : (u64)WAVE;
            u64 q0 = 0, len = 0;
            if ((u64)lane < nb) {
                const u32 b = col[pb + (u64)lane];
                q0 = dense[b];
                len = dense[b + 1] - q0;
            }
            u64 inc = len;
#pragma unroll
            for (int d = 1; d < WAVE; d <<= 1) {
                const u64 t = __shfl_up(inc, d, WAVE);
                if (lane >= d) inc += t;
            }
            const u64 tot = __shfl(inc, WAVE - 1, WAVE), pre = inc - len;
            for (u64 j0 = 0; j0 < tot; j0 += WAVE) {
                const u64 j = j0 + (u64)lane;
                int sidx = 0;   // the last lane whose start is <= j
#pragma unroll
                for (int st = WAVE / 2; st > 0; st >>= 1) {
                    const u64 ps = __shfl(pre, sidx + st, WAVE);
                    if (sidx + st < (int)nb && ps <= j) sidx += st;
                }
                const u64 qs = __shfl(q0, sidx, WAVE), ps = __shfl(pre, sidx, WAVE);
                if (j < tot) J[h0 + filled + j] = tmake(lo + (u32)c, lo + col[qs + (j - ps)]);
            }
            filled += tot;
        }
    }
}

// H = the heavy rows' distinct (c, a), sorted.  First the head of each row's
// run to head[c], then each a to the row's place in cand and the run's length
// to ucnt[c] (from its last element).
__global__ void sj_heavy_heads_k(const u64 *H, u64 nh, u32 lo, u32 *head) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < nh; i += stride)
        if (i == 0 || tkey(H[i]) != tkey(H[i - 1])) head[tkey(H[i]) - lo] = (u32)i;
}
__global__ void sj_heavy_place_k(const u64 *H, u64 nh, u32 lo, const u32 *head, const u64 *roff, u32 *cand,
                                 u32 *ucnt) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < nh; i += stride) {
        const u64 x = H[i];
        const u32 c = tkey(x) - lo, k = (u32)i - head[c];
        cand[roff[c] + k] = tval(x) - lo;
        if (i + 1 == nh || tkey(H[i + 1]) != tkey(x)) ucnt[c] = k + 1;
    }
}

bool self_join_rows(dlg_ctx *ctx, const u64 *rows, u64 n, const u64 *dense, u32 lo, u64 W, DBuf<u64> &out,
                    u64 *nout, u64 *raw) {
    *nout = 0;
    *raw = 0;
    if (n == 0 || W == 0) return true;
    DBuf<u32> col(ctx, n);
    DLG_LAUNCH(ctx, DLG_K_MISC, 12.0 * n, sj_col_k, sgrid(ctx, n), ctx->block_size, 0, rows, n, lo, col.p);
    DBuf<u64> roff(ctx, W + 1), mx(ctx, 2);
    fill_u64(ctx, mx.p, 2, 0ull);
    DLG_LAUNCH(ctx, DLG_K_MISC, 16.0 * W + 12.0 * n, sj_row_raw_k, sgrid(ctx, W), 256, 0, (const u32 *)col.p, dense,
               W, roff.p, (unsigned long long *)mx.p);
    u64 hm[2];
    ctx->read_scalars(mx.p, 2, hm);
    const u64 m = hm[0];
    const bool big_b = hm[1] > (u64)WAVE;   // a row with more b's than the register kernel takes
    scan_u64_async(ctx, roff.p, W, roff.p, roff.p + W);   // roff[W] = all candidates
    u64 total = 0;
    ctx->read_scalars(roff.p + W, 1, &total);
    *raw = total;
    DBuf<u32> cand(ctx, total ? total : 1), ucnt(ctx, W);
    // the wave-per-row kernels: a wave per row, up to eight 4-wave workgroups per CU
    const unsigned g = (unsigned)std::min<u64>((W + SJW_W - 1) / SJW_W, (u64)ctx->num_cus * 8);
    DLG_LAUNCH(ctx, DLG_K_BITMAP_INSERT, 8.0 * total + 8.0 * n + 16.0 * W, sj_row_small_k, g, 256, 0,
               (const u32 *)col.p, dense, (const u64 *)roff.p, W, cand.p, ucnt.p);
    if (m > (u64)SJW_REG * WAVE || big_b)
        DLG_LAUNCH(ctx, DLG_K_BITMAP_INSERT, 16.0 * W, sj_row_sort_k, g, SJW_T, 0, (const u32 *)col.p, dense,
                   (const u64 *)roff.p, W, cand.p, ucnt.p);
    if (m > SJW_CAP) {   // heavy rows: their candidates materialised, sorted and deduplicated
        DBuf<u64> hoff(ctx, W + 1);
        DLG_LAUNCH(ctx, DLG_K_MISC, 16.0 * W, sj_heavy_raw_k, sgrid(ctx, W), ctx->block_size, 0, (const u64 *)roff.p,
                   W, hoff.p);
        scan_u64_async(ctx, hoff.p, W, hoff.p, hoff.p + W);
        u64 nj = 0;
        ctx->read_scalars(hoff.p + W, 1, &nj);
        if (nj > 0xffffffffull) return false;   // so many heavy candidates: the caller's sorted join
        DBuf<u64> J(ctx, nj), tmp(ctx), H(ctx, nj);
        DLG_LAUNCH(ctx, DLG_K_JOIN_EXPAND, 8.0 * nj + 16.0 * W, sj_heavy_gather_k, g, 256, 0, (const u32 *)col.p,
                   dense, (const u64 *)hoff.p, W, lo, J.p);
        const u32 hb[4] = {lo, (u32)(lo + W - 1), lo, (u32)(lo + W - 1)};
        radix_sort(ctx, J, nj, tmp, 0, hb);
        const u64 nh = select(ctx, SEL_UNIQUE, J.p, nj, nullptr, 0, H.p);
        DBuf<u32> head(ctx, W);
        DLG_LAUNCH(ctx, DLG_K_MISC, 16.0 * nh, sj_heavy_heads_k, sgrid(ctx, nh), ctx->block_size, 0,
                   (const u64 *)H.p, nh, lo, head.p);
        DLG_LAUNCH(ctx, DLG_K_MISC, 24.0 * nh, sj_heavy_place_k, sgrid(ctx, nh), ctx->block_size, 0,
                   (const u64 *)H.p, nh, lo, (const u32 *)head.p, (const u64 *)roff.p, cand.p, ucnt.p);
    }
    DBuf<u64> uoff(ctx, W + 1);
    const u64 h = scan_u32_u64(ctx, ucnt.p, W, uoff.p);
    fill_u64(ctx, uoff.p + W, 1, h);
    out.ensure(h ? h : 1);
    if (h)
        DLG_LAUNCH(ctx, DLG_K_MISC, 12.0 * h + 24.0 * W, sj_row_emit_k, g, 256, 0,
                   (const u32 *)cand.p, (const u64 *)roff.p, (const u64 *)uoff.p, W, lo, out.p);
    *nout = h;
    return true;
}

// ------------------------------------------------- many maps, one emission
// The maps each at a whole tile of one scratch (transposed back when
// `transposed`), one count pass and one scan over all the tiles, and one
// emission launch whose workgroups find their tile's map: no launch or
// read-back per map.  Maps in the order given, each in bit order, so the
// output is sorted when the maps' id ranges rise.
u64 bitmap_emit_maps(dlg_ctx *ctx, const u32 *words, const std::vector<MapRef> &maps, bool transposed, u64 *out,
                     u64 expect) {
    const u32 nm = (u32)maps.size();
    if (nm == 0) return 0;
    std::vector<u64> db(nm + 1), nw(nm), W(nm), wpr(nm), sb(nm), cum(nm + 1), mt(nm + 1);
    std::vector<u32> lo(nm);
    u64 t = 0, q = 0;
    for (u32 m = 0; m < nm; m++) {
        const MapRef &r = maps[m];
        mt[m] = t;
        db[m] = t * BE_TILE;
        nw[m] = r.W * r.wpr;
        W[m] = r.W;
        wpr[m] = r.wpr;
        sb[m] = r.base;
        lo[m] = r.lo;
        cum[m] = q;
        if (r.wpr <= BT_SMALL_WPR) q += r.wpr * r.wpr;
        t += (nw[m] + BE_TILE - 1) / BE_TILE;
    }
    mt[nm] = t;
    db[nm] = t * BE_TILE;
    cum[nm] = q;
    if (t == 0) return 0;
    // the tables: db, nw, W, wpr, sb, cum, mt (u64) and lo (u32)
    const size_t L = nm + 1;
    DBuf<u64> tab(ctx, 7 * L + (L + 1) / 2);
    u64 *d_db = tab.p, *d_nw = d_db + L, *d_W = d_nw + L, *d_wpr = d_W + L, *d_sb = d_wpr + L, *d_cum = d_sb + L,
        *d_mt = d_cum + L;
    u32 *d_lo = (u32 *)(d_mt + L);
    std::vector<u64> h(7 * L + (L + 1) / 2, 0);
    std::copy(db.begin(), db.end(), h.begin());
    std::copy(nw.begin(), nw.end(), h.begin() + L);
    std::copy(W.begin(), W.end(), h.begin() + 2 * L);
    std::copy(wpr.begin(), wpr.end(), h.begin() + 3 * L);
    std::copy(sb.begin(), sb.end(), h.begin() + 4 * L);
    std::copy(cum.begin(), cum.end(), h.begin() + 5 * L);
    std::copy(mt.begin(), mt.end(), h.begin() + 6 * L);
    std::memcpy(h.data() + 7 * L, lo.data(), 4 * nm);
    HIPCHK(hipMemcpyAsync(tab.p, h.data(), 8 * h.size(), hipMemcpyHostToDevice, ctx->stream));
    DBuf<u32> sc(ctx, t * BE_TILE);
    if (transposed) {
        if (q)
            DLG_LAUNCH(ctx, DLG_K_MISC, 8.0 * q * 32, bm_transpose_small_multi_k, grid_for(q * 32, 256), 256, 0,
                       words, sc.p, (const u64 *)d_cum, (const u64 *)d_W, (const u64 *)d_wpr, (const u64 *)d_sb,
                       (const u64 *)d_db, nm);
        for (u32 m = 0; m < nm; m++) {
            if (wpr[m] <= BT_SMALL_WPR) continue;
            const u64 nt = (wpr[m] + BT - 1) / BT;
            DLG_LAUNCH(ctx, DLG_K_MISC, 8.0 * nw[m], bm_transpose_k, (unsigned)(nt * nt), 256, 0, words + sb[m],
                       sc.p + db[m], W[m], wpr[m]);
        }
    } else {
        for (u32 m = 0; m < nm; m++)
            if (nw[m])
                HIPCHK(hipMemcpyAsync(sc.p + db[m], words + sb[m], 4 * nw[m], hipMemcpyDeviceToDevice, ctx->stream));
    }
    u64 tails = db[nm];   // the padding words past each map
    for (u32 m = 0; m < nm; m++) tails -= nw[m];
    DLG_LAUNCH(ctx, DLG_K_MISC, 4.0 * tails, bm_zero_tails_k, std::min<u32>(nm, 4096u), 256, 0, sc.p,
               (const u64 *)d_db, (const u64 *)d_nw, nm);
    DBuf<u64> bsum(ctx, t), tot(ctx, 1);
    DLG_LAUNCH(ctx, DLG_K_MISC, 4.0 * t * BE_TILE, bm_count_k, (unsigned)t, BE_T, 0, (const u32 *)sc.p, bsum.p);
    scan_blocks(ctx, bsum.p, t, tot.p);
    u64 total = 0;
    ctx->read_scalars(tot.p, 1, &total);
    if (total != expect) fail(DLG_DEVICE_ERROR, "bitmap emission count mismatch");
    if (total)
        DLG_LAUNCH(ctx, DLG_K_MISC, 4.0 * t * BE_TILE + 8.0 * total, (bm_emit_k<false, false, true>), (unsigned)t,
                   BE_T, 0, sc.p, (const u64 *)bsum.p, 0u, (u64)1, out, (const u64 *)d_mt, (const u32 *)d_lo,
                   (const u64 *)d_wpr, nm);
    return total;
}

}  // namespace dlg
