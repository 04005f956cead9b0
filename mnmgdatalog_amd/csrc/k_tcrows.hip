// k_tcrows.hip -- TC rounds with one wave per row of the bitmap.
//
// With the full relation F as a transposed bitmap (row x = a path's source,
// column z = its end), a TC round (tc.cu:140-314)
//     N = { (z, x) : (y, x) in Δ, (y, z) in E } \ F ;  F = F u N
// only ever writes row x for the probes (y, x) of source x.  So Δ is kept
// grouped by source -- a segment (x, begin, length) over an array of ends y
// -- and one wave takes a whole segment (four per workgroup, no barriers):
//   1 expand: the segment's probes y walk their adjacency in a dense CSR of
//     E over the id range (no hash probe, no count pass);
//   2 combine: the bits of each touched word of row x are ORed together in
//     an LDS hash table;
//   3 set: one plain read and (if a bit is new) one plain write per touched
//     word -- no other wave touches row x in this round, so no atomics
//     (device atomics execute at the memory side, ~1e10/s for the whole
//     chip); kernel boundaries make the rows coherent between rounds;
//   4 emit: the bits that were clear are the segment's new tuples, written
//     contiguously as the next round's segment for x, into a region reserved
//     by a scan of the segments' output bounds (no global counter; a
//     segment whose touched words overflow the table flushes several times
//     into its region).
#include "dlg_internal.h"

#include <algorithm>
#include <vector>

namespace dlg {

constexpr int TR_T = 256;                    // threads per workgroup: four waves, one segment each
constexpr int TR_W = TR_T / WAVE;
constexpr int TR_BATCH = 256;                // rows per batch of the load-balanced path (TR_U per lane)
constexpr u32 TR_EMPTY = 0xffffffffu;
constexpr int TR_SMALL = 8;   // chunks whose probes all have at most this many matches: lane per probe

// The word table of a wave: 2^LOG slots.  A larger table flushes a long
// segment fewer times; a smaller one leaves LDS for more waves per CU (short
// segments).  Insert batches are sized so the table stays <= 3/4 full.
template <int LOG>
struct TrCfg {
    static constexpr int H = 1 << LOG;
    static constexpr int STEP = LOG >= 10 ? 8 : 2;   // matches per lane between two flush checks
    static constexpr int FLUSH = H * 3 / 4 - TR_BATCH;              // before a load-balanced batch
    static constexpr int FLUSH_SMALL = H * 3 / 4 - STEP * WAVE;     // before a lane-per-probe step
    static constexpr int BLOCKS_PER_CU = LOG >= 10 ? 3 : 5;         // what LDS allows
    static_assert(FLUSH > 0 && FLUSH_SMALL > 0 && TR_SMALL % STEP == 0, "table load");
};

template <int LOG>
__device__ __forceinline__ u32 tr_hash(u32 w) { return (w * 0x9e3779b1u) >> (32 - LOG); }

// Memory written by some lanes of a wave and then read by others (the LDS
// table; F's words across two flushes of one segment): a workgroup-scope
// fence orders them (the wave's lanes share one CU and its L1), and the
// wave barrier keeps the compiler from moving accesses across it.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ u64 wave_excl_scan(u64 v, u64 *total) {
    const int lane = threadIdx.x & (WAVE - 1);
    u64 inc = v;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        u64 t = __shfl_up(inc, d, WAVE);
        if (lane >= d) inc += t;
    }
    *total = __shfl(inc, WAVE - 1, WAVE);
    return inc - v;
}

template <int LOG>
struct TrWave {
    u32 key[TrCfg<LOG>::H];
    u32 bits[TrCfg<LOG>::H];
    u16 used[TrCfg<LOG>::H];
    u64 p_off[WAVE + 1];    // row offsets of the chunk's probes (p_off[WAVE] = total)
    u64 p_start[WAVE];      // CSR start of each probe's adjacency
    u32 nused;
};

constexpr int TR_U = 4;   // words / rows per lane with their loads in flight together
constexpr u32 TR_DMAX_BOUND = 16;   // degrees up to this: output regions bounded by new ends * max degree

// OR bit z of the row into the wave's table (z relative to the id range).
template <int LOG>
__device__ __forceinline__ void tr_insert(TrWave<LOG> &s, u32 z) {
    u32 w = z >> 5, mk = 1u << (z & 31);
    u32 h = tr_hash<LOG>(w);
    while (true) {
        u32 k = atomicCAS(&s.key[h], TR_EMPTY, w);
        if (k == TR_EMPTY) {
            s.used[atomicAdd(&s.nused, 1u)] = (u16)h;
            break;
        }
        if (k == w) break;
        h = (h + 1) & (TrCfg<LOG>::H - 1);
    }
    atomicOr(&s.bits[h], mk);
}

// Flush the wave's table into row `rowbase` of F: per occupied slot a read
// and, for new bits, a write of the word (TR_U words per lane in flight);
// the new bits are written as ends z (packed) to dst[0..), in slot order.
// Returns the number written.  The table is left empty.
template <int LOG>
__device__ u64 tr_flush(TrWave<LOG> &s, u32 *F, u64 rowbase, u32 lo, u32 *dst) {
    const int lane = threadIdx.x & (WAVE - 1);
    wave_sync();
    const u32 n = s.nused;
    u64 cnt = 0;
    for (u32 j0 = 0; j0 < n; j0 += TR_U * WAVE) {
        u32 slot[TR_U], w[TR_U], b[TR_U], cur[TR_U];
#pragma unroll
        for (int q = 0; q < TR_U; q++) {
            u32 j = j0 + q * WAVE + lane;
            slot[q] = j < n ? s.used[j] : 0u;
            w[q] = j < n ? s.key[slot[q]] : 0u;
            b[q] = j < n ? s.bits[slot[q]] : 0u;
        }
#pragma unroll
        for (int q = 0; q < TR_U; q++) cur[q] = b[q] ? F[rowbase + w[q]] : 0u;   // the row's owner:
#pragma unroll                                                              // plain read-modify-write
        for (int q = 0; q < TR_U; q++) {
            u32 nb = b[q] & ~cur[q];
            if (nb) F[rowbase + w[q]] = cur[q] | nb;
            if (b[q]) s.bits[slot[q]] = nb;
            cnt += __popc(nb);
        }
    }
    u64 tot;
    u64 off = wave_excl_scan(cnt, &tot);
    for (u32 j = lane; j < n; j += WAVE) {
        u32 slot = s.used[j];
        u32 w = s.key[slot], nb = s.bits[slot];
        while (nb) {
            dst[off++] = lo + w * 32 + (u32)(__ffs(nb) - 1);
            nb &= nb - 1;
        }
        s.key[slot] = TR_EMPTY;
        s.bits[slot] = 0u;
    }
    wave_sync();
    if (lane == 0) s.nused = 0;
    wave_sync();
    return tot;
}

// One round.  Segment s (source seg_x[s], ends dy[seg_beg[s] ..][0, seg_len[s]))
// writes its new ends at ny[seg_out[s] ..] -- seg_out is the exclusive scan
// of the segments' next-round bounds -- and its next length and bound at the
// same index s.  The bound is the segment's next join rows, Σ deg over its
// new ends (exact; `dmax` = 0), or, when every degree is at most dmax,
// min(new ends * dmax, W) (no degree lookups; the output has gaps).  The
// round's join rows are counted on the way (counters[2]).  Workgroups take
// static ranges of segments; the four waves of a workgroup share the range
// through an LDS counter, so no global counter is contended.
template <int LOG>
__global__ __launch_bounds__(TR_T) void tc_rows_k(const u32 *dy, const u32 *seg_x, const u64 *seg_beg,
                                                  const u32 *seg_len, const u64 *seg_out, u64 nseg,
                                                  const u64 *csr_off, const u32 *csr_dst, const u32 *deg,
                                                  u32 dmax, u32 *F, const u64 *row_base, const u32 *col_lo,
                                                  u32 lo, u64 wcap, u32 *ny, u32 *nseg_len, u64 *next_rows,
                                                  unsigned long long *counters) {
    using C = TrCfg<LOG>;
    __shared__ TrWave<LOG> sw[TR_W];
    __shared__ u32 s_next;
    const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
    TrWave<LOG> &s = sw[wv];
    for (int i = lane; i < C::H; i += WAVE) {
        s.key[i] = TR_EMPTY;
        s.bits[i] = 0u;
    }
    if (lane == 0) s.nused = 0;
    const u64 per = (nseg + gridDim.x - 1) / gridDim.x;
    const u64 b0 = (u64)blockIdx.x * per, b1 = b0 + per < nseg ? b0 + per : nseg;
    if (threadIdx.x == 0) s_next = 0;
    __syncthreads();
    u64 wave_new = 0, wave_rows = 0;
    while (true) {
        u32 k0 = 0;
        if (lane == 0) k0 = atomicAdd(&s_next, 1u);
        const u64 sg = b0 + __shfl(k0, 0, WAVE);
        if (sg >= b1) break;
        const u32 len = seg_len[sg];
        u64 rows_next = 0, cnt = 0;
        if (len) {
            const u32 x = seg_x[sg];
            const u64 beg = seg_beg[sg], out0 = seg_out[sg];
            const u64 rowbase = row_base[x - lo];   // x's row and its block's first id
            const u32 clo = col_lo[x - lo];
            // the chunks' ends are loaded one chunk ahead (software pipelining)
            u32 yn[2];
#pragma unroll
            for (int h = 0; h < 2; h++) yn[h] = (u32)(h * WAVE + lane) < len ? dy[beg + h * WAVE + lane] : 0u;
            for (u32 pc = 0; pc < len; pc += 2 * WAVE) {
                // two probes per lane and their adjacency ranges
                u64 st[2] = {0, 0};
                u32 d[2] = {0, 0};
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    if (pc + h * WAVE + lane < len) {
                        u32 y = yn[h] - lo;
                        st[h] = csr_off[y];
                        d[h] = deg[y];
                    }
                }
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    u32 i = pc + 2 * WAVE + h * WAVE + lane;
                    yn[h] = i < len ? dy[beg + i] : 0u;
                }
                wave_rows += d[0] + d[1];
                u32 maxd = max(d[0], d[1]);
#pragma unroll
                for (int o = WAVE / 2; o > 0; o >>= 1) maxd = max(maxd, (u32)__shfl_xor((int)maxd, o, WAVE));
                if (maxd <= (u32)TR_SMALL) {
                    // lane per probe: the matches of both probes gathered at once, then inserted
                    u32 zs[2][TR_SMALL];
#pragma unroll
                    for (int h = 0; h < 2; h++)
#pragma unroll
                        for (int i = 0; i < TR_SMALL; i++) zs[h][i] = (u32)i < d[h] ? csr_dst[st[h] + i] : 0u;
#pragma unroll
                    for (int h = 0; h < 2; h++) {
#pragma unroll
                        for (int i0 = 0; i0 < TR_SMALL; i0 += C::STEP) {
                            if ((u32)i0 < maxd) {   // uniform
                                if (s.nused > (u32)C::FLUSH_SMALL)   // uniform: every lane reads the same LDS word
                                    cnt += tr_flush(s, F, rowbase, clo, ny + out0 + cnt);
#pragma unroll
                                for (int i = i0; i < i0 + C::STEP; i++)
                                    if ((u32)i < d[h]) tr_insert(s, zs[h][i] - clo);
                                wave_sync();
                            }
                        }
                    }
                    continue;
                }
                // load-balanced rows, one half-chunk of 64 probes at a time
#pragma unroll 1
                for (int h = 0; h < 2; h++) {
                    u64 R;
                    u64 ex = wave_excl_scan(d[h], &R);
                    wave_sync();
                    s.p_off[lane] = ex;
                    s.p_start[lane] = st[h];
                    if (lane == 0) s.p_off[WAVE] = R;
                    wave_sync();
                    for (u64 rb = 0; rb < R; rb += TR_BATCH) {
                        if (s.nused > (u32)C::FLUSH)   // uniform: every lane reads the same LDS word
                            cnt += tr_flush(s, F, rowbase, clo, ny + out0 + cnt);
                        const u64 re = R < rb + TR_BATCH ? R : rb + TR_BATCH;
                        u64 ci[TR_U];
#pragma unroll
                        for (int q = 0; q < TR_U; q++) {   // locate TR_U rows per lane ...
                            u64 r = rb + (u64)q * WAVE + lane;
                            ci[q] = ~0ull;
                            if (r < re) {
                                u32 a = 0, b = WAVE;   // last probe with p_off <= r
                                while (b - a > 1) {
                                    u32 m = (a + b) >> 1;
                                    if (s.p_off[m] <= r) a = m; else b = m;
                                }
                                ci[q] = s.p_start[a] + (r - s.p_off[a]);
                            }
                        }
                        u32 zq[TR_U];
#pragma unroll
                        for (int q = 0; q < TR_U; q++) zq[q] = ci[q] != ~0ull ? csr_dst[ci[q]] : 0u;   // ... gather
#pragma unroll
                        for (int q = 0; q < TR_U; q++)   // ... insert
                            if (ci[q] != ~0ull) tr_insert(s, zq[q] - clo);
                        wave_sync();
                    }
                }
            }
            cnt += tr_flush(s, F, rowbase, clo, ny + out0 + cnt);
            if (dmax) {
                rows_next = lane == 0 ? min(cnt * dmax, wcap) : 0;
            } else {
                // the next round's join rows of this segment: Σ deg over its new ends
                wave_sync();
                for (u64 i0 = 0; i0 < cnt; i0 += TR_U * WAVE) {
                    u32 zz[TR_U];
#pragma unroll
                    for (int q = 0; q < TR_U; q++) {
                        u64 i = i0 + (u64)q * WAVE + lane;
                        zz[q] = i < cnt ? ny[out0 + i] : TR_EMPTY;
                    }
#pragma unroll
                    for (int q = 0; q < TR_U; q++)
                        if (zz[q] != TR_EMPTY) rows_next += deg[zz[q] - lo];
                }
            }
        }
#pragma unroll
        for (int o = WAVE / 2; o > 0; o >>= 1) rows_next += __shfl_xor(rows_next, o, WAVE);
        if (lane == 0) {
            nseg_len[sg] = (u32)cnt;
            next_rows[sg] = rows_next;
        }
        wave_new += cnt;
    }
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) wave_rows += __shfl_xor(wave_rows, o, WAVE);
    if (lane == 0 && wave_new) atomicAdd(counters, (unsigned long long)wave_new);
    if (lane == 0 && wave_rows) atomicAdd(counters + 2, (unsigned long long)wave_rows);
}

// Σ deg over each segment's ends (the first round's join rows per segment).
__global__ void tc_rows_segrows_k(const u32 *dy, const u64 *seg_beg, const u32 *seg_len, u64 nseg,
                                  const u64 *csr_off, u32 lo, u64 *rows) {
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x; j < nseg; j += stride) {
        u64 s = 0;
        for (u32 i = 0; i < seg_len[j]; i++) {
            u32 y = dy[seg_beg[j] + i] - lo;
            s += csr_off[y + 1] - csr_off[y];
        }
        rows[j] = s;
    }
}

// ---------------------------------------------------------------- set-up
__global__ void csr_hist_k(const u64 *rows, u64 n, u32 lo, u32 *cnt) {
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        atomicAdd(cnt + (tkey(rows[i]) - lo), 1u);
}
__global__ void csr_dst_k(const u64 *rows, u64 n, u32 *dst) {
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = tval(rows[i]);
}
// Δ0 from the unique sorted edges (src, dst): ends = dst, one segment per src
// run (heads = run starts), and the edges' bits (row src, column dst) in F.
__global__ void tc_rows_init_k(const u64 *uniq, u64 n, const u32 *heads, u64 nh, u32 *dy, u32 *seg_x,
                               u64 *seg_beg, u32 *seg_len, u32 *F, const u64 *row_base, const u32 *col_lo,
                               u32 lo) {
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        u64 t = uniq[i];
        u32 sx = tkey(t) - lo, dz = tval(t) - col_lo[sx];
        dy[i] = tval(t);
        atomicOr(F + row_base[sx] + (dz >> 5), 1u << (dz & 31));
    }
    for (u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x; j < nh; j += stride) {
        u64 b = heads[j], e = j + 1 < nh ? heads[j + 1] : n;
        seg_x[j] = tkey(uniq[b]);
        seg_beg[j] = b;
        seg_len[j] = (u32)(e - b);
    }
}
// Segments back to (z, x) tuples, compacted (the ends have gaps between
// segments), for a round hook.
__global__ void tc_rows_tuples_k(const u32 *dy, const u32 *seg_x, const u64 *seg_beg, const u32 *seg_len,
                                 const u64 *pos, u64 nseg, u64 *out) {
    for (u64 j = blockIdx.x; j < nseg; j += gridDim.x)
        for (u32 i = threadIdx.x; i < seg_len[j]; i += blockDim.x)
            out[pos[j] + i] = tmake(dy[seg_beg[j] + i], seg_x[j]);
}
__global__ void u32_to_u64_k(const u32 *in, u64 n, u64 *out) {
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = in[i];
}

void tc_rows_setup(dlg_ctx *ctx, const Index &ix, RowMap &F, TcRows &tr) {
    const u64 W = F.W, n = ix.n;
    // CSR of E over the id range (duplicates kept, as the hash index keeps them)
    tr.deg.ensure(W + 1);
    fill_u32(ctx, tr.deg.p, W + 1, 0u);
    if (n)
        DLG_LAUNCH(ctx, DLG_K_MISC, 12.0 * n, csr_hist_k, sgrid(ctx, n), ctx->block_size, 0,
                   (const u64 *)ix.rows.p, n, F.lo, tr.deg.p);
    tr.csr_off.ensure(W + 1);
    scan_u32_u64(ctx, tr.deg.p, W + 1, tr.csr_off.p);
    // the largest degree: small enough, and the next rounds' output regions
    // are bounded by it instead of by exact degree sums
    {
        std::vector<u32> hd(W);
        HIPCHK(hipMemcpyAsync(hd.data(), tr.deg.p, sizeof(u32) * W, hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
        u32 m = 0;
        for (u32 v : hd) m = std::max(m, v);
        tr.dmax = m <= TR_DMAX_BOUND ? std::max<u32>(m, 1) : 0;
    }
    tr.csr_dst.ensure(n ? n : 1);
    if (n)
        DLG_LAUNCH(ctx, DLG_K_MISC, 12.0 * n, csr_dst_k, sgrid(ctx, n), ctx->block_size, 0,
                   (const u64 *)ix.rows.p, n, tr.csr_dst.p);
    // Δ0 = unique(E reversed), grouped by source; one segment per source, for good
    DBuf<u64> uniq(ctx, n ? n : 1);
    u64 nu = n ? select(ctx, SEL_UNIQUE, ix.rows.p, n, nullptr, 0, uniq.p) : 0;
    DBuf<u32> heads(ctx, nu ? nu : 1);
    u64 nh = nu ? select(ctx, SEL_KEY_HEAD, uniq.p, nu, nullptr, 0, (u64 *)heads.p) : 0;
    tr.dy.ensure(nu ? nu : 1);
    tr.seg_x.ensure(nh ? nh : 1);
    tr.seg_beg.ensure(nh ? nh : 1);
    tr.seg_len.ensure(nh ? nh : 1);
    tr.nseg_len.ensure(nh ? nh : 1);
    tr.seg_out.ensure(nh ? nh : 1);
    tr.next_rows.ensure(nh ? nh : 1);
    tr.counters.ensure(4);
    tr.nd = nu;
    tr.nseg = nh;
    tr.rows = 0;
    tr.wcap = 0;
    for (const auto &b : F.blocks) tr.wcap = std::max<u64>(tr.wcap, b.W);
    if (!nu) return;
    DLG_LAUNCH(ctx, DLG_K_BITMAP_INSERT, 20.0 * nu, tc_rows_init_k, sgrid(ctx, nu), ctx->block_size, 0,
               (const u64 *)uniq.p, nu, (const u32 *)heads.p, nh, tr.dy.p, tr.seg_x.p, tr.seg_beg.p,
               tr.seg_len.p, F.words.p, (const u64 *)F.row_base.p, (const u32 *)F.col_lo.p, F.lo);
    DLG_LAUNCH(ctx, DLG_K_MISC, 12.0 * nu, tc_rows_segrows_k, sgrid(ctx, nh), ctx->block_size, 0,
               (const u32 *)tr.dy.p, (const u64 *)tr.seg_beg.p, (const u32 *)tr.seg_len.p, nh,
               (const u64 *)tr.csr_off.p, F.lo, tr.seg_out.p);
    scan_u64_async(ctx, tr.seg_out.p, nh, tr.seg_out.p, tr.counters.p + 1);
    u64 h[2];
    ctx->read_scalars(tr.counters.p, 2, h);
    tr.rows = h[1];
}

constexpr u64 TR_LONG_SEGMENTS = 64;   // average probes per segment above which the big table wins

template <int LOG>
static void launch_rows(dlg_ctx *ctx, RowMap &F, TcRows &tr) {
    unsigned g = (unsigned)std::min<u64>((tr.nseg + TR_W - 1) / TR_W, (u64)ctx->num_cus * TrCfg<LOG>::BLOCKS_PER_CU);
    DLG_LAUNCH(ctx, DLG_K_BITMAP_INSERT, 0.0, tc_rows_k<LOG>, g, TR_T, 0, (const u32 *)tr.dy.p, (const u32 *)tr.seg_x.p,
               (const u64 *)tr.seg_beg.p, (const u32 *)tr.seg_len.p, (const u64 *)tr.seg_out.p, tr.nseg,
               (const u64 *)tr.csr_off.p, (const u32 *)tr.csr_dst.p, (const u32 *)tr.deg.p, tr.dmax, F.words.p,
               (const u64 *)F.row_base.p, (const u32 *)F.col_lo.p, F.lo, tr.wcap, tr.ny.p, tr.nseg_len.p,
               tr.next_rows.p, (unsigned long long *)tr.counters.p);
}

u64 tc_rows_round(dlg_ctx *ctx, RowMap &F, TcRows &tr, u64 *raw) {
    if (tr.nd == 0 || tr.rows == 0) {
        tr.nd = 0;
        return 0;
    }
    tr.ny.ensure(tr.rows);
    fill_u64(ctx, tr.counters.p, 3, 0ull);
    // algorithmic bytes, added once the round's join rows are known: 4 B per
    // join row (its matched end) + 20 B per probe (its end, CSR offset, degree)
    const u64 nd0 = tr.nd;
    if (tr.nd >= TR_LONG_SEGMENTS * tr.nseg) launch_rows<10>(ctx, F, tr);   // long segments: big tables
    else launch_rows<9>(ctx, F, tr);                                        // short ones: more waves
    // the new ends sit at the old output offsets; the next offsets scan the next rows
    tr.seg_beg.swap(tr.seg_out);
    scan_u64_async(ctx, tr.next_rows.p, tr.nseg, tr.seg_out.p, tr.counters.p + 1);
    u64 h[3];
    ctx->read_scalars(tr.counters.p, 3, h);
    tr.dy.swap(tr.ny);
    tr.seg_len.swap(tr.nseg_len);
    tr.nd = h[0];
    tr.rows = h[1];
    if (raw) *raw += h[2];
    ctx->prof_add_bytes(DLG_K_BITMAP_INSERT, 4.0 * h[2] + 20.0 * nd0);
    return h[0];
}

u64 rowmap_bits(u32 lo, u32 hi, const std::vector<u32> *starts) {
    if (!starts) {
        u64 W = (u64)(hi - lo) + 1;
        return W * ((W + 31) / 32) * 32;
    }
    u64 bits = 0;
    for (size_t i = 0; i + 1 < starts->size(); i++) {
        u64 e = (*starts)[i + 1] == 0xffffffffu && i + 2 == starts->size() ? 0x100000000ull : (*starts)[i + 1];
        u64 b0 = std::max<u64>((*starts)[i], lo), b1 = std::min<u64>(e, (u64)hi + 1);
        if (b1 > b0) bits += (b1 - b0) * ((b1 - b0 + 31) / 32) * 32;
    }
    return bits;
}

void rowmap_init(dlg_ctx *ctx, RowMap &m, u32 lo, u32 hi, const std::vector<u32> *starts) {
    m.lo = lo;
    m.W = (u64)(hi - lo) + 1;
    m.blocks.clear();
    std::vector<u64> st;
    if (starts) st.assign(starts->begin(), starts->end());
    else st = {lo, (u64)hi + 1};
    if (st.back() == 0xffffffffull) st.back() = 0x100000000ull;   // a range that ends at the last id
    // block i covers the ids of [st[i], st[i+1]) that fall in [lo, hi]
    std::vector<u64> hb(m.W);
    std::vector<u32> hl(m.W);
    u64 base = 0;
    for (size_t i = 0; i + 1 < st.size(); i++) {
        u64 b0 = std::max<u64>(st[i], lo), b1 = std::min<u64>(st[i + 1], (u64)hi + 1);
        if (b1 <= b0) continue;
        RowMap::Block b{(u32)b0, b1 - b0, (b1 - b0 + 31) / 32, base};
        for (u64 r = 0; r < b.W; r++) {
            hb[b0 - lo + r] = base + r * b.wpr;
            hl[b0 - lo + r] = b.lo;
        }
        base += b.W * b.wpr;
        m.blocks.push_back(b);
    }
    m.nwords = base;
    m.words.ensure(m.nwords ? m.nwords : 1);
    fill_u32(ctx, m.words.p, m.nwords ? m.nwords : 1, 0u);
    m.row_base.ensure(m.W);
    m.col_lo.ensure(m.W);
    HIPCHK(hipMemcpyAsync(m.row_base.p, hb.data(), 8 * m.W, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(m.col_lo.p, hl.data(), 4 * m.W, hipMemcpyHostToDevice, ctx->stream));
    ctx->sync();   // the host arrays go out of scope
}

u64 rowmap_emit(dlg_ctx *ctx, RowMap &m, DBuf<u64> &out, u64 expect) {
    out.ensure(expect ? expect : 1);
    DBuf<u32> scratch(ctx);
    u64 pos = 0;
    for (const auto &b : m.blocks)
        pos += bitmap_emit_block(ctx, m.words.p + b.base, b.lo, b.W, b.wpr, true, scratch, out.p + pos);
    if (pos != expect) fail(DLG_DEVICE_ERROR, "bitmap emission count mismatch");
    return pos;
}

void tc_rows_tuples(dlg_ctx *ctx, const TcRows &tr, DBuf<u64> &out) {
    out.ensure(tr.nd ? tr.nd : 1);
    if (!tr.nd) return;
    DBuf<u64> pos(ctx, tr.nseg);
    DBuf<u64> tot(ctx, 1);
    DBuf<u64> lens(ctx, tr.nseg);
    DLG_LAUNCH(ctx, DLG_K_MISC, 12.0 * tr.nseg, u32_to_u64_k, sgrid(ctx, tr.nseg), ctx->block_size, 0,
               (const u32 *)tr.seg_len.p, tr.nseg, lens.p);
    scan_u64_async(ctx, lens.p, tr.nseg, pos.p, tot.p);
    unsigned g = (unsigned)std::min<u64>(tr.nseg, (u64)ctx->num_cus * 8);
    DLG_LAUNCH(ctx, DLG_K_MISC, 16.0 * tr.nd, tc_rows_tuples_k, g, 256, 0, (const u32 *)tr.dy.p,
               (const u32 *)tr.seg_x.p, (const u64 *)tr.seg_beg.p, (const u32 *)tr.seg_len.p,
               (const u64 *)pos.p, tr.nseg, out.p);
}

}  // namespace dlg
