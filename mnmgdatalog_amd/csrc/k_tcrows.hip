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
#include <deque>
#include <type_traits>
#include <vector>

namespace dlg {

constexpr int TR_T = 256;                    // threads per workgroup: four waves, one segment each
constexpr int TR_W = TR_T / WAVE;
constexpr u32 TR_EMPTY = 0xffffffffu;
constexpr int TR_SMALL = 8;   // chunks whose probes all have at most this many matches: lane per probe
constexpr int TR_ELL_K = 8;   // degrees up to this: the join reads an ELL adjacency (one 32-byte row per id)
constexpr int TR_U = 4;       // words / rows per lane with their loads in flight together
constexpr int TR_BATCH = TR_U * WAVE;        // rows per batch of the load-balanced path
constexpr u32 TR_DMAX_BOUND = 16;   // degrees up to this: output regions bounded by new ends * max degree
constexpr u64 TR_ENTRY_COST = 512;  // an entry's fixed cost (loads, final flush) in join rows
constexpr int TR_PART = 6;          // totals per workgroup of the round kernel
// The round state in device memory (TcRows::counters), left by each round's
// compaction for the next round, so that rounds can be queued ahead of the
// host: the round's totals (new ends = the next Δ, Σ bounds = the next output
// span, join rows, the next list's length, words of F read and written),
// then whether the last row kernel ran, and the stall flag (an output span
// above the output buffer: the round did nothing; the host grows the buffer).
enum { TR_ST_ND = 0, TR_ST_SPAN = 1, TR_ST_RAW = 2, TR_ST_NACT = 3, TR_ST_RUN = TR_PART, TR_ST_STALL, TR_ST_WORDS };
// A round's mapped slot: the totals, then its status, then its number.
enum : u64 { TR_RAN = 1, TR_IDLE = 2, TR_STALLED = 3 };
constexpr int TR_QUEUE = 4;   // rounds queued ahead of the host by default
static_assert(TR_PART + 1 <= dlg_ctx::MAPPED_SEQ, "the round's totals and status precede its number in a slot");
// every queued round has a mapped slot of its own (a depth asked for is capped too)
static_assert(TR_QUEUE <= dlg_ctx::MAPPED_SLOTS, "more queued rounds than mapped slots");

// The word table of a wave: 2^LOG slots, linear probing.  A larger table
// flushes a long segment fewer times; a smaller one leaves LDS for more waves
// per CU (short segments).  A step inserts at most NI (lane-per-probe path)
// or TR_U (load-balanced path) entries per lane; the table is flushed before
// a step's entries could take it past LIMIT occupied slots.
template <int LOG>
struct TrCfg {
    static constexpr int H = 1 << LOG;
    static constexpr int NI = 4;
    static constexpr u32 LIMIT = H * 3 / 4;
    static_assert(NI * WAVE <= (int)LIMIT && TR_BATCH <= (int)LIMIT && TR_SMALL % NI == 0 && TR_ELL_K % NI == 0,
                  "table load");
};

template <int LOG>
__device__ __forceinline__ u32 tr_hash(u32 w) { return (w * 0x9e3779b1u) >> (32 - LOG); }

// Memory written by some lanes of a wave and then read by others (the slot
// list; F's words across two flushes of one segment): a workgroup-scope
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
    u16 used[TrCfg<LOG>::LIMIT];   // occupied slots, in insertion order (a flush keeps them <= LIMIT)
    u32 touched, changed;          // words of F the wave's flushes read, and wrote
    u32 pad[WAVE];                 // each lane's dummy CAS target (never TR_EMPTY)
};
// The CSR join's load-balanced path also keeps the chunk's probes.
template <int LOG>
struct TrWaveLB : TrWave<LOG> {
    u64 p_off[WAVE + 1];   // row offsets of the chunk's probes (p_off[WAVE] = total)
    u64 p_start[WAVE];     // CSR start of each probe's adjacency
};

// Workgroups per CU: four, except for the CSR variant with 1024-slot tables,
// whose LDS (four 10.5 KB waves) allows three.  At four workgroups the ELL
// variants are held to 128 registers and spill 24 bytes per lane; with
// 1024-slot tables that is still 4.5 % faster on the 400 x 400 grid than
// three workgroups without spills.
template <int LOG, bool ELL>
constexpr int tr_blocks_per_cu() { return LOG >= 10 && !ELL ? 3 : 4; }

// OR the lane's entries (word w[i], bits mk[i]; mk[i] = 0: no entry) into
// the wave's table, one entry at a time (a CAS claims or finds the word's
// slot, an OR with no result sets the bits), and append the new slots to
// `used`; nused is the wave's (uniform) count.
template <int LOG, int N>
__device__ __forceinline__ void tr_insert(TrWave<LOG> &s, u32 &nused, const u32 (&w)[N], const u32 (&mk)[N]) {
    const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
    for (int i = 0; i < N; i++) {
        const u64 any = __ballot(mk[i] != 0);
        if (!any) continue;   // uniform
        // Linear probing without branches in the loop body: every lane issues
        // the CAS, lanes with no entry (or already placed) on their own pad
        // word, which never holds TR_EMPTY; a lane whose slot holds another
        // word moves on; the loop ends when no lane does.
        u32 h = tr_hash<LOG>(w[i]);
        bool act = mk[i] != 0, fresh = false;
        do {
            const u32 k = atomicCAS(act ? &s.key[h] : &s.pad[lane], TR_EMPTY, w[i]);
            fresh = fresh || (act && k == TR_EMPTY);
            act = act && k != TR_EMPTY && k != w[i];
            h = act ? ((h + 1) & (TrCfg<LOG>::H - 1)) : h;
        } while (__ballot(act));
        atomicOr(&s.bits[h], mk[i]);   // h is a slot for every lane; no entry ORs 0
        const u64 bal = __ballot(fresh);
        if (fresh)
            s.used[nused + __builtin_amdgcn_mbcnt_hi((u32)(bal >> 32), __builtin_amdgcn_mbcnt_lo((u32)bal, 0u))] =
                (u16)h;
        nused += (u32)__popcll(bal);
    }
}

// Flush the wave's table into row `rowbase` of F: per occupied slot a read
// and, for new bits, a write of the word (TR_U words per lane in flight);
// the new bits are written as ends z (packed) to dst[0..).  Returns the
// number written.  The table is left empty.
template <int LOG>
__device__ u64 tr_flush(TrWave<LOG> &s, u32 &nused, u32 *F, u64 rowbase, u32 lo, u32 *dst) {
    const int lane = threadIdx.x & (WAVE - 1);
    wave_sync();
    const u32 n = nused;
    u64 cnt = 0;
    for (u32 j0 = 0; j0 < n; j0 += TR_U * WAVE) {
        u32 slot[TR_U], w[TR_U], b[TR_U], nb[TR_U];
#pragma unroll
        for (int q = 0; q < TR_U; q++) {
            const u32 j = j0 + q * WAVE + lane;
            slot[q] = j < n ? s.used[j] : TR_EMPTY;
        }
#pragma unroll
        for (int q = 0; q < TR_U; q++) {
            w[q] = slot[q] != TR_EMPTY ? s.key[slot[q]] : 0u;
            b[q] = slot[q] != TR_EMPTY ? s.bits[slot[q]] : 0u;
        }
#pragma unroll
        for (int q = 0; q < TR_U; q++) nb[q] = b[q] ? F[rowbase + w[q]] : 0u;   // the row's owner:
        u32 c = 0;                                                           // plain read-modify-write
#pragma unroll
        for (int q = 0; q < TR_U; q++) {
            const u32 cur = nb[q];
            nb[q] = b[q] & ~cur;
            if (nb[q]) F[rowbase + w[q]] = cur | nb[q];
            if (slot[q] != TR_EMPTY) {
                s.key[slot[q]] = TR_EMPTY;
                s.bits[slot[q]] = 0u;
            }
            c += (u32)__popc(nb[q]);
        }
        u32 changed = 0;
#pragma unroll
        for (int q = 0; q < TR_U; q++) changed += (u32)__popcll(__ballot(nb[q] != 0u));
        if (lane == 0) s.changed += changed;
        u32 tot;
        u32 off = wave_excl_scan32(c, &tot);
        u32 *d = dst + cnt;
#pragma unroll
        for (int q = 0; q < TR_U; q++) {
            for (u32 v = nb[q]; v; v &= v - 1) d[off++] = lo + w[q] * 32 + (u32)(__ffs(v) - 1);
        }
        cnt += tot;
    }
    if (lane == 0) s.touched += n;
    wave_sync();
    nused = 0;
    return cnt;
}

// One step: flush first if the step's entries could take the table past
// LIMIT, then insert them.  Returns the ends the flush wrote to dst.
template <int LOG, int N>
__device__ __forceinline__ u64 tr_step(TrWave<LOG> &s, u32 &nused, const u32 (&w)[N], const u32 (&mk)[N], u32 *F,
                                       u64 rowbase, u32 lo, u32 *dst) {
    u32 m = 0;
#pragma unroll
    for (int i = 0; i < N; i++) m += (u32)__popcll(__ballot(mk[i] != 0));
    u64 c = 0;
    if (nused + m > TrCfg<LOG>::LIMIT) c = tr_flush(s, nused, F, rowbase, lo, dst);
    tr_insert(s, nused, w, mk);
    return c;
}

// The lane-per-probe join over an ELL adjacency (every degree <= TR_ELL_K):
// a probe's matches are one 32-byte row of ids relative to lo (TR_EMPTY past
// its degree).  The rows are loaded one chunk ahead and the chunks' ends two
// chunks ahead, so the global loads of a chunk overlap the previous chunk's
// inserts.  Loads are never predicated (a lane past the end reads the last
// end again) and their results are not touched before the next chunk, so no
// wait for them is placed in this chunk; validity travels as flags.
// Returns the ends flushed to dst[0 ..); rows += the join rows.
template <int LOG>
__device__ __forceinline__ u64 tr_ell_rows(TrWave<LOG> &s, u32 &nused, const u32 *dy, u64 beg, u32 len,
                                           const uint4 *ell, u32 lo, u32 *F, u64 rowbase, u32 clo, u32 *dst,
                                           u64 &rows) {
    using C = TrCfg<LOG>;
    const int lane = threadIdx.x & (WAVE - 1);
    const u32 shift = lo - clo;   // ELL ids are relative to lo, columns to the row's block
    const u64 last = beg + len - 1;
    u64 cnt = 0;
    u32 yr[2];        // ends of the next chunk (raw ids)
    bool yv[2], ev[2];
    uint4 ea[2][2];   // rows of this chunk
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const u64 i = beg + h * WAVE + lane;
        ev[h] = i <= last;
        yr[h] = dy[i <= last ? i : last];
    }
#pragma unroll
    for (int h = 0; h < 2; h++) {
        ea[h][0] = ell[2 * (u64)(yr[h] - lo)];
        ea[h][1] = ell[2 * (u64)(yr[h] - lo) + 1];
    }
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const u64 i = beg + 2 * WAVE + h * WAVE + lane;
        yv[h] = i <= last;
        yr[h] = dy[i <= last ? i : last];
    }
    for (u32 pc = 0; pc < len; pc += 2 * WAVE) {
        uint4 eb[2][2];   // the next chunk's rows, in flight while this chunk inserts
        bool evn[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            eb[h][0] = ell[2 * (u64)(yr[h] - lo)];
            eb[h][1] = ell[2 * (u64)(yr[h] - lo) + 1];
            evn[h] = yv[h];
        }
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const u64 i = beg + pc + 4 * WAVE + h * WAVE + lane;
            yv[h] = i <= last;
            yr[h] = dy[i <= last ? i : last];
        }
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const u32 z[TR_ELL_K] = {ea[h][0].x, ea[h][0].y, ea[h][0].z, ea[h][0].w,
                                     ea[h][1].x, ea[h][1].y, ea[h][1].z, ea[h][1].w};
            u32 w[TR_ELL_K], mk[TR_ELL_K];
#pragma unroll
            for (int i = 0; i < TR_ELL_K; i++) {
                const bool ok = ev[h] && z[i] != TR_EMPTY;
                const u32 c = z[i] + shift;
                w[i] = c >> 5;
                mk[i] = ok ? 1u << (c & 31) : 0u;
                rows += ok;
            }
            // a probe's matches are sorted: runs in one word merge into one entry
#pragma unroll
            for (int i = TR_ELL_K - 1; i > 0; i--)
                if (mk[i] && mk[i - 1] && w[i] == w[i - 1]) {
                    mk[i - 1] |= mk[i];
                    mk[i] = 0u;
                }
#pragma unroll
            for (int g0 = 0; g0 < TR_ELL_K; g0 += C::NI) {
                u32 gw[C::NI], gm[C::NI];
#pragma unroll
                for (int i = 0; i < C::NI; i++) {
                    gw[i] = w[g0 + i];
                    gm[i] = mk[g0 + i];
                }
                cnt += tr_step(s, nused, gw, gm, F, rowbase, clo, dst + cnt);
            }
        }
#pragma unroll
        for (int h = 0; h < 2; h++) {
            ea[h][0] = eb[h][0];
            ea[h][1] = eb[h][1];
            ev[h] = evn[h];
        }
    }
    return cnt;
}

// The cost share a workgroup takes: workgroups go to the XCDs in turn
// (workgroup b to XCD b % 8), so XCD x gets the x-th eighth of the shares --
// a contiguous range of the active list, whose sources (and, in a locality
// order, their ends and those ends' adjacency rows) sit together in that
// XCD's L2.  A bijection of [0, G); the totals and ranges are kept by share.
constexpr u32 TR_XCD = 8;
__device__ __forceinline__ u32 tr_share(u32 b, u32 G) {
    const u32 x = b % TR_XCD, q = G / TR_XCD, r = G % TR_XCD;
    return x * q + (x < r ? x : r) + b / TR_XCD;
}

// One round over the active list (the sources with ends): entry i (source
// ax[i], ends dy[abeg[i] ..][0, alen[i])) writes its new ends at
// ny[aoff[i] ..] -- aoff is the exclusive scan of the entries' output
// bounds, span its total -- and its new end count and next bound at rcnt[i],
// rbound[i].  The bound is the source's next join rows, Σ deg over its new
// ends (exact; `dmax` = 0), or, when every degree is at most dmax,
// min(new ends * dmax, W) (no degree lookups; the output has gaps).
// Workgroup b leaves its totals at bpart[TR_PART b ..] (entries with new
// ends, their bounds, new ends, join rows, words of F read, words of F
// written) and its first entry at brange[b].  An
// entry costs about its bound (join rows) plus a fixed TR_ENTRY_COST, so
// with cost prefix c(i) = aoff[i] + TR_ENTRY_COST * i, workgroup b takes the
// entries with c(i) in [b, b + 1) * c(nact) / grid (a binary search; no
// global counter), and its four waves share them through an LDS counter.
template <int LOG, bool ELL>
__global__ __launch_bounds__(TR_T, (tr_blocks_per_cu<LOG, ELL>())) void tc_rows_k(
    const u32 *dy, const u32 *ax, const u64 *abeg, const u32 *alen, const u64 *aoff, u64 *st, u64 cap,
    const u64 *csr_off, const u32 *csr_dst, const u32 *deg, const uint4 *ell, u32 dmax, u32 *F, const u64 *row_base,
    const u32 *col_lo, u32 lo, u64 wcap, u32 *ny, u32 *rcnt, u64 *rbound, u64 *bpart, u64 *brange) {
    using C = TrCfg<LOG>;
    // The round's gate, read from the state the previous round left (every
    // block reads the same): no ends, no join rows or no entries -- nothing
    // to do; an output span above the output buffer's capacity -- nothing
    // done, and the stall flag set for the host.  Block 0 records whether the
    // round ran, for the compaction.
    const u64 nact = st[TR_ST_NACT], span = st[TR_ST_SPAN];
    const bool idle = st[TR_ST_STALL] != 0 || st[TR_ST_ND] == 0 || span == 0 || nact == 0;
    const bool stall = !idle && span > cap;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st[TR_ST_RUN] = idle || stall ? 0 : 1;
        if (stall) st[TR_ST_STALL] = 1;
    }
    if (idle || stall) return;   // block-uniform
    using SW = std::conditional_t<ELL, TrWave<LOG>, TrWaveLB<LOG>>;
    __shared__ SW sw[TR_W];
    __shared__ u64 s_rng[2];
    __shared__ u32 s_next;
    const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
    SW &sl = sw[wv];
    TrWave<LOG> &s = sl;
    for (int i = lane; i < C::H; i += WAVE) {
        s.key[i] = TR_EMPTY;
        s.bits[i] = 0u;
    }
    if (lane == 0) s.touched = s.changed = 0u;
    s.pad[lane] = 0u;
    if (threadIdx.x < WAVE) {   // the first wave finds this share's first entry and the next one's
        const u64 tot = span + TR_ENTRY_COST * nact, sh = tr_share(blockIdx.x, gridDim.x);
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const u64 b = sh + (u64)e;
            const u64 key = (b * tot + gridDim.x - 1) / gridDim.x;
            const u64 l = wave_partition_index(0, nact, [aoff, key](u64 m) { return aoff[m] + TR_ENTRY_COST * m >= key; });
            if (threadIdx.x == 0) s_rng[e] = b >= gridDim.x ? nact : l;
        }
    }
    if (threadIdx.x == 0) s_next = 0;
    __syncthreads();
    const u64 r0 = s_rng[0], r1 = s_rng[1];
    u32 nused = 0;   // the wave's occupied slots (uniform)
    u64 wave_new = 0, wave_rows = 0, wave_act = 0, wave_bound = 0;
    while (true) {
        u32 k0 = 0;
        if (lane == 0) k0 = atomicAdd(&s_next, 1u);
        const u64 ai = r0 + __shfl(k0, 0, WAVE);
        if (ai >= r1) break;
        {
            const u32 len = alen[ai];
            u64 rows_next = 0, cnt = 0;
            if (len) {
                const u32 x = ax[ai];
                const u64 beg = abeg[ai], out0 = aoff[ai];
                const u64 rowbase = row_base[x - lo];   // x's row and its block's first id
                const u32 clo = col_lo[x - lo];
                if constexpr (ELL) {
                    cnt += tr_ell_rows(s, nused, dy, beg, len, ell, lo, F, rowbase, clo, ny + out0, wave_rows);
                } else {
                    // the chunks' ends are loaded one chunk ahead (software pipelining;
                    // unpredicated: a lane past the end reads the last end again)
                    const u64 last = beg + len - 1;
                    u32 yn[2];
#pragma unroll
                    for (int h = 0; h < 2; h++) yn[h] = dy[min(beg + h * WAVE + lane, last)];
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
                        for (int h = 0; h < 2; h++) yn[h] = dy[min(beg + pc + 2 * WAVE + h * WAVE + lane, last)];
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
                                for (int g0 = 0; g0 < TR_SMALL; g0 += C::NI) {
                                    if ((u32)g0 < maxd) {   // uniform
                                        u32 w[C::NI], mk[C::NI];
#pragma unroll
                                        for (int i = 0; i < C::NI; i++) {
                                            const u32 z = zs[h][g0 + i] - clo;
                                            w[i] = z >> 5;
                                            mk[i] = (u32)(g0 + i) < d[h] ? 1u << (z & 31) : 0u;
                                        }
                                        // a probe's matches are sorted: runs in one word merge into one entry
#pragma unroll
                                        for (int i = C::NI - 1; i > 0; i--)
                                            if (mk[i] && mk[i - 1] && w[i] == w[i - 1]) {
                                                mk[i - 1] |= mk[i];
                                                mk[i] = 0u;
                                            }
                                        cnt += tr_step(s, nused, w, mk, F, rowbase, clo, ny + out0 + cnt);
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
                            sl.p_off[lane] = ex;
                            sl.p_start[lane] = st[h];
                            if (lane == 0) sl.p_off[WAVE] = R;
                            wave_sync();
                            for (u64 rb = 0; rb < R; rb += TR_BATCH) {
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
                                            if (sl.p_off[m] <= r) a = m; else b = m;
                                        }
                                        ci[q] = sl.p_start[a] + (r - sl.p_off[a]);
                                    }
                                }
                                u32 w[TR_U], mk[TR_U];
#pragma unroll
                                for (int q = 0; q < TR_U; q++) {   // ... gather ...
                                    const u32 z = (ci[q] != ~0ull ? csr_dst[ci[q]] : clo) - clo;
                                    w[q] = z >> 5;
                                    mk[q] = ci[q] != ~0ull ? 1u << (z & 31) : 0u;
                                }
                                cnt += tr_step(s, nused, w, mk, F, rowbase, clo, ny + out0 + cnt);   // ... insert
                            }
                        }
                    }
                }
                cnt += tr_flush(s, nused, F, rowbase, clo, ny + out0 + cnt);
                if (dmax) {
                    rows_next = lane == 0 ? min(cnt * dmax, wcap) : 0;
                } else {
                    // the next round's join rows of this segment: Σ deg over its new ends
                    wave_sync();
                    for (u64 i0 = 0; i0 < cnt; i0 += TR_U * WAVE) {
                        u32 zz[TR_U];
#pragma unroll
                        for (int q = 0; q < TR_U; q++) {
                            const u64 i = i0 + (u64)q * WAVE + lane;
                            zz[q] = i < cnt ? ny[out0 + i] : lo;
                        }
#pragma unroll
                        for (int q = 0; q < TR_U; q++)   // every id is an end, the range's last one included
                            if (i0 + (u64)q * WAVE + lane < cnt) rows_next += deg[zz[q] - lo];
                    }
                }
            }
#pragma unroll
            for (int o = WAVE / 2; o > 0; o >>= 1) rows_next += __shfl_xor(rows_next, o, WAVE);
            if (lane == 0) {
                rcnt[ai] = (u32)cnt;
                rbound[ai] = rows_next;
            }
            wave_new += cnt;
            wave_act += cnt != 0;
            wave_bound += rows_next;   // zero without new ends
        }
    }
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) wave_rows += __shfl_xor(wave_rows, o, WAVE);
    __shared__ u64 s_tot[TR_PART][TR_W];
    wave_sync();
    if (lane == 0) {
        s_tot[0][wv] = wave_act;
        s_tot[1][wv] = wave_bound;
        s_tot[2][wv] = wave_new;
        s_tot[3][wv] = wave_rows;
        s_tot[4][wv] = s.touched;
        s_tot[5][wv] = s.changed;
    }
    __syncthreads();
    if (threadIdx.x < TR_PART) {   // the workgroup's totals, for the compaction
        u64 t = 0;
        for (int w = 0; w < TR_W; w++) t += s_tot[threadIdx.x][w];
        bpart[TR_PART * (u64)tr_share(blockIdx.x, gridDim.x) + threadIdx.x] = t;
    }
    if (threadIdx.x == 0) {   // and its range of entries
        const u32 sh = tr_share(blockIdx.x, gridDim.x);
        brange[sh] = r0;
        if (sh == gridDim.x - 1) brange[gridDim.x] = nact;
    }
}

// Σ deg over each entry's ends (the first round's join rows per source).
__global__ void tc_rows_segrows_k(const u32 *dy, const u64 *abeg, const u32 *alen, u64 nact, const u64 *csr_off,
                                  u32 lo, u64 *rows) {
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x; j < nact; j += stride) {
        u64 s = 0;
        for (u32 i = 0; i < alen[j]; i++) {
            u32 y = dy[abeg[j] + i] - lo;
            s += csr_off[y + 1] - csr_off[y];
        }
        rows[j] = s;
    }
}

// The next active list: the entries with new ends, in order, each with its
// ends' place (its output offset this round), their count and its next
// output offset (the exclusive scan of the bounds).  One block per block of
// the round's row kernel, over the same entries: it adds the totals of the
// blocks before its own and compacts its entries TA_CHUNK at a time.  The
// last block writes the round's totals (new ends, Σ bounds, join rows, the
// list's length, words of F read and written) into counters and into the
// context's mapped host words.
constexpr int TA_T = 256, TA_I = 8;
constexpr u64 TA_CHUNK = (u64)TA_T * TA_I;

__global__ __launch_bounds__(TA_T) void tc_rows_compact_k(const u32 *ax, const u64 *aoff, const u32 *rcnt,
                                                          const u64 *rbound, const u64 *bpart, const u64 *brange,
                                                          u32 *nx, u64 *nbeg, u32 *nlen, u64 *noff, u64 *counters,
                                                          u64 *hcounters, u64 seq) {
    __shared__ u64 scr[TA_T / WAVE + 1];
    const bool last = blockIdx.x == gridDim.x - 1;
    if (counters[TR_ST_RUN] == 0) {   // the row kernel did nothing: the state stays, the slot says why
        if (last && threadIdx.x == 0) {
            for (int q = 0; q < TR_PART; q++) hcounters[q] = 0;
            hcounters[TR_ST_SPAN] = counters[TR_ST_SPAN];
            hcounters[TR_PART] = counters[TR_ST_STALL] ? TR_STALLED : TR_IDLE;
            __threadfence_system();
            *(volatile u64 *)(hcounters + dlg_ctx::MAPPED_SEQ) = seq;
            __threadfence_system();
        }
        return;
    }
    // the blocks before this one (all of them, for the last block's totals)
    u64 pc = 0, ps = 0, pn = 0, pr = 0, pt = 0, pw = 0;
    for (u32 j = threadIdx.x; j < blockIdx.x; j += TA_T) {   // the two columns every block needs
        pc += bpart[TR_PART * j];
        ps += bpart[TR_PART * j + 1];
    }
    if (last) {   // uniform: the other four only for the round's totals
        for (u32 j = threadIdx.x; j < gridDim.x; j += TA_T) {
            pn += bpart[TR_PART * j + 2];
            pr += bpart[TR_PART * j + 3];
            pt += bpart[TR_PART * j + 4];
            pw += bpart[TR_PART * j + 5];
        }
    }
    u64 ec0, es0, tn = 0, trw = 0, ttw = 0, tcw = 0;
    block_exclusive_scan(pc, scr, &ec0);
    block_exclusive_scan(ps, scr, &es0);
    if (last) {   // uniform
        block_exclusive_scan(pn, scr, &tn);
        block_exclusive_scan(pr, scr, &trw);
        block_exclusive_scan(pt, scr, &ttw);
        block_exclusive_scan(pw, scr, &tcw);
    }
    const u64 e0 = brange[blockIdx.x], e1 = brange[blockIdx.x + 1];
    for (u64 t0 = e0; t0 < e1; t0 += TA_CHUNK) {   // uniform
        const u64 b0 = t0 + (u64)threadIdx.x * TA_I;   // TA_I consecutive entries
        u32 c = 0;
        u64 sb = 0;
#pragma unroll
        for (int k = 0; k < TA_I; k++) {
            const u64 i = b0 + k;
            if (i < e1 && rcnt[i]) {
                c++;
                sb += rbound[i];
            }
        }
        u64 tc, ts;
        u64 ec = block_exclusive_scan((u64)c, scr, &tc) + ec0;
        u64 es = block_exclusive_scan(sb, scr, &ts) + es0;
#pragma unroll
        for (int k = 0; k < TA_I; k++) {
            const u64 i = b0 + k;
            if (i < e1 && rcnt[i]) {
                nx[ec] = ax[i];
                nbeg[ec] = aoff[i];
                nlen[ec] = rcnt[i];
                noff[ec] = es;
                ec++;
                es += rbound[i];
            }
        }
        ec0 += tc;
        es0 += ts;
    }
    if (last && threadIdx.x == 0) {   // the round's totals, also into mapped host memory, then its number
        const u64 v[TR_PART] = {tn, es0, trw, ec0, ttw, tcw};
        for (int q = 0; q < TR_PART; q++) {
            counters[q] = v[q];
            hcounters[q] = v[q];
        }
        hcounters[TR_PART] = TR_RAN;
        __threadfence_system();
        *(volatile u64 *)(hcounters + dlg_ctx::MAPPED_SEQ) = seq;
        __threadfence_system();
    }
}

// ---------------------------------------------------------------- set-up
__global__ void csr_hist_k(const u64 *rows, u64 n, u32 lo, u32 *cnt) {
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        atomicAdd(cnt + (tkey(rows[i]) - lo), 1u);
}
// ELL rows: the matches of id s at ell[s * TR_ELL_K ..], ids relative to lo.
__global__ void ell_fill_k(const u64 *rows, u64 n, const u64 *csr_off, u32 lo, u32 *ell) {
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const u32 s = tkey(rows[i]) - lo;
        ell[(u64)s * TR_ELL_K + (i - csr_off[s])] = tval(rows[i]) - lo;
    }
}
__global__ void csr_dst_k(const u64 *rows, u64 n, u32 *dst) {
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = tval(rows[i]);
}
// Δ0 from the unique sorted edges (src, dst): ends = dst, one active entry
// per src run (heads = run starts), and the edges' bits (row src, column
// dst) in F.
__global__ void tc_rows_init_k(const u64 *uniq, u64 n, const u32 *heads, u64 nh, u32 *dy, u32 *ax, u64 *abeg,
                               u32 *alen, u32 *F, const u64 *row_base, const u32 *col_lo, u32 lo) {
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        u64 t = uniq[i];
        u32 sx = tkey(t) - lo, dz = tval(t) - col_lo[sx];
        dy[i] = tval(t);
        atomicOr(F + row_base[sx] + (dz >> 5), 1u << (dz & 31));
    }
    for (u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x; j < nh; j += stride) {
        u64 b = heads[j], e = j + 1 < nh ? heads[j + 1] : n;
        ax[j] = tkey(uniq[b]);
        abeg[j] = b;
        alen[j] = (u32)(e - b);
    }
}
// The active list back to (z, x) tuples, compacted (the ends have gaps
// between entries), for a round hook.
__global__ void tc_rows_tuples_k(const u32 *dy, const u32 *ax, const u64 *abeg, const u32 *alen, const u64 *pos,
                                 u64 nact, u64 *out) {
    for (u64 j = blockIdx.x; j < nact; j += gridDim.x)
        for (u32 i = threadIdx.x; i < alen[j]; i += blockDim.x) out[pos[j] + i] = tmake(dy[abeg[j] + i], ax[j]);
}

// The first round's state: Δ0's ends and entries (its span is already in
// place), no totals, not stalled.
__global__ void tc_rows_state_k(u64 *st, u64 nd, u64 nact) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        st[TR_ST_ND] = nd;
        st[TR_ST_RAW] = 0;
        st[TR_ST_NACT] = nact;
        st[4] = st[5] = 0;
        st[TR_ST_RUN] = 0;
        st[TR_ST_STALL] = 0;
    }
}
__global__ void tc_rows_unstall_k(u64 *st) {
    if (threadIdx.x == 0 && blockIdx.x == 0) st[TR_ST_STALL] = 0;
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
    // every degree <= TR_ELL_K: the join reads ELL rows instead (one load level less)
    tr.ell.release();
    if (n && tr.dmax && tr.dmax <= (u32)TR_ELL_K) {
        tr.ell.ensure(W * TR_ELL_K);
        fill_u32(ctx, tr.ell.p, W * TR_ELL_K, TR_EMPTY);
        DLG_LAUNCH(ctx, DLG_K_MISC, 20.0 * n, ell_fill_k, sgrid(ctx, n), ctx->block_size, 0, (const u64 *)ix.rows.p,
                   n, (const u64 *)tr.csr_off.p, F.lo, tr.ell.p);
    }
    // Δ0 = unique(E reversed), grouped by source; one segment per source, for good
    DBuf<u64> uniq(ctx, n ? n : 1);
    u64 nu = n ? select(ctx, SEL_UNIQUE, ix.rows.p, n, nullptr, 0, uniq.p) : 0;
    DBuf<u32> heads(ctx, nu ? nu : 1);
    u64 nh = nu ? select(ctx, SEL_KEY_HEAD, uniq.p, nu, nullptr, 0, (u64 *)heads.p) : 0;
    tr.dy.ensure(nu ? nu : 1);
    tr.ax.ensure(nh ? nh : 1);
    tr.abeg.ensure(nh ? nh : 1);
    tr.alen.ensure(nh ? nh : 1);
    tr.aoff.ensure(nh ? nh : 1);
    tr.counters.ensure(TR_ST_WORDS);
    tr.nd = nu;
    tr.nact = nh;
    tr.rows = 0;
    tr.wcap = 0;
    for (const auto &b : F.blocks) tr.wcap = std::max<u64>(tr.wcap, b.W);
    if (!nu) return;
    DLG_LAUNCH(ctx, DLG_K_MISC, 20.0 * nu, tc_rows_init_k, sgrid(ctx, nu), ctx->block_size, 0,
               (const u64 *)uniq.p, nu, (const u32 *)heads.p, nh, tr.dy.p, tr.ax.p, tr.abeg.p, tr.alen.p,
               F.words.p, (const u64 *)F.row_base.p, (const u32 *)F.col_lo.p, F.lo);
    DLG_LAUNCH(ctx, DLG_K_MISC, 12.0 * nu, tc_rows_segrows_k, sgrid(ctx, nh), ctx->block_size, 0,
               (const u32 *)tr.dy.p, (const u64 *)tr.abeg.p, (const u32 *)tr.alen.p, nh, (const u64 *)tr.csr_off.p,
               F.lo, tr.aoff.p);
    scan_u64_async(ctx, tr.aoff.p, nh, tr.aoff.p, tr.counters.p + TR_ST_SPAN);
    DLG_LAUNCH(ctx, DLG_K_MISC, 0.0, tc_rows_state_k, 1, 64, 0, tr.counters.p, nu, nh);
    u64 h[2];
    ctx->read_scalars(tr.counters.p, 2, h);
    tr.rows = h[1];
    // the active list never lengthens: its buffers are sized once
    for (DBuf<u32> *b : {&tr.rcnt, &tr.nx, &tr.nlen}) b->ensure(nh);
    for (DBuf<u64> *b : {&tr.rbound, &tr.nbeg, &tr.noff}) b->ensure(nh);
}

constexpr u64 TR_LONG_SEGMENTS = 64;   // average probes per segment above which the big table wins

template <int LOG, bool ELL>
static unsigned launch_rows(dlg_ctx *ctx, RowMap &F, TcRows &tr) {
    unsigned g = (unsigned)std::min<u64>((tr.nact + TR_W - 1) / TR_W, (u64)ctx->num_cus * tr_blocks_per_cu<LOG, ELL>());
    tr.bpart.ensure(TR_PART * (u64)g);
    tr.brange.ensure((u64)g + 1);
    DLG_LAUNCH(ctx, DLG_K_BITMAP_INSERT, 0.0, (tc_rows_k<LOG, ELL>), g, TR_T, 0, (const u32 *)tr.dy.p,
               (const u32 *)tr.ax.p, (const u64 *)tr.abeg.p, (const u32 *)tr.alen.p, (const u64 *)tr.aoff.p,
               tr.counters.p, tr.ny.cap, (const u64 *)tr.csr_off.p, (const u32 *)tr.csr_dst.p, (const u32 *)tr.deg.p,
               (const uint4 *)tr.ell.p, tr.dmax, F.words.p,
               (const u64 *)F.row_base.p, (const u32 *)F.col_lo.p, F.lo, tr.wcap, tr.ny.p, tr.rcnt.p, tr.rbound.p,
               tr.bpart.p, tr.brange.p);
    return g;
}

static void tc_rows_swap(TcRows &tr) {
    tr.dy.swap(tr.ny);
    tr.ax.swap(tr.nx);
    tr.abeg.swap(tr.nbeg);
    tr.alen.swap(tr.nlen);
    tr.aoff.swap(tr.noff);
}

// One round queued: the row kernel and the compaction, whose totals go to
// the mapped slot of the returned number.  Sizes come from the last round
// the host has read (the rounds queued since can only shorten the list);
// the round itself reads its own from the device state.  The buffers of Δ
// and of the list swap for the next round.  The output buffer grows to the
// last known span only when no round is queued (`may_grow`): behind a queued
// round it is that round's input, which a stall must find in place again.
static u64 tc_rows_enqueue(dlg_ctx *ctx, RowMap &F, TcRows &tr, bool may_grow) {
    const u64 na = tr.nact ? tr.nact : 1;
    // half again the last known span: the rounds queued behind it rarely outgrow it (a larger
    // span stalls the round)
    if (may_grow || !tr.ny.p) tr.ny.ensure(tr.rows ? tr.rows + tr.rows / 2 : 1);
    const bool big = tr.nd >= TR_LONG_SEGMENTS * na;   // long segments: big tables; short ones: more waves
    unsigned g;
    if (tr.ell.p) g = big ? launch_rows<10, true>(ctx, F, tr) : launch_rows<9, true>(ctx, F, tr);
    else g = big ? launch_rows<10, false>(ctx, F, tr) : launch_rows<9, false>(ctx, F, tr);
    const u64 seq = ++ctx->mapped_seq;
    DLG_LAUNCH(ctx, DLG_K_MISC, 48.0 * na, tc_rows_compact_k, g, TA_T, 0, (const u32 *)tr.ax.p,
               (const u64 *)tr.aoff.p, (const u32 *)tr.rcnt.p, (const u64 *)tr.rbound.p, (const u64 *)tr.bpart.p,
               (const u64 *)tr.brange.p, tr.nx.p, tr.nbeg.p, tr.nlen.p, tr.noff.p, tr.counters.p,
               ctx->mapped_dev(seq), seq);
    tc_rows_swap(tr);
    return seq;
}

// The totals of round `seq` (waits for them).  A stalled round did nothing:
// the rounds queued after it did nothing either (the flag holds), so after
// they drain the buffers swap back to the stalled round's (no buffer was
// replaced while they were queued), its output buffer grows to twice the
// span, its input buffer (Δ copied over) to the same size, the flag clears
// and the caller queues again from there.
// Returns false for a stalled round.
static bool tc_rows_collect(dlg_ctx *ctx, TcRows &tr, u64 seq, size_t queued_after, u64 *h) {
    ctx->wait_mapped(seq);
    const volatile u64 *m = ctx->mapped_host(seq);
    for (int q = 0; q <= TR_PART; q++) h[q] = m[q];
    if (h[TR_PART] != TR_STALLED) return true;
    ctx->sync();
    if ((queued_after + 1) % 2) tc_rows_swap(tr);
    tr.ny.ensure(std::max<u64>(2 * h[TR_ST_SPAN], 2 * tr.ny.cap));
    if (tr.dy.cap < tr.ny.cap) {   // the other buffer of the pair as large: Δ moves into a copy
        DBuf<u32> big(ctx, tr.ny.cap);
        HIPCHK(hipMemcpyAsync(big.p, tr.dy.p, sizeof(u32) * tr.dy.cap, hipMemcpyDeviceToDevice, ctx->stream));
        tr.dy.swap(big);
    }
    DLG_LAUNCH(ctx, DLG_K_MISC, 0.0, tc_rows_unstall_k, 1, 64, 0, tr.counters.p);
    return false;
}

// Book-keeping of a round the host has read: returns |N|.
static u64 tc_rows_account(dlg_ctx *ctx, TcRows &tr, const u64 *h, u64 nd_in, u64 *raw) {
    if (h[TR_PART] == TR_IDLE) {   // no ends or no join rows: nothing new
        tr.nd = 0;
        return 0;
    }
    tr.nd = h[TR_ST_ND];
    tr.rows = h[TR_ST_SPAN];
    tr.nact = h[TR_ST_NACT];
    if (raw) *raw += h[TR_ST_RAW];
    // algorithmic bytes, added once the round's join rows are known: 4 B per
    // join row (its matched end) + 20 B per probe (its end, CSR offset,
    // degree) + 4 B per new end (written as the next round's Δ) + 4 B per
    // word of F read and 4 B per word of F written (the set difference and
    // merge of the bitmap layout)
    ctx->prof_add_bytes(DLG_K_BITMAP_INSERT,
                        4.0 * h[TR_ST_RAW] + 20.0 * nd_in + 4.0 * h[TR_ST_ND] + 4.0 * h[4] + 4.0 * h[5]);
    return h[TR_ST_ND];
}

u64 tc_rows_round(dlg_ctx *ctx, RowMap &F, TcRows &tr, u64 *raw) {
    if (tr.nd == 0 || tr.rows == 0) {
        tr.nd = 0;
        return 0;
    }
    const u64 nd_in = tr.nd;
    u64 h[TR_PART + 1];
    while (!tc_rows_collect(ctx, tr, tc_rows_enqueue(ctx, F, tr, true), 0, h)) {
    }
    return tc_rows_account(ctx, tr, h, nd_in, raw);
}

int tc_rows_run(dlg_ctx *ctx, RowMap &F, TcRows &tr, int max_it, int queue, u64 *total, u64 *raw) {
    int it = 0;
    const int depth = queue > 0 ? std::min(queue, (int)dlg_ctx::MAPPED_SLOTS) : TR_QUEUE;
    if (tr.nd == 0 || tr.rows == 0) {   // one round that finds nothing
        tr.nd = 0;
        return 1;
    }
    std::deque<u64> q;   // numbers of the queued rounds, oldest first
    u64 nd_in = tr.nd;
    while (true) {
        while ((int)q.size() < depth && (max_it <= 0 || it + (int)q.size() < max_it))
            q.push_back(tc_rows_enqueue(ctx, F, tr, q.empty()));
        u64 h[TR_PART + 1];
        if (!tc_rows_collect(ctx, tr, q.front(), q.size() - 1, h)) {
            q.clear();
            continue;
        }
        q.pop_front();
        it++;
        const u64 nn = tc_rows_account(ctx, tr, h, nd_in, raw);
        nd_in = nn;
        *total += nn;
        if (nn == 0 || (max_it > 0 && it >= max_it)) {
            // the rounds still queued find no ends and do nothing; tr keeps this round's buffers
            if (q.size() % 2) tc_rows_swap(tr);
            return it;
        }
    }
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
    if (m.blocks.size() > 1) {   // many blocks (a batch's instances): one emission for all
        std::vector<MapRef> maps;
        maps.reserve(m.blocks.size());
        for (const auto &b : m.blocks) maps.push_back({b.lo, b.W, b.wpr, b.base});
        return bitmap_emit_maps(ctx, m.words.p, maps, true, out.p, expect);
    }
    DBuf<u32> scratch(ctx);
    u64 pos = 0;
    for (const auto &b : m.blocks)
        pos += bitmap_emit_block(ctx, m.words.p + b.base, b.lo, b.W, b.wpr, true, scratch, out.p + pos);
    if (pos != expect) fail(DLG_DEVICE_ERROR, "bitmap emission count mismatch");
    return pos;
}

// ---------------------------------------------------------------- locality order
__global__ void relabel_rows_k(const int32_t *in, u64 m, const u32 *sigma, u32 lo, int32_t *out) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        const u32 k = (u32)in[i] ^ 0x80000000u;
        out[i] = (int32_t)((sigma[k - lo] + lo) ^ 0x80000000u);
    }
}

void relabel_rows(dlg_ctx *ctx, const int32_t *in, u64 n, const u32 *sigma, u32 lo, int32_t *out) {
    if (n)
        DLG_LAUNCH(ctx, DLG_K_MISC, 16.0 * n, relabel_rows_k, sgrid(ctx, 2 * n), ctx->block_size, 0, in, 2 * n, sigma,
                   lo, out);
}

u64 rowmap_emit_renumbered(dlg_ctx *ctx, RowMap &m, DBuf<u64> &out, u64 expect, const u32 *sigma) {
    if (m.blocks.size() != 1) fail(DLG_DEVICE_ERROR, "renumbered emission: a map of one block expected");
    out.ensure(expect ? expect : 1);
    DBuf<u32> scratch(ctx);
    const auto &b = m.blocks[0];
    const u64 pos = bitmap_emit_block_renumbered(ctx, m.words.p + b.base, b.lo, b.W, b.wpr, sigma, scratch, out.p);
    if (pos != expect) fail(DLG_DEVICE_ERROR, "bitmap emission count mismatch");
    return pos;
}

void tc_rows_tuples(dlg_ctx *ctx, const TcRows &tr, DBuf<u64> &out) {
    out.ensure(tr.nd ? tr.nd : 1);
    if (!tr.nd) return;
    DBuf<u64> pos(ctx, tr.nact);
    DBuf<u64> tot(ctx, 1);
    scan_u32_u64_async(ctx, tr.alen.p, tr.nact, pos.p, tot.p);
    unsigned g = (unsigned)std::min<u64>(tr.nact, (u64)ctx->num_cus * 8);
    DLG_LAUNCH(ctx, DLG_K_MISC, 16.0 * tr.nd, tc_rows_tuples_k, g, 256, 0, (const u32 *)tr.dy.p,
               (const u32 *)tr.ax.p, (const u64 *)tr.abeg.p, (const u32 *)tr.alen.p, (const u64 *)pos.p, tr.nact,
               out.p);
}

}  // namespace dlg
