// k_sgrows.hip -- a round of SG as row operations on bitmaps.
//
// A round of SG (sg.cu:300-375) derives
//     T(b, x) = Δ(a, b) ⋈ edge(a, x)            (the round's T map R, row-major: row b)
//     N(x, y) = T(b, x) ⋈ edge(b, y) \ F        (new tuples)
// with F and D transposed (row y = the x's of N(x, y); D holds the round's
// new tuples, i.e. the next round's Δ, row b = the a's of Δ(a, b)).  Both
// joins become work on whole rows, with one owner per row and no atomics in
// device memory:
//   T row b = OR over a in D row b of a's successors: one workgroup per row b
//     ORs the successors' bits into an LDS copy of the row (LDS atomics with
//     no result), clears D row b as it reads it, and writes R row b (empty
//     before) and its flag;
//   N row y = (OR over b with edge(b, y) of R row b) & ~(F row y): one
//     workgroup per row y ORs its predecessors' flagged R rows word by word
//     (coalesced) and sets the new bits in F and D with plain stores.
// The raw join rows are Σ deg(a) over Δ and the popcounts of the R words the
// second step reads (each R row b is read once per y with edge(b, y)).  The
// flagged R rows are cleared afterwards, so R is empty again.
#include "dlg_internal.h"

#include <algorithm>
#include <vector>

namespace dlg {

constexpr int SR_T = 256;   // threads per workgroup of the row OR
// The round's totals are added by every workgroup; workgroup b adds into
// slice b % SG_SLICES of the counters (one 64-byte line each), and the host
// sums the slices: device atomics from 2,048 workgroups on one line serialise.
constexpr int SG_SLICES = 64, SG_SLICE_WORDS = 8;
__device__ __forceinline__ unsigned long long *sg_slice(unsigned long long *counters) {
    return counters + SG_SLICE_WORDS * (blockIdx.x % SG_SLICES);
}

// predecessor lists: pred[poff[y] ..] = the b of every edge(b, y), relative to lo
__global__ void sg_pred_hist_k(const u64 *rows, u64 n, u32 lo, u32 *cnt) {
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        atomicAdd(cnt + (tval(rows[i]) - lo), 1u);
}
__global__ void sg_pred_fill_k(const u64 *rows, u64 n, u32 lo, const u64 *poff, u32 *cur, u32 *pred) {
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const u32 y = tval(rows[i]) - lo;
        pred[poff[y] + atomicAdd(cur + y, 1u)] = tkey(rows[i]) - lo;
    }
}

// successor lists: succ[soff[a] ..] = the x of every edge(a, x), relative to lo
__global__ void sg_succ_hist_k(const u64 *rows, u64 n, u32 lo, u32 *cnt) {
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        atomicAdd(cnt + (tkey(rows[i]) - lo), 1u);
}
__global__ void sg_succ_fill_k(const u64 *rows, u64 n, u32 lo, u32 *succ) {
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) succ[i] = tval(rows[i]) - lo;
}

// T row b = OR over a in D row b of a's successors (row b of D emptied).  The
// row is built in LDS (acc: wpr words of dynamic shared memory) and R row b
// (empty before) is written and flagged.  D row b is taken SG_KW x 256 words
// at a time; a block scan of the words' popcounts packs their a's into an LDS
// list (SG_CAP at a time), and the threads take the list's a's in turn, two
// at a time, so a word dense in a's is spread over the workgroup.  ELL: a's
// successors are one row of SG_ELL_K ids (SG_EMPTY past the degree), else
// a CSR list.  In the workgroup's slice of the counters: [4] += Σ deg(a)
// (the join's raw rows), [5] += the words of D read.
constexpr int SG_ELL_K = 16;
constexpr u32 SG_EMPTY = 0xffffffffu;   // ids relative to lo stay below it
constexpr int SG_CAP = 2048;
constexpr int SG_KW = 8;   // D words per thread per scan of a row
template <bool ELL>
__global__ __launch_bounds__(SR_T) void sg_tjoin_k(u32 *D, const u64 *soff, const u32 *succ, const uint4 *ell,
                                                   u64 W, u64 wpr, u32 *R, u32 *rflag,
                                                   unsigned long long *counters) {
    extern __shared__ u32 acc[];
    __shared__ u32 alist[SG_CAP];
    __shared__ u64 s_scan[SR_T / WAVE + 1];
    const int t = threadIdx.x;
    u64 raw = 0, dwords = 0;
    for (u64 b = blockIdx.x; b < W; b += gridDim.x) {
        u32 *drow = D + b * wpr;
        for (u64 j = t; j < wpr; j += SR_T) acc[j] = 0u;
        u64 rowbits = 0;
        for (u64 j0 = 0; j0 < wpr; j0 += SG_KW * SR_T) {   // uniform: SG_KW words per thread
            u32 v[SG_KW], pc = 0;
#pragma unroll
            for (int q = 0; q < SG_KW; q++) {
                const u64 j = j0 + (u64)q * SR_T + t;
                v[q] = j < wpr ? drow[j] : 0u;
                dwords += j < wpr;
                pc += (u32)__popc(v[q]);
            }
#pragma unroll
            for (int q = 0; q < SG_KW; q++)
                if (v[q]) drow[j0 + (u64)q * SR_T + t] = 0u;   // the row's owner
            u64 tot;
            const u64 off = block_exclusive_scan((u64)pc, s_scan, &tot);   // syncs: acc zeroed too
            rowbits += tot;
            for (u64 w0 = 0; w0 < tot; w0 += SG_CAP) {   // uniform
                u64 k = off;
#pragma unroll
                for (int q = 0; q < SG_KW; q++) {
                    const u64 j = j0 + (u64)q * SR_T + t;
                    for (u32 m = v[q]; m; m &= m - 1, k++)
                        if (k >= w0 && k < w0 + SG_CAP) alist[k - w0] = (u32)(j * 32 + (u32)(__ffs(m) - 1));
                }
                __syncthreads();
                const u32 n = (u32)(tot - w0 < (u64)SG_CAP ? tot - w0 : (u64)SG_CAP);
                for (u32 i = t; i < n; i += 2 * SR_T) {
                    const bool two = i + SR_T < n;
                    const u32 a0 = alist[i], a1 = two ? alist[i + SR_T] : a0;
                    if constexpr (ELL) {
                        uint4 e[2][SG_ELL_K / 4];   // both rows' loads in flight
#pragma unroll
                        for (int q = 0; q < SG_ELL_K / 4; q++) {
                            e[0][q] = ell[(u64)a0 * (SG_ELL_K / 4) + q];
                            e[1][q] = ell[(u64)a1 * (SG_ELL_K / 4) + q];
                        }
#pragma unroll
                        for (int r = 0; r < 2; r++) {
                            if (r == 1 && !two) break;
#pragma unroll
                            for (int q = 0; q < SG_ELL_K / 4; q++) {
                                const u32 z[4] = {e[r][q].x, e[r][q].y, e[r][q].z, e[r][q].w};
#pragma unroll
                                for (int c = 0; c < 4; c++)
                                    if (z[c] != SG_EMPTY) {
                                        atomicOr(&acc[z[c] >> 5], 1u << (z[c] & 31));
                                        raw++;
                                    }
                            }
                        }
                    } else {
#pragma unroll
                        for (int r = 0; r < 2; r++) {
                            if (r == 1 && !two) break;
                            const u32 a = r ? a1 : a0;
                            const u64 e0 = soff[a], e1 = soff[a + 1];
                            raw += e1 - e0;
                            for (u64 e = e0; e < e1; e++) {
                                const u32 x = succ[e];
                                atomicOr(&acc[x >> 5], 1u << (x & 31));
                            }
                        }
                    }
                }
                __syncthreads();   // the list is refilled by the next window
            }
        }
        if (rowbits) {   // uniform: write R row b and its flag
            u32 wrote = 0;
            for (u64 j = t; j < wpr; j += SR_T) {
                const u32 v = acc[j];
                if (v) {
                    R[b * wpr + j] = v;
                    wrote = 1;
                }
            }
            if (wrote) rflag[b] = 1u;
        }
        __syncthreads();   // acc is zeroed again for the next row
    }
    u64 rt, dt;
    block_exclusive_scan(raw, s_scan, &rt);
    block_exclusive_scan(dwords, s_scan, &dt);
    if (t == 0) {
        if (rt) atomicAdd(sg_slice(counters) + 4, (unsigned long long)rt);
        if (dt) atomicAdd(sg_slice(counters) + 5, (unsigned long long)dt);
    }
}

// SG's ELL rows: the successors of id a at ell[a * SG_ELL_K ..], relative to lo.
__global__ void sg_ell_fill_k(const u64 *rows, u64 n, const u64 *soff, u32 lo, u32 *ell) {
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const u32 a = tkey(rows[i]) - lo;
        ell[(u64)a * SG_ELL_K + (i - soff[a])] = tval(rows[i]) - lo;
    }
}

// clear the flagged rows and their flags (one wave per row)
__global__ void bm_clear_rows_k(u32 *words, u64 W, u64 wpr, u32 *flag) {
    const int lane = threadIdx.x & (WAVE - 1);
    const u64 nw = (u64)gridDim.x * blockDim.x / WAVE;
    for (u64 r = ((u64)blockIdx.x * blockDim.x + threadIdx.x) / WAVE; r < W; r += nw) {
        if (!flag[r]) continue;   // uniform per wave
        for (u64 j = lane; j < wpr; j += WAVE) words[r * wpr + j] = 0u;
        if (lane == 0) flag[r] = 0u;
    }
}

// N row y = OR of the flagged R rows of y's predecessors, minus F row y;
// new bits go into F and D.  In the workgroup's slice of the counters:
// [0] += the raw join rows (popcounts of the R words read), [1] += R words
// read, [2] += F words read, [3] += new bits.
__global__ __launch_bounds__(SR_T) void sg_rowor_k(const u32 *R, const u32 *rflag, const u64 *poff,
                                                   const u32 *pred, u64 W, u64 wpr, u32 *F, u32 *D,
                                                   unsigned long long *counters) {
    __shared__ u32 s_act[SR_T];
    __shared__ u32 s_n;
    __shared__ u64 s_red[4][SR_T / WAVE];
    u64 raw = 0, rwords = 0, fwords = 0, nnew = 0;
    for (u64 y = blockIdx.x; y < W; y += gridDim.x) {
        const u64 p0 = poff[y], p1 = poff[y + 1];
        for (u64 pb = p0; pb < p1; pb += SR_T) {   // predecessors in batches of SR_T
            if (threadIdx.x == 0) s_n = 0;
            __syncthreads();
            const u64 k = pb + threadIdx.x;
            if (k < p1) {
                const u32 b = pred[k];
                if (rflag[b]) s_act[atomicAdd(&s_n, 1u)] = b;
            }
            __syncthreads();
            const u32 na = s_n;
            if (na) {   // uniform
                u32 *fy = F + y * wpr, *dw = D + y * wpr;
                if (threadIdx.x < wpr) rwords += (u64)na * ((wpr - threadIdx.x + SR_T - 1) / SR_T);
                for (u64 j = threadIdx.x; j < wpr; j += SR_T) {
                    u32 acc = 0;
                    u32 i = 0;
                    for (; i + 4 <= na; i += 4) {   // four rows' loads in flight
                        const u32 v0 = R[(u64)s_act[i] * wpr + j], v1 = R[(u64)s_act[i + 1] * wpr + j];
                        const u32 v2 = R[(u64)s_act[i + 2] * wpr + j], v3 = R[(u64)s_act[i + 3] * wpr + j];
                        acc |= v0 | v1 | v2 | v3;
                        raw += (u64)(__popc(v0) + __popc(v1) + __popc(v2) + __popc(v3));
                    }
                    for (; i < na; i++) {
                        const u32 v = R[(u64)s_act[i] * wpr + j];
                        acc |= v;
                        raw += (u64)__popc(v);
                    }
                    if (acc) {
                        fwords++;
                        const u32 f = fy[j], nb = acc & ~f;   // the row's owner: plain read-modify-write
                        if (nb) {
                            fy[j] = f | nb;
                            dw[j] |= nb;
                            nnew += (u64)__popc(nb);
                        }
                    }
                }
            }
            __syncthreads();   // s_act and s_n are refilled by the next batch
        }
    }
    const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        raw += __shfl_xor(raw, o, WAVE);
        rwords += __shfl_xor(rwords, o, WAVE);
        fwords += __shfl_xor(fwords, o, WAVE);
        nnew += __shfl_xor(nnew, o, WAVE);
    }
    if (lane == 0) {
        s_red[0][wv] = raw;
        s_red[1][wv] = rwords;
        s_red[2][wv] = fwords;
        s_red[3][wv] = nnew;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        u64 t = 0;
        for (int w = 0; w < SR_T / WAVE; w++) t += s_red[threadIdx.x][w];
        if (t) atomicAdd(sg_slice(counters) + threadIdx.x, (unsigned long long)t);
    }
}

void sg_rows_setup(dlg_ctx *ctx, const Index &ix, const Bitmap &F, SgRows &sr) {
    const u64 W = F.W, n = ix.n;
    DBuf<u32> cnt(ctx, W + 1);
    fill_u32(ctx, cnt.p, W + 1, 0u);
    if (n)
        DLG_LAUNCH(ctx, DLG_K_MISC, 12.0 * n, sg_pred_hist_k, sgrid(ctx, n), ctx->block_size, 0,
                   (const u64 *)ix.rows.p, n, F.lo, cnt.p);
    sr.poff.ensure(W + 1);
    scan_u32_u64(ctx, cnt.p, W + 1, sr.poff.p);
    sr.pred.ensure(n ? n : 1);
    fill_u32(ctx, cnt.p, W + 1, 0u);
    if (n)
        DLG_LAUNCH(ctx, DLG_K_MISC, 20.0 * n, sg_pred_fill_k, sgrid(ctx, n), ctx->block_size, 0,
                   (const u64 *)ix.rows.p, n, F.lo, (const u64 *)sr.poff.p, cnt.p, sr.pred.p);
    // successors: the index's rows are sorted by (source, target), so the
    // lists are the rows' targets at the rows' offsets by source
    fill_u32(ctx, cnt.p, W + 1, 0u);
    if (n)
        DLG_LAUNCH(ctx, DLG_K_MISC, 12.0 * n, sg_succ_hist_k, sgrid(ctx, n), ctx->block_size, 0,
                   (const u64 *)ix.rows.p, n, F.lo, cnt.p);
    sr.soff.ensure(W + 1);
    scan_u32_u64(ctx, cnt.p, W + 1, sr.soff.p);
    sr.succ.ensure(n ? n : 1);
    if (n)
        DLG_LAUNCH(ctx, DLG_K_MISC, 12.0 * n, sg_succ_fill_k, sgrid(ctx, n), ctx->block_size, 0,
                   (const u64 *)ix.rows.p, n, F.lo, sr.succ.p);
    // every out-degree <= SG_ELL_K: successors as ELL rows (one load level, no offsets)
    {
        std::vector<u64> ho(W + 1);
        HIPCHK(hipMemcpyAsync(ho.data(), sr.soff.p, 8 * (W + 1), hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
        u64 dm = 0;
        for (u64 i = 0; i < W; i++) dm = std::max(dm, ho[i + 1] - ho[i]);
        sr.ell.release();
        if (n && dm <= (u64)SG_ELL_K) {
            sr.ell.ensure(W * SG_ELL_K);
            fill_u32(ctx, sr.ell.p, W * SG_ELL_K, SG_EMPTY);
            DLG_LAUNCH(ctx, DLG_K_MISC, 20.0 * n, sg_ell_fill_k, sgrid(ctx, n), ctx->block_size, 0,
                       (const u64 *)ix.rows.p, n, (const u64 *)sr.soff.p, F.lo, sr.ell.p);
        }
    }
    sr.rflag.ensure(W);
    fill_u32(ctx, sr.rflag.p, W, 0u);
    sr.counters.ensure(SG_SLICES * SG_SLICE_WORDS);
    ctx->sync();   // cnt goes out of scope
}

u64 sg_rows_round(dlg_ctx *ctx, SgRows &sr, Bitmap &R, Bitmap &F, Bitmap &D, u64 *raw) {
    if (R.tr || !F.tr || !D.tr || R.W != F.W || R.W != D.W || R.lo != F.lo || R.lo != D.lo)
        fail(DLG_INVALID_INPUT, "sg_rows_round: R row-major, F and D transposed, one geometry");
    const u64 W = R.W, wpr = R.wpr;
    if ((u64)ctx->lds_per_block < 4 * wpr + 4 * SG_CAP + 4096)
        fail(DLG_INVALID_INPUT, "sg_rows_round: a row of the map does not fit the LDS of a workgroup");
    fill_u64(ctx, sr.counters.p, SG_SLICES * SG_SLICE_WORDS, 0ull);
    const unsigned g = (unsigned)std::min<u64>(W, (u64)ctx->num_cus * 8);
    if (sr.ell.p)
        DLG_LAUNCH(ctx, DLG_K_BITMAP_INSERT, 0.0, sg_tjoin_k<true>, g, SR_T, (unsigned)(4 * wpr), D.words.p,
                   (const u64 *)sr.soff.p, (const u32 *)sr.succ.p, (const uint4 *)sr.ell.p, W, wpr, R.words.p,
                   sr.rflag.p, (unsigned long long *)sr.counters.p);
    else
        DLG_LAUNCH(ctx, DLG_K_BITMAP_INSERT, 0.0, sg_tjoin_k<false>, g, SR_T, (unsigned)(4 * wpr), D.words.p,
                   (const u64 *)sr.soff.p, (const u32 *)sr.succ.p, (const uint4 *)sr.ell.p, W, wpr, R.words.p,
                   sr.rflag.p, (unsigned long long *)sr.counters.p);
    DLG_LAUNCH(ctx, DLG_K_BITMAP_INSERT, 0.0, sg_rowor_k, g, SR_T, 0, (const u32 *)R.words.p,
               (const u32 *)sr.rflag.p, (const u64 *)sr.poff.p, (const u32 *)sr.pred.p, W, wpr, F.words.p,
               D.words.p, (unsigned long long *)sr.counters.p);
    const unsigned gw = (unsigned)std::min<u64>((W + 3) / 4, (u64)ctx->num_cus * 16);   // a wave per row
    DLG_LAUNCH(ctx, DLG_K_MISC, 4.0 * W, bm_clear_rows_k, gw, 256, 0, R.words.p, W, wpr, sr.rflag.p);
    u64 sl[SG_SLICES * SG_SLICE_WORDS], h[6] = {0, 0, 0, 0, 0, 0};
    HIPCHK(hipMemcpyAsync(sl, sr.counters.p, sizeof(sl), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < SG_SLICES; k++)
        for (int q = 0; q < 6; q++) h[q] += sl[SG_SLICE_WORDS * k + q];
    // algorithmic bytes: D words read, 4 B per successor, R words read, touched F words read (+ written)
    ctx->prof_add_bytes(DLG_K_BITMAP_INSERT, 4.0 * h[5] + 4.0 * h[4] + 4.0 * h[1] + 8.0 * h[2]);
    if (raw) *raw += h[4] + h[0];
    return h[3];
}

}  // namespace dlg
