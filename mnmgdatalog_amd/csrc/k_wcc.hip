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
// edge with a binary search in the frontier offsets.
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

constexpr int WR_T = 256, WR_I = 8, WR_TILE = WR_T * WR_I;

__global__ __launch_bounds__(WR_T) void wcc_relax_k(const u32 *flabel, u64 nf, const u32 *pstart,
                                                    const u64 *offs, u64 total, const u32 *nbr,
                                                    u32 *L, u32 *mark, u32 round, u32 *next_node,
                                                    u32 *next_count) {
    __shared__ u64 s_lo, s_hi;
    const u64 base = (u64)blockIdx.x * WR_TILE;
    const int lane = threadIdx.x & (WAVE - 1);
    if (threadIdx.x == 0) {
        u64 last = (base + WR_TILE < total ? base + WR_TILE : total) - 1;
        u64 lo = upper_bound_u64(offs, 0, nf, base) - 1;
        s_lo = lo;
        s_hi = upper_bound_u64(offs, lo, nf, last);
    }
    __syncthreads();
    const u64 lo = s_lo, hi = s_hi;
    const u64 lt = lanemask_lt();
#pragma unroll
    for (int r = 0; r < WR_I; r++) {
        u64 o = base + (u64)r * WR_T + threadIdx.x;
        bool app = false;
        u32 m = 0;
        if (o < total) {
            u64 i = upper_bound_u64(offs, lo, hi, o) - 1;
            m = nbr[(u64)pstart[i] + (o - offs[i])];
            u32 l = flabel[i];
            if (L[m] > l) {
                u32 old = atomicMin(&L[m], l);
                if (old > l) app = atomicExch(&mark[m], round) != round;
            }
        }
        u64 bm = __ballot(app);
        if (bm) {
            int leader = __ffsll((unsigned long long)bm) - 1;
            u32 b = 0;
            if (lane == leader) b = atomicAdd(next_count, (u32)__popcll(bm));
            b = __shfl(b, leader, WAVE);
            if (app) next_node[b + (u32)__popcll(bm & lt)] = m;
        }
    }
}

__global__ void wcc_gather_k(const u32 *node, u64 n, const u32 *L, u32 *label) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < n; i += stride) label[i] = L[node[i]];
}

__global__ void wcc_comp_count_k(const u32 *L, u64 V, u32 *ccount) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < V; i += stride) atomicAdd(&ccount[L[i]], 1u);
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

// One round.  Returns |Δ'| (the new frontier size); the frontier is replaced.
u64 wcc_round(dlg_ctx *ctx, WccState &s, const Index &ix, const u32 *nbr, u64 nf, u32 round) {
    s.cnt.ensure(nf);
    s.pstart.ensure(nf);
    s.offs.ensure(nf + 1);
    DLG_LAUNCH(ctx, DLG_K_WCC_RELAX, 16.0 * nf, wcc_degree_k, sgrid(ctx, nf), ctx->block_size, 0,
               (const u32 *)s.fnode.p, nf, (const u32 *)ix.gstart.p, s.cnt.p, s.pstart.p);
    u64 total = scan_u32_u64(ctx, s.cnt.p, nf, s.offs.p);
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
    DLG_LAUNCH(ctx, DLG_K_WCC_FINAL, 12.0 * V, wcc_comp_count_k, sgrid(ctx, V), ctx->block_size, 0,
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
