// k_hashset.hip -- device hash set of tuples: membership in the full relation.
//
// The set difference N = unique(J) \ F (tc.cu:260-263) needs one membership
// test per distinct candidate.  Against sorted runs that is a binary search
// per run; against this open-addressed set it is one or two slot reads,
// independent of |F|.  Slots are u64 tuples, linear probing from a 64-bit
// mix of the tuple, EMPTY = all ones; the one tuple equal to EMPTY
// ((INT32_MAX, INT32_MAX)) is kept in a flag instead of a slot.
#include "dlg_internal.h"

#include <algorithm>
#include <cmath>

namespace dlg {

__device__ __forceinline__ u64 mix64(u64 x) {
    // SplitMix64 finaliser (Steele, Lea, Flood 2014)
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

__device__ __forceinline__ bool hs_contains(const u64 *slots, u64 mask, const u32 *flag, u64 x) {
    if (x == EMPTY_SLOT) return *flag != 0;
    u64 p = mix64(x) & mask;
    while (true) {
        u64 s = slots[p];
        if (s == x) return true;
        if (s == EMPTY_SLOT) return false;
        p = (p + 1) & mask;
    }
}

// Insert tuples known to be absent and pairwise distinct.
__global__ void hs_insert_k(const u64 *x, u64 n, u64 *slots, u64 mask, u32 *flag) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        u64 v = x[i];
        if (v == EMPTY_SLOT) {
            *flag = 1u;
            continue;
        }
        u64 p = mix64(v) & mask;
        while (true) {
            unsigned long long prev = atomicCAS((unsigned long long *)&slots[p],
                                                (unsigned long long)EMPTY_SLOT, (unsigned long long)v);
            if (prev == EMPTY_SLOT) break;
            p = (p + 1) & mask;
        }
    }
}

void hashset_reserve(dlg_ctx *ctx, HashSet &hs, u64 need, double lf, const u64 *all, u64 n_all) {
    if (!(lf > 0.0)) lf = ctx->load_factor;
    lf = std::min(0.9, std::max(0.1, lf));
    if (hs.cap && (double)need <= lf * (double)hs.cap) return;
    u64 want = (u64)std::ceil((double)need / lf);
    u64 cap = hs.cap ? hs.cap : 1024;
    while (cap < want || cap <= need) cap <<= 1;
    hs.slots.ensure(cap);
    hs.cap = cap;
    if (!hs.flag.p) hs.flag.ensure(1);
    fill_u64(ctx, hs.slots.p, cap, EMPTY_SLOT);
    fill_u32(ctx, hs.flag.p, 1, 0u);
    hs.count = 0;
    if (n_all) hashset_insert(ctx, hs, all, n_all);   // rehash what the set already held
}

void hashset_insert(dlg_ctx *ctx, HashSet &hs, const u64 *x, u64 n) {
    if (!n) return;
    // a tuple read, one slot line and a CAS per element
    DLG_LAUNCH(ctx, DLG_K_HSET_INSERT, 24.0 * n, hs_insert_k, sgrid(ctx, n), ctx->block_size, 0, x, n,
               hs.slots.p, hs.cap - 1, hs.flag.p);
    hs.count += n;
}

// ---------------------------------------------------------------- selection
constexpr int HS_T = 256, HS_I = 8, HS_TILE = HS_T * HS_I;

__global__ __launch_bounds__(HS_T) void select_count_hash_k(const u64 *in, u64 n, const u64 *slots,
                                                            u64 mask, const u32 *flag, u64 *flags,
                                                            u64 *bsum) {
    __shared__ u64 sc[HS_T / WAVE + 1];
    const u64 base = (u64)blockIdx.x * HS_TILE;
    u64 c = 0;
#pragma unroll
    for (int r = 0; r < HS_I; r++) {
        u64 idx = base + (u64)r * HS_T + threadIdx.x;
        bool keep = false;
        if (idx < n) {
            u64 x = in[idx];
            keep = (idx == 0 || in[idx - 1] != x) && !hs_contains(slots, mask, flag, x);
        }
        u64 m = __ballot(keep);
        u64 w0 = base + (u64)r * HS_T + (threadIdx.x & ~(WAVE - 1));
        if ((threadIdx.x & (WAVE - 1)) == 0 && w0 < n) flags[w0 / WAVE] = m;
        c += keep ? 1 : 0;
    }
    u64 tot;
    block_exclusive_scan<u64>(c, sc, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

u64 select_unique_not_in_hash(dlg_ctx *ctx, const u64 *in, u64 n, const HashSet &hs, u64 *out) {
    if (n == 0) return 0;
    if (hs.cap == 0) return select(ctx, SEL_UNIQUE, in, n, nullptr, 0, out);
    u64 nb = (n + HS_TILE - 1) / HS_TILE;
    DBuf<u64> flags(ctx, (n + WAVE - 1) / WAVE + 1);
    DBuf<u64> bsum(ctx, nb);
    DBuf<u64> tot(ctx, 1);
    DLG_LAUNCH(ctx, DLG_K_SELECT, 8.0 * n + 8.0 * ((n + 63) / 64), select_count_hash_k, (unsigned)nb,
               HS_T, 0, in, n, (const u64 *)hs.slots.p, hs.cap - 1, (const u32 *)hs.flag.p, flags.p,
               bsum.p);
    return select_scatter(ctx, in, n, flags.p, bsum, tot, out);
}

}  // namespace dlg
