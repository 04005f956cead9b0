// k_prims.hip -- data-parallel primitives for the fixpoint engine on gfx950.
//
// Replaces the Thrust calls of the reference (sort/unique/set_difference/
// merge/reduce/exclusive_scan; sites in tc.cu:177-283, sg.cu:150-368,
// wcc.cu:163-463, single_join.cu:165-264, common/join.cu:72-74) with
// hand-written kernels:
//   * exclusive scan       reduce -> scan of block sums -> apply
//   * radix sort           LSD, reduce-then-scan per digit, wave-match ranking
//                          (64-lane ballots), LDS staging for coalesced scatter;
//                          digits run over a compacted key (hi - min_hi, lo - min_lo)
//                          so only the bits that vary are sorted
//   * select (compaction)  unique / unique minus diagonal / unique minus a sorted
//                          set / set difference / key-run heads; flags are
//                          ballot words, offsets are a block-sum scan
//   * merge                rank-by-binary-search scatter (a first on ties)
#include "dlg_internal.h"

#include <algorithm>

namespace dlg {

// =============================================================== fills / copies
__global__ void fill_u64_k(u64 *p, u64 n, u64 v) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}
__global__ void fill_u32_k(u32 *p, u64 n, u32 v) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}
__global__ void copy_u64_k(u64 *dst, const u64 *src, u64 n) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

void fill_u64(dlg_ctx *ctx, u64 *p, u64 n, u64 v) {
    if (!n) return;
    DLG_LAUNCH(ctx, DLG_K_MISC, 8.0 * n, fill_u64_k, sgrid(ctx, n), ctx->block_size, 0, p, n, v);
}
void fill_u32(dlg_ctx *ctx, u32 *p, u64 n, u32 v) {
    if (!n) return;
    DLG_LAUNCH(ctx, DLG_K_MISC, 4.0 * n, fill_u32_k, sgrid(ctx, n), ctx->block_size, 0, p, n, v);
}
void copy_u64(dlg_ctx *ctx, u64 *dst, const u64 *src, u64 n) {
    if (!n) return;
    DLG_LAUNCH(ctx, DLG_K_MISC, 16.0 * n, copy_u64_k, sgrid(ctx, n), ctx->block_size, 0, dst, src, n);
}

// =============================================================== pack / unpack
// Tuple t = ((col_a ^ 0x80000000) << 32) | (col_b ^ 0x80000000): u64 order is
// the signed lexicographic order of the reference's set_cmp.
__global__ void pack_k(const int2 *rows, u64 n, int orient, u64 *out) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        int2 r = rows[i];
        u32 a = (u32)r.x ^ 0x80000000u, b = (u32)r.y ^ 0x80000000u;
        out[i] = orient ? tmake(b, a) : tmake(a, b);
    }
}
// Each (a, b) row packed as (a, b) into out and as (b, a) into out_rev and,
// in the same read, the bounds of both columns (as flipped ids) and the
// count of adjacent rows out of signed lexicographic order; one set of
// atomics per workgroup.
__global__ __launch_bounds__(256) void pack_sym_scan_k(const int2 *rows, u64 n, u64 *out, u64 *out_rev, u32 *mm,
                                                       unsigned long long *desc) {
    u32 mna = 0xffffffffu, mxa = 0u, mnb = 0xffffffffu, mxb = 0u;
    u64 dc = 0;
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int2 r = rows[i];
        const u32 a = (u32)r.x ^ 0x80000000u, b = (u32)r.y ^ 0x80000000u;
        out[i] = tmake(a, b);
        out_rev[i] = tmake(b, a);
        mna = min(mna, a); mxa = max(mxa, a); mnb = min(mnb, b); mxb = max(mxb, b);
        if (i + 1 < n) {
            const int2 q = rows[i + 1];
            dc += r.x > q.x || (r.x == q.x && r.y > q.y);
        }
    }
#pragma unroll
    for (int d = WAVE / 2; d > 0; d >>= 1) {
        mna = min(mna, (u32)__shfl_xor((int)mna, d, WAVE));
        mxa = max(mxa, (u32)__shfl_xor((int)mxa, d, WAVE));
        mnb = min(mnb, (u32)__shfl_xor((int)mnb, d, WAVE));
        mxb = max(mxb, (u32)__shfl_xor((int)mxb, d, WAVE));
        dc += __shfl_xor(dc, d, WAVE);
    }
    __shared__ u32 sm[4][256 / WAVE];
    __shared__ u64 sd[256 / WAVE];
    const int wave = threadIdx.x / WAVE;
    if ((threadIdx.x & (WAVE - 1)) == 0) {
        sm[0][wave] = mna; sm[1][wave] = mxa; sm[2][wave] = mnb; sm[3][wave] = mxb;
        sd[wave] = dc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x / WAVE); w++) {
            mna = min(mna, sm[0][w]); mxa = max(mxa, sm[1][w]);
            mnb = min(mnb, sm[2][w]); mxb = max(mxb, sm[3][w]);
            dc += sd[w];
        }
        // an atomic only where it can still move the bound: after the first
        // workgroups most skip it (a stale read only costs an atomic)
        if (mna < __hip_atomic_load(&mm[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMin(&mm[0], mna);
        if (mxa > __hip_atomic_load(&mm[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(&mm[1], mxa);
        if (mnb < __hip_atomic_load(&mm[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMin(&mm[2], mnb);
        if (mxb > __hip_atomic_load(&mm[3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(&mm[3], mxb);
        if (dc) atomicAdd(desc, (unsigned long long)dc);
    }
}
__global__ void unpack_k(const u64 *t, u64 n, int orient, int2 *rows) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        u64 x = t[i];
        int k = (int)(tkey(x) ^ 0x80000000u), v = (int)(tval(x) ^ 0x80000000u);
        rows[i] = orient ? make_int2(v, k) : make_int2(k, v);
    }
}

void pack_rows(dlg_ctx *ctx, const int32_t *rows, u64 n, int orient, u64 *out) {
    if (!n) return;
    DLG_LAUNCH(ctx, DLG_K_PACK, 16.0 * n, pack_k, sgrid(ctx, n), ctx->block_size, 0,
               (const int2 *)rows, n, orient, out);
}
__global__ void pack_scan_init_k(u64 *w) {   // bounds words (min a, max a, min b, max b), descents
    if (threadIdx.x == 0) {
        u32 *mm = (u32 *)w;
        mm[0] = 0xffffffffu; mm[1] = 0u; mm[2] = 0xffffffffu; mm[3] = 0u;
        w[2] = 0ull;
    }
}
u64 pack_sym_rows_scan(dlg_ctx *ctx, const int32_t *rows, u64 n, u64 *out, u64 *out_rev, u32 *cb) {
    cb[0] = cb[1] = cb[2] = cb[3] = 0u;
    if (!n) return 0;
    DBuf<u64> w(ctx, 3);
    DLG_LAUNCH(ctx, DLG_K_MISC, 0.0, pack_scan_init_k, 1, 64, 0, w.p);
    const unsigned g = (unsigned)std::min<u64>((n + 255) / 256, (u64)ctx->num_cus * 8);
    DLG_LAUNCH(ctx, DLG_K_PACK, 24.0 * n, pack_sym_scan_k, g, 256, 0, (const int2 *)rows, n, out, out_rev,
               (u32 *)w.p, (unsigned long long *)(w.p + 2));
    u64 h[3];
    ctx->read_scalars(w.p, 3, h);
    std::memcpy(cb, h, 16);
    return h[2];
}
void unpack_rows(dlg_ctx *ctx, const u64 *t, u64 n, int orient, int32_t *rows) {
    if (!n) return;
    DLG_LAUNCH(ctx, DLG_K_PACK, 16.0 * n, unpack_k, sgrid(ctx, n), ctx->block_size, 0, t, n, orient,
               (int2 *)rows);
}

// =============================================================== scan
constexpr int SCAN_T = 256, SCAN_I = 8, SCAN_TILE = SCAN_T * SCAN_I;

template <typename T>
__global__ __launch_bounds__(SCAN_T) void scan_reduce_k(const T *in, u64 n, u64 *bsum) {
    __shared__ u64 sc[SCAN_T / WAVE + 1];
    u64 base = (u64)blockIdx.x * SCAN_TILE + (u64)threadIdx.x * SCAN_I;
    u64 s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_I; i++)
        if (base + i < n) s += in[base + i];
    u64 tot;
    block_exclusive_scan<u64>(s, sc, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Block sums scanned in place (exclusive), total to *d_total, by one
// workgroup: each thread sums a contiguous segment of ceil(nb / 1024)
// entries, one block scan of those sums, then each thread rewrites its
// segment from its prefix (the segments stay in cache between the passes).
__global__ __launch_bounds__(1024) void scan_blocks_k(u64 *bsum, u64 nb, u64 *d_total) {
    __shared__ u64 sc[1024 / WAVE + 1];
    const u64 seg = (nb + 1023) / 1024;
    const u64 b0 = (u64)threadIdx.x * seg < nb ? (u64)threadIdx.x * seg : nb;
    const u64 b1 = b0 + seg < nb ? b0 + seg : nb;
    u64 s = 0;
#pragma unroll 8
    for (u64 i = b0; i < b1; i++) s += bsum[i];
    u64 tot;
    u64 run = block_exclusive_scan<u64>(s, sc, &tot);
#pragma unroll 8
    for (u64 i = b0; i < b1; i++) {
        const u64 v = bsum[i];
        bsum[i] = run;
        run += v;
    }
    if (threadIdx.x == 0 && d_total) *d_total = tot;
}

template <typename T>
__global__ __launch_bounds__(SCAN_T) void scan_apply_k(const T *in, u64 n, const u64 *bpre, u64 *out) {
    __shared__ u64 sc[SCAN_T / WAVE + 1];
    u64 base = (u64)blockIdx.x * SCAN_TILE + (u64)threadIdx.x * SCAN_I;
    T v[SCAN_I];
    u64 s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_I; i++) {
        v[i] = base + i < n ? in[base + i] : 0;
        s += v[i];
    }
    u64 ex = block_exclusive_scan<u64>(s, sc, nullptr) + bpre[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SCAN_I; i++) {
        if (base + i < n) out[base + i] = ex;
        ex += v[i];
    }
}

// Block sums scanned in place (exclusive), total to *d_total: one workgroup
// up to SCAN_ONE_WG entries, else reduce -> scan of the partials -> apply.
constexpr u64 SCAN_ONE_WG = 16384;
void scan_blocks(dlg_ctx *ctx, u64 *bsum, u64 nb, u64 *d_total) {
    if (nb <= SCAN_ONE_WG) {
        DLG_LAUNCH(ctx, DLG_K_SCAN, 16.0 * nb, scan_blocks_k, 1, 1024, 0, bsum, nb, d_total);
        return;
    }
    const u64 np = (nb + SCAN_TILE - 1) / SCAN_TILE;
    DBuf<u64> part(ctx, np);
    DLG_LAUNCH(ctx, DLG_K_SCAN, 8.0 * nb + 8.0 * np, scan_reduce_k<u64>, (unsigned)np, SCAN_T, 0,
               (const u64 *)bsum, nb, part.p);
    scan_blocks(ctx, part.p, np, d_total);
    DLG_LAUNCH(ctx, DLG_K_SCAN, 16.0 * nb, scan_apply_k<u64>, (unsigned)np, SCAN_T, 0, (const u64 *)bsum, nb,
               (const u64 *)part.p, bsum);
}

u64 scan_partials(u64 n) {
    u64 nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    return nb ? nb : 1;
}

// The three launches of a scan over caller-provided partials (scan_partials(n) u64s).
template <typename T>
static void scan_to_u64_with(dlg_ctx *ctx, const T *in, u64 n, u64 *out, u64 *bsum, u64 *d_total) {
    const u64 nb = scan_partials(n);
    DLG_LAUNCH(ctx, DLG_K_SCAN, (double)sizeof(T) * n + 8.0 * nb, scan_reduce_k<T>, (unsigned)nb, SCAN_T, 0,
               in, n, bsum);
    scan_blocks(ctx, bsum, nb, d_total);
    DLG_LAUNCH(ctx, DLG_K_SCAN, (sizeof(T) + 8.0) * n, scan_apply_k<T>, (unsigned)nb, SCAN_T, 0, in, n,
               (const u64 *)bsum, out);
}

template <typename T>
static void scan_to_u64_async(dlg_ctx *ctx, const T *in, u64 n, u64 *out, u64 *d_total) {
    DBuf<u64> bsum(ctx, scan_partials(n));
    scan_to_u64_with<T>(ctx, in, n, out, bsum.p, d_total);
}

void scan_u32_u64_with(dlg_ctx *ctx, const u32 *in, u64 n, u64 *out, u64 *bsum, u64 *d_total) {
    scan_to_u64_with<u32>(ctx, in, n, out, bsum, d_total);
}

void scan_u32_u64_async(dlg_ctx *ctx, const u32 *in, u64 n, u64 *out, u64 *d_total) {
    scan_to_u64_async<u32>(ctx, in, n, out, d_total);
}

void scan_u64_async(dlg_ctx *ctx, const u64 *in, u64 n, u64 *out, u64 *d_total) {
    scan_to_u64_async<u64>(ctx, in, n, out, d_total);
}

u64 scan_u32_u64(dlg_ctx *ctx, const u32 *in, u64 n, u64 *out) {
    DBuf<u64> tot(ctx, 1);
    scan_u32_u64_async(ctx, in, n, out, tot.p);
    u64 h = 0;
    ctx->read_scalars(tot.p, 1, &h);
    return h;
}

// Exclusive scan of u32 counts into u32 offsets (radix digit counts; sums < 2^32).
__global__ __launch_bounds__(SCAN_T) void scan_reduce32_k(const u32 *in, u64 n, u64 *bsum) {
    __shared__ u64 sc[SCAN_T / WAVE + 1];
    u64 base = (u64)blockIdx.x * SCAN_TILE + (u64)threadIdx.x * SCAN_I;
    u64 s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_I; i++)
        if (base + i < n) s += in[base + i];
    u64 tot;
    block_exclusive_scan<u64>(s, sc, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}
__global__ __launch_bounds__(SCAN_T) void scan_apply32_k(u32 *data, u64 n, const u64 *bpre) {
    __shared__ u64 sc[SCAN_T / WAVE + 1];
    u64 base = (u64)blockIdx.x * SCAN_TILE + (u64)threadIdx.x * SCAN_I;
    u32 v[SCAN_I];
    u64 s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_I; i++) {
        v[i] = base + i < n ? data[base + i] : 0;
        s += v[i];
    }
    u64 ex = block_exclusive_scan<u64>(s, sc, nullptr) + bpre[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SCAN_I; i++) {
        if (base + i < n) data[base + i] = (u32)ex;
        ex += v[i];
    }
}
static void scan_u32_inplace(dlg_ctx *ctx, u32 *data, u64 n, DBuf<u64> &bsum) {
    u64 nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nb == 0) nb = 1;
    bsum.ensure(nb);
    DLG_LAUNCH(ctx, DLG_K_SCAN, 4.0 * n, scan_reduce32_k, (unsigned)nb, SCAN_T, 0, data, n,
               bsum.p);
    scan_blocks(ctx, bsum.p, nb, nullptr);
    DLG_LAUNCH(ctx, DLG_K_SCAN, 8.0 * n, scan_apply32_k, (unsigned)nb, SCAN_T, 0, data, n,
               bsum.p);
}

// =============================================================== radix sort
// Tile shapes: 256 threads x 16 keys (4096) for small inputs, 512 threads x
// 32 keys (16384, 128 KiB of LDS staging) for large ones, where longer runs
// per digit make the scatter's writes closer to whole cache lines.
constexpr int RS_SMALL_T = 256, RS_SMALL_I = 16;
constexpr int RS_BIG_T = 512, RS_BIG_I = 32;
constexpr u64 RS_BIG_MIN = 1ull << 21;

struct RadixKey {
    u32 min_hi, min_lo;
    u32 vb;     // bits of (lo - min_lo)
    u32 shift;  // digit position in the compacted key
};

// vb = 0: the value half takes no bits -- every value equal (then lo - min_lo
// is 0 anyway), or a sort by key alone (radix_sort's key_only).
__device__ __forceinline__ u32 radix_digit(u64 x, const RadixKey &k, u32 mask) {
    u64 hi = (u64)(tkey(x) - k.min_hi);
    u64 lo = k.vb ? (u64)(tval(x) - k.min_lo) : 0ull;
    u64 c = (hi << k.vb) | lo;  // vb <= 32, (hi bits + vb) <= 64
    return (u32)(c >> k.shift) & mask;
}

// The same 8-bit digit in 32-bit operations (the branches are uniform): from
// the value half when it lies there, from the key half when it lies there,
// from both when it straddles them.  shift - vb < 32 (the key half has at
// most 32 bits); shift < vb <= 32 otherwise.
__device__ __forceinline__ u32 radix_digit8_hl(u32 hi, u32 lo, const RadixKey &k) {
    if (k.shift >= k.vb) return __builtin_amdgcn_ubfe(hi, k.shift - k.vb, 8);
    u32 d = lo >> k.shift;
    if (k.vb - k.shift < 8) d |= hi << (k.vb - k.shift);
    return d & 0xffu;
}
__device__ __forceinline__ u32 radix_digit8(u64 x, const RadixKey &k) {
    return radix_digit8_hl(tkey(x) - k.min_hi, tval(x) - k.min_lo, k);
}

// min/max of the key and value halves: out[0]=min_hi out[1]=max_hi out[2]=min_lo out[3]=max_lo.
// Each workgroup folds a grid-stride share through the wave and LDS and
// stores its four partials (part[4 b ..]); minmax_final_k folds the
// workgroups' partials.  Device atomics on the four shared words serialised:
// ~100 µs for 2,048 workgroups.
__device__ __forceinline__ void minmax_block(u32 &mnh, u32 &mxh, u32 &mnl, u32 &mxl) {
#pragma unroll
    for (int d = WAVE / 2; d > 0; d >>= 1) {
        mnh = min(mnh, (u32)__shfl_xor((int)mnh, d, WAVE));
        mxh = max(mxh, (u32)__shfl_xor((int)mxh, d, WAVE));
        mnl = min(mnl, (u32)__shfl_xor((int)mnl, d, WAVE));
        mxl = max(mxl, (u32)__shfl_xor((int)mxl, d, WAVE));
    }
    __shared__ u32 sm[4][256 / WAVE];
    const int wave = threadIdx.x / WAVE;
    if ((threadIdx.x & (WAVE - 1)) == 0) {
        sm[0][wave] = mnh; sm[1][wave] = mxh; sm[2][wave] = mnl; sm[3][wave] = mxl;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x / WAVE); w++) {
            mnh = min(mnh, sm[0][w]); mxh = max(mxh, sm[1][w]);
            mnl = min(mnl, sm[2][w]); mxl = max(mxl, sm[3][w]);
        }
    }
}

__global__ __launch_bounds__(256) void minmax_k(const u64 *in, u64 n, u32 *part) {
    u32 mnh = 0xffffffffu, mxh = 0, mnl = 0xffffffffu, mxl = 0;
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        u64 x = in[i];
        u32 h = tkey(x), l = tval(x);
        mnh = min(mnh, h); mxh = max(mxh, h); mnl = min(mnl, l); mxl = max(mxl, l);
    }
    minmax_block(mnh, mxh, mnl, mxl);
    if (threadIdx.x == 0) {
        u32 *p = part + 4 * (u64)blockIdx.x;
        p[0] = mnh; p[1] = mxh; p[2] = mnl; p[3] = mxl;
    }
}

// One workgroup: the fold of nb workgroups' partials (none: the empty bounds).
__global__ __launch_bounds__(256) void minmax_final_k(const u32 *part, u32 nb, u32 *out) {
    u32 mnh = 0xffffffffu, mxh = 0, mnl = 0xffffffffu, mxl = 0;
    for (u32 b = threadIdx.x; b < nb; b += blockDim.x) {
        const u32 *p = part + 4 * (u64)b;
        mnh = min(mnh, p[0]); mxh = max(mxh, p[1]); mnl = min(mnl, p[2]); mxl = max(mxl, p[3]);
    }
    minmax_block(mnh, mxh, mnl, mxl);
    if (threadIdx.x == 0) {
        out[0] = mnh; out[1] = mxh; out[2] = mnl; out[3] = mxl;
    }
}

template <int BITS, int T, int I>
__global__ __launch_bounds__(T) void rs_upsweep_k(const u64 *in, u64 n, RadixKey key, u32 *counts,
                                                  u32 nb) {
    constexpr int BINS = 1 << BITS, W = T / WAVE, TILE = T * I;
    __shared__ u32 h[W][BINS];
    for (int i = threadIdx.x; i < W * BINS; i += T) (&h[0][0])[i] = 0;
    __syncthreads();
    const int wave = threadIdx.x / WAVE;
    u64 base = (u64)blockIdx.x * TILE;
#pragma unroll 4
    for (int r = 0; r < I; r++) {
        u64 idx = base + (u64)r * T + threadIdx.x;
        if (idx < n) atomicAdd(&h[wave][radix_digit(in[idx], key, BINS - 1)], 1u);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < BINS; d += T) {
        u32 s = 0;
#pragma unroll
        for (int w = 0; w < W; w++) s += h[w][d];
        counts[(u64)d * nb + blockIdx.x] = s;
    }
}

// Stable scatter of one tile: wave-match ranking (one ballot per digit bit),
// per-wave running counts in LDS, keys staged in LDS in digit order, then
// written to offs[digit][tile] + rank.  Wave w owns keys
// [tile0 + w*64*I, tile0 + (w+1)*64*I), consumed 64 at a time in order, so
// the (wave, round, lane) order is the input order and the sort is stable.
template <int BITS, int T, int I>
__global__ __launch_bounds__(T) void rs_downsweep_k(const u64 *in, u64 *out, u64 n, RadixKey key,
                                                    const u32 *offs, u32 nb) {
    constexpr int BINS = 1 << BITS, W = T / WAVE, TILE = T * I;
    __shared__ u64 stage[TILE];
    __shared__ u32 wcnt[W][BINS];
    __shared__ u32 bstart[BINS];
    __shared__ u32 gofs[BINS];
    __shared__ u32 sc[T / WAVE + 1];

    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
    for (int i = threadIdx.x; i < W * BINS; i += T) (&wcnt[0][0])[i] = 0;
    for (int d = threadIdx.x; d < BINS; d += T) gofs[d] = offs[(u64)d * nb + blockIdx.x];
    __syncthreads();

    const u64 tile0 = (u64)blockIdx.x * TILE;
    const u64 wbase = tile0 + (u64)wave * (WAVE * I);
    const u64 lt = lanemask_lt();
    u64 keys[I];
    u32 pk[I];  // digit | (rank within the wave's digit run << 8)

#pragma unroll
    for (int r = 0; r < I; r++) {
        u64 idx = wbase + (u64)r * WAVE + lane;
        keys[r] = idx < n ? in[idx] : 0ull;
    }
#pragma unroll
    for (int r = 0; r < I; r++) {
        u64 idx = wbase + (u64)r * WAVE + lane;
        bool valid = idx < n;
        u32 d = radix_digit(keys[r], key, BINS - 1);
        u64 m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < BITS; b++) {
            bool bit = (d >> b) & 1u;
            u64 bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        u32 before = __popcll(m & lt);
        u32 cnt = wcnt[wave][d];
        // The lowest lane of each digit group publishes the new running count;
        // every lane of the wave has read the old value in the same LDS op.
        if (valid && before == 0) wcnt[wave][d] = cnt + (u32)__popcll(m);
        pk[r] = d | ((cnt + before) << 8);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();

    // per digit: exclusive prefix over waves, then over digits
    u32 tot = 0;
    if (threadIdx.x < BINS) {
        u32 s = 0;
#pragma unroll
        for (int w = 0; w < W; w++) {
            u32 c = wcnt[w][threadIdx.x];
            wcnt[w][threadIdx.x] = s;
            s += c;
        }
        tot = s;
    }
    u32 ex = block_exclusive_scan<u32>(tot, sc, nullptr);
    if (threadIdx.x < BINS) bstart[threadIdx.x] = ex;
    __syncthreads();

#pragma unroll
    for (int r = 0; r < I; r++) {
        u64 idx = wbase + (u64)r * WAVE + lane;
        if (idx < n) {
            u32 d = pk[r] & 0xffu;
            stage[bstart[d] + wcnt[wave][d] + (pk[r] >> 8)] = keys[r];
        }
    }
    __syncthreads();

    u64 rem = n - tile0;
    u32 tn = rem < (u64)TILE ? (u32)rem : (u32)TILE;
    for (u32 i = threadIdx.x; i < tn; i += T) {
        u64 x = stage[i];
        u32 d = radix_digit(x, key, BINS - 1);
        out[(u64)gofs[d] + (i - bstart[d])] = x;
    }
}

static int bit_width(u32 v) { return v ? 32 - __builtin_clz(v) : 0; }

template <int BITS, int T, int I>
static void radix_passes(dlg_ctx *ctx, u64 *&src, u64 *&dst, u64 n, RadixKey key, int total_bits) {
    constexpr int BINS = 1 << BITS, TILE = T * I;
    u32 nb = (u32)((n + TILE - 1) / TILE);
    DBuf<u32> counts(ctx, (size_t)BINS * nb);
    DBuf<u64> bsum(ctx);
    for (int shift = 0; shift < total_bits; shift += BITS) {
        key.shift = (u32)shift;
        DLG_LAUNCH(ctx, DLG_K_SORT_UPSWEEP, 8.0 * n + 4.0 * BINS * nb, (rs_upsweep_k<BITS, T, I>), nb,
                   T, 0, (const u64 *)src, n, key, counts.p, nb);
        scan_u32_inplace(ctx, counts.p, (u64)BINS * nb, bsum);
        DLG_LAUNCH(ctx, DLG_K_SORT_DOWNSWEEP, 16.0 * n + 4.0 * BINS * nb, (rs_downsweep_k<BITS, T, I>),
                   nb, T, 0, (const u64 *)src, dst, n, key, (const u32 *)counts.p, nb);
        std::swap(src, dst);
    }
}

// ---------------------------------------------------------- one-sweep passes
// For large inputs with 8-bit digits every pass reads its keys once: one
// histogram kernel counts the digits of all passes in a single read, and
// each pass is one kernel whose tiles take tickets in launch order, rank
// their keys as rs_downsweep_k does, and find where their digit runs go by
// decoupled look-back: a tile publishes its digit counts (AGGREGATE), adds
// the counts of the tiles before it, walking back until a tile that has
// published its inclusive prefix (PREFIX), and publishes its own prefix.  A
// tile only waits for tiles with smaller tickets, which are all running or
// done, so every wait ends; a bounded spin sets an error flag instead of
// hanging should that ever fail.  Status words: tag (8 bits: 2 * pass + 1 for
// AGGREGATE, + 2 for PREFIX, so one zeroed array serves every pass) and
// a 56-bit count.
// Tiles of 512 x 16 keys (64 KiB of LDS staging, two workgroups per CU): on
// 3.4e8 keys 1.82 ms per pass against 2.04 ms for 512 x 32, 2.49 ms for
// 256 x 16 and 1.95-2.53 ms for 256 x 32 (two runs).
constexpr int OS_T = 512, OS_I = 16, OS_TILE = OS_T * OS_I;
constexpr u32 OS_SPIN_LIMIT = 1u << 22;
constexpr u64 OS_VALUE = (1ull << 56) - 1;

// Digits of every pass, counted in one LDS histogram per wave (8 passes x
// 256 bins x 4 waves = 32 KiB), then added into the global counts.  Lanes
// that hit one LDS address in one atomic are serialised, and clustered input
// (sorted runs, hub keys) puts a whole wave on one digit: a wave whose lanes
// share the digit adds its count with one atomic, other waves one per lane.
// A pass whose digits span fewer than 8 bits (the top one) has few distinct
// digits: its count goes through ballot matching, one atomic per distinct
// digit per wave.  Each thread loads OH_I keys of a tile before counting
// any.  The kernel is bound by its vector instructions, so the digits come
// from 32-bit bit-field extracts of the two halves.
constexpr int OH_T = 256, OH_I = 8, OH_TILE = OH_T * OH_I;

__global__ __launch_bounds__(OH_T) void os_hist_k(const u64 *in, u64 n, RadixKey key, u32 passes, u32 top_bits,
                                                  unsigned long long *hist) {
    constexpr int BINS = 256, W = OH_T / WAVE;
    __shared__ u32 h[W][8][BINS];
    for (int i = threadIdx.x; i < W * 8 * BINS; i += OH_T) (&h[0][0][0])[i] = 0;
    __syncthreads();
    const int wave = threadIdx.x / WAVE;
    const u32 p = passes - 1;
    for (u64 base = (u64)blockIdx.x * OH_TILE; base < n; base += (u64)gridDim.x * OH_TILE) {   // block-uniform
        u64 x[OH_I];
#pragma unroll
        for (int r = 0; r < OH_I; r++) {
            const u64 i = base + (u64)r * OH_T + threadIdx.x;
            x[r] = i < n ? in[i] : 0ull;
        }
#pragma unroll
        for (int r = 0; r < OH_I; r++) {
            const bool valid = base + (u64)r * OH_T + threadIdx.x < n;   // valid lanes: a prefix of the wave
            const u64 vm = __ballot(valid);
            if (vm == 0) continue;                                     // wave-uniform
            const u32 hi = tkey(x[r]) - key.min_hi, lo = tval(x[r]) - key.min_lo;
            const u32 nvalid = (u32)__popcll(vm);
            RadixKey kq = key;
            for (u32 q = 0; q < p; q++) {
                kq.shift = 8 * q;
                const u32 dq = radix_digit8_hl(hi, lo, kq);
                const u32 d0 = __builtin_amdgcn_readfirstlane(dq);   // lane 0: valid
                if ((__ballot(dq == d0) & vm) == vm) {
                    if (threadIdx.x % WAVE == 0) atomicAdd(&h[wave][q][d0], nvalid);
                } else if (valid) {
                    atomicAdd(&h[wave][q][dq], 1u);
                }
            }
            kq.shift = 8 * p;
            const u32 d = radix_digit8_hl(hi, lo, kq), nd = ~d;
            u32 mlo = (u32)vm, mhi = (u32)(vm >> 32);
            for (u32 b = 0; b < top_bits; b++) {
                const u32 clr = (u32)__builtin_amdgcn_sbfe((int)nd, b, 1);
                const u64 set = __ballot(clr == 0u);
                mlo &= (u32)set ^ clr;
                mhi &= (u32)(set >> 32) ^ clr;
            }
            const u32 below = __builtin_amdgcn_mbcnt_hi(mhi, __builtin_amdgcn_mbcnt_lo(mlo, 0u));
            if (valid && below == 0) atomicAdd(&h[wave][p][d], (u32)__popc(mlo) + (u32)__popc(mhi));
        }
    }
    __syncthreads();
    for (u32 i = threadIdx.x; i < passes * BINS; i += OH_T) {
        u32 v = 0;   // a workgroup's share of one digit fits 32 bits
#pragma unroll
        for (int w = 0; w < W; w++) v += (&h[w][0][0])[i];
        if (v) atomicAdd(&hist[i], (unsigned long long)v);
    }
}

// Exclusive scan of each pass's digit counts into its digit starts (64-bit:
// one buffer may hold more than 2^32 keys).
__global__ __launch_bounds__(256) void os_starts_k(const u64 *hist, u32 passes, u64 *start) {
    __shared__ u64 sc[256 / WAVE + 1];
    for (u32 p = 0; p < passes; p++) {
        const u64 v = hist[p * 256 + threadIdx.x];
        start[p * 256 + threadIdx.x] = block_exclusive_scan<u64>(v, sc, nullptr);
        __syncthreads();   // sc is reused by the next pass
    }
}

__device__ __forceinline__ u64 os_load(const u64 *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void os_store(u64 *p, u64 v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int T, int I>
__device__ __forceinline__ bool os_load_tile(const u64 *in, u64 n, u64 tile0, u64 wbase, int lane, u64 (&keys)[I]) {
    constexpr int TILE = T * I;
    const bool full = tile0 + TILE <= n;   // every tile but the last: no per-key bounds checks
    if (full) {
#pragma unroll
        for (int r = 0; r < I; r++) keys[r] = in[wbase + (u64)r * WAVE + lane];
    } else {
#pragma unroll
        for (int r = 0; r < I; r++) {
            const u64 idx = wbase + (u64)r * WAVE + lane;
            keys[r] = idx < n ? in[idx] : 0ull;
        }
    }
    return full;
}

template <int T, int I>
__global__ __launch_bounds__(T) void os_pass_k(const u64 *in, u64 *out, u64 n, RadixKey key, const u64 *dstart,
                                               u64 *status, u32 *ticket, u32 pass, u32 *err) {
    constexpr int BINS = 256, W = T / WAVE, TILE = T * I;
    __shared__ u64 stage[TILE];
    __shared__ u32 wcnt[W][BINS];
    __shared__ u32 bstart[BINS];
    __shared__ u64 gofs[BINS];   // output position of stage slot i of digit d: gofs[d] + i (mod 2^64)
    __shared__ u32 sc[T / WAVE + 1];
    __shared__ u32 s_tile;
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    for (int i = threadIdx.x; i < W * BINS; i += T) (&wcnt[0][0])[i] = 0;
    __syncthreads();
    const u32 tile = s_tile;
    const u64 tile0 = (u64)tile * TILE;
    const u64 wbase = tile0 + (u64)wave * (WAVE * I);
    u64 keys[I];
    u32 pk[I];   // digit | (rank within the wave's digit run << 8)
    const bool full = os_load_tile<T, I>(in, n, tile0, wbase, lane, keys);
    // Lanes of equal digit by ballot matching: per bit, clr = -1 when the lane's
    // bit is clear, the ballot of clr == 0 is the lanes with it set, and the
    // lane keeps (ballot ^ clr) -- the lanes that agree with it on that bit.
#pragma unroll
    for (int r = 0; r < I; r++) {
        const bool valid = full || wbase + (u64)r * WAVE + lane < n;
        const u32 d = radix_digit8(keys[r], key), nd = ~d;
        const u64 vm = full ? ~0ull : __ballot(valid);
        u32 mlo = (u32)vm, mhi = (u32)(vm >> 32);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const u32 clr = (u32)__builtin_amdgcn_sbfe((int)nd, b, 1);
            const u64 set = __ballot(clr == 0u);
            mlo &= (u32)set ^ clr;
            mhi &= (u32)(set >> 32) ^ clr;
        }
        const u32 before = __builtin_amdgcn_mbcnt_hi(mhi, __builtin_amdgcn_mbcnt_lo(mlo, 0u));
        const u32 cnt = wcnt[wave][d];
        if (valid && before == 0) wcnt[wave][d] = cnt + (u32)__popc(mlo) + (u32)__popc(mhi);
        pk[r] = d | ((cnt + before) << 8);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    u32 tot = 0;   // the tile's count of digit threadIdx.x; wcnt becomes the exclusive prefix over waves
    if (threadIdx.x < BINS) {
        u32 s = 0;
#pragma unroll
        for (int w = 0; w < W; w++) {
            const u32 c = wcnt[w][threadIdx.x];
            wcnt[w][threadIdx.x] = s;
            s += c;
        }
        tot = s;
    }
    const u32 ex = block_exclusive_scan<u32>(tot, sc, nullptr);
    if (threadIdx.x < BINS) {
        const u32 d = threadIdx.x;
        bstart[d] = ex;
        const u64 agg = (u64)(2 * pass + 1) << 56, pre = (u64)(2 * pass + 2) << 56;
        u64 *mine = status + (u64)tile * BINS + d;
        u64 excl = 0;
        if (tile == 0) {
            os_store(mine, pre | tot);
        } else {
            os_store(mine, agg | tot);
            for (u64 t = tile; t-- > 0;) {
                u64 v = 0;
                for (u32 spin = 0;; spin++) {
                    v = os_load(status + t * BINS + d);
                    const u64 tag = v >> 56;
                    if (tag == 2 * pass + 1 || tag == 2 * pass + 2) break;
                    if (spin >= OS_SPIN_LIMIT) {   // never expected: give up on this digit, flag the sort
                        atomicOr(err, 1u);
                        v = pre;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                excl += v & OS_VALUE;
                if ((v >> 56) != 2 * pass + 1) break;   // an inclusive prefix (or the flagged give-up)
            }
            os_store(mine, pre | (excl + tot));
        }
        gofs[d] = dstart[d] + excl - (u64)ex;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < I; r++) {
        if (full || wbase + (u64)r * WAVE + lane < n) {
            const u32 d = pk[r] & 0xffu;
            stage[bstart[d] + wcnt[wave][d] + (pk[r] >> 8)] = keys[r];
        }
    }
    __syncthreads();
    const u64 rem = n - tile0;
    const u32 tn = rem < (u64)TILE ? (u32)rem : (u32)TILE;
    for (u32 i = threadIdx.x; i < tn; i += T) {
        const u64 x = stage[i];
        out[gofs[radix_digit8(x, key)] + (u64)i] = x;
    }
}

// All passes of an 8-bit sort: histograms of every pass in one read, then one
// kernel per pass (one read, one write of the keys).
static void radix_onesweep(dlg_ctx *ctx, u64 *&src, u64 *&dst, u64 n, RadixKey key, int total_bits) {
    constexpr int OT = OS_T, OI = OS_I, OTILE = OS_TILE;
    const u32 passes = (u32)((total_bits + 7) / 8);
    const u64 tiles = (n + OTILE - 1) / OTILE;
    if (tiles > 0x7fffffffull) fail(DLG_OVERFLOW, "radix sort: too many tiles");
    DBuf<u64> hist(ctx, 8 * 256 * 2);   // histograms, digit starts (64-bit)
    DBuf<u32> tick(ctx, 16);             // tickets, error flag
    DBuf<u64> status(ctx, tiles * 256);
    u64 *starts = hist.p + 8 * 256;
    u32 *tickets = tick.p, *err = tickets + 8;
    fill_u64(ctx, hist.p, 8 * 256 * 2, 0ull);
    fill_u32(ctx, tick.p, 16, 0u);
    fill_u64(ctx, status.p, tiles * 256, 0ull);
    key.shift = 0;
    const u32 top_bits = (u32)(total_bits - 8 * (int)(passes - 1));
    // 4 blocks per CU (32 KiB of LDS each), each looping over tiles
    const unsigned hg = (unsigned)std::min<u64>((n + OH_TILE - 1) / OH_TILE, (u64)ctx->num_cus * 4);
    DLG_LAUNCH(ctx, DLG_K_SORT_UPSWEEP, 8.0 * n, os_hist_k, hg, OH_T, 0, (const u64 *)src, n, key, passes, top_bits,
               (unsigned long long *)hist.p);
    DLG_LAUNCH(ctx, DLG_K_SORT_UPSWEEP, 8.0 * 256 * passes, os_starts_k, 1, 256, 0, (const u64 *)hist.p, passes,
               starts);
    for (u32 p = 0; p < passes; p++) {
        key.shift = 8 * p;
        DLG_LAUNCH(ctx, DLG_K_SORT_DOWNSWEEP, 16.0 * n, (os_pass_k<OT, OI>), (unsigned)tiles, OT, 0,
                   (const u64 *)src, dst, n, key, (const u64 *)(starts + 256 * p), status.p, tickets + p, p, err);
        std::swap(src, dst);
    }
    u32 e = 0;
    HIPCHK(hipMemcpyAsync(ctx->h_scalars, err, 4, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    e = *(u32 *)ctx->h_scalars;
    if (e) fail(DLG_DEVICE_ERROR, "radix sort: a tile's look-back did not complete");
}

template <int BITS>
static void radix_passes_sized(dlg_ctx *ctx, u64 *&src, u64 *&dst, u64 n, RadixKey key, int total_bits) {
    if (BITS == 8 && n >= RS_BIG_MIN)
        radix_onesweep(ctx, src, dst, n, key, total_bits);
    else if (n >= RS_BIG_MIN)
        radix_passes<BITS, RS_BIG_T, RS_BIG_I>(ctx, src, dst, n, key, total_bits);
    else
        radix_passes<BITS, RS_SMALL_T, RS_SMALL_I>(ctx, src, dst, n, key, total_bits);
}

void tuple_bounds(dlg_ctx *ctx, const u64 *data, u64 n, u32 *h) {
    const unsigned g = n ? (unsigned)std::min<u64>((n + 255) / 256, (u64)ctx->num_cus * 8) : 0u;
    DBuf<u32> mm(ctx, 4 + 4 * (u64)g);   // the bounds, then the workgroups' partials
    if (g) DLG_LAUNCH(ctx, DLG_K_SORT_HIST, 8.0 * n, minmax_k, g, 256, 0, data, n, mm.p + 4);
    DLG_LAUNCH(ctx, DLG_K_MISC, 16.0 * g, minmax_final_k, 1, 256, 0, (const u32 *)(mm.p + 4), g, mm.p);
    HIPCHK(hipMemcpyAsync(ctx->h_scalars, mm.p, 16, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    std::memcpy(h, ctx->h_scalars, 16);
}

// Per block: the digest sum and the order breaks of a grid-stride share,
// reduced through the wave and LDS, then one atomic each.
__global__ __launch_bounds__(256) void rows_digest_k(const u64 *d, u64 n, int orient, u32 sub0, u32 sub1,
                                                     unsigned long long *acc) {
    __shared__ u64 s_sum[256 / WAVE], s_brk[256 / WAVE];
    u64 sum = 0, brk = 0;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        u64 t = d[i];
        u32 k = tkey(t) ^ 0x80000000u, v = tval(t) ^ 0x80000000u;   // the int32 bits
        u64 r0 = (u32)((orient == DLG_ORIENT_KV ? k : v) - sub0), r1 = (u32)((orient == DLG_ORIENT_KV ? v : k) - sub1);
        sum += r0 * 0x9E3779B97F4A7C15ull + r1 * 0xC2B2AE3D27D4EB4Full + r0 * r1 + 0x632BE59BD9B4E019ull;
        if (i + 1 < n && t >= d[i + 1]) brk++;
    }
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        sum += __shfl_xor(sum, o, WAVE);
        brk += __shfl_xor(brk, o, WAVE);
    }
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    if (lane == 0) {
        s_sum[w] = sum;
        s_brk[w] = brk;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int j = 1; j < (int)(blockDim.x / WAVE); j++) {
            sum += s_sum[j];
            brk += s_brk[j];
        }
        atomicAdd(acc, (unsigned long long)sum);
        atomicAdd(acc + 1, (unsigned long long)brk);
    }
}

void rows_digest(dlg_ctx *ctx, const u64 *d, u64 n, int orient, u64 *sum, u64 *breaks, u32 sub0, u32 sub1) {
    DBuf<u64> acc(ctx, 2);
    fill_u64(ctx, acc.p, 2, 0ull);
    if (n) {
        unsigned g = (unsigned)std::min<u64>((n + 255) / 256, (u64)ctx->num_cus * 8);
        DLG_LAUNCH(ctx, DLG_K_MISC, 8.0 * n, rows_digest_k, g, 256, 0, d, n, orient, sub0, sub1,
                   (unsigned long long *)acc.p);
    }
    u64 h[2];
    ctx->read_scalars(acc.p, 2, h);
    *sum = h[0];
    *breaks = h[1];
}

void radix_sort(dlg_ctx *ctx, DBuf<u64> &data, u64 n, DBuf<u64> &tmp, int bits, const u32 *bounds,
                bool key_only) {
    if (n < 2) return;
    if (bits <= 0) bits = ctx->radix_bits;
    if (bits < 4) bits = 4;
    if (bits > 8) bits = 8;
    // The 8-bit onesweep counts digits in 64 bits; the narrower digit passes
    // keep 32-bit digit offsets.
    if (n > 0xffffffffull && bits != 8)
        fail(DLG_OVERFLOW, "radix_sort: more than 2^32-1 tuples need 8-bit digits");
    u32 h[4];
    if (bounds) std::memcpy(h, bounds, sizeof(h));
    else tuple_bounds(ctx, data.p, n, h);
    RadixKey key;
    key.min_hi = h[0];
    key.min_lo = h[2];
    key.vb = key_only ? 0u : (u32)bit_width(h[3] - h[2]);   // key_only: a stable sort by key alone
    key.shift = 0;
    int total_bits = (int)key.vb + bit_width(h[1] - h[0]);
    if (total_bits == 0) return;  // all equal
    tmp.ensure(n);
    u64 *src = data.p, *dst = tmp.p;
    switch (bits) {
        case 4: radix_passes_sized<4>(ctx, src, dst, n, key, total_bits); break;
        case 5: radix_passes_sized<5>(ctx, src, dst, n, key, total_bits); break;
        case 6: radix_passes_sized<6>(ctx, src, dst, n, key, total_bits); break;
        case 7: radix_passes_sized<7>(ctx, src, dst, n, key, total_bits); break;
        default: radix_passes_sized<8>(ctx, src, dst, n, key, total_bits); break;
    }
    if (src != data.p) data.swap(tmp);   // result is in the other buffer: swap ownership
}

void radix_sort(dlg_ctx *ctx, u64 *data, u64 n, DBuf<u64> &tmp, int bits, bool key_only) {
    DBuf<u64> view(ctx);
    view.p = data;          // borrowed: released without returning it to the pool
    view.cap = n;
    radix_sort(ctx, view, n, tmp, bits, nullptr, key_only);
    if (view.p != data) {   // sorted copy ended in tmp's old block; bring it back
        copy_u64(ctx, data, view.p, n);
        tmp.swap(view);     // tmp owns that block again
    }
    view.p = nullptr;
    view.cap = 0;
}

// =============================================================== select
constexpr int SEL_T = 256, SEL_I = 8, SEL_TILE = SEL_T * SEL_I;

__device__ __forceinline__ bool found_sorted(const u64 *a, u64 lo, u64 hi, u64 x) {
    u64 p = lower_bound_u64(a, lo, hi, x);
    return p < hi && a[p] == x;
}

template <int MODE>
__global__ __launch_bounds__(SEL_T) void select_count_k(const u64 *in, u64 n, const u64 *other,
                                                        u64 no, u64 *flags, u64 *bsum) {
    __shared__ u64 s_lo, s_hi;
    __shared__ u64 sc[SEL_T / WAVE + 1];
    const u64 base = (u64)blockIdx.x * SEL_TILE;
    constexpr bool NEED_OTHER = (MODE == SEL_UNIQUE_NOT_IN || MODE == SEL_NOT_IN);
    if (NEED_OTHER) {
        if (threadIdx.x < WAVE) {   // the first wave finds the tile's range of `other`
            const u64 last = base + SEL_TILE - 1 < n ? base + SEL_TILE - 1 : n - 1;
            const u64 lo = wave_lower_bound(other, 0, no, in[base]);
            const u64 hi = wave_upper_bound(other, lo, no, in[last]);
            if (threadIdx.x == 0) {
                s_lo = lo;
                s_hi = hi;
            }
        }
        __syncthreads();
    }
    u64 c = 0;
#pragma unroll
    for (int r = 0; r < SEL_I; r++) {
        u64 idx = base + (u64)r * SEL_T + threadIdx.x;
        bool keep = false;
        if (idx < n) {
            u64 x = in[idx];
            if (MODE == SEL_KEY_HEAD) {
                keep = idx == 0 || tkey(in[idx - 1]) != tkey(x);
            } else if (MODE == SEL_NOT_IN) {
                keep = !found_sorted(other, s_lo, s_hi, x);
            } else if (MODE == SEL_NODIAG) {
                keep = tkey(x) != tval(x);
            } else {
                keep = idx == 0 || in[idx - 1] != x;
                if (MODE == SEL_UNIQUE_NODIAG) keep = keep && tkey(x) != tval(x);
                if (MODE == SEL_UNIQUE_NOT_IN && keep) keep = !found_sorted(other, s_lo, s_hi, x);
            }
        }
        u64 m = __ballot(keep);
        u64 w0 = base + (u64)r * SEL_T + (threadIdx.x & ~(WAVE - 1));
        if ((threadIdx.x & (WAVE - 1)) == 0 && w0 < n) flags[w0 / WAVE] = m;
        c += keep ? 1 : 0;
    }
    u64 tot;
    block_exclusive_scan<u64>(c, sc, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

template <bool OUT_INDEX>
__global__ __launch_bounds__(SEL_T) void select_scatter_k(const u64 *in, u64 n, const u64 *flags,
                                                          const u64 *bpre, void *out) {
    __shared__ u32 wc[SEL_I][SEL_T / WAVE];
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
    const u64 base = (u64)blockIdx.x * SEL_TILE;
    u64 words[SEL_I];
#pragma unroll
    for (int r = 0; r < SEL_I; r++) {
        u64 w0 = base + (u64)r * SEL_T + (u64)wave * WAVE;
        words[r] = w0 < n ? flags[w0 / WAVE] : 0ull;
        if (lane == 0) wc[r][wave] = (u32)__popcll(words[r]);
    }
    __syncthreads();
    u64 pre = bpre[blockIdx.x];
    const u64 lt = lanemask_lt();
#pragma unroll
    for (int r = 0; r < SEL_I; r++) {
        u32 p = 0;
        for (int rr = 0; rr < SEL_I; rr++)
            for (int ww = 0; ww < SEL_T / WAVE; ww++)
                if (rr < r || (rr == r && ww < wave)) p += wc[rr][ww];
        if ((words[r] >> lane) & 1ull) {
            u64 idx = base + (u64)r * SEL_T + threadIdx.x;
            u64 pos = pre + p + (u64)__popcll(words[r] & lt);
            if (OUT_INDEX) ((u32 *)out)[pos] = (u32)idx;
            else ((u64 *)out)[pos] = in[idx];
        }
    }
}

u64 select(dlg_ctx *ctx, int mode, const u64 *in, u64 n, const u64 *other, u64 n_other,
           void *out) {
    if (n == 0) return 0;
    if ((mode == SEL_NOT_IN || mode == SEL_UNIQUE_NOT_IN) && n_other == 0)
        mode = mode == SEL_NOT_IN ? -1 : SEL_UNIQUE;
    if (mode == -1) {  // nothing to remove
        if (out != (const void *)in) copy_u64(ctx, (u64 *)out, in, n);
        return n;
    }
    u64 nb = (n + SEL_TILE - 1) / SEL_TILE;
    DBuf<u64> flags(ctx, (n + WAVE - 1) / WAVE + 1);
    DBuf<u64> bsum(ctx, nb);
    DBuf<u64> tot(ctx, 1);
    double rb = 8.0 * n + 8.0 * ((n + 63) / 64);
    switch (mode) {
        case SEL_UNIQUE:
            DLG_LAUNCH(ctx, DLG_K_SELECT, rb, select_count_k<SEL_UNIQUE>, (unsigned)nb, SEL_T, 0,
                       in, n, other, n_other, flags.p, bsum.p);
            break;
        case SEL_UNIQUE_NODIAG:
            DLG_LAUNCH(ctx, DLG_K_SELECT, rb, select_count_k<SEL_UNIQUE_NODIAG>, (unsigned)nb,
                       SEL_T, 0, in, n, other, n_other, flags.p, bsum.p);
            break;
        case SEL_UNIQUE_NOT_IN:
            DLG_LAUNCH(ctx, DLG_K_SELECT, rb, select_count_k<SEL_UNIQUE_NOT_IN>, (unsigned)nb,
                       SEL_T, 0, in, n, other, n_other, flags.p, bsum.p);
            break;
        case SEL_NOT_IN:
            DLG_LAUNCH(ctx, DLG_K_SELECT, rb, select_count_k<SEL_NOT_IN>, (unsigned)nb, SEL_T, 0,
                       in, n, other, n_other, flags.p, bsum.p);
            break;
        case SEL_KEY_HEAD:
            DLG_LAUNCH(ctx, DLG_K_SELECT, rb, select_count_k<SEL_KEY_HEAD>, (unsigned)nb, SEL_T,
                       0, in, n, other, n_other, flags.p, bsum.p);
            break;
        case SEL_NODIAG:
            DLG_LAUNCH(ctx, DLG_K_SELECT, rb, select_count_k<SEL_NODIAG>, (unsigned)nb, SEL_T, 0,
                       in, n, other, n_other, flags.p, bsum.p);
            break;
        default: fail(DLG_INVALID_INPUT, "select: bad mode");
    }
    scan_blocks(ctx, bsum.p, nb, tot.p);
    u64 total = 0;
    ctx->read_scalars(tot.p, 1, &total);
    if (total == 0) return 0;
    if (mode == SEL_KEY_HEAD)
        DLG_LAUNCH(ctx, DLG_K_SELECT, 8.0 * ((n + 63) / 64) + 12.0 * total,
                   select_scatter_k<true>, (unsigned)nb, SEL_T, 0, in, n, flags.p, bsum.p, out);
    else
        DLG_LAUNCH(ctx, DLG_K_SELECT, 8.0 * ((n + 63) / 64) + 16.0 * total,
                   select_scatter_k<false>, (unsigned)nb, SEL_T, 0, in, n, flags.p, bsum.p, out);
    return total;
}

// First of each equal run of sorted `in` and absent from every run of `rs`.
// Each block narrows every run to the range spanned by its own tile (one
// thread per run), so a candidate's searches stay inside small ranges.
__global__ __launch_bounds__(SEL_T) void select_count_runs_k(const u64 *in, u64 n, RunSet rs,
                                                             u64 *flags, u64 *bsum) {
    __shared__ u64 s_lo[MAX_RUNS], s_hi[MAX_RUNS];
    __shared__ u64 sc[SEL_T / WAVE + 1];
    const u64 base = (u64)blockIdx.x * SEL_TILE;
    if ((int)threadIdx.x < rs.k) {
        const int j = threadIdx.x;
        u64 last = base + SEL_TILE - 1 < n ? base + SEL_TILE - 1 : n - 1;
        u64 lo = lower_bound_u64(rs.p[j], 0, rs.n[j], in[base]);
        s_lo[j] = lo;
        s_hi[j] = upper_bound_u64(rs.p[j], lo, rs.n[j], in[last]);
    }
    __syncthreads();
    u64 c = 0;
#pragma unroll
    for (int r = 0; r < SEL_I; r++) {
        u64 idx = base + (u64)r * SEL_T + threadIdx.x;
        bool keep = false;
        if (idx < n) {
            u64 x = in[idx];
            keep = idx == 0 || in[idx - 1] != x;
            for (int j = 0; keep && j < rs.k; j++)
                if (s_lo[j] < s_hi[j]) keep = !found_sorted(rs.p[j], s_lo[j], s_hi[j], x);
        }
        u64 m = __ballot(keep);
        u64 w0 = base + (u64)r * SEL_T + (threadIdx.x & ~(WAVE - 1));
        if ((threadIdx.x & (WAVE - 1)) == 0 && w0 < n) flags[w0 / WAVE] = m;
        c += keep ? 1 : 0;
    }
    u64 tot;
    block_exclusive_scan<u64>(c, sc, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

u64 select_scatter(dlg_ctx *ctx, const u64 *in, u64 n, const u64 *flags, DBuf<u64> &bsum, DBuf<u64> &tot,
                   u64 *out) {
    u64 nb = (n + SEL_TILE - 1) / SEL_TILE;
    scan_blocks(ctx, bsum.p, nb, tot.p);
    u64 total = 0;
    ctx->read_scalars(tot.p, 1, &total);
    if (total == 0) return 0;
    DLG_LAUNCH(ctx, DLG_K_SELECT, 8.0 * ((n + 63) / 64) + 16.0 * total, select_scatter_k<false>,
               (unsigned)nb, SEL_T, 0, in, n, flags, (const u64 *)bsum.p, (void *)out);
    return total;
}

u64 select_unique_not_in_runs(dlg_ctx *ctx, const u64 *in, u64 n, const RunSet &rs, u64 *out) {
    if (n == 0) return 0;
    if (rs.k == 0) return select(ctx, SEL_UNIQUE, in, n, nullptr, 0, out);
    u64 nb = (n + SEL_TILE - 1) / SEL_TILE;
    DBuf<u64> flags(ctx, (n + WAVE - 1) / WAVE + 1);
    DBuf<u64> bsum(ctx, nb);
    DBuf<u64> tot(ctx, 1);
    DLG_LAUNCH(ctx, DLG_K_SELECT, 8.0 * n + 8.0 * ((n + 63) / 64), select_count_runs_k,
               (unsigned)nb, SEL_T, 0, in, n, rs, flags.p, bsum.p);
    scan_blocks(ctx, bsum.p, nb, tot.p);
    u64 total = 0;
    ctx->read_scalars(tot.p, 1, &total);
    if (total == 0) return 0;
    DLG_LAUNCH(ctx, DLG_K_SELECT, 8.0 * ((n + 63) / 64) + 16.0 * total, select_scatter_k<false>,
               (unsigned)nb, SEL_T, 0, in, n, flags.p, bsum.p, (void *)out);
    return total;
}

// =============================================================== merge
// Merge path: output tile t of MP_TILE elements takes a[split[t], split[t+1])
// and the matching stretch of b; the split of diagonal d is the number of
// a-elements among the first d outputs of the stable merge (a first on ties).
constexpr int MP_T = 256, MP_I = 8, MP_TILE = MP_T * MP_I, MP_SPLIT_T = 256;

// One thread per diagonal d = t * MP_TILE, t in [0, tiles]: a plain binary search
// (a wave's 64 pivots per step would read 64 lines per step for every diagonal).
__global__ __launch_bounds__(MP_SPLIT_T) void merge_split_k(const u64 *a, u64 na, const u64 *b, u64 nb,
                                                            u64 tiles, u64 *split) {
    const u64 t = (u64)blockIdx.x * MP_SPLIT_T + threadIdx.x;
    if (t > tiles) return;
    const u64 n = na + nb, d = t * MP_TILE < n ? t * MP_TILE : n;
    u64 lo = d > nb ? d - nb : 0, hi = d < na ? d : na;
    while (lo < hi) {   // a[m] > b[d-1-m] turns true at the split; d-1-m stays in [0, nb)
        const u64 m = (lo + hi) >> 1;
        if (a[m] > b[d - 1 - m]) hi = m; else lo = m + 1;
    }
    split[t] = lo;
}

// Tile blockIdx.x of the merge: a's stretch and b's side by side in s, then
// this thread's MP_I outputs (tile positions [t MP_I, t MP_I + MP_I), valid
// below the returned count) merged into v.  s is still being read when it
// returns: a barrier must come before s is written again.
__device__ __forceinline__ int mp_tile(const u64 *a, u64 na, const u64 *b, u64 nb, const u64 *split, u64 *s,
                                       u64 (&v)[MP_I]) {
    const u64 n = na + nb, d0 = (u64)blockIdx.x * MP_TILE;
    const int cnt = (int)(n - d0 < (u64)MP_TILE ? n - d0 : (u64)MP_TILE);
    const u64 i0 = split[blockIdx.x], j0 = d0 - i0;
    const int ca = (int)(split[blockIdx.x + 1] - i0), cb = cnt - ca;
#pragma unroll
    for (int r = 0; r < MP_I; r++) {
        const int k = r * MP_T + (int)threadIdx.x;
        if (k < cnt) s[k] = k < ca ? a[i0 + (u64)k] : b[j0 + (u64)(k - ca)];
    }
    __syncthreads();
    const int dt = (int)threadIdx.x * MP_I;
    if (dt < cnt) {
        int lo = dt > cb ? dt - cb : 0, hi = dt < ca ? dt : ca;
        while (lo < hi) {   // this thread's split inside the tile
            const int m = (lo + hi) >> 1;
            if (s[m] <= s[ca + dt - 1 - m]) lo = m + 1; else hi = m;
        }
        int ia = lo, ib = dt - lo;
        u64 va = ia < ca ? s[ia] : 0, vb = ib < cb ? s[ca + ib] : 0;
#pragma unroll
        for (int r = 0; r < MP_I; r++) {
            const bool take_a = ia < ca && (ib >= cb || va <= vb);
            v[r] = take_a ? va : vb;
            if (take_a) {
                ia++;
                va = ia < ca ? s[ia] : 0;
            } else {
                ib++;
                vb = ib < cb ? s[ca + ib] : 0;
            }
        }
    }
    return cnt;
}

__global__ __launch_bounds__(MP_T) void merge_path_k(const u64 *a, u64 na, const u64 *b, u64 nb,
                                                     const u64 *split, u64 *out) {
    __shared__ u64 s[MP_TILE];
    u64 v[MP_I];
    const int cnt = mp_tile(a, na, b, nb, split, s, v);
    const int dt = (int)threadIdx.x * MP_I;
    __syncthreads();
    if (dt < cnt) {
#pragma unroll
        for (int r = 0; r < MP_I; r++)
            if (dt + r < cnt) s[dt + r] = v[r];
    }
    __syncthreads();
    const u64 d0 = (u64)blockIdx.x * MP_TILE;
#pragma unroll
    for (int r = 0; r < MP_I; r++) {
        const int k = r * MP_T + (int)threadIdx.x;
        if (k < cnt) out[d0 + (u64)k] = s[k];
    }
}

// Merge keeping one copy of each value: an output is kept when it differs
// from the one before it (the tile's first compares with the element before
// the tile: the larger of a[i0 - 1] and b[j0 - 1]).  Pass 1 (SCATTER false)
// writes each tile's count to tcnt; pass 2 reads the tiles' exclusive
// offsets there, compacts the kept outputs in LDS and stores them coalesced.
template <bool SCATTER>
__global__ __launch_bounds__(MP_T) void merge_unique_k(const u64 *a, u64 na, const u64 *b, u64 nb,
                                                       const u64 *split, u64 *tcnt, u64 *out) {
    __shared__ u64 s[MP_TILE];
    __shared__ u64 s_last[MP_T];
    __shared__ u64 sc[MP_T / WAVE + 1];
    u64 v[MP_I];
    const int cnt = mp_tile(a, na, b, nb, split, s, v);
    const int dt = (int)threadIdx.x * MP_I;
    const int m = dt < cnt ? (cnt - dt < MP_I ? cnt - dt : MP_I) : 0;   // this thread's valid outputs
    if (m) s_last[threadIdx.x] = v[m - 1];
    u64 prev = 0;
    bool has_prev = threadIdx.x > 0;
    if (threadIdx.x == 0 && blockIdx.x > 0) {   // the element before the tile
        const u64 i0 = split[blockIdx.x], j0 = (u64)blockIdx.x * MP_TILE - i0;
        const u64 pa = i0 ? a[i0 - 1] : 0, pb = j0 ? b[j0 - 1] : 0;
        prev = pa > pb ? pa : pb;
        has_prev = true;
    }
    __syncthreads();
    if (threadIdx.x > 0) prev = s_last[threadIdx.x - 1];   // a thread with outputs follows a full one
    u32 keep = 0, nk = 0;
#pragma unroll
    for (int r = 0; r < MP_I; r++) {
        if (r < m) {
            const bool k = !has_prev || v[r] != prev;
            keep |= (u32)k << r;
            nk += k;
            prev = v[r];
            has_prev = true;
        }
    }
    u64 tot;
    u64 pos = block_exclusive_scan<u64>(nk, sc, &tot);   // its barriers also end the tile's reads of s
    if (!SCATTER) {
        if (threadIdx.x == 0) tcnt[blockIdx.x] = tot;
        return;
    }
#pragma unroll
    for (int r = 0; r < MP_I; r++)
        if ((keep >> r) & 1u) s[pos++] = v[r];
    __syncthreads();
    const u64 o0 = tcnt[blockIdx.x];
    for (u64 k = threadIdx.x; k < tot; k += MP_T) out[o0 + k] = s[k];
}

static u64 merge_splits(dlg_ctx *ctx, const u64 *a, u64 na, const u64 *b, u64 nb, DBuf<u64> &split) {
    const u64 tiles = (na + nb + MP_TILE - 1) / MP_TILE;
    split.ensure(tiles + 1);
    DLG_LAUNCH(ctx, DLG_K_MERGE, 16.0 * (tiles + 1), merge_split_k, grid_for(tiles + 1, MP_SPLIT_T), MP_SPLIT_T, 0,
               a, na, b, nb, tiles, split.p);
    return tiles;
}

void merge(dlg_ctx *ctx, const u64 *a, u64 na, const u64 *b, u64 nb, u64 *out) {
    const u64 n = na + nb;
    if (n == 0) return;
    DBuf<u64> split(ctx);
    const u64 tiles = merge_splits(ctx, a, na, b, nb, split);
    DLG_LAUNCH(ctx, DLG_K_MERGE, 16.0 * n, merge_path_k, (unsigned)tiles, MP_T, 0, a, na, b, nb,
               (const u64 *)split.p, out);
}

u64 merge_unique(dlg_ctx *ctx, const u64 *a, u64 na, const u64 *b, u64 nb, u64 *out) {
    const u64 n = na + nb;
    if (n == 0) return 0;
    DBuf<u64> split(ctx), tcnt(ctx), tot(ctx, 1);
    const u64 tiles = merge_splits(ctx, a, na, b, nb, split);
    tcnt.ensure(tiles);
    DLG_LAUNCH(ctx, DLG_K_MERGE, 8.0 * n, merge_unique_k<false>, (unsigned)tiles, MP_T, 0, a, na, b, nb,
               (const u64 *)split.p, tcnt.p, (u64 *)nullptr);
    scan_blocks(ctx, tcnt.p, tiles, tot.p);
    u64 total = 0;
    ctx->read_scalars(tot.p, 1, &total);
    if (total == 0) return 0;
    DLG_LAUNCH(ctx, DLG_K_MERGE, 8.0 * n + 8.0 * total, merge_unique_k<true>, (unsigned)tiles, MP_T, 0, a, na, b,
               nb, (const u64 *)split.p, tcnt.p, out);
    return total;
}

}  // namespace dlg
