// dlg_internal.h -- runtime shared by the HIP kernels and the C ABI of libdlg.so.
//
// One dlg_ctx = one device, one HIP stream (default priority), one caching
// device-memory pool and an optional per-launch event profiler.  Errors are
// C++ exceptions inside the library and become dlg_status at the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/dlg.h"

namespace dlg {

typedef uint64_t u64;
typedef uint32_t u32;
typedef uint16_t u16;
typedef int64_t i64;

constexpr int WAVE = 64;          // CDNA wavefront width
constexpr u64 EMPTY_SLOT = ~0ull; // hash-table empty marker

struct Error {
    dlg_status status;
    std::string msg;
};

[[noreturn]] void fail(dlg_status s, const std::string &msg);
void check_hip(hipError_t e, const char *what, const char *file, int line);
hipEvent_t timing_event();   // an event for timing only (no system-scope fence)
// Host blocks handed to callers (result rows): page-locked while they stay
// under the caps below, so that the device-to-host copy runs at the link's
// rate (57 GB/s measured, against 7-15 GB/s into pageable memory); pageable
// (malloc) otherwise.  host_block_free frees either kind (and any malloc'd
// block of the library).  Null on failure.
void *host_block_alloc(size_t bytes);
void host_block_free(void *p);
// Looks up, once per device and with that device current, the address of the
// no-allocation entry points' static scratch array (noalloc.hip); called by
// dlg_ctx_create.  A failure is kept and reported by those entry points.
void noalloc_init(int device);
#define HIPCHK(call) ::dlg::check_hip((call), #call, __FILE__, __LINE__)

// ------------------------------------------------------------------ pool
// Caching allocator.  Sizes round up to 2^k * {1, 1.25, 1.5, 1.75} (waste
// under 25 %); freed blocks are kept per size class and reused by later
// requests on the same stream, so fixpoint loops do no hipMalloc once warm.
class Pool {
public:
    explicit Pool(int device) : device_(device) {}
    ~Pool();
    void *get(size_t bytes);
    void put(void *p);
    void trim();                      // release cached blocks to the runtime
    void trim_to(size_t keep);        // release cached blocks until at most `keep` bytes remain
    size_t live_bytes() const { return live_bytes_; }
    size_t cached_bytes() const { return cached_bytes_; }
    size_t peak_bytes() const { return peak_bytes_; }
    void set_budget(size_t b) { budget_ = b; }   // 0 = no limit besides the device
    size_t budget() const { return budget_; }
    static size_t round_size(size_t bytes);

private:
    int device_;
    std::map<size_t, std::vector<void *>> free_;
    std::unordered_map<void *, size_t> live_;
    size_t live_bytes_ = 0, cached_bytes_ = 0, peak_bytes_ = 0, budget_ = 0;
};

// ------------------------------------------------------------- profiler
struct KStat {
    i64 launches = 0;
    double ms = 0.0;
    double bytes = 0.0;
};

struct ProfRec {
    int kid;
    hipEvent_t a, b;
    double bytes;
};

}  // namespace dlg

struct dlg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    dlg::Pool *pool = nullptr;
    std::string last_error;
    int num_cus = 256;
    int64_t hbm_bytes = 0;
    int lds_per_block = 65536;
    std::string arch;
    // tuned defaults
    double load_factor = 0.5;
    int block_size = 256;
    int radix_bits = 8;
    // profiler
    bool prof = false;
    dlg::KStat kstat[DLG_K_COUNT];
    std::vector<dlg::ProfRec> pending;
    std::vector<hipEvent_t> event_free;
    hipEvent_t cur_a = nullptr;
    // pinned host scalar area for size read-backs
    uint64_t *h_scalars = nullptr;
    // pinned host words that kernels write directly (mapped), as a ring of
    // MAPPED_SLOTS slots of MAPPED_WORDS words: a fixpoint round's totals go
    // to slot seq % MAPPED_SLOTS, then (after a system-scope fence) the
    // round's sequence number to that slot's word MAPPED_SEQ; the host waits
    // for that number, with no copy.  Rounds queued ahead of the host each
    // have a slot of their own.
    uint64_t *h_mapped = nullptr, *d_mapped = nullptr;
    uint64_t mapped_seq = 0;
    static constexpr int MAPPED_SEQ = 7, MAPPED_WORDS = 8, MAPPED_SLOTS = 8;
    static constexpr int MAPPED_SPIN_US = 2000;   // busy-wait this long, then yield between checks
    uint64_t *mapped_host(uint64_t seq) const { return h_mapped + MAPPED_WORDS * (seq % MAPPED_SLOTS); }
    uint64_t *mapped_dev(uint64_t seq) const { return d_mapped + MAPPED_WORDS * (seq % MAPPED_SLOTS); }

    void prof_begin();
    void prof_end(int kid, double bytes);
    void prof_flush();
    // Algorithmic bytes of a launch known only after it ran (profiling on).
    void prof_add_bytes(int kid, double bytes) {
        if (prof) kstat[kid].bytes += bytes;
    }
    void sync();
    // Wait until a kernel has written `seq` to word MAPPED_SEQ of its slot
    // (spinning: no sleep between the kernel's write and the host's next
    // launch); fails if the stream runs dry without it.
    void wait_mapped(uint64_t seq);
    // read `count` u64 scalars from device memory into host (synchronises)
    void read_scalars(const uint64_t *d, int count, uint64_t *h);
};

namespace dlg {

// Device buffer owned by a context's pool.  Move-only.
template <typename T>
struct DBuf {
    dlg_ctx *ctx = nullptr;
    T *p = nullptr;
    size_t cap = 0;  // elements

    DBuf() = default;
    explicit DBuf(dlg_ctx *c) : ctx(c) {}
    DBuf(dlg_ctx *c, size_t n) : ctx(c) { ensure(n); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : ctx(o.ctx), p(o.p), cap(o.cap) { o.p = nullptr; o.cap = 0; }
    DBuf &operator=(DBuf &&o) noexcept {
        if (this != &o) {
            release();
            ctx = o.ctx; p = o.p; cap = o.cap;
            o.p = nullptr; o.cap = 0;
        }
        return *this;
    }
    ~DBuf() { release(); }
    // Make room for n elements; contents are NOT preserved when it grows.
    void ensure(size_t n) {
        if (n <= cap && p) return;
        release();
        size_t bytes = (n ? n : 1) * sizeof(T);
        p = static_cast<T *>(ctx->pool->get(bytes));
        cap = Pool::round_size(bytes) / sizeof(T);
    }
    void release() {
        if (p && ctx) ctx->pool->put(p);
        p = nullptr;
        cap = 0;
    }
    void swap(DBuf &o) {
        std::swap(ctx, o.ctx); std::swap(p, o.p); std::swap(cap, o.cap);
    }
};

// ------------------------------------------------------------- launches
inline unsigned grid_for(u64 n, u64 per_block) {
    u64 g = (n + per_block - 1) / per_block;
    return (unsigned)(g ? g : 1);
}
// Grid-stride kernels run ctx->block_size threads per block (tunable) and at
// most 4096 threads per CU.
inline unsigned sgrid(const dlg_ctx *ctx, u64 n) {
    u64 bs = (u64)ctx->block_size;
    u64 g = (n + bs - 1) / bs;
    u64 cap = (u64)ctx->num_cus * 4096 / bs;
    if (g > cap) g = cap;
    return (unsigned)(g ? g : 1);
}

#define DLG_LAUNCH(ctx, kid, bytes, kern, grid, block, shmem, ...)                        \
    do {                                                                                  \
        if ((ctx)->prof) (ctx)->prof_begin();                                             \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(block), shmem, (ctx)->stream, __VA_ARGS__); \
        HIPCHK(hipGetLastError());                                                        \
        if ((ctx)->prof) (ctx)->prof_end(kid, (double)(bytes));                           \
    } while (0)

// ------------------------------------------------------------- device helpers
__device__ __forceinline__ u32 fmix32(u32 h) {
    // MurmurHash3 finaliser (Austin Appleby, public domain), unsigned form.
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}
__device__ __forceinline__ u32 tkey(u64 t) { return (u32)(t >> 32); }
__device__ __forceinline__ u32 tval(u64 t) { return (u32)t; }
__device__ __forceinline__ u64 tmake(u32 k, u32 v) { return ((u64)k << 32) | v; }
__device__ __forceinline__ u64 lanemask_lt() {
    return (1ull << (threadIdx.x & (WAVE - 1))) - 1ull;
}

// cnt[key] += 1 for every active lane, one atomic per distinct key of the
// wave: lanes with the first pending lane's key are counted together (a
// label shared by most nodes, such as a giant component's, would otherwise
// serialise one atomic per node on one address).  Every lane of the wave
// must call it.
__device__ __forceinline__ void wave_count_by_key(u32 *cnt, u32 key, bool active) {
    const int lane = threadIdx.x & (WAVE - 1);
    u64 pending = __ballot(active);
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const u32 k = __shfl(key, leader, WAVE);
        const u64 same = __ballot(active && key == k) & pending;
        if (lane == leader) atomicAdd(&cnt[k], (u32)__popcll(same));
        pending &= ~same;
    }
}

// Block-wide exclusive scan (blockDim.x a multiple of 64, <= 1024).
// `scratch` holds >= blockDim.x/64 + 1 elements of T in LDS.
template <typename T>
__device__ __forceinline__ T block_exclusive_scan(T v, T *scratch, T *total) {
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
    const int nw = blockDim.x / WAVE;
    T inc = v;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        T t = __shfl_up(inc, d, WAVE);
        if (lane >= d) inc += t;
    }
    if (lane == WAVE - 1) scratch[wave] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        T s = 0;
        for (int w = 0; w < nw; w++) {
            T c = scratch[w];
            scratch[w] = s;
            s += c;
        }
        scratch[nw] = s;
    }
    __syncthreads();
    T r = scratch[wave] + inc - v;
    if (total) *total = scratch[nw];
    __syncthreads();
    return r;
}

__device__ __forceinline__ u32 wave_excl_scan32(u32 v, u32 *total) {
    const int lane = threadIdx.x & (WAVE - 1);
    u32 inc = v;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        u32 t = __shfl_up(inc, d, WAVE);
        if (lane >= d) inc += t;
    }
    *total = __shfl(inc, WAVE - 1, WAVE);
    return inc - v;
}

// Binary searches over sorted u64 arrays in global memory.
__device__ __forceinline__ u64 lower_bound_u64(const u64 *a, u64 lo, u64 hi, u64 x) {
    while (lo < hi) {
        u64 mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ u64 upper_bound_u64(const u64 *a, u64 lo, u64 hi, u64 x) {
    while (lo < hi) {
        u64 mid = (lo + hi) >> 1;
        if (a[mid] <= x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// The same searches by one whole wave (every lane calls with the same
// arguments; every lane gets the result): 64 pivots per step, so about
// log64(n) dependent loads instead of log2(n) -- for a block's range in a
// large offsets array, found before the block can start (a 10M-entry array:
// 4 loads instead of 23).  The first index in [lo, hi) whose element
// satisfies `past` (false, then true, along the array), or hi.
template <typename P>
__device__ __forceinline__ u64 wave_partition_index(u64 lo, u64 hi, P past) {   // past(i), i in [lo, hi)
    const int lane = threadIdx.x & (WAVE - 1);
    while (hi - lo > (u64)WAVE) {
        const u64 step = (hi - lo + WAVE - 1) / WAVE;
        const u64 p = lo + (u64)lane * step;
        const int c = __popcll(__ballot(p < hi && !past(p)));   // pivots before the point: a prefix
        if (c == 0) return lo;
        const u64 top = lo + (u64)c * step;
        lo = lo + (u64)(c - 1) * step + 1;
        hi = top < hi ? top : hi;
    }
    const u64 p = lo + (u64)lane;
    const u64 m = __ballot(p < hi && past(p));
    return m ? lo + (u64)(__ffsll((unsigned long long)m) - 1) : hi;
}
template <typename T, typename P>
__device__ __forceinline__ u64 wave_partition_point(const T *a, u64 lo, u64 hi, P past) {
    return wave_partition_index(lo, hi, [a, past](u64 i) { return past(a[i]); });
}
template <typename T>
__device__ __forceinline__ u64 wave_upper_bound(const T *a, u64 lo, u64 hi, u64 x) {
    return wave_partition_point(a, lo, hi, [x](T v) { return (u64)v > x; });
}
template <typename T>
__device__ __forceinline__ u64 wave_lower_bound(const T *a, u64 lo, u64 hi, u64 x) {
    return wave_partition_point(a, lo, hi, [x](T v) { return (u64)v >= x; });
}

// ------------------------------------------------------------- primitives (k_prims.hip)
// Exclusive scan of u32 counts into u64 offsets; returns the total (synchronises).
u64 scan_u32_u64(dlg_ctx *ctx, const u32 *in, u64 n, u64 *out);
// Same without reading the total back; total left at d_total (device).
void scan_u32_u64_async(dlg_ctx *ctx, const u32 *in, u64 n, u64 *out, u64 *d_total);
// Partials (u64s) a scan of n elements uses.
u64 scan_partials(u64 n);
// Same over caller-provided partials bsum[scan_partials(n)] (allocates nothing).
void scan_u32_u64_with(dlg_ctx *ctx, const u32 *in, u64 n, u64 *out, u64 *bsum, u64 *d_total);
// Exclusive scan of u64 values (out may alias in); total left at d_total.
void scan_u64_async(dlg_ctx *ctx, const u64 *in, u64 n, u64 *out, u64 *d_total);
// Stable LSD radix sort of u64 keys; result in `data`.  `tmp` is scratch of >= n.
// key_only: a stable sort by the key half alone (equal keys keep their order).
void radix_sort(dlg_ctx *ctx, u64 *data, u64 n, DBuf<u64> &tmp, int bits = 0, bool key_only = false);
// Same on a pool buffer: the sorted keys may end in the other buffer, in which
// case `data` and `tmp` swap ownership (no copy).  `bounds` (min key, max key,
// min value, max value of the u64 halves) may be given when known to cover
// every element; otherwise a pass over the data computes them.
void radix_sort(dlg_ctx *ctx, DBuf<u64> &data, u64 n, DBuf<u64> &tmp, int bits = 0,
                const u32 *bounds = nullptr, bool key_only = false);
// Digest of tuples as output rows in `orient` (dlg_result.output_sum) and the
// number of i with d[i] >= d[i + 1] (synchronises).
// sub0 / sub1 are subtracted (mod 2^32) from the rows' two columns first: the
// batched form's id shift of an instance.
void rows_digest(dlg_ctx *ctx, const u64 *d, u64 n, int orient, u64 *sum, u64 *breaks, u32 sub0 = 0,
                 u32 sub1 = 0);
// Ranges of the two u64 halves over data[0..n): out = {min_hi, max_hi, min_lo, max_lo}.
void tuple_bounds(dlg_ctx *ctx, const u64 *data, u64 n, u32 *out);
// Compaction modes.
enum SelectMode {
    SEL_UNIQUE = 0,         // keep first of each run of equal elements
    SEL_UNIQUE_NODIAG = 1,  // ... and drop key == value
    SEL_UNIQUE_NOT_IN = 2,  // ... and drop elements present in sorted `other`
    SEL_NOT_IN = 3,         // drop elements present in sorted `other`
    SEL_KEY_HEAD = 4,       // output INDEX of first element of each equal-key run
    SEL_NODIAG = 5,         // keep every element with key != value (any order, duplicates kept)
};
// Writes kept elements (or indices for SEL_KEY_HEAD, as u32) to out; returns count.
// `out` must not alias `in`.
u64 select(dlg_ctx *ctx, int mode, const u64 *in, u64 n, const u64 *other, u64 n_other,
           void *out);
// A set of sorted, pairwise disjoint runs (the full relation kept in tiers).
constexpr int MAX_RUNS = 24;
struct RunSet {
    const u64 *p[MAX_RUNS];
    u64 n[MAX_RUNS];
    int k;
};
// Unique elements of sorted `in` that no run of `rs` contains.
u64 select_unique_not_in_runs(dlg_ctx *ctx, const u64 *in, u64 n, const RunSet &rs, u64 *out);
// Second half of a compaction whose count kernel (tiles of 2048, flags as
// ballot words, per-tile counts in bsum) already ran: scan, read the total
// back, scatter.  Returns the kept count.
u64 select_scatter(dlg_ctx *ctx, const u64 *in, u64 n, const u64 *flags, DBuf<u64> &bsum, DBuf<u64> &tot,
                   u64 *out);

// Exclusive scan of bsum[0..nb) in place by one block; the total to *d_total.
void scan_blocks(dlg_ctx *ctx, u64 *bsum, u64 nb, u64 *d_total);

// ------------------------------------------------------------- bitmap (k_bitmap.hip)
struct Index;
struct JoinScratch;
// The full relation as a bitmap over the id range [lo, lo + W): W rows of
// wpr = ceil(W / 32) words (S = 32 * wpr bits); bit (k - lo) * S + (v - lo)
// for tuple (k, v), or (v - lo) * S + (k - lo) when transposed.
struct Bitmap {
    DBuf<u32> words;
    DBuf<u64> count;   // append counter of the insert kernels
    u32 lo = 0;
    u64 W = 0, wpr = 0, S = 0, nbits = 0, nwords = 0;
    bool tr = false;
    explicit Bitmap(dlg_ctx *c) : words(c), count(c) {}
};
// Geometry for ids [lo, hi]; the map is cleared unless `clear` is false (its
// padding always is).
void bitmap_init(dlg_ctx *ctx, Bitmap &bm, u32 lo, u32 hi, bool transposed, bool clear = true);
// The single join R(a, b), R(b, c) straight into a row-major map (row c, bit
// a), from `rows` = R keyed by its second column, sorted, n rows, with dense
// offsets over the map's id range (W + 1 entries).  Every row is written.
constexpr u64 SJ_ROW_LDS_MAX = 64 * 1024;   // bytes of one map row in LDS
void bitmap_self_join(dlg_ctx *ctx, const u64 *rows, u64 n, const u64 *dense, Bitmap &bm);
// The raw size of that join, sum over b of |{a}| x |{c}| (synchronises).
u64 self_join_raw(dlg_ctx *ctx, const u64 *rows, u64 n, const u64 *dense, u32 lo);
// The same single join by rows: each output row's candidates sorted and
// deduplicated in LDS by one wave, the rows placed by a scan, (c, a) tuples
// in order into out (resized), *nout of them, and the raw join's size into
// *raw.  Rows with more than SJW_CAP candidates are materialised, sorted and
// deduplicated together.  Returns false, with nothing written, when those
// exceed 2^32 - 1 tuples: the caller's sorted join takes such inputs.
// Synchronises.
constexpr u32 SJW_CAP = 2048;
bool self_join_rows(dlg_ctx *ctx, const u64 *rows, u64 n, const u64 *dense, u32 lo, u64 W, DBuf<u64> &out,
                    u64 *nout, u64 *raw);
// probe ⋈ index, each row inserted; rows that set a clear bit go to `out`
// (unordered).  Returns their number; *raw += |J|.
u64 bitmap_join_insert(dlg_ctx *ctx, const Index &ix, const u64 *probe, u64 n, JoinScratch &js, int proj,
                       bool nodiag, Bitmap &bm, DBuf<u64> &out, u64 *raw);
// probe ⋈ index, each row inserted; the bits that were clear are also set in
// `mark` (same geometry), or recorded nowhere when `mark` is null.
void bitmap_join_mark(dlg_ctx *ctx, const Index &ix, const u64 *probe, u64 n, JoinScratch &js, int proj,
                      bool nodiag, Bitmap &bm, Bitmap *mark, u64 *raw);
// All set bits in bit order (as (column, row) tuples when `swap`), then the
// map cleared (unless `clear` is false).  Returns their number.
u64 bitmap_drain(dlg_ctx *ctx, Bitmap &bm, DBuf<u64> &out, bool swap, bool clear = true);
// The number of set bits (synchronises).
u64 bitmap_count(dlg_ctx *ctx, const Bitmap &bm);
// Insert a list; the tuples that were absent go to `out`.
u64 bitmap_insert_list(dlg_ctx *ctx, const u64 *in, u64 n, bool nodiag, Bitmap &bm, u64 *out);
// Clear the bits of a list whose bits are the only ones set in their words.
void bitmap_clear_list(dlg_ctx *ctx, const u64 *in, u64 n, Bitmap &bm);
// All set bits as sorted tuples; `expect` = the number of set bits.  A
// transposed map is transposed back into a scratch map first -- unless it
// holds a symmetric relation, which equals its transpose and is emitted as it
// lies.
u64 bitmap_emit(dlg_ctx *ctx, Bitmap &bm, DBuf<u64> &out, u64 expect, bool symmetric);

// Several maps emitted at once, in the order given, each in bit order (ids
// from lo, W rows of wpr words at words + base; transposed back when
// `transposed`): one count, one scan and one emission for all of them.
// `expect` = their set bits (checked).  Returns the count.
struct MapRef {
    u32 lo;
    u64 W, wpr, base;
};
u64 bitmap_emit_maps(dlg_ctx *ctx, const u32 *words, const std::vector<MapRef> &maps, bool transposed, u64 *out,
                     u64 expect);
// A W-row block of a map, emitted in tuple order at out through a padded
// scratch copy (transposed back when `transposed`).  Returns its set bits.
u64 bitmap_emit_block(dlg_ctx *ctx, const u32 *words, u32 lo, u64 W, u64 wpr, bool transposed, DBuf<u32> &scratch,
                      u64 *out);
// The same for a transposed block whose ids were renumbered (row sigma(x),
// column sigma(z) holds tuple (z, x)): rows gathered back into input order,
// the map transposed, its rows gathered again, then emitted.  `words` is
// clobbered.
u64 bitmap_emit_block_renumbered(dlg_ctx *ctx, u32 *words, u32 lo, u64 W, u64 wpr, const u32 *sigma,
                                 DBuf<u32> &scratch, u64 *out);

// ------------------------------------------------------------- TC rows (k_tcrows.hip)
// The full TC relation as a block-diagonal set of transposed bitmaps: one
// block per id range [starts[i], starts[i+1]) (one block for a plain run,
// one per instance in the batched form, whose instances own disjoint id
// ranges); row x of a block holds the ends z of the paths from x, columns
// relative to the block's first id.  row_base / col_lo give each id's row
// and its block's first id.
struct RowMap {
    struct Block {
        u32 lo;
        u64 W, wpr, base;   // ids, words per row, first word
    };
    DBuf<u32> words;
    DBuf<u64> row_base;
    DBuf<u32> col_lo;
    std::vector<Block> blocks;
    u32 lo = 0;
    u64 W = 0, nwords = 0;
    explicit RowMap(dlg_ctx *c) : words(c), row_base(c), col_lo(c) {}
};
// Bits a RowMap over [lo, hi] with the given block starts (null: one block) needs.
u64 rowmap_bits(u32 lo, u32 hi, const std::vector<u32> *starts);
void rowmap_init(dlg_ctx *ctx, RowMap &m, u32 lo, u32 hi, const std::vector<u32> *starts);
// All bits in tuple order (blocks in id order, each transposed back).
u64 rowmap_emit(dlg_ctx *ctx, RowMap &m, DBuf<u64> &out, u64 expect);

// TC over a transposed bitmap F with Δ grouped by source: segment s holds
// the ends dy[seg_beg[s] ..][0, seg_len[s]) of source seg_x[s] (one segment
// per source, fixed for the whole fixpoint; empty ones stay), plus the dense
// CSR of E over F's id range.  seg_out = exclusive scan of the segments'
// output bounds = where their new ends go.
struct TcRows {
    DBuf<u64> csr_off;
    DBuf<u32> csr_dst, deg, ell;   // ell: ELL rows when every degree is small (else unallocated)
    // Δ: ends dy grouped by source, and the active list (the sources with
    // ends): source ax, its ends at dy[abeg ..][0, alen), its output offset
    // aoff (exclusive scan of the bounds); n*: the next round's.  rcnt,
    // rbound: each entry's new ends and next bound; bpart, brange: the row
    // kernel's totals and entry range per workgroup.
    DBuf<u32> dy, ny, ax, nx, alen, nlen, rcnt;
    DBuf<u64> abeg, nbeg, aoff, noff, rbound, bpart, brange, counters;
    u64 nd = 0, nact = 0, rows = 0;   // rows: the round's output bound (Σ of the entries')
    u32 dmax = 0;                     // nonzero: every degree <= dmax, bounds = new ends * dmax
    u64 wcap = 0;                     // the widest block: no source has more new ends
    explicit TcRows(dlg_ctx *c)
        : csr_off(c), csr_dst(c), deg(c), ell(c), dy(c), ny(c), ax(c), nx(c), alen(c), nlen(c), rcnt(c), abeg(c), nbeg(c),
          aoff(c), noff(c), rbound(c), bpart(c), brange(c), counters(c) {}
};
// CSR of the index's rows; Δ0 = unique(E reversed) grouped by source, its
// bits set in F (F empty).
void tc_rows_setup(dlg_ctx *ctx, const Index &ix, RowMap &F, TcRows &tr);
// One round: N = (Δ ⋈ E) \ F into F, Δ = N (grouped).  Returns |N|; *raw += |Δ ⋈ E|.
u64 tc_rows_round(dlg_ctx *ctx, RowMap &F, TcRows &tr, u64 *raw);
// Every round to the fixpoint (or max_it > 0 rounds), with up to `queue`
// rounds (0: the default) queued ahead of the host; returns the round count,
// *total += Σ |N|.
int tc_rows_run(dlg_ctx *ctx, RowMap &F, TcRows &tr, int max_it, int queue, u64 *total, u64 *raw);
// The current Δ as (z, x) tuples.
void tc_rows_tuples(dlg_ctx *ctx, const TcRows &tr, DBuf<u64> &out);
// Locality order of the ids 0 .. W-1 of the edges (src[i], dst[i]) (order.cpp):
// sigma[old] = new.
std::vector<u32> locality_order(const std::vector<u32> &src, const std::vector<u32> &dst, u32 W);
// out = the int32 rows with every id renumbered: key k -> sigma[k - lo] + lo
// (k the id as a flipped key, so the id range is kept).
void relabel_rows(dlg_ctx *ctx, const int32_t *in, u64 n, const u32 *sigma, u32 lo, int32_t *out);
// rowmap_emit of a one-block map whose ids were renumbered by sigma (the map
// holds row sigma(x), column sigma(z) for tuple (z, x)): the tuples in the
// input's ids and order.  The map's words are clobbered.
u64 rowmap_emit_renumbered(dlg_ctx *ctx, RowMap &m, DBuf<u64> &out, u64 expect, const u32 *sigma);
// Fraction of the index's rows (k, v) whose reverse (v, k) is also a row.
double reciprocal_fraction(dlg_ctx *ctx, const Index &ix);

// ------------------------------------------------------------- SG row ORs (k_sgrows.hip)
struct SgRows {
    DBuf<u64> poff, soff, counters;   // predecessor / successor list offsets (W + 1); step counters
    DBuf<u32> pred, succ, rflag;      // predecessors, successors (relative to lo); nonempty rows of R
    DBuf<u32> ell;                    // successors as ELL rows when every out-degree is small
    explicit SgRows(dlg_ctx *c) : poff(c), soff(c), counters(c), pred(c), succ(c), rflag(c), ell(c) {}
};
// Predecessor and successor lists of the index's edges over F's id range.
void sg_rows_setup(dlg_ctx *ctx, const Index &ix, const Bitmap &F, SgRows &sr);
// One SG round on the maps: T = Δ ⋈ E from Δ in D (D emptied) into the
// row-major map R, then N = (T ⋈ E) \ F: new bits into F and D (both
// transposed); R is left empty.  Returns |N|; *raw += both joins' rows.
u64 sg_rows_round(dlg_ctx *ctx, SgRows &sr, Bitmap &R, Bitmap &F, Bitmap &D, u64 *raw);

// ------------------------------------------------------------- hash set (k_hashset.hip)
struct HashSet {
    DBuf<u64> slots;   // open addressing, EMPTY_SLOT = free
    DBuf<u32> flag;    // holds the one tuple equal to EMPTY_SLOT
    u64 cap = 0, count = 0;
    explicit HashSet(dlg_ctx *c) : slots(c), flag(c) {}
};
// Make room for `need` elements at load <= lf; a new table is refilled from all[0..n_all).
void hashset_reserve(dlg_ctx *ctx, HashSet &hs, u64 need, double lf, const u64 *all, u64 n_all);
// Insert tuples that are absent and pairwise distinct.
void hashset_insert(dlg_ctx *ctx, HashSet &hs, const u64 *x, u64 n);
// First of each equal run of sorted `in` that the set does not contain.
u64 select_unique_not_in_hash(dlg_ctx *ctx, const u64 *in, u64 n, const HashSet &hs, u64 *out);
// Merge of sorted a, b into out (a first on ties).
void merge(dlg_ctx *ctx, const u64 *a, u64 na, const u64 *b, u64 nb, u64 *out);
// The merge with one copy of each value kept (inputs sorted, duplicates
// allowed); returns the count written to out (synchronises).
u64 merge_unique(dlg_ctx *ctx, const u64 *a, u64 na, const u64 *b, u64 nb, u64 *out);
void pack_rows(dlg_ctx *ctx, const int32_t *rows, u64 n, int orient, u64 *out);
void unpack_rows(dlg_ctx *ctx, const u64 *t, u64 n, int orient, int32_t *rows);
// E into out[0, n) and E⁻¹ into out_rev[0, n), row order kept, also returning
// the count of adjacent input rows out of order and, in cb, the flipped
// bounds of both columns (min a, max a, min b, max b) -- one read
// (synchronises).
u64 pack_sym_rows_scan(dlg_ctx *ctx, const int32_t *rows, u64 n, u64 *out, u64 *out_rev, u32 *cb);
void fill_u64(dlg_ctx *ctx, u64 *p, u64 n, u64 v);
void fill_u32(dlg_ctx *ctx, u32 *p, u64 n, u32 v);
void copy_u64(dlg_ctx *ctx, u64 *dst, const u64 *src, u64 n);

// ------------------------------------------------------------- index + join (k_join.hip)
struct Index {
    DBuf<u64> rows;      // relation sorted by (key, value); duplicates kept
    DBuf<u32> gstart;    // start row of each key group, plus gstart[G] = n
    DBuf<u64> slots;     // open-addressed (key << 32 | group) slots, EMPTY_SLOT = free
    DBuf<u64> dense;     // optional: first row of each key of [dense_lo, dense_lo + dense_w), plus n
    u64 n = 0, groups = 0, nslots = 0, dense_w = 0;
    u32 dense_lo = 0;
    bool sorted = false; // probes binary-search `rows` instead of the slots (no slots built); a WCC
                         // graph's index may also have neither (value lookups by rank directory)
    explicit Index(dlg_ctx *c) : rows(c), gstart(c), slots(c), dense(c) {}
};
// Build from a device relation (copied); `presorted` skips the sort;
// `sorted_lookup` builds no hash slots (probes binary-search the rows).
void index_build(dlg_ctx *ctx, Index &ix, const u64 *rel, u64 n, double load_factor,
                 bool presorted, bool sorted_lookup = false);
// Dense row offsets by key over [lo, lo + W) (index_build drops them); the
// hash-index count phase then reads an offset pair per probe instead of
// probing the slots.
// The hash slots of an index built with sorted_lookup (index_build's last
// step), for key lookups by hash probe.
void index_hash(dlg_ctx *ctx, Index &ix, double lf);
// Whether index_lookup_values takes the rank directory (no slots needed).
bool index_rank_lookup_applies(dlg_ctx *ctx, const Index &ix);
void index_dense(dlg_ctx *ctx, Index &ix, u32 lo, u64 W);
// Count phase: per probe, the matching group's row start and count.  Returns |J|.
struct JoinScratch {
    DBuf<u32> cnt, pstart;
    DBuf<u64> offs;
    explicit JoinScratch(dlg_ctx *c) : cnt(c), pstart(c), offs(c) {}
};
u64 join_count(dlg_ctx *ctx, const Index &ix, const u64 *probe, u64 n, JoinScratch &js);
void join_expand(dlg_ctx *ctx, const Index &ix, const u64 *probe, u64 n, const JoinScratch &js,
                 u64 total, int proj, u64 *out);
// The same two phases over caller-provided scratch (no allocation): cnt and
// pstart of n u32s, offs of n + 1 u64s, bsum of scan_partials(n) u64s and
// one u64 at d_total.  join_count_with returns |J| (synchronises).
u64 join_count_with(dlg_ctx *ctx, const Index &ix, const u64 *probe, u64 n, u32 *cnt, u32 *pstart, u64 *offs,
                    u64 *bsum, u64 *d_total);
void join_expand_with(dlg_ctx *ctx, const Index &ix, const u64 *probe, u64 n, const u32 *pstart, const u64 *offs,
                      u64 total, int proj, u64 *out);
// Dense lookup of every value of `ix.rows` as a key: out[i] = group id (u32) or ~0u.
void index_lookup_values(dlg_ctx *ctx, const Index &ix, u32 *out);

// ------------------------------------------------------------- WCC (k_wcc.hip)
struct WccState {
    u64 V = 0;
    DBuf<u32> L, Ln, mark, fnode, flabel, nnode, cnt, pstart, counter;
    DBuf<u64> offs;
    explicit WccState(dlg_ctx *c)
        : L(c), Ln(c), mark(c), fnode(c), flabel(c), nnode(c), cnt(c), pstart(c), counter(c), offs(c) {}
};
void wcc_setup(dlg_ctx *ctx, WccState &s, u64 V);
// WCC round forms: WCC_AUTO pulls when the frontier's edges are at least half
// of the graph's, WCC_PUSH walks the frontier's edges, WCC_PULL every node's.
enum { WCC_AUTO = 0, WCC_PUSH = 1, WCC_PULL = 2 };
// One round over frontier of size nf; returns the new frontier size.
u64 wcc_round(dlg_ctx *ctx, WccState &s, const Index &ix, const u32 *nbr, u64 nf, u32 round, int form = WCC_PUSH);
// Round one from the identity labels (every node in the frontier) when each
// node's neighbour list is sorted: returns the new frontier size.
u64 wcc_first_round(dlg_ctx *ctx, WccState &s, const Index &ix, const u32 *nbr);
// The reference's round on tuples (dlg_wcc_relax): cc (one (node, label)
// tuple per group of `ix`, in group order) lowered in place to the per-node
// least label of cc and Δ ⋈ E; returns |Δ'| with Δ' = cc' minus cc (sorted) in
// `next`.  nbr[r] = the group of row r's value.
u64 wcc_relax_tuples(dlg_ctx *ctx, const Index &ix, const u32 *nbr, const u64 *delta, u64 nd, u64 *cc, u64 ncc,
                     u64 *next);
// #components, largest component size and the (label, 0) tuples in `out`.
u64 wcc_finalize(dlg_ctx *ctx, WccState &s, const Index &ix, DBuf<u64> &out, i64 *largest);
// All rounds through replays of a captured graph of `kr` rounds; returns the
// iteration count (labels in s.L).  `ne` = edges of the index.
int wcc_rounds_graph(dlg_ctx *ctx, WccState &s, const Index &ix, const u32 *nbr, u64 ne, int max_it,
                     int kr);

}  // namespace dlg

// Exception -> dlg_status translation used by every C ABI entry point.
#define API_BEGIN(ctx)       \
    try {                    \
        if (ctx) (void)hipSetDevice((ctx)->device);
#define API_END(ctx)                                                   \
    }                                                                  \
    catch (const dlg::Error &e) {                                      \
        if (ctx) (ctx)->last_error = e.msg;                            \
        return e.status;                                               \
    }                                                                  \
    catch (const std::bad_alloc &) {                                   \
        if (ctx) (ctx)->last_error = "host allocation failed";         \
        return DLG_OUT_OF_MEMORY;                                      \
    }                                                                  \
    return DLG_OK;
// Same, running `cleanup` before an error status is returned.
#define API_END_CLEANUP(ctx, cleanup)                                  \
    }                                                                  \
    catch (const dlg::Error &e) {                                      \
        cleanup;                                                       \
        if (ctx) (ctx)->last_error = e.msg;                            \
        return e.status;                                               \
    }                                                                  \
    catch (const std::bad_alloc &) {                                   \
        cleanup;                                                       \
        if (ctx) (ctx)->last_error = "host allocation failed";         \
        return DLG_OUT_OF_MEMORY;                                      \
    }                                                                  \
    return DLG_OK;

