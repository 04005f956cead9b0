// noalloc.hip -- dlg_join_noalloc: the hash join of dlg_join with no device
// allocation at all, for callers that cannot allocate (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The caller passes the index, the probes and the output buffer.  The join's
// scratch -- per-probe match counts and row starts, the exclusive scan of the
// counts, the scan's per-tile partials and its total -- lives in `scratch`, a
// file-scope device array of DLG_NOALLOC_SCRATCH_BYTES that every GPU the
// library runs on holds until the process ends.  Host code never passes the
// array's name to a kernel (in host code it is only a stand-in for the
// symbol): dlg_ctx_create looks up each device's own address of the array
// once, with that device current, and the entry point hands that address to
// the kernels.  Every call on one GPU shares the array and leaves its
// contents behind; each kernel writes every scratch word it later reads
// (counts, starts, offsets, partials, total), so nothing needs clearing.
// Calls that overlap on one GPU race on the array and must be serialised.
#include "engine.h"

#include <cstdio>
#include <mutex>

namespace dlg {

static_assert(DLG_NOALLOC_SCRATCH_BYTES % 256 == 0, "scratch size");
static __device__ __align__(256) unsigned char scratch[DLG_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int NA_MAX_DEVICES = 64;
std::mutex na_mu;
bool na_done[NA_MAX_DEVICES];
void *na_ptr[NA_MAX_DEVICES];
hipError_t na_err[NA_MAX_DEVICES];

constexpr u64 na_align(u64 b) { return (b + 255) & ~255ull; }

// Byte offsets of the join's scratch for n probes: offsets (n + 1 u64s),
// counts and row starts (n u32s each), scan partials, scan total.
struct JoinLayout {
    u64 offs, cnt, pstart, bsum, total, end;
    explicit JoinLayout(u64 n) {
        offs = 0;
        cnt = na_align(offs + 8 * (n + 1));
        pstart = na_align(cnt + 4 * n);
        bsum = na_align(pstart + 4 * n);
        total = na_align(bsum + 8 * scan_partials(n));
        end = total + 8;
    }
};
static_assert(DLG_JOIN_NOALLOC_MAX_PROBES > 0, "probe limit");

// Saves the context's profiling switch and turns it off for the call (the
// event profiler creates events and grows a host list).
struct NoProfScope {
    dlg_ctx *c;
    bool saved;
    explicit NoProfScope(dlg_ctx *ctx) : c(ctx), saved(ctx->prof) { ctx->prof = false; }
    ~NoProfScope() { c->prof = saved; }
};
}  // namespace

void noalloc_init(int device) {
    if (device < 0 || device >= NA_MAX_DEVICES) return;
    std::lock_guard<std::mutex> lk(na_mu);
    if (na_done[device]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) (void)hipGetLastError();
    na_err[device] = e;
    na_ptr[device] = e == hipSuccess ? p : nullptr;
    na_done[device] = true;
}

// This device's address of the scratch array (looked up at context creation).
static unsigned char *device_scratch(int device) {
    if (device < 0 || device >= NA_MAX_DEVICES) fail(DLG_DEVICE_ERROR, "no-allocation scratch: device id too large");
    std::lock_guard<std::mutex> lk(na_mu);
    if (!na_done[device]) fail(DLG_DEVICE_ERROR, "no-allocation scratch: device not initialised");
    if (!na_ptr[device]) {
        char buf[160];
        std::snprintf(buf, sizeof(buf), "no-allocation scratch: hipGetSymbolAddress failed: %s",
                      hipGetErrorString(na_err[device]));
        fail(DLG_DEVICE_ERROR, buf);
    }
    return static_cast<unsigned char *>(na_ptr[device]);
}

}  // namespace dlg

using namespace dlg;

extern "C" dlg_status dlg_join_noalloc(dlg_ctx *c, const dlg_hash *h, const uint64_t *d_probe, int64_t n, int proj,
                                       uint64_t *d_out, int64_t out_cap, int64_t *out_n) {
    if (!c || !h || !out_n || n < 0 || out_cap < 0 || (n && !d_probe) || (out_cap && !d_out) ||
        (proj != DLG_PROJ_HV_PV && proj != DLG_PROJ_PV_HV))
        return DLG_INVALID_INPUT;
    *out_n = 0;
    API_BEGIN(c)
    if (n > DLG_JOIN_NOALLOC_MAX_PROBES) {
        char buf[160];
        std::snprintf(buf, sizeof(buf), "dlg_join_noalloc: %lld probes exceed the limit of %lld",
                      (long long)n, (long long)DLG_JOIN_NOALLOC_MAX_PROBES);
        fail(DLG_INVALID_INPUT, buf);
    }
    unsigned char *base = device_scratch(c->device);
    if (n && h->ix.n) {
        const JoinLayout L((u64)n);
        if (L.end > (u64)DLG_NOALLOC_SCRATCH_BYTES) fail(DLG_DEVICE_ERROR, "no-allocation scratch too small");
        NoProfScope np(c);
        u64 *offs = reinterpret_cast<u64 *>(base + L.offs);
        u32 *cnt = reinterpret_cast<u32 *>(base + L.cnt);
        u32 *pstart = reinterpret_cast<u32 *>(base + L.pstart);
        u64 *bsum = reinterpret_cast<u64 *>(base + L.bsum);
        u64 *total = reinterpret_cast<u64 *>(base + L.total);
        const u64 rows = join_count_with(c, h->ix, d_probe, (u64)n, cnt, pstart, offs, bsum, total);
        *out_n = (int64_t)rows;
        if (rows > (u64)out_cap) {
            char buf[160];
            std::snprintf(buf, sizeof(buf), "dlg_join_noalloc: %llu join rows exceed out_cap %lld",
                          (unsigned long long)rows, (long long)out_cap);
            fail(DLG_OVERFLOW, buf);
        }
        join_expand_with(c, h->ix, d_probe, (u64)n, pstart, offs, rows, proj, d_out);
        c->sync();
    }
    API_END(c)
}
