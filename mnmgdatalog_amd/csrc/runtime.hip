// runtime.hip -- context, stream, memory pool and event profiler of libdlg.so.
#include "dlg_internal.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <thread>

namespace dlg {

void fail(dlg_status s, const std::string &msg) { throw Error{s, msg}; }

void check_hip(hipError_t e, const char *what, const char *file, int line) {
    if (e == hipSuccess) return;
    char buf[512];
    std::snprintf(buf, sizeof(buf), "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    fail(e == hipErrorOutOfMemory ? DLG_OUT_OF_MEMORY : DLG_DEVICE_ERROR, buf);
}

size_t Pool::round_size(size_t bytes) {
    if (bytes < 256) return 256;
    int k = 63 - __builtin_clzll((unsigned long long)bytes);  // 2^k <= bytes
    size_t base = (size_t)1 << k;
    if (k < 4) return base * 2;
    size_t q = base >> 2;                                      // quarter steps
    size_t r = (bytes + q - 1) / q * q;
    return r;
}

Pool::~Pool() {
    (void)hipSetDevice(device_);
    trim();
    for (auto &kv : live_) (void)hipFree(kv.first);
    live_.clear();
}

void *Pool::get(size_t bytes) {
    size_t sz = round_size(bytes);
    auto it = free_.find(sz);
    if (it != free_.end() && !it->second.empty()) {
        void *p = it->second.back();
        it->second.pop_back();
        cached_bytes_ -= sz;
        live_[p] = sz;
        live_bytes_ += sz;
        peak_bytes_ = std::max(peak_bytes_, live_bytes_);
        return p;
    }
    if (budget_ && live_bytes_ + cached_bytes_ + sz > budget_) {
        // the budget covers cached blocks too: release cache before giving up
        trim_to(budget_ > live_bytes_ + sz ? budget_ - live_bytes_ - sz : 0);
        if (live_bytes_ + sz > budget_) {
            char buf[160];
            std::snprintf(buf, sizeof(buf),
                          "device memory budget exceeded: need %zu more bytes, budget %zu", sz, budget_);
            fail(DLG_OUT_OF_MEMORY, buf);
        }
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, sz);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        trim();
        e = hipMalloc(&p, sz);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            char buf[160];
            std::snprintf(buf, sizeof(buf), "hipMalloc(%zu) failed: %s", sz, hipGetErrorString(e));
            fail(DLG_OUT_OF_MEMORY, buf);
        }
    }
    live_[p] = sz;
    live_bytes_ += sz;
    peak_bytes_ = std::max(peak_bytes_, live_bytes_);
    return p;
}

void Pool::put(void *p) {
    auto it = live_.find(p);
    if (it == live_.end()) return;
    size_t sz = it->second;
    live_.erase(it);
    live_bytes_ -= sz;
    free_[sz].push_back(p);
    cached_bytes_ += sz;
}

void Pool::trim() {
    for (auto &kv : free_)
        for (void *p : kv.second) (void)hipFree(p);
    free_.clear();
    cached_bytes_ = 0;
}

void Pool::trim_to(size_t keep) {
    // free the largest cached blocks first
    while (cached_bytes_ > keep && !free_.empty()) {
        auto it = std::prev(free_.end());
        if (it->second.empty()) {
            free_.erase(it);
            continue;
        }
        (void)hipFree(it->second.back());
        it->second.pop_back();
        cached_bytes_ -= it->first;
        if (it->second.empty()) free_.erase(it);
    }
}

namespace {
constexpr size_t PINNED_BLOCK_MAX = 4ull << 30;    // larger blocks stay pageable
constexpr size_t PINNED_TOTAL_MAX = 32ull << 30;   // page-locked host memory held at once
std::mutex pinned_mu;
std::unordered_map<void *, size_t> pinned_blocks;
size_t pinned_total = 0;
}  // namespace

void *host_block_alloc(size_t bytes) {
    if (bytes <= PINNED_BLOCK_MAX) {
        std::lock_guard<std::mutex> lk(pinned_mu);
        if (pinned_total + bytes <= PINNED_TOTAL_MAX) {
            void *p = nullptr;
            if (hipHostMalloc(&p, bytes, hipHostMallocDefault) == hipSuccess && p) {
                pinned_blocks[p] = bytes;
                pinned_total += bytes;
                return p;
            }
            (void)hipGetLastError();
        }
    }
    return std::malloc(bytes);
}

void host_block_free(void *p) {
    if (!p) return;
    {
        std::lock_guard<std::mutex> lk(pinned_mu);
        auto it = pinned_blocks.find(p);
        if (it != pinned_blocks.end()) {
            pinned_total -= it->second;
            pinned_blocks.erase(it);
            (void)hipHostFree(p);
            return;
        }
    }
    std::free(p);
}

// Timing events only (read by hipEventElapsedTime, never used to order
// memory): no system-scope fence, so recording one does not write back and
// invalidate the caches between two kernels.
hipEvent_t timing_event() {
    hipEvent_t e;
    HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    return e;
}

}  // namespace dlg

using namespace dlg;

void dlg_ctx::prof_begin() {
    if (event_free.empty()) {
        hipEvent_t e = timing_event();
        event_free.push_back(e);
    }
    cur_a = event_free.back();
    event_free.pop_back();
    HIPCHK(hipEventRecord(cur_a, stream));
}

void dlg_ctx::prof_end(int kid, double bytes) {
    if (event_free.empty()) {
        hipEvent_t e = timing_event();
        event_free.push_back(e);
    }
    hipEvent_t b = event_free.back();
    event_free.pop_back();
    HIPCHK(hipEventRecord(b, stream));
    pending.push_back(ProfRec{kid, cur_a, b, bytes});
    if (pending.size() >= 4096) prof_flush();
}

void dlg_ctx::prof_flush() {
    if (pending.empty()) return;
    HIPCHK(hipEventSynchronize(pending.back().b));
    for (auto &r : pending) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
        kstat[r.kid].launches++;
        kstat[r.kid].ms += ms;
        kstat[r.kid].bytes += r.bytes;
        event_free.push_back(r.a);
        event_free.push_back(r.b);
    }
    pending.clear();
}

void dlg_ctx::sync() { HIPCHK(hipStreamSynchronize(stream)); }

void dlg_ctx::wait_mapped(uint64_t seq) {
    volatile uint64_t *w = mapped_host(seq) + MAPPED_SEQ;
    // Spin for about one long round (the next launch goes out the moment the
    // number lands), then yield the core between checks so that other host
    // threads (the served form's slots and admission) keep running.
    const auto t0 = std::chrono::steady_clock::now();
    bool yielding = false;
    for (unsigned spin = 1; *w != seq; spin++) {
        if (spin % 1024 == 0) {   // now and then: has the stream finished without writing it?
            const hipError_t e = hipStreamQuery(stream);
            if (e == hipSuccess) {
                if (*w == seq) break;
                fail(DLG_DEVICE_ERROR, "a round's totals never reached the host");
            }
            if (e != hipErrorNotReady) HIPCHK(e);
            if (!yielding)
                yielding = std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(MAPPED_SPIN_US);
        }
        if (yielding) {
            std::this_thread::yield();
        } else {
#if defined(__x86_64__)
            __builtin_ia32_pause();
#endif
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);   // the totals were written before the number
}

void dlg_ctx::read_scalars(const uint64_t *d, int count, uint64_t *h) {
    HIPCHK(hipMemcpyAsync(h_scalars, d, sizeof(uint64_t) * count, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    std::memcpy(h, h_scalars, sizeof(uint64_t) * count);
}

// ------------------------------------------------------------------ C ABI
extern "C" {

dlg_status dlg_ctx_create(int device, dlg_ctx **out) {
    if (!out) return DLG_INVALID_INPUT;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return DLG_DEVICE_ERROR;
    if (device < 0 || device >= ndev) return DLG_INVALID_INPUT;
    dlg_ctx *c = new dlg_ctx();
    try {
        c->device = device;
        HIPCHK(hipSetDevice(device));
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, device));
        c->num_cus = prop.multiProcessorCount;
        c->hbm_bytes = (int64_t)prop.totalGlobalMem;
        c->lds_per_block = (int)prop.sharedMemPerBlock;
        c->arch = prop.gcnArchName;
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->pool = new Pool(device);
        HIPCHK(hipHostMalloc((void **)&c->h_scalars, 64 * sizeof(uint64_t), hipHostMallocDefault));
        const size_t mw = (size_t)dlg_ctx::MAPPED_WORDS * dlg_ctx::MAPPED_SLOTS;
        HIPCHK(hipHostMalloc((void **)&c->h_mapped, mw * sizeof(uint64_t), hipHostMallocMapped));
        std::memset(c->h_mapped, 0, mw * sizeof(uint64_t));   // no sequence number yet
        HIPCHK(hipHostGetDevicePointer((void **)&c->d_mapped, c->h_mapped, 0));
        noalloc_init(device);
    } catch (const Error &e) {
        fprintf(stderr, "dlg_ctx_create: %s\n", e.msg.c_str());
        dlg_ctx_destroy(c);
        return e.status;
    }
    *out = c;
    return DLG_OK;
}

void dlg_ctx_destroy(dlg_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &r : c->pending) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto e : c->event_free) (void)hipEventDestroy(e);
    delete c->pool;
    if (c->h_scalars) (void)hipHostFree(c->h_scalars);
    if (c->h_mapped) (void)hipHostFree(c->h_mapped);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *dlg_last_error(const dlg_ctx *c) { return c ? c->last_error.c_str() : "null context"; }

void dlg_host_free(void *p) { host_block_free(p); }

dlg_status dlg_ctx_sync(dlg_ctx *c) {
    if (!c) return DLG_INVALID_INPUT;
    API_BEGIN(c)
    c->sync();
    API_END(c)
}

dlg_status dlg_device_info(dlg_ctx *c, int32_t *num_cus, int64_t *hbm, int32_t *lds,
                           char *arch, int32_t arch_len) {
    if (!c) return DLG_INVALID_INPUT;
    if (num_cus) *num_cus = c->num_cus;
    if (hbm) *hbm = c->hbm_bytes;
    if (lds) *lds = c->lds_per_block;
    if (arch && arch_len > 0) {
        std::strncpy(arch, c->arch.c_str(), (size_t)arch_len - 1);
        arch[arch_len - 1] = 0;
    }
    return DLG_OK;
}

dlg_status dlg_device_pci_bus_id(dlg_ctx *c, char *bdf, int32_t len) {
    if (!c || !bdf || len < 13) return DLG_INVALID_INPUT;
    API_BEGIN(c)
    HIPCHK(hipDeviceGetPCIBusId(bdf, len, c->device));
    API_END(c)
}

dlg_status dlg_profiling(dlg_ctx *c, int enable) {
    if (!c) return DLG_INVALID_INPUT;
    API_BEGIN(c)
    if (!enable) c->prof_flush();
    c->prof = enable != 0;
    API_END(c)
}

const char *dlg_kernel_name(int k) {
    static const char *names[DLG_K_COUNT] = {
        "pack", "hash_build", "join_count", "scan", "join_expand", "sort_minmax",
        "sort_upsweep", "sort_downsweep", "select", "merge", "wcc_relax", "wcc_final",
        "hashset_insert", "bitmap_insert", "misc"};
    return (k >= 0 && k < DLG_K_COUNT) ? names[k] : "?";
}

dlg_status dlg_kernel_stats(dlg_ctx *c, int k, int64_t *launches, double *ms, double *bytes) {
    if (!c || k < 0 || k >= DLG_K_COUNT) return DLG_INVALID_INPUT;
    API_BEGIN(c)
    c->prof_flush();
    if (launches) *launches = c->kstat[k].launches;
    if (ms) *ms = c->kstat[k].ms;
    if (bytes) *bytes = c->kstat[k].bytes;
    API_END(c)
}

dlg_status dlg_kernel_stats_reset(dlg_ctx *c) {
    if (!c) return DLG_INVALID_INPUT;
    API_BEGIN(c)
    c->prof_flush();
    for (auto &s : c->kstat) s = KStat();
    API_END(c)
}

dlg_status dlg_dev_alloc(dlg_ctx *c, size_t bytes, void **p) {
    if (!c || !p) return DLG_INVALID_INPUT;
    API_BEGIN(c)
    *p = c->pool->get(bytes);
    API_END(c)
}

dlg_status dlg_dev_free(dlg_ctx *c, void *p) {
    if (!c) return DLG_INVALID_INPUT;
    if (!p) return DLG_OK;
    API_BEGIN(c)
    c->sync();
    c->pool->put(p);
    API_END(c)
}

dlg_status dlg_copy_h2d(dlg_ctx *c, void *d, const void *h, size_t bytes) {
    if (!c || (bytes && (!d || !h))) return DLG_INVALID_INPUT;
    API_BEGIN(c)
    if (bytes) {
        HIPCHK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c->stream));
        c->sync();
    }
    API_END(c)
}

dlg_status dlg_copy_d2h(dlg_ctx *c, void *h, const void *d, size_t bytes) {
    if (!c || (bytes && (!d || !h))) return DLG_INVALID_INPUT;
    API_BEGIN(c)
    if (bytes) {
        HIPCHK(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, c->stream));
        c->sync();
    }
    API_END(c)
}

}  // extern "C"
