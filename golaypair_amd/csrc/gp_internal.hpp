// Host-side internals shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <mutex>
#include <string>
#include <vector>

#include "gp_common.hpp"
#include "gp_host.hpp"
#include "golay_gpu.h"

namespace gp {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct ArgError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define GP_HIP(expr)                                                                              \
    do {                                                                                          \
        hipError_t _e = (expr);                                                                   \
        if (_e != hipSuccess)                                                                     \
            throw ::gp::HipError(std::string(#expr) + ": " + hipGetErrorString(_e) + " at " +   \
                                 __FILE__ + ":" + std::to_string(__LINE__));                      \
    } while (0)

// Growable device allocation owned by the context.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { if (p) (void)hipFree(p); }
    void* ensure(size_t n) {
        if (n <= bytes) return p;
        if (p) { GP_HIP(hipFree(p)); p = nullptr; bytes = 0; }
        size_t want = n < 256 ? 256 : n;
        GP_HIP(hipMalloc(&p, want));
        bytes = want;
        return p;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// One PSD table per word length: MFMA A fragments, FP32 and FP64 twiddles.
struct PsdTables {
    std::vector<PsdTableDesc> desc;       // index = table id
    std::map<int, int> by_n;              // n -> table id
    std::vector<uint4> frags;             // all tables' A fragments
    std::vector<float> trig32;            // per table: cos[n], sin[n]
    std::vector<double> trig64;           // per table: cos[n], sin[n]
    DevBuf d_desc, d_frags, d_trig32, d_trig64;
    bool dirty = true;
    int get(int n);                       // builds the table on first use
    void upload(hipStream_t s);
};

// Per-table filter thresholds for one call: accept if max < lo, reject if
// max > hi, FP64 recheck otherwise; thr = 2*order - paf + 0.001.
struct FilterThr {
    float lo, hi;
    double thr;
};

struct TileDesc {
    int32_t table;
    int32_t rows;
    int64_t first_row;
};

// Host wall-clock phases, printed to stderr when GOLAY_PROFILE is set.
struct Phase {
    const char* name;
    std::chrono::steady_clock::time_point t0;
    explicit Phase(const char* n) : name(n), t0(std::chrono::steady_clock::now()) {}
    ~Phase() { report(); }
    // ends this phase and starts the next one
    void next(const char* n) {
        report();
        name = n;
        t0 = std::chrono::steady_clock::now();
    }
    void report() const {
        static const bool on = std::getenv("GOLAY_PROFILE") != nullptr;
        static const auto epoch = std::chrono::steady_clock::now();
        if (on) {
            const auto t1 = std::chrono::steady_clock::now();
            std::fprintf(stderr, "[golay] %-28s %9.3f ms  (ends at %.3f ms)\n", name,
                         std::chrono::duration<double, std::milli>(t1 - t0).count(),
                         std::chrono::duration<double, std::milli>(t1 - epoch).count());
        }
    }
};

// dst += src, field by field.
inline void add_stats(gp_stats* dst, const gp_stats& src) {
    if (!dst) return;
    dst->words += src.words;
    dst->psd_decisions += src.psd_decisions;
    dst->psd_pass += src.psd_pass;
    dst->band_rechecks += src.band_rechecks;
    dst->key_rechecks += src.key_rechecks;
    dst->canonical += src.canonical;
    dst->pair_checks += src.pair_checks;
    dst->pairs += src.pairs;
    dst->tie_band += src.tie_band;
    dst->kernel_ms += src.kernel_ms;
    dst->psd_kernel_ms += src.psd_kernel_ms;
    dst->band_kernel_ms += src.band_kernel_ms;
    dst->psd_launches += src.psd_launches;
}

// Pinned host buffers for staging uploads, shared by host threads.  `upload`
// (created on first use) is recorded on the stream that copies the buffer to
// the device; acquire() waits for it, so a buffer handed back while its copy is
// still in flight (a batch dropped by an exception) is not overwritten early.
struct PinnedBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool busy = false;
    hipEvent_t upload = nullptr;
    bool upload_pending = false;
    // marks the buffer as read by work queued on `st` up to now
    void record_upload(hipStream_t st) {
        if (!upload) GP_HIP(hipEventCreateWithFlags(&upload, hipEventDisableTiming));
        GP_HIP(hipEventRecord(upload, st));
        upload_pending = true;
    }
};
struct PinnedPool {
    std::mutex mu;
    std::vector<std::unique_ptr<PinnedBuf>> bufs;
    size_t high = 0;  // largest request so far: a buffer that must grow grows to it
    PinnedBuf* acquire(size_t n) {
        std::lock_guard<std::mutex> lk(mu);
        high = std::max(high, n);
        PinnedBuf* pick = nullptr;  // the smallest free buffer that fits
        for (auto& b : bufs)
            if (!b->busy && b->bytes >= n && (!pick || b->bytes < pick->bytes)) pick = b.get();
        if (!pick)
            for (auto& b : bufs)
                if (!b->busy) { pick = b.get(); break; }
        if (!pick) {
            bufs.emplace_back(new PinnedBuf());
            pick = bufs.back().get();
        }
        if (pick->bytes < n) {
            if (pick->p) GP_HIP(hipHostFree(pick->p));
            pick->p = nullptr;
            pick->bytes = 0;
            // sized for the largest request seen, so that requests of varying
            // sizes stop reallocating once every buffer has grown
            const size_t want = high + high / 4 + 4096;
            GP_HIP(hipHostMalloc(&pick->p, want, hipHostMallocDefault));
            pick->bytes = want;
        }
        pick->busy = true;
        if (pick->upload_pending) {  // its last copy to the device must be done before it is refilled
            GP_HIP(hipEventSynchronize(pick->upload));
            pick->upload_pending = false;
        }
        return pick;
    }
    void release(PinnedBuf* b) {
        std::lock_guard<std::mutex> lk(mu);
        if (b) b->busy = false;
    }
    ~PinnedPool() {
        for (auto& b : bufs) {
            if (b->upload) {
                (void)hipEventSynchronize(b->upload);
                (void)hipEventDestroy(b->upload);
            }
            if (b->p) (void)hipHostFree(b->p);
        }
    }
};

}  // namespace gp

struct gp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t copy_stream = nullptr;  // read-backs that must not wait for later work on `stream`
    hipEvent_t slot_ev[2][5] = {};      // per batch slot: filter start, filter end, band end, batch end, tables uploaded
    int64_t unc_ecap = 1 << 16;         // emitted-pair rows per in-flight batch (grows on overflow)
    hipGraphExec_t unc_graph[2] = {};   // per batch slot: the band recheck + re-match chain
    std::map<std::string, int64_t> tuning;  // measured launch / batch parameters (tuning.hip)
    uint64_t unc_graph_launches = 0, unc_graph_instantiations = 0;
    std::vector<unsigned char> unc_side_buf;  // storage of the uncompress sides (reused across calls)
    int psd_mode = GP_PSD_SPLIT_BF16;
    std::string err;
    gp::PsdTables tables;
    // scratch owned by the context
    gp::DevBuf words, decisions, counters, thr, aux0, aux1, aux2, aux3, aux4, aux5;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
    int num_cus = 256;
    // result of the last gp_uncompress that failed with GP_ERR_CAPACITY
    bool unc_cache_valid = false;
    uint64_t unc_cache_key = 0;
    gp::PairSet unc_cache_out;
    gp::CandSet unc_cache_ta, unc_cache_tb;
    gp_stats unc_cache_stats{};
    gp::PinnedPool pinned;  // staging for batch uploads
    // named scratch buffers of the pipeline stages
    std::map<std::string, std::unique_ptr<gp::DevBuf>> pool;
    gp::DevBuf& buf(const std::string& name) {
        auto& p = pool[name];
        if (!p) p.reset(new gp::DevBuf());
        return *p;
    }
    template <class T> T* dev(const std::string& name, size_t count) {
        return static_cast<T*>(buf(name).ensure(count * sizeof(T) + 64));
    }
};

namespace gp {

// Tuning table (tuning.hip): loaded from the tuning file when a context is
// created; tuned(ctx, key, dflt) is the table's value or dflt.
void load_tuning(gp_ctx* ctx);
int64_t tuned(const gp_ctx* ctx, const std::string& key, int64_t dflt);

// Looks up (once per device, with that device current) the device address of
// the static scratch array used by gp_psd_filter_batch_static.
void psd_static_setup(int device);

// Error bound on Re/Im of the split-BF16 contraction for words with
// sum |x| <= s1 using table `d`.
double split_bf16_E(const PsdTableDesc& d, double s1);
// Same for the plain FP32 direct DFT.
double fp32_E(int n, double s1);
// Thresholds for `n` with threshold thr, error bound E.
FilterThr make_thr(double thr, double E);

// Filters rows [0, nrows) of `d_words` (stride bytes per row, rows zero
// padded to 16*ceil(n/16)).  `segs` lists the rows as per-table segments (at
// most 8); the split kernel cuts them into 32-row tiles on the device.
// Writes decisions (kReject/kAccept) after the FP64 recheck of band rows.
// Returns PSD-kernel device time in ms; stats counts band rechecks and ties.
// Rows [first_row, first_row + rows) of one PSD batch, all of one word length:
// row first_row + i is at byte byte0 + i * stride of the words.  stride 0 takes
// the batch's stride, with byte0 = first_row * that stride.
struct BatchSeg {
    int32_t table;
    int32_t stride;
    int64_t first_row;
    int64_t rows;
    int64_t byte0;
};
double run_psd_filter(gp_ctx* ctx, const int8_t* d_words, int stride, int64_t nrows,
                      const std::vector<BatchSeg>& segs, const std::vector<FilterThr>& thr_by_table,
                      uint8_t* d_decision, gp_stats* stats, float* d_maxpsd = nullptr);

}  // namespace gp
