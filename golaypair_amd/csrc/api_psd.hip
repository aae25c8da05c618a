// C ABI: context management and the batched PSD filter entry points.
#include <algorithm>
#include <cstring>
#include <map>
#include <numeric>
#include <set>
#include <tuple>

#include "gp_internal.hpp"
#include "psd_core.hpp"

using namespace gp;

namespace {

template <class F>
int guarded(gp_ctx* ctx, F&& f) {
    try {
        return f();
    } catch (const ArgError& e) {
        if (ctx) ctx->err = e.what();
        return GP_ERR_ARG;
    } catch (const HipError& e) {
        if (ctx) ctx->err = e.what();
        return GP_ERR_HIP;
    } catch (const std::exception& e) {
        if (ctx) ctx->err = e.what();
        return GP_ERR_UNSUPPORTED;
    }
}

// Microbenchmark interleave: slot(i) = (a i + b) mod count, a odd and coprime
// to count; kinds occupy contiguous slot ranges in mix order.
struct MbPerm {
    uint64_t a = 1, a_inv = 1, b = 0;
};

uint64_t modinv(uint64_t a, uint64_t m) {
    __int128 t = 0, nt = 1, r = (__int128)m, nr = (__int128)a;
    while (nr != 0) {
        __int128 q = r / nr;
        __int128 tmp = t - q * nt; t = nt; nt = tmp;
        tmp = r - q * nr; r = nr; nr = tmp;
    }
    if (t < 0) t += m;
    return (uint64_t)t;
}

MbPerm mb_perm(uint64_t seed, uint64_t count) {
    MbPerm p;
    if (count <= 1) return p;
    uint64_t a = (splitmix64(seed) % count) | 1ull;
    if (a >= count) a = 1;
    while (std::gcd(a, count) != 1) { a += 2; if (a >= count) a = 1; }
    p.a = a;
    p.a_inv = modinv(a, count);
    p.b = splitmix64(seed ^ 1ull) % count;
    return p;
}

// One word length of the microbenchmark: slots [slot_begin, slot_end), their
// rows `stride` bytes apart from byte byte0 of the words.
struct MbKind {
    int32_t n, c, alpha, table;
    int64_t slot_begin, slot_end;
    int32_t order, paf;
    int64_t byte0;
    int32_t stride, pad;
};

__global__ void k_mb_gen(uint64_t seed, uint64_t count, uint64_t a_inv, uint64_t b, const MbKind* __restrict__ kinds,
                         int nkind, int8_t* __restrict__ words, int64_t* __restrict__ cand_of_slot) {
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < count;
         s += (uint64_t)gridDim.x * blockDim.x) {
        int k = 0;
        while (k + 1 < nkind && (int64_t)s >= kinds[k + 1].slot_begin) ++k;
        const MbKind kd = kinds[k];
        const uint64_t i = (a_inv * ((s + count - b) % count)) % count;
        int8_t* w = words + kd.byte0 + (int64_t)(s - kd.slot_begin) * kd.stride;
        for (int j = 0; j < kd.stride; ++j) {
            int v = 0;
            if (j < kd.n) v = -kd.c + 2 * mb_letter(seed, i, j, kd.alpha);
            w[j] = (int8_t)v;
        }
        cand_of_slot[s] = (int64_t)i;
    }
}

__global__ void k_mb_digest(uint64_t seed, uint64_t count, const MbKind* __restrict__ kinds, int nkind,
                            const uint8_t* __restrict__ decision, const int64_t* __restrict__ cand_of_slot,
                            unsigned long long* __restrict__ accepted, unsigned long long* __restrict__ digest,
                            uint8_t* __restrict__ pass_by_cand, int64_t pass_cap) {
    // accepted slots counted per kind in registers (kinds are slot ranges, at most 8)
    unsigned long long dsum = 0, nacc[8] = {};
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < count;
         s += (uint64_t)gridDim.x * blockDim.x) {
        const int64_t i = cand_of_slot[s];
        const bool acc = decision[s] == kAccept;
        if (pass_by_cand && i < pass_cap) pass_by_cand[i] = acc ? 1 : 0;
        if (acc) {
            dsum += splitmix64((uint64_t)i ^ seed);
#pragma unroll
            for (int k = 0; k < 8; ++k)
                nacc[k] += k < nkind && (int64_t)s >= kinds[k].slot_begin && (int64_t)s < kinds[k].slot_end;
        }
    }
    block_atomic_add(digest, dsum, nkind > 0 ? &accepted[0] : nullptr, nacc[0], nkind > 1 ? &accepted[1] : nullptr,
                     nacc[1]);
    block_atomic_add(nkind > 2 ? &accepted[2] : nullptr, nacc[2], nkind > 3 ? &accepted[3] : nullptr, nacc[3],
                     nkind > 4 ? &accepted[4] : nullptr, nacc[4]);
    block_atomic_add(nkind > 5 ? &accepted[5] : nullptr, nacc[5], nkind > 6 ? &accepted[6] : nullptr, nacc[6],
                     nkind > 7 ? &accepted[7] : nullptr, nacc[7]);
}

}  // namespace

extern "C" {

int gp_create(gp_ctx** out, int device) {
    if (!out) return GP_ERR_ARG;
    *out = nullptr;
    gp_ctx* ctx = new gp_ctx();
    int rc = guarded(ctx, [&] {
        int n = 0;
        GP_HIP(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) throw ArgError("no HIP device " + std::to_string(device));
        GP_HIP(hipSetDevice(device));
        ctx->device = device;
        hipDeviceProp_t prop;
        GP_HIP(hipGetDeviceProperties(&prop, device));
        ctx->num_cus = prop.multiProcessorCount;
        psd_static_setup(device);  // device address of the static scratch of gp_psd_filter_batch_static
        load_tuning(ctx);
        GP_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        GP_HIP(hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
        for (auto& sl : ctx->slot_ev)
            for (auto& e : sl) GP_HIP(hipEventCreate(&e));
        GP_HIP(hipEventCreate(&ctx->ev0));
        GP_HIP(hipEventCreate(&ctx->ev1));
        GP_HIP(hipEventCreate(&ctx->ev2));
        GP_HIP(hipEventCreate(&ctx->ev3));
        return GP_OK;
    });
    if (rc != GP_OK) {
        delete ctx;
        return rc;
    }
    *out = ctx;
    return GP_OK;
}

void gp_destroy(gp_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->copy_stream) (void)hipStreamSynchronize(ctx->copy_stream);
    for (hipEvent_t e : {ctx->ev0, ctx->ev1, ctx->ev2, ctx->ev3})
        if (e) (void)hipEventDestroy(e);
    for (auto& sl : ctx->slot_ev)
        for (hipEvent_t e : sl)
            if (e) (void)hipEventDestroy(e);
    for (hipGraphExec_t g : ctx->unc_graph)
        if (g) (void)hipGraphExecDestroy(g);
    hipStream_t s = ctx->stream, s2 = ctx->copy_stream;
    delete ctx;
    if (s) (void)hipStreamDestroy(s);
    if (s2) (void)hipStreamDestroy(s2);
}

const char* gp_last_error(const gp_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int gp_set_psd_mode(gp_ctx* ctx, int mode) {
    if (!ctx) return GP_ERR_ARG;
    if (mode != GP_PSD_SPLIT_BF16 && mode != GP_PSD_FP32) {
        ctx->err = "unknown PSD mode";
        return GP_ERR_ARG;
    }
    ctx->psd_mode = mode;
    return GP_OK;
}

int gp_psd_filter_batch(gp_ctx* ctx, const int8_t* words, int stride, const int32_t* lens, const int32_t* orders,
                        const int32_t* pafs, int64_t count, uint8_t* pass, float* max_psd, gp_stats* stats) {
    if (!ctx) return GP_ERR_ARG;
    return guarded(ctx, [&] {
        GP_HIP(hipSetDevice(ctx->device));
        if (count < 0 || (count > 0 && (!words || !lens || !orders || !pafs || !pass)))
            throw ArgError("gp_psd_filter_batch: null buffer");
        if (count == 0) return GP_OK;
        // Group rows by (n, order, paf): one segment per group, its rows at 16
        // bytes per block of 16 letters (zero padded).
        std::map<std::tuple<int, int, int>, std::vector<int64_t>> groups;
        for (int64_t i = 0; i < count; ++i) {
            if (lens[i] < 2 || lens[i] > kMaxN || lens[i] > stride) throw ArgError("word length out of range");
            groups[{lens[i], orders[i], pafs[i]}].push_back(i);
        }
        int64_t bytes = 0;
        for (auto& [key, rows] : groups) bytes += (int64_t)rows.size() * 16 * ((std::get<0>(key) + 15) / 16);
        std::vector<int8_t> packed((size_t)bytes, 0);
        std::vector<int64_t> row_of(count);
        std::vector<BatchSeg> segs;
        std::vector<FilterThr> thr(ctx->tables.desc.size() + groups.size());
        int64_t r = 0, off = 0;
        std::vector<FilterThr> seg_thr;
        for (auto& [key, rows] : groups) {
            auto [n, order, paf] = key;
            const int table = ctx->tables.get(n);
            BatchSeg sg{};
            sg.first_row = r;
            sg.rows = (int64_t)rows.size();
            sg.stride = 16 * ctx->tables.desc[table].nb;
            sg.byte0 = off;
            double s1 = 0;
            for (int64_t i : rows) {
                double a = 0;
                for (int p = 0; p < n; ++p) {
                    int8_t v = words[i * stride + p];
                    packed[off + p] = v;
                    a += std::abs((int)v);
                }
                s1 = std::max(s1, a);
                row_of[i] = r++;
                off += sg.stride;
            }
            const PsdTableDesc& d = ctx->tables.desc[table];
            const double th = 2.0 * order - paf + 0.001;
            const double E = ctx->psd_mode == GP_PSD_SPLIT_BF16 ? split_bf16_E(d, s1) : fp32_E(n, s1);
            // Each segment gets its own pseudo-table id so thresholds may differ
            // between segments of one length.
            seg_thr.push_back(make_thr(th, E));
            sg.table = table;
            segs.push_back(sg);
        }
        // Thresholds are indexed by table id; with several (order, paf) per n we
        // run one batch per distinct threshold set.
        std::vector<uint8_t> dec((size_t)count);
        std::vector<float> mx(max_psd ? (size_t)count : 0);
        int8_t* d_words = static_cast<int8_t*>(ctx->words.ensure((size_t)bytes));
        uint8_t* d_dec = static_cast<uint8_t*>(ctx->decisions.ensure((size_t)count));
        float* d_max = max_psd ? static_cast<float*>(ctx->aux4.ensure((size_t)count * sizeof(float))) : nullptr;
        GP_HIP(hipMemcpyAsync(d_words, packed.data(), packed.size(), hipMemcpyHostToDevice, ctx->stream));
        size_t si = 0;
        while (si < segs.size()) {
            // collect a run of segments with distinct tables
            std::vector<BatchSeg> run;
            std::vector<FilterThr> tt(ctx->tables.desc.size());
            std::set<int> used;
            size_t sj = si;
            for (; sj < segs.size() && sj - si < 8 && !used.count(segs[sj].table); ++sj) {
                used.insert(segs[sj].table);
                run.push_back(segs[sj]);
                tt[segs[sj].table] = seg_thr[sj];
            }
            run_psd_filter(ctx, d_words, 0, (int64_t)count, run, tt, d_dec, stats, d_max);
            si = sj;
        }
        GP_HIP(hipMemcpyAsync(dec.data(), d_dec, (size_t)count, hipMemcpyDeviceToHost, ctx->stream));
        if (max_psd)
            GP_HIP(hipMemcpyAsync(mx.data(), d_max, (size_t)count * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        GP_HIP(hipStreamSynchronize(ctx->stream));
        uint64_t npass = 0;
        for (int64_t i = 0; i < count; ++i) {
            pass[i] = dec[row_of[i]] == kAccept ? 1 : 0;
            npass += pass[i];
            if (max_psd) max_psd[i] = mx[row_of[i]];
        }
        if (stats) stats->psd_pass += npass;
        return GP_OK;
    });
}

int gp_microbench_words(uint64_t seed, int64_t count, const int32_t* mix, int nmix, int64_t first, int64_t rows,
                        int32_t* kinds, int8_t* out, int stride) {
    if (!mix || nmix <= 0 || count <= 0 || first < 0 || rows < 0 || first + rows > count || !out) return GP_ERR_ARG;
    int maxn = 0;
    for (int k = 0; k < nmix; ++k) maxn = std::max(maxn, (int)mix[5 * k]);
    if (stride < maxn) return GP_ERR_ARG;
    MbPerm p = mb_perm(seed, (uint64_t)count);
    for (int64_t r = 0; r < rows; ++r) {
        const uint64_t i = (uint64_t)(first + r);
        uint64_t s = (uint64_t)(((unsigned __int128)p.a * i + p.b) % (uint64_t)count);
        int64_t acc = 0;
        int k = 0;
        for (; k < nmix; ++k) {
            if ((int64_t)s < acc + mix[5 * k + 4]) break;
            acc += mix[5 * k + 4];
        }
        if (k == nmix) return GP_ERR_ARG;
        const int n = mix[5 * k], c = mix[5 * k + 1];
        if (kinds) kinds[r] = k;
        int8_t* w = out + r * stride;
        for (int j = 0; j < stride; ++j) w[j] = j < n ? (int8_t)(-c + 2 * mb_letter(seed, i, j, c + 1)) : 0;
    }
    return GP_OK;
}

int gp_psd_microbench(gp_ctx* ctx, uint64_t seed, int64_t count, const int32_t* mix, int nmix, uint64_t* accepted,
                      uint64_t* digest, uint8_t* pass_out, int64_t pass_cap, gp_stats* stats) {
    if (!ctx) return GP_ERR_ARG;
    return guarded(ctx, [&] {
        GP_HIP(hipSetDevice(ctx->device));
        if (!mix || nmix <= 0 || nmix > 8 || count <= 0) throw ArgError("gp_psd_microbench: bad mix");
        int64_t total = 0;
        for (int k = 0; k < nmix; ++k) total += mix[5 * k + 4];
        if (total != count) throw ArgError("gp_psd_microbench: shares must sum to count");
        MbPerm p = mb_perm(seed, (uint64_t)count);
        std::vector<MbKind> kinds(nmix);
        std::vector<BatchSeg> segs;
        std::vector<FilterThr> tt;
        // rows of each length at 16 bytes per block of 16 letters, one region per length
        int64_t acc = 0, bytes = 0;
        for (int k = 0; k < nmix; ++k) {
            MbKind& kd = kinds[k];
            kd.n = mix[5 * k]; kd.c = mix[5 * k + 1]; kd.order = mix[5 * k + 2]; kd.paf = mix[5 * k + 3];
            kd.alpha = kd.c + 1;
            kd.table = ctx->tables.get(kd.n);
            kd.slot_begin = acc;
            acc += mix[5 * k + 4];
            kd.slot_end = acc;
            kd.stride = 16 * ctx->tables.desc[kd.table].nb;
            kd.byte0 = bytes;
            bytes += (kd.slot_end - kd.slot_begin) * kd.stride;
        }
        tt.resize(ctx->tables.desc.size());
        std::set<int> used;
        for (int k = 0; k < nmix; ++k) {
            const PsdTableDesc& d = ctx->tables.desc[kinds[k].table];
            if (used.count(kinds[k].table)) throw ArgError("gp_psd_microbench: one kind per word length");
            used.insert(kinds[k].table);
            const double s1 = (double)kinds[k].n * kinds[k].c;
            const double th = 2.0 * kinds[k].order - kinds[k].paf + 0.001;
            const double E = ctx->psd_mode == GP_PSD_SPLIT_BF16 ? split_bf16_E(d, s1) : fp32_E(kinds[k].n, s1);
            tt[kinds[k].table] = make_thr(th, E);
            segs.push_back(BatchSeg{kinds[k].table, kinds[k].stride, kinds[k].slot_begin,
                                    kinds[k].slot_end - kinds[k].slot_begin, kinds[k].byte0});
        }
        hipStream_t st = ctx->stream;
        int8_t* d_words = static_cast<int8_t*>(ctx->words.ensure((size_t)bytes));
        uint8_t* d_dec = static_cast<uint8_t*>(ctx->decisions.ensure((size_t)count));
        int64_t* d_cand = static_cast<int64_t*>(ctx->aux0.ensure((size_t)count * sizeof(int64_t)));
        MbKind* d_kinds = static_cast<MbKind*>(ctx->aux1.ensure(sizeof(MbKind) * nmix));
        unsigned long long* d_cnt = static_cast<unsigned long long*>(ctx->aux2.ensure(sizeof(unsigned long long) * 9));
        uint8_t* d_pass = pass_out ? static_cast<uint8_t*>(ctx->aux3.ensure((size_t)std::max<int64_t>(pass_cap, 1))) : nullptr;
        GP_HIP(hipMemcpyAsync(d_kinds, kinds.data(), sizeof(MbKind) * nmix, hipMemcpyHostToDevice, st));
        GP_HIP(hipMemsetAsync(d_cnt, 0, sizeof(unsigned long long) * 9, st));
        hipLaunchKernelGGL(k_mb_gen, dim3(ctx->num_cus * 8), dim3(256), 0, st, seed, (uint64_t)count, p.a_inv, p.b,
                           d_kinds, nmix, d_words, d_cand);
        GP_HIP(hipGetLastError());
        run_psd_filter(ctx, d_words, 0, count, segs, tt, d_dec, stats);
        hipLaunchKernelGGL(k_mb_digest, dim3(ctx->num_cus * 4), dim3(256), 0, st, seed, (uint64_t)count, d_kinds,
                           nmix, d_dec, d_cand, d_cnt, d_cnt + 8, d_pass, pass_cap);
        GP_HIP(hipGetLastError());
        unsigned long long h[9];
        GP_HIP(hipMemcpyAsync(h, d_cnt, sizeof h, hipMemcpyDeviceToHost, st));
        if (pass_out) GP_HIP(hipMemcpyAsync(pass_out, d_pass, (size_t)pass_cap, hipMemcpyDeviceToHost, st));
        GP_HIP(hipStreamSynchronize(st));
        uint64_t npass = 0;
        for (int k = 0; k < nmix; ++k) { if (accepted) accepted[k] = h[k]; npass += h[k]; }
        if (digest) *digest = h[8];
        if (stats) stats->psd_pass += npass;
        return GP_OK;
    });
}

}  // extern "C"
