// C ABI: generation, sort, match and key entry points.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <string>

#include "gp_host.hpp"
#include "gp_internal.hpp"
#include "gp_sort.hpp"
#include "psd_core.hpp"
#include "textfmt.hpp"

using namespace gp;

namespace gp {
int guarded_call(gp_ctx* ctx, const std::function<int()>& f) {
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

// Keys of arbitrary words in FP64 (k_keys_fp64 semantics without a filter).
__global__ void k_keys_all(const int8_t* __restrict__ words, int stride, int64_t m, int n,
                           const double* __restrict__ trig, int bound, int inverted, uint8_t* __restrict__ keys,
                           int nkey, unsigned int* __restrict__ bad) {
    double psd[kMaxK];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        fp64_psd(words + i * stride, n, trig, trig + n, psd);
        for (int k = 0; k < n / 2; ++k) {
            int v = (int)rint(psd[k]);
            if (inverted) v = bound - v;
            if (v < 0 || v > 255) { atomicAdd(bad, 1u); v = 0; }
            keys[i * nkey + k] = (uint8_t)v;
        }
    }
}
}  // namespace gp

namespace {
struct GenCacheKey {
    int order, compress, paf, id, num, mode;
    bool operator==(const GenCacheKey& o) const {
        return order == o.order && compress == o.compress && paf == o.paf && id == o.id && num == o.num &&
               mode == o.mode;
    }
};
thread_local GenCacheKey g_gen_key{-1, 0, 0, 0, 0, 0};
thread_local GenOutput g_gen_out;
thread_local gp_stats g_gen_stats;
}  // namespace

extern "C" {

int gp_generate(gp_ctx* ctx, int order, int compress, int paf, int proc_id, int proc_num, gp_cands* out_a,
                gp_cands* out_b, gp_stats* stats) {
    if (!ctx) return GP_ERR_ARG;
    return guarded_call(ctx, [&] {
        GP_HIP(hipSetDevice(ctx->device));
        if (order <= 0 || compress <= 0 || proc_num <= 0 || proc_id < 0 || proc_id >= proc_num)
            throw ArgError("gp_generate: bad arguments");
        GenCacheKey key{order, compress, paf, proc_id, proc_num, ctx->psd_mode};
        if (!(key == g_gen_key)) {
            GenOutput o;
            gp_stats s{};
            int rc = generate_level(ctx, order, compress, paf, proc_id, proc_num, o, &s);
            if (rc != GP_OK) return rc;
            g_gen_out = std::move(o);
            g_gen_stats = s;
            g_gen_key = key;
        }
        int ra = export_cands(g_gen_out.a, out_a);
        int rb = export_cands(g_gen_out.b, out_b);
        if (ra == GP_ERR_CAPACITY || rb == GP_ERR_CAPACITY) return GP_ERR_CAPACITY;
        if (ra != GP_OK) return ra;
        if (rb != GP_OK) return rb;
        add_stats(stats, g_gen_stats);  // reported once, by the call that returns the candidates
        g_gen_key = GenCacheKey{-1, 0, 0, 0, 0, 0};
        g_gen_out = GenOutput{};
        return GP_OK;
    });
}

int gp_sort_candidates(gp_ctx* ctx, gp_cands* c) {
    if (!ctx) return GP_ERR_ARG;
    return guarded_call(ctx, [&] {
        GP_HIP(hipSetDevice(ctx->device));
        if (!c) throw ArgError("gp_sort_candidates: null argument");
        const int64_t m = c->count;
        if (m <= 0) return GP_OK;
        const int n = c->n, K = c->nkey;
        if (K != n / 2) throw ArgError("gp_sort_candidates: nkey must be n/2");
        if (!c->words || !c->keys) throw ArgError("gp_sort_candidates: null buffers");
        // records in, sorted and unique records out, on the device
        hipStream_t st = ctx->stream;
        int8_t* dw = ctx->dev<int8_t>("s_w", (size_t)m * n);
        uint8_t* dk = ctx->dev<uint8_t>("s_k", (size_t)m * K);
        int8_t* ow = ctx->dev<int8_t>("s_ow", (size_t)m * n);
        uint8_t* ok = ctx->dev<uint8_t>("s_ok", (size_t)m * K + 1);
        GP_HIP(hipMemcpyAsync(dw, c->words, (size_t)m * n, hipMemcpyHostToDevice, st));
        GP_HIP(hipMemcpyAsync(dk, c->keys, (size_t)m * K, hipMemcpyHostToDevice, st));
        const int64_t kept = sort_uniq_cands(ctx, dw, dk, m, n, K, ow, ok);
        GP_HIP(hipMemcpyAsync(c->words, ow, (size_t)kept * n, hipMemcpyDeviceToHost, st));
        if (K > 0) GP_HIP(hipMemcpyAsync(c->keys, ok, (size_t)kept * K, hipMemcpyDeviceToHost, st));
        GP_HIP(hipStreamSynchronize(st));
        c->count = kept;
        return GP_OK;
    });
}

int gp_match(gp_ctx* ctx, int order, int compress, int paf, const gp_cands* a_sorted, const gp_cands* b_sorted,
             gp_pairs* out, gp_stats* stats) {
    if (!ctx) return GP_ERR_ARG;
    return guarded_call(ctx, [&] {
        GP_HIP(hipSetDevice(ctx->device));
        if (!a_sorted || !b_sorted || !out) throw ArgError("gp_match: null argument");
        if (a_sorted->count == 0 || b_sorted->count == 0) return GP_NO_INPUT;  // match_pairs.cpp:32-38
        const int n = order / compress;
        if (a_sorted->n != n || b_sorted->n != n) throw ArgError("gp_match: word length != order/compress");
        if (a_sorted->nkey != n / 2 || b_sorted->nkey != n / 2) throw ArgError("gp_match: nkey must be n/2");
        hipStream_t st = ctx->stream;
        const int64_t na = a_sorted->count, nb = b_sorted->count;
        int8_t* wa = ctx->dev<int8_t>("m_wa", (size_t)na * n);
        int8_t* wb = ctx->dev<int8_t>("m_wb", (size_t)nb * n);
        uint8_t* ka = ctx->dev<uint8_t>("m_ka", (size_t)na * (n / 2));
        uint8_t* kb = ctx->dev<uint8_t>("m_kb", (size_t)nb * (n / 2));
        GP_HIP(hipMemcpyAsync(wa, a_sorted->words, (size_t)na * n, hipMemcpyHostToDevice, st));
        GP_HIP(hipMemcpyAsync(wb, b_sorted->words, (size_t)nb * n, hipMemcpyHostToDevice, st));
        GP_HIP(hipMemcpyAsync(ka, a_sorted->keys, (size_t)na * (n / 2), hipMemcpyHostToDevice, st));
        GP_HIP(hipMemcpyAsync(kb, b_sorted->keys, (size_t)nb * (n / 2), hipMemcpyHostToDevice, st));
        JoinResult jr = join_pairs(ctx, wa, ka, na, wb, kb, nb, n, n / 2, n, n / 2, paf * compress, stats);
        PairSet ps;
        ps.n = n;
        ps.a.resize(jr.ia.size() * n);
        ps.b.resize(jr.ia.size() * n);
        for (size_t i = 0; i < jr.ia.size(); ++i) {
            std::memcpy(&ps.a[i * n], a_sorted->words + (int64_t)jr.ia[i] * n, n);
            std::memcpy(&ps.b[i * n], b_sorted->words + (int64_t)jr.ib[i] * n, n);
        }
        if (stats) stats->pairs += jr.ia.size();
        return export_pairs(ps, out, true);
    });
}

int gp_psd_keys(gp_ctx* ctx, const int8_t* words, int stride, int n, int64_t count, int bound, int inverted,
                uint8_t* keys, gp_stats* stats) {
    if (!ctx) return GP_ERR_ARG;
    return guarded_call(ctx, [&] {
        GP_HIP(hipSetDevice(ctx->device));
        if (count < 0 || n < 2 || n > kMaxN || stride < n || (count > 0 && (!words || !keys)))
            throw ArgError("gp_psd_keys: bad arguments");
        if (count == 0) return GP_OK;
        hipStream_t st = ctx->stream;
        const int table = ctx->tables.get(n);
        ctx->tables.upload(st);
        const PsdTableDesc& d = ctx->tables.desc[table];
        int8_t* dw = ctx->dev<int8_t>("k_w", (size_t)count * stride);
        uint8_t* dk = ctx->dev<uint8_t>("k_k", (size_t)count * (n / 2));
        unsigned int* bad = ctx->dev<unsigned int>("k_bad", 1);
        GP_HIP(hipMemcpyAsync(dw, words, (size_t)count * stride, hipMemcpyHostToDevice, st));
        GP_HIP(hipMemsetAsync(bad, 0, 4, st));
        hipLaunchKernelGGL(k_keys_all, dim3(ctx->num_cus * 4), dim3(256), 0, st, dw, stride, count, n,
                           ctx->tables.d_trig64.as<double>() + d.trig_off, bound, inverted, dk, n / 2, bad);
        GP_HIP(hipGetLastError());
        unsigned int hb = 0;
        GP_HIP(hipMemcpyAsync(keys, dk, (size_t)count * (n / 2), hipMemcpyDeviceToHost, st));
        GP_HIP(hipMemcpyAsync(&hb, bad, 4, hipMemcpyDeviceToHost, st));
        GP_HIP(hipStreamSynchronize(st));
        if (hb) throw ArgError("gp_psd_keys: key digit outside [0, 255]");
        (void)stats;
        return GP_OK;
    });
}

}  // extern "C"
