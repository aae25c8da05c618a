// Candidate join + exact PAF check (replaces the merge-join of match_pairs,
// src/core/match/match_pairs.cpp:15-121, and check_if_pair, golay.cpp:27-44).
//
// The reference merge-joins two files sorted on the key text and runs
// check_if_pair on every (A, B) with equal key text.  Key text equality and
// key tuple equality select the same PAF-passing pairs: if (A, B) passes the
// PAF check, PSD_A(k) + PSD_B(k) = Q for all k != 0 with Q an integer, so the
// inverted B key is the A key shifted by the constant BOUND - Q, and two
// digit strings of tuples that differ by a nonzero constant shift cannot be
// equal (the first component keeps or changes its decimal length, and either
// way the strings differ).  The join therefore hashes the key tuples, sorts B
// by hash (LSD radix sort), and for every A scans the equal-hash range of B,
// verifying tuple equality and then the integer PAF condition.
#include <algorithm>

#include "gp_internal.hpp"
#include "gp_sort.hpp"
#include "psd_core.hpp"

namespace gp {

__device__ __forceinline__ uint64_t key_hash(const uint8_t* __restrict__ k, int nkey) {
    uint64_t h = 0x243F6A8885A308D3ull;
    for (int i = 0; i < nkey; i += 8) {
        uint64_t v = 0;
        for (int j = 0; j < 8 && i + j < nkey; ++j) v |= (uint64_t)k[i + j] << (8 * j);
        h = splitmix64(h ^ v);
    }
    return h;
}

__global__ void k_hash_keys(const uint8_t* __restrict__ keys, int64_t m, int nkey, int kstride,
                            uint64_t* __restrict__ h, uint32_t* __restrict__ idx) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        h[i] = key_hash(keys + i * kstride, nkey);
        if (idx) idx[i] = (uint32_t)i;
    }
}

__device__ bool paf_ok(const int8_t* __restrict__ a, const int8_t* __restrict__ b, int n, int target) {
    for (int s = 1; s <= n / 2; ++s) {
        int acc = 0;
        int j = s;
        for (int p = 0; p < n; ++p) {
            acc += (int)a[p] * (int)a[j] + (int)b[p] * (int)b[j];
            if (++j == n) j = 0;
        }
        if (acc != target) return false;
    }
    return true;
}

struct JoinCtl {
    unsigned long long out;
    unsigned long long checks;
};

__global__ void k_join(const int8_t* __restrict__ wa, const uint8_t* __restrict__ ka, int64_t na,
                       const int8_t* __restrict__ wb, const uint8_t* __restrict__ kb, int64_t nb,
                       const uint64_t* __restrict__ sh, const uint32_t* __restrict__ sidx, int n, int nkey,
                       int wstride, int kstride, int target, uint32_t* __restrict__ out_a,
                       uint32_t* __restrict__ out_b, int64_t cap, JoinCtl* ctl) {
    unsigned long long checks = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < na; i += (int64_t)gridDim.x * blockDim.x) {
        const uint8_t* k = ka + i * kstride;
        const uint64_t h = key_hash(k, nkey);
        int64_t lo = 0, hi = nb;
        while (lo < hi) {
            int64_t mid = (lo + hi) >> 1;
            if (sh[mid] < h) lo = mid + 1; else hi = mid;
        }
        for (int64_t j = lo; j < nb && sh[j] == h; ++j) {
            const uint32_t ib = sidx[j];
            const uint8_t* k2 = kb + (int64_t)ib * kstride;
            bool eq = true;
            for (int q = 0; q < nkey && eq; ++q) eq = k[q] == k2[q];
            if (!eq) continue;
            ++checks;
            if (paf_ok(wa + i * wstride, wb + (int64_t)ib * wstride, n, target)) {
                unsigned long long o = atomicAdd(&ctl->out, 1ull);
                if ((int64_t)o < cap) { out_a[o] = (uint32_t)i; out_b[o] = ib; }
            }
        }
    }
    block_atomic_add(&ctl->checks, checks);
}

JoinResult join_pairs(gp_ctx* ctx, const int8_t* d_wa, const uint8_t* d_ka, int64_t na, const int8_t* d_wb,
                      const uint8_t* d_kb, int64_t nb, int n, int nkey, int wstride, int kstride, int paf_target,
                      gp_stats* stats) {
    JoinResult res;
    if (na <= 0 || nb <= 0) return res;
    if (na > 0xFFFFFFFFll || nb > 0xFFFFFFFFll) throw ArgError("join: more than 2^32 candidates");
    hipStream_t st = ctx->stream;
    uint64_t* hb = ctx->dev<uint64_t>("j_hb", nb);
    uint32_t* ib = ctx->dev<uint32_t>("j_ib", nb);
    uint64_t* tk = ctx->dev<uint64_t>("j_tk", nb);
    uint32_t* tv = ctx->dev<uint32_t>("j_tv", nb);
    uint32_t* hist = static_cast<uint32_t*>(ctx->buf("j_hist").ensure(radix_hist_bytes(nb)));
    JoinCtl* ctl = ctx->dev<JoinCtl>("j_ctl", 1);
    GP_HIP(hipEventRecord(ctx->ev2, st));
    hipLaunchKernelGGL(k_hash_keys, dim3(ctx->num_cus * 4), dim3(256), 0, st, d_kb, nb, nkey, kstride, hb, ib);
    GP_HIP(hipGetLastError());
    radix_sort_pairs(st, hb, ib, tk, tv, hist, nb, 0, 64);
    int64_t cap = std::max<int64_t>(1024, 2 * std::max(na, nb));
    JoinCtl h{};
    for (int attempt = 0; attempt < 2; ++attempt) {
        uint32_t* oa = ctx->dev<uint32_t>("j_oa", cap);
        uint32_t* ob = ctx->dev<uint32_t>("j_ob", cap);
        GP_HIP(hipMemsetAsync(ctl, 0, sizeof(JoinCtl), st));
        hipLaunchKernelGGL(k_join, dim3((unsigned)std::min<int64_t>((na + 255) / 256, ctx->num_cus * 8)), dim3(256), 0,
                           st, d_wa, d_ka, na, d_wb, d_kb, nb, hb, ib, n, nkey, wstride, kstride, paf_target, oa, ob,
                           cap, ctl);
        GP_HIP(hipGetLastError());
        GP_HIP(hipMemcpyAsync(&h, ctl, sizeof h, hipMemcpyDeviceToHost, st));
        GP_HIP(hipStreamSynchronize(st));
        if ((int64_t)h.out <= cap) break;
        cap = (int64_t)h.out;
    }
    GP_HIP(hipEventRecord(ctx->ev3, st));
    res.ia.resize(h.out);
    res.ib.resize(h.out);
    if (h.out) {
        GP_HIP(hipMemcpyAsync(res.ia.data(), ctx->buf("j_oa").p, h.out * 4, hipMemcpyDeviceToHost, st));
        GP_HIP(hipMemcpyAsync(res.ib.data(), ctx->buf("j_ob").p, h.out * 4, hipMemcpyDeviceToHost, st));
    }
    GP_HIP(hipStreamSynchronize(st));
    float ms = 0;
    GP_HIP(hipEventElapsedTime(&ms, ctx->ev2, ctx->ev3));
    res.checks = h.checks;
    if (stats) { stats->pair_checks += h.checks; stats->kernel_ms += ms; }
    // emission order of the reference: A position outer, B position inner
    std::vector<size_t> order(h.out);
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](size_t x, size_t y) {
        return res.ia[x] != res.ia[y] ? res.ia[x] < res.ia[y] : res.ib[x] < res.ib[y];
    });
    JoinResult sorted;
    sorted.checks = res.checks;
    sorted.ia.resize(order.size());
    sorted.ib.resize(order.size());
    for (size_t i = 0; i < order.size(); ++i) { sorted.ia[i] = res.ia[order[i]]; sorted.ib[i] = res.ib[order[i]]; }
    return sorted;
}

}  // namespace gp
