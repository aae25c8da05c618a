// Device sort + uniq of candidate records in the byte order of their text lines
// (`sort | uniq` under LC_ALL=C, src/core/sort/sort.cpp:7-23, of the lines that
// src/utils/io.cpp:14-36 writes: "PSD" d_1 .. d_{n-1} " " v_0 " " .. v_{n-1} " ").
//
// Line order without the text.  Each record gets a fixed-width key of symbols,
// packed most significant first into W 64-bit words (column-major, word w of
// record i at keyw[w * m + i]):
//   3 (n - 1) text symbols of 4 bits: the key digits' decimal text with '0'..'9'
//       as 1..10, followed by zeros up to the full width;
//   n letter symbols of `lbits` bits: each letter's rank among the letters
//       present, ordered as their token strings "v " compare bytewise.
// Comparing keys word by word compares the lines bytewise: where one key text
// is a proper prefix of the other, that line has ' ' (below every digit) where
// the other has a digit, as the zero symbol is below every digit symbol; with
// equal texts both lines continue with " " and then the tokens, which are
// prefix-free (each ends in its only space), so the lines compare as the
// sequences of the tokens' ranks.  Equal keys are equal lines.
//
// Sort: LSD radix sort (k_rs_hist, a scan of the block histograms, k_rs_scatter
// with digit ranks from ballots and LDS counts), least significant key word first, each word
// gathered through the current permutation; only the key bytes that differ
// between some two records get a pass (their mask is an OR / AND reduction of
// the keys).  Uniq: a record whose key equals its predecessor's is dropped;
// kept records are numbered by a scan and gathered in order.
//
// HBM traffic (algorithmic, 2 N R in the sense of one read and one write of
// each record; R = n + n/2 bytes here): the key build reads each record once
// and writes W words; each radix pass reads 12 B twice and writes 12 B per
// record; the compaction reads and writes each kept record once.
#include <algorithm>
#include <array>
#include <cstring>
#include <string>
#include <vector>

#include "gp_internal.hpp"
#include "gp_sort.hpp"

namespace gp {

constexpr int kSortMaxW = 24;  // 4 * 3 * 63 + 8 * 64 bits < 24 * 64

__device__ __forceinline__ void put_sym(uint64_t& cur, int& used, uint32_t sym, int bits, uint64_t* __restrict__ col,
                                        int64_t m, int64_t i, int& w) {
    // append `bits` bits; a full word is flushed to its column
    const int room = 64 - used;
    if (bits <= room) {
        cur = (cur << bits) | sym;
        used += bits;
    } else {
        const int hi = bits - room;  // bits that spill into the next word
        cur = (cur << room) | (uint64_t)(sym >> hi);
        col[(int64_t)w * m + i] = cur;
        ++w;
        cur = sym & ((1u << hi) - 1u);
        used = hi;
    }
    if (used == 64) {
        col[(int64_t)w * m + i] = cur;
        ++w;
        cur = 0;
        used = 0;
    }
}

// One record per thread: the key words of record i.
__global__ void __launch_bounds__(256) k_cand_keys(const int8_t* __restrict__ words, const uint8_t* __restrict__ keys,
                                                   int64_t m, int n, int K, int lbits,
                                                   const uint8_t* __restrict__ lrank, uint64_t* __restrict__ keyw) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    uint64_t cur = 0;
    int used = 0, w = 0, nsym = 0;
    const uint8_t* kd = keys + i * K;
    for (int k = 1; k < n; ++k) {
        const uint32_t d = kd[(k <= K ? k : n - k) - 1];
        if (d >= 100) { put_sym(cur, used, d / 100 + 1, 4, keyw, m, i, w); ++nsym; }
        if (d >= 10) { put_sym(cur, used, (d / 10) % 10 + 1, 4, keyw, m, i, w); ++nsym; }
        put_sym(cur, used, d % 10 + 1, 4, keyw, m, i, w);
        ++nsym;
    }
    for (; nsym < 3 * (n - 1); ++nsym) put_sym(cur, used, 0, 4, keyw, m, i, w);
    const int8_t* wd = words + i * n;
    for (int p = 0; p < n; ++p) put_sym(cur, used, lrank[(uint8_t)wd[p]], lbits, keyw, m, i, w);
    if (used > 0) keyw[(int64_t)w * m + i] = cur << (64 - used);  // last word, left aligned
}

// OR and AND of every key word over the records (bits that differ between some
// two records are OR ^ AND): per-thread accumulation over a grid-stride loop,
// wave and block reductions, one atomic per block and word.
__global__ void __launch_bounds__(256) k_key_masks(const uint64_t* __restrict__ keyw, int W, int64_t m,
                                                   unsigned long long* __restrict__ or_mask,
                                                   unsigned long long* __restrict__ and_mask) {
    __shared__ unsigned long long so[4], sa[4];
    for (int w = 0; w < W; ++w) {
        unsigned long long o = 0, a = ~0ull;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
            const unsigned long long x = keyw[(int64_t)w * m + i];
            o |= x;
            a &= x;
        }
        for (int off = 32; off > 0; off >>= 1) {
            o |= __shfl_xor(o, off, 64);
            a &= __shfl_xor(a, off, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            so[threadIdx.x >> 6] = o;
            sa[threadIdx.x >> 6] = a;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int q = 1; q < (int)(blockDim.x >> 6); ++q) {
                so[0] |= so[q];
                sa[0] &= sa[q];
            }
            atomicOr(&or_mask[w], so[0]);
            atomicAnd(&and_mask[w], sa[0]);
        }
        __syncthreads();
    }
}

__global__ void k_sort_iota(uint32_t* __restrict__ idx, int64_t m) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
        idx[i] = (uint32_t)i;
}

__global__ void k_sort_gather(const uint64_t* __restrict__ col, const uint32_t* __restrict__ idx, int64_t m,
                              uint64_t* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = col[idx[i]];
}

// flag[i] = 1 when sorted record i differs from record i - 1 (or i == 0);
// pos[i] starts as a copy of flag[i] and is scanned afterwards.
__global__ void k_sort_uniq(const uint64_t* __restrict__ keyw, int W, int64_t m, const uint32_t* __restrict__ idx,
                            uint32_t* __restrict__ flag, uint32_t* __restrict__ pos) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t f = 1;
        if (i > 0) {
            const uint32_t a = idx[i], b = idx[i - 1];
            f = 0;
            for (int w = 0; w < W; ++w)
                if (keyw[(int64_t)w * m + a] != keyw[(int64_t)w * m + b]) { f = 1; break; }
        }
        flag[i] = f;
        pos[i] = f;
    }
}

__global__ void k_sort_compact(const int8_t* __restrict__ words, const uint8_t* __restrict__ keys, int n, int K,
                               int64_t m, const uint32_t* __restrict__ idx, const uint32_t* __restrict__ flag,
                               const uint32_t* __restrict__ pos, int8_t* __restrict__ out_w,
                               uint8_t* __restrict__ out_k) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        if (!flag[i]) continue;
        const int64_t src = idx[i], dst = pos[i];
        for (int p = 0; p < n; ++p) out_w[dst * n + p] = words[src * n + p];
        for (int p = 0; p < K; ++p) out_k[dst * K + p] = keys[src * K + p];
    }
}

// Bit set of the letters present (bit u of word u / 32 for the byte u of the
// letter): letters -32..31 collect in a per-lane mask, others go to LDS at once.
__global__ void k_letters_present(const int8_t* __restrict__ words, int64_t bytes, unsigned int* __restrict__ present) {
    __shared__ unsigned int s[8];
    if (threadIdx.x < 8) s[threadIdx.x] = 0;
    __syncthreads();
    unsigned long long mask = 0;  // bit v + 32 for letter v in [-32, 32)
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < bytes; i += (int64_t)gridDim.x * blockDim.x) {
        const int v = words[i];
        if (v >= -32 && v < 32) {
            mask |= 1ull << (v + 32);
        } else {
            const unsigned u = (uint8_t)v;
            atomicOr(&s[u >> 5], 1u << (u & 31));
        }
    }
    for (int off = 32; off > 0; off >>= 1) mask |= __shfl_xor(mask, off, 64);
    if ((threadIdx.x & 63) == 0) {
        atomicOr(&s[0], (unsigned)(mask >> 32));  // letters 0..31: bytes 0..31
        atomicOr(&s[7], (unsigned)mask);          // letters -32..-1: bytes 224..255
    }
    __syncthreads();
    if (threadIdx.x < 8 && s[threadIdx.x]) atomicOr(&present[threadIdx.x], s[threadIdx.x]);
}

namespace {
std::string token(int v) { return std::to_string(v) + " "; }
}  // namespace

// Sorts and uniqs the m records (words n bytes, keys K bytes each) held in
// d_words / d_keys on the context's stream; writes the kept records in line
// order to d_out_words / d_out_keys and returns their number.
int64_t sort_uniq_cands(gp_ctx* ctx, const int8_t* d_words, const uint8_t* d_keys, int64_t m, int n, int K,
                        int8_t* d_out_words, uint8_t* d_out_keys) {
    if (m <= 0) return 0;
    if (m > 0xFFFFFFFFll) throw ArgError("gp_sort_candidates: more than 2^32 records");
    if (n < 2 || n > kMaxN || K != n / 2) throw ArgError("gp_sort_candidates: bad record shape");
    hipStream_t st = ctx->stream;
    const int blocks = (int)std::min<int64_t>((m + 255) / 256, (int64_t)ctx->num_cus * 8);
    // letters present, ranked by token string
    unsigned int* present = ctx->dev<unsigned int>("s_present", 8);
    GP_HIP(hipMemsetAsync(present, 0, 32, st));
    hipLaunchKernelGGL(k_letters_present, dim3(blocks), dim3(256), 0, st, d_words, m * n, present);
    GP_HIP(hipGetLastError());
    unsigned int hp[8];
    GP_HIP(hipMemcpyAsync(hp, present, 32, hipMemcpyDeviceToHost, st));
    GP_HIP(hipStreamSynchronize(st));
    std::vector<int> letters;
    for (int v = -128; v < 128; ++v) {
        const unsigned u = (uint8_t)(int8_t)v;
        if (hp[u >> 5] >> (u & 31) & 1u) letters.push_back(v);
    }
    std::sort(letters.begin(), letters.end(), [](int a, int b) { return token(a) < token(b); });
    std::array<uint8_t, 256> lr{};
    for (size_t r = 0; r < letters.size(); ++r) lr[(uint8_t)(int8_t)letters[r]] = (uint8_t)r;
    int lbits = 1;
    while ((1u << lbits) < letters.size()) ++lbits;
    const int bits = 4 * 3 * (n - 1) + lbits * n;
    const int W = (bits + 63) / 64;
    uint8_t* d_lr = ctx->dev<uint8_t>("s_lrank", 256);
    uint64_t* keyw = ctx->dev<uint64_t>("s_keyw", (size_t)W * m);
    unsigned long long* masks = ctx->dev<unsigned long long>("s_masks", 2 * kSortMaxW);
    std::vector<unsigned long long> hm(2 * kSortMaxW);
    for (int w = 0; w < kSortMaxW; ++w) { hm[w] = 0; hm[kSortMaxW + w] = ~0ull; }
    GP_HIP(hipMemcpyAsync(d_lr, lr.data(), 256, hipMemcpyHostToDevice, st));
    GP_HIP(hipMemcpyAsync(masks, hm.data(), hm.size() * 8, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_cand_keys, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, d_words, d_keys, m, n, K,
                       lbits, d_lr, keyw);
    hipLaunchKernelGGL(k_key_masks, dim3(blocks), dim3(256), 0, st, keyw, W, m, masks, masks + kSortMaxW);
    GP_HIP(hipGetLastError());
    GP_HIP(hipMemcpyAsync(hm.data(), masks, hm.size() * 8, hipMemcpyDeviceToHost, st));
    GP_HIP(hipStreamSynchronize(st));
    // LSD over the key words, least significant first; only varying bytes
    uint32_t* idx = ctx->dev<uint32_t>("s_idx", (size_t)m);
    uint32_t* tidx = ctx->dev<uint32_t>("s_tidx", (size_t)m);
    uint64_t* kb = ctx->dev<uint64_t>("s_kb", (size_t)m);
    uint64_t* tkb = ctx->dev<uint64_t>("s_tkb", (size_t)m);
    uint32_t* hist = ctx->dev<uint32_t>("s_hist", radix_hist_bytes(m) / 4);
    hipLaunchKernelGGL(k_sort_iota, dim3(blocks), dim3(256), 0, st, idx, m);
    GP_HIP(hipGetLastError());
    for (int w = W - 1; w >= 0; --w) {
        const unsigned long long vary = hm[w] ^ hm[kSortMaxW + w];
        int shifts[8], ns = 0;
        for (int b = 0; b < 8; ++b)
            if ((vary >> (8 * b)) & 0xFFull) shifts[ns++] = 8 * b;
        if (ns == 0) continue;
        hipLaunchKernelGGL(k_sort_gather, dim3(blocks), dim3(256), 0, st, keyw + (int64_t)w * m, idx, m, kb);
        GP_HIP(hipGetLastError());
        radix_sort_pairs_shifts(st, kb, idx, tkb, tidx, hist, m, shifts, ns);
    }
    uint32_t* flag = ctx->dev<uint32_t>("s_flag", (size_t)m);
    uint32_t* pos = ctx->dev<uint32_t>("s_pos", (size_t)m);
    hipLaunchKernelGGL(k_sort_uniq, dim3(blocks), dim3(256), 0, st, keyw, W, m, idx, flag, pos);
    GP_HIP(hipGetLastError());
    scan_exclusive_u32(st, pos, m, ctx->dev<uint32_t>("s_scan", (size_t)scan_sums_count(m)));
    hipLaunchKernelGGL(k_sort_compact, dim3(blocks), dim3(256), 0, st, d_words, d_keys, n, K, m, idx, flag, pos,
                       d_out_words, d_out_keys);
    GP_HIP(hipGetLastError());
    uint32_t last[2];
    GP_HIP(hipMemcpyAsync(&last[0], pos + (m - 1), 4, hipMemcpyDeviceToHost, st));
    GP_HIP(hipMemcpyAsync(&last[1], flag + (m - 1), 4, hipMemcpyDeviceToHost, st));
    GP_HIP(hipStreamSynchronize(st));
    return (int64_t)last[0] + last[1];
}

}  // namespace gp
