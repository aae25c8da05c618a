// LSD radix sort of (uint64 key, uint32 value) pairs, 8 bits per pass.
//   k_rs_hist     per-block digit histogram in LDS
//   k_scan_*      digit-major exclusive scan of all block histograms
//   k_rs_scatter  stable scatter: each block walks its chunk in 256-element
//                 tiles; ranks inside a wave come from 8 ballots (one per digit
//                 bit), wave offsets from LDS counts, block offsets from the scan.
// HBM traffic per pass: read keys+values twice (histogram, scatter) and write
// them once: 3 * 12 bytes per element.
#include <algorithm>

#include "gp_internal.hpp"
#include "gp_sort.hpp"

namespace gp {

constexpr int kRsThreads = 256;
constexpr int kRsItems = 16;
constexpr int kRsChunk = kRsThreads * kRsItems;

__global__ void __launch_bounds__(kRsThreads) k_rs_hist(const uint64_t* __restrict__ keys, int64_t n, int shift,
                                                        uint32_t* __restrict__ hist, int nblocks) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kRsChunk;
    for (int i = 0; i < kRsItems; ++i) {
        const int64_t idx = base + (int64_t)i * kRsThreads + threadIdx.x;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of m uint32 values in place, in three launches: the sum of
// every chunk of kRsChunk values (k_scan_sums), the exclusive scan of those
// sums by one block (k_scan_offsets), and the scan of each chunk from its
// offset (k_scan_apply).  A chunk is read as kRsItems rows of kRsThreads
// consecutive values, so every load and store is coalesced.

// Inclusive scan of x over the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    return x;
}

__global__ void __launch_bounds__(kRsThreads) k_scan_sums(const uint32_t* __restrict__ v, int64_t m,
                                                          uint32_t* __restrict__ sums) {
    __shared__ uint32_t ws[kRsThreads / 64];
    const int64_t base = (int64_t)blockIdx.x * kRsChunk;
    uint32_t s = 0;
    for (int k = 0; k < kRsItems; ++k) {
        const int64_t i = base + (int64_t)k * kRsThreads + threadIdx.x;
        if (i < m) s += v[i];
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kRsThreads / 64; ++w) t += ws[w];
        sums[blockIdx.x] = t;
    }
}

// Exclusive scan of the nb chunk sums by one block, 1024 at a time.
__global__ void __launch_bounds__(1024) k_scan_offsets(uint32_t* __restrict__ sums, int64_t nb) {
    __shared__ uint32_t ws[16];
    __shared__ uint32_t carry_s;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) carry_s = 0;
    __syncthreads();
    for (int64_t b0 = 0; b0 < nb; b0 += 1024) {
        const int64_t i = b0 + threadIdx.x;
        const uint32_t x = i < nb ? sums[i] : 0u;
        const uint32_t s = wave_incl_scan(x, lane);
        if (lane == 63) ws[wave] = s;
        __syncthreads();
        uint32_t before = carry_s, tot = 0;
        for (int w = 0; w < 16; ++w) {
            if (w < wave) before += ws[w];
            tot += ws[w];
        }
        if (i < nb) sums[i] = before + s - x;
        __syncthreads();
        if (threadIdx.x == 0) carry_s += tot;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kRsThreads) k_scan_apply(uint32_t* __restrict__ v, int64_t m,
                                                           const uint32_t* __restrict__ offs) {
    __shared__ uint32_t ws[kRsThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * kRsChunk;
    uint32_t carry = offs[blockIdx.x];
    for (int k = 0; k < kRsItems; ++k) {
        const int64_t i = base + (int64_t)k * kRsThreads + threadIdx.x;
        const uint32_t x = i < m ? v[i] : 0u;
        const uint32_t s = wave_incl_scan(x, lane);
        if (lane == 63) ws[wave] = s;
        __syncthreads();
        uint32_t before = carry, tot = 0;
        for (int w = 0; w < kRsThreads / 64; ++w) {
            if (w < wave) before += ws[w];
            tot += ws[w];
        }
        if (i < m) v[i] = before + s - x;
        carry += tot;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kRsThreads) k_rs_scatter(const uint64_t* __restrict__ keys_in,
                                                           const uint32_t* __restrict__ vals_in, int64_t n, int shift,
                                                           const uint32_t* __restrict__ offs, int nblocks,
                                                           uint64_t* __restrict__ keys_out,
                                                           uint32_t* __restrict__ vals_out) {
    __shared__ uint32_t run[256];              // running output position per digit
    __shared__ uint32_t wcnt[kRsThreads / 64][256];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    run[threadIdx.x] = offs[(int64_t)threadIdx.x * nblocks + blockIdx.x];
    const int64_t base = (int64_t)blockIdx.x * kRsChunk;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int it = 0; it < kRsItems; ++it) {
        for (int w = 0; w < kRsThreads / 64; ++w) wcnt[w][threadIdx.x] = 0;
        __syncthreads();
        const int64_t idx = base + (int64_t)it * kRsThreads + threadIdx.x;
        const bool valid = idx < n;
        uint64_t k = 0;
        uint32_t v = 0;
        uint32_t dg = 0;
        if (valid) { k = keys_in[idx]; v = vals_in[idx]; dg = (uint32_t)((k >> shift) & 255u); }
        unsigned long long peers = __ballot(valid);
        for (int bit = 0; bit < 8; ++bit) {
            const unsigned long long b = __ballot((dg >> bit) & 1u);
            peers &= ((dg >> bit) & 1u) ? b : ~b;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & below);
        if (valid && rank == 0) wcnt[wave][dg] = (uint32_t)__popcll(peers);
        __syncthreads();
        uint32_t off = 0;
        if (valid) {
            for (int w = 0; w < wave; ++w) off += wcnt[w][dg];
            vals_out[run[dg] + off + rank] = v;
            keys_out[run[dg] + off + rank] = k;
        }
        __syncthreads();
        uint32_t tot = 0;
        for (int w = 0; w < kRsThreads / 64; ++w) tot += wcnt[w][threadIdx.x];
        run[threadIdx.x] += tot;
        __syncthreads();
    }
}

// keys/vals hold the input; the result ends in keys/vals (ping-pong through
// tmp).  One 8-bit pass per entry of `shifts`, in the order given (least
// significant digit first).
void radix_sort_pairs_shifts(hipStream_t st, uint64_t* keys, uint32_t* vals, uint64_t* tmp_keys, uint32_t* tmp_vals,
                             uint32_t* hist, int64_t n, const int* shifts, int nshift) {
    if (n <= 1 || nshift <= 0) return;
    if (n > 0xFFFFFFFFll) throw ArgError("radix sort: more than 2^32 elements");
    const int nblocks = (int)((n + kRsChunk - 1) / kRsChunk);
    uint64_t* ki = keys;
    uint32_t* vi = vals;
    uint64_t* ko = tmp_keys;
    uint32_t* vo = tmp_vals;
    for (int p = 0; p < nshift; ++p) {
        const int shift = shifts[p];
        hipLaunchKernelGGL(k_rs_hist, dim3(nblocks), dim3(kRsThreads), 0, st, ki, n, shift, hist, nblocks);
        scan_exclusive_u32(st, hist, (int64_t)nblocks * 256, hist + (size_t)(nblocks + 1) * 256);
        hipLaunchKernelGGL(k_rs_scatter, dim3(nblocks), dim3(kRsThreads), 0, st, ki, vi, n, shift, hist, nblocks, ko,
                           vo);
        GP_HIP(hipGetLastError());
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    if (nshift & 1) {
        GP_HIP(hipMemcpyAsync(keys, ki, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
        GP_HIP(hipMemcpyAsync(vals, vi, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
    }
}

void radix_sort_pairs(hipStream_t st, uint64_t* keys, uint32_t* vals, uint64_t* tmp_keys, uint32_t* tmp_vals,
                      uint32_t* hist, int64_t n, int begin_bit, int end_bit) {
    int shifts[8];
    int ns = 0;
    for (int shift = begin_bit; shift < end_bit && ns < 8; shift += 8) shifts[ns++] = shift;
    radix_sort_pairs_shifts(st, keys, vals, tmp_keys, tmp_vals, hist, n, shifts, ns);
}

// Exclusive scan of m uint32 values in place; `sums` holds scan_sums_count(m)
// values of scratch.
void scan_exclusive_u32(hipStream_t st, uint32_t* v, int64_t m, uint32_t* sums) {
    if (m <= 0) return;
    const int64_t nb = scan_sums_count(m);
    if (nb > 0x7FFFFFFF) throw ArgError("scan: too many values");
    hipLaunchKernelGGL(k_scan_sums, dim3((unsigned)nb), dim3(kRsThreads), 0, st, v, m, sums);
    hipLaunchKernelGGL(k_scan_offsets, dim3(1), dim3(1024), 0, st, sums, nb);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(kRsThreads), 0, st, v, m, sums);
    GP_HIP(hipGetLastError());
}

int64_t scan_sums_count(int64_t m) { return (m + kRsChunk - 1) / kRsChunk; }

// The histograms of n elements ((blocks + 1) x 256 values) followed by the
// scratch of their scan.
size_t radix_hist_bytes(int64_t n) {
    const int64_t h = ((n + kRsChunk - 1) / kRsChunk + 1) * 256;
    return (size_t)(h + scan_sums_count(h)) * 4;
}

}  // namespace gp
