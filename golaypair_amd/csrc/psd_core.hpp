// Device-side PSD evaluators for gfx950 (one wave = 64 lanes).
//
// Split-BF16 MFMA tile (transposed layout, v_mfma_f32_32x32x16_bf16):
//   D[ch][c] = sum_e W[ch][e] B[e][c]   for 32 candidates c (N dim) and
//   channels ch = (Re|Im, k) (M dim), reduction e over word positions, each
//   used twice (BF16 hi part of W, then lo part).  Entries x are integers with
//   |x| <= 127, exact in BF16, and every product is exact in FP32.
// Register layouts (gfx950 MFMA 32x32x16, wave64):
//   A (32x16): lane l holds row l%32, k = 8(l/32) + 0..7
//   B (16x32): lane l holds col l%32, k = 8(l/32) + 0..7
//   D (32x32): lane l holds col l%32, reg r holds row 8(r/4) + 4(l/32) + r%4
// With rows 0..15 = Re X_k and 16..31 = Im X_k (k = 16t + row%16 + 1), reg r
// and reg r+8 (r < 8) hold Re and Im of the same k in the same lane, so
// PSD_k = D[r]^2 + D[r+8]^2 needs no cross-lane traffic.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gp_common.hpp"

namespace gp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// Exact int8 -> BF16 bit pattern for |x| <= 127.
__device__ __forceinline__ uint32_t bf16_bits(int x) { return __float_as_uint((float)x) >> 16; }

// 8 int8 letters (packed little-endian in two uint32) -> 8 BF16 (uint4).
__device__ __forceinline__ uint4 pack8_bf16(uint32_t w0, uint32_t w1) {
    uint4 r;
    r.x = bf16_bits((int8_t)(w0 & 0xff)) | (bf16_bits((int8_t)((w0 >> 8) & 0xff)) << 16);
    r.y = bf16_bits((int8_t)((w0 >> 16) & 0xff)) | (bf16_bits((int8_t)(w0 >> 24)) << 16);
    r.z = bf16_bits((int8_t)(w1 & 0xff)) | (bf16_bits((int8_t)((w1 >> 8) & 0xff)) << 16);
    r.w = bf16_bits((int8_t)((w1 >> 16) & 0xff)) | (bf16_bits((int8_t)(w1 >> 24)) << 16);
    return r;
}

// Contraction for one tile: NB position blocks, MT M-tiles.  `afrag` points at
// the table's pre-swizzled A fragments; `bop[b]` is this lane's B operand for
// block b.  `acc` is initialised by the caller (zero or a prefix sum).
template <int NB, int MT>
__device__ __forceinline__ void mfma_tile(const uint4* __restrict__ afrag, const uint4 (&bop)[NB],
                                          f32x16 (&acc)[MT], int lane) {
#pragma unroll
    for (int t = 0; t < MT; ++t) {
#pragma unroll
        for (int s = 0; s < 2 * NB; ++s) {
            uint4 a = afrag[((t * 2 * NB) + s) * 64 + lane];
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(a), as_bf16x8(bop[s % NB]), acc[t], 0, 0, 0);
        }
    }
}

// Same contraction with the A fragments already held in registers.
template <int NB, int MT>
__device__ __forceinline__ void mfma_tile_reg(const uint4 (&a)[MT][2 * NB], const uint4 (&bop)[NB],
                                              f32x16 (&acc)[MT]) {
#pragma unroll
    for (int s = 0; s < 2 * NB; ++s) {
#pragma unroll
        for (int t = 0; t < MT; ++t)
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(a[t][s]), as_bf16x8(bop[s % NB]), acc[t], 0, 0, 0);
    }
}

// Max over this lane's k of PSD_k, combined with the partner lane (l ^ 32):
// the result is max_k PSD_k of candidate l % 32 (channels with k > K carry
// zero weights and contribute 0).
template <int MT>
__device__ __forceinline__ float tile_max_psd(const f32x16 (&acc)[MT]) {
    float m = 0.0f;
#pragma unroll
    for (int t = 0; t < MT; ++t) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            float re = acc[t][r], im = acc[t][r + 8];
            m = fmaxf(m, __builtin_fmaf(re, re, im * im));
        }
    }
    float o = __shfl_xor(m, 32);
    return fmaxf(m, o);
}

// Maxima of two one-M-tile tiles in one lane-half swap: lane l < 32 returns
// max_k PSD_k of candidate l of tile 0, lane l >= 32 that of candidate l - 32
// of tile 1 (v_permlane32_swap(a, b) exchanges lanes 32..63 of a with lanes
// 0..31 of b; |X|^2 >= 0, so maxima are taken on the float bit patterns).
__device__ __forceinline__ float tile_pair_max_psd1(const f32x16& acc0, const f32x16& acc1) {
    uint32_t m0 = 0u, m1 = 0u;
#pragma unroll
    for (int r = 0; r < 8; ++r) m0 = max(m0, __float_as_uint(__builtin_fmaf(acc0[r], acc0[r], acc0[r + 8] * acc0[r + 8])));
#pragma unroll
    for (int r = 0; r < 8; ++r) m1 = max(m1, __float_as_uint(__builtin_fmaf(acc1[r], acc1[r], acc1[r + 8] * acc1[r + 8])));
    const auto sw = __builtin_amdgcn_permlane32_swap(m0, m1, false, false);
    return __uint_as_float(max((uint32_t)sw[0], (uint32_t)sw[1]));
}

// The k (1-based) held by register r (< 8) of M-tile t in lane `lane`.
__device__ __forceinline__ int tile_k(int t, int r, int lane) {
    return 16 * t + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3) + 1;
}

// FP64 max_{1<=k<=n/2} PSD_k of a word held in memory (direct DFT with the
// table's FP64 twiddles c[m] = cos(2 pi m/n), s[m] = sin(2 pi m/n)).
__device__ inline double fp64_max_psd(const int8_t* __restrict__ w, int n, const double* __restrict__ c,
                                      const double* __restrict__ s) {
    double mx = 0.0;
    const int K = n >> 1;
    for (int k = 1; k <= K; ++k) {
        double re = 0.0, im = 0.0;
        int m = 0;
        for (int p = 0; p < n; ++p) {
            double x = (double)w[p];
            re = __builtin_fma(x, c[m], re);
            im = __builtin_fma(x, s[m], im);
            m += k;
            if (m >= n) m -= n;
        }
        double psd = re * re + im * im;
        mx = psd > mx ? psd : mx;
    }
    return mx;
}

// Adds the sum of `v` over the lanes of the calling wave to *p with one
// atomic (none when the sum is 0).  Every lane of the wave must call it.
__device__ inline void wave_atomic_add(unsigned long long* p, unsigned long long v) {
    for (int o = 32; o; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(p, v);
}

// Adds the sums of v0 .. v2 over the calling block's threads to *p0 .. *p2 (a
// null pointer skips its value) with one atomic each (none for a zero sum).
// Every thread of the block must call it.  Counters that every wave of a
// large grid bumps once at its end are best summed per block first: atomics on
// one address are applied one at a time, about 10 ns apart here, so a grid of
// thousands of waves spends tens of microseconds on them.
__device__ inline void block_atomic_add(unsigned long long* p0, unsigned long long v0,
                                        unsigned long long* p1 = nullptr, unsigned long long v1 = 0,
                                        unsigned long long* p2 = nullptr, unsigned long long v2 = 0) {
    __shared__ unsigned long long part[3][16];
    for (int o = 32; o; o >>= 1) {
        v0 += __shfl_xor(v0, o);
        v1 += __shfl_xor(v1, o);
        v2 += __shfl_xor(v2, o);
    }
    const int w = threadIdx.x >> 6, nw = (int)((blockDim.x + 63) >> 6);
    __syncthreads();  // an earlier call may still be reading `part`
    if ((threadIdx.x & 63) == 0) {
        part[0][w] = v0;
        part[1][w] = v1;
        part[2][w] = v2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        unsigned long long* p = threadIdx.x == 0 ? p0 : threadIdx.x == 1 ? p1 : p2;
        unsigned long long t = 0;
        for (int i = 0; i < nw; ++i) t += part[threadIdx.x][i];
        if (p && t) atomicAdd(p, t);
    }
}

// FP64 PSD_k of one frequency k, summed in the order of fp64_psd and
// fp64_max_psd (so all three agree bit for bit).
__device__ inline double fp64_psd_k(const int8_t* __restrict__ w, int n, const double* __restrict__ c,
                                    const double* __restrict__ s, int k) {
    double re = 0.0, im = 0.0;
    int m = 0;
    for (int p = 0; p < n; ++p) {
        double x = (double)w[p];
        re = __builtin_fma(x, c[m], re);
        im = __builtin_fma(x, s[m], im);
        m += k;
        if (m >= n) m -= n;
    }
    return re * re + im * im;
}

// FP64 PSD_k for k = 1..K into out[0..K-1].
__device__ inline void fp64_psd(const int8_t* __restrict__ w, int n, const double* __restrict__ c,
                                const double* __restrict__ s, double* out) {
    const int K = n >> 1;
    for (int k = 1; k <= K; ++k) {
        double re = 0.0, im = 0.0;
        int m = 0;
        for (int p = 0; p < n; ++p) {
            double x = (double)w[p];
            re = __builtin_fma(x, c[m], re);
            im = __builtin_fma(x, s[m], im);
            m += k;
            if (m >= n) m -= n;
        }
        out[k - 1] = re * re + im * im;
    }
}

}  // namespace gp
