// Shared host/device definitions of the gfx950 GolayPair kernels.
//
// Words are int8 letters (|x| <= 127).  The PSD filter of a length-n word
// evaluates X_k = sum_p x_p e^{-2 pi i kp/n} for k = 1..K, K = floor(n/2), and
// PSD_k = |X_k|^2 (src/utils/fourier.cpp:25-52 semantics: reject iff some
// PSD_k > 2*order - paf + 0.001).
#pragma once

#include <stdint.h>

namespace gp {

constexpr int kMaxN = 64;              // longest word the PSD kernels take
constexpr int kMaxK = kMaxN / 2;       // frequencies k = 1..K
constexpr int kTileCols = 32;          // candidates per MFMA tile (N dim)
constexpr int kMaxGroup = 64;          // max 16-entry reduction blocks per table

// Reduction layout of one PSD table (the MFMA contraction
//   D[ch][cand] = sum_e W[ch][e] * B[e][cand]):
//   channels ch: M-tile t (0..mt-1), row rho (0..31):
//       rho < 16  -> Re X_k,  rho >= 16 -> Im X_k,  k = 16 t + (rho % 16) + 1
//   entries e:   `nb` 16-entry blocks of positions (p = 16 b + j), used twice:
//                k-step s < nb with the BF16 "hi" part of W, s >= nb with "lo".
// A-operand fragments are stored pre-swizzled for v_mfma_f32_32x32x16_bf16:
//   frag[((t * 2 nb) + s) * 64 + lane] = 8 bf16 of row (lane % 32),
//   entries 16 (s % nb) + 8 (lane / 32) + 0..7.
struct PsdTableDesc {
    int32_t n;       // word length
    int32_t K;       // floor(n/2)
    int32_t nb;      // ceil(n/16) reduction blocks
    int32_t mt;      // ceil(K/16) M-tiles
    float rho_max;   // max |w - hi - lo| over the table's cos/sin entries
    int32_t frag_off;   // offset (in uint4 units) of this table's A fragments
    int32_t trig_off;   // offset (in floats/doubles) of this table's n-entry cos/sin
    int32_t pad;
};

// Decision codes written per candidate by the filter kernels.
enum : uint8_t { kReject = 0, kAccept = 1, kBand = 2 };

// Error bound eps (absolute, on PSD) for a filter evaluation near the
// threshold `thr`, given the per-entry DFT error bound E on Re and Im.
// |PSD' - PSD| <= 2E(|Re'|+|Im'|) + 2E^2 + 2.0001 u PSD'  with
// |Re'|+|Im'| <= sqrt(2 PSD') and PSD' <= thr + 1 in the band.
#if defined(__HIPCC__)
__host__ __device__
#endif
inline double psd_eps(double E, double thr) {
    const double P = thr + 1.0;
    double s = 2.0 * P;
    // sqrt without <cmath> on device: one Newton polish of a float estimate
    double r = (double)__builtin_sqrtf((float)s);
    r = 0.5 * (r + s / r);
    r *= 1.0000001;
    return 2.0 * E * r + 2.0 * E * E + 2.0001 * 5.9604644775390625e-08 * P;
}

// splitmix64 (Steele, Lea, Flood 2014), the microbenchmark's counter generator.
#if defined(__HIPCC__)
__host__ __device__
#endif
inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

// Letter index in [0, alpha) of entry j of microbenchmark candidate i:
// floor(splitmix64(seed ^ phi * (i * 64 + j)) * alpha / 2^64).
#if defined(__HIPCC__)
__host__ __device__
#endif
inline int mb_letter(uint64_t seed, uint64_t i, int j, int alpha) {
    uint64_t h = splitmix64(seed ^ (kGolden * (i * 64ull + (uint64_t)j)));
    return (int)(((unsigned __int128)h * (unsigned)alpha) >> 64);
}

}  // namespace gp
