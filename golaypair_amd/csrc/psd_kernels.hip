// Batched PSD filter kernels (gfx950):
//   k_psd_mfma    split-BF16 MFMA contraction, 32 candidates per tile
//   k_psd_fp32    plain FP32 direct DFT, one candidate per lane
//   k_psd_fp64    FP64 direct DFT for candidates inside the error band
// Semantics: src/utils/fourier.cpp:25-52 (reject iff PSD_k > 2*order - paf +
// 0.001 for some 1 <= k <= n/2).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>

#include "gp_internal.hpp"
#include "psd_core.hpp"

namespace gp {

// ---------------------------------------------------------------------------
// host: tables

namespace {
uint16_t bf16_rne(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}
double bf16_to_double(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return (double)f;
}
}  // namespace

int PsdTables::get(int n) {
    auto it = by_n.find(n);
    if (it != by_n.end()) return it->second;
    if (n < 2 || n > kMaxN) throw ArgError("word length " + std::to_string(n) + " outside [2, 64]");
    PsdTableDesc d{};
    d.n = n;
    d.K = n / 2;
    d.nb = (n + 15) / 16;
    d.mt = (d.K + 15) / 16;
    // FP64 twiddles, symmetric in m <-> n-m (psd[k] == psd[n-k] bit-exactly).
    std::vector<double> c(n), s(n);
    for (int m = 0; 2 * m <= n; ++m) {
        double a = 2.0 * M_PI * (double)m / (double)n;
        double cm = std::cos(a), sm = std::sin(a);
        if (m == 0) { cm = 1.0; sm = 0.0; }
        if (2 * m == n) { cm = -1.0; sm = 0.0; }
        c[m] = cm; s[m] = sm;
        if (m != 0 && 2 * m != n) { c[n - m] = cm; s[n - m] = -sm; }
    }
    d.trig_off = (int32_t)trig64.size();
    for (int m = 0; m < n; ++m) trig64.push_back(c[m]);
    for (int m = 0; m < n; ++m) trig64.push_back(s[m]);
    for (int m = 0; m < n; ++m) trig32.push_back((float)c[m]);
    for (int m = 0; m < n; ++m) trig32.push_back((float)s[m]);
    // A fragments.
    d.frag_off = (int32_t)frags.size();
    double rho = 0.0;
    for (int t = 0; t < d.mt; ++t)
        for (int st = 0; st < 2 * d.nb; ++st)
            for (int lane = 0; lane < 64; ++lane) {
                uint16_t h[8];
                for (int i = 0; i < 8; ++i) {
                    int row = lane % 32;
                    int e = 16 * (st % d.nb) + 8 * (lane / 32) + i;
                    int part = st / d.nb;
                    int k = 16 * t + row % 16 + 1;
                    bool im = row >= 16;
                    uint16_t v = 0;
                    if (e < n && k <= d.K) {
                        int m = (int)(((long)k * e) % n);
                        double w = im ? s[m] : c[m];
                        uint16_t hi = bf16_rne((float)w);
                        double r1 = w - bf16_to_double(hi);
                        uint16_t lo = bf16_rne((float)r1);
                        double r2 = std::fabs(r1 - bf16_to_double(lo));
                        rho = std::max(rho, r2);
                        v = part ? lo : hi;
                    }
                    h[i] = v;
                }
                uint4 q;
                q.x = h[0] | ((uint32_t)h[1] << 16);
                q.y = h[2] | ((uint32_t)h[3] << 16);
                q.z = h[4] | ((uint32_t)h[5] << 16);
                q.w = h[6] | ((uint32_t)h[7] << 16);
                frags.push_back(q);
            }
    d.rho_max = (float)(rho * 1.0000001);
    int id = (int)desc.size();
    desc.push_back(d);
    by_n[n] = id;
    dirty = true;
    return id;
}

void PsdTables::upload(hipStream_t st) {
    if (!dirty) return;
    d_desc.ensure(desc.size() * sizeof(PsdTableDesc));
    d_frags.ensure(frags.size() * sizeof(uint4));
    d_trig32.ensure(trig32.size() * sizeof(float));
    d_trig64.ensure(trig64.size() * sizeof(double));
    GP_HIP(hipMemcpyAsync(d_desc.p, desc.data(), desc.size() * sizeof(PsdTableDesc), hipMemcpyHostToDevice, st));
    GP_HIP(hipMemcpyAsync(d_frags.p, frags.data(), frags.size() * sizeof(uint4), hipMemcpyHostToDevice, st));
    GP_HIP(hipMemcpyAsync(d_trig32.p, trig32.data(), trig32.size() * sizeof(float), hipMemcpyHostToDevice, st));
    GP_HIP(hipMemcpyAsync(d_trig64.p, trig64.data(), trig64.size() * sizeof(double), hipMemcpyHostToDevice, st));
    GP_HIP(hipStreamSynchronize(st));
    dirty = false;
}

double split_bf16_E(const PsdTableDesc& d, double s1) {
    // representation error of W (rho) + FP32 accumulation over the 2*16*nb
    // exact products (+1 for the initial accumulator), each add taken as a
    // truncation (2^-23 relative of a partial sum bounded by sum |products|).
    const double adds = 32.0 * d.nb + 1.0;
    return s1 * ((double)d.rho_max + 1.004 * adds * 0x1p-23);
}

double fp32_E(int n, double s1) {
    // FP32 twiddles (2^-24 relative) and n FMAs per sum (2^-24 each, on partial
    // sums bounded by s1).
    return s1 * ((double)(n + 2) * 0x1p-24) * 1.01;
}

FilterThr make_thr(double thr, double E) {
    const double eps = psd_eps(E, thr);
    FilterThr t;
    t.thr = thr;
    t.hi = std::nextafter((float)(thr + eps), INFINITY);
    t.lo = std::nextafter((float)(thr - eps), -INFINITY);
    return t;
}

// ---------------------------------------------------------------------------
// device

// Counts of the FP64 recheck.  The filter kernels leave a band row's decision
// at kBand and count nothing: the recheck finds the band rows by scanning the
// decisions (a counter shared by every wave of the filter serialises its
// atomics once a few tiles in a hundred hold a band row).
struct BandCtl {
    unsigned long long count;  // band rows rechecked in FP64
    unsigned long long ties;   // FP64 values within 1e-9 of the threshold
};

// The work of one launch (one operand class): up to 8 row segments, passed by
// value.  Its waves are shared out between the segments in proportion to their
// 32-row tiles: wave w of segment s (wave0[s] <= w < wave0[s + 1]) filters the
// contiguous tiles (w - wave0[s]) tpw[s] .. + tpw[s] - 1 of s, so that a wave
// reads its table, threshold and A fragments once and streams its rows.
struct SegTiles {
    int32_t nseg;
    int32_t table[8];
    int32_t stride[8];    // bytes per row
    int32_t tpw[8];       // tiles per wave
    int32_t wave0[9];     // first wave of each segment (and the wave count)
    int64_t first_row[8];
    int64_t rows[8];
    int64_t byte0[8];     // byte offset of the segment's first row
};

// Shares at most `waves` waves out between T's segments (at least one each,
// and at least 4 tiles per wave); fills tpw and wave0, returns the waves used.
inline int64_t seg_waves(SegTiles& T, int64_t waves) {
    int64_t total = 0;
    for (int i = 0; i < T.nseg; ++i) total += (T.rows[i] + 31) / 32;
    waves = std::max<int64_t>(1, std::min<int64_t>(waves, (total + 3) / 4));
    int64_t used = 0;
    for (int i = 0; i < T.nseg; ++i) {
        const int64_t tiles = (T.rows[i] + 31) / 32;
        T.wave0[i] = (int32_t)used;
        T.tpw[i] = 1;
        if (tiles == 0) continue;
        const int64_t ws = std::max<int64_t>(1, waves * tiles / std::max<int64_t>(total, 1));
        const int64_t tpw = (tiles + ws - 1) / ws;
        T.tpw[i] = (int32_t)tpw;
        used += (tiles + tpw - 1) / tpw;
    }
    T.wave0[T.nseg] = (int32_t)used;
    return used;
}

// Rows of segments that share one operand class (NB position blocks, MT
// M-tiles): one instantiation per class, so each sizes its registers alone.
// A wave keeps D tiles of letters in flight (a ring refilled D tiles ahead as
// each tile is taken): the kernel streams its rows.  Lanes read bytes
// 16 bl + 8 (lane / 32) .. + 7 of their row: a row holds 16 NB bytes (letters
// past n meet zero weights in the table); lanes past the segment's last row
// read that row again and store no decision.
template <int NB, int MT>
__global__ void __launch_bounds__(256) k_psd_mfma(const int8_t* __restrict__ words, const SegTiles T,
                                                  const PsdTableDesc* __restrict__ tabs,
                                                  const uint4* __restrict__ frags,
                                                  const FilterThr* __restrict__ thr,
                                                  uint8_t* __restrict__ decision, float* __restrict__ maxpsd) {
    constexpr int D = NB == 1 ? 8 : NB == 2 ? 4 : 2;
    const int lane = threadIdx.x & 63, col = lane & 31, half = lane >> 5;
    const int wave = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
    if (wave >= T.wave0[T.nseg]) return;
    int sg = 0;
    while (sg + 1 < T.nseg && wave >= T.wave0[sg + 1]) ++sg;
    const int64_t rows = T.rows[sg];
    const int64_t q0 = (int64_t)(wave - T.wave0[sg]) * T.tpw[sg];
    const int64_t q1 = min<int64_t>(q0 + T.tpw[sg], (rows + 31) / 32);
    if (q0 >= q1) return;
    const int8_t* __restrict__ base = words + T.byte0[sg] + 8 * half;
    const int64_t stride = T.stride[sg], row_first = T.first_row[sg];
    const PsdTableDesc d = tabs[T.table[sg]];
    const FilterThr th = thr[T.table[sg]];
    uint4 a[MT][2 * NB];
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int s2 = 0; s2 < 2 * NB; ++s2) a[t][s2] = frags[d.frag_off + ((t * 2 * NB) + s2) * 64 + lane];
    auto load = [&](int64_t q, uint2 (&raw)[NB]) {
        const int8_t* w = base + min(q * 32 + col, rows - 1) * stride;
#pragma unroll
        for (int bl = 0; bl < NB; ++bl) raw[bl] = *reinterpret_cast<const uint2*>(w + 16 * bl);
    };
    uint2 ring[D][NB];
#pragma unroll
    for (int g = 0; g < D; ++g) load(q0 + g, ring[g]);
    if constexpr (MT == 1) {
        // Tiles in pairs (q, q + 1): two independent MFMA chains, one lane-half
        // swap for both maxima (tile q in lanes 0..31, tile q + 1 in lanes
        // 32..63), so lane l decides row 32 q + l and one store covers 64 rows.
        static_assert(D % 2 == 0, "tile pairs");
        for (int64_t qq = q0; qq < q1; qq += D) {
#pragma unroll
            for (int g = 0; g < D; g += 2) {
                const int64_t q = qq + g;
                if (q >= q1) break;
                uint4 b0[NB], b1[NB];
#pragma unroll
                for (int bl = 0; bl < NB; ++bl) {
                    b0[bl] = pack8_bf16(ring[g][bl].x, ring[g][bl].y);
                    b1[bl] = pack8_bf16(ring[g + 1][bl].x, ring[g + 1][bl].y);
                }
                load(q + D, ring[g]);
                load(q + 1 + D, ring[g + 1]);
                f32x16 acc0[1], acc1[1];
                acc0[0] = f32x16{};
                acc1[0] = f32x16{};
#pragma unroll
                for (int s2 = 0; s2 < 2 * NB; ++s2) {
                    acc0[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(a[0][s2]), as_bf16x8(b0[s2 % NB]),
                                                                     acc0[0], 0, 0, 0);
                    acc1[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(a[0][s2]), as_bf16x8(b1[s2 % NB]),
                                                                     acc1[0], 0, 0, 0);
                }
                const float m = tile_pair_max_psd1(acc0[0], acc1[0]);
                const int64_t r = q * 32 + lane;
                const bool mine = r < rows && (lane < 32 || q + 1 < q1);
                const uint8_t dec = m > th.hi ? kReject : (m < th.lo ? kAccept : kBand);
                if (mine) {
                    decision[row_first + r] = dec;
                    if (maxpsd) maxpsd[row_first + r] = m;
                }
            }
        }
        return;
    }
    for (int64_t qq = q0; qq < q1; qq += D) {
#pragma unroll
        for (int g = 0; g < D; ++g) {
            const int64_t q = qq + g;
            if (q >= q1) break;
            uint4 bop[NB];
#pragma unroll
            for (int bl = 0; bl < NB; ++bl) bop[bl] = pack8_bf16(ring[g][bl].x, ring[g][bl].y);
            load(q + D, ring[g]);
            f32x16 acc[MT];
#pragma unroll
            for (int t = 0; t < MT; ++t) acc[t] = f32x16{};
            mfma_tile_reg<NB, MT>(a, bop, acc);
            const float m = tile_max_psd<MT>(acc);
            const int64_t r = q * 32 + col;
            const bool mine = half == 0 && r < rows;
            const uint8_t dec = m > th.hi ? kReject : (m < th.lo ? kAccept : kBand);
            const int64_t row = row_first + r;
            if (mine) {
                decision[row] = dec;
                if (maxpsd) maxpsd[row] = m;
            }
        }
    }
}

// Launches the rows of T, one kernel per operand class among its segments
// (at most every resident wave slot of the class).
void launch_psd_mfma(const SegTiles& all, const std::vector<PsdTableDesc>& desc, int num_cus, hipStream_t st,
                     const int8_t* words, const PsdTableDesc* tabs, const uint4* frags,
                     const FilterThr* thr, uint8_t* decision, float* maxpsd) {
    auto cls = [&](int sg) {
        const PsdTableDesc& d = desc[all.table[sg]];
        return d.nb == 1 ? 0 : d.nb == 2 && d.mt == 1 ? 1 : d.nb == 3 && d.mt == 1 ? 2 : d.nb == 3 ? 3 : 4;
    };
    for (int c = 0; c < 5; ++c) {
        SegTiles T{};
        for (int sg = 0; sg < all.nseg; ++sg)
            if (cls(sg) == c) {
                T.table[T.nseg] = all.table[sg];
                T.stride[T.nseg] = all.stride[sg];
                T.first_row[T.nseg] = all.first_row[sg];
                T.rows[T.nseg] = all.rows[sg];
                T.byte0[T.nseg] = all.byte0[sg];
                ++T.nseg;
            }
        if (T.nseg == 0) continue;
        // at most every resident wave slot of the class (blocks of 4 waves per CU
        // from the kernel's register and LDS use): a larger grid would run its
        // last waves after the first ones
        static const void* const kFn[5] = {
            reinterpret_cast<const void*>(&k_psd_mfma<1, 1>), reinterpret_cast<const void*>(&k_psd_mfma<2, 1>),
            reinterpret_cast<const void*>(&k_psd_mfma<3, 1>), reinterpret_cast<const void*>(&k_psd_mfma<3, 2>),
            reinterpret_cast<const void*>(&k_psd_mfma<4, 2>)};
        static std::atomic<int> blocks_per_cu[5] = {};
        if (blocks_per_cu[c].load() == 0) {
            int nb = 0;
            GP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kFn[c], 256, 0));
            blocks_per_cu[c].store(std::max(nb, 1));
        }
        const int64_t waves = seg_waves(T, (int64_t)num_cus * 4 * blocks_per_cu[c].load());
        if (waves == 0) continue;
        const dim3 grid((unsigned)std::max<int64_t>(1, (waves + 3) / 4));
#define GP_PSD_CASE(C, NB, MT)                                                                                   \
    case C:                                                                                                      \
        hipLaunchKernelGGL((k_psd_mfma<NB, MT>), grid, dim3(256), 0, st, words, T, tabs, frags, thr, decision,  \
                           maxpsd);                                                                              \
        break;
        switch (c) {
            GP_PSD_CASE(0, 1, 1)
            GP_PSD_CASE(1, 2, 1)
            GP_PSD_CASE(2, 3, 1)
            GP_PSD_CASE(3, 3, 2)
            GP_PSD_CASE(4, 4, 2)
        }
#undef GP_PSD_CASE
        GP_HIP(hipGetLastError());
    }
}

// Plain FP32 direct DFT: one candidate per lane, rows in per-table segments.
template <int KMAX>
__device__ __forceinline__ float fp32_max_psd(const int8_t* __restrict__ w, int n, int K,
                                              const float* __restrict__ c, const float* __restrict__ s) {
    float re[KMAX], im[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) { re[k] = 0.f; im[k] = 0.f; }
    for (int p = 0; p < n; ++p) {
        const float x = (float)w[p];
        int m = 0;
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            m += p;
            if (m >= n) m -= n;
            re[k] = __builtin_fmaf(x, c[m], re[k]);
            im[k] = __builtin_fmaf(x, s[m], im[k]);
        }
    }
    float mx = 0.f;
#pragma unroll
    for (int k = 0; k < KMAX; ++k)
        if (k < K) mx = fmaxf(mx, __builtin_fmaf(re[k], re[k], im[k] * im[k]));
    return mx;
}

__global__ void __launch_bounds__(256) k_psd_fp32(const int8_t* __restrict__ words,
                                                  const BatchSeg* __restrict__ segs, int nseg,
                                                  const PsdTableDesc* __restrict__ tabs,
                                                  const float* __restrict__ trig32,
                                                  const FilterThr* __restrict__ thr,
                                                  uint8_t* __restrict__ decision, float* __restrict__ maxpsd) {
    __shared__ float lds[8][2 * kMaxN];
    __shared__ int seg_table[8];
    if (threadIdx.x < 8 && (int)threadIdx.x < nseg) seg_table[threadIdx.x] = segs[threadIdx.x].table;
    __syncthreads();
    for (int sg = 0; sg < nseg && sg < 8; ++sg) {
        const PsdTableDesc d = tabs[seg_table[sg]];
        for (int i = threadIdx.x; i < 2 * d.n; i += blockDim.x) lds[sg][i] = trig32[d.trig_off + i];
    }
    __syncthreads();
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t gsz = (int64_t)gridDim.x * blockDim.x;
    for (int sg = 0; sg < nseg && sg < 8; ++sg) {
        const BatchSeg S = segs[sg];
        const PsdTableDesc d = tabs[S.table];
        const FilterThr th = thr[S.table];
        const float* c = lds[sg];
        const float* s = lds[sg] + d.n;
        for (int64_t r = gid; r < S.rows; r += gsz) {
            const int64_t row = S.first_row + r;
            const int8_t* w = words + S.byte0 + r * S.stride;
            float m = d.K <= 16 ? fp32_max_psd<16>(w, d.n, d.K, c, s) : fp32_max_psd<32>(w, d.n, d.K, c, s);
            decision[row] = m > th.hi ? kReject : (m < th.lo ? kAccept : kBand);
            if (maxpsd) maxpsd[row] = m;
        }
    }
}

// FP64 recheck of the rows left in the band (decision kBand), segment by
// segment.  A wave scans 512 decisions at a time (8 per lane) and settles each
// band row it finds with all its lanes: lane q < n/2 computes PSD_{q+1} in the
// summation order of fp64_psd_k / fp64_max_psd, from the segments' FP64
// twiddles and the row's letters staged in LDS; the wave takes the max.
__global__ void __launch_bounds__(256) k_psd_fp64(const int8_t* __restrict__ words,
                                                  const BatchSeg* __restrict__ segs, int nseg,
                                                  const PsdTableDesc* __restrict__ tabs,
                                                  const double* __restrict__ trig64,
                                                  const FilterThr* __restrict__ thr,
                                                  uint8_t* __restrict__ decision, BandCtl* ctl) {
    __shared__ double tw[8][2 * kMaxN];  // per segment: cos then sin of 2 pi m / n
    __shared__ int8_t wl[4][kMaxN];      // per wave: the band row's letters
    for (int sg = 0; sg < nseg && sg < 8; ++sg) {
        const PsdTableDesc d = tabs[segs[sg].table];
        for (int i = threadIdx.x; i < 2 * d.n; i += blockDim.x) tw[sg][i] = trig64[d.trig_off + i];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + wv;
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
    unsigned long long nb = 0, nt = 0;
    for (int sg = 0; sg < nseg && sg < 8; ++sg) {
        const BatchSeg S = segs[sg];
        const int n = tabs[S.table].n, K = n >> 1;
        const double* c = tw[sg];
        const double* s = tw[sg] + n;
        const double t = thr[S.table].thr;
        uint8_t* __restrict__ dec = decision + S.first_row;
        for (int64_t b = wave * 512; b < S.rows; b += nwaves * 512) {
            uint8_t v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int64_t r = b + 64 * j + lane;
                v[j] = r < S.rows ? dec[r] : (uint8_t)kAccept;
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                unsigned long long bb = __ballot(v[j] == kBand);
                while (bb) {
                    const int64_t r = b + 64 * j + (__ffsll((long long)bb) - 1);
                    bb &= bb - 1;
                    if (lane < n) wl[wv][lane] = words[S.byte0 + r * S.stride + lane];
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    double m = 0.0;
                    if (lane < K) {
                        const int k = lane + 1;
                        double re = 0.0, im = 0.0;
                        int q = 0;
                        for (int p = 0; p < n; ++p) {
                            const double x = (double)wl[wv][p];
                            re = __builtin_fma(x, c[q], re);
                            im = __builtin_fma(x, s[q], im);
                            q += k;
                            if (q >= n) q -= n;
                        }
                        m = re * re + im * im;
                    }
                    for (int o = 32; o; o >>= 1) m = fmax(m, __shfl_xor(m, o));
                    if (lane == 0) {
                        ++nb;
                        nt += fabs(m - t) < 1e-9;
                        dec[r] = m > t ? kReject : kAccept;
                    }
                    __builtin_amdgcn_wave_barrier();  // wl is rewritten for the next band row
                }
            }
        }
    }
    block_atomic_add(&ctl->count, nb, &ctl->ties, nt);
}

// ---------------------------------------------------------------------------
// host driver

double run_psd_filter(gp_ctx* ctx, const int8_t* d_words, int stride, int64_t nrows,
                      const std::vector<BatchSeg>& segs_in, const std::vector<FilterThr>& thr_by_table,
                      uint8_t* d_decision, gp_stats* stats, float* d_maxpsd) {
    if (segs_in.size() > 8) throw ArgError("at most 8 word lengths per PSD batch");
    std::vector<BatchSeg> segs(segs_in);
    for (BatchSeg& sg : segs) {
        if (sg.stride == 0) {
            sg.stride = stride;
            sg.byte0 = sg.first_row * stride;
        }
        if (sg.stride < 16 * ctx->tables.desc[sg.table].nb || sg.stride % 8 || sg.byte0 % 8)
            throw ArgError("PSD batch: rows must be 8-byte aligned, 16 bytes per block of 16 letters");
    }
    hipStream_t st = ctx->stream;
    ctx->tables.upload(st);
    SegTiles T{};
    T.nseg = (int32_t)segs.size();
    for (size_t i = 0; i < segs.size(); ++i) {
        T.table[i] = segs[i].table;
        T.stride[i] = segs[i].stride;
        T.first_row[i] = segs[i].first_row;
        T.rows[i] = segs[i].rows;
        T.byte0[i] = segs[i].byte0;
    }
    int64_t nrow_total = 0;
    for (const BatchSeg& sg : segs) nrow_total += sg.rows;
    BandCtl* ctl = static_cast<BandCtl*>(ctx->counters.ensure(sizeof(BandCtl)));
    FilterThr* d_thr = static_cast<FilterThr*>(ctx->thr.ensure(thr_by_table.size() * sizeof(FilterThr) + 64));
    BatchSeg* d_segs = static_cast<BatchSeg*>(ctx->aux5.ensure(segs.size() * sizeof(BatchSeg) + 64));
    GP_HIP(hipMemcpyAsync(d_thr, thr_by_table.data(), thr_by_table.size() * sizeof(FilterThr), hipMemcpyHostToDevice, st));
    GP_HIP(hipMemcpyAsync(d_segs, segs.data(), segs.size() * sizeof(BatchSeg), hipMemcpyHostToDevice, st));
    GP_HIP(hipMemsetAsync(ctl, 0, sizeof(BandCtl), st));
    GP_HIP(hipEventRecord(ctx->ev0, st));
    if (ctx->psd_mode == GP_PSD_SPLIT_BF16) {
        if (nrow_total > 0)
            launch_psd_mfma(T, ctx->tables.desc, ctx->num_cus, st, d_words, ctx->tables.d_desc.as<PsdTableDesc>(),
                            ctx->tables.d_frags.as<uint4>(), d_thr, d_decision, d_maxpsd);
    } else {
        int blocks = (int)std::min<int64_t>((nrow_total + 255) / 256, (int64_t)ctx->num_cus * 8);
        if (nrow_total > 0)
            hipLaunchKernelGGL(k_psd_fp32, dim3(std::max(blocks, 1)), dim3(256), 0, st, d_words, d_segs,
                               (int)segs.size(), ctx->tables.d_desc.as<PsdTableDesc>(),
                               ctx->tables.d_trig32.as<float>(), d_thr, d_decision, d_maxpsd);
    }
    GP_HIP(hipGetLastError());
    GP_HIP(hipEventRecord(ctx->ev1, st));
    const int fblocks = (int)std::min<int64_t>((nrow_total + 2047) / 2048, (int64_t)ctx->num_cus * 8);
    hipLaunchKernelGGL(k_psd_fp64, dim3(std::max(fblocks, 1)), dim3(256), 0, st, d_words, d_segs, (int)segs.size(),
                       ctx->tables.d_desc.as<PsdTableDesc>(), ctx->tables.d_trig64.as<double>(), d_thr, d_decision,
                       ctl);
    GP_HIP(hipGetLastError());
    GP_HIP(hipEventRecord(ctx->ev2, st));
    BandCtl h{};
    GP_HIP(hipMemcpyAsync(&h, ctl, sizeof(BandCtl), hipMemcpyDeviceToHost, st));
    GP_HIP(hipStreamSynchronize(st));
    float ms_main = 0, ms_band = 0;
    GP_HIP(hipEventElapsedTime(&ms_main, ctx->ev0, ctx->ev1));
    GP_HIP(hipEventElapsedTime(&ms_band, ctx->ev1, ctx->ev2));
    if (stats) {
        stats->band_rechecks += h.count;
        stats->tie_band += h.ties;
        for (const BatchSeg& sg : segs) stats->psd_decisions += (uint64_t)sg.rows;
        stats->psd_kernel_ms += ms_main;
        stats->band_kernel_ms += ms_band;
        stats->psd_launches += 1;
        stats->kernel_ms += ms_main + ms_band;
    }
    return ms_main;
}

}  // namespace gp

// ---------------------------------------------------------------------------
// The batched PSD filter without allocation (gp_psd_filter_batch_static).
//
// Everything the call needs beyond the caller's words and pass flags lives in
// `scratch`, one file-scope device array per GPU, held until the process ends:
//   [0, 256)                  NaHead: band / tie / pass counters, the row-sum
//                             bound S1, the thresholds and the table descriptor
//   [256, 256 + 16 KiB)       the split-BF16 A fragments of the word length
//   [.., + 1 KiB)             its FP64 twiddles.
// Host code refers to it only through the address hipGetSymbolAddress gives
// (kept per device by psd_static_setup); the kernels take that address.  Every
// call on one GPU shares the array: calls that overlap on one GPU race.

namespace gp {

struct NaHead {
    BandCtl ctl;                  // band rows rechecked by k_na_band, FP64 ties
    unsigned long long passed;    // rows accepted
    unsigned int s1;              // max over the rows of sum |x_p|
    unsigned int pad;
    FilterThr thr;
    PsdTableDesc desc;            // frag_off = trig_off = 0: the regions below
};
constexpr size_t kNaHeadBytes = 256;
constexpr size_t kNaFragCount = 2 * 8 * 64;  // M-tiles (<= 2) x 2 nb (<= 8) x 64 lanes, n <= 64
constexpr size_t kNaFragOff = kNaHeadBytes;
constexpr size_t kNaTrigOff = kNaFragOff + kNaFragCount * sizeof(uint4);
constexpr size_t kNaBytes = kNaTrigOff + 2 * kMaxN * sizeof(double);
static_assert(sizeof(NaHead) <= kNaHeadBytes, "scratch header");
constexpr int kNaMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNaBytes];

namespace {
std::mutex g_na_mu;
unsigned char* g_na_ptr[kNaMaxDevices] = {};  // device address of `scratch`, per device
int g_na_n[kNaMaxDevices] = {};               // word length whose table `scratch` holds
}  // namespace

void psd_static_setup(int device) {
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (device < 0 || device >= kNaMaxDevices) throw ArgError("device ordinal beyond the static scratch table");
    if (g_na_ptr[device]) return;
    void* p = nullptr;
    GP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_na_ptr[device] = static_cast<unsigned char*>(p);
}

__global__ void k_na_reset(NaHead* h) {
    if (threadIdx.x == 0) {
        h->ctl.count = 0;
        h->ctl.ties = 0;
        h->passed = 0;
        h->s1 = 0;
    }
}

// S1 = max over the rows of sum |x_p| (the row-sum bound of the error analysis).
__global__ void __launch_bounds__(256) k_na_s1(const int8_t* __restrict__ words, int stride, int n, int64_t count,
                                               NaHead* h) {
    unsigned int best = 0;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < count; r += (int64_t)gridDim.x * blockDim.x) {
        const int8_t* w = words + r * stride;
        unsigned int a = 0;
        for (int p = 0; p < n; ++p) a += (unsigned int)abs((int)w[p]);
        best = max(best, a);
    }
    for (int o = 32; o; o >>= 1) best = max(best, (unsigned int)__shfl_xor((int)best, o));
    if ((threadIdx.x & 63) == 0) atomicMax(&h->s1, best);
}

// Thresholds from S1 as the host path makes them (split_bf16_E, make_thr).
__global__ void k_na_thr(NaHead* h, double thr) {
    if (threadIdx.x != 0) return;
    const double adds = 32.0 * h->desc.nb + 1.0;
    const double E = (double)h->s1 * ((double)h->desc.rho_max + 1.004 * adds * 0x1p-23);
    const double eps = psd_eps(E, thr);
    h->thr.thr = thr;
    h->thr.hi = nextafterf((float)(thr + eps), INFINITY);
    h->thr.lo = nextafterf((float)(thr - eps), -INFINITY);
}

// Rows left in the band by k_psd_mfma are settled in FP64 in place; accepted
// rows are counted.
__global__ void __launch_bounds__(256) k_na_band(const int8_t* __restrict__ words, int stride, int64_t count,
                                                 const double* __restrict__ trig64, NaHead* h,
                                                 uint8_t* __restrict__ pass) {
    const int n = h->desc.n;
    const double t = h->thr.thr;
    unsigned long long nb = 0, nt = 0, npass = 0;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < count; base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = base + threadIdx.x;
        unsigned long long acc = 0;
        if (r < count) {
            uint8_t d = pass[r];
            if (d == kBand) {
                const double m = fp64_max_psd(words + r * stride, n, trig64, trig64 + n);
                ++nb;
                nt += fabs(m - t) < 1e-9;
                d = m > t ? kReject : kAccept;
                pass[r] = d;
            }
            acc = d == kAccept;
        }
        npass += acc;
    }
    block_atomic_add(&h->passed, npass, &h->ctl.count, nb, &h->ctl.ties, nt);
}

}  // namespace gp

extern "C" int gp_psd_filter_batch_static(gp_ctx* ctx, const int8_t* d_words, int stride, int n, int order, int paf,
                                          int64_t count, uint8_t* d_pass, gp_stats* stats) {
    using namespace gp;
    if (!ctx) return GP_ERR_ARG;
    try {
        GP_HIP(hipSetDevice(ctx->device));
        if (n < 2 || n > kMaxN)
            throw ArgError("gp_psd_filter_batch_static: word length " + std::to_string(n) +
                           " outside [2, 64] (the static scratch holds one table of at most 64 entries)");
        const int nb = (n + 15) / 16;
        if (count < 0 || (count > 0 && (!d_words || !d_pass)))
            throw ArgError("gp_psd_filter_batch_static: bad buffers");
        if (stride < 16 * nb || stride % 8 || reinterpret_cast<uintptr_t>(d_words) % 8)
            throw ArgError("gp_psd_filter_batch_static: rows must be 8-byte aligned with stride >= 16*ceil(n/16)");
        unsigned char* base = nullptr;
        {
            std::lock_guard<std::mutex> lk(g_na_mu);
            base = g_na_ptr[ctx->device];
        }
        if (!base) throw ArgError("gp_psd_filter_batch_static: static scratch not set up on this device");
        if (count == 0) return GP_OK;
        hipStream_t st = ctx->stream;
        NaHead* h = reinterpret_cast<NaHead*>(base);
        const uint4* frags = reinterpret_cast<const uint4*>(base + kNaFragOff);
        const double* trig64 = reinterpret_cast<const double*>(base + kNaTrigOff);
        {
            // the table of this word length, copied in when the scratch holds another
            std::lock_guard<std::mutex> lk(g_na_mu);
            if (g_na_n[ctx->device] != n) {
                const int id = ctx->tables.get(n);
                PsdTableDesc d = ctx->tables.desc[id];
                const size_t nfrag = (size_t)d.mt * 2 * d.nb * 64;
                if (nfrag > kNaFragCount) throw ArgError("gp_psd_filter_batch_static: table beyond the scratch");
                GP_HIP(hipMemcpyAsync(base + kNaFragOff, ctx->tables.frags.data() + d.frag_off, nfrag * sizeof(uint4),
                                      hipMemcpyHostToDevice, st));
                GP_HIP(hipMemcpyAsync(base + kNaTrigOff, ctx->tables.trig64.data() + d.trig_off,
                                      2 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, st));
                d.frag_off = 0;
                d.trig_off = 0;
                GP_HIP(hipMemcpyAsync(&h->desc, &d, sizeof d, hipMemcpyHostToDevice, st));
                GP_HIP(hipStreamSynchronize(st));
                g_na_n[ctx->device] = n;
            }
        }
        const int blocks = (int)std::min<int64_t>((count + 255) / 256, (int64_t)ctx->num_cus * 8);
        hipLaunchKernelGGL(k_na_reset, dim3(1), dim3(64), 0, st, h);
        hipLaunchKernelGGL(k_na_s1, dim3(blocks), dim3(256), 0, st, d_words, stride, n, count, h);
        hipLaunchKernelGGL(k_na_thr, dim3(1), dim3(64), 0, st, h, 2.0 * order - paf + 0.001);
        GP_HIP(hipGetLastError());
        SegTiles T{};
        T.nseg = 1;
        T.table[0] = 0;
        T.stride[0] = stride;
        T.first_row[0] = 0;
        T.rows[0] = count;
        T.byte0[0] = 0;
        GP_HIP(hipEventRecord(ctx->ev0, st));
        {
            // the one table in the scratch, at index 0 of its descriptor array
            std::vector<PsdTableDesc> dd(1, ctx->tables.desc[ctx->tables.get(n)]);
            launch_psd_mfma(T, dd, ctx->num_cus, st, d_words, &h->desc, frags, &h->thr, d_pass, nullptr);
        }
        GP_HIP(hipEventRecord(ctx->ev1, st));
        hipLaunchKernelGGL(k_na_band, dim3(blocks), dim3(256), 0, st, d_words, stride, count, trig64, h, d_pass);
        GP_HIP(hipGetLastError());
        GP_HIP(hipEventRecord(ctx->ev2, st));
        NaHead hh;
        GP_HIP(hipMemcpyAsync(&hh, h, sizeof hh, hipMemcpyDeviceToHost, st));
        GP_HIP(hipStreamSynchronize(st));
        if (stats) {
            float ms_main = 0, ms_band = 0;
            GP_HIP(hipEventElapsedTime(&ms_main, ctx->ev0, ctx->ev1));
            GP_HIP(hipEventElapsedTime(&ms_band, ctx->ev1, ctx->ev2));
            stats->psd_decisions += (uint64_t)count;
            stats->psd_pass += hh.passed;
            stats->band_rechecks += hh.ctl.count;
            stats->tie_band += hh.ctl.ties;
            stats->psd_kernel_ms += ms_main;
            stats->band_kernel_ms += ms_band;
            stats->psd_launches += 1;
            stats->kernel_ms += ms_main + ms_band;
        }
        return GP_OK;
    } catch (const ArgError& e) {
        ctx->err = e.what();
        return GP_ERR_ARG;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return GP_ERR_HIP;
    }
}
