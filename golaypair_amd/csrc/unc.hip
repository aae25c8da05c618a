// Uncompression driver: tables, filter launches, and the batched re-match.
//
// Re-match (match_pairs on the two expanded candidate files of each pair,
// uncompression.cpp:71-76).  For a pair (A, B) that passes check_if_pair at the
// finer compression c', PSD_A(k) + PSD_B(k) = Q for every k != 0 with
// Q = sum A^2 + sum B^2 - paf c'; with the keys of the expansion stage (A
// inverted, B plain, uncompression.cpp:199-204) the key texts are equal exactly
// when Q = 2v - paf (the argument of match.hip, roles exchanged).  So the pairs
// the reference emits are the PSD-filter survivors (A, B) of one input pair with
//   PAF_A(s) + PAF_B(s) = paf c'  (s = 1..N/2)  and  sum A^2 + sum B^2 = 2v - paf + paf c'.
// Every survivor gets a hash of its exact integer PAF vector (B: of paf c' - PAF
// and of 2v - paf + paf c' - sum B^2).  In a batch on the device the rows go to
// hash partitions, A and B apart, and each partition is joined in LDS (B rows
// in a table, A rows probing it); the synchronous path uses one global
// open-addressing table.  Every candidate is verified exactly before it is
// emitted.
// FP64 keys are computed for the emitted pairs only (their output order is the
// text order of the key lines, match_pairs.cpp:86-93), and the FP64 key equality
// of each emitted pair is checked (counted in tie_band when it fails).
#include <array>
#include <functional>
#include <numeric>
#include <queue>
#include <tuple>
#include <map>
#include <stdexcept>
#include <thread>
#include <unordered_map>

#include "threadpool.hpp"
#include "unc_kernels.inc"  // opens namespace gp (closed at the end of this file)

// ---------------------------------------------------------------------------
// FP32 (plain) filter: one leaf per lane, VALU FMAs over the suffix entries.

__global__ void __launch_bounds__(256) k_unc_filter_fp32(UncTabs tb, int nsides, int64_t nitems, UncCtl* ctl,
                                                         uint64_t* __restrict__ surv, int64_t scap,
                                                         uint64_t* __restrict__ band, int64_t bcap) {
    const UncSide* __restrict__ sides = tb.sides;
    const int32_t* __restrict__ radix = tb.radix;
    const float* __restrict__ c32 = tb.c32;
    const float* __restrict__ mid = tb.mid;
    const int8_t* __restrict__ ent = tb.ent;
    const float* __restrict__ w32 = tb.w32;
    __shared__ float xpre[4][64];
    __shared__ uint64_t wbuf_all[4][kWaveBuf];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint64_t* wbuf = wbuf_all[wv];
    int wcount = 0;
    const int64_t wave = (int64_t)blockIdx.x * 4 + wv;
    const int64_t nwaves = (int64_t)gridDim.x * 4;
    const unsigned long long below = (1ull << lane) - 1ull;
    int sid = 0;
    for (int64_t it = wave; it < nitems; it += nwaves) {
        while (sid + 1 < nsides && it >= sides[sid + 1].item_base) ++sid;
        const UncSide& S = sides[sid];
        const int64_t o = it - S.item_base;
        const int mt = (S.nch + 15) / 16;
        const int nch = mt * 32;
        const int nent = S.nent;
        // X_outer, one channel per lane: the single-split constant plus the outer digits
        float xo = lane < nch ? tb.fix32[S.fix_off + lane] : 0.f;
        {
            int64_t rem = o;
            for (int x = S.nout - 1; x >= 0; --x) {
                const int j = tb.olist[S.olist_off + x];
                const int R = radix[S.radix_off + j];
                const int d = (int)(rem % R);
                rem /= R;
                if (lane < nch) xo += c32[tb.pos32[S.c32_off + j] + (int64_t)d * nch + lane];
            }
        }
        for (int64_t m = 0; m < S.p_mid; ++m) {
            if (lane < nch) xpre[wv][lane] = xo + mid[S.mid_off + m * nch + lane];
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            const int64_t leaf0 = (o * S.p_mid + m) * S.p_suf;
            for (int64_t s0 = 0; s0 < S.p_suf; s0 += 64) {
                const int64_t s = s0 + lane;
                const bool valid = s < S.p_suf;
                float mx = 0.f;
                if (valid) {
                    const int8_t* e = ent + S.ent_off + s * nent;
                    for (int k = 0; k < S.nch; ++k) {
                        const int t = k >> 4, row = k & 15;
                        float re = xpre[wv][t * 32 + row], im = xpre[wv][t * 32 + 16 + row];
                        for (int q = 0; q < nent; ++q) {
                            const float x = (float)e[q];
                            re = __builtin_fmaf(x, w32[S.w32_off + (int64_t)q * nch + t * 32 + row], re);
                            im = __builtin_fmaf(x, w32[S.w32_off + (int64_t)q * nch + t * 32 + 16 + row], im);
                        }
                        mx = fmaxf(mx, __builtin_fmaf(re, re, im * im));
                    }
                }
                const bool acc_ok = valid && mx < S.thr_lo;
                const bool in_band = valid && !(mx > S.thr_hi) && !acc_ok;
                const unsigned long long ba = __ballot(acc_ok);
                const uint64_t rec = ((uint64_t)sid << 40) | (uint64_t)(leaf0 + s);
                if (ba) {
                    if (acc_ok) wbuf[wcount + __popcll(ba & below)] = rec;
                    wcount += __popcll(ba);
                    if (wcount > kWaveBuf - 64) flush_wave(wbuf, wcount, lane, ctl, surv, scap);
                }
                if (in_band) {
                    const unsigned long long i = atomicAdd(&ctl->band, 1ull);
                    if ((int64_t)i < bcap) band[i] = rec;
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    if (wcount) flush_wave(wbuf, wcount, lane, ctl, surv, scap);
}

// ---------------------------------------------------------------------------
// survivor rows: PAF-vector hash

__device__ __forceinline__ uint64_t mix_in(uint64_t h, int64_t v) {
    return (h ^ (uint64_t)(v + 0x100000)) * 0x100000001B3ull + 0x9E3779B97F4A7C15ull;
}

// floor(2^64 / R) + 1 for R = 2 .. kDivMagicMax.  For a 32-bit n the quotient
// n / R is the high word of n * magic[R]: with 2^64 = R Q + rho the product is
// n 2^64 / R + n (R - rho) / R, whose second term adds less than n / 2^64 < 1/R
// to n / R, too little to reach the next integer.
constexpr int kDivMagicMax = 256;
struct DivMagic {
    uint64_t m[kDivMagicMax + 1];
    constexpr DivMagic() : m() {
        for (int r = 2; r <= kDivMagicMax; ++r) m[r] = ~0ull / (uint64_t)r + 1ull;
    }
};
__constant__ DivMagic kDivMagic = DivMagic();

// Decodes a record into its word (N int8) and, when the side has +1 masks,
// the word's +1 bit mask.
// Digit of `leaf` in radix R (R >= 1); `leaf` becomes the quotient.
__device__ __forceinline__ int next_digit(uint64_t& leaf, uint32_t R) {
    if (R == 1) return 0;  // single-split position: digit 0, index unchanged
    if (R == 2) {
        const int d = (int)(leaf & 1);
        leaf >>= 1;
        return d;
    }
    if (leaf < 0x100000000ull && R <= (uint32_t)kDivMagicMax) {  // multiply-high division
        const uint32_t l32 = (uint32_t)leaf;
        const uint32_t q = (uint32_t)__umul64hi((uint64_t)l32, kDivMagic.m[R]);
        leaf = q;
        return (int)(l32 - q * R);
    }
    const int d = (int)(leaf % R);
    leaf /= R;
    return d;
}

// Positions are taken 8 at a time (last position first): their radixes load
// together, then their digits, then their tuple and mask loads, so that a
// survivor waits for a few memory round trips instead of two per position.
__device__ __forceinline__ uint64_t decode_word(const UncSide& S, uint64_t leaf, const int32_t* __restrict__ radix,
                                                const int8_t* __restrict__ tup, const uint64_t* __restrict__ pm1,
                                                int8_t* w) {
    uint64_t mask = 0;
    const bool masks = pm1 && S.pm1_off >= 0;
    for (int j0 = S.L - 1; j0 >= 0; j0 -= 8) {
        uint32_t R[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) R[k] = j0 - k >= 0 ? (uint32_t)radix[S.radix_off + j0 - k] : 1u;
        int d[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) d[k] = next_digit(leaf, R[k]);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int j = j0 - k;
            if (j < 0) break;
            if (w) {
                const int8_t* t = tup + S.tup_off + ((int64_t)j * S.rmax + d[k]) * S.r;
                for (int i = 0; i < S.r; ++i) w[j + S.L * i] = t[i];
            }
            if (masks) mask |= pm1[S.pm1_off + (int64_t)j * S.rmax + d[k]];
        }
    }
    return mask;
}

struct JoinCtl2 {
    unsigned long long out;       // emitted pairs
    unsigned long long checks;    // full PAF checks
    unsigned long long ncand;     // candidate pairs (equal hashes)
    unsigned long long overflow;  // a join sub-bucket held more than kSubCap rows
};

// The batched join partitions the survivors by hash: P partitions (a power of
// two, 64 <= P <= pmax, about 512 rows or fewer each on average), each with
// kPartCap slots for A rows and kPartCap for B rows, records
// (hash bits 0..31 << 32) | survivor index.  The partition is hash bits 40..
// A partition's slots of one role are kPartSub sub-buckets of kSubCap, one per
// XCD (blocks go to the 8 XCDs round-robin by block index): each sub-bucket's
// cache lines are then written through one XCD's L2 only, and leave it whole
// (appends from several XCDs to one line wrote back a partial line each).
constexpr int kPartCap = 1024;
constexpr int kPartSub = 8;
constexpr int kSubCap = kPartCap / kPartSub;
__host__ __device__ inline int64_t part_count(int64_t n, int64_t pmax) {
    int64_t P = 64;
    while (P < pmax && P * 512 < n) P <<= 1;
    return P;
}

// Clears two small device records (the join counters and the tie count of a
// batch) with vector stores: a kernel node, so that the batch's captured graph
// holds no memset node.
__global__ void k_zero_ctl(unsigned int* __restrict__ a, int na, unsigned int* __restrict__ b, int nb) {
    const int t = threadIdx.x;
    if (t < na) a[t] = 0u;
    if (t < nb) b[t] = 0u;
}

// Survivor rows -> PAF-vector hashes: into `hashes` (the row's role in bit 0),
// or, with `part`, as records into the join partitions (pcnt: A then B row
// counts per partition, zero on entry).
// kAllMasks: every side of the batch carries +1 masks (every tuple entry is
// +-1), so no word is built (and the kernel needs no private array).
template <bool kAllMasks>
__global__ void k_unc_rows(const UncSide* __restrict__ sides, const int32_t* __restrict__ radix,
                           const int8_t* __restrict__ tup, const uint64_t* __restrict__ pm1,
                           const uint64_t* __restrict__ surv, int64_t m, int paf_target, int q_target,
                           uint64_t* __restrict__ hashes, const unsigned long long* dcount, int64_t dcap,
                           JoinCtl2* jctl, uint64_t* __restrict__ part, uint32_t* __restrict__ pcnt, int64_t pmax) {
    int8_t w[kAllMasks ? 1 : kMaxN];
    m = list_count(dcount, m, dcap);
    const uint64_t pmask = part ? (uint64_t)part_count(m, pmax) - 1 : 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t rec = surv[i];
        const UncSide& S = sides[rec >> 40];
        const int N = S.N;
        const bool fast = kAllMasks || S.pm1_off >= 0;
        const uint64_t mask = decode_word(S, rec & ((1ull << 40) - 1), radix, tup, pm1, fast ? nullptr : w);
        const uint64_t full = N >= 64 ? ~0ull : ((1ull << N) - 1);
        int64_t sq = 0;
        if (fast) sq = N;
        else for (int p = 0; p < N; ++p) sq += (int)w[p] * (int)w[p];
        uint64_t h = 0xC6A4A7935BD1E995ull ^ (uint64_t)S.group;
        h = mix_in(h, S.role ? (int64_t)q_target - sq : sq);
        for (int s = 1; s <= N / 2; ++s) {
            int64_t v;
            if (fast) {
                const uint64_t rot = ((mask >> s) | (mask << (N - s))) & full;
                v = N - 2 * (int)__popcll(mask ^ rot);  // __popcll is unsigned
            } else {
                v = 0;
                int j = s;
                for (int p = 0; p < N; ++p) {
                    v += (int)w[p] * (int)w[j];
                    if (++j == N) j = 0;
                }
            }
            h = mix_in(h, S.role ? (int64_t)paf_target - v : v);
        }
        const uint64_t hv = splitmix64(h);
        if (part) {
            const uint64_t q = ((uint64_t)(S.role == 1) * pmax + ((hv >> 40) & pmask)) * kPartSub +
                               blockIdx.x % kPartSub;
            const uint32_t slot = atomicAdd(&pcnt[q], 1u);
            if (slot < (uint32_t)kSubCap) part[q * kSubCap + slot] = (hv << 32) | (uint64_t)(uint32_t)i;
            else jctl->overflow = 1ull;
        } else {
            // the row's role in bit 0, so that insert and probe read one array
            hashes[i] = (hv << 1) | (uint64_t)(S.role == 1);
        }
    }
}

__global__ void k_zero_u32(uint32_t* __restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// Join of each partition in LDS: its B records go into an open-addressing table
// of 2 kPartCap slots (at most half full), its A records probe it; every B row
// with the A row's 32-bit tag in the same partition is a candidate pair,
// checked exactly by k_pair_check (tags that collide are dropped there).  A
// block loads its partition's records (kPartCap / 256 slots of each role per
// thread, slot i in sub-bucket i / kSubCap) before it clears the table, so the
// loads overlap the clearing.
__global__ void __launch_bounds__(256) k_part_join(const uint64_t* __restrict__ part, const uint32_t* __restrict__ pcnt,
                                                   int64_t pmax, const unsigned long long* dcount, int64_t dcap,
                                                   uint32_t* __restrict__ cand_a, uint32_t* __restrict__ cand_b,
                                                   int64_t ccap, JoinCtl2* ctl) {
    constexpr uint32_t T = 2 * kPartCap;
    constexpr int kPer = kPartCap / 256;
    __shared__ unsigned long long tab[T];
    const int64_t P = part_count(list_count(dcount, 0, dcap), pmax);
    for (int64_t p = blockIdx.x; p < P; p += gridDim.x) {
        const uint32_t* ca = pcnt + p * kPartSub;
        const uint32_t* cb = pcnt + (pmax + p) * kPartSub;
        const uint64_t* __restrict__ ra = part + p * kPartCap;
        const uint64_t* __restrict__ rb = part + (pmax + p) * kPartCap;
        unsigned long long ea[kPer], eb[kPer];
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const uint32_t i = q * 256 + threadIdx.x, x = i / kSubCap, slot = i % kSubCap;
            ea[q] = slot < ca[x] ? ra[i] : ~0ull;
            eb[q] = slot < cb[x] ? rb[i] : ~0ull;
        }
        for (uint32_t e = threadIdx.x; e < T; e += blockDim.x) tab[e] = ~0ull;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            if (eb[q] == ~0ull) continue;
            uint32_t s = (uint32_t)(eb[q] >> 32) & (T - 1);
            while (atomicCAS(&tab[s], ~0ull, eb[q]) != ~0ull) s = (s + 1) & (T - 1);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            if (ea[q] == ~0ull) continue;
            const uint32_t tag = (uint32_t)(ea[q] >> 32);
            for (uint32_t s = tag & (T - 1);; s = (s + 1) & (T - 1)) {
                const unsigned long long x = tab[s];
                if (x == ~0ull) break;
                if ((uint32_t)(x >> 32) == tag) {
                    const unsigned long long o = atomicAdd(&ctl->ncand, 1ull);
                    if ((int64_t)o < ccap) { cand_a[o] = (uint32_t)ea[q]; cand_b[o] = (uint32_t)x; }
                }
            }
        }
        __syncthreads();  // the table is cleared for the next partition
    }
}

constexpr uint64_t kEmpty = ~0ull;

// Global-table join (the synchronous path): B rows (bit 0 of the hash set) go
// into the table, tagged with the upper hash bits.
__global__ void k_ht_insert(const uint64_t* __restrict__ hashes, int64_t m, unsigned long long* __restrict__ table,
                            uint64_t tmask) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t hr = hashes[i];
        if (!(hr & 1)) continue;  // B rows only
        const uint64_t h = hr >> 1;
        const unsigned long long e = ((h & 0xFFFFFFFFull) << 32) | (uint64_t)(uint32_t)i;
        uint64_t slot = (h >> 24) & tmask;
        while (atomicCAS(&table[slot], kEmpty, e) != kEmpty) slot = (slot + 1) & tmask;
    }
}

// A rows probe the table of B rows: every B row with the same PAF hash is a
// candidate pair (A row i, B row j), checked (group and PAF, exactly) by
// k_pair_check.  Hashes carry the row's role in bit 0.
__global__ void k_ht_probe(const uint64_t* __restrict__ hashes, int64_t m, const unsigned long long* __restrict__ table,
                           uint64_t tmask, uint32_t* __restrict__ cand_a, uint32_t* __restrict__ cand_b, int64_t ccap,
                           JoinCtl2* ctl) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t hr = hashes[i];
        if (hr & 1) continue;  // A rows only
        const uint64_t h = hr >> 1;
        uint64_t slot = (h >> 24) & tmask;
        for (;;) {
            const unsigned long long e = table[slot];
            if (e == kEmpty) break;
            if ((e >> 32) == (h & 0xFFFFFFFFull)) {
                const uint32_t j = (uint32_t)(e & 0xFFFFFFFFull);
                if ((hashes[j] >> 1) == h) {
                    const unsigned long long o = atomicAdd(&ctl->ncand, 1ull);
                    if ((int64_t)o < ccap) { cand_a[o] = (uint32_t)i; cand_b[o] = j; }
                }
            }
            slot = (slot + 1) & tmask;
        }
    }
}

// Exact check of the candidate pairs: sum a_p^2 + b_p^2 = q_target and
// PAF_a(s) + PAF_b(s) = paf_target for s = 1..N/2.  Blocks of one wave, one
// pair per half-wave: the two words are decoded into LDS, lane q checks
// s = q + 1 (lane 0 also the squares); passing pairs are appended to out.
__global__ void __launch_bounds__(64) k_pair_check(const UncSide* __restrict__ sides,
                                                   const int32_t* __restrict__ radix,
                                                   const int8_t* __restrict__ tup, const uint64_t* __restrict__ surv,
                                                   const uint32_t* __restrict__ cand_a,
                                                   const uint32_t* __restrict__ cand_b, int64_t ccap, int paf_target,
                                                   int q_target, uint32_t* __restrict__ out_a,
                                                   uint32_t* __restrict__ out_b, int64_t cap, JoinCtl2* ctl) {
    __shared__ int8_t wl[2][2][kMaxN];  // [half][A, B]
    const int64_t nc = list_count(&ctl->ncand, 0, ccap);
    const int h = threadIdx.x >> 5, q = threadIdx.x & 31;
    unsigned long long checks = 0;
    for (int64_t i0 = 2 * (int64_t)blockIdx.x; i0 < nc; i0 += 2 * (int64_t)gridDim.x) {
        const int64_t i = i0 + h;
        const bool live = i < nc;
        int N = 0;
        bool same = false;  // both rows from one input pair
        if (live) {
            const uint64_t rec = surv[q == 1 ? cand_b[i] : cand_a[i]];
            const UncSide& S = sides[rec >> 40];
            N = S.N;
            same = S.group == sides[surv[q == 1 ? cand_a[i] : cand_b[i]] >> 40].group;
            if (q < 2 && same) decode_word(S, rec & ((1ull << 40) - 1), radix, tup, nullptr, wl[h][q]);
        }
        __syncthreads();
        bool ok = true;
        if (live && same) {
            const int8_t* wa = wl[h][0];
            const int8_t* wb = wl[h][1];
            const int s = q + 1;
            if (s <= N / 2) {
                int acc = 0, k = s;
                for (int p = 0; p < N; ++p) {
                    acc += (int)wa[p] * wa[k] + (int)wb[p] * wb[k];
                    if (++k == N) k = 0;
                }
                ok = acc == paf_target;
            }
            if (q == 0) {
                int sq = 0;
                for (int p = 0; p < N; ++p) sq += (int)wa[p] * wa[p] + (int)wb[p] * wb[p];
                ok = ok && sq == q_target;
            }
        }
        const uint64_t bad = __ballot(!ok);
        if (live && same && q == 0) {
            ++checks;
            if (((bad >> (32 * h)) & 0xFFFFFFFFull) == 0) {
                const unsigned long long o = atomicAdd(&ctl->out, 1ull);
                if ((int64_t)o < cap) { out_a[o] = cand_a[i]; out_b[o] = cand_b[i]; }
            }
        }
        __syncthreads();  // wl is rewritten by the next iteration
    }
    wave_atomic_add(&ctl->checks, checks);
}

// Emitted rows -> words + FP64 key digits (A inverted, B plain) + group of
// rows i < ngroups.  Blocks of one wave, one row per half-wave: the row's word
// is decoded into LDS, lane q of the half computes PSD_{q+1} (the order of
// fp64_psd) and writes bytes q and q + 32 of the word.
template <class RecOf>
__device__ __forceinline__ void emit_rows_halfwave(int64_t m, int64_t ngroups, RecOf rec_of,
                                                   const UncSide* __restrict__ sides,
                                                   const int32_t* __restrict__ radix, const int8_t* __restrict__ tup,
                                                   const double* __restrict__ trig, int bound,
                                                   int8_t* __restrict__ words, uint8_t* __restrict__ keys,
                                                   int32_t* __restrict__ groups, unsigned int* __restrict__ ties) {
    __shared__ int8_t wl[2][kMaxN];
    const int h = threadIdx.x >> 5, q = threadIdx.x & 31;
    for (int64_t i0 = 2 * (int64_t)blockIdx.x; i0 < m; i0 += 2 * (int64_t)gridDim.x) {
        const int64_t i = i0 + h;
        const bool live = i < m;
        const uint64_t rec = live ? rec_of(i) : 0;
        const UncSide& S = sides[rec >> 40];
        if (live && q == 0) decode_word(S, rec & ((1ull << 40) - 1), radix, tup, nullptr, wl[h]);
        __syncthreads();
        if (live) {
            const int N = S.N;
            if (q < N) words[i * N + q] = wl[h][q];
            if (q + 32 < N) words[i * N + q + 32] = wl[h][q + 32];
            if (q == 0 && i < ngroups) groups[i] = S.group;
            if (q < S.K) {
                const double p = fp64_psd_k(wl[h], N, trig, trig + N, q + 1);
                const double f = p - floor(p);
                if (fabs(f - 0.5) < 1e-9) atomicAdd(ties, 1u);
                const int v = (int)rint(p);
                keys[i * S.K + q] = (uint8_t)(S.seqflag ? v : bound - v);
            }
        }
        __syncthreads();  // wl is rewritten by the next iteration
    }
}

// Emitted pairs (device count) -> rows [0, ne) the A words/keys, [ne, 2 ne) the
// B ones, and the pair's group.
__global__ void __launch_bounds__(64) k_emit_pairs(const UncSide* __restrict__ sides,
                                                   const int32_t* __restrict__ radix,
                                                   const int8_t* __restrict__ tup, const uint64_t* __restrict__ surv,
                                                   const uint32_t* __restrict__ oa, const uint32_t* __restrict__ ob,
                                                   const unsigned long long* __restrict__ dcount, int64_t cap,
                                                   const double* __restrict__ trig, int bound,
                                                   int8_t* __restrict__ words, uint8_t* __restrict__ keys,
                                                   int32_t* __restrict__ groups, unsigned int* __restrict__ ties) {
    const int64_t ne = list_count(dcount, 0, cap);
    emit_rows_halfwave(
        2 * ne, ne, [&](int64_t i) { return surv[i < ne ? oa[i] : ob[i - ne]]; }, sides, radix, tup, trig, bound,
        words, keys, groups, ties);
}

// Emitted rows (surv[rows[i]]) -> words, FP64 keys and groups.
__global__ void __launch_bounds__(64) k_emit_rows(const UncSide* __restrict__ sides,
                                                  const int32_t* __restrict__ radix,
                                                  const int8_t* __restrict__ tup, const uint64_t* __restrict__ surv,
                                                  const uint32_t* __restrict__ rows, int64_t m,
                                                  const double* __restrict__ trig, int bound,
                                                  int8_t* __restrict__ words, uint8_t* __restrict__ keys,
                                                  int32_t* __restrict__ groups, unsigned int* __restrict__ ties) {
    emit_rows_halfwave(
        m, m, [&](int64_t i) { return surv[rows[i]]; }, sides, radix, tup, trig, bound, words, keys, groups, ties);
}

// One suffix operand table of the split filter: the B fragments of every
// combination of the suffix positions' tuples (mixed-radix order, last
// position fastest), laid out [tile][block][lane] as uint4 of 8 BF16 (entry
// e = 16 block + 8 (lane / 32) + i of combination 32 tile + lane % 32, entry
// e = x r + i being element i of suffix position x's tuple).  The tile count
// is rounded up to even (the filter loads tiles in pairs); padding is zero.  `side` is a
// table side with these positions and letters (its tuple rows are used).
struct SufJob {
    int64_t out_off;  // in uint4
    int32_t side, npos, pos_off, pad;
};

__global__ void __launch_bounds__(256) k_build_suf(const SufJob* __restrict__ jobs, const int32_t* __restrict__ pos,
                                                   const UncSide* __restrict__ sides,
                                                   const int32_t* __restrict__ radix,
                                                   const int8_t* __restrict__ tup, uint4* __restrict__ sufb) {
    const SufJob J = jobs[blockIdx.x];
    const UncSide& S = sides[J.side];
    const int r = S.r, nent = J.npos * r, nbs = nent > 16 ? (nent + 15) / 16 : 1;
    const int32_t* sp = pos + J.pos_off;
    int64_t ps = 1;
    for (int x = 0; x < J.npos; ++x) ps *= radix[S.radix_off + sp[x]];
    const int64_t ntiles = (ps + 63) / 64 * 2;
    uint4* out = sufb + J.out_off;
    for (int64_t u = threadIdx.x; u < ntiles * nbs * 64; u += blockDim.x) {
        const int lane = (int)(u & 63), blk = (int)((u >> 6) % nbs);
        const int64_t tile = (u >> 6) / nbs;
        const int64_t s0 = tile * 32 + (lane & 31);
        uint32_t hv[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (s0 < ps) {
            int dg[kMaxN];
            int64_t rest = s0;
            for (int x = J.npos - 1; x >= 0; --x) {
                const int R = radix[S.radix_off + sp[x]];
                dg[x] = (int)(rest % R);
                rest /= R;
            }
            for (int i8 = 0; i8 < 8; ++i8) {
                const int e = 16 * blk + 8 * (lane >> 5) + i8;
                if (e < nent) {
                    const int x = e / r, i = e % r;
                    const float v = (float)tup[S.tup_off + ((int64_t)sp[x] * S.rmax + dg[x]) * r + i];
                    hv[i8] = __float_as_uint(v) >> 16;  // small integers are exact in BF16
                }
            }
        }
        out[u] = make_uint4(hv[0] | (hv[1] << 16), hv[2] | (hv[3] << 16), hv[4] | (hv[5] << 16),
                            hv[6] | (hv[7] << 16));
    }
}

// Records of one table side (unordered) -> compact list.
__global__ void k_select_side(const uint64_t* __restrict__ surv, int64_t m, uint64_t sid,
                              unsigned long long* __restrict__ count, uint64_t* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t r = surv[i];
        if ((r >> 40) == sid) out[atomicAdd(count, 1ull)] = r;
    }
}

// ---------------------------------------------------------------------------
// host: tables

namespace {

uint16_t bf16_rne_u(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
double bf16_d(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// Per-side tables (concatenated over the sides of a call), views into the
// batch's staging buffer.
struct Tables {
    int64_t nsides = 0;
    UncSide* sides = nullptr;
    int32_t* radix = nullptr;
    int8_t* tup = nullptr;
    uint64_t* pm1 = nullptr;
    int32_t* pos32 = nullptr;  // per side and position: block offsets into Shared
    int32_t* pos64 = nullptr;
    int32_t* olist = nullptr;  // per side: outer positions (active, in order)
    float* fix32 = nullptr;    // per side: FP32 sum of the single-split positions
    int64_t items = 0;
};

// Tables shared by every side of a call, content-addressed: the per-(position,
// letter) DFT contribution blocks (FP32 channel layout and FP64 (Re, Im) pairs),
// the MFMA A fragments and FP32 weights of each suffix start b, the mid tables of
// each (a, b, letters at a..b-1) and the suffix operands of each (b, letters at b..L-1).
struct Shared {
    std::vector<int> block;  // [position j][letter + 128] -> block id, -1 if absent
    std::vector<int32_t> off32, off64;
    std::vector<float> c32;
    std::vector<double> c64;
    // suffix weight sets (one per distinct suffix position set): positions,
    // offset into frags (split path) or w32 (FP32 path), max BF16 split residual
    std::vector<std::vector<int>> sets;
    std::vector<int64_t> set_off;
    std::vector<double> set_rho;
    int64_t n_frags = 0, n_w32 = 0, n_mid = 0, n_suf = 0;
    // views into the staging buffer
    uint4* frags = nullptr;  // MFMA A fragments per suffix position set (split path)
    float* w32 = nullptr;    // FP32 suffix weights (FP32 path)
    float* mid = nullptr;    // mid tables
    uint4* sufb = nullptr;   // suffix B operands (split path)
    int8_t* ent = nullptr;   // suffix entries (FP32 path)
    // FP32 filter channels: slot q (row q % 16 of M-tile q / 16) holds k = chan[q]
    std::vector<int> chan;   // 0: an empty slot of the last M-tile (zero weights)
    int lr = 8;              // registers per lane read in the last M-tile (8, 4 or 2)
    int mt() const { return ((int)chan.size() + 15) / 16; }
    int id(int j, int letter) const {
        const int b = block[(size_t)j * 256 + (letter + 128)];
        if (b < 0) throw std::logic_error("uncompress: missing position block");
        return b;
    }
};

// The split of a side's leaf index over its active positions (more than one
// split; single-split positions add a constant and digit 0 of radix 1, so the
// leaf index of the DFS order is the mixed-radix number of the active digits):
// suffix = the last active positions (p_suf >= 64, grown while the entries fit
// the K blocks already used and p_suf stays <= 1024), mid = the active
// positions before it (16 <= p_mid <= 4096 where possible), outer = the rest.
struct Split {
    std::vector<int> out, mid, suf;  // position indices, ascending
    int64_t po = 1, pm = 1, ps = 1;
    int nent = 0;
};

Split side_split(const SideHost& S) {
    Split X;
    std::vector<int> act;
    for (int j = 0; j < S.L; ++j)
        if (S.R[j] > 1) act.push_back(j);
    int k = (int)act.size();
    while (k > 0) {
        const int j = act[k - 1];
        if (X.ps >= 64) {
            const int nbs_now = std::max(1, (X.nent + 15) / 16);
            if (X.nent + S.r > 16 * nbs_now || X.ps * S.R[j] > 1024) break;
        }
        X.ps *= S.R[j];
        X.nent += S.r;
        --k;
    }
    X.suf.assign(act.begin() + k, act.end());
    int a = k;
    while (a > 0 && X.pm < 16 && X.pm * S.R[act[a - 1]] <= 4096) X.pm *= S.R[act[--a]];
    X.mid.assign(act.begin() + a, act.begin() + k);
    X.out.assign(act.begin(), act.begin() + a);
    for (int j : X.out) X.po *= S.R[j];
    return X;
}

// Content keys: the positions, then the letters at them, one byte each.
std::string pos_key(const SideHost& S, const std::vector<int>& pos) {
    std::string k;
    k.reserve(2 * pos.size());
    for (int j : pos) k.push_back((char)j);
    for (int j : pos) k.push_back((char)(S.orig[j] + 128));
    return k;
}

// Mid table of (mid positions, letters): FP64 sums over the mid digits, rounded
// once to FP32, [p_mid][mt*32] (last position fastest).
void build_mid(const Shared& SH, const SideHost& S, const std::vector<int>& mp, float* out) {
    const int K = S.N / 2, mt = SH.mt();
    int64_t pm = 1;
    for (int j : mp) pm *= (int64_t)S.R[j];
    std::fill(out, out + (size_t)pm * mt * 32, 0.f);
    std::vector<double> acc((size_t)2 * K);
    std::vector<int> dg(mp.size(), 0);
    for (int64_t m = 0; m < pm; ++m) {
        std::fill(acc.begin(), acc.end(), 0.0);
        for (size_t x = 0; x < mp.size(); ++x) {
            const double* c = &SH.c64[SH.off64[SH.id(mp[x], S.orig[mp[x]])] + (int64_t)dg[x] * K * 2];
            for (int q = 0; q < 2 * K; ++q) acc[q] += c[q];
        }
        for (size_t q = 0; q < SH.chan.size(); ++q) {
            const int k = SH.chan[q], t = (int)q / 16, row = (int)q % 16;
            if (k <= 0) continue;  // empty slot
            out[m * mt * 32 + t * 32 + row] = (float)acc[2 * (k - 1)];
            out[m * mt * 32 + t * 32 + 16 + row] = (float)acc[2 * (k - 1) + 1];
        }
        for (int x = (int)mp.size() - 1; x >= 0; --x) {
            if (++dg[x] < S.R[mp[x]]) break;
            dg[x] = 0;
        }
    }
}

// Suffix operands of (suffix positions, letters): BF16 B fragments
// [ntiles][nbs][64] (split path) or int8 entries [p_suf][nent] (FP32 path).
// Entry e = x r + i is element i of the tuple of suffix position x; suffix
// combinations in mixed-radix order, last position fastest.
void build_suf(const SideHost& S, const std::vector<int>& sp, bool fp32, uint4* sufb, int8_t* ent) {
    const int r = S.r, nent = (int)sp.size() * r;
    int64_t ps = 1;
    for (int j : sp) ps *= (int64_t)S.R[j];
    std::vector<int> sdg(sp.size(), 0);
    std::vector<int8_t> vals((size_t)ps * std::max(nent, 1), 0);
    for (int64_t s = 0; s < ps; ++s) {
        for (int e = 0; e < nent; ++e) vals[s * nent + e] = S.tuple(sp[e / r], sdg[e / r])[e % r];
        for (int x = (int)sp.size() - 1; x >= 0; --x) {
            if (++sdg[x] < S.R[sp[x]]) break;
            sdg[x] = 0;
        }
    }
    if (fp32) {
        std::copy(vals.begin(), vals.end(), ent);
        return;
    }
    // BF16 of every int8 value (exact: |v| <= 128 has at most 8 significant bits)
    static const std::array<uint16_t, 256> bf = [] {
        std::array<uint16_t, 256> t{};
        for (int v = -128; v < 128; ++v) t[v + 128] = bf16_rne_u((float)v);
        return t;
    }();
    const int nbs = std::max(1, (nent + 15) / 16);
    const int ntiles = (int)((ps + 63) / 64 * 2);  // even: tiles go in pairs
    for (int tile = 0; tile < ntiles; ++tile)
        for (int blk = 0; blk < nbs; ++blk)
            for (int lane = 0; lane < 64; ++lane) {
                uint16_t hv[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                const int64_t s = (int64_t)tile * 32 + lane % 32;
                if (s < ps)
                    for (int i8 = 0; i8 < 8; ++i8) {
                        const int e = 16 * blk + 8 * (lane / 32) + i8;
                        if (e < nent) hv[i8] = bf[vals[s * nent + e] + 128];
                    }
                uint4 q;
                q.x = hv[0] | ((uint32_t)hv[1] << 16);
                q.y = hv[2] | ((uint32_t)hv[3] << 16);
                q.z = hv[4] | ((uint32_t)hv[5] << 16);
                q.w = hv[6] | ((uint32_t)hv[7] << 16);
                *sufb++ = q;
            }
}

// Element counts of the mid and suffix tables of a split (their layouts above).
int64_t mid_size(const Split& X, int mt) { return X.pm * mt * 32; }
int64_t suf_size(const Split& X, bool fp32) {
    if (fp32) return X.ps * std::max(X.nent, 1);
    return ((X.ps + 63) / 64 * 2) * std::max(1, (X.nent + 15) / 16) * 64;
}

// FP32 suffix DFT weights of the suffix positions, [nent][mt*32] (FP32 path).
// Registers a new suffix position set `sp`: its offset and size in w32 (FP32
// path) or frags (split path); returns its index.
int add_set(Shared& SH, const std::vector<int>& sp, int r, bool fp32) {
    const int nent = (int)sp.size() * r, nbs = std::max(1, (nent + 15) / 16);
    SH.sets.push_back(sp);
    SH.set_rho.push_back(0.0);
    if (fp32) {
        SH.set_off.push_back(SH.n_w32);
        SH.n_w32 += (int64_t)nent * SH.mt() * 32;
    } else {
        SH.set_off.push_back(SH.n_frags);
        SH.n_frags += (int64_t)SH.mt() * 2 * nbs * 64;
    }
    return (int)SH.sets.size() - 1;
}

// Writes the FP32 weights of set `set`.
void build_w32(const Shared& SH, int set, int L, int r, int N, const std::vector<double>& cs,
               const std::vector<double>& sn) {
    const std::vector<int>& sp = SH.sets[set];
    const int mt = SH.mt(), nent = (int)sp.size() * r;
    float* out = SH.w32 + SH.set_off[set];
    for (int e = 0; e < nent; ++e) {
        const int j = sp[e / r], i = e % r, p = j + L * i;
        for (int t = 0; t < mt; ++t)
            for (int row = 0; row < 32; ++row) {
                const size_t q = 16 * t + row % 16;
                float v = 0.f;
                if (q < SH.chan.size() && SH.chan[q] > 0) {
                    const int mm = (int)(((int64_t)SH.chan[q] * p) % N);
                    v = (float)(row >= 16 ? sn[mm] : cs[mm]);
                }
                *out++ = v;
            }
    }
}

// Blocks of every position j < L and every letter of the split table: the DFT
// contributions of each of the position's tuples (FP64 (Re, Im) for all k, and
// the FP32 filter channels).
void add_blocks(Shared& SH, const SplitTable& X, int L, int N, const std::vector<double>& cs,
                const std::vector<double>& sn) {
    const int r = X.r, K = N / 2, mt = SH.mt();
    SH.block.assign((size_t)L * 256, -1);
    for (int j = 0; j < L; ++j)
        for (const auto& [letter, lid] : X.letter_id) {
            const int R = j == 0 ? X.nfirst[lid] : X.nsplit[lid];
            const int8_t* tp = j == 0 ? X.first[lid].data() : X.split[lid].data();
            SH.block[(size_t)j * 256 + (letter + 128)] = (int)SH.off32.size();
            SH.off32.push_back((int32_t)SH.c32.size());
            SH.off64.push_back((int32_t)SH.c64.size());
            SH.c32.resize(SH.c32.size() + (size_t)std::max(R, 1) * mt * 32, 0.f);
            SH.c64.resize(SH.c64.size() + (size_t)std::max(R, 1) * 2 * K, 0.0);
            float* c32 = &SH.c32[SH.off32.back()];
            double* c64 = &SH.c64[SH.off64.back()];
            for (int d = 0; d < R; ++d)
                for (int k = 1; k <= K; ++k) {
                    double re = 0, im = 0;
                    for (int i = 0; i < r; ++i) {
                        const int p = j + L * i;
                        const int mm = (int)(((int64_t)k * p) % N);
                        re += tp[d * r + i] * cs[mm];
                        im += tp[d * r + i] * sn[mm];
                    }
                    c64[((int64_t)d * K + (k - 1)) * 2] = re;
                    c64[((int64_t)d * K + (k - 1)) * 2 + 1] = im;
                }
            for (int d = 0; d < R; ++d)
                for (size_t q = 0; q < SH.chan.size(); ++q) {
                    const int k = SH.chan[q], t = (int)q / 16, row = (int)q % 16;
                    if (k <= 0) continue;  // empty slot
                    c32[(int64_t)d * mt * 32 + t * 32 + row] = (float)c64[((int64_t)d * K + (k - 1)) * 2];
                    c32[(int64_t)d * mt * 32 + t * 32 + 16 + row] = (float)c64[((int64_t)d * K + (k - 1)) * 2 + 1];
                }
        }
}

// Writes the MFMA A fragments of set `set`; returns the largest BF16 split
// residual |w - hi - lo| of its weights.  32x32x16 layout: lane l holds
// A[row l % 32][entry 8 (l / 32) + i] (rows 0..15 Re, 16..31 Im of the tile's
// channel slots).  The last tile of the small-tile class (unc_small_last_tile)
// is a 16x16x32 operand, lane l holding A[m = l % 16][k = 8 (l / 16) + i]: its
// B operand is the 32x32 one read as 16x16x32, k = 8 g + i (g = l / 16) being
// entry 8 (g / 2) + i of leaf 16 (g % 2) + n, so row m = 4 G + c (G = m / 4)
// holds, for leaves 16 (G % 2) + n only (k groups g with g % 2 == G % 2), the
// Re (c even) or Im (c odd) weights of channel slot 4 (G / 2) + c / 2.  Output
// lane l then holds leaf l % 32's two channels of slots 4 (l / 32) + {0, 1}.
double build_frags(const Shared& SH, int set, int L, int r, int N, const std::vector<double>& cs,
                   const std::vector<double>& sn) {
    const std::vector<int>& sp = SH.sets[set];
    const int mt = SH.mt(), nent = (int)sp.size() * r, nbs = std::max(1, (nent + 15) / 16);
    const bool small = unc_small_last_tile(nbs, mt, SH.lr);
    uint4* out = SH.frags + SH.set_off[set];
    double rho = 0;
    for (int t = 0; t < mt; ++t)
        for (int st = 0; st < 2 * nbs; ++st)
            for (int lane = 0; lane < 64; ++lane) {
                uint16_t hv[8];
                for (int i8 = 0; i8 < 8; ++i8) {
                    int e, slot;
                    bool im;
                    if (small && t == mt - 1) {
                        const int m = lane % 16, G = m / 4, c = m % 4, g = lane / 16;
                        e = (G % 2 == g % 2) ? 8 * (g / 2) + i8 : nent;  // other half of the leaves: zero
                        slot = 4 * (G / 2) + c / 2;
                        im = (c % 2) != 0;
                    } else {
                        const int row = lane % 32;
                        e = 16 * (st % nbs) + 8 * (lane / 32) + i8;
                        slot = row % 16;
                        im = row >= 16;
                    }
                    const int part = st / nbs;
                    const size_t qs = 16 * t + slot;
                    uint16_t v = 0;
                    if (e < nent && qs < SH.chan.size() && SH.chan[qs] > 0) {
                        const int k = SH.chan[qs];
                        const int j = sp[e / r], i = e % r;
                        const int p = j + L * i;
                        const int mm = (int)(((int64_t)k * p) % N);
                        const double w = im ? sn[mm] : cs[mm];
                        const uint16_t hi = bf16_rne_u((float)w);
                        const double r1 = w - bf16_d(hi);
                        const uint16_t lo = bf16_rne_u((float)r1);
                        rho = std::max(rho, std::fabs(r1 - bf16_d(lo)));
                        v = part ? lo : hi;
                    }
                    hv[i8] = v;
                }
                uint4 q;
                q.x = hv[0] | ((uint32_t)hv[1] << 16);
                q.y = hv[2] | ((uint32_t)hv[3] << 16);
                q.z = hv[4] | ((uint32_t)hv[5] << 16);
                q.w = hv[6] | ((uint32_t)hv[7] << 16);
                *out++ = q;
            }
    return rho;
}

// Where one side's rows go in the batch tables (filled by add_side on host threads).
struct SideSlot {
    int sid = 0;
    bool pm1 = false;
    int64_t item_base = 0, radix_off = 0, tup_off = 0, pm1_off = -1, pos_off = 0, olist_off = 0, fix_off = 0;
};

// Whether every tuple entry of the side is +-1 (words then carry +1 bit masks).
bool side_pm1(const SideHost& S) {
    if (S.N > 64) return false;
    for (int j = 0; j < S.L; ++j)
        for (int d = 0; d < S.R[j]; ++d)
            for (int i = 0; i < S.r; ++i)
                if (S.tuple(j, d)[i] != 1 && S.tuple(j, d)[i] != -1) return false;
    return true;
}

void add_side(Tables& T, const SideSlot& Z, const Shared& SH, const SideHost& S, const Split& X, int rmax,
              int order, int paf, int group, bool fp32, int64_t mid_off, int64_t suf_off, int set) {
    UncSide D{};
    D.L = S.L; D.r = S.r; D.N = S.N; D.K = S.N / 2; D.seqflag = S.seqflag; D.rmax = rmax;
    D.group = group;
    D.role = S.seqflag ? 1 : 0;
    D.nch = (int32_t)SH.chan.size();
    const int L = S.L, r = S.r, N = S.N;
    const int64_t ps = X.ps, pm = X.pm, po = X.po;
    D.nout = (int32_t)X.out.size();
    D.nent = X.nent;
    D.p_suf = ps; D.p_mid = pm; D.p_out = po; D.leaves = (int64_t)S.leaves;
    D.nbs = std::max(1, (X.nent + 15) / 16);
    D.ntiles = (int)((ps + 31) / 32);
    D.item_base = Z.item_base;
    D.olist_off = Z.olist_off;
    for (size_t x = 0; x < X.out.size(); ++x) T.olist[Z.olist_off + x] = X.out[x];
    // single-split positions: one constant, FP64 sum rounded once to FP32
    {
        const int K = N / 2;
        double acc[2 * kMaxK] = {};
        for (int j = 0; j < L; ++j)
            if (S.R[j] == 1) {
                const double* c = &SH.c64[SH.off64[SH.id(j, S.orig[j])]];
                for (int q = 0; q < 2 * K; ++q) acc[q] += c[q];
            }
        D.fix_off = Z.fix_off;
        for (size_t q = 0; q < SH.chan.size(); ++q) {
            const int k = SH.chan[q], t = (int)q / 16, row = (int)q % 16;
            if (k <= 0) continue;  // empty slot
            T.fix32[D.fix_off + t * 32 + row] = (float)acc[2 * (k - 1)];
            T.fix32[D.fix_off + t * 32 + 16 + row] = (float)acc[2 * (k - 1) + 1];
        }
    }
    D.radix_off = Z.radix_off;
    for (int j = 0; j < L; ++j) T.radix[Z.radix_off + j] = S.R[j];
    D.tup_off = Z.tup_off;
    for (int j = 0; j < L; ++j)
        for (int d = 0; d < S.R[j]; ++d)
            for (int i = 0; i < r; ++i) T.tup[D.tup_off + ((int64_t)j * rmax + d) * r + i] = S.tuple(j, d)[i];
    D.pm1_off = Z.pm1 ? Z.pm1_off : -1;
    if (Z.pm1)
        for (int j = 0; j < L; ++j)
            for (int d = 0; d < S.R[j]; ++d) {
                uint64_t msk = 0;
                for (int i = 0; i < r; ++i)
                    if (S.tuple(j, d)[i] == 1) msk |= 1ull << (j + L * i);
                T.pm1[D.pm1_off + (int64_t)j * rmax + d] = msk;
            }
    // per-position contribution blocks (shared between sides)
    D.c32_off = Z.pos_off;
    D.c64_off = Z.pos_off;
    for (int j = 0; j < L; ++j) {
        const int id = SH.id(j, S.orig[j]);
        T.pos32[Z.pos_off + j] = SH.off32[id];
        T.pos64[Z.pos_off + j] = SH.off64[id];
    }
    D.mid_off = mid_off;
    D.sufb_off = fp32 ? 0 : suf_off;
    D.ent_off = fp32 ? suf_off : 0;
    D.w32_off = fp32 ? SH.set_off[set] : 0;
    // A fragments (suffix DFT weights, BF16 hi and lo), shared per suffix position set
    D.frag_off = fp32 ? 0 : SH.set_off[set];
    const double rho = fp32 ? 0.0 : SH.set_rho[set];
    bool in_suf[kMaxN] = {};
    for (int j : X.suf) in_suf[j] = true;
    // error bound of the filter's Re/Im
    double s1_pre = 0, s1_suf = 0;
    for (int j = 0; j < L; ++j) {
        double mxabs = 0;
        for (int d = 0; d < S.R[j]; ++d) {
            double ab = 0;
            for (int i = 0; i < r; ++i) ab += std::abs(S.tuple(j, d)[i]);
            mxabs = std::max(mxabs, ab);
        }
        (in_suf[j] ? s1_suf : s1_pre) += mxabs;
    }
    const double u24 = 0x1p-24, u23 = 0x1p-23;
    double E;
    if (!fp32)
        E = (D.nout + 4) * u24 * s1_pre * 1.01                              // fixed + X_outer + Mid rounding (FP32)
            + s1_suf * rho * 1.0000001                                      // BF16 split residual
            + 1.004 * (32.0 * D.nbs + 1.0) * u23 * (s1_pre + s1_suf);       // MFMA accumulation
    else
        E = (D.nout + 4) * u24 * s1_pre * 1.01 + s1_suf * u24 * 1.01         // FP32 weights
            + (X.nent + 1.0) * u24 * (s1_pre + s1_suf) * 1.01;              // FP32 FMAs
    const double thr = 2.0 * order - paf + 0.001;
    const double eps = psd_eps(E, thr);
    D.thr = thr;
    D.thr_hi = std::nextafter((float)(thr + eps), INFINITY);
    D.thr_lo = std::nextafter((float)(thr - eps), -INFINITY);
    T.sides[Z.sid] = D;
}

template <class V>
V* upload(gp_ctx* ctx, const char* name, const std::vector<V>& v) {
    void* p = ctx->buf(name).ensure(v.size() * sizeof(V) + 64);
    if (!v.empty()) GP_HIP(hipMemcpyAsync(p, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice, ctx->stream));
    return static_cast<V*>(p);
}

}  // namespace

// ---------------------------------------------------------------------------
// host: the three stages

// Sides of one operand class of the split filter kernel, contiguous in table order.
struct UncClass {
    int nbs = 0, mt = 0;
    int lr = 8;  // registers per lane read in the last M-tile (8, 4 or 2)
    int sid_lo = 0, sid_hi = 0;
    int64_t item_lo = 0, item_hi = 0;
};

// Host tables of one batch (built without the device, so that the next batch's
// tables can be prepared while the current batch runs).
struct UncPrep {
    int N = 0, bound = 0, rmax = 1;
    bool fp32 = false;
    bool all_pm1 = false;  // every live side's tuple entries are +-1
    int64_t trig_off = 0;
    Tables T;
    Shared SH;
    std::vector<int> side_of;
    std::vector<UncClass> classes;
    uint64_t total_leaves = 0;
    uint64_t dead_leaves = 0;  // leaves of sides rejected by their fixed channels
    int64_t n_suf_jobs = 0, n_sufb = 0;  // split path: suffix operand jobs and their size (uint4)
    // every table in one staging buffer (one host-to-device copy): pinned
    // (`stage`) or, when prepared without a pool, `host` until unc_stage
    PinnedBuf* stage = nullptr;
    PinnedPool* stage_pool = nullptr;
    std::unique_ptr<char[]> host;
    char* base = nullptr;
    size_t stage_bytes = 0;
    std::vector<size_t> stage_off;
    ~UncPrep() {
        if (stage_pool) stage_pool->release(stage);
    }
};

namespace {

// Staging buffer regions, in upload order (sizes in bytes).
enum StageRegion {
    kStSides, kStRadix, kStTup, kStPm1, kStC64, kStPos64, kStPos32, kStOlist, kStFix32, kStC32,
    kStMid, kStSufJobs, kStSufPos, kStFrags, kStEnt, kStW32, kStToff, kStCount
};

}  // namespace

void unc_stage(UncPrep& U, PinnedPool& pool) {
    if (U.stage || U.T.nsides == 0) return;
    U.stage = pool.acquire(U.stage_bytes);
    U.stage_pool = &pool;
    std::memcpy(U.stage->p, U.base, U.stage_bytes);
    U.base = static_cast<char*>(U.stage->p);
    U.host.reset();
}

template <int NBS, int MT, int LR>
void launch_filter_t(const UncClass& c, int64_t rounds, int cls, int nsub, dim3 grid, hipStream_t st,
                     const UncTabs& tb, UncCtl* ctl, uint64_t* surv, int64_t scap, uint64_t* band, int64_t bcap) {
    hipLaunchKernelGGL((k_unc_filter<NBS, MT, LR>), grid, dim3(256), 0, st, tb, c.sid_lo, c.sid_hi, c.item_lo,
                       c.item_hi, rounds, cls, nsub, ctl, surv, scap, band, bcap);
}

void launch_filter_class(const UncClass& c, int64_t rounds, int cls, int nsub, dim3 grid, hipStream_t st,
                         const UncTabs& tb, UncCtl* ctl, uint64_t* surv, int64_t scap, uint64_t* band, int64_t bcap) {
#define GP_FILTER_CASE(NB, M, LR)                \
    if (c.nbs == NB && c.mt == M && c.lr == LR) \
        return launch_filter_t<NB, M, LR>(c, rounds, cls, nsub, grid, st, tb, ctl, surv, scap, band, bcap);
#define GP_FILTER_LR(NB, M) GP_FILTER_CASE(NB, M, 8) GP_FILTER_CASE(NB, M, 4) GP_FILTER_CASE(NB, M, 2)
    GP_FILTER_LR(1, 1) GP_FILTER_LR(2, 1) GP_FILTER_LR(3, 1) GP_FILTER_LR(4, 1)
    GP_FILTER_LR(1, 2) GP_FILTER_LR(2, 2) GP_FILTER_LR(3, 2) GP_FILTER_LR(4, 2)
#undef GP_FILTER_LR
#undef GP_FILTER_CASE
    throw ArgError("uncompress: no filter kernel for this operand class");
}

UncTrig unc_trig(gp_ctx* ctx, int N) {
    if (N > kMaxN || N < 2) throw ArgError("uncompress: expanded length outside [2, 64]");
    const int table = ctx->tables.get(N);
    ctx->tables.upload(ctx->stream);
    const PsdTableDesc& td = ctx->tables.desc[table];
    UncTrig t;
    t.N = N;
    t.trig_off = td.trig_off;
    t.cs.assign(ctx->tables.trig64.begin() + td.trig_off, ctx->tables.trig64.begin() + td.trig_off + N);
    t.sn.assign(ctx->tables.trig64.begin() + td.trig_off + N, ctx->tables.trig64.begin() + td.trig_off + 2 * N);
    return t;
}

std::shared_ptr<UncPrep> unc_prepare(std::span<const SideHost> sides, std::span<const int> group, int order, int paf,
                                     bool fp32, const UncTrig& trig, PinnedPool* pinned,
                                     const std::function<void()>& before_fill) {
    auto prep = std::make_shared<UncPrep>();
    UncPrep& U = *prep;
    if (sides.empty()) return prep;
    const int N = sides[0].N;
    if (N != trig.N) throw ArgError("uncompress: twiddle table length mismatch");
    if ((int64_t)sides.size() >= (1ll << 23)) throw ArgError("uncompress: too many sequences in one batch");
    const int bound = 2 * order - paf;
    if (bound < 0 || bound > 255) throw ArgError("uncompress: keys need 0 <= 2*order - paf <= 255");
    int rmax = 1;
    for (const auto& s : sides)
        for (int j = 0; j < s.L; ++j) rmax = std::max(rmax, (int)s.R[j]);
    U.N = N;
    U.bound = bound;
    U.rmax = rmax;
    U.fp32 = fp32;
    U.trig_off = trig.trig_off;
    const std::vector<double>& cs = trig.cs;
    const std::vector<double>& sn = trig.sn;
    Tables& T = U.T;
    Shared& SH = U.SH;
    uint64_t& total_leaves = U.total_leaves;
    std::vector<int>& side_of = U.side_of;
    {
        Phase ph("unc_filter: side tables");
        // Channels k = r m are fixed by the compressed word c: the expanded word x
        // has x[j + L i] = t_j[i] with sum_i t_j[i] = c_j, so X(r m) = sum_j c_j
        // e^{-2 pi i m j / L} for every leaf.  A side whose fixed channels all
        // lie clearly below the threshold needs only the other channels per leaf;
        // one clearly above it rejects every leaf; a value within 1e-6 of the
        // threshold keeps every channel in the per-leaf filter (whole batch).
        const int r = sides[0].r, K = N / 2;
        const double thr = 2.0 * order - paf + 0.001;
        bool all_channels = r <= 1 || K < r;
        auto parallel = [](size_t count, const std::function<void(size_t)>& f) { pool().parallel_for(count, f); };
        // per side: bit 0 = dead (a fixed channel above the threshold), bit 1 = a
        // fixed channel within 1e-6 of it
        std::vector<uint8_t> fixed_state(sides.size(), 0);
        if (r > 1)
            parallel(sides.size(), [&](size_t i) {
                if (sides[i].leaves == 0) return;
                uint8_t st = 0;
                for (int k = r; k <= K; k += r) {
                    double re = 0, im = 0;
                    int mm = 0;  // k j mod N
                    for (int j = 0; j < sides[i].L; ++j) {
                        re += sides[i].orig[j] * cs[mm];
                        im += sides[i].orig[j] * sn[mm];
                        mm += k;
                        if (mm >= N) mm -= N;
                    }
                    const double psd = re * re + im * im;
                    if (psd > thr + 1e-6) st |= 1;
                    else if (psd >= thr - 1e-6) st |= 2;
                }
                fixed_state[i] = st;
            });
        std::vector<size_t> live;
        for (size_t i = 0; i < sides.size(); ++i) {
            if (sides[i].leaves == 0) continue;
            if (fixed_state[i] & 2) all_channels = true;
            if (fixed_state[i] & 1) {
                U.dead_leaves += sides[i].leaves;
                continue;
            }
            live.push_back(i);
        }
        if (all_channels) U.dead_leaves = 0, live.clear();
        if (all_channels)
            for (size_t i = 0; i < sides.size(); ++i)
                if (sides[i].leaves != 0) live.push_back(i);
        // filter channels by slot: full M-tiles in order; in the last M-tile its
        // c real channels go to rows (i mod lr) + 4 (i / lr), lr the fewest
        // registers per lane (2, 4 or 8) that hold them, and the other slots hold
        // 0 (no channel: zero weights, never read)
        std::vector<int> real;
        for (int k = 1; k <= K; ++k)
            if (all_channels || k % r != 0) real.push_back(k);
        const int mtr = ((int)real.size() + 15) / 16;
        const int last = (int)real.size() - 16 * (mtr - 1);
        SH.lr = last <= 4 ? 2 : last <= 8 ? 4 : 8;
        SH.chan.assign(real.begin(), real.begin() + 16 * (mtr - 1));
        if (SH.lr == 8) {
            SH.chan.insert(SH.chan.end(), real.begin() + 16 * (mtr - 1), real.end());
        } else {
            std::vector<int> tile(8, 0);
            for (int i = 0; i < last; ++i) tile[(i % SH.lr) + 4 * (i / SH.lr)] = real[16 * (mtr - 1) + i];
            while (!tile.empty() && tile.back() == 0) tile.pop_back();
            SH.chan.insert(SH.chan.end(), tile.begin(), tile.end());
        }
        // shared per-(position, letter) blocks; per side its split and table keys
        Phase ph1("tables: splits + keys");
        if (!live.empty()) add_blocks(SH, *sides[live[0]].tab, sides[live[0]].L, N, cs, sn);
        std::vector<Split> spl(live.size());
        std::vector<std::string> kmid(live.size()), ksuf(live.size());
        std::vector<char> pm1(live.size());
        parallel(live.size(), [&](size_t q) {
            const SideHost& S = sides[live[q]];
            spl[q] = side_split(S);
            kmid[q] = pos_key(S, spl[q].mid);
            ksuf[q] = pos_key(S, spl[q].suf);
            pm1[q] = side_pm1(S);
        });
        U.all_pm1 = std::all_of(pm1.begin(), pm1.end(), [](char b) { return b != 0; });
        ph1.next("tables: layout");
        // content-addressed suffix weight sets (key: the suffix positions, the
        // first half of the suffix key), mid and suffix tables (built below, in
        // parallel): one per distinct key
        std::unordered_map<std::string_view, int> set_of, mid_job, suf_job;
        set_of.reserve(64);
        mid_job.reserve(live.size());
        suf_job.reserve(live.size());
        std::vector<size_t> mid_rep, suf_rep;  // representative side (index into live) per job
        std::vector<int> side_mid(live.size()), side_suf(live.size()), side_set(live.size());
        for (size_t q = 0; q < live.size(); ++q) {
            const std::string_view kpos(ksuf[q].data(), ksuf[q].size() / 2);
            auto it = set_of.find(kpos);
            if (it == set_of.end()) it = set_of.emplace(kpos, add_set(SH, spl[q].suf, sides[live[q]].r, fp32)).first;
            side_set[q] = it->second;
            auto im = mid_job.emplace(kmid[q], (int)mid_rep.size());
            if (im.second) mid_rep.push_back(q);
            side_mid[q] = im.first->second;
            auto is = suf_job.emplace(ksuf[q], (int)suf_rep.size());
            if (is.second) suf_rep.push_back(q);
            side_suf[q] = is.first->second;
        }
        if (std::getenv("GOLAY_PROFILE"))
            std::fprintf(stderr, "[golay] prep: %zu sides, %zu mid tables, %zu suffix tables, %zu fragment sets\n",
                         live.size(), mid_rep.size(), suf_rep.size(), SH.sets.size());
        std::vector<int64_t> mid_off(mid_rep.size()), suf_off(suf_rep.size());
        for (size_t i = 0; i < mid_rep.size(); ++i) {
            mid_off[i] = SH.n_mid;
            SH.n_mid += mid_size(spl[mid_rep[i]], SH.mt());
        }
        for (size_t i = 0; i < suf_rep.size(); ++i) {
            suf_off[i] = SH.n_suf;
            SH.n_suf += suf_size(spl[suf_rep[i]], fp32);
        }
        // table order: sides grouped by operand class (one kernel instantiation each)
        const int mtc = SH.mt();
        auto cls = [&](size_t q) { return std::max(1, (spl[q].nent + 15) / 16); };
        std::vector<size_t> torder(live.size());
        std::iota(torder.begin(), torder.end(), 0);
        // within a class, sides by decreasing cost of one work item (mid rows x suffix tile
        // pairs): the kernel hands the items out in this order to the first free wave
        auto item_cost = [&](size_t q) { return spl[q].pm * ((spl[q].ps + 63) / 64); };
        std::stable_sort(torder.begin(), torder.end(), [&](size_t x, size_t y) {
            if (cls(x) != cls(y)) return cls(x) < cls(y);
            return item_cost(x) > item_cost(y);
        });
        std::vector<SideSlot> slot(live.size());
        int64_t n_radix = 0, n_tup = 0, n_pm1 = 0, n_pos = 0, n_olist = 0, n_fix = 0;
        for (size_t x = 0; x < torder.size(); ++x) {
            const size_t q = torder[x];
            const SideHost& S = sides[live[q]];
            const int nbs = cls(q);
            if (U.classes.empty() || U.classes.back().nbs != nbs) {
                UncClass c;
                c.nbs = nbs;
                c.mt = mtc;
                c.lr = SH.lr;
                c.sid_lo = c.sid_hi = (int)x;
                c.item_lo = c.item_hi = T.items;
                U.classes.push_back(c);
            }
            SideSlot& Z = slot[q];
            Z.sid = (int)x;
            Z.item_base = T.items;
            T.items += spl[q].po;
            Z.radix_off = n_radix;
            n_radix += S.L;
            Z.tup_off = n_tup;
            n_tup += (int64_t)S.L * rmax * S.r;
            Z.pm1 = pm1[q];
            if (Z.pm1) {
                Z.pm1_off = n_pm1;
                n_pm1 += (int64_t)S.L * rmax;
            }
            Z.pos_off = n_pos;
            n_pos += S.L;
            Z.olist_off = n_olist;
            n_olist += (int64_t)spl[q].out.size();
            Z.fix_off = n_fix;
            n_fix += (int64_t)mtc * 32;
            side_of.push_back((int)live[q]);
            total_leaves += S.leaves;
            U.classes.back().sid_hi = (int)x + 1;
            U.classes.back().item_hi = T.items;
        }
        T.nsides = (int64_t)live.size();
        if (T.nsides == 0) return prep;
        // one staging buffer, regions 256-byte aligned
        size_t bytes[kStCount] = {};
        bytes[kStSides] = T.nsides * sizeof(UncSide);
        bytes[kStRadix] = n_radix * sizeof(int32_t);
        bytes[kStTup] = n_tup;
        bytes[kStPm1] = n_pm1 * sizeof(uint64_t);
        bytes[kStC64] = SH.c64.size() * sizeof(double);
        bytes[kStPos64] = n_pos * sizeof(int32_t);
        bytes[kStPos32] = n_pos * sizeof(int32_t);
        bytes[kStOlist] = n_olist * sizeof(int32_t);
        bytes[kStFix32] = n_fix * sizeof(float);
        bytes[kStC32] = SH.c32.size() * sizeof(float);
        bytes[kStMid] = SH.n_mid * sizeof(float);
        // split path: suffix operands are built on the device from a job list
        int64_t n_sufpos = 0;
        if (!fp32)
            for (size_t i = 0; i < suf_rep.size(); ++i) n_sufpos += (int64_t)spl[suf_rep[i]].suf.size();
        bytes[kStSufJobs] = fp32 ? 0 : suf_rep.size() * sizeof(SufJob);
        bytes[kStSufPos] = n_sufpos * sizeof(int32_t);
        bytes[kStFrags] = SH.n_frags * sizeof(uint4);
        bytes[kStEnt] = fp32 ? SH.n_suf : 0;
        bytes[kStW32] = SH.n_w32 * sizeof(float);
        bytes[kStToff] = sizeof(int64_t);
        U.stage_off.assign(kStCount, 0);
        size_t total = 0;
        for (int i = 0; i < kStCount; ++i) {
            U.stage_off[i] = total;
            total += (bytes[i] + 255) & ~(size_t)255;
        }
        U.stage_bytes = total;
        if (std::getenv("GOLAY_PROFILE"))
            std::fprintf(stderr, "[golay] staging: %.2f MB (suffix operands %.2f MB, mid %.2f MB, sides %.2f MB)\n",
                         total / 1e6, bytes[kStEnt] / 1e6, bytes[kStMid] / 1e6,
                         bytes[kStSides] / 1e6);
        if (pinned) {
            U.stage = pinned->acquire(total);
            U.stage_pool = pinned;
            U.base = static_cast<char*>(U.stage->p);
        } else {
            U.host.reset(new char[total]);
            U.base = U.host.get();
        }
        auto at = [&](int i) { return static_cast<void*>(U.base + U.stage_off[i]); };
        T.sides = static_cast<UncSide*>(at(kStSides));
        T.radix = static_cast<int32_t*>(at(kStRadix));
        T.tup = static_cast<int8_t*>(at(kStTup));
        T.pm1 = static_cast<uint64_t*>(at(kStPm1));
        T.pos64 = static_cast<int32_t*>(at(kStPos64));
        T.pos32 = static_cast<int32_t*>(at(kStPos32));
        T.olist = static_cast<int32_t*>(at(kStOlist));
        T.fix32 = static_cast<float*>(at(kStFix32));
        SH.mid = static_cast<float*>(at(kStMid));
        U.n_suf_jobs = fp32 ? 0 : (int64_t)suf_rep.size();
        U.n_sufb = fp32 ? 0 : SH.n_suf;
        if (!fp32) {
            auto* jobs = static_cast<SufJob*>(at(kStSufJobs));
            auto* spos = static_cast<int32_t*>(at(kStSufPos));
            int64_t po = 0;
            for (size_t i = 0; i < suf_rep.size(); ++i) {
                const Split& X = spl[suf_rep[i]];
                jobs[i] = SufJob{suf_off[i], slot[suf_rep[i]].sid, (int32_t)X.suf.size(), (int32_t)po, 0};
                for (int j : X.suf) spos[po++] = j;
            }
        }
        SH.frags = static_cast<uint4*>(at(kStFrags));
        SH.ent = static_cast<int8_t*>(at(kStEnt));
        SH.w32 = static_cast<float*>(at(kStW32));
        std::memcpy(at(kStC64), SH.c64.data(), bytes[kStC64]);
        std::memcpy(at(kStC32), SH.c32.data(), bytes[kStC32]);
        std::memcpy(at(kStToff), &U.trig_off, sizeof(int64_t));
        // rows that add_side leaves unwritten: tuples and masks d >= R[j] (never
        // read) and the fixed-sum padding channels (read as 0)
        std::memset(T.tup, 0, bytes[kStTup]);
        std::memset(T.pm1, 0, bytes[kStPm1]);
        std::memset(T.fix32, 0, bytes[kStFix32]);
        if (before_fill) {
            ph1.next("tables: wait");
            before_fill();
        }
        ph1.next("tables: fill");
        // every table built in place on host threads: weight sets, mid and suffix
        // tables, then the per-side rows (which need the sets' BF16 residuals)
        const size_t nset = SH.sets.size();
        const int r0 = sides[live[0]].r, L0 = sides[live[0]].L;
        auto fill_one = [&](size_t i) {
            if (i < nset) {
                if (fp32) build_w32(SH, (int)i, L0, r0, N, cs, sn);
                else SH.set_rho[i] = build_frags(SH, (int)i, L0, r0, N, cs, sn);
            } else if (i < nset + mid_rep.size()) {
                const size_t k = i - nset, q = mid_rep[k];
                build_mid(SH, sides[live[q]], spl[q].mid, SH.mid + mid_off[k]);
            } else {
                const size_t k = i - nset - mid_rep.size(), q = suf_rep[k];
                build_suf(sides[live[q]], spl[q].suf, true, nullptr, SH.ent + suf_off[k]);
            }
        };
        parallel(nset + mid_rep.size() + (fp32 ? suf_rep.size() : 0), fill_one);
        parallel(live.size(), [&](size_t q) {
            add_side(T, slot[q], SH, sides[live[q]], spl[q], rmax, order, paf, group[live[q]], fp32,
                     mid_off[side_mid[q]], suf_off[side_suf[q]], side_set[q]);
        });
    }
    return prep;
}

namespace {

void batch_header(const UncPrep& U, UncBatch& B) {
    B.N = U.N;
    B.K = U.N / 2;
    B.trig_off = U.trig_off;
    B.bound = U.bound;
    B.side_of = U.side_of;
    B.nsides_table = (int)U.T.nsides;
    B.all_pm1 = U.all_pm1;
}

// Survivor and band list capacities for a batch of `leaves` leaves.
int64_t surv_cap(uint64_t leaves) { return (int64_t)std::min<uint64_t>(leaves * 3 / 5 + (1 << 20), (uint64_t)3 << 30); }
int64_t band_cap(uint64_t leaves) { return (int64_t)std::min<uint64_t>(leaves / 64 + (1 << 16), (uint64_t)1 << 28); }

// Uploads the staged tables into the device buffer `name` (one copy: on the
// context's stream, or, with `uploaded`, on the copy stream, which the
// context's stream then waits for, so that the copy overlaps earlier kernels),
// points B's table fields there and returns the filter kernels' view of them.
UncTabs upload_tables(gp_ctx* ctx, UncPrep& U, const std::string& name, UncBatch& B, const int64_t** d_toff,
                      hipEvent_t uploaded = nullptr) {
    unc_stage(U, ctx->pinned);
    char* base = static_cast<char*>(ctx->buf(name).ensure(U.stage_bytes));
    if (uploaded) {
        GP_HIP(hipMemcpyAsync(base, U.base, U.stage_bytes, hipMemcpyHostToDevice, ctx->copy_stream));
        GP_HIP(hipEventRecord(uploaded, ctx->copy_stream));
        GP_HIP(hipStreamWaitEvent(ctx->stream, uploaded, 0));
        if (U.stage) U.stage->record_upload(ctx->copy_stream);
    } else {
        GP_HIP(hipMemcpyAsync(base, U.base, U.stage_bytes, hipMemcpyHostToDevice, ctx->stream));
        if (U.stage) U.stage->record_upload(ctx->stream);
    }
    auto at = [&](int i) { return static_cast<void*>(base + U.stage_off[i]); };
    B.d_sides = static_cast<const UncSide*>(at(kStSides));
    B.d_radix = static_cast<const int32_t*>(at(kStRadix));
    B.d_tup = static_cast<const int8_t*>(at(kStTup));
    B.d_pm1 = static_cast<const uint64_t*>(at(kStPm1));
    B.d_c64 = static_cast<const double*>(at(kStC64));
    B.d_pos64 = static_cast<const int32_t*>(at(kStPos64));
    *d_toff = static_cast<const int64_t*>(at(kStToff));
    // split path: the suffix operands, built on the device from the job list
    uint4* sufb = nullptr;
    if (U.n_suf_jobs > 0) {
        sufb = static_cast<uint4*>(ctx->buf(name + ".sufb").ensure((size_t)U.n_sufb * sizeof(uint4)));
        hipLaunchKernelGGL(k_build_suf, dim3((unsigned)U.n_suf_jobs), dim3(256), 0, ctx->stream,
                           static_cast<const SufJob*>(at(kStSufJobs)), static_cast<const int32_t*>(at(kStSufPos)),
                           B.d_sides, B.d_radix, B.d_tup, sufb);
        GP_HIP(hipGetLastError());
    }
    return UncTabs{B.d_sides,
                   B.d_radix,
                   static_cast<const int32_t*>(at(kStPos32)),
                   static_cast<const int32_t*>(at(kStOlist)),
                   static_cast<const float*>(at(kStC32)),
                   static_cast<const float*>(at(kStMid)),
                   static_cast<const float*>(at(kStFix32)),
                   sufb,
                   static_cast<const uint4*>(at(kStFrags)),
                   static_cast<const int8_t*>(at(kStEnt)),
                   static_cast<const float*>(at(kStW32))};
}

// The filter kernels of one batch on the context's stream.
void launch_filters(gp_ctx* ctx, const UncPrep& U, const UncTabs& tb, UncCtl* ctl, uint64_t* d_surv, int64_t scap,
                    uint64_t* d_band, int64_t bcap) {
    hipStream_t st = ctx->stream;
    if (!U.fp32) {
        if (U.classes.size() > 4) throw ArgError("uncompress: more than 4 operand classes in a batch");
        for (size_t ci = 0; ci < U.classes.size(); ++ci) {
            const UncClass& c = U.classes[ci];
            // as many waves as the class keeps resident (its occupancy hint, 4 SIMDs per CU)
            const int cap = unc_waves_per_simd(c.nbs, c.mt, c.lr);
            const int per_simd = (int)std::clamp<int64_t>(tuned(ctx, "unc_waves/" + std::to_string(U.N), cap), 1, cap);
            const int64_t waves = std::min<int64_t>(c.item_hi - c.item_lo, (int64_t)ctx->num_cus * 4 * per_simd);
            const dim3 grid((unsigned)std::max<int64_t>(1, (waves + 3) / 4));
            // items by wave index for all but the last `tail` rounds, which go out through
            // the work counters (by default every item goes through them)
            const int64_t W = (int64_t)grid.x * 4;
            const int64_t tail = tuned(ctx, "unc_counter_rounds/" + std::to_string(U.N), (int64_t)1 << 40);
            const int64_t rounds = std::max<int64_t>(0, (c.item_hi - c.item_lo) / W - tail);
            const int64_t nsub = std::clamp<int64_t>(tuned(ctx, "unc_counter_streams/" + std::to_string(U.N),
                                                           kSubCounters), 1, std::min<int64_t>(kSubCounters, grid.x));
            if (std::getenv("GOLAY_PROFILE")) {
                // leaf-tiles per wave as the kernel hands items out (the fixed rounds, then
                // each counter stream's items to whichever of its waves is free first):
                // mean and heaviest wave
                std::vector<double> cost;
                for (int sd = c.sid_lo; sd < c.sid_hi; ++sd) {
                    const UncSide& S = U.T.sides[sd];
                    cost.insert(cost.end(), (size_t)S.p_out, (double)S.p_mid * (double)(2 * ((S.ntiles + 1) / 2)));
                }
                std::vector<double> fixed((size_t)W, 0.0);
                for (int64_t k = 0; k < rounds; ++k)
                    for (int64_t w = 0; w < W; ++w) fixed[(size_t)w] += cost[(size_t)(k * W + ((k & 1) ? W - 1 - w : w))];
                double total = 0, mx = 0;
                for (size_t k = 0; k < cost.size(); ++k) total += cost[k];
                for (int64_t j = 0; j < nsub; ++j) {
                    std::priority_queue<double, std::vector<double>, std::greater<double>> load;
                    for (int64_t w = 0; w < W; ++w)
                        if ((w / 4) % nsub == j) load.push(fixed[(size_t)w]);
                    for (int64_t k = rounds * W + j; k < (int64_t)cost.size(); k += nsub) {
                        const double l = cost[(size_t)k] + load.top();
                        load.pop();
                        load.push(l);
                    }
                    while (!load.empty()) {
                        mx = std::max(mx, load.top());
                        load.pop();
                    }
                }
                std::fprintf(stderr, "[golay] filter class <%d,%d,%d>: %lld items on %lld waves, tiles per wave mean "
                             "%.0f, max %.0f (x%.3f)\n", c.nbs, c.mt, c.lr, (long long)(c.item_hi - c.item_lo),
                             (long long)W, total / (double)W, mx, mx * (double)W / std::max(total, 1.0));
            }
            launch_filter_class(c, rounds, (int)ci, (int)nsub, grid, st, tb, ctl, d_surv, scap, d_band, bcap);
        }
    } else {
        const int64_t waves = std::min<int64_t>(U.T.items, (int64_t)ctx->num_cus * 8);
        const dim3 grid((unsigned)std::max<int64_t>(1, (waves + 3) / 4));
        hipLaunchKernelGGL(k_unc_filter_fp32, grid, dim3(256), 0, st, tb, (int)U.T.nsides, U.T.items, ctl, d_surv,
                           scap, d_band, bcap);
    }
    GP_HIP(hipGetLastError());
}

void add_filter_stats(gp_stats* stats, const UncPrep& U, const UncCtl& h, float ms_filter, float ms_band) {
    if (!stats) return;
    stats->words += U.total_leaves;
    stats->psd_decisions += U.total_leaves;
    stats->psd_pass += h.surv;
    stats->band_rechecks += h.band;
    stats->tie_band += h.ties;
    stats->psd_kernel_ms += ms_filter;
    stats->band_kernel_ms += ms_band;
    stats->psd_launches += U.fp32 ? 1 : U.classes.size();
    stats->kernel_ms += ms_filter + ms_band;
}

void add_dead_stats(gp_stats* stats, const UncPrep& U) {
    if (!stats) return;  // sides rejected by their fixed channels: every leaf decided, none passed
    stats->words += U.dead_leaves;
    stats->psd_decisions += U.dead_leaves;
}

}  // namespace

UncBatch unc_launch(gp_ctx* ctx, UncPrep& U, gp_stats* stats) {
    UncBatch B;
    hipStream_t st = ctx->stream;
    batch_header(U, B);
    add_dead_stats(stats, U);
    if (U.T.nsides == 0) return B;
    Phase ph2("unc_filter: upload");
    const int64_t* d_toff = nullptr;
    const UncTabs tb = upload_tables(ctx, U, "u_stage", B, &d_toff);
    UncCtl* ctl = ctx->dev<UncCtl>("u_ctl", 1);
    int64_t scap = surv_cap(U.total_leaves);
    int64_t bcap = band_cap(U.total_leaves);
    UncCtl h{};
    float ms_filter = 0, ms_band = 0;
    ph2.next("unc_filter: kernels");
    for (int attempt = 0; attempt < 3; ++attempt) {
        uint64_t* d_surv = ctx->dev<uint64_t>("u_surv", (size_t)scap);
        uint64_t* d_band = ctx->dev<uint64_t>("u_band", (size_t)bcap);
        GP_HIP(hipMemsetAsync(ctl, 0, sizeof(UncCtl), st));
        GP_HIP(hipEventRecord(ctx->ev0, st));
        launch_filters(ctx, U, tb, ctl, d_surv, scap, d_band, bcap);
        GP_HIP(hipEventRecord(ctx->ev1, st));
        GP_HIP(hipMemcpyAsync(&h, ctl, sizeof h, hipMemcpyDeviceToHost, st));
        GP_HIP(hipStreamSynchronize(st));
        GP_HIP(hipEventElapsedTime(&ms_filter, ctx->ev0, ctx->ev1));
        if ((int64_t)h.band > bcap) { bcap = (int64_t)h.band + 1024; continue; }
        if (h.band) {
            hipLaunchKernelGGL(k_unc_band, dim3(ctx->num_cus * 2), dim3(256), 0, st, B.d_sides, B.d_radix, B.d_tup,
                               ctx->tables.d_trig64.as<double>(), d_toff, d_band, ctl, d_surv, scap, (int64_t)h.band,
                               nullptr, (int64_t)0);
            GP_HIP(hipGetLastError());
            GP_HIP(hipEventRecord(ctx->ev2, st));
            GP_HIP(hipMemcpyAsync(&h, ctl, sizeof h, hipMemcpyDeviceToHost, st));
            GP_HIP(hipStreamSynchronize(st));
            GP_HIP(hipEventElapsedTime(&ms_band, ctx->ev1, ctx->ev2));
        }
        if ((int64_t)h.surv > scap) { scap = (int64_t)h.surv + 1024; continue; }
        break;
    }
    B.nsurv = (int64_t)h.surv;
    B.d_surv = ctx->buf("u_surv").as<uint64_t>();
    add_filter_stats(stats, U, h, ms_filter, ms_band);
    return B;
}

// Stream capture of one batch's band recheck and re-match chain into the graph
// of its buffer slot.  The first batch of a slot instantiates the graph; later
// batches update it in place (hipGraphExecUpdate), or instantiate it again if
// the update is refused.  GOLAY_UNC_GRAPH=0 launches the kernels directly.
// A capture that ends by an exception is closed and dropped.
struct BatchGraph {
    gp_ctx* ctx;
    int slot;
    bool on;
    bool capturing = false;
    hipStream_t st = nullptr;
    BatchGraph(gp_ctx* c, int s) : ctx(c), slot(s) {
        const char* e = std::getenv("GOLAY_UNC_GRAPH");
        on = !(e && e[0] == '0');
    }
    void begin(hipStream_t s) {
        if (!on) return;
        st = s;
        GP_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        capturing = true;
    }
    void end_and_launch(hipStream_t s) {
        hipGraph_t g = nullptr;
        capturing = false;
        GP_HIP(hipStreamEndCapture(s, &g));
        hipGraphExec_t& ex = ctx->unc_graph[slot];
        bool ok = false;
        if (ex) {
            hipGraphNode_t err_node = nullptr;
            hipGraphExecUpdateResult res;
            ok = hipGraphExecUpdate(ex, g, &err_node, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
            if (!ok) {
                (void)hipGetLastError();
                (void)hipGraphExecDestroy(ex);
                ex = nullptr;
            }
        }
        if (!ok) {
            const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                (void)hipGraphDestroy(g);
                ex = nullptr;
                GP_HIP(e);
            }
            ++ctx->unc_graph_instantiations;
        }
        (void)hipGraphDestroy(g);
        GP_HIP(hipGraphLaunch(ex, s));
        ++ctx->unc_graph_launches;
    }
    ~BatchGraph() {
        if (capturing) {  // an exception left the capture open
            hipGraph_t g = nullptr;
            if (hipStreamEndCapture(st, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
            (void)hipGetLastError();
        }
    }
};

// A batch whose filter, band recheck and join are queued on the context's
// stream with device-side counts, in the buffers of one of two slots.
struct UncInflight {
    std::shared_ptr<UncPrep> prep;  // keeps the staged tables for a synchronous rerun
    int slot = 0;
    UncBatch B;
    int64_t scap = 0, bcap = 0, rcap = 0, ecap = 0, ccap = 0;
    UncCtl* ctl = nullptr;
    JoinCtl2* jctl = nullptr;
    unsigned int* ties = nullptr;
    int8_t* w = nullptr;
    uint8_t* k = nullptr;
    int32_t* grp = nullptr;
};

std::shared_ptr<UncInflight> unc_enqueue(gp_ctx* ctx, std::shared_ptr<UncPrep> prep, int slot, int paf_target) {
    auto F = std::make_shared<UncInflight>();
    F->prep = std::move(prep);
    F->slot = slot;
    UncPrep& U = *F->prep;
    UncBatch& B = F->B;
    batch_header(U, B);
    if (U.T.nsides == 0) return F;
    hipStream_t st = ctx->stream;
    hipEvent_t* ev = ctx->slot_ev[slot];
    const std::string sfx = "#" + std::to_string(slot);
    const int64_t* d_toff = nullptr;
    const UncTabs tb = upload_tables(ctx, U, "u_stage" + sfx, B, &d_toff, ev[4]);
    F->scap = surv_cap(U.total_leaves);
    F->bcap = band_cap(U.total_leaves);
    F->ctl = ctx->dev<UncCtl>("u_ctl" + sfx, 1);
    uint64_t* d_surv = ctx->dev<uint64_t>("u_surv" + sfx, (size_t)F->scap);
    uint64_t* d_band = ctx->dev<uint64_t>("u_band" + sfx, (size_t)F->bcap);
    B.d_surv = d_surv;
    GP_HIP(hipMemsetAsync(F->ctl, 0, sizeof(UncCtl), st));
    GP_HIP(hipEventRecord(ev[0], st));
    launch_filters(ctx, U, tb, F->ctl, d_surv, F->scap, d_band, F->bcap);
    GP_HIP(hipEventRecord(ev[1], st));
    // join over at most rcap survivors (the batch is joined again synchronously
    // when more survive); the part of the table in use is sized on the device
    // from the B rows counted by k_unc_rows and stays at most half full
    F->rcap = std::min<int64_t>(F->scap, (int64_t)(U.total_leaves / 32) + (1 << 20));
    int64_t pmax = 64;
    while (pmax * 512 < F->rcap) pmax <<= 1;
    const int q_target = B.bound + paf_target;
    uint64_t* part = ctx->dev<uint64_t>("u_part" + sfx, (size_t)2 * pmax * kPartCap);
    uint32_t* pcnt = ctx->dev<uint32_t>("u_pcnt" + sfx, (size_t)2 * pmax * kPartSub);
    F->jctl = ctx->dev<JoinCtl2>("u_jctl" + sfx, 1);
    F->ecap = ctx->unc_ecap;
    if (const char* e = std::getenv("GOLAY_UNC_EMIT_CAP"))  // fixed emitted-pair capacity (tests)
        F->ecap = std::max<int64_t>(1, std::strtoll(e, nullptr, 10));
    F->ccap = 2 * F->ecap;
    uint32_t* ca = ctx->dev<uint32_t>("u_ca" + sfx, (size_t)F->ccap);
    uint32_t* cb = ctx->dev<uint32_t>("u_cb" + sfx, (size_t)F->ccap);
    uint32_t* oa = ctx->dev<uint32_t>("u_oa" + sfx, (size_t)F->ecap);
    uint32_t* ob = ctx->dev<uint32_t>("u_ob" + sfx, (size_t)F->ecap);
    F->w = ctx->dev<int8_t>("u_ew" + sfx, (size_t)2 * F->ecap * B.N);
    F->k = ctx->dev<uint8_t>("u_ek" + sfx, (size_t)2 * F->ecap * B.K);
    F->grp = ctx->dev<int32_t>("u_egrp" + sfx, (size_t)F->ecap);
    F->ties = ctx->dev<unsigned int>("u_eties" + sfx, 1);
    // The band recheck and the re-match chain have the same nodes in every
    // batch: they are captured into a graph, which updates the slot's
    // instantiated graph (kernel arguments and grids change per batch) and is
    // launched as one unit.  Everything they use is allocated above.
    BatchGraph graph(ctx, slot);
    if (graph.on) GP_HIP(hipEventRecord(ev[2], st));  // the band time is counted with the join's
    graph.begin(st);
    hipLaunchKernelGGL(k_unc_band, dim3(ctx->num_cus * 2), dim3(256), 0, st, B.d_sides, B.d_radix, B.d_tup,
                       ctx->tables.d_trig64.as<double>(), d_toff, d_band, F->ctl, d_surv, F->scap, (int64_t)0,
                       &F->ctl->band, F->bcap);
    GP_HIP(hipGetLastError());
    if (!graph.on) GP_HIP(hipEventRecord(ev[2], st));
    static_assert(sizeof(JoinCtl2) % 4 == 0);
    hipLaunchKernelGGL(k_zero_ctl, dim3(1), dim3(64), 0, st, reinterpret_cast<unsigned int*>(F->jctl),
                       (int)(sizeof(JoinCtl2) / 4), F->ties, 1);
    GP_HIP(hipGetLastError());
    const unsigned long long* nsurv = &F->ctl->surv;
    const unsigned grid = (unsigned)std::min<int64_t>((F->rcap + 255) / 256, (int64_t)ctx->num_cus * 16);
    hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)std::min<int64_t>((2 * pmax * kPartSub + 255) / 256, ctx->num_cus * 4)),
                       dim3(256), 0, st, pcnt, 2 * pmax * kPartSub);
    if (B.all_pm1)
        hipLaunchKernelGGL(k_unc_rows<true>, dim3(grid), dim3(256), 0, st, B.d_sides, B.d_radix, B.d_tup, B.d_pm1,
                           d_surv, (int64_t)0, paf_target, q_target, nullptr, nsurv, F->rcap, F->jctl, part, pcnt, pmax);
    else
        hipLaunchKernelGGL(k_unc_rows<false>, dim3(grid), dim3(256), 0, st, B.d_sides, B.d_radix, B.d_tup, B.d_pm1,
                           d_surv, (int64_t)0, paf_target, q_target, nullptr, nsurv, F->rcap, F->jctl, part, pcnt, pmax);
    GP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_part_join, dim3((unsigned)std::min<int64_t>(pmax, (int64_t)ctx->num_cus * 8)), dim3(256), 0,
                       st, part, pcnt, pmax, nsurv, F->rcap, ca, cb, F->ccap, F->jctl);
    GP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_pair_check, dim3(ctx->num_cus * 4), dim3(64), 0, st, B.d_sides, B.d_radix, B.d_tup, d_surv,
                       ca, cb, F->ccap, paf_target, q_target, oa, ob, F->ecap, F->jctl);
    GP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_emit_pairs, dim3(ctx->num_cus * 4), dim3(64), 0, st, B.d_sides, B.d_radix, B.d_tup, d_surv,
                       oa, ob, &F->jctl->out, F->ecap, ctx->tables.d_trig64.as<double>() + B.trig_off, B.bound, F->w,
                       F->k, F->grp, F->ties);
    GP_HIP(hipGetLastError());
    if (graph.on) graph.end_and_launch(st);
    GP_HIP(hipEventRecord(ev[3], st));
    return F;
}

UncBatch unc_collect(gp_ctx* ctx, UncInflight& F, int paf_target, UncEmitted& E, gp_stats* stats) {
    UncPrep& U = *F.prep;
    UncBatch B = F.B;
    E = UncEmitted{};
    E.N = B.N;
    E.K = B.K;
    if (U.T.nsides == 0) {
        add_dead_stats(stats, U);
        return B;
    }
    hipEvent_t* ev = ctx->slot_ev[F.slot];
    GP_HIP(hipEventSynchronize(ev[3]));
    UncCtl h{};
    JoinCtl2 hj{};
    unsigned int ht = 0;
    hipStream_t cs = ctx->copy_stream;  // does not wait for the next batch, queued on ctx->stream
    GP_HIP(hipMemcpyAsync(&h, F.ctl, sizeof h, hipMemcpyDeviceToHost, cs));
    GP_HIP(hipMemcpyAsync(&hj, F.jctl, sizeof hj, hipMemcpyDeviceToHost, cs));
    GP_HIP(hipMemcpyAsync(&ht, F.ties, 4, hipMemcpyDeviceToHost, cs));
    GP_HIP(hipStreamSynchronize(cs));
    if ((int64_t)h.band > F.bcap || (int64_t)h.surv > F.scap) {
        // a list overflowed: run the whole batch again on the synchronous path
        B = unc_launch(ctx, U, stats);
        E = unc_join(ctx, B, paf_target, stats);
        return B;
    }
    float ms_filter = 0, ms_band = 0, ms_join = 0;
    GP_HIP(hipEventElapsedTime(&ms_filter, ev[0], ev[1]));
    GP_HIP(hipEventElapsedTime(&ms_band, ev[1], ev[2]));
    GP_HIP(hipEventElapsedTime(&ms_join, ev[2], ev[3]));
    add_dead_stats(stats, U);
    add_filter_stats(stats, U, h, ms_filter, ms_band);
    if (std::getenv("GOLAY_PROFILE"))
        std::fprintf(stderr,
                     "[golay] batch: %llu leaves, %llu survivors, %llu candidate pairs; device filter %.3f, band "
                     "%.3f, join %.3f ms\n",
                     (unsigned long long)U.total_leaves, (unsigned long long)h.surv, hj.ncand, ms_filter, ms_band,
                     ms_join);
    B.nsurv = (int64_t)h.surv;
    if ((int64_t)h.surv > F.rcap || (int64_t)hj.ncand > F.ccap || (int64_t)hj.out > F.ecap || hj.overflow) {
        // the join's lists or a partition were too small: join again with exact sizes
        ctx->unc_ecap = std::max<int64_t>(ctx->unc_ecap, 2 * (int64_t)std::max<uint64_t>(hj.out, hj.ncand / 2));
        E = unc_join(ctx, B, paf_target, stats);
        return B;
    }
    const int64_t ne = (int64_t)hj.out;
    E.count = ne;
    if (ne > 0) {
        E.words.resize((size_t)2 * ne * B.N);
        E.keys.resize((size_t)2 * ne * B.K);
        E.group.resize((size_t)ne);
        GP_HIP(hipMemcpyAsync(E.words.data(), F.w, E.words.size(), hipMemcpyDeviceToHost, cs));
        GP_HIP(hipMemcpyAsync(E.keys.data(), F.k, E.keys.size(), hipMemcpyDeviceToHost, cs));
        GP_HIP(hipMemcpyAsync(E.group.data(), F.grp, E.group.size() * 4, hipMemcpyDeviceToHost, cs));
        GP_HIP(hipStreamSynchronize(cs));
    }
    if (stats) {
        stats->tie_band += ht;
        stats->pair_checks += hj.checks;
        stats->kernel_ms += ms_join;
    }
    return B;
}

UncBatch unc_filter(gp_ctx* ctx, std::span<const SideHost> sides, std::span<const int> group, int order, int paf,
                    gp_stats* stats) {
    if (sides.empty()) return UncBatch{};
    const UncTrig trig = unc_trig(ctx, sides[0].N);
    auto prep = unc_prepare(sides, group, order, paf, ctx->psd_mode == GP_PSD_FP32, trig, &ctx->pinned);
    return unc_launch(ctx, *prep, stats);
}

UncEmitted unc_join(gp_ctx* ctx, const UncBatch& B, int paf_target, gp_stats* stats) {
    UncEmitted E;
    E.N = B.N;
    E.K = B.K;
    const int64_t m = B.nsurv;
    if (m == 0) return E;
    if (m > 0xFFFFFFFFll) throw ArgError("uncompress: more than 2^32 survivors in one batch");
    hipStream_t st = ctx->stream;
    const int q_target = B.bound + paf_target;  // sum A^2 + sum B^2 of an emitted pair
    uint64_t* hashes = ctx->dev<uint64_t>("u_hash", (size_t)m);
    uint64_t T = 1;
    while (T < (uint64_t)m * 2) T <<= 1;  // at most half full even if every row is a B row
    unsigned long long* table = ctx->dev<unsigned long long>("u_ht", (size_t)T);
    JoinCtl2* ctl = ctx->dev<JoinCtl2>("u_jctl", 1);
    GP_HIP(hipEventRecord(ctx->ev0, st));
    GP_HIP(hipMemsetAsync(table, 0xFF, T * sizeof(unsigned long long), st));
    const unsigned grid = (unsigned)std::min<int64_t>((m + 255) / 256, (int64_t)ctx->num_cus * 16);
    hipLaunchKernelGGL(k_unc_rows<false>, dim3(grid), dim3(256), 0, st, B.d_sides, B.d_radix, B.d_tup, B.d_pm1,
                       B.d_surv, m, paf_target, q_target, hashes, nullptr, (int64_t)0, nullptr, nullptr, nullptr,
                       (int64_t)0);
    GP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_ht_insert, dim3(grid), dim3(256), 0, st, hashes, m, table, T - 1);
    GP_HIP(hipGetLastError());
    // candidate pairs (the list is grown once if it overflows), then the exact
    // check, whose output cannot exceed the candidates
    int64_t ccap = 1 << 16;
    JoinCtl2 h{};
    for (int attempt = 0; attempt < 2; ++attempt) {
        uint32_t* ca = ctx->dev<uint32_t>("u_ca", (size_t)ccap);
        uint32_t* cb = ctx->dev<uint32_t>("u_cb", (size_t)ccap);
        GP_HIP(hipMemsetAsync(ctl, 0, sizeof(JoinCtl2), st));
        hipLaunchKernelGGL(k_ht_probe, dim3(grid), dim3(256), 0, st, hashes, m, table, T - 1, ca, cb, ccap, ctl);
        GP_HIP(hipGetLastError());
        GP_HIP(hipMemcpyAsync(&h, ctl, sizeof h, hipMemcpyDeviceToHost, st));
        GP_HIP(hipStreamSynchronize(st));
        if ((int64_t)h.ncand <= ccap) break;
        ccap = (int64_t)h.ncand;
    }
    const int64_t cap = std::max<int64_t>(1, (int64_t)h.ncand);
    {
        uint32_t* oa = ctx->dev<uint32_t>("u_oa", (size_t)cap);
        uint32_t* ob = ctx->dev<uint32_t>("u_ob", (size_t)cap);
        hipLaunchKernelGGL(k_pair_check, dim3(ctx->num_cus * 4), dim3(64), 0, st, B.d_sides, B.d_radix, B.d_tup,
                           B.d_surv, ctx->buf("u_ca").as<uint32_t>(), ctx->buf("u_cb").as<uint32_t>(), ccap,
                           paf_target, q_target, oa, ob, cap, ctl);
        GP_HIP(hipGetLastError());
        GP_HIP(hipMemcpyAsync(&h, ctl, sizeof h, hipMemcpyDeviceToHost, st));
        GP_HIP(hipStreamSynchronize(st));
    }
    const int64_t ne = (int64_t)h.out;
    E.count = ne;
    if (ne > 0) {
        // rows of the emitted pairs: A rows then B rows
        uint32_t* rows = ctx->dev<uint32_t>("u_erows", (size_t)2 * ne);
        GP_HIP(hipMemcpyAsync(rows, ctx->buf("u_oa").p, ne * 4, hipMemcpyDeviceToDevice, st));
        GP_HIP(hipMemcpyAsync(rows + ne, ctx->buf("u_ob").p, ne * 4, hipMemcpyDeviceToDevice, st));
        int8_t* w = ctx->dev<int8_t>("u_ew", (size_t)2 * ne * B.N);
        uint8_t* k = ctx->dev<uint8_t>("u_ek", (size_t)2 * ne * B.K);
        int32_t* grp = ctx->dev<int32_t>("u_egrp", (size_t)2 * ne);
        unsigned int* ties = ctx->dev<unsigned int>("u_eties", 1);
        GP_HIP(hipMemsetAsync(ties, 0, 4, st));
        hipLaunchKernelGGL(k_emit_rows, dim3((unsigned)std::min<int64_t>(ne, 4096)), dim3(64), 0, st,
                           B.d_sides, B.d_radix, B.d_tup, B.d_surv, rows, 2 * ne,
                           ctx->tables.d_trig64.as<double>() + B.trig_off, B.bound, w, k, grp, ties);
        GP_HIP(hipGetLastError());
        E.words.resize((size_t)2 * ne * B.N);
        E.keys.resize((size_t)2 * ne * B.K);
        std::vector<int32_t> g2((size_t)2 * ne);
        unsigned int ht = 0;
        GP_HIP(hipMemcpyAsync(E.words.data(), w, E.words.size(), hipMemcpyDeviceToHost, st));
        GP_HIP(hipMemcpyAsync(E.keys.data(), k, E.keys.size(), hipMemcpyDeviceToHost, st));
        GP_HIP(hipMemcpyAsync(g2.data(), grp, g2.size() * 4, hipMemcpyDeviceToHost, st));
        GP_HIP(hipMemcpyAsync(&ht, ties, 4, hipMemcpyDeviceToHost, st));
        GP_HIP(hipStreamSynchronize(st));
        E.group.assign(g2.begin(), g2.begin() + ne);
        if (stats) stats->tie_band += ht;
    }
    GP_HIP(hipEventRecord(ctx->ev1, st));
    GP_HIP(hipStreamSynchronize(st));
    float ms = 0;
    GP_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (stats) { stats->pair_checks += h.checks; stats->kernel_ms += ms; }
    return E;
}

CandSet unc_side_cands(gp_ctx* ctx, const UncBatch& B, int input_side, uint64_t job_suffix, int proc_id, int proc_num) {
    CandSet c;
    c.n = B.N;
    c.nkey = B.K;
    int tsid = -1;
    for (size_t t = 0; t < B.side_of.size(); ++t)
        if (B.side_of[t] == input_side) tsid = (int)t;
    if (tsid < 0 || B.nsurv == 0) return c;
    hipStream_t st = ctx->stream;
    unsigned long long* cnt = ctx->dev<unsigned long long>("u_scnt", 1);
    uint64_t* sel = ctx->dev<uint64_t>("u_sel", (size_t)B.nsurv);
    GP_HIP(hipMemsetAsync(cnt, 0, 8, st));
    hipLaunchKernelGGL(k_select_side, dim3(ctx->num_cus * 4), dim3(256), 0, st, B.d_surv, B.nsurv, (uint64_t)tsid, cnt,
                       sel);
    GP_HIP(hipGetLastError());
    unsigned long long m = 0;
    GP_HIP(hipMemcpyAsync(&m, cnt, 8, hipMemcpyDeviceToHost, st));
    GP_HIP(hipStreamSynchronize(st));
    std::vector<uint64_t> recs(m);
    if (m) GP_HIP(hipMemcpyAsync(recs.data(), sel, m * 8, hipMemcpyDeviceToHost, st));
    GP_HIP(hipStreamSynchronize(st));
    std::sort(recs.begin(), recs.end());  // DFS order
    if (proc_num > 1) {
        const uint64_t leaf_mask = ((uint64_t)1 << 40) - 1;
        std::erase_if(recs, [&](uint64_t r) {
            return ((r & leaf_mask) / job_suffix) % (uint64_t)proc_num != (uint64_t)proc_id;
        });
        m = recs.size();
    }
    if (m == 0) return c;
    GP_HIP(hipMemcpyAsync(sel, recs.data(), m * 8, hipMemcpyHostToDevice, st));
    int8_t* w = ctx->dev<int8_t>("u_cw", (size_t)m * B.N);
    uint8_t* k = ctx->dev<uint8_t>("u_ck", (size_t)m * B.K);
    UncCtl* ctl = ctx->dev<UncCtl>("u_ctl", 1);
    hipLaunchKernelGGL(k_unc_keys, dim3(ctx->num_cus * 4), dim3(128), 0, st, B.d_sides, B.d_radix, B.d_tup, B.d_pos64, B.d_c64,
                       sel, (int64_t)m, B.bound, w, k, ctl);
    GP_HIP(hipGetLastError());
    c.words.resize((size_t)m * B.N);
    c.keys.resize((size_t)m * B.K);
    GP_HIP(hipMemcpyAsync(c.words.data(), w, c.words.size(), hipMemcpyDeviceToHost, st));
    GP_HIP(hipMemcpyAsync(c.keys.data(), k, c.keys.size(), hipMemcpyDeviceToHost, st));
    GP_HIP(hipStreamSynchronize(st));
    return c;
}

}  // namespace gp
