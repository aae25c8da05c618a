// Candidate generation (replaces generate_hybrid, src/core/generate/generate_hybrid.cpp:107-163).
//
// The reference walks all necklaces in FKM order (lyndon.cpp:47-121) and keeps
// a necklace w when rowsum(w) matches a decomposition, the PSD filter passes and
// w is lexicographically smallest among its images under the index group
// (orderly_equivalence.cpp:53-68): AGL(1,n) = {i -> u i + t} for A, the dihedral
// group {i -> +-i + t} for B.  Lex-min under a group containing the rotations
// implies a necklace, so the kept set is
//   {w : rowsum(w) = x, w <= w o g for all g in G, PSD filter passes}.
// Here that set is enumerated directly: words of fixed letter-index sum are
// ranked in lexicographic order, each wavefront takes one contiguous rank range
// (one contiguous sub-range per lane, successor steps in between), rejects
// non-canonical words in registers, and appends the survivors; the PSD filter
// (split-BF16 MFMA + FP64 band) and FP64 keys follow on the survivors.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <set>
#include <thread>
#include <type_traits>

#include "gp_internal.hpp"
#include "gp_host.hpp"
#include "gp_sort.hpp"
#include "linesort.hpp"
#include "psd_core.hpp"

namespace gp {

typedef unsigned __int128 u128;

// One segment: the words of letter-index sum J whose first letter d0 is their
// least letter (the rest lie in d0..k-1; a necklace always has this form).
// Its words are d0 followed by the n-1 letter suffixes d0 + s, s over the
// alphabet 0..ks-1 (ks = k - d0) with index sum Js = J - n d0, ranked in
// lexicographic order with count table cnt + cnt_off.
struct GenSeg {
    int32_t J;
    int32_t side;     // bit 0: A candidates (AGL), bit 1: B candidates (dihedral)
    int64_t count;    // words of the segment
    int32_t d0, ks;   // first letter; suffix alphabet size
    int32_t Js;       // suffix index sum (shifted letters)
    int32_t pad;
    int64_t cnt_off;  // count table of alphabet size ks
};
struct GenTask {
    int32_t seg;
    int32_t per_lane; // ranks per lane
    int64_t start;    // first rank of the wave
    int64_t end;      // one past the last rank of the segment range
};
struct GenParams {
    int32_t n, k;
    int32_t nunits_a;          // multipliers u != 1 for A (units of n)
    int32_t nunits_b;          // multipliers u != 1 for B (n-1 when n > 2)
    int32_t units_a[64];
    int32_t units_b[2];
    int32_t shard_len;         // prefix length of the job split (0 = none)
    int32_t pad;
    int64_t capacity;          // survivor slots
};
struct GenCtl {
    unsigned long long survivors;
    unsigned long long words;
};

// Words are packed 4 bits per letter index, first letter most significant:
// uint64_t for n <= 16, u128 up to n = 32.  Lexicographic order = unsigned order.
template <class Wd>
__device__ __forceinline__ int digit_at(Wd w, int n, int p) { return (int)((w >> (4 * (n - 1 - p))) & 15); }

template <class Wd>
__device__ __forceinline__ Wd mask_n(int n) {
    return 4 * n >= (int)(8 * sizeof(Wd)) ? ~(Wd)0 : (((Wd)1 << (4 * n)) - 1);
}

template <class Wd>
__device__ __forceinline__ Wd rot(Wd w, int n, int t, Wd mask) {
    if (t == 0) return w;
    return ((w << (4 * t)) | (w >> (4 * (n - t)))) & mask;
}

// w <= every rotation of v (v may equal w).  A rotation can only be smaller if
// it starts with a letter <= w's first letter, so only those are compared.
template <class Wd>
__device__ __forceinline__ bool le_all_rotations(Wd w, Wd v, int n, Wd mask, int t0) {
    const int w0 = digit_at(w, n, 0);
    for (int t = t0; t < n; ++t) {
        if (digit_at(v, n, t) > w0) continue;
        if (rot(v, n, t, mask) < w) return false;
    }
    return true;
}

template <class Wd>
__device__ __forceinline__ Wd decimate(Wd w, int n, int u) {
    Wd r = 0;
    int idx = 0;
    for (int i = 0; i < n; ++i) {
        r = (r << 4) | (Wd)digit_at(w, n, idx);
        idx += u;
        if (idx >= n) idx -= n;
    }
    return r;
}

// Unrank rank r of the words with index sum J (lexicographic order).
template <class Wd>
__device__ Wd unrank(int64_t r, int J, int n, int k, const uint64_t* __restrict__ cnt, int cstride) {
    Wd w = 0;
    for (int p = 0; p < n; ++p) {
        const int rem = n - 1 - p;
        int d = 0;
        for (; d < k; ++d) {
            if (J - d < 0) break;
            const int64_t c = (int64_t)cnt[rem * cstride + (J - d)];
            if (r < c) break;
            r -= c;
        }
        w = (w << 4) | (Wd)d;
        J -= d;
    }
    return w;
}

// Next word with the same digit sum in lexicographic order (false at the end).
template <class Wd>
__device__ bool successor(Wd& w, int n, int k) {
    int S = 0;
    for (int p = n - 2; p >= 0; --p) {
        S += digit_at(w, n, p + 1);
        const int d = digit_at(w, n, p);
        if (d < k - 1 && S >= 1) {
            // digits 0..p-1 unchanged, digit p + 1, suffix = smallest arrangement of S - 1
            const Wd head = p > 0 ? w >> (4 * (n - p)) : (Wd)0;  // digits 0..p-1
            const Wd out = (head << 4) | (Wd)(d + 1);
            const int rem = S - 1;
            const int L = n - 1 - p;
            Wd suf = 0;
            const int full = rem / (k - 1);
            const int part = rem - full * (k - 1);
            // suffix (left to right): zeros, then `part` (if full < L), then full x (k-1)
            for (int q = 0; q < L; ++q) {
                int v;
                const int from_right = L - 1 - q;
                if (from_right < full) v = k - 1;
                else if (from_right == full) v = part;
                else v = 0;
                suf = (suf << 4) | (Wd)v;
            }
            w = (out << (4 * L)) | suf;
            return true;
        }
    }
    return false;
}

template <class Wd>
__global__ void __launch_bounds__(256) k_gen_enum(GenParams P, const GenSeg* __restrict__ segs,
                                                  const GenTask* __restrict__ tasks, int64_t ntasks,
                                                  const uint64_t* __restrict__ cnt, int cstride,
                                                  const uint32_t* __restrict__ shard_bits,
                                                  u128* __restrict__ surv_w, uint8_t* __restrict__ surv_side,
                                                  GenCtl* ctl) {
    // Necklaces (about 1 word in n) wait in a per-wave LDS queue; the group
    // tests (dihedral, AGL) then run on 64 of them at once, so the lanes of a
    // wave do not wait on the few that hold a necklace at each step.
    __shared__ Wd queue_all[4][128];
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    Wd* const queue = queue_all[wv];
    const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + wv;
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
    const int n = P.n, k = P.k;
    const Wd mask = mask_n<Wd>(n);
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned long long nwords = 0;
    // group tests of queue[0 .. cnt) (cnt <= 64, one per lane) and the append of the kept words
    auto drain = [&](int cnt, int side) {
        const bool have = lane < cnt;
        const Wd w = have ? queue[lane] : (Wd)0;
        bool keepA = false, keepB = false;
        if (have) {
            bool dih = true;
            for (int j = 0; j < P.nunits_b && dih; ++j)
                dih = le_all_rotations(w, decimate(w, n, P.units_b[j]), n, mask, 0);
            keepB = dih && (side & 2);
            if (dih && (side & 1)) {
                bool agl = true;
                for (int j = 0; j < P.nunits_a && agl; ++j) {
                    const int u = P.units_a[j];
                    if (u == n - 1) continue;  // covered by the dihedral test
                    agl = le_all_rotations(w, decimate(w, n, u), n, mask, 0);
                }
                keepA = agl;
            }
        }
        const unsigned long long ba = __ballot(keepA), bb = __ballot(keepB);
        const int tot = __popcll(ba) + __popcll(bb);
        if (tot) {
            unsigned long long base = 0;
            if (lane == 0) base = atomicAdd(&ctl->survivors, (unsigned long long)tot);
            base = __shfl(base, 0);
            if (keepA) {
                const unsigned long long i = base + __popcll(ba & below);
                if ((int64_t)i < P.capacity) { surv_w[i] = (u128)w; surv_side[i] = 0; }
            }
            if (keepB) {
                const unsigned long long i = base + __popcll(ba) + __popcll(bb & below);
                if ((int64_t)i < P.capacity) { surv_w[i] = (u128)w; surv_side[i] = 1; }
            }
        }
    };
    for (int64_t ti = wave; ti < ntasks; ti += nwaves) {
        const GenTask T = tasks[ti];
        const GenSeg S = segs[T.seg];
        const int side = __builtin_amdgcn_readfirstlane(S.side);
        int64_t r0 = T.start + (int64_t)lane * T.per_lane;
        int64_t r1 = min(r0 + (int64_t)T.per_lane, T.end);
        // word = d0, then the suffix (n - 1 letters) with d0 added to each letter
        const Wd head = (Wd)S.d0 << (4 * (n - 1));
        Wd add = 0;
        for (int p = 0; p < n - 1; ++p) add = (add << 4) | (Wd)S.d0;
        Wd sfx = 0, w = 0;
        bool live = r0 < r1;
        if (live) {
            sfx = unrank<Wd>(r0, S.Js, n - 1, S.ks, cnt + S.cnt_off, cstride);
            w = head | (sfx + add);
        }
        int64_t r = r0;
        int qn = 0;  // queued necklaces (wave-uniform, < 64 between steps)
        // every lane steps until the whole wave is done
        while (__any(live)) {
            bool neck = false;
            if (live) {
                ++nwords;
                bool ok = true;
                if (P.shard_len > 0) {
                    uint32_t idx = 0;
                    for (int p = 0; p < P.shard_len; ++p) idx = idx * (uint32_t)k + (uint32_t)digit_at(w, n, p);
                    ok = (shard_bits[idx >> 5] >> (idx & 31)) & 1u;
                }
                // rotations (u = 1) first: the necklace condition
                neck = ok && le_all_rotations(w, w, n, mask, 1);
            }
            const unsigned long long bn = __ballot(neck);
            if (neck) queue[qn + __popcll(bn & below)] = w;
            qn += __popcll(bn);
            if (qn >= 64) {
                drain(64, side);
                // the rest (fewer than 64) moves to the front
                const int rest = qn - 64;
                const Wd t = lane < rest ? queue[64 + lane] : (Wd)0;
                if (lane < rest) queue[lane] = t;
                qn = rest;
            }
            if (live) {
                ++r;
                live = r < r1 && successor(sfx, n - 1, S.ks);
                w = head | (sfx + add);
            }
        }
        if (qn > 0) drain(qn, side);
    }
    block_atomic_add(&ctl->words, nwords);
}

// Fused gen (split-BF16 path): the enumeration and group tests of k_gen_enum,
// then, with no word or row leaving the chip, the PSD filter of the canonical
// words 32 at a time, the FP64 settling of the words inside the error band and
// the FP64 keys of the accepted ones, appended as (word, key, side) rows.
//   - canonical words (with their sides: bit 0 A, bit 1 B) wait in a second
//     per-wave LDS queue; 32 of them make one MFMA tile, whose B operand is
//     unpacked in registers from the 4-bit letter indices (x = a0 + 2 d);
//   - the length-n table's A fragments stay in registers, its FP64 twiddles in
//     LDS; one word per lane of the lower half-wave settles the band
//     (fp64_max_psd) and computes the keys (fp64_psd_k, rint) of an accepted
//     word, one row per side it was kept for (A: rint PSD_k, B: bound - rint).
// The rows come out in no particular order (the host puts them in file order).
struct GenFuse {
    FilterThr thr;   // thresholds of the length-n table (S1 = n c)
    int32_t a0;      // letter of index 0 (-c)
    int32_t bound;   // 2 order - paf (inverted B keys)
    int64_t cap;     // output rows
};
struct GenFuseCtl {
    unsigned long long words;  // words visited
    unsigned long long canon;  // canonical (word, side) rows, the PSD decisions
    unsigned long long rows;   // accepted rows (slots used; may exceed cap)
    unsigned long long band;   // words settled in FP64
    unsigned int ties;         // FP64 values within 1e-9 of the threshold or of a rounding half-point
    unsigned int pad;
};

template <class Wd, int NB>
__global__ void __launch_bounds__(256) k_gen_fused(GenParams P, GenFuse F, const GenSeg* __restrict__ segs,
                                                   const GenTask* __restrict__ tasks, int64_t ntasks,
                                                   const uint64_t* __restrict__ cnt, int cstride,
                                                   const uint32_t* __restrict__ shard_bits,
                                                   const uint4* __restrict__ afrag, const double* __restrict__ trig64,
                                                   int8_t* __restrict__ ow, uint8_t* __restrict__ okeys,
                                                   uint8_t* __restrict__ oside, GenFuseCtl* ctl) {
    __shared__ Wd queue_all[4][128];      // necklaces waiting for the group tests
    __shared__ Wd canon_all[4][128];      // canonical words waiting for the PSD tile
    __shared__ uint8_t cside_all[4][128];
    __shared__ int8_t wl_all[4][32][32];  // letters of the tile's words (FP64 work)
    __shared__ double tl[2 * 32];         // FP64 twiddles: cos[n], sin[n]
    const int n = P.n, k = P.k, K = n >> 1;
    for (int t = threadIdx.x; t < 2 * n; t += blockDim.x) tl[t] = trig64[t];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    Wd* const queue = queue_all[wv];
    Wd* const canon = canon_all[wv];
    uint8_t* const cside = cside_all[wv];
    const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + wv;
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
    const Wd mask = mask_n<Wd>(n);
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned long long nwords = 0, ncanon = 0, nband = 0;
    unsigned int nties = 0;
    uint4 af[1][2 * NB];
#pragma unroll
    for (int s = 0; s < 2 * NB; ++s) af[0][s] = afrag[s * 64 + lane];
    int cq = 0;  // canonical words queued (wave-uniform)
    // PSD tile of canon[0 .. m) (m <= 32), decisions, FP64 work, row append
    auto psd_tile = [&](int m) {
        const int col = lane & 31, half = lane >> 5;
        const bool valid = col < m;
        const Wd w = valid ? canon[col] : (Wd)0;
        uint4 bop[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            uint32_t h[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int p0 = 16 * b + 8 * half + 2 * i;
                const int x0 = valid && p0 < n ? F.a0 + 2 * digit_at(w, n, p0) : 0;
                const int x1 = valid && p0 + 1 < n ? F.a0 + 2 * digit_at(w, n, p0 + 1) : 0;
                h[i] = bf16_bits(x0) | (bf16_bits(x1) << 16);
            }
            bop[b] = make_uint4(h[0], h[1], h[2], h[3]);
        }
        f32x16 acc[1];
        acc[0] = f32x16{};
        mfma_tile_reg<NB, 1>(af, bop, acc);
        const float mx = tile_max_psd<1>(acc);  // both halves: max_k PSD_k of word col
        bool acc_ok = false;
        if (half == 0 && valid && !(mx > F.thr.hi)) {
            int8_t* wl = wl_all[wv][col];
            for (int p = 0; p < n; ++p) wl[p] = (int8_t)(F.a0 + 2 * digit_at(w, n, p));
            if (mx < F.thr.lo) {
                acc_ok = true;
            } else {  // inside the error band: FP64 (the order of k_psd_fp64)
                ++nband;
                const double m64 = fp64_max_psd(wl, n, tl, tl + n);
                if (fabs(m64 - F.thr.thr) < 1e-9) ++nties;
                acc_ok = !(m64 > F.thr.thr);
            }
        }
        const int sd = valid ? cside[col] : 0;
        const bool rowA = acc_ok && (sd & 1), rowB = acc_ok && (sd & 2);
        const unsigned long long ba = __ballot(rowA), bb = __ballot(rowB);
        const int tot = __popcll(ba) + __popcll(bb);
        if (tot) {
            unsigned long long base = 0;
            if (lane == 0) base = atomicAdd(&ctl->rows, (unsigned long long)tot);
            base = __shfl(base, 0);
            if (acc_ok && (rowA || rowB)) {
                const int8_t* wl = wl_all[wv][col];
                const unsigned long long sa = base + __popcll(ba & below);
                const unsigned long long sb = base + __popcll(ba) + __popcll(bb & below);
                for (int j = 1; j <= K; ++j) {
                    const double pk = fp64_psd_k(wl, n, tl, tl + n, j);
                    const double f = pk - floor(pk);
                    if (fabs(f - 0.5) < 1e-9) ++nties;
                    const int v = (int)rint(pk);
                    if (rowA && (int64_t)sa < F.cap) okeys[sa * K + j - 1] = (uint8_t)v;
                    if (rowB && (int64_t)sb < F.cap) okeys[sb * K + j - 1] = (uint8_t)(F.bound - v);
                }
                for (int p = 0; p < n; ++p) {
                    if (rowA && (int64_t)sa < F.cap) ow[sa * n + p] = wl[p];
                    if (rowB && (int64_t)sb < F.cap) ow[sb * n + p] = wl[p];
                }
                if (rowA && (int64_t)sa < F.cap) oside[sa] = 0;
                if (rowB && (int64_t)sb < F.cap) oside[sb] = 1;
            }
        }
    };
    // group tests of queue[0 .. qc) (qc <= 64, one per lane); kept words join the canonical queue
    auto drain = [&](int qc, int side) {
        const bool have = lane < qc;
        const Wd w = have ? queue[lane] : (Wd)0;
        bool keepA = false, keepB = false;
        if (have) {
            bool dih = true;
            for (int j = 0; j < P.nunits_b && dih; ++j)
                dih = le_all_rotations(w, decimate(w, n, P.units_b[j]), n, mask, 0);
            keepB = dih && (side & 2);
            if (dih && (side & 1)) {
                bool agl = true;
                for (int j = 0; j < P.nunits_a && agl; ++j) {
                    const int u = P.units_a[j];
                    if (u == n - 1) continue;  // covered by the dihedral test
                    agl = le_all_rotations(w, decimate(w, n, u), n, mask, 0);
                }
                keepA = agl;
            }
        }
        const bool keep = keepA || keepB;
        ncanon += (unsigned long long)keepA + (unsigned long long)keepB;
        const unsigned long long bk = __ballot(keep);
        if (keep) {
            const int at = cq + __popcll(bk & below);
            canon[at] = w;
            cside[at] = (uint8_t)((keepA ? 1 : 0) | (keepB ? 2 : 0));
        }
        cq += __popcll(bk);
        while (cq >= 32) {
            psd_tile(32);
            // the rest (fewer than 96) moves to the front
            const int rest = cq - 32;
            Wd t0 = 0, t1 = 0;
            uint8_t s0 = 0, s1 = 0;
            if (lane < rest) { t0 = canon[32 + lane]; s0 = cside[32 + lane]; }
            if (lane + 64 < rest) { t1 = canon[96 + lane]; s1 = cside[96 + lane]; }
            if (lane < rest) { canon[lane] = t0; cside[lane] = s0; }
            if (lane + 64 < rest) { canon[64 + lane] = t1; cside[64 + lane] = s1; }
            cq = rest;
        }
    };
    for (int64_t ti = wave; ti < ntasks; ti += nwaves) {
        const GenTask T = tasks[ti];
        const GenSeg S = segs[T.seg];
        const int side = __builtin_amdgcn_readfirstlane(S.side);
        int64_t r0 = T.start + (int64_t)lane * T.per_lane;
        int64_t r1 = min(r0 + (int64_t)T.per_lane, T.end);
        const Wd head = (Wd)S.d0 << (4 * (n - 1));
        Wd add = 0;
        for (int p = 0; p < n - 1; ++p) add = (add << 4) | (Wd)S.d0;
        Wd sfx = 0, w = 0;
        bool live = r0 < r1;
        if (live) {
            sfx = unrank<Wd>(r0, S.Js, n - 1, S.ks, cnt + S.cnt_off, cstride);
            w = head | (sfx + add);
        }
        int64_t r = r0;
        int qn = 0;
        while (__any(live)) {
            bool neck = false;
            if (live) {
                ++nwords;
                bool ok = true;
                if (P.shard_len > 0) {
                    uint32_t idx = 0;
                    for (int p = 0; p < P.shard_len; ++p) idx = idx * (uint32_t)k + (uint32_t)digit_at(w, n, p);
                    ok = (shard_bits[idx >> 5] >> (idx & 31)) & 1u;
                }
                neck = ok && le_all_rotations(w, w, n, mask, 1);
            }
            const unsigned long long bn = __ballot(neck);
            if (neck) queue[qn + __popcll(bn & below)] = w;
            qn += __popcll(bn);
            if (qn >= 64) {
                drain(64, side);
                const int rest = qn - 64;
                const Wd t = lane < rest ? queue[64 + lane] : (Wd)0;
                if (lane < rest) queue[lane] = t;
                qn = rest;
            }
            if (live) {
                ++r;
                live = r < r1 && successor(sfx, n - 1, S.ks);
                w = head | (sfx + add);
            }
        }
        if (qn > 0) drain(qn, side);
    }
    if (cq > 0) psd_tile(cq);
    block_atomic_add(&ctl->words, nwords, &ctl->canon, ncanon, &ctl->band, nband);
    for (int o = 32; o; o >>= 1) nties += (unsigned int)__shfl_xor((int)nties, o);
    if (lane == 0 && nties) atomicAdd(&ctl->ties, nties);
}

__global__ void k_gen_unpack(const u128* __restrict__ surv_w, int64_t m, int n, int a0, int stride,
                             int8_t* __restrict__ words) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const u128 w = surv_w[i];
        int8_t* o = words + i * stride;
        for (int p = 0; p < stride; ++p) o[p] = p < n ? (int8_t)(a0 + 2 * digit_at(w, n, p)) : 0;
    }
}

// Accepted rows, compacted: word (n bytes), FP64 key (K bytes: A plain rint(PSD_k),
// B inverted bound - rint(PSD_k)) and side, each at its own slot (unordered).
// Blocks of 256; the twiddles and each thread's word are staged in LDS, and
// PSD_k is summed in the order of fp64_psd.
__global__ void __launch_bounds__(256) k_gen_emit(const int8_t* __restrict__ words, int stride, int64_t m, int n,
                                                  const double* __restrict__ trig, const uint8_t* __restrict__ decision,
                                                  const uint8_t* __restrict__ side, int bound,
                                                  unsigned long long* __restrict__ count, int8_t* __restrict__ ow,
                                                  uint8_t* __restrict__ okeys, uint8_t* __restrict__ oside,
                                                  unsigned int* __restrict__ ties) {
    __shared__ double tl[2 * kMaxN];
    __shared__ int8_t wl[256][kMaxN];
    for (int t = threadIdx.x; t < 2 * n; t += blockDim.x) tl[t] = trig[t];
    __syncthreads();
    const int K = n >> 1;
    const int lane = threadIdx.x & 63;
    int8_t* w = wl[threadIdx.x];
    unsigned int t = 0;
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < m; i0 += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = i0 + threadIdx.x;
        const bool acc = i < m && decision[i] == kAccept;
        // one atomic per wave for the wave's accepted rows (slot < m)
        const unsigned long long bal = __ballot(acc);
        if (!bal) continue;
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(count, (unsigned long long)__popcll(bal));
        base = __shfl(base, 0);
        if (!acc) continue;
        const unsigned long long slot = base + __popcll(bal & ((1ull << lane) - 1ull));
        for (int j = 0; j < n; ++j) {
            w[j] = words[i * stride + j];
            ow[slot * n + j] = w[j];
        }
        for (int k = 1; k <= K; ++k) {
            const double p = fp64_psd_k(w, n, tl, tl + n, k);
            const double f = p - floor(p);
            if (fabs(f - 0.5) < 1e-9) ++t;
            const int v = (int)rint(p);
            okeys[slot * K + k - 1] = (uint8_t)(side[i] ? bound - v : v);
        }
        oside[slot] = side[i];
    }
    if (t) atomicAdd(ties, t);
}

// Sort keys of the accepted rows in the reference's file order, A rows first:
// side << 63 | the letter indices (v - a0) / 2, 4 bits each, first letter
// most significant (n <= 15).
__global__ void k_gen_sortkey(const int8_t* __restrict__ ow, const uint8_t* __restrict__ oside, int64_t m, int n,
                              int a0, uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t k = 0;
        for (int j = 0; j < n; ++j) k = (k << 4) | (uint64_t)((ow[i * n + j] - a0) >> 1);
        keys[i] = ((uint64_t)(oside[i] != 0) << 63) | k;
        vals[i] = (uint32_t)i;
    }
}

// Rows permuted into sorted order.
__global__ void k_gen_gather(const int8_t* __restrict__ ow, const uint8_t* __restrict__ okeys,
                             const uint32_t* __restrict__ order, int64_t m, int n, int K, int8_t* __restrict__ w2,
                             uint8_t* __restrict__ k2) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = order[i];
        for (int j = 0; j < n; ++j) w2[i * n + j] = ow[s * n + j];
        for (int k = 0; k < K; ++k) k2[i * K + k] = okeys[s * K + k];
    }
}

// ---------------------------------------------------------------------------
// host

namespace {

std::vector<int> units_of(int n) {
    std::vector<int> u;
    for (int i = 2; i < n; ++i)
        if (std::gcd(i, n) == 1) u.push_back(i);
    return u;
}

}  // namespace

int generate_level(gp_ctx* ctx, int order, int compress, int paf, int proc_id, int proc_num, GenOutput& out,
                   gp_stats* stats) {
    hipStream_t st = ctx->stream;
    Phase ph("gen: setup");
    const int n = order / compress;
    if (n < 2 || n > 32) throw ArgError("gen: word length " + std::to_string(n) + " outside [2, 32]");
    const std::vector<int> alpha = getalphabet(compress);
    const int k = (int)alpha.size();
    if (k > 16) throw ArgError("gen: alphabet larger than 16 letters");
    const int a0 = alpha[0];
    const auto decomps = getdecomps(order * 2 + (order - 1) * paf);
    if (decomps.empty()) return GP_NO_INPUT;
    const int bound = order * 2 - paf;
    if (bound > 255 || bound < 0) throw ArgError("gen: keys need 0 <= 2*order - paf <= 255");
    // row sum x <-> index sum J = (x - n a0) / 2
    std::map<int, int> side_of_J;
    auto addJ = [&](int x, int bit) {
        int v = x - n * a0;
        if (v < 0 || (v & 1)) return;
        int J = v / 2;
        if (J > n * (k - 1)) return;
        side_of_J[J] |= bit;
    };
    for (auto& d : decomps) { addJ(d.first, 1); addJ(d.second, 2); }
    // counts cnt_a[m][j]: words of m letters over the alphabet 0..a-1 with index
    // sum j, one table per alphabet size a = 1..k (table a at (a - 1) * tsize)
    const int jmax = n * (k - 1);
    const int cstride = jmax + 1;
    const size_t tsize = (size_t)(n + 1) * cstride;
    std::vector<uint64_t> cnt(tsize * k, 0);
    const uint64_t cap = (uint64_t)1 << 62;
    for (int a = 1; a <= k; ++a) {
        uint64_t* c = &cnt[(a - 1) * tsize];
        c[0] = 1;
        for (int m = 1; m <= n; ++m)
            for (int j = 0; j <= jmax; ++j) {
                unsigned __int128 s = 0;
                for (int d = 0; d < a && d <= j; ++d) s += c[(m - 1) * cstride + (j - d)];
                c[m * cstride + j] = s > cap ? cap : (uint64_t)s;
            }
    }
    // a kept word is <= its rotations, so its first letter is its least letter:
    // one segment per (J, first letter d0) over the suffixes with letters >= d0
    std::vector<GenSeg> segs;
    std::vector<GenTask> tasks;
    int64_t total = 0;
    for (auto& [J, side] : side_of_J)
        for (int d0 = 0; d0 < k; ++d0) {
            const int ks = k - d0, Js = J - n * d0;
            if (Js < 0 || Js > (n - 1) * (ks - 1)) continue;
            const size_t off = (size_t)(ks - 1) * tsize;
            GenSeg s{J, side, (int64_t)cnt[off + (size_t)(n - 1) * cstride + Js], d0, ks, Js, 0, (int64_t)off};
            if ((uint64_t)s.count >= cap) throw ArgError("gen: search space too large");
            if (s.count == 0) continue;
            segs.push_back(s);
            total += s.count;
        }
    // ranks per lane: about 8 wave tasks per resident wave (so that the last
    // round of tasks is short), at least 16 ranks to amortise each lane's unrank
    const int64_t target_tasks = (int64_t)ctx->num_cus * 32 * 8;
    const int heuristic = (int)std::clamp<int64_t>((total + 64 * target_tasks - 1) / (64 * target_tasks), 16, 128);
    const int per_lane = (int)std::clamp<int64_t>(
        tuned(ctx, "gen_ranks_per_lane/" + std::to_string(n) + "/" + std::to_string(k), heuristic), 4, 4096);
    for (size_t id = 0; id < segs.size(); ++id)
        for (int64_t r = 0; r < segs[id].count; r += 64 * per_lane)
            tasks.push_back(GenTask{(int32_t)id, per_lane, r, segs[id].count});
    GenParams P{};
    P.n = n;
    P.k = k;
    auto ua = units_of(n);
    if (ua.size() > 64) throw ArgError("gen: too many units");
    P.nunits_a = (int)ua.size();
    for (size_t i = 0; i < ua.size(); ++i) P.units_a[i] = ua[i];
    P.nunits_b = n > 2 ? 1 : 0;
    P.units_b[0] = n - 1;
    // job split (lyndon.cpp:87-121): prefix shard table
    std::vector<uint32_t> bits;
    int shard_len = 0;
    if (proc_num > 1) {
        shard_len = shard_prefix_len(n, k, proc_num);
        if (shard_len > 0) {
            uint64_t space = 1;
            for (int i = 0; i < shard_len; ++i) space *= (uint64_t)k;
            if (space > ((uint64_t)1 << 31)) throw ArgError("gen: job-split prefix table too large");
            bits.assign((size_t)(space + 31) / 32, 0);
            for (uint64_t idx : shard_prefixes(shard_len, alpha, proc_id, proc_num)) bits[idx >> 5] |= 1u << (idx & 31);
        }
    }
    P.shard_len = shard_len;
    DevBuf& d_segs = ctx->buf("g_segs");
    DevBuf& d_tasks = ctx->buf("g_tasks");
    DevBuf& d_cnt = ctx->buf("g_cnt");
    DevBuf& d_bits = ctx->buf("g_bits");
    DevBuf& d_ctl = ctx->buf("g_ctl");
    GP_HIP(hipMemcpyAsync(d_segs.ensure(segs.size() * sizeof(GenSeg)), segs.data(), segs.size() * sizeof(GenSeg),
                          hipMemcpyHostToDevice, st));
    GP_HIP(hipMemcpyAsync(d_tasks.ensure(tasks.size() * sizeof(GenTask) + 16), tasks.data(),
                          tasks.size() * sizeof(GenTask), hipMemcpyHostToDevice, st));
    GP_HIP(hipMemcpyAsync(d_cnt.ensure(cnt.size() * 8), cnt.data(), cnt.size() * 8, hipMemcpyHostToDevice, st));
    if (!bits.empty())
        GP_HIP(hipMemcpyAsync(d_bits.ensure(bits.size() * 4), bits.data(), bits.size() * 4, hipMemcpyHostToDevice, st));
    const int K = n / 2;
    const int table = ctx->tables.get(n);
    ctx->tables.upload(st);
    const PsdTableDesc& d = ctx->tables.desc[table];
    const double E = ctx->psd_mode == GP_PSD_SPLIT_BF16 ? split_bf16_E(d, (double)n * compress)
                                                        : fp32_E(n, (double)n * compress);
    const FilterThr thr = make_thr(2.0 * order - paf + 0.001, E);
    int64_t m = 0;        // canonical (word, side) rows: the PSD decisions
    int64_t acc_n = 0;    // accepted rows
    unsigned int ties = 0;
    int8_t* d_ow = nullptr;
    uint8_t* d_ok = nullptr;
    uint8_t* d_os = nullptr;
    const int64_t waves = std::min<int64_t>((int64_t)tasks.size(), (int64_t)ctx->num_cus * 32);
    const dim3 grid((unsigned)std::max<int64_t>(1, (waves + 3) / 4));
    const uint32_t* d_shard = bits.empty() ? nullptr : d_bits.as<uint32_t>();
    if (ctx->psd_mode == GP_PSD_SPLIT_BF16) {
        // one fused kernel: enumerate, group tests, PSD filter, FP64 band and keys
        ph.next("gen: fused enumerate + psd + keys");
        int64_t cap = std::max<int64_t>(1 << 16, total / 64 + 4096);
        GenFuseCtl h{};
        for (int attempt = 0; attempt < 2; ++attempt) {
            d_ow = ctx->dev<int8_t>("g_ow", (size_t)cap * n);
            d_ok = ctx->dev<uint8_t>("g_keys", (size_t)cap * std::max(K, 1));
            d_os = ctx->dev<uint8_t>("g_os", (size_t)cap);
            GenFuseCtl* d_fc = static_cast<GenFuseCtl*>(d_ctl.ensure(sizeof(GenFuseCtl)));
            GP_HIP(hipMemsetAsync(d_fc, 0, sizeof(GenFuseCtl), st));
            GenFuse F{thr, a0, bound, cap};
            const uint4* af = ctx->tables.d_frags.as<uint4>() + d.frag_off;
            const double* t64 = ctx->tables.d_trig64.as<double>() + d.trig_off;
            GP_HIP(hipEventRecord(ctx->ev0, st));
            if (!tasks.empty()) {
                const GenSeg* sg = d_segs.as<GenSeg>();
                const GenTask* tk = d_tasks.as<GenTask>();
                const uint64_t* cn = d_cnt.as<uint64_t>();
                const int64_t nt = (int64_t)tasks.size();
                if (n <= 16 && d.nb == 1)
                    hipLaunchKernelGGL((k_gen_fused<uint64_t, 1>), grid, dim3(256), 0, st, P, F, sg, tk, nt, cn, cstride,
                                       d_shard, af, t64, d_ow, d_ok, d_os, d_fc);
                else if (n <= 16)
                    hipLaunchKernelGGL((k_gen_fused<uint64_t, 2>), grid, dim3(256), 0, st, P, F, sg, tk, nt, cn, cstride,
                                       d_shard, af, t64, d_ow, d_ok, d_os, d_fc);
                else
                    hipLaunchKernelGGL((k_gen_fused<u128, 2>), grid, dim3(256), 0, st, P, F, sg, tk, nt, cn, cstride,
                                       d_shard, af, t64, d_ow, d_ok, d_os, d_fc);
            }
            GP_HIP(hipGetLastError());
            GP_HIP(hipEventRecord(ctx->ev1, st));
            GP_HIP(hipMemcpyAsync(&h, d_fc, sizeof h, hipMemcpyDeviceToHost, st));
            GP_HIP(hipStreamSynchronize(st));
            float ms = 0;
            GP_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            if (stats) { stats->kernel_ms += ms; stats->psd_kernel_ms += ms; stats->psd_launches += 1; }
            if ((int64_t)h.rows <= cap) break;
            cap = (int64_t)h.rows + 1024;
        }
        m = (int64_t)h.canon;
        acc_n = (int64_t)h.rows;
        ties = h.ties;
        if (stats) {
            stats->words += h.words;
            stats->canonical += (uint64_t)m;
            stats->psd_decisions += (uint64_t)m;
            stats->band_rechecks += h.band;
        }
    } else {
        // plain FP32 path: canonical rows materialised, filtered, keyed
        int64_t capacity = std::max<int64_t>(1 << 20, total / 4 + 1024);
        GenCtl h{};
        ph.next("gen: enumerate");
        for (int attempt = 0; attempt < 2; ++attempt) {
            P.capacity = capacity;
            u128* d_w = static_cast<u128*>(ctx->buf("g_w").ensure((size_t)capacity * sizeof(u128)));
            uint8_t* d_side = static_cast<uint8_t*>(ctx->buf("g_side").ensure((size_t)capacity));
            GP_HIP(hipMemsetAsync(d_ctl.ensure(sizeof(GenCtl)), 0, sizeof(GenCtl), st));
            GP_HIP(hipEventRecord(ctx->ev0, st));
            if (!tasks.empty()) {
                if (n <= 16)
                    hipLaunchKernelGGL(k_gen_enum<uint64_t>, grid, dim3(256), 0, st, P, d_segs.as<GenSeg>(),
                                       d_tasks.as<GenTask>(), (int64_t)tasks.size(), d_cnt.as<uint64_t>(), cstride,
                                       d_shard, d_w, d_side, d_ctl.as<GenCtl>());
                else
                    hipLaunchKernelGGL(k_gen_enum<u128>, grid, dim3(256), 0, st, P, d_segs.as<GenSeg>(),
                                       d_tasks.as<GenTask>(), (int64_t)tasks.size(), d_cnt.as<uint64_t>(), cstride,
                                       d_shard, d_w, d_side, d_ctl.as<GenCtl>());
            }
            GP_HIP(hipGetLastError());
            GP_HIP(hipEventRecord(ctx->ev1, st));
            GP_HIP(hipMemcpyAsync(&h, d_ctl.p, sizeof h, hipMemcpyDeviceToHost, st));
            GP_HIP(hipStreamSynchronize(st));
            float ms = 0;
            GP_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            if (stats) stats->kernel_ms += ms;
            if ((int64_t)h.survivors <= capacity) break;
            capacity = (int64_t)h.survivors + 1024;
        }
        m = (int64_t)h.survivors;
        if (stats) { stats->words += h.words; stats->canonical += (uint64_t)m; }
        ph.next("gen: psd + keys");
        const int stride = 16 * d.nb;
        int8_t* d_words = static_cast<int8_t*>(ctx->words.ensure((size_t)std::max<int64_t>(m, 1) * stride));
        uint8_t* d_dec = static_cast<uint8_t*>(ctx->decisions.ensure((size_t)std::max<int64_t>(m, 1)));
        if (m > 0) {
            hipLaunchKernelGGL(k_gen_unpack, dim3(ctx->num_cus * 4), dim3(256), 0, st, ctx->buf("g_w").as<u128>(), m, n,
                               a0, stride, d_words);
            GP_HIP(hipGetLastError());
            std::vector<FilterThr> tt(ctx->tables.desc.size());
            tt[table] = thr;
            std::vector<BatchSeg> bs{BatchSeg{table, 0, 0, m}};
            run_psd_filter(ctx, d_words, stride, m, bs, tt, d_dec, stats);
            d_ow = ctx->dev<int8_t>("g_ow", (size_t)m * n);
            d_ok = ctx->dev<uint8_t>("g_keys", (size_t)m * std::max(K, 1));
            d_os = ctx->dev<uint8_t>("g_os", (size_t)m);
            auto* d_cnt_e = ctx->dev<unsigned long long>("g_emit_count", 1);
            unsigned int* d_ties = ctx->dev<unsigned int>("g_ties", 1);
            GP_HIP(hipMemsetAsync(d_ties, 0, sizeof(unsigned int), st));
            GP_HIP(hipMemsetAsync(d_cnt_e, 0, sizeof(unsigned long long), st));
            hipLaunchKernelGGL(k_gen_emit, dim3(ctx->num_cus * 4), dim3(256), 0, st, d_words, stride, m, n,
                               ctx->tables.d_trig64.as<double>() + d.trig_off, d_dec, ctx->buf("g_side").as<uint8_t>(),
                               bound, d_cnt_e, d_ow, d_ok, d_os, d_ties);
            GP_HIP(hipGetLastError());
            unsigned long long cnt_h = 0;
            GP_HIP(hipMemcpyAsync(&cnt_h, d_cnt_e, sizeof cnt_h, hipMemcpyDeviceToHost, st));
            GP_HIP(hipMemcpyAsync(&ties, d_ties, sizeof ties, hipMemcpyDeviceToHost, st));
            GP_HIP(hipStreamSynchronize(st));
            if ((int64_t)cnt_h > m) throw std::runtime_error("gen: emitted more rows than candidates");
            acc_n = (int64_t)cnt_h;
        }
    }
    // accepted rows in the reference's file order (n <= 15 on the device)
    bool sorted_on_device = false;  // rows already in file order, A rows first
    std::vector<int8_t> words_h;   // accepted rows only: [acc_n][n]
    std::vector<uint8_t> keys_h, side_h;
    words_h.resize((size_t)acc_n * n);
    keys_h.resize((size_t)acc_n * K);
    side_h.resize((size_t)acc_n);
    if (acc_n > 0) {
        const int8_t* src_w = d_ow;
        const uint8_t* src_k = d_ok;
        if (acc_n > 1 && n <= 15) {
            // file order on the device: radix sort of (side, word) keys, rows gathered
            uint64_t* d_sk = ctx->dev<uint64_t>("g_sk", (size_t)acc_n);
            uint32_t* d_sv = ctx->dev<uint32_t>("g_sv", (size_t)acc_n);
            uint64_t* d_tk = ctx->dev<uint64_t>("g_tk", (size_t)acc_n);
            uint32_t* d_tv = ctx->dev<uint32_t>("g_tv", (size_t)acc_n);
            uint32_t* d_hist = static_cast<uint32_t*>(ctx->buf("g_hist").ensure(radix_hist_bytes(acc_n)));
            const unsigned g = (unsigned)std::min<int64_t>((acc_n + 255) / 256, (int64_t)ctx->num_cus * 8);
            hipLaunchKernelGGL(k_gen_sortkey, dim3(g), dim3(256), 0, st, d_ow, d_os, acc_n, n, a0, d_sk, d_sv);
            GP_HIP(hipGetLastError());
            radix_sort_pairs(st, d_sk, d_sv, d_tk, d_tv, d_hist, acc_n, 0, 64);
            int8_t* d_w2 = ctx->dev<int8_t>("g_ow2", (size_t)acc_n * n);
            uint8_t* d_k2 = ctx->dev<uint8_t>("g_ok2", (size_t)acc_n * std::max(K, 1));
            hipLaunchKernelGGL(k_gen_gather, dim3(g), dim3(256), 0, st, d_ow, d_ok, d_sv, acc_n, n, K, d_w2, d_k2);
            GP_HIP(hipGetLastError());
            src_w = d_w2;
            src_k = d_k2;
            sorted_on_device = true;
        }
        GP_HIP(hipMemcpyAsync(words_h.data(), src_w, words_h.size(), hipMemcpyDeviceToHost, st));
        GP_HIP(hipMemcpyAsync(keys_h.data(), src_k, keys_h.size(), hipMemcpyDeviceToHost, st));
        GP_HIP(hipMemcpyAsync(side_h.data(), d_os, side_h.size(), hipMemcpyDeviceToHost, st));
        GP_HIP(hipStreamSynchronize(st));
    }
    if (stats) stats->tie_band += ties;
    // collect A and B in the reference's file order (lexicographic order of the words)
    ph.next("gen: order");
    std::vector<int64_t> ia, ib;
    for (int64_t i = 0; i < acc_n; ++i) (side_h[i] ? ib : ia).push_back(i);
    if (sorted_on_device) {  // side_h is in emission order: only the A/B split counts are used
        const int64_t na = (int64_t)ia.size();
        std::iota(ia.begin(), ia.end(), (int64_t)0);
        std::iota(ib.begin(), ib.end(), na);
    }
    if (stats) stats->psd_pass += ia.size() + ib.size();
    // packed big-endian keys (byte v + 128 per letter): unsigned order = signed
    // lexicographic order; (key, row) records sorted directly, A and B concurrently
    const int W = (n + 7) / 8;
    auto sort_side = [&](std::vector<int64_t>& idx) {
        auto run = [&](auto tag) {
            constexpr int KW = decltype(tag)::value;
            struct Rec {
                uint64_t k[KW];
                int64_t i;
            };
            std::vector<Rec> rec(idx.size());
            parallel_chunks((int64_t)idx.size(), [&](int64_t x0, int64_t x1) {
                for (int64_t x = x0; x < x1; ++x) {
                    Rec& R = rec[x];
                    for (int w = 0; w < KW; ++w) R.k[w] = 0;
                    R.i = idx[x];
                    for (int j = 0; j < n; ++j)
                        R.k[j / 8] |= (uint64_t)(uint8_t)(words_h[R.i * n + j] + 128) << (56 - 8 * (j % 8));
                }
            });
            parallel_sort(rec, [](const Rec& x, const Rec& y) {
                for (int w = 0; w < KW; ++w)
                    if (x.k[w] != y.k[w]) return x.k[w] < y.k[w];
                return false;
            });
            for (size_t x = 0; x < idx.size(); ++x) idx[x] = rec[x].i;
        };
        if (W == 1) run(std::integral_constant<int, 1>{});
        else if (W == 2) run(std::integral_constant<int, 2>{});
        else run(std::integral_constant<int, 4>{});
    };
    if (!sorted_on_device) {
        std::thread ta([&] { sort_side(ia); });
        sort_side(ib);
        ta.join();
    }
    auto emit = [&](const std::vector<int64_t>& idx, CandSet& cs) {
        cs.n = n;
        cs.nkey = K;
        cs.words.resize(idx.size() * n);
        cs.keys.resize(idx.size() * K);
        parallel_chunks((int64_t)idx.size(), [&](int64_t r0, int64_t r1) {
            for (int64_t r = r0; r < r1; ++r) {
                std::memcpy(&cs.words[r * n], &words_h[idx[r] * n], n);
                std::memcpy(&cs.keys[r * K], &keys_h[idx[r] * K], K);
            }
        });
    };
    emit(ia, out.a);
    emit(ib, out.b);
    return GP_OK;
}

}  // namespace gp
