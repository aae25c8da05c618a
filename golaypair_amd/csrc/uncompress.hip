// Uncompression (replaces uncompress_recursive / uncompress_gpu,
// src/core/uncompression/uncompression.cpp:81-210, gpu/uncompress_gpu.cpp:17-92).
//
// One compressed sequence of length L at compression c expands to words of
// length N = L r (r = c/c') over the finer alphabet: position j takes an r-tuple
// t from its split list T_j and writes t[i] to index j + L i.  T_0 holds the
// dihedral-orbit minima of the first letter's splits (:133-142); the letter order
// follows the reference's in-place rotation loop (:113-131).  Leaves are
// numbered in the reference's DFS order (position 0 most significant digit).
//
// Device pipeline per batch of sides (a side = one sequence + its key style):
//   k_unc_filter   work item = (side, outer digits).  X_pre = X_outer (FP32 sum
//                  of per-position contributions) + Mid[m] (FP32 table over
//                  the middle digits) seeds the MFMA accumulator; the split-BF16
//                  contraction runs only over the suffix positions, 32 suffix
//                  combinations per tile.  Survivors go through per-wave LDS
//                  buffers to global memory; band candidates to a band list.
//   k_unc_band     FP64 direct DFT of band candidates.
//   radix sort     survivors by (side, leaf): grouped per side, DFS order.
//   k_unc_keys     word + FP64 key digits of every survivor (sum over positions
//                  of FP64 per-position contributions).
// then per pair the candidate join with the exact PAF check (match.hip).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <set>

#include "gp_host.hpp"
#include "gp_internal.hpp"
#include "gp_sort.hpp"
#include "gp_unc.hpp"
#include "psd_core.hpp"
#include "textfmt.hpp"

namespace gp {

// ---------------------------------------------------------------------------
// host: side descriptors (uncompression.cpp:85-142)

namespace {

// array.cpp:131-160  nondecreasing len-tuples over the alphabet, lex order
void combinations(int len, const std::vector<int>& alpha, size_t from, std::vector<int>& cur,
                  std::vector<std::vector<int>>& out) {
    if ((int)cur.size() == len) { out.push_back(cur); return; }
    for (size_t i = from; i < alpha.size(); ++i) {
        cur.push_back(alpha[i]);
        combinations(len, alpha, i, cur, out);
        cur.pop_back();
    }
}

// equivalence.cpp:26-44 orbit under rotation + reversal; returns the minimum.
std::vector<int> dihedral_min(const std::vector<int>& t) {
    std::vector<int> best = t, v = t;
    const size_t r = t.size();
    for (int flip = 0; flip < 2; ++flip) {
        for (size_t s = 0; s < r; ++s) {
            std::rotate(v.begin(), v.begin() + 1, v.end());
            if (v < best) best = v;
        }
        std::reverse(v.begin(), v.end());
    }
    return best;
}

}  // namespace

SideHost build_side(const int8_t* orig_in, int L, int compress, int new_compress, int seqflag) {
    SideHost S;
    S.L = L;
    S.r = compress / new_compress;
    S.N = L * S.r;
    S.seqflag = seqflag;
    const std::vector<int> alphabet = getalphabet(compress);
    const std::vector<int> newalpha = getalphabet(new_compress);
    std::vector<std::vector<int>> parts;
    std::vector<int> cur;
    combinations(S.r, newalpha, 0, cur, parts);
    std::map<int, std::vector<std::vector<int>>> partitions;  // :97-111
    for (int letter : alphabet) {
        std::vector<std::vector<int>> part;
        for (std::vector<int> p : parts) {
            int sum = 0;
            for (int v : p) sum += v;
            if (sum == letter)
                do { part.push_back(p); } while (std::next_permutation(p.begin(), p.end()));
        }
        partitions[letter] = part;
    }
    std::vector<int> orig(orig_in, orig_in + L);
    for (int v : orig)
        if (!partitions.count(v)) throw ArgError("uncompress: letter " + std::to_string(v) + " not in the alphabet");
    // :113-131 best letter and the in-place rotation loop
    std::set<int> seta(orig.begin(), orig.end());
    size_t mx = 0;
    int best = orig[0];
    for (int e : seta)
        if (partitions.at(e).size() > mx) { mx = partitions.at(e).size(); best = e; }
    for (int i = 0; i < L; ++i)
        if (orig[i] == best) std::rotate(orig.begin(), orig.begin() + i, orig.end());
    S.orig = orig;
    // :133-142 first position: dihedral minima, sorted and unique
    std::set<std::vector<int>> perma;
    for (const auto& p : partitions.at(orig[0])) perma.insert(dihedral_min(p));
    S.T.resize(L);
    S.T[0].assign(perma.begin(), perma.end());
    for (int j = 1; j < L; ++j) S.T[j] = partitions.at(orig[j]);
    S.leaves = 1;
    for (int j = 0; j < L; ++j) {
        const uint64_t R = S.T[j].size();
        if (R == 0) { S.leaves = 0; break; }
        if (S.leaves > (((uint64_t)1 << 40) / R)) throw ArgError("uncompress: more than 2^40 leaves in one sequence");
        S.leaves *= R;
    }
    return S;
}

// ---------------------------------------------------------------------------
// device

struct UncSide {
    int32_t L, r, N, K;
    int32_t nbs, a, b, seqflag;
    int32_t ntiles, rmax, pad0, pad1;
    int64_t p_out, p_mid, p_suf, leaves;
    int64_t item_base;   // first work item of this side
    int64_t radix_off;   // int32 radix[L]
    int64_t tup_off;     // int8 tuples [L][rmax][r]
    int64_t c32_off;     // float [L][rmax][mt*32]
    int64_t mid_off;     // float [p_mid][mt*32]
    int64_t sufb_off;    // uint4 [ntiles][nbs][64]
    int64_t frag_off;    // uint4 [mt][2 nbs][64]
    int64_t c64_off;     // double [L][rmax][K][2]
    float thr_lo, thr_hi;
    double thr;
};

struct UncCtl {
    unsigned long long surv;
    unsigned long long band;
    unsigned long long leaves;
    unsigned int ties;
    unsigned int pad;
};

constexpr int kWaveBuf = 448;  // survivor slots per wave in LDS

__device__ __forceinline__ void flush_wave(uint64_t* wbuf, int& wcount, int lane, UncCtl* ctl,
                                           uint64_t* __restrict__ surv, int64_t cap) {
    // LDS writes of this wave must be visible to its other lanes
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(&ctl->surv, (unsigned long long)wcount);
    base = __shfl(base, 0);
    for (int i = lane; i < wcount; i += 64) {
        const unsigned long long o = base + i;
        if ((int64_t)o < cap) surv[o] = wbuf[i];
    }
    wcount = 0;
}

template <int NBS, int MT>
__device__ void unc_item(const UncSide& S, int sid, int64_t o, const int32_t* __restrict__ radix,
                         const float* __restrict__ c32, const float* __restrict__ mid, const uint4* __restrict__ sufb,
                         const uint4* __restrict__ frags, uint64_t* wbuf, int& wcount, int lane, UncCtl* ctl,
                         uint64_t* __restrict__ surv, int64_t scap, uint64_t* __restrict__ band, int64_t bcap) {
    const int h = lane >> 5, col = lane & 31;
    // A operands (suffix DFT weights) in registers
    uint4 af[MT][2 * NBS];
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int s = 0; s < 2 * NBS; ++s) af[t][s] = frags[S.frag_off + ((t * 2 * NBS) + s) * 64 + lane];
    // X_outer: per-position FP32 contributions of the outer digits
    float xo[MT][16];
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int q = 0; q < 16; ++q) xo[t][q] = 0.f;
    {
        int64_t rem = o;
        for (int j = S.a - 1; j >= 0; --j) {
            const int R = radix[S.radix_off + j];
            const int d = (int)(rem % R);
            rem /= R;
            const float* src = c32 + S.c32_off + ((int64_t)j * S.rmax + d) * (MT * 32);
#pragma unroll
            for (int t = 0; t < MT; ++t)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float4 v = *reinterpret_cast<const float4*>(src + t * 32 + 8 * q + 4 * h);
                    xo[t][4 * q + 0] += v.x;
                    xo[t][4 * q + 1] += v.y;
                    xo[t][4 * q + 2] += v.z;
                    xo[t][4 * q + 3] += v.w;
                }
        }
    }
    const float thr_lo = S.thr_lo, thr_hi = S.thr_hi;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int64_t m = 0; m < S.p_mid; ++m) {
        f32x16 c0[MT];
        const float* mrow = mid + S.mid_off + m * (MT * 32);
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float4 v = *reinterpret_cast<const float4*>(mrow + t * 32 + 8 * q + 4 * h);
                c0[t][4 * q + 0] = xo[t][4 * q + 0] + v.x;
                c0[t][4 * q + 1] = xo[t][4 * q + 1] + v.y;
                c0[t][4 * q + 2] = xo[t][4 * q + 2] + v.z;
                c0[t][4 * q + 3] = xo[t][4 * q + 3] + v.w;
            }
        const int64_t leaf0 = (o * S.p_mid + m) * S.p_suf;
        for (int tile = 0; tile < S.ntiles; ++tile) {
            uint4 bop[NBS];
#pragma unroll
            for (int b = 0; b < NBS; ++b) bop[b] = sufb[S.sufb_off + ((int64_t)tile * NBS + b) * 64 + lane];
            f32x16 acc[MT];
#pragma unroll
            for (int t = 0; t < MT; ++t) acc[t] = c0[t];
            mfma_tile_reg<NBS, MT>(af, bop, acc);
            const float mx = tile_max_psd<MT>(acc);
            const int64_t s = (int64_t)tile * 32 + col;
            const bool valid = h == 0 && s < S.p_suf;
            const bool acc_ok = valid && mx < thr_lo;
            const bool in_band = valid && !(mx > thr_hi) && !acc_ok;
            const unsigned long long ba = __ballot(acc_ok);
            const uint64_t rec = ((uint64_t)sid << 40) | (uint64_t)(leaf0 + s);
            if (ba) {
                if (acc_ok) wbuf[wcount + __popcll(ba & below)] = rec;
                wcount += __popcll(ba);
                if (wcount > kWaveBuf - 32) flush_wave(wbuf, wcount, lane, ctl, surv, scap);
            }
            if (in_band) {
                const unsigned long long i = atomicAdd(&ctl->band, 1ull);
                if ((int64_t)i < bcap) band[i] = rec;
            }
        }
    }
}

__global__ void __launch_bounds__(256) k_unc_filter(const UncSide* __restrict__ sides, int nsides,
                                                    int64_t nitems, const int32_t* __restrict__ radix,
                                                    const float* __restrict__ c32, const float* __restrict__ mid,
                                                    const uint4* __restrict__ sufb, const uint4* __restrict__ frags,
                                                    UncCtl* ctl, uint64_t* __restrict__ surv, int64_t scap,
                                                    uint64_t* __restrict__ band, int64_t bcap) {
    __shared__ uint64_t wbuf_all[4][kWaveBuf];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint64_t* wbuf = wbuf_all[wv];
    int wcount = 0;
    const int64_t wave = (int64_t)blockIdx.x * 4 + wv;
    const int64_t nwaves = (int64_t)gridDim.x * 4;
    int sid = 0;
    for (int64_t it = wave; it < nitems; it += nwaves) {
        // side of this item (items are grouped by side in order)
        while (sid + 1 < nsides && it >= sides[sid + 1].item_base) ++sid;
        while (sid > 0 && it < sides[sid].item_base) --sid;
        const UncSide& S = sides[sid];
        const int64_t o = it - S.item_base;
        const int nbs = S.nbs, mt = (S.K + 15) / 16;
        if (mt == 1) {
            if (nbs == 1) unc_item<1, 1>(S, sid, o, radix, c32, mid, sufb, frags, wbuf, wcount, lane, ctl, surv, scap, band, bcap);
            else if (nbs == 2) unc_item<2, 1>(S, sid, o, radix, c32, mid, sufb, frags, wbuf, wcount, lane, ctl, surv, scap, band, bcap);
            else if (nbs == 3) unc_item<3, 1>(S, sid, o, radix, c32, mid, sufb, frags, wbuf, wcount, lane, ctl, surv, scap, band, bcap);
            else unc_item<4, 1>(S, sid, o, radix, c32, mid, sufb, frags, wbuf, wcount, lane, ctl, surv, scap, band, bcap);
        } else {
            if (nbs == 1) unc_item<1, 2>(S, sid, o, radix, c32, mid, sufb, frags, wbuf, wcount, lane, ctl, surv, scap, band, bcap);
            else if (nbs == 2) unc_item<2, 2>(S, sid, o, radix, c32, mid, sufb, frags, wbuf, wcount, lane, ctl, surv, scap, band, bcap);
            else if (nbs == 3) unc_item<3, 2>(S, sid, o, radix, c32, mid, sufb, frags, wbuf, wcount, lane, ctl, surv, scap, band, bcap);
            else unc_item<4, 2>(S, sid, o, radix, c32, mid, sufb, frags, wbuf, wcount, lane, ctl, surv, scap, band, bcap);
        }
    }
    if (wcount) flush_wave(wbuf, wcount, lane, ctl, surv, scap);
}

// word of leaf `leaf` of side S
__device__ __forceinline__ void build_word(const UncSide& S, uint64_t leaf, const int32_t* __restrict__ radix,
                                           const int8_t* __restrict__ tup, int8_t* w) {
    for (int j = S.L - 1; j >= 0; --j) {
        const uint64_t R = (uint64_t)radix[S.radix_off + j];
        const int d = (int)(leaf % R);
        leaf /= R;
        const int8_t* t = tup + S.tup_off + ((int64_t)j * S.rmax + d) * S.r;
        for (int i = 0; i < S.r; ++i) w[j + S.L * i] = t[i];
    }
}

// FP64 recheck of band candidates; accepted ones join the survivors.
__global__ void k_unc_band(const UncSide* __restrict__ sides, const int32_t* __restrict__ radix,
                           const int8_t* __restrict__ tup, const double* __restrict__ trig64_all,
                           const int64_t* __restrict__ trig_off_of_n, const uint64_t* __restrict__ band, UncCtl* ctl,
                           uint64_t* __restrict__ surv, int64_t scap, int64_t bcount) {
    int8_t w[kMaxN];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < bcount; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t rec = band[i];
        const int sid = (int)(rec >> 40);
        const UncSide& S = sides[sid];
        build_word(S, rec & ((1ull << 40) - 1), radix, tup, w);
        const double* c = trig64_all + trig_off_of_n[0];
        const double m = fp64_max_psd(w, S.N, c, c + S.N);
        if (fabs(m - S.thr) < 1e-9) atomicAdd(&ctl->ties, 1u);
        if (m <= S.thr) {
            const unsigned long long o = atomicAdd(&ctl->surv, 1ull);
            if ((int64_t)o < scap) surv[o] = rec;
        }
    }
}

// Survivor -> word (N int8) + FP64 key digits (K uint8).
__global__ void k_unc_keys(const UncSide* __restrict__ sides, const int32_t* __restrict__ radix,
                           const int8_t* __restrict__ tup, const double* __restrict__ c64,
                           const uint64_t* __restrict__ surv, int64_t m, int bound, int8_t* __restrict__ words,
                           uint8_t* __restrict__ keys, UncCtl* ctl) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t rec = surv[i];
        const int sid = (int)(rec >> 40);
        const UncSide& S = sides[sid];
        uint64_t leaf = rec & ((1ull << 40) - 1);
        int8_t* w = words + i * S.N;
        double re[kMaxK], im[kMaxK];
        for (int k = 0; k < S.K; ++k) { re[k] = 0.0; im[k] = 0.0; }
        for (int j = S.L - 1; j >= 0; --j) {
            const uint64_t R = (uint64_t)radix[S.radix_off + j];
            const int d = (int)(leaf % R);
            leaf /= R;
            const int8_t* t = tup + S.tup_off + ((int64_t)j * S.rmax + d) * S.r;
            for (int q = 0; q < S.r; ++q) w[j + S.L * q] = t[q];
            const double* cc = c64 + S.c64_off + ((int64_t)j * S.rmax + d) * (2 * S.K);
            for (int k = 0; k < S.K; ++k) { re[k] += cc[2 * k]; im[k] += cc[2 * k + 1]; }
        }
        unsigned int ties = 0;
        for (int k = 0; k < S.K; ++k) {
            const double p = re[k] * re[k] + im[k] * im[k];
            const double f = p - floor(p);
            if (fabs(f - 0.5) < 1e-9) ++ties;
            const int v = (int)rint(p);
            keys[i * S.K + k] = (uint8_t)(S.seqflag ? v : bound - v);
        }
        if (ties) atomicAdd(&ctl->ties, ties);
    }
}

// ---------------------------------------------------------------------------
// host driver

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

struct Tables {
    std::vector<UncSide> sides;
    std::vector<int32_t> radix;
    std::vector<int8_t> tup;
    std::vector<float> c32, mid;
    std::vector<uint4> sufb, frags;
    std::vector<double> c64;
    int64_t items = 0;
};

void add_side(Tables& T, const SideHost& S, int rmax, int order, int paf, const std::vector<double>& cs,
              const std::vector<double>& sn, double& s1_out) {
    UncSide D{};
    D.L = S.L; D.r = S.r; D.N = S.N; D.K = S.N / 2; D.seqflag = S.seqflag; D.rmax = rmax;
    const int L = S.L, r = S.r, N = S.N, K = D.K;
    const int mt = (K + 15) / 16;
    std::vector<int64_t> R(L);
    for (int j = 0; j < L; ++j) R[j] = (int64_t)S.T[j].size();
    // split: suffix [b, L) with p_suf >= 64 (or everything), mid [a, b) with p_mid >= 16 (<= 4096)
    int b = L;
    int64_t ps = 1;
    while (b > 0 && ps < 64 && (L - (b - 1)) * r <= 64) { --b; ps *= R[b]; }
    int a = b;
    int64_t pm = 1;
    while (a > 0 && pm < 16 && pm * R[a - 1] <= 4096) { --a; pm *= R[a]; }
    int64_t po = 1;
    for (int j = 0; j < a; ++j) po *= R[j];
    D.a = a; D.b = b; D.p_suf = ps; D.p_mid = pm; D.p_out = po; D.leaves = (int64_t)S.leaves;
    D.nbs = std::max(1, ((L - b) * r + 15) / 16);
    D.ntiles = (int)((ps + 31) / 32);
    D.item_base = T.items;
    T.items += po;
    // radices + tuples
    D.radix_off = (int64_t)T.radix.size();
    for (int j = 0; j < L; ++j) T.radix.push_back((int32_t)R[j]);
    D.tup_off = (int64_t)T.tup.size();
    T.tup.resize(T.tup.size() + (size_t)L * rmax * r, 0);
    for (int j = 0; j < L; ++j)
        for (int64_t d = 0; d < R[j]; ++d)
            for (int i = 0; i < r; ++i) T.tup[D.tup_off + ((int64_t)j * rmax + d) * r + i] = (int8_t)S.T[j][d][i];
    // FP64 per-position contributions (k = 1..K; Re, Im with +sin) and FP32 channel copies
    D.c64_off = (int64_t)T.c64.size();
    T.c64.resize(T.c64.size() + (size_t)L * rmax * 2 * K, 0.0);
    D.c32_off = (int64_t)T.c32.size();
    T.c32.resize(T.c32.size() + (size_t)L * rmax * mt * 32, 0.f);
    std::vector<double> contrib((size_t)L * rmax * 2 * K, 0.0);
    double s1 = 0;
    for (int j = 0; j < L; ++j) {
        double mxabs = 0;
        for (int64_t d = 0; d < R[j]; ++d) {
            double ab = 0;
            for (int i = 0; i < r; ++i) ab += std::abs(S.T[j][d][i]);
            mxabs = std::max(mxabs, ab);
            for (int k = 1; k <= K; ++k) {
                double re = 0, im = 0;
                for (int i = 0; i < r; ++i) {
                    const int p = j + L * i;
                    const int mm = (int)(((int64_t)k * p) % N);
                    re += S.T[j][d][i] * cs[mm];
                    im += S.T[j][d][i] * sn[mm];
                }
                double* c = &T.c64[D.c64_off + (((int64_t)j * rmax + d) * K + (k - 1)) * 2];
                c[0] = re; c[1] = im;
                const int t = (k - 1) / 16, row = (k - 1) % 16;
                T.c32[D.c32_off + ((int64_t)j * rmax + d) * mt * 32 + t * 32 + row] = (float)re;
                T.c32[D.c32_off + ((int64_t)j * rmax + d) * mt * 32 + t * 32 + 16 + row] = (float)im;
            }
        }
        s1 += mxabs;
    }
    s1_out = s1;
    // mid table: FP64 sums over the middle digits, rounded once to FP32
    D.mid_off = (int64_t)T.mid.size();
    T.mid.resize(T.mid.size() + (size_t)pm * mt * 32, 0.f);
    {
        std::vector<double> acc((size_t)2 * K);
        for (int64_t m = 0; m < pm; ++m) {
            std::fill(acc.begin(), acc.end(), 0.0);
            int64_t rem = m;
            for (int j = b - 1; j >= a; --j) {
                const int d = (int)(rem % R[j]);
                rem /= R[j];
                const double* c = &T.c64[D.c64_off + ((int64_t)j * rmax + d) * K * 2];
                for (int q = 0; q < 2 * K; ++q) acc[q] += c[q];
            }
            for (int k = 1; k <= K; ++k) {
                const int t = (k - 1) / 16, row = (k - 1) % 16;
                T.mid[D.mid_off + m * mt * 32 + t * 32 + row] = (float)acc[2 * (k - 1)];
                T.mid[D.mid_off + m * mt * 32 + t * 32 + 16 + row] = (float)acc[2 * (k - 1) + 1];
            }
        }
    }
    // suffix B operands (BF16 words of the suffix combinations)
    D.sufb_off = (int64_t)T.sufb.size();
    const int nent = (L - b) * r;
    for (int tile = 0; tile < D.ntiles; ++tile)
        for (int blk = 0; blk < D.nbs; ++blk)
            for (int lane = 0; lane < 64; ++lane) {
                uint16_t hv[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                const int64_t s = (int64_t)tile * 32 + lane % 32;
                if (s < ps) {
                    // digits of suffix combination s
                    std::vector<int> dg(L, 0);
                    int64_t rem = s;
                    for (int j = L - 1; j >= b; --j) { dg[j] = (int)(rem % R[j]); rem /= R[j]; }
                    for (int i8 = 0; i8 < 8; ++i8) {
                        const int e = 16 * blk + 8 * (lane / 32) + i8;
                        if (e >= nent) continue;
                        const int j = b + e / r, i = e % r;
                        hv[i8] = bf16_rne_u((float)S.T[j][dg[j]][i]);
                    }
                }
                uint4 q;
                q.x = hv[0] | ((uint32_t)hv[1] << 16);
                q.y = hv[2] | ((uint32_t)hv[3] << 16);
                q.z = hv[4] | ((uint32_t)hv[5] << 16);
                q.w = hv[6] | ((uint32_t)hv[7] << 16);
                T.sufb.push_back(q);
            }
    // A fragments: suffix DFT weights, BF16 hi and lo
    D.frag_off = (int64_t)T.frags.size();
    double rho = 0;
    for (int t = 0; t < mt; ++t)
        for (int st = 0; st < 2 * D.nbs; ++st)
            for (int lane = 0; lane < 64; ++lane) {
                uint16_t hv[8];
                for (int i8 = 0; i8 < 8; ++i8) {
                    const int row = lane % 32;
                    const int e = 16 * (st % D.nbs) + 8 * (lane / 32) + i8;
                    const int part = st / D.nbs;
                    const int k = 16 * t + row % 16 + 1;
                    uint16_t v = 0;
                    if (e < nent && k <= K) {
                        const int j = b + e / r, i = e % r;
                        const int p = j + L * i;
                        const int mm = (int)(((int64_t)k * p) % N);
                        const double w = row >= 16 ? sn[mm] : cs[mm];
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
                T.frags.push_back(q);
            }
    // error bound: prefix FP32 sums + suffix split-BF16 contraction
    double s1_pre = 0, s1_suf = 0;
    for (int j = 0; j < L; ++j) {
        double mxabs = 0;
        for (int64_t d = 0; d < R[j]; ++d) {
            double ab = 0;
            for (int i = 0; i < r; ++i) ab += std::abs(S.T[j][d][i]);
            mxabs = std::max(mxabs, ab);
        }
        (j < b ? s1_pre : s1_suf) += mxabs;
    }
    const double u24 = 0x1p-24, u23 = 0x1p-23;
    const double E = (a + 3) * u24 * s1_pre * 1.01                       // X_outer + Mid rounding
                     + s1_suf * rho * 1.0000001                          // BF16 split residual
                     + 1.004 * (32.0 * D.nbs + 1.0) * u23 * (s1_pre + s1_suf);  // MFMA accumulation
    const double thr = 2.0 * order - paf + 0.001;
    const double eps = psd_eps(E, thr);
    D.thr = thr;
    D.thr_hi = std::nextafter((float)(thr + eps), INFINITY);
    D.thr_lo = std::nextafter((float)(thr - eps), -INFINITY);
    T.sides.push_back(D);
}

template <class V>
void* upload(gp_ctx* ctx, const char* name, const std::vector<V>& v) {
    void* p = ctx->buf(name).ensure(v.size() * sizeof(V) + 64);
    if (!v.empty()) GP_HIP(hipMemcpyAsync(p, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice, ctx->stream));
    return p;
}

}  // namespace

// Expands, filters and keys every side; survivors come back grouped per side
// in DFS order as device arrays (words N bytes, keys K bytes per row).
UncResult expand_sides(gp_ctx* ctx, const std::vector<SideHost>& sides, int order, int paf, gp_stats* stats) {
    UncResult res;
    if (sides.empty()) return res;
    hipStream_t st = ctx->stream;
    const int N = sides[0].N, K = N / 2;
    if (N > kMaxN || N < 2) throw ArgError("uncompress: expanded length outside [2, 64]");
    if ((int64_t)sides.size() >= (1ll << 23)) throw ArgError("uncompress: too many sequences in one batch");
    const int bound = 2 * order - paf;
    if (bound < 0 || bound > 255) throw ArgError("uncompress: keys need 0 <= 2*order - paf <= 255");
    int rmax = 1;
    for (const auto& s : sides)
        for (const auto& t : s.T) rmax = std::max<int>(rmax, (int)t.size());
    // FP64 twiddles of length N (symmetric)
    const int table = ctx->tables.get(N);
    ctx->tables.upload(st);
    const PsdTableDesc& td = ctx->tables.desc[table];
    std::vector<double> cs(ctx->tables.trig64.begin() + td.trig_off, ctx->tables.trig64.begin() + td.trig_off + N);
    std::vector<double> sn(ctx->tables.trig64.begin() + td.trig_off + N,
                           ctx->tables.trig64.begin() + td.trig_off + 2 * N);
    Tables T;
    uint64_t total_leaves = 0;
    std::vector<int> side_of;  // table index -> input side
    for (size_t i = 0; i < sides.size(); ++i) {
        if (sides[i].leaves == 0) continue;
        double s1;
        add_side(T, sides[i], rmax, order, paf, cs, sn, s1);
        side_of.push_back((int)i);
        total_leaves += sides[i].leaves;
    }
    res.side_begin.assign(sides.size() + 1, 0);
    if (T.sides.empty()) return res;
    auto* d_sides = static_cast<UncSide*>(upload(ctx, "u_sides", T.sides));
    auto* d_radix = static_cast<int32_t*>(upload(ctx, "u_radix", T.radix));
    auto* d_tup = static_cast<int8_t*>(upload(ctx, "u_tup", T.tup));
    auto* d_c32 = static_cast<float*>(upload(ctx, "u_c32", T.c32));
    auto* d_mid = static_cast<float*>(upload(ctx, "u_mid", T.mid));
    auto* d_sufb = static_cast<uint4*>(upload(ctx, "u_sufb", T.sufb));
    auto* d_frags = static_cast<uint4*>(upload(ctx, "u_frags", T.frags));
    auto* d_c64 = static_cast<double*>(upload(ctx, "u_c64", T.c64));
    std::vector<int64_t> toff{td.trig_off};
    auto* d_toff = static_cast<int64_t*>(upload(ctx, "u_toff", toff));
    UncCtl* ctl = ctx->dev<UncCtl>("u_ctl", 1);
    int64_t scap = (int64_t)std::min<uint64_t>(total_leaves / 4 + (1 << 20), (uint64_t)1 << 30);
    int64_t bcap = (int64_t)std::min<uint64_t>(total_leaves / 64 + (1 << 16), (uint64_t)1 << 28);
    UncCtl h{};
    float ms_filter = 0, ms_band = 0, ms_keys = 0;
    for (int attempt = 0; attempt < 2; ++attempt) {
        uint64_t* d_surv = ctx->dev<uint64_t>("u_surv", (size_t)scap);
        uint64_t* d_band = ctx->dev<uint64_t>("u_band", (size_t)bcap);
        GP_HIP(hipMemsetAsync(ctl, 0, sizeof(UncCtl), st));
        GP_HIP(hipEventRecord(ctx->ev0, st));
        const int64_t waves = std::min<int64_t>(T.items, (int64_t)ctx->num_cus * 8);
        hipLaunchKernelGGL(k_unc_filter, dim3((unsigned)std::max<int64_t>(1, (waves + 3) / 4)), dim3(256), 0, st,
                           d_sides, (int)T.sides.size(), T.items, d_radix, d_c32, d_mid, d_sufb, d_frags, ctl, d_surv,
                           scap, d_band, bcap);
        GP_HIP(hipGetLastError());
        GP_HIP(hipEventRecord(ctx->ev1, st));
        GP_HIP(hipMemcpyAsync(&h, ctl, sizeof h, hipMemcpyDeviceToHost, st));
        GP_HIP(hipStreamSynchronize(st));
        GP_HIP(hipEventElapsedTime(&ms_filter, ctx->ev0, ctx->ev1));
        if ((int64_t)h.band > bcap) { bcap = (int64_t)h.band + 1024; continue; }
        if (h.band) {
            hipLaunchKernelGGL(k_unc_band, dim3(ctx->num_cus * 2), dim3(128), 0, st, d_sides, d_radix, d_tup,
                               ctx->tables.d_trig64.as<double>(), d_toff, d_band, ctl, d_surv, scap, (int64_t)h.band);
            GP_HIP(hipGetLastError());
            GP_HIP(hipEventRecord(ctx->ev2, st));
            GP_HIP(hipMemcpyAsync(&h, ctl, sizeof h, hipMemcpyDeviceToHost, st));
            GP_HIP(hipStreamSynchronize(st));
            GP_HIP(hipEventElapsedTime(&ms_band, ctx->ev1, ctx->ev2));
        }
        if ((int64_t)h.surv > scap) {
            scap = (int64_t)h.surv + 1024;
            bcap = std::max<int64_t>(bcap, (int64_t)h.band + 1024);
            continue;
        }
        break;
    }
    const int64_t m = (int64_t)h.surv;
    // group per side, DFS order
    uint64_t* d_surv = ctx->buf("u_surv").as<uint64_t>();
    if (m > 1) {
        uint32_t* vals = ctx->dev<uint32_t>("u_sv", (size_t)m);
        uint64_t* tk = ctx->dev<uint64_t>("u_tk", (size_t)m);
        uint32_t* tv = ctx->dev<uint32_t>("u_tv", (size_t)m);
        uint32_t* hist = static_cast<uint32_t*>(ctx->buf("u_hist").ensure(radix_hist_bytes(m)));
        GP_HIP(hipMemsetAsync(vals, 0, (size_t)m * 4, st));
        int top = 40;
        while ((1ll << (top - 40)) < (int64_t)T.sides.size()) ++top;
        radix_sort_pairs(st, d_surv, vals, tk, tv, hist, m, 0, ((top + 7) / 8) * 8);
    }
    int8_t* d_words = ctx->dev<int8_t>("u_words", (size_t)std::max<int64_t>(m, 1) * N);
    uint8_t* d_keys = ctx->dev<uint8_t>("u_keys", (size_t)std::max<int64_t>(m, 1) * K);
    GP_HIP(hipEventRecord(ctx->ev2, st));
    if (m > 0)
        hipLaunchKernelGGL(k_unc_keys, dim3(ctx->num_cus * 4), dim3(128), 0, st, d_sides, d_radix, d_tup, d_c64, d_surv,
                           m, bound, d_words, d_keys, ctl);
    GP_HIP(hipGetLastError());
    GP_HIP(hipEventRecord(ctx->ev3, st));
    std::vector<uint64_t> recs(m);
    if (m) GP_HIP(hipMemcpyAsync(recs.data(), d_surv, (size_t)m * 8, hipMemcpyDeviceToHost, st));
    GP_HIP(hipMemcpyAsync(&h, ctl, sizeof h, hipMemcpyDeviceToHost, st));
    GP_HIP(hipStreamSynchronize(st));
    GP_HIP(hipEventElapsedTime(&ms_keys, ctx->ev2, ctx->ev3));
    // side ranges
    std::vector<int64_t> cnt(sides.size(), 0);
    for (uint64_t r : recs) cnt[side_of[r >> 40]]++;
    for (size_t i = 0; i < sides.size(); ++i) res.side_begin[i + 1] = res.side_begin[i] + cnt[i];
    res.leaf.resize(m);
    for (int64_t i = 0; i < m; ++i) res.leaf[i] = recs[i] & ((1ull << 40) - 1);
    res.d_words = d_words;
    res.d_keys = d_keys;
    res.N = N;
    res.K = K;
    res.filter_ms = ms_filter;
    if (stats) {
        stats->words += total_leaves;
        stats->psd_decisions += total_leaves;
        stats->psd_pass += (uint64_t)m;
        stats->band_rechecks += h.band;
        stats->tie_band += h.ties;
        stats->psd_kernel_ms += ms_filter;
        stats->band_kernel_ms += ms_band;
        stats->psd_launches += 1;
        stats->kernel_ms += ms_filter + ms_band + ms_keys;
    }
    return res;
}

}  // namespace gp
