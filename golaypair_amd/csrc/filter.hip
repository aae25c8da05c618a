// Equivalence filter (replaces cache_filter, src/core/filter/filter.cpp:18-163).
//
// The reference builds the pair group as the closure of an index template under
// shift_pair (rotate b), swap_pair, reverse_pair (reverse b), decimate_pair (the
// same unit on a and b) and, when COMPRESS == 1 and LEN is even, altnegative_pair
// (negate the odd positions of a and b) (equivalence.cpp:46-95,146-267,
// filter.cpp:80-87).  Every element of that closure acts as
//   (X, Y) = swap ? (b, a) : (a, b)
//   a'[i] = sa * (-1)^(kappa i) * X[(u i + ta) mod n]
//   b'[i] = sb * (-1)^(kappa i) * Y[(e u i + tb) mod n]
// with u a unit mod n, e = +-1, ta, tb in Z_n, and (kappa, sa, sb) = (0, +1, +1)
// without altnegative, free in {0,1} x {+-1}^2 with it; the set of transforms the
// reference applies in generateClassPairs (equivalence.cpp:97-144, including
// its swapped-template branch) is this whole group, so its classes are orbits.
// Walking the pairs in (a, b) order and erasing each survivor's class
// (filter.cpp:117-144) therefore keeps the lexicographically smallest input
// member of every orbit.  Here every pair's orbit minimum is computed on the GPU
// (one pair per wave, as an order-preserving packed key); pairs are then grouped
// by orbit minimum on the host.
#include <algorithm>
#include <cstring>
#include <functional>
#include <numeric>

#include "gp_host.hpp"
#include "gp_internal.hpp"
#include "linesort.hpp"

namespace gp {

struct FilterParams {
    int32_t n;
    int32_t alt;
    int32_t nunits;
    int32_t units[64];
    int32_t mx, sh, B, E;  // value v packs as (v + mx) >> sh in B bits, E = 64 / B values per word
};

// Packed key of the image s * (-1)^(kappa i) * X[(u i + t) mod n], i = 0..n-1:
// B bits per value, first value in the most significant bits of word 0, so
// that unsigned word-by-word order is the signed lexicographic order of images.
template <int W>
__device__ __forceinline__ void img_key(const int8_t* X, const FilterParams& P, int u, int t, int kappa, int sg,
                                        uint64_t (&k)[W]) {
    const int n = P.n;
    int idx = t, i = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        uint64_t acc = 0;
        const int iend = min(n, (w + 1) * P.E);
        const int cnt = max(0, iend - i);
        for (; i < iend; ++i) {
            int v = X[idx] * sg;
            if (kappa & i & 1) v = -v;
            acc = (acc << P.B) | (uint64_t)((v + P.mx) >> P.sh);
            idx += u;
            if (idx >= n) idx -= n;
        }
        k[w] = cnt ? acc << (P.B * (P.E - cnt)) : 0ull;
    }
}

template <int W>
__device__ __forceinline__ bool key_less(const uint64_t (&x)[W], const uint64_t (&y)[W]) {
#pragma unroll
    for (int w = 0; w < W; ++w)
        if (x[w] != y[w]) return x[w] < y[w];
    return false;
}

template <int W>
__device__ __forceinline__ bool key_eq(const uint64_t (&x)[W], const uint64_t (&y)[W]) {
    bool e = true;
#pragma unroll
    for (int w = 0; w < W; ++w) e = e && x[w] == y[w];
    return e;
}

// minimum over the 64 lanes of the wave, returned to every lane
template <int W>
__device__ __forceinline__ void wave_min(uint64_t (&k)[W]) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        uint64_t o[W];
#pragma unroll
        for (int w = 0; w < W; ++w) o[w] = __shfl_xor(k[w], off);
        if (key_less(o, k))
#pragma unroll
            for (int w = 0; w < W; ++w) k[w] = o[w];
    }
}

// Least rotation (the start r of the lexicographically smallest rotation) of
// the lane's sequence z[0..n) held in LDS at z[j * 64] (signed order); the
// two-pointer minimum-expression method, O(n).
__device__ __forceinline__ int least_rotation(const int8_t* z, int n) {
    int i = 0, j = 1, k = 0;
    while (i < n && j < n && k < n) {
        int x = i + k, y = j + k;
        if (x >= n) x -= n;
        if (y >= n) y -= n;
        const int a = z[x * 64], b = z[y * 64];
        if (a == b) {
            ++k;
            continue;
        }
        if (a > b) i += k + 1;
        else j += k + 1;
        if (i == j) ++j;
        k = 0;
    }
    return i < j ? i : j;
}

// Writes z[j] = sigma * (-1)^(kappa j) * X[(u j) mod n] (lane's column of the
// LDS sequence area); kappa is used only for even n (altnegative).
__device__ __forceinline__ void decimated(const int8_t* X, int n, int u, int kappa, int sigma, int8_t* z) {
    int idx = 0;
    for (int j = 0; j < n; ++j) {
        int v = X[idx] * sigma;
        if (kappa & j & 1) v = -v;
        z[j * 64] = (int8_t)v;
        idx += u;
        if (idx >= n) idx -= n;
    }
}

// Packed key of the rotation of z starting at r (layout of img_key).
template <int W>
__device__ __forceinline__ void rot_key(const int8_t* z, const FilterParams& P, int r, uint64_t (&k)[W]) {
    const int n = P.n;
    int idx = r, i = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        uint64_t acc = 0;
        const int iend = min(n, (w + 1) * P.E);
        const int cnt = max(0, iend - i);
        for (; i < iend; ++i) {
            acc = (acc << P.B) | (uint64_t)((z[idx * 64] + P.mx) >> P.sh);
            if (++idx == n) idx = 0;
        }
        k[w] = cnt ? acc << (P.B * (P.E - cnt)) : 0ull;
    }
}

// One pair per wave; the orbit minimum by construction.  Every transform acts
// as image[i] = s (-1)^(kappa i) X[(u i + t) mod n] on each half.  With
// t = u s' this is a rotation by s' of z[j] = (-1)^(kappa j) X[(u j) mod n]
// (for kappa = 1, n is even and (-1)^(kappa i) = (-1)^(s') (-1)^(kappa (i + s')),
// so the rotation's sign is s (-1)^(s'), both signs reachable since s is free
// whenever kappa is).  So for each (swap, u, kappa, sign) the minimum over t
// is the least rotation of +-z.  Phase 1: lanes over the a-part variants
// (swap, u, kappa, sign); the wave minimum is the a key.  Phase 2: every lane
// whose variant reaches it takes the least rotations of its b-part variants
// (unit e u with e = +-1, the same kappa, sign free with altnegative), which
// do not depend on the a-part's shift or sign; the wave minimum is the b key.
// canon[q] = (a key, b key), as the minimum over the whole group.
template <int W>
__global__ void __launch_bounds__(256) k_orbit_min(FilterParams P, const int8_t* __restrict__ pa,
                                                   const int8_t* __restrict__ pb, int64_t m,
                                                   uint64_t* __restrict__ canon) {
    __shared__ int8_t sx[4][2][kMaxN];
    __shared__ int8_t zs[4][kMaxN * 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int n = P.n;
    const int nk = P.alt ? 2 : 1, ns = P.alt ? 2 : 1;
    const int nvar = 2 * P.nunits * nk * ns;
    int8_t* z = &zs[wv][lane];
    for (int64_t q = (int64_t)blockIdx.x * 4 + wv; q < m; q += (int64_t)gridDim.x * 4) {
        if (lane < n) {
            sx[wv][0][lane] = pa[q * n + lane];
            sx[wv][1][lane] = pb[q * n + lane];
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // variant c -> (swap, unit index, kappa, sign)
        auto decode = [&](int c, int& sw, int& ui, int& kappa, int& sg) {
            sg = (c % ns) ? -1 : 1;
            c /= ns;
            kappa = c % nk;
            c /= nk;
            ui = c % P.nunits;
            sw = c / P.nunits;
        };
        uint64_t ka[W], kb[W], lmin[W];
#pragma unroll
        for (int w = 0; w < W; ++w) ka[w] = kb[w] = ~0ull;
        for (int c = lane; c < nvar; c += 64) {
            int sw, ui, kappa, sg;
            decode(c, sw, ui, kappa, sg);
            decimated(sx[wv][sw], n, P.units[ui], kappa, sg, z);
            uint64_t k[W];
            rot_key<W>(z, P, least_rotation(z, n), k);
            if (key_less(k, ka))
#pragma unroll
                for (int w = 0; w < W; ++w) ka[w] = k[w];
        }
#pragma unroll
        for (int w = 0; w < W; ++w) lmin[w] = ka[w];
        wave_min<W>(ka);
        const bool mine = key_eq(lmin, ka);
        // phase 2: the lanes holding a minimising variant (re-found by recomputing)
        for (int c = lane; c < nvar && mine; c += 64) {
            int sw, ui, kappa, sg;
            decode(c, sw, ui, kappa, sg);
            decimated(sx[wv][sw], n, P.units[ui], kappa, sg, z);
            uint64_t k[W];
            rot_key<W>(z, P, least_rotation(z, n), k);
            if (!key_eq(k, ka)) continue;
            const int u = P.units[ui];
            for (int e = 0; e < 2; ++e) {
                const int ub = e ? (n - u) % n : u;
                for (int sb = 0; sb < ns; ++sb) {
                    decimated(sx[wv][sw ? 0 : 1], n, n == 1 ? 0 : ub, kappa, sb ? -1 : 1, z);
                    rot_key<W>(z, P, least_rotation(z, n), k);
                    if (key_less(k, kb))
#pragma unroll
                        for (int w = 0; w < W; ++w) kb[w] = k[w];
                }
            }
        }
        wave_min<W>(kb);
        if (lane == 0) {
#pragma unroll
            for (int w = 0; w < W; ++w) {
                canon[q * 2 * W + w] = ka[w];
                canon[q * 2 * W + W + w] = kb[w];
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

int filter_pairs(gp_ctx* ctx, int order, int compress, const PairSet& in, PairSet& out, gp_stats* stats) {
    const int n = order / compress;
    out.n = n;
    out.a.clear();
    out.b.clear();
    const int64_t m = in.size();
    if (m == 0) return GP_OK;
    if (in.n != n) throw ArgError("filter: pair length != order/compress");
    if (n > kMaxN) throw ArgError("filter: length above 64");
    Phase ph("filter: packing setup");
    FilterParams P{};
    P.n = n;
    P.alt = !(compress > 1 || n % 2 == 1);  // filter.cpp:83-87
    for (int u = 1; u < n || (n == 1 && u == 1); ++u)
        if (std::gcd(u, n) == 1) P.units[P.nunits++] = u;
    if (n == 1) { P.nunits = 1; P.units[0] = 0; }
    // key packing: the values and their negations lie in [-mx, mx]; halve when all have mx's parity
    int mx = 0, odd_or = 0, odd_and = 1;  // max |v|; whether some / every value is odd
    {
        const int nc = 16;
        std::vector<int> cmx(nc, 0), cor(nc, 0), cand(nc, 1);
        const int64_t tot = m * n;
        pool().parallel_for(nc, [&](size_t c) {
            int lmx = 0, lor = 0, land = 1;
            for (int64_t i = tot * (int64_t)c / nc; i < tot * (int64_t)(c + 1) / nc; ++i) {
                const int a = in.a[i], b = in.b[i];
                lmx = std::max(lmx, std::max(a < 0 ? -a : a, b < 0 ? -b : b));
                lor |= (a | b) & 1;
                land &= a & b & 1;
            }
            cmx[c] = lmx;
            cor[c] = lor;
            cand[c] = land;
        });
        for (int c = 0; c < nc; ++c) {
            mx = std::max(mx, cmx[c]);
            odd_or |= cor[c];
            odd_and &= cand[c];
        }
    }
    // every v + mx even: all values odd with mx odd, or all even with mx even
    const bool same_parity = (mx & 1) ? odd_and == 1 : odd_or == 0;
    P.mx = mx;
    P.sh = same_parity ? 1 : 0;
    const int range = (2 * mx) >> P.sh;
    P.B = range <= 1 ? 1 : range <= 3 ? 2 : range <= 15 ? 4 : 8;
    P.E = 64 / P.B;
    const int wneed = (n + P.E - 1) / P.E;
    const int W = wneed <= 1 ? 1 : wneed <= 2 ? 2 : wneed <= 4 ? 4 : 8;
    hipStream_t st = ctx->stream;
    ph.next("filter: device");
    int8_t* da = ctx->dev<int8_t>("f_a", (size_t)m * n);
    int8_t* db = ctx->dev<int8_t>("f_b", (size_t)m * n);
    uint64_t* dc = ctx->dev<uint64_t>("f_c", (size_t)m * 2 * W);
    GP_HIP(hipMemcpyAsync(da, in.a.data(), (size_t)m * n, hipMemcpyHostToDevice, st));
    GP_HIP(hipMemcpyAsync(db, in.b.data(), (size_t)m * n, hipMemcpyHostToDevice, st));
    GP_HIP(hipEventRecord(ctx->ev0, st));
    const dim3 grid((unsigned)std::min<int64_t>((m + 3) / 4, (int64_t)ctx->num_cus * 16));
    if (W == 1) hipLaunchKernelGGL(k_orbit_min<1>, grid, dim3(256), 0, st, P, da, db, m, dc);
    else if (W == 2) hipLaunchKernelGGL(k_orbit_min<2>, grid, dim3(256), 0, st, P, da, db, m, dc);
    else if (W == 4) hipLaunchKernelGGL(k_orbit_min<4>, grid, dim3(256), 0, st, P, da, db, m, dc);
    else hipLaunchKernelGGL(k_orbit_min<8>, grid, dim3(256), 0, st, P, da, db, m, dc);
    GP_HIP(hipGetLastError());
    GP_HIP(hipEventRecord(ctx->ev1, st));
    std::vector<uint64_t> canon((size_t)m * 2 * W);
    GP_HIP(hipMemcpyAsync(canon.data(), dc, canon.size() * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    GP_HIP(hipStreamSynchronize(st));
    float ms = 0;
    GP_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (stats) stats->kernel_ms += ms;
    ph.next("filter: group + order");
    // Pairs with equal canonical keys form one class; its survivor is the pair
    // that comes first in (a, b) lexicographic order (golay.h:25-31).  Both orders
    // use packed big-endian keys (byte v + 128 per entry, a then b).
    // one row per pair: the canonical key (2W words) then the pair key (PW
    // words), so that (class, pair order) is one lexicographic word compare
    const int PW = (2 * n + 7) / 8, CW = 2 * W, RW = CW + PW;
    std::vector<uint64_t> rk((size_t)m * RW);
    parallel_chunks(m, [&](int64_t i0, int64_t i1) {
        for (int64_t i = i0; i < i1; ++i) {
            uint64_t* r = &rk[i * RW];
            for (int w = 0; w < CW; ++w) r[w] = canon[i * CW + w];
            const int8_t* pa = &in.a[i * n];
            const int8_t* pb = &in.b[i * n];
            for (int w = 0; w < PW; ++w) {
                uint64_t acc = 0;
                for (int j = 8 * w; j < 8 * w + 8; ++j) {
                    const int v = j < n ? pa[j] : j < 2 * n ? pb[j - n] : -128;  // -128 packs as 0
                    acc = (acc << 8) | (uint64_t)(uint8_t)(v + 128);
                }
                r[CW + w] = acc;
            }
        }
    }, 1024);
    auto words_less = [&](int64_t x, int64_t y, int w0, int w1) {
        for (int w = w0; w < w1; ++w)
            if (rk[x * RW + w] != rk[y * RW + w]) return rk[x * RW + w] < rk[y * RW + w];
        return false;
    };
    auto pair_less = [&](int64_t x, int64_t y) { return words_less(x, y, CW, RW); };
    auto canon_eq = [&](int64_t x, int64_t y) {
        for (int w = 0; w < CW; ++w)
            if (rk[x * RW + w] != rk[y * RW + w]) return false;
        return true;
    };
    // Classes by hashing the canonical key, each keeping its least pair: rows go
    // to 16 partitions by hash (a counting pass), each partition has its own
    // open-addressing table on a host thread; the kept pairs are then sorted.
    constexpr int kParts = 16;
    std::vector<uint64_t> hsh((size_t)m);
    parallel_chunks(m, [&](int64_t i0, int64_t i1) {
        for (int64_t i = i0; i < i1; ++i) {
            uint64_t h = 0x9e3779b97f4a7c15ull;
            for (int w = 0; w < CW; ++w) h = (h ^ rk[i * RW + w]) * 0xff51afd7ed558ccdull;
            hsh[i] = h ^ (h >> 29);
        }
    }, 4096);
    std::vector<int64_t> part_off(kParts + 1, 0), part_rows((size_t)m);
    for (int64_t i = 0; i < m; ++i) ++part_off[(hsh[i] >> 60) + 1];
    for (int q = 0; q < kParts; ++q) part_off[q + 1] += part_off[q];
    {
        std::vector<int64_t> fill(part_off.begin(), part_off.end() - 1);
        for (int64_t i = 0; i < m; ++i) part_rows[fill[hsh[i] >> 60]++] = i;
    }
    std::vector<std::vector<int64_t>> part_keep(kParts);
    pool().parallel_for(kParts, [&](size_t q) {
        const int64_t r0 = part_off[q], r1 = part_off[q + 1];
        size_t T = 16;
        while (T < 2 * (size_t)(r1 - r0)) T <<= 1;
        std::vector<int64_t> slot(T, -1);
        for (int64_t r = r0; r < r1; ++r) {
            const int64_t i = part_rows[r];
            size_t s = hsh[i] & (T - 1);
            while (slot[s] >= 0 && !canon_eq(slot[s], i)) s = (s + 1) & (T - 1);
            if (slot[s] < 0 || pair_less(i, slot[s])) slot[s] = i;
        }
        for (size_t s = 0; s < T; ++s)
            if (slot[s] >= 0) part_keep[q].push_back(slot[s]);
    });
    std::vector<int64_t> keep;
    for (const auto& k : part_keep) keep.insert(keep.end(), k.begin(), k.end());
    parallel_sort(keep, pair_less);
    out.a.resize(keep.size() * n);
    out.b.resize(keep.size() * n);
    for (size_t r = 0; r < keep.size(); ++r) {
        std::memcpy(&out.a[r * n], &in.a[keep[r] * n], n);
        std::memcpy(&out.b[r * n], &in.b[keep[r] * n], n);
    }
    if (stats) stats->pairs += keep.size();
    return GP_OK;
}

}  // namespace gp

using namespace gp;
namespace gp {
int guarded_call(gp_ctx* ctx, const std::function<int()>& f);
}

extern "C" int gp_filter(gp_ctx* ctx, int order, int compress, const gp_pairs* in, gp_pairs* out, gp_stats* stats) {
    if (!ctx) return GP_ERR_ARG;
    return guarded_call(ctx, [&] {
        GP_HIP(hipSetDevice(ctx->device));
        if (!in || !out || order <= 0 || compress <= 0) throw ArgError("gp_filter: bad arguments");
        PairSet p = import_pairs(in), o;
        int rc = filter_pairs(ctx, order, compress, p, o, stats);
        if (rc != GP_OK) return rc;
        return export_pairs(o, out, false);
    });
}
