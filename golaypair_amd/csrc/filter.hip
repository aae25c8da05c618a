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
// (one pair per lane); pairs are then grouped by orbit minimum on the host.
#include <algorithm>
#include <cstring>
#include <functional>
#include <numeric>

#include "gp_host.hpp"
#include "gp_internal.hpp"

namespace gp {

struct FilterParams {
    int32_t n;
    int32_t alt;
    int32_t nunits;
    int32_t units[64];
};

// lexicographic compare of transform image against best[]: -1 less, 0 equal, 1 greater
__device__ __forceinline__ int cmp_img(const int8_t* X, int n, int u, int t, int kappa, int sg, const int8_t* best) {
    int idx = t;
    for (int i = 0; i < n; ++i) {
        int v = X[idx] * sg;
        if (kappa && (i & 1)) v = -v;
        if (v != best[i]) return v < best[i] ? -1 : 1;
        idx += u;
        if (idx >= n) idx -= n;
    }
    return 0;
}

__device__ __forceinline__ void write_img(const int8_t* X, int n, int u, int t, int kappa, int sg, int8_t* out) {
    int idx = t;
    for (int i = 0; i < n; ++i) {
        int v = X[idx] * sg;
        if (kappa && (i & 1)) v = -v;
        out[i] = (int8_t)v;
        idx += u;
        if (idx >= n) idx -= n;
    }
}

__global__ void k_orbit_min(FilterParams P, const int8_t* __restrict__ pa, const int8_t* __restrict__ pb, int64_t m,
                            int8_t* __restrict__ canon) {
    const int n = P.n;
    int8_t besta[kMaxN], bestb[kMaxN];
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < m; q += (int64_t)gridDim.x * blockDim.x) {
        const int8_t* a = pa + q * n;
        const int8_t* b = pb + q * n;
        const int nk = P.alt ? 2 : 1, ns = P.alt ? 2 : 1;
        // phase 1: minimum a-part
        bool have = false;
        for (int sw = 0; sw < 2; ++sw) {
            const int8_t* X = sw ? b : a;
            for (int ui = 0; ui < P.nunits; ++ui)
                for (int kappa = 0; kappa < nk; ++kappa)
                    for (int si = 0; si < ns; ++si)
                        for (int t = 0; t < n; ++t) {
                            const int sg = si ? -1 : 1;
                            if (!have || cmp_img(X, n, P.units[ui], t, kappa, sg, besta) < 0) {
                                write_img(X, n, P.units[ui], t, kappa, sg, besta);
                                have = true;
                            }
                        }
        }
        // phase 2: minimum b-part over the transforms reaching the minimum a-part
        bool haveb = false;
        for (int sw = 0; sw < 2; ++sw) {
            const int8_t* X = sw ? b : a;
            const int8_t* Y = sw ? a : b;
            for (int ui = 0; ui < P.nunits; ++ui)
                for (int kappa = 0; kappa < nk; ++kappa)
                    for (int si = 0; si < ns; ++si)
                        for (int t = 0; t < n; ++t) {
                            const int u = P.units[ui];
                            if (cmp_img(X, n, u, t, kappa, si ? -1 : 1, besta) != 0) continue;
                            for (int e = 0; e < 2; ++e) {
                                const int ub = e ? n - u : u;
                                for (int sb = 0; sb < ns; ++sb)
                                    for (int tb = 0; tb < n; ++tb) {
                                        const int sgb = sb ? -1 : 1;
                                        if (!haveb || cmp_img(Y, n, ub, tb, kappa, sgb, bestb) < 0) {
                                            write_img(Y, n, ub, tb, kappa, sgb, bestb);
                                            haveb = true;
                                        }
                                    }
                            }
                        }
        }
        int8_t* o = canon + q * 2 * n;
        for (int i = 0; i < n; ++i) { o[i] = besta[i]; o[n + i] = bestb[i]; }
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
    FilterParams P{};
    P.n = n;
    P.alt = !(compress > 1 || n % 2 == 1);  // filter.cpp:83-87
    for (int u = 1; u < n || (n == 1 && u == 1); ++u)
        if (std::gcd(u, n) == 1) P.units[P.nunits++] = u;
    if (n == 1) { P.nunits = 1; P.units[0] = 0; }
    hipStream_t st = ctx->stream;
    int8_t* da = ctx->dev<int8_t>("f_a", (size_t)m * n);
    int8_t* db = ctx->dev<int8_t>("f_b", (size_t)m * n);
    int8_t* dc = ctx->dev<int8_t>("f_c", (size_t)m * 2 * n);
    GP_HIP(hipMemcpyAsync(da, in.a.data(), (size_t)m * n, hipMemcpyHostToDevice, st));
    GP_HIP(hipMemcpyAsync(db, in.b.data(), (size_t)m * n, hipMemcpyHostToDevice, st));
    GP_HIP(hipEventRecord(ctx->ev0, st));
    hipLaunchKernelGGL(k_orbit_min, dim3((unsigned)std::min<int64_t>((m + 63) / 64, ctx->num_cus * 8)), dim3(64), 0,
                       st, P, da, db, m, dc);
    GP_HIP(hipGetLastError());
    GP_HIP(hipEventRecord(ctx->ev1, st));
    std::vector<int8_t> canon((size_t)m * 2 * n);
    GP_HIP(hipMemcpyAsync(canon.data(), dc, canon.size(), hipMemcpyDeviceToHost, st));
    GP_HIP(hipStreamSynchronize(st));
    float ms = 0;
    GP_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (stats) stats->kernel_ms += ms;
    // lexicographic order on (a, b) as signed ints (golay.h:25-31)
    auto pair_less = [&](int64_t x, int64_t y) {
        for (int i = 0; i < n; ++i)
            if (in.a[x * n + i] != in.a[y * n + i]) return in.a[x * n + i] < in.a[y * n + i];
        for (int i = 0; i < n; ++i)
            if (in.b[x * n + i] != in.b[y * n + i]) return in.b[x * n + i] < in.b[y * n + i];
        return false;
    };
    auto canon_cmp = [&](int64_t x, int64_t y) {
        for (int i = 0; i < 2 * n; ++i)
            if (canon[x * 2 * n + i] != canon[y * 2 * n + i]) return canon[x * 2 * n + i] < canon[y * 2 * n + i] ? -1 : 1;
        return 0;
    };
    std::vector<int64_t> idx(m);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int64_t x, int64_t y) {
        int c = canon_cmp(x, y);
        return c != 0 ? c < 0 : pair_less(x, y);
    });
    std::vector<int64_t> keep;
    for (int64_t r = 0; r < m; ++r)
        if (r == 0 || canon_cmp(idx[r], idx[r - 1]) != 0) keep.push_back(idx[r]);
    std::sort(keep.begin(), keep.end(), pair_less);
    for (int64_t i : keep) {
        out.a.insert(out.a.end(), &in.a[i * n], &in.a[i * n] + n);
        out.b.insert(out.b.end(), &in.b[i * n], &in.b[i * n] + n);
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
