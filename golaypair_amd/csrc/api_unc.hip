// C ABI: uncompression entry points (uncompression.cpp:32-210).
#include <algorithm>
#include <cstring>
#include <functional>
#include <numeric>
#include <string>

#include "gp_host.hpp"
#include "gp_internal.hpp"
#include "gp_sort.hpp"
#include "gp_unc.hpp"
#include "textfmt.hpp"

using namespace gp;

namespace gp {
int guarded_call(gp_ctx* ctx, const std::function<int()>& f);
}

namespace {

constexpr uint64_t kLeafBudget = 1ull << 30;  // leaves per device batch

// Rows [begin, end) of the batch's device survivors, copied to the host.
void fetch_rows(gp_ctx* ctx, const UncResult& u, int64_t begin, int64_t end, CandSet& out) {
    const int64_t m = end - begin;
    out.n = u.N;
    out.nkey = u.K;
    out.words.resize((size_t)m * u.N);
    out.keys.resize((size_t)m * u.K);
    if (m > 0) {
        GP_HIP(hipMemcpyAsync(out.words.data(), u.d_words + begin * u.N, (size_t)m * u.N, hipMemcpyDeviceToHost,
                              ctx->stream));
        GP_HIP(hipMemcpyAsync(out.keys.data(), u.d_keys + begin * u.K, (size_t)m * u.K, hipMemcpyDeviceToHost,
                              ctx->stream));
    }
    GP_HIP(hipStreamSynchronize(ctx->stream));
}

// Text-order permutation of a candidate set (C-locale line order).
std::vector<int64_t> text_order(const CandSet& c) {
    const int64_t m = c.size();
    std::string text;
    std::vector<size_t> off(m + 1);
    for (int64_t i = 0; i < m; ++i) {
        off[i] = text.size();
        append_cand_line(text, &c.words[i * c.n], &c.keys[i * c.nkey], c.n);
    }
    off[m] = text.size();
    std::vector<int64_t> idx(m);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int64_t x, int64_t y) {
        return std::string_view(text.data() + off[x], off[x + 1] - off[x]) <
               std::string_view(text.data() + off[y], off[y + 1] - off[y]);
    });
    return idx;
}

CandSet permuted(const CandSet& c, const std::vector<int64_t>& idx) {
    CandSet o;
    o.n = c.n;
    o.nkey = c.nkey;
    o.words.resize(idx.size() * c.n);
    o.keys.resize(idx.size() * c.nkey);
    for (size_t r = 0; r < idx.size(); ++r) {
        std::memcpy(&o.words[r * c.n], &c.words[idx[r] * c.n], c.n);
        std::memcpy(&o.keys[r * c.nkey], &c.keys[idx[r] * c.nkey], c.nkey);
    }
    return o;
}

}  // namespace

extern "C" {

int gp_uncompress_side(gp_ctx* ctx, int8_t* orig, int len, int compress, int new_compress, int paf, int seqflag,
                       gp_cands* out, gp_stats* stats) {
    if (!ctx) return GP_ERR_ARG;
    return guarded_call(ctx, [&] {
        GP_HIP(hipSetDevice(ctx->device));
        if (!orig || len <= 0 || compress <= 1 || new_compress <= 0 || compress % new_compress || !out)
            throw ArgError("gp_uncompress_side: bad arguments");
        SideHost S = build_side(orig, len, compress, new_compress, seqflag);
        for (int i = 0; i < len; ++i) orig[i] = (int8_t)S.orig[i];  // the reference rotates in place
        const int order = len * compress;
        std::vector<SideHost> v{S};
        UncResult u = expand_sides(ctx, v, order, paf, stats);
        CandSet c;
        c.n = S.N;
        c.nkey = S.N / 2;
        if (!u.side_begin.empty() && u.side_begin.back() > 0) fetch_rows(ctx, u, 0, u.side_begin.back(), c);
        return export_cands(c, out);
    });
}

int gp_uncompress(gp_ctx* ctx, int order, int compress, int new_compress, int paf, const gp_pairs* in,
                  gp_pairs* out, gp_cands* tmp_a, gp_cands* tmp_b, gp_stats* stats) {
    if (!ctx) return GP_ERR_ARG;
    return guarded_call(ctx, [&] {
        GP_HIP(hipSetDevice(ctx->device));
        if (compress <= 1) return GP_OK;  // uncompression.cpp:43
        if (!in || !out) throw ArgError("gp_uncompress: null argument");
        if (new_compress <= 0 || compress % new_compress) throw ArgError("gp_uncompress: bad compression pair");
        const int L = order / compress;
        if (in->count > 0 && in->n != L) throw ArgError("gp_uncompress: pair length != order/compress");
        PairSet inp = import_pairs(in);
        const int64_t P = inp.size();
        const int N = order / new_compress;
        PairSet result;
        result.n = N;
        CandSet last_a, last_b;
        last_a.n = last_b.n = N;
        last_a.nkey = last_b.nkey = N / 2;
        int64_t q0 = 0;
        while (q0 < P) {
            // batch of pairs by leaf budget
            std::vector<SideHost> sides;
            uint64_t leaves = 0;
            int64_t q1 = q0;
            while (q1 < P) {
                SideHost a = build_side(&inp.a[q1 * L], L, compress, new_compress, 0);
                SideHost b = build_side(&inp.b[q1 * L], L, compress, new_compress, 1);
                const uint64_t add = a.leaves + b.leaves;
                if (q1 > q0 && leaves + add > kLeafBudget) break;
                leaves += add;
                sides.push_back(std::move(a));
                sides.push_back(std::move(b));
                ++q1;
            }
            UncResult u = expand_sides(ctx, sides, order, paf, stats);
            for (int64_t q = q0; q < q1; ++q) {
                const int64_t sa = 2 * (q - q0), sb = sa + 1;
                const int64_t a0 = u.side_begin[sa], a1 = u.side_begin[sa + 1];
                const int64_t b0 = u.side_begin[sb], b1 = u.side_begin[sb + 1];
                if (q == P - 1 && (tmp_a || tmp_b)) {
                    fetch_rows(ctx, u, a0, a1, last_a);
                    fetch_rows(ctx, u, b0, b1, last_b);
                }
                if (a1 == a0 || b1 == b0) continue;  // match_pairs.cpp:32-38
                JoinResult jr = join_pairs(ctx, u.d_words + a0 * N, u.d_keys + a0 * (N / 2), a1 - a0,
                                           u.d_words + b0 * N, u.d_keys + b0 * (N / 2), b1 - b0, N, N / 2, N, N / 2,
                                           paf * new_compress, stats);
                if (jr.ia.empty()) continue;
                // emission order: (A line, B line) in text order
                std::vector<CandSet> rows(2);
                for (int s = 0; s < 2; ++s) {
                    rows[s].n = N;
                    rows[s].nkey = N / 2;
                    rows[s].words.resize(jr.ia.size() * N);
                    rows[s].keys.resize(jr.ia.size() * (N / 2));
                }
                for (size_t i = 0; i < jr.ia.size(); ++i) {
                    const int64_t ra = a0 + jr.ia[i], rb = b0 + jr.ib[i];
                    GP_HIP(hipMemcpyAsync(&rows[0].words[i * N], u.d_words + ra * N, N, hipMemcpyDeviceToHost, ctx->stream));
                    GP_HIP(hipMemcpyAsync(&rows[0].keys[i * (N / 2)], u.d_keys + ra * (N / 2), N / 2, hipMemcpyDeviceToHost, ctx->stream));
                    GP_HIP(hipMemcpyAsync(&rows[1].words[i * N], u.d_words + rb * N, N, hipMemcpyDeviceToHost, ctx->stream));
                    GP_HIP(hipMemcpyAsync(&rows[1].keys[i * (N / 2)], u.d_keys + rb * (N / 2), N / 2, hipMemcpyDeviceToHost, ctx->stream));
                }
                GP_HIP(hipStreamSynchronize(ctx->stream));
                std::vector<std::string> la(jr.ia.size()), lb(jr.ia.size());
                for (size_t i = 0; i < jr.ia.size(); ++i) {
                    append_cand_line(la[i], &rows[0].words[i * N], &rows[0].keys[i * (N / 2)], N);
                    append_cand_line(lb[i], &rows[1].words[i * N], &rows[1].keys[i * (N / 2)], N);
                }
                std::vector<size_t> ord(jr.ia.size());
                std::iota(ord.begin(), ord.end(), 0);
                std::sort(ord.begin(), ord.end(), [&](size_t x, size_t y) {
                    return la[x] != la[y] ? la[x] < la[y] : lb[x] < lb[y];
                });
                for (size_t i : ord) {
                    result.a.insert(result.a.end(), &rows[0].words[i * N], &rows[0].words[i * N] + N);
                    result.b.insert(result.b.end(), &rows[1].words[i * N], &rows[1].words[i * N] + N);
                }
                if (stats) stats->pairs += jr.ia.size();
            }
            q0 = q1;
        }
        if (P > 0) {
            if (tmp_a) { int rc = export_cands(last_a, tmp_a); if (rc) return rc; }
            if (tmp_b) { int rc = export_cands(last_b, tmp_b); if (rc) return rc; }
        } else {
            if (tmp_a) tmp_a->count = 0;
            if (tmp_b) tmp_b->count = 0;
        }
        return export_pairs(result, out, true);
    });
}

}  // extern "C"
