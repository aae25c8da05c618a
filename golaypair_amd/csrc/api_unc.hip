// C ABI: uncompression entry points (uncompression.cpp:32-210).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <future>
#include <numeric>
#include <string>
#include <thread>

#include "gp_host.hpp"
#include "gp_internal.hpp"
#include "gp_unc.hpp"
#include "candkey.hpp"
#include "linesort.hpp"
#include "textfmt.hpp"
#include "threadpool.hpp"

using namespace gp;

namespace gp {
int guarded_call(gp_ctx* ctx, const std::function<int()>& f);
}

namespace {

constexpr uint64_t kLeafBudget = 1ull << 30;  // leaves per device batch

}  // namespace

extern "C" {

int gp_uncompress_side(gp_ctx* ctx, int8_t* orig, int len, int compress, int new_compress, int paf, int proc_id,
                       int proc_num, int seqflag, gp_cands* out, gp_stats* stats) {
    if (!ctx) return GP_ERR_ARG;
    return guarded_call(ctx, [&] {
        GP_HIP(hipSetDevice(ctx->device));
        if (!orig || len <= 0 || compress <= 1 || new_compress <= 0 || compress % new_compress || !out)
            throw ArgError("gp_uncompress_side: bad arguments");
        if (proc_num < 1) throw ArgError("gp_uncompress_side: proc_num must be at least 1");
        const auto tab = make_splits(compress, new_compress);
        const SideHost S = build_side(*tab, orig, len, seqflag);
        for (int i = 0; i < len; ++i) orig[i] = (int8_t)S.orig[i];  // the reference rotates in place
        // uncompression.cpp:167-191: the job split depth is the first number of
        // leading positions whose partial words reach 1000 * proc_num (or all of
        // them); job proc_id takes the partials of index proc_id mod proc_num.
        // The device filters the whole side; the job's leaves are selected from
        // its survivors by that index.
        uint64_t job_suffix = 1;
        if (proc_num > 1) {
            uint64_t partials = 1;
            int depth = 0;
            while (depth < len && partials < 1000ull * (uint64_t)proc_num) partials *= (uint64_t)S.R[depth++];
            for (int j = depth; j < len; ++j) job_suffix *= (uint64_t)S.R[j];
        }
        const int order = len * compress;
        const int g0 = 0;
        UncBatch B = unc_filter(ctx, std::span<const SideHost>(&S, 1), std::span<const int>(&g0, 1), order, paf, stats);
        CandSet c = unc_side_cands(ctx, B, 0, job_suffix, proc_id, proc_num);
        c.n = S.N;
        c.nkey = S.N / 2;
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
        Phase ph_all("uncompress: total");
        PairSet inp = import_pairs(in);
        const int64_t P = inp.size();
        const int N = order / new_compress;
        // a call repeated after GP_ERR_CAPACITY returns the stored result
        uint64_t key = 1469598103934665603ull;
        for (int v : {order, compress, new_compress, paf, ctx->psd_mode, (int)P, tmp_a ? 1 : 0, tmp_b ? 1 : 0})
            key = (key ^ (uint64_t)(uint32_t)v) * 1099511628211ull;
        for (int8_t v : inp.a) key = (key ^ (uint8_t)v) * 1099511628211ull;
        for (int8_t v : inp.b) key = (key ^ (uint8_t)v) * 1099511628211ull;
        if (ctx->unc_cache_valid && ctx->unc_cache_key == key) {
            ctx->unc_cache_valid = false;
            int rc = GP_OK;
            if (P > 0 && tmp_a) rc = export_cands(ctx->unc_cache_ta, tmp_a);
            if (rc == GP_OK && P > 0 && tmp_b) rc = export_cands(ctx->unc_cache_tb, tmp_b);
            if (rc == GP_OK) rc = export_pairs(ctx->unc_cache_out, out, true);
            if (rc == GP_ERR_CAPACITY) ctx->unc_cache_valid = true;
            else if (rc == GP_OK) add_stats(stats, ctx->unc_cache_stats);
            return rc;
        }
        gp_stats local{};  // reported by the call that returns the pairs
        PairSet result;
        result.n = N;
        CandSet last_a, last_b;
        last_a.n = last_b.n = N;
        last_a.nkey = last_b.nkey = N / 2;
        // Sides are built on host threads, the first batch's first, so that its
        // tables are prepared while the other sides are built.  Batches by leaf
        // budget; the host tables of batch i + 2 are prepared on another thread
        // while the device runs batch i, one batch at a time (each preparation
        // has every host thread); the first three batches are an eighth, a
        // quarter and half of the budget, so that the device starts early and
        // the host keeps ahead of it, and the last is at most a quarter.
        static_assert(std::is_trivially_copyable_v<SideHost> && std::is_trivially_destructible_v<SideHost>);
        if (ctx->unc_side_buf.size() < (size_t)(2 * P) * sizeof(SideHost))
            ctx->unc_side_buf.resize((size_t)(2 * P) * sizeof(SideHost));
        SideHost* const all = reinterpret_cast<SideHost*>(ctx->unc_side_buf.data());
        std::vector<int> group((size_t)(2 * P));
        for (size_t i = 0; i < group.size(); ++i) group[i] = (int)(i / 2);
        uint64_t leaf_budget = (uint64_t)std::clamp<int64_t>(
            tuned(ctx, "unc_leaf_budget/" + std::to_string(N), (int64_t)kLeafBudget), 1 << 16, (int64_t)1 << 34);
        if (const char* e = std::getenv("GOLAY_UNC_BATCH_LEAVES"))  // smaller batches (tests)
            leaf_budget = std::max<uint64_t>(4, std::strtoull(e, nullptr, 10));
        UncTrig trig;
        if (P > 0) trig = unc_trig(ctx, N);
        const bool fp32 = ctx->psd_mode == GP_PSD_FP32;
        auto prepare = [&, L](int64_t q0, int64_t q1, const std::function<void()>& before_fill) {
            const std::span<const SideHost> sides(all + 2 * q0, (size_t)(2 * (q1 - q0)));
            const std::span<const int> grp(group.data(), sides.size());
            GP_HIP(hipSetDevice(ctx->device));  // the pinned staging buffer
            return unc_prepare(sides, grp, order, paf, fp32, trig, &ctx->pinned, before_fill);
        };
        const auto tab = make_splits(compress, new_compress);
        auto build_range = [&](int64_t q0, int64_t q1) {
            pool().parallel_for((size_t)(q1 - q0), [&](size_t i) {
                const int64_t q = q0 + (int64_t)i;
                all[2 * q] = build_side(*tab, &inp.a[q * L], L, 0);
                all[2 * q + 1] = build_side(*tab, &inp.b[q * L], L, 1);
            });
        };
        auto pair_leaves = [&](int64_t q) { return all[2 * q].leaves + all[2 * q + 1].leaves; };
        std::vector<std::pair<int64_t, int64_t>> batches;
        using PrepFuture = std::shared_future<std::shared_ptr<UncPrep>>;
        PrepFuture next, after;
        // a batch's splits and layout (mostly one thread) overlap the previous
        // batch's table fill; its own fill (all host threads) waits for that one
        auto prepare_after = [&](PrepFuture prev, int64_t q0, int64_t q1) {
            return std::async(std::launch::async, [&prepare, prev, q0, q1]() mutable {
                       auto P = prepare(q0, q1, [&prev] {
                           if (prev.valid()) prev.wait();
                           prev = PrepFuture();  // the previous batch's tables are released with its own future
                       });
                       prev = PrepFuture();
                       return P;
                   }).share();
        };
        {
            Phase ph("uncompress: build sides");
            int64_t built = 0, q1 = 0;
            uint64_t leaves = 0;
            bool closed = P == 0;
            while (!closed && built < P) {  // the first batch, in chunks of pairs
                const int64_t c1 = std::min<int64_t>(P, built + 512);
                build_range(built, c1);
                built = c1;
                for (; q1 < built; ++q1) {
                    const uint64_t add = pair_leaves(q1);
                    if (q1 > 0 && leaves + add > leaf_budget / 8) {
                        closed = true;
                        break;
                    }
                    leaves += add;
                }
                closed = closed || q1 == P;
            }
            if (P > 0) {
                batches.emplace_back(0, q1);
                next = prepare_after(PrepFuture(), 0, q1);
            }
            build_range(built, P);
            for (int64_t q0 = q1; q0 < P;) {
                uint64_t lv = 0;
                int64_t e = q0;
                while (e < P) {
                    const uint64_t add = pair_leaves(e);
                    // the second and third batches are a quarter and half the budget:
                    // their tables are ready sooner
                    const uint64_t cap = batches.size() == 1   ? leaf_budget / 4
                                         : batches.size() == 2 ? leaf_budget / 2
                                                               : leaf_budget;
                    if (e > q0 && lv + add > cap) break;
                    lv += add;
                    ++e;
                }
                batches.emplace_back(q0, e);
                q0 = e;
            }
            // a last batch of at most a quarter of the budget: its pairs are
            // ordered after the device is done, the others' while it works
            uint64_t last_leaves = 0;
            if (!batches.empty())
                for (int64_t q = batches.back().first; q < batches.back().second; ++q) last_leaves += pair_leaves(q);
            if (batches.size() > 1 && last_leaves > leaf_budget / 4) {
                const auto [l0, l1] = batches.back();
                uint64_t tail = 0;
                int64_t cut = l1;
                while (cut > l0 + 1 && tail + pair_leaves(cut - 1) <= leaf_budget / 4) tail += pair_leaves(--cut);
                if (cut > l0 && cut < l1) {
                    batches.back().second = cut;
                    batches.emplace_back(cut, l1);
                }
            }
            if (batches.size() > 1) after = prepare_after(next, batches[1].first, batches[1].second);
        }
        // Emitted pairs of one batch in the reference's order (input pair, then A
        // line text, then B line text), appended to `result`.  Runs on a host
        // thread while the device works on the next batch; tasks are chained so
        // batches append in order.
        uint64_t ord_ties = 0, ord_pairs = 0;
        auto order_batch = [&, N](UncEmitted& E) {
            const int K = N / 2;
            Phase ph("order: keys");
            // fixed-width keys in line order (candkey.hpp) of the A and B rows
            const CandKeyer ck(N);
            const int W = ck.W;
            std::vector<uint64_t> ka((size_t)E.count * W), kb((size_t)E.count * W);
            parallel_chunks(E.count, [&](int64_t i0, int64_t i1) {
                for (int64_t i = i0; i < i1; ++i) {
                    ck.key(&E.words[i * N], &E.keys[i * K], &ka[(size_t)i * W]);
                    ck.key(&E.words[(E.count + i) * N], &E.keys[(E.count + i) * K], &kb[(size_t)i * W]);
                }
            }, 256);
            ph.next("order: sort");
            std::vector<int64_t> ord;
            for (int64_t i = 0; i < E.count; ++i) {
                // key texts must agree (they do unless an FP64 PSD sits on a rounding tie)
                if (std::memcmp(&E.keys[i * K], &E.keys[(E.count + i) * K], K) != 0) {
                    ++ord_ties;
                    continue;
                }
                ord.push_back(i);
            }
            auto kless = [W](const uint64_t* x, const uint64_t* y) {
                for (int q = 0; q < W; ++q)
                    if (x[q] != y[q]) return x[q] < y[q];
                return false;
            };
            std::sort(ord.begin(), ord.end(), [&](int64_t x, int64_t y) {
                if (E.group[x] != E.group[y]) return E.group[x] < E.group[y];
                const uint64_t* ax = &ka[(size_t)x * W];
                const uint64_t* ay = &ka[(size_t)y * W];
                if (kless(ax, ay)) return true;
                if (kless(ay, ax)) return false;
                return kless(&kb[(size_t)x * W], &kb[(size_t)y * W]);
            });
            ph.next("order: append");
            for (int64_t i : ord) {
                result.a.insert(result.a.end(), &E.words[i * N], &E.words[i * N] + N);
                result.b.insert(result.b.end(), &E.words[(E.count + i) * N], &E.words[(E.count + i) * N] + N);
            }
            ord_pairs += ord.size();
        };
        std::future<void> ordered;
        // batch i + 1 is queued on the device (in the other buffer slot) before
        // batch i is collected, so the device never waits for the host between
        // batches
        auto collect = [&](UncInflight& F, size_t bi) {
            const auto [q0, q1] = batches[bi];
            auto E = std::make_shared<UncEmitted>();
            UncBatch B;
            {
                Phase ph("uncompress: collect batch");
                B = unc_collect(ctx, F, paf * new_compress, *E, &local);
            }
            if (q1 == P && (tmp_a || tmp_b)) {  // the last pair's candidate files
                last_a = unc_side_cands(ctx, B, (int)(2 * (P - 1 - q0)));
                last_b = unc_side_cands(ctx, B, (int)(2 * (P - 1 - q0) + 1));
            }
            ordered = std::async(std::launch::async, [&order_batch, E, prev = std::move(ordered)]() mutable {
                if (prev.valid()) prev.get();
                Phase ph("uncompress: order batch");
                order_batch(*E);
            });
        };
        std::shared_ptr<UncInflight> inflight;
        for (size_t bi = 0; bi < batches.size(); ++bi) {
            std::shared_ptr<UncPrep> prep = next.get();
            next = std::move(after);
            after = PrepFuture();
            if (bi + 2 < batches.size()) after = prepare_after(next, batches[bi + 2].first, batches[bi + 2].second);
            std::shared_ptr<UncInflight> F;
            {
                Phase ph("uncompress: enqueue batch");
                F = unc_enqueue(ctx, std::move(prep), (int)(bi & 1), paf * new_compress);
            }
            if (inflight) collect(*inflight, bi - 1);
            inflight = std::move(F);
        }
        if (inflight) collect(*inflight, batches.size() - 1);
        Phase ph_end("uncompress: last order");
        if (ordered.valid()) ordered.get();
        ph_end.next("uncompress: release");
        inflight.reset();
        ph_end.next("uncompress: export");
        local.tie_band += ord_ties;
        local.pairs += ord_pairs;
        int rc = GP_OK;
        if (P > 0) {
            if (tmp_a) rc = export_cands(last_a, tmp_a);
            if (rc == GP_OK && tmp_b) rc = export_cands(last_b, tmp_b);
        } else {
            if (tmp_a) tmp_a->count = 0;
            if (tmp_b) tmp_b->count = 0;
        }
        const int64_t base = out->count;
        if (rc == GP_OK) rc = export_pairs(result, out, true);
        if (rc == GP_ERR_CAPACITY) {
            if (out->count < base + result.size()) out->count = base + result.size();
            // report every size the caller must grow, then keep the result for the retry
            if (tmp_a && P > 0) tmp_a->count = last_a.size();
            if (tmp_b && P > 0) tmp_b->count = last_b.size();
            ctx->unc_cache_key = key;
            ctx->unc_cache_out = std::move(result);
            ctx->unc_cache_ta = std::move(last_a);
            ctx->unc_cache_tb = std::move(last_b);
            ctx->unc_cache_valid = true;
            ctx->unc_cache_stats = local;
        } else if (rc == GP_OK) {
            add_stats(stats, local);
        }
        return rc;
    });
}

}  // extern "C"
