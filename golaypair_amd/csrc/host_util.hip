// Host-side helpers: alphabet, decompositions, job-split prefixes, record copies.
#include <algorithm>
#include <cstring>
#include <functional>
#include <numeric>
#include <set>

#include "gp_host.hpp"
#include "gp_internal.hpp"

namespace gp {

std::vector<int> getalphabet(int compress) {
    std::set<int> s;
    for (int i = (compress % 2 == 0) ? 0 : 1; i <= compress; i += 2) { s.insert(i); s.insert(-i); }
    return std::vector<int>(s.begin(), s.end());
}

std::vector<std::pair<int, int>> getdecomps(int k) {
    std::vector<std::pair<int, int>> out;
    for (int x = 0; x * x <= k; ++x)
        for (int y = x; x * x + y * y <= k; ++y)
            if (x * x + y * y == k) out.emplace_back(x, y);
    return out;
}

namespace {
// (sum_{i=1..len} k^gcd(i,len)) / len, saturating far above any job count.
uint64_t necklace_count_sat(int len, int k) {
    const unsigned __int128 cap = (unsigned __int128)1 << 100;
    unsigned __int128 sum = 0;
    for (int i = 1; i <= len; ++i) {
        unsigned __int128 p = 1;
        for (int j = 0, e = std::gcd(i, len); j < e && p < cap; ++j) p *= (unsigned)k;
        sum += p;
        if (sum > cap) return UINT64_MAX;
    }
    unsigned __int128 r = sum / (unsigned)len;
    return r > UINT64_MAX ? UINT64_MAX : (uint64_t)r;
}
}  // namespace

int shard_prefix_len(int n, int k, int proc_num) {
    if (proc_num <= 1) return 0;
    int i = 1;
    for (; i <= n; ++i)
        if (necklace_count_sat(i, k) >= (uint64_t)proc_num * 10000ull) break;
    return i < n ? i : 0;
}

std::vector<uint64_t> shard_prefixes(int len, const std::vector<int>& alphabet, int proc_id, int proc_num) {
    // FKM prenecklace recursion over letter indices, lexicographic order.
    const int k = (int)alphabet.size();
    std::vector<int> d(len, 0);
    std::vector<uint64_t> out;
    uint64_t count = 0;
    std::function<void(int, int)> rec = [&](int index, int period) {
        if (index == len) {
            ++count;
            if ((int64_t)(count % (uint64_t)proc_num) == proc_id) {
                uint64_t v = 0;
                for (int p = 0; p < len; ++p) v = v * (uint64_t)k + (uint64_t)d[p];
                out.push_back(v);
            }
            return;
        }
        d[index] = d[index - period];
        rec(index + 1, period);
        for (int c = d[index - period] + 1; c < k; ++c) {
            d[index] = c;
            rec(index + 1, index + 1);
        }
    };
    for (int c = 0; c < k; ++c) {
        d[0] = c;
        rec(1, 1);
    }
    return out;
}

int export_cands(const CandSet& s, gp_cands* out) {
    if (!out) return GP_ERR_ARG;
    const int64_t m = s.size();
    out->n = s.n;
    out->nkey = s.nkey;
    if (m > out->capacity || (m > 0 && (!out->words || !out->keys))) {
        out->count = m;
        return GP_ERR_CAPACITY;
    }
    if (m) {
        std::memcpy(out->words, s.words.data(), s.words.size());
        std::memcpy(out->keys, s.keys.data(), s.keys.size());
    }
    out->count = m;
    return GP_OK;
}

int export_pairs(const PairSet& s, gp_pairs* out, bool append) {
    if (!out) return GP_ERR_ARG;
    const int64_t base = append ? out->count : 0;
    const int64_t m = s.size();
    if (append && base > 0 && out->n != s.n && m > 0) return GP_ERR_ARG;
    if (m > 0) out->n = s.n;
    if (base + m > out->capacity || (m > 0 && (!out->a || !out->b))) {
        out->count = base + m;
        return GP_ERR_CAPACITY;
    }
    if (m) {
        std::memcpy(out->a + base * s.n, s.a.data(), s.a.size());
        std::memcpy(out->b + base * s.n, s.b.data(), s.b.size());
    }
    out->count = base + m;
    return GP_OK;
}

CandSet import_cands(const gp_cands* in) {
    CandSet s;
    if (!in) throw ArgError("null candidate set");
    if (in->count < 0 || (in->count > 0 && (!in->words || !in->keys))) throw ArgError("bad candidate set");
    s.n = in->n;
    s.nkey = in->nkey;
    s.words.assign(in->words, in->words + in->count * in->n);
    s.keys.assign(in->keys, in->keys + in->count * in->nkey);
    return s;
}

PairSet import_pairs(const gp_pairs* in) {
    PairSet s;
    if (!in) throw ArgError("null pair set");
    if (in->count < 0 || (in->count > 0 && (!in->a || !in->b))) throw ArgError("bad pair set");
    s.n = in->n;
    s.a.assign(in->a, in->a + in->count * in->n);
    s.b.assign(in->b, in->b + in->count * in->n);
    return s;
}

}  // namespace gp
