// Bytewise (C locale) sort of text lines held in one buffer, as `sort` does
// under LC_ALL=C (src/core/sort/sort.cpp:17): each line gets a 16-byte
// big-endian prefix key (zero padded, so a proper prefix sorts first, every
// text byte being >= 0x20), and ties on the key fall back to comparing the rest
// of the lines.  Sorted on host threads by a sample sort.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <string>
#include <string_view>
#include <thread>
#include <vector>

#include "threadpool.hpp"

namespace gp {

struct LineSet {
    const std::string* text = nullptr;
    const std::vector<size_t>* off = nullptr;  // line i = text[off[i], off[i+1])
    size_t skip = 0;                           // bytes every line shares (not keyed)
    std::string_view line(int64_t i) const {
        return std::string_view(text->data() + (*off)[i], (*off)[i + 1] - (*off)[i]);
    }
};

// Builds m lines into one buffer on host threads: append(str, i) appends line i.
template <class F>
void build_lines(int64_t m, F append, std::string& text, std::vector<size_t>& off, int threads = 16) {
    const int nt = std::max(1, std::min<int>(threads, (int)(m / 256)));
    std::vector<std::string> part(nt);
    std::vector<std::vector<size_t>> poff(nt);
    pool().parallel_for(nt, [&](size_t t) {
        const int64_t a = m * (int64_t)t / nt, b = m * (int64_t)(t + 1) / nt;
        for (int64_t i = a; i < b; ++i) {
            poff[t].push_back(part[t].size());
            append(part[t], i);
        }
    });
    size_t total = 0;
    for (auto& p : part) total += p.size();
    text.clear();
    text.reserve(total);
    off.clear();
    off.reserve(m + 1);
    for (int t = 0; t < nt; ++t) {
        const size_t base = text.size();
        for (size_t o : poff[t]) off.push_back(base + o);
        text += part[t];
    }
    off.push_back(text.size());
}

// Runs f(a, b) over [0, m) split into chunks of at least `grain` on host threads.
template <class F>
void parallel_chunks(int64_t m, F f, int64_t grain = 4096, int threads = 16) {
    const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(threads, m / grain));
    pool().parallel_for(nt, [&](size_t t) { f(m * (int64_t)t / nt, m * (int64_t)(t + 1) / nt); });
}

// Sorts v with `less` on host threads (sample sort): p - 1 splitters from a
// regular sample cut v into p buckets (counted and scattered per chunk), and
// the buckets are sorted concurrently, so no pass is serial over all of v.
// Not stable: elements that compare equal may come out in any order.
template <class T, class Less>
void parallel_sort(std::vector<T>& v, const Less& less, int threads = 16) {
    const int64_t m = (int64_t)v.size();
    const int p = (int)std::max<int64_t>(1, std::min<int64_t>(threads, m / 4096));
    if (p == 1) {
        std::sort(v.begin(), v.end(), less);
        return;
    }
    // splitters: every 32nd element of a sorted sample of 32 p evenly spaced elements
    const int per = 32;
    std::vector<T> sample;
    sample.reserve((size_t)per * p);
    for (int s = 0; s < per * p; ++s) sample.push_back(v[(size_t)(m * (int64_t)s / (per * p))]);
    std::sort(sample.begin(), sample.end(), less);
    std::vector<T> split;
    for (int b = 1; b < p; ++b) split.push_back(sample[(size_t)b * per]);
    auto bucket_of = [&](const T& x) {
        return (int)(std::upper_bound(split.begin(), split.end(), x, less) - split.begin());
    };
    // per chunk and bucket counts, then the scatter offsets
    const int nc = p;
    std::vector<int64_t> cut(nc + 1);
    for (int c = 0; c <= nc; ++c) cut[c] = m * c / nc;
    std::vector<int> bk((size_t)m);
    std::vector<int64_t> cnt((size_t)nc * p, 0);
    pool().parallel_for(nc, [&](size_t c) {
        for (int64_t i = cut[c]; i < cut[c + 1]; ++i) {
            bk[i] = bucket_of(v[i]);
            ++cnt[c * p + bk[i]];
        }
    });
    std::vector<int64_t> start((size_t)p + 1, 0), pos((size_t)nc * p);
    for (int b = 0; b < p; ++b) {
        int64_t acc = start[b];
        for (int c = 0; c < nc; ++c) {
            pos[(size_t)c * p + b] = acc;
            acc += cnt[(size_t)c * p + b];
        }
        start[b + 1] = acc;
    }
    std::vector<T> out((size_t)m);
    pool().parallel_for(nc, [&](size_t c) {
        for (int64_t i = cut[c]; i < cut[c + 1]; ++i) out[pos[c * p + bk[i]]++] = v[i];
    });
    pool().parallel_for(p, [&](size_t b) { std::sort(out.begin() + start[b], out.begin() + start[b + 1], less); });
    v.swap(out);
}

inline void line_keys(const LineSet& L, int64_t m, std::vector<uint64_t>& hi, std::vector<uint64_t>& lo) {
    hi.assign(m, 0);
    lo.assign(m, 0);
    parallel_chunks(m, [&](int64_t i0, int64_t i1) {
    for (int64_t i = i0; i < i1; ++i) {
        const std::string_view s = L.line(i);
        uint64_t h = 0, l = 0;
        for (size_t b = 0; b < 16; ++b) {
            const size_t p = L.skip + b;
            const uint64_t c = p < s.size() ? (uint8_t)s[p] : 0;
            if (b < 8) h |= c << (56 - 8 * b);
            else l |= c << (56 - 8 * (b - 8));
        }
        hi[i] = h;
        lo[i] = l;
    }
    });
}

// idx: a permutation of [0, m); sorted into line order.
inline void sort_lines(const LineSet& L, std::vector<int64_t>& idx, int threads = 16) {
    const int64_t m = (int64_t)idx.size();
    std::vector<uint64_t> hi, lo;
    line_keys(L, m, hi, lo);
    auto less = [&](int64_t x, int64_t y) {
        if (hi[x] != hi[y]) return hi[x] < hi[y];
        if (lo[x] != lo[y]) return lo[x] < lo[y];
        const std::string_view a = L.line(x), b = L.line(y);
        const size_t k = std::min(L.skip + 16, std::min(a.size(), b.size()));
        return a.substr(k) < b.substr(k);
    };
    parallel_sort(idx, less, threads);
}

}  // namespace gp
