// Bytewise (C locale) sort of text lines held in one buffer, as `sort` does
// under LC_ALL=C (src/core/sort/sort.cpp:17): each line gets a 16-byte
// big-endian prefix key (zero padded, so a proper prefix sorts first, every
// text byte being >= 0x20), and ties on the key fall back to comparing the rest
// of the lines.  Chunks are sorted on host threads and merged pairwise.
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

// Sorts idx with `less` on host threads: chunks sorted concurrently, then merged pairwise.
template <class T, class Less>
void parallel_sort(std::vector<T>& idx, const Less& less, int threads = 16) {
    const int64_t m = (int64_t)idx.size();
    const int nt = std::max(1, std::min<int>(threads, (int)(m / 4096)));
    std::vector<int64_t> cut(nt + 1);
    for (int t = 0; t <= nt; ++t) cut[t] = m * t / nt;
    pool().parallel_for(nt, [&](size_t t) { std::sort(idx.begin() + cut[t], idx.begin() + cut[t + 1], less); });
    for (int w = 1; w < nt; w *= 2) {
        const int pairs = (nt - w + 2 * w - 1) / (2 * w);
        pool().parallel_for(pairs, [&](size_t q) {
            const int t = (int)q * 2 * w;
            const int64_t a = cut[t], b = cut[t + w], e = cut[std::min(t + 2 * w, nt)];
            std::inplace_merge(idx.begin() + a, idx.begin() + b, idx.begin() + e, less);
        });
    }
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
