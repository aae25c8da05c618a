// Fixed-width sort keys of candidate lines (src/utils/io.cpp:14-36 text form
// "PSD" d_1 .. d_{n-1} " " v_0 " " .. v_{n-1} " "), ordered as the lines are
// under LC_ALL=C `sort` (src/core/sort/sort.cpp:17), without building the text:
//   bytes [0, 3 (n-1)): the key digits' decimal text, zero padded;
//   bytes [3 (n-1), 3 (n-1) + n): each letter's rank among the token strings
//   "v " in byte order;
// packed big-endian into W 64-bit words, so comparing keys word by word is
// comparing the lines bytewise.  Why: every text byte is >= 0x20, so where one
// key text is a proper prefix of the other the line with the shorter one has
// ' ' against a digit and sorts first, as the zero padding does; with equal key
// texts the lines continue with the same ' ' and then the tokens, which are
// prefix-free strings (each ends in its only space), so their concatenations
// compare as the sequences of their ranks.  Equal keys are equal lines.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <array>
#include <string>
#include <vector>

#include "textfmt.hpp"

namespace gp {

struct CandKeyer {
    int n = 0, K = 0, W = 0, kt = 0;
    std::array<uint8_t, 256> rank{};  // letter + 128 -> rank of its token string

    explicit CandKeyer(int n_) : n(n_), K(n_ / 2), kt(3 * (n_ - 1)) {
        W = (kt + n + 7) / 8;
        std::array<std::pair<std::string, int>, 256> tok;
        for (int v = -128; v < 128; ++v) {
            std::string s;
            append_int(s, v);
            s.push_back(' ');
            tok[v + 128] = {s, v + 128};
        }
        std::sort(tok.begin(), tok.end());
        for (int r = 0; r < 256; ++r) rank[tok[r].second] = (uint8_t)r;
    }

    // key words of (word w, key digits d[0..K)) into out[0..W)
    void key(const int8_t* w, const uint8_t* d, uint64_t* out) const {
        const NumText& T = num_text();
        uint8_t b[8 * 32] = {};  // W <= 32 for n <= 64
        int p = 0;
        for (int k = 1; k < n; ++k) {
            const uint8_t v = d[(k <= K ? k : n - k) - 1];
            for (int c = 0; c < T.key_len[v]; ++c) b[p++] = (uint8_t)T.key[v][c];
        }
        for (int i = 0; i < n; ++i) b[kt + i] = rank[(uint8_t)(w[i] + 128)];
        for (int x = 0; x < W; ++x) {
            uint64_t u = 0;
            for (int y = 0; y < 8; ++y) u = (u << 8) | b[8 * x + y];
            out[x] = u;
        }
    }
};

}  // namespace gp
