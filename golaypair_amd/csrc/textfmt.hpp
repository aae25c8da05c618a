// Text forms of the reference's files (src/utils/io.cpp:14-36 candidate lines,
// io.cpp:75-95 pair lines).  Header-only; used by the library and the CLI.
#pragma once

#include <stdint.h>

#include <string>

namespace gp {

inline void append_int(std::string& s, int v) {
    char buf[16];
    int n = 0;
    unsigned u = v < 0 ? (unsigned)(-(long)v) : (unsigned)v;
    do { buf[n++] = (char)('0' + u % 10); u /= 10; } while (u);
    if (v < 0) s.push_back('-');
    while (n) s.push_back(buf[--n]);
}

// "PSD" + d_1 .. d_{n-1} (d_k = key[k-1] for k <= n/2, d_{n-k} = d_k) + " "
inline void append_key_text(std::string& s, const uint8_t* key, int n) {
    s += "PSD";
    const int K = n / 2;
    for (int k = 1; k < n; ++k) append_int(s, key[(k <= K ? k : n - k) - 1]);
}

// Candidate line without the trailing newline.
inline void append_cand_line(std::string& s, const int8_t* w, const uint8_t* key, int n) {
    append_key_text(s, key, n);
    s.push_back(' ');
    for (int i = 0; i < n; ++i) { append_int(s, w[i]); s.push_back(' '); }
}

// Pair line without the trailing newline: "a0 a1 .. " + " " + "b0 b1 .. "
inline void append_pair_line(std::string& s, const int8_t* a, const int8_t* b, int n) {
    for (int i = 0; i < n; ++i) { append_int(s, a[i]); s.push_back(' '); }
    s.push_back(' ');
    for (int i = 0; i < n; ++i) { append_int(s, b[i]); s.push_back(' '); }
}

}  // namespace gp
