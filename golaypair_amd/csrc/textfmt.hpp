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

// Decimal text of 0..255 (key digits) and of -128..127 followed by a space
// (sequence tokens), built once.
struct NumText {
    char key[256][3];
    uint8_t key_len[256];
    char tok[256][5];
    uint8_t tok_len[256];
    NumText() {
        for (int v = 0; v < 256; ++v) {
            std::string a;
            append_int(a, v);
            key_len[v] = (uint8_t)a.size();
            for (size_t i = 0; i < a.size(); ++i) key[v][i] = a[i];
            std::string b;
            append_int(b, v - 128);
            b.push_back(' ');
            tok_len[v] = (uint8_t)b.size();
            for (size_t i = 0; i < b.size(); ++i) tok[v][i] = b[i];
        }
    }
};
inline const NumText& num_text() {
    static const NumText t;
    return t;
}

// "PSD" + d_1 .. d_{n-1} (d_k = key[k-1] for k <= n/2, d_{n-k} = d_k) + " "
inline void append_key_text(std::string& s, const uint8_t* key, int n) {
    const NumText& T = num_text();
    const int K = n / 2;
    s += "PSD";
    for (int k = 1; k < n; ++k) {
        const uint8_t v = key[(k <= K ? k : n - k) - 1];
        s.append(T.key[v], T.key_len[v]);
    }
}

// Candidate line without the trailing newline, written with one resize.
inline void append_cand_line(std::string& s, const int8_t* w, const uint8_t* key, int n) {
    const NumText& T = num_text();
    const int K = n / 2;
    size_t need = 4;
    for (int k = 1; k < n; ++k) need += T.key_len[key[(k <= K ? k : n - k) - 1]];
    for (int i = 0; i < n; ++i) need += T.tok_len[(uint8_t)(w[i] + 128)];
    const size_t old = s.size();
    s.resize(old + need);
    char* p = &s[old];
    *p++ = 'P';
    *p++ = 'S';
    *p++ = 'D';
    for (int k = 1; k < n; ++k) {
        const uint8_t v = key[(k <= K ? k : n - k) - 1];
        for (int c = 0; c < T.key_len[v]; ++c) *p++ = T.key[v][c];
    }
    *p++ = ' ';
    for (int i = 0; i < n; ++i) {
        const uint8_t v = (uint8_t)(w[i] + 128);
        for (int c = 0; c < T.tok_len[v]; ++c) *p++ = T.tok[v][c];
    }
}

// Pair line without the trailing newline: "a0 a1 .. " + " " + "b0 b1 .. "
inline void append_pair_line(std::string& s, const int8_t* a, const int8_t* b, int n) {
    for (int i = 0; i < n; ++i) { append_int(s, a[i]); s.push_back(' '); }
    s.push_back(' ');
    for (int i = 0; i < n; ++i) { append_int(s, b[i]); s.push_back(' '); }
}

}  // namespace gp
