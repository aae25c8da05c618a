"""Fixed-width candidate sort keys (golaypair_amd/csrc/candkey.hpp) order
candidates exactly as the C-locale byte order of their text lines
(src/utils/io.cpp:14-36, src/core/sort/sort.cpp:17), and equal keys are equal
lines.  Random words over alphabets with negative and two-digit letters and
random key values of one to three digits, with duplicates; checked against a
plain std::string sort of the lines built by textfmt.hpp."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROG = r'''
#include <algorithm>
#include <cstdio>
#include <numeric>
#include <random>
#include <string>
#include <vector>
#include "candkey.hpp"
int main() {
    std::mt19937_64 rng(12345);
    for (int n : {5, 8, 9, 15, 18, 45, 64}) {
        const int K = n / 2;
        gp::CandKeyer ck(n);
        for (int alpha : {1, 3, 10, 127}) {
            const int m = 3000;
            std::vector<int8_t> w((size_t)m * n);
            std::vector<uint8_t> d((size_t)m * K);
            for (int i = 0; i < m; ++i) {
                const int src = (i % 5 == 4) ? (int)(rng() % i) : -1;  // some duplicates
                for (int j = 0; j < n; ++j)
                    w[i * n + j] = src >= 0 ? w[src * n + j] : (int8_t)((int)(rng() % (2 * alpha + 1)) - alpha);
                for (int k = 0; k < K; ++k) {
                    const int r = (int)(rng() % 3);  // one, two or three digits
                    d[i * K + k] = src >= 0 ? d[src * K + k] : (uint8_t)(r == 0 ? rng() % 10 : r == 1 ? 10 + rng() % 90 : 100 + rng() % 156);
                }
            }
            std::vector<std::string> line(m);
            for (int i = 0; i < m; ++i) gp::append_cand_line(line[i], &w[i * n], &d[i * K], n);
            std::vector<uint64_t> key((size_t)m * ck.W);
            for (int i = 0; i < m; ++i) ck.key(&w[i * n], &d[i * K], &key[(size_t)i * ck.W]);
            auto kless = [&](int x, int y) {
                for (int q = 0; q < ck.W; ++q)
                    if (key[(size_t)x * ck.W + q] != key[(size_t)y * ck.W + q]) return key[(size_t)x * ck.W + q] < key[(size_t)y * ck.W + q];
                return false;
            };
            for (int t = 0; t < 20000; ++t) {
                const int x = (int)(rng() % m), y = (int)(rng() % m);
                const bool tl = line[x] < line[y], teq = line[x] == line[y];
                const bool kl = kless(x, y), keq = !kless(x, y) && !kless(y, x);
                if (tl != kl || teq != keq) { std::printf("n=%d alpha=%d: %s | %s\n", n, alpha, line[x].c_str(), line[y].c_str()); return 1; }
            }
        }
    }
    std::puts("ok");
    return 0;
}
'''


def test_candkey_order_equals_text_order(tmp_path):
    src = tmp_path / "ck.cpp"
    exe = tmp_path / "ck"
    src.write_text(PROG)
    subprocess.run(["g++", "-O2", "-std=c++20", "-I", os.path.join(ROOT, "golaypair_amd", "csrc"), str(src), "-o",
                    str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
