"""Host sorting helpers (golaypair_amd/csrc/linesort.hpp): parallel_sort (a
sample sort on the host thread pool) gives std::sort's result on random inputs
with many duplicates and skewed values, at sizes below and above its serial
cutoff; sort_lines orders text lines bytewise, as LC_ALL=C sort does."""
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
#include "linesort.hpp"
int main() {
    std::mt19937_64 rng(7);
    for (int64_t m : {0, 1, 4095, 8192, 100000, 1000000}) {
        for (int mode = 0; mode < 3; ++mode) {
            std::vector<int64_t> v(m);
            for (auto& x : v) x = mode == 0 ? (int64_t)(rng() >> 1) : mode == 1 ? (int64_t)(rng() % 17) : (int64_t)(rng() % 1000) * (int64_t)(rng() % 1000);
            std::vector<int64_t> want = v;
            std::sort(want.begin(), want.end());
            gp::parallel_sort(v, [](int64_t a, int64_t b) { return a < b; });
            if (v != want) { std::printf("parallel_sort m=%ld mode=%d\n", (long)m, mode); return 1; }
        }
    }
    // text lines with shared prefixes longer than the 16-byte key
    std::vector<std::string> lines;
    for (int i = 0; i < 30000; ++i) {
        std::string s = "PSD" + std::string(10 + rng() % 12, '1') + std::to_string(rng() % 50) + " " + std::to_string((int)(rng() % 21) - 10) + " ";
        lines.push_back(s);
    }
    std::string text;
    std::vector<size_t> off;
    for (auto& l : lines) { off.push_back(text.size()); text += l; }
    off.push_back(text.size());
    std::vector<int64_t> idx(lines.size());
    std::iota(idx.begin(), idx.end(), 0);
    gp::sort_lines(gp::LineSet{&text, &off, 3}, idx);
    for (size_t i = 1; i < idx.size(); ++i)
        if (lines[idx[i - 1]] > lines[idx[i]]) { std::puts("sort_lines"); return 1; }
    std::puts("ok");
    return 0;
}
'''


def test_parallel_sort_and_sort_lines(tmp_path):
    src = tmp_path / "ls.cpp"
    exe = tmp_path / "ls"
    src.write_text(PROG)
    subprocess.run(["g++", "-O2", "-std=c++20", "-pthread", "-I", os.path.join(ROOT, "golaypair_amd", "csrc"),
                    str(src), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
