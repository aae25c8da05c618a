"""Host thread pool (golaypair_amd/csrc/threadpool.hpp): every index runs once,
nested and concurrent calls complete, an exception reaches the caller."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROG = r'''
#include <atomic>
#include <cstdio>
#include <stdexcept>
#include <thread>
#include <vector>
#include "threadpool.hpp"
int main() {
    std::vector<std::atomic<int>> hit(100000);
    gp::pool().parallel_for(hit.size(), [&](size_t i) { hit[i]++; });
    for (auto& h : hit) if (h != 1) { std::puts("coverage"); return 1; }
    // nested calls from every index, from several host threads at once
    std::atomic<long> sum{0};
    std::vector<std::thread> th;
    for (int t = 0; t < 4; ++t)
        th.emplace_back([&] {
            gp::pool().parallel_for(64, [&](size_t i) {
                gp::pool().parallel_for(100, [&](size_t j) { sum += (long)(i * 100 + j) % 7; });
            });
        });
    for (auto& x : th) x.join();
    long want = 0;
    for (int i = 0; i < 64; ++i) for (int j = 0; j < 100; ++j) want += (i * 100 + j) % 7;
    if (sum != 4 * want) { std::puts("nested"); return 1; }
    bool caught = false;
    try {
        gp::pool().parallel_for(1000, [&](size_t i) { if (i == 517) throw std::runtime_error("x"); });
    } catch (const std::runtime_error&) { caught = true; }
    if (!caught) { std::puts("exception"); return 1; }
    std::puts("ok");
    return 0;
}
'''


def test_thread_pool(tmp_path):
    src = tmp_path / "pool.cpp"
    src.write_text(PROG)
    exe = tmp_path / "pool"
    subprocess.run(["g++", "-O2", "-std=c++20", "-pthread", "-I", os.path.join(ROOT, "golaypair_amd", "csrc"),
                    str(src), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
