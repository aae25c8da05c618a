// C entry points of the oracle for ctypes callers (tests/ and bench.py's
// cpu_baseline).  TEST INFRASTRUCTURE ONLY -- never loaded by golaypair_amd/.
#include <chrono>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "refcpu.hpp"

using refcpu::Seq;

namespace {
std::mutex g_mu;
std::map<std::pair<int, int>, std::set<Seq>> g_gens;

const std::set<Seq>& gens(int flag, int n) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto key = std::make_pair(flag, n);
    auto it = g_gens.find(key);
    if (it == g_gens.end()) it = g_gens.emplace(key, refcpu::construct_generators(flag, n)).first;
    return it->second;
}

int64_t copy_out(const std::string& s, char* out, int64_t cap) {
    if (out && cap > 0) {
        int64_t m = std::min<int64_t>(cap - 1, (int64_t)s.size());
        std::memcpy(out, s.data(), (size_t)m);
        out[m] = 0;
    }
    return (int64_t)s.size();
}
}  // namespace

extern "C" {

// fourier.cpp:25-43
int ref_psd(const int32_t* seq, int n, double* out) {
    std::vector<int> s(seq, seq + n);
    refcpu::psd(s.data(), n, out);
    return 0;
}

// Batched PSD + filter for arbitrary words (row i: lens[i] int8 entries at
// words + i*stride).  pass[i] = psd_filter(...); maxpsd[i] = max_{1<=k<=n/2} psd[k].
int ref_psd_filter_batch(const int8_t* words, int stride, const int32_t* lens, const int32_t* orders,
                         const int32_t* pafs, int64_t count, uint8_t* pass, double* maxpsd,
                         double* psd_out, int psd_stride) {
    std::vector<int> s;
    std::vector<double> p;
    for (int64_t i = 0; i < count; ++i) {
        int n = lens[i];
        s.assign(words + i * stride, words + i * stride + n);
        p.resize(n);
        refcpu::psd(s.data(), n, p.data());
        if (pass) pass[i] = refcpu::psd_filter(p, orders[i], pafs[i]) ? 1 : 0;
        if (maxpsd) {
            double m = 0;
            for (int k = 1; k <= n / 2; ++k) m = std::max(m, p[k]);
            maxpsd[i] = m;
        }
        if (psd_out)
            for (int k = 0; k < n; ++k) psd_out[i * psd_stride + k] = p[k];
    }
    return 0;
}

// Threaded variant for the CPU baseline: returns wall seconds.
double ref_psd_filter_batch_timed(const int8_t* words, int stride, const int32_t* lens, const int32_t* orders,
                                  const int32_t* pafs, int64_t count, uint8_t* pass, int threads) {
    auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> th;
    for (int t = 0; t < threads; ++t) {
        th.emplace_back([=] {
            int64_t b = count * t / threads, e = count * (t + 1) / threads;
            ref_psd_filter_batch(words + b * stride, stride, lens + b, orders + b, pafs + b, e - b, pass + b,
                                 nullptr, nullptr, 0);
        });
    }
    for (auto& x : th) x.join();
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// orderly_equivalence.cpp:53-68 with constructGenerators(flag, n).
int ref_is_canonical(const int32_t* seq, int n, int flag) {
    Seq s(seq, seq + n);
    return refcpu::is_canonical(s, gens(flag, n)) ? 1 : 0;
}

int64_t ref_group_size(int flag, int n) { return (int64_t)gens(flag, n).size(); }

// Generator permutations (1-based entries), row-major |G| x n.
int64_t ref_group_perms(int flag, int n, int32_t* out, int64_t cap_rows) {
    const auto& g = gens(flag, n);
    int64_t r = 0;
    for (const Seq& s : g) {
        if (r < cap_rows)
            for (int i = 0; i < n; ++i) out[r * n + i] = s[i];
        ++r;
    }
    return r;
}

// equivalence.cpp:84-95 generator templates: rows of 2n signed entries.
int64_t ref_pair_group(int len, int altnegative, int32_t* out, int64_t cap_rows) {
    auto g = refcpu::construct_pair_generators(len, altnegative != 0);
    int64_t r = 0;
    for (const auto& it : g) {
        if (r < cap_rows) {
            for (int i = 0; i < len; ++i) out[r * 2 * len + i] = it.a[i];
            for (int i = 0; i < len; ++i) out[r * 2 * len + len + i] = it.b[i];
        }
        ++r;
    }
    return r;
}

// equivalence.cpp:97-144 applied to one pair with a freshly built group;
// rows of 2n ints (a then b) in set order.
int64_t ref_class_pairs(int len, int altnegative, const int32_t* a, const int32_t* b, int32_t* out,
                        int64_t cap_rows) {
    auto g = refcpu::construct_pair_generators(len, altnegative != 0);
    refcpu::PairItem p{Seq(a, a + len), Seq(b, b + len)};
    auto cls = refcpu::generate_class_pairs(g, p);
    int64_t r = 0;
    for (const auto& it : cls) {
        if (r < cap_rows) {
            for (int i = 0; i < len; ++i) out[r * 2 * len + i] = it.a[i];
            for (int i = 0; i < len; ++i) out[r * 2 * len + len + i] = it.b[i];
        }
        ++r;
    }
    return r;
}

int ref_check_if_pair(const int32_t* a, const int32_t* b, int n, int paf, int compress) {
    return refcpu::check_if_pair(Seq(a, a + n), Seq(b, b + n), paf, compress) ? 1 : 0;
}

// generate_hybrid.cpp:107-163 -> text of the A and B candidate files.
int64_t ref_generate(int order, int compress, int paf, int job, int njobs, char* out_a, int64_t cap_a,
                     char* out_b, int64_t cap_b, int64_t* sizes) {
    std::vector<std::string> la, lb;
    refcpu::GenStats st;
    int rc = refcpu::generate(order, compress, paf, job, njobs, la, lb, &st, false);
    std::string ta, tb;
    for (auto& l : la) { ta += l; ta += "\n"; }
    for (auto& l : lb) { tb += l; tb += "\n"; }
    if (sizes) { sizes[0] = copy_out(ta, out_a, cap_a); sizes[1] = copy_out(tb, out_b, cap_b); sizes[2] = (int64_t)st.orderly_count; }
    return rc;
}

// uncompression.cpp:81-210 on one sequence, job proc_id of proc_num; returns
// the candidate-file text size (lines in leaf DFS order); *leaves = leaves examined.
int64_t ref_uncompress_side(const int32_t* orig, int len, int compress, int new_compress, int paf, int seqflag,
                            int proc_id, int proc_num, char* out, int64_t cap, uint64_t* leaves) {
    Seq o(orig, orig + len);
    std::vector<std::string> lines;
    uint64_t nl = 0;
    refcpu::uncompress_recursive(o, compress, new_compress, paf, seqflag, lines, &nl, proc_id, proc_num);
    if (leaves) *leaves = nl;
    std::string t;
    for (auto& l : lines) { t += l; t += "\n"; }
    return copy_out(t, out, cap);
}

uint64_t ref_tie_band(void) { return refcpu::tie_band_count(); }
void ref_tie_band_reset(void) { refcpu::tie_band_reset(); }

}  // extern "C"
