// golaypair -- drop-in for the reference's `main` (src/main.cpp:14-251).
//
//   golaypair [all|gen|sort|match|uncompress|filter]... <order>
//             [-c c0 c1 ...] [--paf p] [-j id] [-n count] [-d dir]
//   gen -o prefix | sort -i prefix -o prefix | match -i prefix -o prefix -a bool |
//   uncompress -i -o --internal --sorted -l -b -e -s -p prefix -d cpu|gpu | filter -i -o
//
// Same subcommands, flags, results-directory names, file formats and exit
// codes.  Every stage's compute runs through the C ABI (include/golay_gpu.h) on
// the GPU; this file only parses arguments and reads/writes the text files.
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <string_view>
#include <vector>

#include "golay_gpu.h"
#include "textfmt.hpp"

namespace {

struct Options {  // main.cpp:14-25
    int order = 0;
    bool have_order = false;
    int paf = 0;
    int job_id = 0;
    int job_count = 1;
    std::vector<int> compress{1};
    std::string dir = "results";
    std::string work_dir() const { return dir + "/order-" + std::to_string(order); }
};

gp_ctx* g_ctx = nullptr;
bool g_stats = false;

[[noreturn]] void die_gp(const char* what, int rc) {
    std::cerr << what << " failed (" << rc << "): " << gp_last_error(g_ctx) << "\n";
    std::exit(1);
}

void report(const char* stage, const gp_stats& s) {
    if (!g_stats) return;
    std::fprintf(stderr,
                 "{\"stage\": \"%s\", \"words\": %llu, \"psd_decisions\": %llu, \"psd_pass\": %llu, "
                 "\"band_rechecks\": %llu, \"canonical\": %llu, \"pair_checks\": %llu, \"pairs\": %llu, "
                 "\"tie_band\": %llu, \"kernel_ms\": %.3f, \"psd_kernel_ms\": %.3f}\n",
                 stage, (unsigned long long)s.words, (unsigned long long)s.psd_decisions,
                 (unsigned long long)s.psd_pass, (unsigned long long)s.band_rechecks, (unsigned long long)s.canonical,
                 (unsigned long long)s.pair_checks, (unsigned long long)s.pairs, (unsigned long long)s.tie_band,
                 s.kernel_ms, s.psd_kernel_ms);
}

bool is_int(const std::string& s) {
    if (s.empty()) return false;
    size_t i = (s[0] == '-' || s[0] == '+') ? 1 : 0;
    if (i >= s.size()) return false;
    for (; i < s.size(); ++i)
        if (s[i] < '0' || s[i] > '9') return false;
    return true;
}

bool read_file(const std::string& path, std::string& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f.good()) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

void write_file(const std::string& path, const std::string& text, bool append = false) {
    std::ofstream f(path, std::ios::binary | (append ? std::ios::app : std::ios::trunc));
    f << text;
}

int mkdirs(const std::string& path) {
    std::string cur;
    for (size_t i = 0; i <= path.size(); ++i) {
        if (i == path.size() || path[i] == '/')
            if (!cur.empty() && mkdir(cur.c_str(), 0777) != 0 && errno != EEXIST) return -1;
        if (i < path.size()) cur += path[i];
    }
    return 0;
}

std::vector<std::string_view> split_lines(const std::string& t) {
    std::vector<std::string_view> out;
    size_t pos = 0;
    while (pos < t.size()) {
        size_t e = t.find('\n', pos);
        if (e == std::string::npos) { out.emplace_back(t.data() + pos, t.size() - pos); break; }
        out.emplace_back(t.data() + pos, e - pos);
        pos = e + 1;
    }
    return out;
}

// --- host record sets -------------------------------------------------------

struct Cands {
    int n = 0;
    std::vector<int8_t> words;
    std::vector<uint8_t> keys;
    int64_t size() const { return n ? (int64_t)words.size() / n : 0; }
    gp_cands view() {
        gp_cands c{};
        c.n = n;
        c.nkey = n / 2;
        c.count = size();
        c.capacity = size();
        c.words = words.data();
        c.keys = keys.data();
        return c;
    }
};

struct Pairs {
    int n = 0;
    std::vector<int8_t> a, b;
    int64_t size() const { return n ? (int64_t)a.size() / n : 0; }
};

std::string cands_text(const Cands& c) {
    std::string s;
    s.reserve((size_t)c.size() * (4 * c.n + 8));
    for (int64_t i = 0; i < c.size(); ++i) {
        gp::append_cand_line(s, &c.words[i * c.n], &c.keys[i * (c.n / 2)], c.n);
        s.push_back('\n');
    }
    return s;
}

std::string pairs_text(const Pairs& p) {
    std::string s;
    s.reserve((size_t)p.size() * (6 * p.n + 4));
    for (int64_t i = 0; i < p.size(); ++i) {
        gp::append_pair_line(s, &p.a[i * p.n], &p.b[i * p.n], p.n);
        s.push_back('\n');
    }
    return s;
}

// Candidate lines "PSD<digits> v v v ... " -> records; the key digits are
// recomputed on the GPU and checked against the line's key text.
bool parse_cands(const std::string& text, int n, int bound, bool prefer_inverted, Cands& out, std::string& err) {
    auto lines = split_lines(text);
    out.n = n;
    out.words.assign(lines.size() * n, 0);
    std::vector<std::string_view> keytext(lines.size());
    size_t r = 0;
    for (auto l : lines) {
        size_t p = 0;
        while (p < l.size() && l[p] == ' ') ++p;
        if (p >= l.size()) continue;
        size_t e = l.find(' ', p);
        if (e == std::string_view::npos) { err = "malformed candidate line"; return false; }
        keytext[r] = l.substr(p, e - p);
        int got = 0;
        const char* s = l.data() + e;
        const char* end = l.data() + l.size();
        while (got < n) {
            while (s < end && *s == ' ') ++s;
            if (s >= end) break;
            char* q = nullptr;
            long v = std::strtol(s, &q, 10);
            if (q == s) break;
            out.words[r * n + got++] = (int8_t)v;
            s = q;
        }
        if (got != n) { err = "candidate line with fewer than n values"; return false; }
        ++r;
    }
    out.words.resize(r * n);
    keytext.resize(r);
    out.keys.assign(r * (n / 2), 0);
    if (r == 0) return true;
    std::vector<uint8_t> alt(r * (n / 2));
    gp_stats st{};
    int rc = gp_psd_keys(g_ctx, out.words.data(), n, n, (int64_t)r, bound, prefer_inverted ? 1 : 0, out.keys.data(), &st);
    int rc2 = gp_psd_keys(g_ctx, out.words.data(), n, n, (int64_t)r, bound, prefer_inverted ? 0 : 1, alt.data(), &st);
    std::string buf;
    for (size_t i = 0; i < r; ++i) {
        buf.clear();
        if (rc == GP_OK) gp::append_key_text(buf, &out.keys[i * (n / 2)], n);
        if (rc == GP_OK && buf == keytext[i]) continue;
        buf.clear();
        if (rc2 == GP_OK) gp::append_key_text(buf, &alt[i * (n / 2)], n);
        if (rc2 == GP_OK && buf == keytext[i]) {
            std::memcpy(&out.keys[i * (n / 2)], &alt[i * (n / 2)], n / 2);
            continue;
        }
        err = "key text does not match the PSD of its sequence: " + std::string(keytext[i]);
        return false;
    }
    return true;
}

// Compact binary candidate files (--binary): "GPCB", format version 1, n,
// K = n/2, a zero byte, the record count (8 bytes, little endian), then per
// candidate its n letters (int8) and its K key values (uint8, the numbers the
// text form prints after "PSD").  About a third of the text size at n = 15;
// the text form stays the default (byte-identical to the reference's files).
constexpr char kBinMagic[4] = {'G', 'P', 'C', 'B'};

std::string cands_binary(const Cands& c) {
    const int K = c.n / 2;
    const uint64_t m = (uint64_t)c.size();
    std::string s(16 + m * (c.n + K), '\0');
    std::memcpy(&s[0], kBinMagic, 4);
    s[4] = 1;
    s[5] = (char)c.n;
    s[6] = (char)K;
    for (int b = 0; b < 8; ++b) s[8 + b] = (char)((m >> (8 * b)) & 0xFF);
    char* p = &s[16];
    for (uint64_t i = 0; i < m; ++i) {
        std::memcpy(p, &c.words[i * c.n], c.n);
        std::memcpy(p + c.n, &c.keys[i * K], K);
        p += c.n + K;
    }
    return s;
}

bool is_binary_cands(const std::string& data) { return data.size() >= 16 && std::memcmp(data.data(), kBinMagic, 4) == 0; }

bool parse_binary_cands(const std::string& data, int n, Cands& out, std::string& err) {
    const int K = n / 2;
    uint64_t m = 0;
    for (int b = 0; b < 8; ++b) m |= (uint64_t)(uint8_t)data[8 + b] << (8 * b);
    if (data[4] != 1 || (uint8_t)data[5] != n || (uint8_t)data[6] != K || data.size() != 16 + m * (n + K)) {
        err = "binary candidate file: bad header or size";
        return false;
    }
    out.n = n;
    out.words.resize(m * n);
    out.keys.resize(m * K);
    const char* p = data.data() + 16;
    for (uint64_t i = 0; i < m; ++i) {
        std::memcpy(&out.words[i * n], p, n);
        std::memcpy(&out.keys[i * K], p + n, K);
        p += n + K;
    }
    return true;
}

// Either form of a candidate file.
bool parse_any_cands(const std::string& data, int n, int bound, bool prefer_inverted, Cands& out, std::string& err) {
    if (is_binary_cands(data)) return parse_binary_cands(data, n, out, err);
    return parse_cands(data, n, bound, prefer_inverted, out, err);
}

// PairReader (io.cpp:39-73): one pair per line, exactly 2*len integers.
bool read_pairs_lines(const std::string& path, int len, Pairs& out) {
    out.n = len;
    std::string text;
    if (!read_file(path, text)) return true;  // a missing file reads as no pairs
    size_t pos = 0, line_no = 0;
    std::vector<int> vals;
    while (pos < text.size()) {
        size_t e = text.find('\n', pos);
        bool terminated = e != std::string::npos;
        std::string_view line(text.data() + pos, (terminated ? e : text.size()) - pos);
        pos = terminated ? e + 1 : text.size();
        vals.clear();
        std::string tmp(line);
        const char* s = tmp.c_str();
        while (*s) {
            while (*s == ' ' || *s == '\t' || *s == '\r') ++s;
            if (!*s) break;
            char* q = nullptr;
            long v = std::strtol(s, &q, 10);
            if (q == s) break;
            vals.push_back((int)v);
            s = q;
        }
        if ((int)vals.size() != 2 * len) {
            std::cerr << "terminate called after throwing an instance of 'std::runtime_error'\n  what():  PairReader: "
                      << "input line " << line_no << ": length is " << vals.size() << ". Expected " << 2 * len << "\n";
            std::abort();
        }
        ++line_no;
        if (!terminated) break;
        for (int i = 0; i < len; ++i) out.a.push_back((int8_t)vals[i]);
        for (int i = 0; i < len; ++i) out.b.push_back((int8_t)vals[len + i]);
    }
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    Options o;
    bool s_all = false, s_gen = false, s_sort = false, s_match = false, s_unc = false, s_filter = false;
    std::string gen_prefix = "filtered";
    std::string sort_in = "filtered", sort_out = "filtered_sorted";
    std::string match_in = "filtered_sorted", match_out = "pairs";
    bool match_append = false;
    bool binary = false;  // --binary: compact binary candidate files
    std::string unc_in = "pairs", unc_out = "pairs", unc_tmp = "";
    // input lines to uncompress (main.cpp:110-113,192-196 declare these; the
    // reference ignores them): --line (1-indexed) or the 0-indexed range
    // [begin, end) with step, as script/uncompress.sh uses them to shard
    long long unc_line = 0;
    unsigned long long unc_begin = 0, unc_end = ~0ull, unc_step = 1;
    std::string ctx_name;
    std::vector<std::string> args(argv + 1, argv + argc);
    for (size_t i = 0; i < args.size(); ++i) {
        const std::string& a = args[i];
        auto need = [&](std::string& dst) {
            if (i + 1 >= args.size()) { std::cerr << a << " requires an argument\n"; std::exit(106); }
            dst = args[++i];
        };
        std::string v;
        if (a == "-h" || a == "--help") {
            std::cout << "Complementary Pairs Pipeline (MI355X)\nUsage: golaypair [SUBCOMMANDS] order [-c c0 c1 ..] "
                         "[--paf p] [-j id] [-n count] [-d dir]\nSubcommands: all gen sort match uncompress filter\n";
            return 0;
        }
        if (a == "all" || a == "gen" || a == "sort" || a == "match" || a == "uncompress" || a == "filter") {
            ctx_name = a;
            if (a == "all") s_all = true; else if (a == "gen") s_gen = true; else if (a == "sort") s_sort = true;
            else if (a == "match") s_match = true; else if (a == "uncompress") s_unc = true; else s_filter = true;
        } else if (a == "-c" || a == "--compress") {
            o.compress.clear();
            while (i + 1 < args.size() && is_int(args[i + 1])) o.compress.push_back(std::stoi(args[++i]));
            if (o.compress.empty()) { std::cerr << "--compress requires at least one value\n"; return 106; }
        } else if (a == "--paf") { need(v); o.paf = std::stoi(v); }
        else if (a == "--binary") binary = true;
        else if (a == "-j" || a == "--jobid") { need(v); o.job_id = std::stoi(v); }
        else if (a == "-n" || a == "--numjob") { need(v); o.job_count = std::stoi(v); }
        else if ((a == "-d" || a == "--device") && ctx_name == "uncompress") {
            need(v);
            if (v != "cpu" && v != "gpu" && v != "CPU" && v != "GPU") { std::cerr << "--device: cpu or gpu\n"; return 105; }
        } else if (a == "-d" || a == "--dir") { need(o.dir); }
        else if ((a == "-o" || a == "--output") && ctx_name == "gen") need(gen_prefix);
        else if ((a == "-i" || a == "--input") && ctx_name == "sort") need(sort_in);
        else if ((a == "-o" || a == "--output") && ctx_name == "sort") need(sort_out);
        else if ((a == "-i" || a == "--input") && ctx_name == "match") need(match_in);
        else if ((a == "-o" || a == "--output") && ctx_name == "match") need(match_out);
        else if ((a == "-a" || a == "--append") && ctx_name == "match") { need(v); match_append = (v == "1" || v == "true"); }
        else if ((a == "-i" || a == "--input") && (ctx_name == "uncompress" || ctx_name == "filter")) need(unc_in);
        else if ((a == "-o" || a == "--output") && (ctx_name == "uncompress" || ctx_name == "filter")) need(unc_out);
        else if ((a == "-p" || a == "--prefix") && ctx_name == "uncompress") need(unc_tmp);
        else if ((a == "-l" || a == "--line") && ctx_name == "uncompress") { need(v); unc_line = std::stoll(v); }
        else if ((a == "-b" || a == "--range_begin") && ctx_name == "uncompress") { need(v); unc_begin = std::stoull(v); }
        else if ((a == "-e" || a == "--range_end") && ctx_name == "uncompress") { need(v); unc_end = std::stoull(v); }
        else if ((a == "-s" || a == "--range_step") && ctx_name == "uncompress") { need(v); unc_step = std::stoull(v); }
        else if ((a == "--internal" || a == "--sorted") && ctx_name == "uncompress") { need(v); }
        else if (is_int(a) && !o.have_order) { o.order = std::stoi(a); o.have_order = true; }
        else { std::cerr << "The following argument was not expected: " << a << "\n"; return 109; }
    }
    if (!o.have_order) { std::cerr << "order is required\n"; return 106; }
    if (unc_step == 0 || unc_line < 0) { std::cerr << "--range_step must be >= 1 and --line >= 1\n"; return 105; }
    if (!(s_all || s_gen || s_sort || s_match || s_unc || s_filter)) { std::cerr << "A subcommand is required\n"; return 106; }
    g_stats = std::getenv("GOLAY_STATS") != nullptr;
    if (mkdirs(o.work_dir()) != 0) { std::cerr << "Failed to create temp directories\n"; return 1; }
    int rc = gp_create(&g_ctx, 0);
    if (rc != GP_OK) { std::cerr << "golaypair: no usable GPU (gp_create returned " << rc << ")\n"; return 1; }
    // GOLAY_PSD=fp32 selects the plain FP32 PSD filter (same outputs, slower)
    if (const char* m = std::getenv("GOLAY_PSD"); m && std::string(m) == "fp32") gp_set_psd_mode(g_ctx, GP_PSD_FP32);
    const std::string wd = o.work_dir();
    const std::string V = std::to_string(o.order);
    auto cand = [&](const std::string& prefix, char ab, int j) { return wd + "/" + V + "-" + prefix + "-" + ab + "#" + std::to_string(j); };
    auto sorted_name = [&](const std::string& prefix, char ab) { return wd + "/" + V + "-" + prefix + "-" + ab; };
    auto match_name = [&](const std::string& prefix, size_t ci) { return wd + "/" + V + "-" + prefix + "-" + std::to_string(o.compress[ci]); };
    const int bound = 2 * o.order - o.paf;

    if (s_gen || s_all) {  // main.cpp:47-54, generate_hybrid.cpp:107-163
        const int n = o.order / o.compress[0];
        Cands A, B;
        A.n = B.n = n;
        gp_stats st{};
        int64_t cap = 1 << 16;
        for (;;) {
            A.words.resize(cap * n); A.keys.resize(cap * (n / 2));
            B.words.resize(cap * n); B.keys.resize(cap * (n / 2));
            gp_cands ca{n, n / 2, 0, cap, A.words.data(), A.keys.data()};
            gp_cands cb{n, n / 2, 0, cap, B.words.data(), B.keys.data()};
            rc = gp_generate(g_ctx, o.order, o.compress[0], o.paf, o.job_id, o.job_count, &ca, &cb, &st);
            if (rc == GP_ERR_CAPACITY) { cap = std::max(ca.count, cb.count) + 1; continue; }
            if (rc == GP_NO_INPUT) {
                write_file(cand(gen_prefix, 'a', o.job_id), "");
                write_file(cand(gen_prefix, 'b', o.job_id), "");
                std::cerr << "Order " << o.order << " cannot be decomposed into a sum of squares\n" << o.order << "*2\n";
                std::cerr << "Hence there are no solutions.\n";
                return 1;
            }
            if (rc < 0) die_gp("gp_generate", rc);
            A.words.resize(ca.count * n); A.keys.resize(ca.count * (n / 2));
            B.words.resize(cb.count * n); B.keys.resize(cb.count * (n / 2));
            break;
        }
        write_file(cand(gen_prefix, 'a', o.job_id), binary ? cands_binary(A) : cands_text(A));
        write_file(cand(gen_prefix, 'b', o.job_id), binary ? cands_binary(B) : cands_text(B));
        std::cout << "Passed PSD filter: " << (A.size() + B.size()) << "\n";
        report("gen", st);
    }
    if (s_sort || s_all) {  // main.cpp:69-79, sort.cpp:7-23
        std::cout << "Sorting Candidates\n";
        for (char ab : {'a', 'b'}) {
            // shards of either form -> records -> sorted, unique on the GPU in
            // the byte order of the text lines -> one file in the chosen form
            const int n = o.order / o.compress[0];
            Cands all;
            all.n = n;
            for (int j = 0; j < o.job_count; ++j) {
                std::string t, err;
                Cands c;
                if (!read_file(cand(sort_in, ab, j), t) || !parse_any_cands(t, n, bound, ab == 'b', c, err)) {
                    std::cerr << "Sort failed\n";
                    return 1;
                }
                all.words.insert(all.words.end(), c.words.begin(), c.words.end());
                all.keys.insert(all.keys.end(), c.keys.begin(), c.keys.end());
            }
            gp_cands v = all.view();
            if (all.size() > 0 && (rc = gp_sort_candidates(g_ctx, &v)) < 0) die_gp("gp_sort_candidates", rc);
            all.words.resize(v.count * n);
            all.keys.resize(v.count * (n / 2));
            write_file(sorted_name(sort_out, ab), binary ? cands_binary(all) : cands_text(all));
        }
    }
    if (s_match || s_all) {  // main.cpp:92-102, match_pairs.cpp:15-121
        std::cout << "Matching Candidates\n";
        const int n = o.order / o.compress[0];
        std::string ta, tb, err;
        read_file(sorted_name(match_in, 'a'), ta);
        read_file(sorted_name(match_in, 'b'), tb);
        Cands A, B;
        if (!parse_any_cands(ta, n, bound, false, A, err) || !parse_any_cands(tb, n, bound, true, B, err)) {
            std::cerr << "match: " << err << "\n";
            return 1;
        }
        const std::string out_path = match_name(match_out, 0);
        if (!match_append) write_file(out_path, "");
        if (A.size() == 0 || B.size() == 0) return 1;  // match_pairs.cpp:32-38
        Pairs P;
        P.n = n;
        gp_stats st{};
        int64_t cap = std::max<int64_t>(1024, 2 * (A.size() + B.size()));
        for (;;) {
            P.a.resize(cap * n); P.b.resize(cap * n);
            gp_pairs out{n, 0, cap, P.a.data(), P.b.data()};
            gp_cands ca = A.view(), cb = B.view();
            rc = gp_match(g_ctx, o.order, o.compress[0], o.paf, &ca, &cb, &out, &st);
            if (rc == GP_ERR_CAPACITY) { cap = out.count + 1; continue; }
            if (rc < 0) die_gp("gp_match", rc);
            P.a.resize(out.count * n); P.b.resize(out.count * n);
            break;
        }
        write_file(out_path, pairs_text(P), true);
        report("match", st);
    }
    if (s_unc || s_all) {  // main.cpp:117-139, uncompression.cpp:32-79
        for (size_t i = 0; i + 1 < o.compress.size(); ++i) {
            const std::string prefix = i == 0 ? unc_in : unc_out;
            const int L = o.order / o.compress[i];
            const int N = o.order / o.compress[i + 1];
            Pairs in;
            read_pairs_lines(match_name(prefix, i), L, in);
            if (i == 0 && (unc_line > 0 || unc_begin > 0 || unc_end != ~0ull || unc_step != 1)) {
                // the selected lines of the first level's input, in file order
                Pairs sel;
                sel.n = in.n;
                for (int64_t q = 0; q < in.size(); ++q) {
                    const uint64_t u = (uint64_t)q;
                    const bool take = unc_line > 0 ? q + 1 == unc_line
                                                   : u >= unc_begin && u < unc_end && (u - unc_begin) % unc_step == 0;
                    if (!take) continue;
                    sel.a.insert(sel.a.end(), in.a.begin() + q * L, in.a.begin() + (q + 1) * L);
                    sel.b.insert(sel.b.end(), in.b.begin() + q * L, in.b.begin() + (q + 1) * L);
                }
                in = std::move(sel);
            }
            const std::string out_path = match_name(unc_out, i + 1);
            write_file(out_path, "");
            if (o.compress[i] <= 1) continue;
            Pairs P;
            P.n = N;
            Cands ta, tb;
            ta.n = tb.n = N;
            gp_stats st{};
            int64_t cap = std::max<int64_t>(1024, 4 * in.size()), tcap = 1 << 16;
            for (;;) {
                P.a.resize(cap * N); P.b.resize(cap * N);
                ta.words.resize(tcap * N); ta.keys.resize(tcap * (N / 2));
                tb.words.resize(tcap * N); tb.keys.resize(tcap * (N / 2));
                gp_pairs pin{L, in.size(), in.size(), in.a.data(), in.b.data()};
                gp_pairs out{N, 0, cap, P.a.data(), P.b.data()};
                gp_cands ca{N, N / 2, 0, tcap, ta.words.data(), ta.keys.data()};
                gp_cands cb{N, N / 2, 0, tcap, tb.words.data(), tb.keys.data()};
                st = gp_stats{};
                rc = gp_uncompress(g_ctx, o.order, o.compress[i], o.compress[i + 1], o.paf, &pin, &out, &ca, &cb, &st);
                if (rc == GP_ERR_CAPACITY) {
                    cap = std::max(cap, out.count + 1);
                    tcap = std::max({tcap, ca.count + 1, cb.count + 1});
                    continue;
                }
                if (rc < 0) die_gp("gp_uncompress", rc);
                P.a.resize(out.count * N); P.b.resize(out.count * N);
                ta.words.resize(ca.count * N); ta.keys.resize(ca.count * (N / 2));
                tb.words.resize(cb.count * N); tb.keys.resize(cb.count * (N / 2));
                break;
            }
            write_file(out_path, pairs_text(P));
            if (in.size() > 0) {  // uncompression.cpp:45-48: the last pair's files stay behind
                const std::string base = wd + "/" + unc_tmp + "uncompress_tmp_";
                if (binary) {
                    write_file(base + "a", cands_binary(ta));
                    write_file(base + "b", cands_binary(tb));
                    for (Cands* c : {&ta, &tb}) {
                        gp_cands v = c->view();
                        if (c->size() > 0 && (rc = gp_sort_candidates(g_ctx, &v)) < 0) die_gp("gp_sort_candidates", rc);
                        c->words.resize(v.count * N);
                        c->keys.resize(v.count * (N / 2));
                    }
                    write_file(base + "a.sorted", cands_binary(ta));
                    write_file(base + "b.sorted", cands_binary(tb));
                } else {
                    write_file(base + "a", cands_text(ta));
                    write_file(base + "b", cands_text(tb));
                    for (Cands* c : {&ta, &tb}) {
                        gp_cands v = c->view();
                        if (c->size() > 0 && (rc = gp_sort_candidates(g_ctx, &v)) < 0) die_gp("gp_sort_candidates", rc);
                        c->words.resize(v.count * N);
                        c->keys.resize(v.count * (N / 2));
                    }
                    write_file(base + "a.sorted", cands_text(ta));
                    write_file(base + "b.sorted", cands_text(tb));
                }
            }
            report("uncompress", st);
        }
    }
    if (s_filter || s_all) {  // main.cpp:151-159, filter.cpp:18-163
        std::cout << "Filtering pairs of compression size " << o.compress.back() << "\n";
        const std::string in_name = match_name("pairs", o.compress.size() - 1);
        const std::string out_name = wd + "/" + V + "-unique";
        std::cout << in_name << "\n";
        std::printf("Filtering Equivalent Pairs...\n");
        std::fflush(stdout);
        write_file(out_name, "");
        std::string text;
        if (!read_file(in_name, text)) { std::printf("Bad File.\n"); return 0; }
        const int n = o.order / o.compress.back();
        // token stream: n values of a, then n of b (filter.cpp:46-63)
        std::vector<long> toks;
        {
            const char* s = text.c_str();
            while (*s) {
                while (*s && std::isspace((unsigned char)*s)) ++s;
                if (!*s) break;
                char* q = nullptr;
                long v = std::strtol(s, &q, 10);
                if (q == s) break;
                toks.push_back(v);
                s = q;
            }
        }
        if (toks.empty()) {  // the reference's stoi("") on an empty file
            std::cerr << "terminate called after throwing an instance of 'std::invalid_argument'\n  what():  stoi\n";
            std::abort();
        }
        Pairs in;
        in.n = n;
        for (size_t r = 0; (r + 1) * 2 * n <= toks.size(); ++r) {
            for (int i = 0; i < n; ++i) in.a.push_back((int8_t)toks[r * 2 * n + i]);
            for (int i = 0; i < n; ++i) in.b.push_back((int8_t)toks[r * 2 * n + n + i]);
        }
        std::printf("%lld sequences loaded.\n", (long long)in.size());
        Pairs P;
        P.n = n;
        P.a.resize(std::max<int64_t>(1, in.size()) * n);
        P.b.resize(std::max<int64_t>(1, in.size()) * n);
        gp_pairs pin{n, in.size(), in.size(), in.a.data(), in.b.data()};
        gp_pairs out{n, 0, std::max<int64_t>(1, in.size()), P.a.data(), P.b.data()};
        gp_stats st{};
        rc = gp_filter(g_ctx, o.order, o.compress.back(), &pin, &out, &st);
        if (rc < 0) die_gp("gp_filter", rc);
        P.a.resize(out.count * n);
        P.b.resize(out.count * n);
        std::printf("%lld unique sequences found.\n", (long long)P.size());
        write_file(out_name, pairs_text(P));
        report("filter", st);
    }
    gp_destroy(g_ctx);
    return 0;
}
