// refcpu CLI -- the oracle pipeline behind the reference's command line.
//
// TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline).  Restates
// src/main.cpp:14-251: same subcommands, flags, results-directory file names
// and exit codes.  Text stays in memory between stages exactly as the
// reference's files hold it; every file the reference writes is written.
#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "refcpu.hpp"

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

std::vector<std::string> split_lines(const std::string& t) {
    std::vector<std::string> out;
    size_t pos = 0;
    while (pos < t.size()) {
        size_t e = t.find('\n', pos);
        if (e == std::string::npos) { out.push_back(t.substr(pos)); break; }
        out.push_back(t.substr(pos, e - pos));
        pos = e + 1;
    }
    return out;
}

std::string join(const std::vector<std::string>& v) {
    std::string o;
    for (auto& l : v) { o += l; o += "\n"; }
    return o;
}

int mkdirs(const std::string& path) {
    std::string cur;
    for (size_t i = 0; i <= path.size(); ++i) {
        if (i == path.size() || path[i] == '/') {
            if (!cur.empty() && mkdir(cur.c_str(), 0777) != 0 && errno != EEXIST) return -1;
        }
        if (i < path.size()) cur += path[i];
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    Options o;
    bool s_all = false, s_gen = false, s_sort = false, s_match = false, s_unc = false, s_filter = false;
    std::string gen_prefix = "filtered";                          // main.cpp:27-29
    std::string sort_in = "filtered", sort_out = "filtered_sorted";  // :56-59
    std::string match_in = "filtered_sorted", match_out = "pairs";   // :81-85
    bool match_append = false;
    std::string unc_in = "pairs", unc_out = "pairs", unc_tmp = "";   // :104-115
    std::string ctx;  // current subcommand context
    std::vector<std::string> args(argv + 1, argv + argc);
    for (size_t i = 0; i < args.size(); ++i) {
        const std::string& a = args[i];
        auto need = [&](std::string& dst) {
            if (i + 1 >= args.size()) { std::cerr << a << " requires an argument\n"; std::exit(106); }
            dst = args[++i];
        };
        std::string v;
        if (a == "all" || a == "gen" || a == "sort" || a == "match" || a == "uncompress" || a == "filter") {
            ctx = a;
            if (a == "all") s_all = true; else if (a == "gen") s_gen = true; else if (a == "sort") s_sort = true;
            else if (a == "match") s_match = true; else if (a == "uncompress") s_unc = true; else s_filter = true;
        } else if (a == "-c" || a == "--compress") {
            o.compress.clear();
            while (i + 1 < args.size() && is_int(args[i + 1])) o.compress.push_back(std::stoi(args[++i]));
        } else if (a == "--paf") { need(v); o.paf = std::stoi(v); }
        else if (a == "-j" || a == "--jobid") { need(v); o.job_id = std::stoi(v); }
        else if (a == "-n" || a == "--numjob") { need(v); o.job_count = std::stoi(v); }
        else if ((a == "-d" || a == "--device") && ctx == "uncompress") { need(v); }  // device: CPU only here
        else if (a == "-d" || a == "--dir") { need(o.dir); }
        else if ((a == "-o" || a == "--output") && ctx == "gen") need(gen_prefix);
        else if ((a == "-i" || a == "--input") && ctx == "sort") need(sort_in);
        else if ((a == "-o" || a == "--output") && ctx == "sort") need(sort_out);
        else if ((a == "-i" || a == "--input") && ctx == "match") need(match_in);
        else if ((a == "-o" || a == "--output") && ctx == "match") need(match_out);
        else if ((a == "-a" || a == "--append") && ctx == "match") { need(v); match_append = (v == "1" || v == "true"); }
        else if ((a == "-i" || a == "--input") && (ctx == "uncompress" || ctx == "filter")) need(unc_in);  // :208 binds filter -i here
        else if ((a == "-o" || a == "--output") && (ctx == "uncompress" || ctx == "filter")) need(unc_out);
        else if ((a == "-p" || a == "--prefix") && ctx == "uncompress") need(unc_tmp);
        else if ((a == "--internal" || a == "--sorted" || a == "-l" || a == "--line" || a == "-b" ||
                  a == "--range_begin" || a == "-e" || a == "--range_end" || a == "-s" || a == "--range_step") &&
                 ctx == "uncompress") { need(v); }  // parsed and ignored (main.cpp:117-139)
        else if (is_int(a) && !o.have_order) { o.order = std::stoi(a); o.have_order = true; }
        else { std::cerr << "The following argument was not expected: " << a << "\n"; return 109; }
    }
    if (!o.have_order) { std::cerr << "order is required\n"; return 106; }
    if (!(s_all || s_gen || s_sort || s_match || s_unc || s_filter)) { std::cerr << "A subcommand is required\n"; return 106; }
    if (mkdirs(o.work_dir()) != 0) { std::cerr << "Failed to create temp directories\n"; return 1; }
    const std::string wd = o.work_dir();
    const std::string V = std::to_string(o.order);
    auto cand = [&](const std::string& prefix, char ab, int j) {  // main.cpp:32-44
        return wd + "/" + V + "-" + prefix + "-" + ab + "#" + std::to_string(j);
    };
    auto sorted_name = [&](const std::string& prefix, char ab) { return wd + "/" + V + "-" + prefix + "-" + ab; };  // :62-67
    auto match_name = [&](const std::string& prefix, size_t ci) {  // :88-90
        return wd + "/" + V + "-" + prefix + "-" + std::to_string(o.compress[ci]);
    };
    refcpu::tie_band_reset();

    if (s_gen || s_all) {  // main.cpp:47-54
        std::vector<std::string> la, lb;
        refcpu::GenStats st;
        int rc = refcpu::generate(o.order, o.compress[0], o.paf, o.job_id, o.job_count, la, lb, &st, true);
        write_file(cand(gen_prefix, 'a', o.job_id), join(la));
        write_file(cand(gen_prefix, 'b', o.job_id), join(lb));
        if (rc > 0) return 1;
    }
    if (s_sort || s_all) {  // main.cpp:69-79
        std::cout << "Sorting Candidates\n";
        for (char ab : {'a', 'b'}) {
            std::vector<std::string> lines;
            for (int j = 0; j < o.job_count; ++j) {
                std::string t;
                if (!read_file(cand(sort_in, ab, j), t)) { std::cerr << "Sort failed\n"; return 1; }
                auto l = split_lines(t);
                lines.insert(lines.end(), l.begin(), l.end());
            }
            write_file(sorted_name(sort_out, ab), join(refcpu::sort_uniq(std::move(lines))));
        }
    }
    if (s_match || s_all) {  // main.cpp:92-102
        std::cout << "Matching Candidates\n";
        std::string ta, tb, out;
        read_file(sorted_name(match_in, 'a'), ta);
        read_file(sorted_name(match_in, 'b'), tb);
        int rc = refcpu::match_pairs(o.order, o.compress[0], o.paf, ta, tb, out);
        write_file(match_name(match_out, 0), out, match_append);
        if (rc) return 1;
    }
    if (s_unc || s_all) {  // main.cpp:117-139
        for (size_t i = 0; i + 1 < o.compress.size(); ++i) {
            std::string prefix = i == 0 ? unc_in : unc_out;
            std::string in, out;
            read_file(match_name(prefix, i), in);
            write_file(match_name(unc_out, i + 1), "");
            std::vector<std::string> tmp;
            refcpu::uncompress_pipeline(o.order, o.compress[i], o.compress[i + 1], o.paf, in, out, nullptr, &tmp);
            write_file(match_name(unc_out, i + 1), out);
            if (tmp.size() == 4) {  // uncompression.cpp:45-48, last pair's content
                const std::string base = wd + "/" + unc_tmp + "uncompress_tmp_";
                write_file(base + "a", tmp[0]);
                write_file(base + "b", tmp[1]);
                write_file(base + "a.sorted", tmp[2]);
                write_file(base + "b.sorted", tmp[3]);
            }
        }
    }
    if (s_filter || s_all) {  // main.cpp:151-159 (filter reads the default "pairs"/"unique" names)
        std::cout << "Filtering pairs of compression size " << o.compress.back() << "\n";
        std::string in, out;
        const std::string in_name = match_name("pairs", o.compress.size() - 1);
        std::cout << in_name << "\n";
        const std::string out_name = wd + "/" + V + "-unique";
        write_file(out_name, "");
        if (!read_file(in_name, in)) { std::cout << "Bad File.\n"; return 0; }
        refcpu::FilterStats fs;
        refcpu::cache_filter(o.order, o.compress.back(), in, out, &fs);
        std::cout << fs.unique << " unique sequences found.\n";
        write_file(out_name, out);
    }
    if (refcpu::tie_band_count())
        std::cerr << "warning: " << refcpu::tie_band_count() << " PSD values inside the 1e-9 tie band\n";
    return 0;
}
