// refcpu -- CPU restatement of the GolayPair search pipeline (oracle).
//
// TEST INFRASTRUCTURE ONLY: the checker for tests/, smoke() and bench.py's
// cpu_baseline.  Never linked into golaypair_amd/.  Each function cites the
// reference file:line (relative to the tylerlumsden/GolayPair checkout) that it
// restates.
#include "refcpu.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <map>
#include <numeric>
#include <stdexcept>

namespace refcpu {

// ---------------------------------------------------------------------------
// small math utilities

std::vector<int> getalphabet(int compress) {  // golay.cpp:60-75
    std::set<int> s;
    if (compress % 2 == 0) {
        for (int i = 0; i <= compress; i += 2) { s.insert(i); s.insert(-i); }
    } else {
        for (int i = 1; i <= compress; i += 2) { s.insert(i); s.insert(-i); }
    }
    return std::vector<int>(s.begin(), s.end());
}

int sum_constant(int order, int paf) { return order * 2 + (order - 1) * paf; }  // generate_hybrid.cpp:45-47

std::vector<std::pair<int, int>> getdecomps(int k) {  // decomps.cpp:6-17
    std::vector<std::pair<int, int>> out;
    for (int x = 0; x <= k; x++)
        for (int y = x; y <= k; y++)
            if (x * x + y * y == k) out.emplace_back(x, y);
    return out;
}

std::vector<int> getcoprimes(int k) {  // coprimes.cpp:12-33 (MAX_COPRIME_N = 400, coprimes.h:2)
    if (k > 400) {
        std::cerr << "ERROR: Tried to access coprimes to the integer " << k << "\n";
        std::abort();
    }
    std::vector<int> out;
    for (int i = 1; i < k; ++i)
        if (std::gcd(i, k) == 1) out.push_back(i);
    return out;
}

namespace {
// Minimal base-1e9 unsigned bignum for printed counts only.
struct Big {
    std::vector<uint32_t> d;  // little-endian limbs, base 1e9
    explicit Big(uint64_t v = 0) { while (v) { d.push_back(v % 1000000000u); v /= 1000000000u; } }
    void mul(uint32_t m) {
        uint64_t carry = 0;
        for (auto& x : d) { uint64_t t = (uint64_t)x * m + carry; x = t % 1000000000u; carry = t / 1000000000u; }
        while (carry) { d.push_back(carry % 1000000000u); carry /= 1000000000u; }
    }
    void add(const Big& o) {
        uint64_t carry = 0;
        for (size_t i = 0; i < std::max(d.size(), o.d.size()) || carry; ++i) {
            if (i == d.size()) d.push_back(0);
            uint64_t t = (uint64_t)d[i] + carry + (i < o.d.size() ? o.d[i] : 0);
            d[i] = t % 1000000000u; carry = t / 1000000000u;
        }
    }
    void div(uint32_t m) {
        uint64_t rem = 0;
        for (size_t i = d.size(); i-- > 0;) { uint64_t cur = d[i] + rem * 1000000000ull; d[i] = cur / m; rem = cur % m; }
        while (!d.empty() && d.back() == 0) d.pop_back();
    }
    std::string str() const {
        if (d.empty()) return "0";
        std::string s = std::to_string(d.back());
        char buf[16];
        for (size_t i = d.size() - 1; i-- > 0;) { snprintf(buf, sizeof buf, "%09u", d[i]); s += buf; }
        return s;
    }
};

// necklace_count with saturation, for the job-split threshold (lyndon.cpp:38-45, :91-93).
unsigned __int128 necklace_count_sat(int length, int k) {
    const unsigned __int128 cap = (unsigned __int128)1 << 100;
    unsigned __int128 sum = 0;
    for (int i = 1; i <= length; i++) {
        int e = std::gcd(i, length);
        unsigned __int128 p = 1;
        for (int j = 0; j < e && p < cap; j++) p *= (unsigned)k;
        sum += p;
        if (sum > cap) return cap;
    }
    return sum / length;
}
}  // namespace

std::string necklace_count_str(int length, int alphabet_size) {  // lyndon.cpp:38-45
    Big sum;
    for (int i = 1; i <= length; i++) {
        Big p(1);
        for (int j = 0, e = std::gcd(i, length); j < e; j++) p.mul((uint32_t)alphabet_size);
        sum.add(p);
    }
    sum.div((uint32_t)length);
    return sum.str();
}

// ---------------------------------------------------------------------------
// orderly generation

std::set<Seq> construct_generators(int flag, int len) {  // orderly_equivalence.cpp:17-51
    Seq base(len);
    for (int i = 1; i <= len; i++) base[i - 1] = i;
    const std::vector<int> coprimes = getcoprimes(len);
    std::set<Seq> map{base};
    size_t size = 0;
    // Closure until no new element appears.  The reference's early `return 0`
    // on `newseq < base` (:103-105, :128-130) cannot fire: base is the
    // identity, the lexicographically smallest permutation.
    while (map.size() != size) {
        size = map.size();
        if (flag == 0) {  // decimation_equivalence :122-135
            std::vector<Seq> snap(map.begin(), map.end());
            for (const Seq& seq : snap)
                for (int c : coprimes) {
                    Seq ns(len);
                    for (int i = 0; i < len; i++) ns[i] = seq[(i * c) % len];
                    map.insert(ns);
                }
        }
        {  // shift_equivalence :88-97
            std::vector<Seq> snap(map.begin(), map.end());
            for (Seq seq : snap)
                for (int i = 0; i < len; i++) {
                    std::rotate(seq.begin(), seq.begin() + 1, seq.end());
                    map.insert(seq);
                }
        }
        {  // reverse_equivalence :99-110
            std::vector<Seq> snap(map.begin(), map.end());
            for (Seq seq : snap) {
                std::reverse(seq.begin(), seq.end());
                map.insert(seq);
            }
        }
    }
    return map;
}

bool is_canonical(const Seq& seq, const std::set<Seq>& generators) {  // orderly_equivalence.cpp:53-68
    const int len = (int)seq.size();
    Seq ns(len);
    for (const Seq& item : generators) {
        for (int i = 0; i < len; i++) ns[i] = seq[item[i] - 1];
        if (ns < seq) return false;
    }
    return true;
}

namespace {
// lyndon.cpp:57-72  the recursive `generate` lambda.
void prenecklace_rec(int len, const std::vector<int>& alpha, Seq& seq, int index, int period,
                     const std::function<void(const Seq&, int)>& cb) {
    if (index == len) { cb(seq, period); return; }
    seq[index] = seq[index - period];
    prenecklace_rec(len, alpha, seq, index + 1, period, cb);
    auto it = std::find(alpha.begin(), alpha.end(), seq[index - period]);
    if (it != alpha.end()) ++it;
    for (; it != alpha.end(); ++it) {
        seq[index] = *it;
        prenecklace_rec(len, alpha, seq, index + 1, index + 1, cb);
    }
}

// lyndon.cpp:47-85  generate_necklaces_prefix.
void necklaces_prefix(int len, const std::vector<int>& alpha, Seq& seq, int i, int p,
                      const std::function<void(const Seq&, int)>& cb) {
    if (len <= i || alpha.empty()) return;
    seq.resize(len);
    if (i == 0 && p == 0) {
        for (int c : alpha) { seq[0] = c; prenecklace_rec(len, alpha, seq, 1, 1, cb); }
    } else {
        prenecklace_rec(len, alpha, seq, i, p, cb);
    }
}
}  // namespace

void generate_prenecklaces(int len, const std::vector<int>& alphabet,
                           const std::function<void(const Seq&, int)>& cb) {
    Seq seq;
    necklaces_prefix(len, alphabet, seq, 0, 0, cb);
}

void generate_necklaces(int len, const std::vector<int>& alphabet, int proc_id, int proc_num,
                        const std::function<void(const Seq&)>& cb) {  // lyndon.cpp:87-121
    int i = 1;
    for (; i <= len; ++i)
        if (necklace_count_sat(i, (int)alphabet.size()) >= (unsigned __int128)proc_num * 10000) break;
    std::vector<std::pair<Seq, int>> work;
    if (proc_num > 1 && i < len) {
        uint64_t count = 0;
        generate_prenecklaces(i, alphabet, [&](const Seq& seq, int period) {
            ++count;
            if ((int64_t)(count % (uint64_t)proc_num) == proc_id) work.push_back({seq, period});
        });
    } else {
        work.push_back({Seq{}, 0});
    }
    for (auto& [seq, period] : work) {
        Seq s = seq;
        necklaces_prefix(len, alphabet, s, (int)seq.size(), period, [&](const Seq& full, int per) {
            if (len % per == 0) cb(full);
        });
    }
}

// ---------------------------------------------------------------------------
// PSD

namespace {
std::atomic<uint64_t> g_tie_band{0};

struct Twiddle {
    int n = 0;
    std::vector<double> c, s;  // e^{-2 pi i m/n} = c[m] - i s[m]
    void init(int len) {
        n = len; c.assign(n, 0.0); s.assign(n, 0.0);
        // Built symmetric (m and n-m share one evaluation) so that
        // psd[k] == psd[n-k] bit-exactly.
        for (int m = 0; 2 * m <= n; ++m) {
            double a = 2.0 * M_PI * (double)m / (double)n;
            double cm = std::cos(a), sm = std::sin(a);
            if (m == 0) { cm = 1.0; sm = 0.0; }
            if (2 * m == n) { cm = -1.0; sm = 0.0; }
            c[m] = cm; s[m] = sm;
            if (m != 0 && 2 * m != n) { c[n - m] = cm; s[n - m] = -sm; }
        }
    }
};
thread_local Twiddle g_tw;

inline void note_tie(double v, double threshold) {
    if (std::fabs(v - threshold) < 1e-9) g_tie_band.fetch_add(1, std::memory_order_relaxed);
}
inline void note_round_tie(double v) {
    double f = v - std::floor(v);
    if (std::fabs(f - 0.5) < 1e-9) g_tie_band.fetch_add(1, std::memory_order_relaxed);
}
}  // namespace

uint64_t tie_band_count() { return g_tie_band.load(); }
void tie_band_reset() { g_tie_band = 0; }

void psd(const int* seq, int n, double* out) {  // fourier.cpp:25-43 (direct DFT for FFTW)
    if (g_tw.n != n) g_tw.init(n);
    for (int k = 0; k < n; ++k) {
        double re = 0.0, im = 0.0;
        int m = 0;
        for (int j = 0; j < n; ++j) {
            re += seq[j] * g_tw.c[m];
            im -= seq[j] * g_tw.s[m];
            m += k; if (m >= n) m -= n;
        }
        out[k] = re * re + im * im;
    }
}

std::vector<double> psd(const Seq& seq) {
    std::vector<double> out(seq.size());
    psd(seq.data(), (int)seq.size(), out.data());
    return out;
}

bool psd_filter(const std::vector<double>& p, int order, int paf) {  // fourier.cpp:45-52
    const double thr = 2 * order - paf + 0.001;
    for (size_t i = 1; i <= p.size() / 2; i++) {
        note_tie(p[i], thr);
        if (p[i] > thr) return false;
    }
    return true;
}

std::string seq_psd_line(const Seq& seq, const std::vector<double>& p) {  // io.cpp:14-24
    std::string out = "PSD";
    for (size_t i = 1; i < p.size(); i++) { note_round_tie(p[i]); out += std::to_string((int)std::rint(p[i])); }
    out += " ";
    for (int v : seq) { out += std::to_string(v); out += " "; }
    return out;
}

std::string seq_psd_line_invert(const Seq& seq, const std::vector<double>& p, int bound) {  // io.cpp:26-36
    std::string out = "PSD";
    for (size_t i = 1; i < p.size(); i++) { note_round_tie(p[i]); out += std::to_string(bound - (int)std::rint(p[i])); }
    out += " ";
    for (int v : seq) { out += std::to_string(v); out += " "; }
    return out;
}

std::string pair_line(const Seq& a, const Seq& b) {  // io.cpp:75-95
    std::string out;
    for (int v : a) { out += std::to_string(v); out += " "; }
    out += " ";
    for (int v : b) { out += std::to_string(v); out += " "; }
    return out;
}

// ---------------------------------------------------------------------------
// gen

int generate(int order, int compress, int paf, int proc_id, int proc_num,
             std::vector<std::string>& out_a, std::vector<std::string>& out_b,
             GenStats* stats, bool verbose) {  // generate_hybrid.cpp:107-163
    const int len = order / compress;
    const auto decomps = getdecomps(sum_constant(order, paf));
    if (decomps.empty()) {
        std::cerr << "Order " << order << " cannot be decomposed into a sum of squares\n" << order << "*2\n";
        std::cerr << "Hence there are no solutions.\n";
        return 1;
    }
    if (verbose)
        for (auto& d : decomps) std::cout << d.first << " " << d.second << "\n";
    const std::vector<int> alpha = getalphabet(compress);
    const std::set<Seq> gen_a = construct_generators(0, len);
    const std::set<Seq> gen_b = construct_generators(1, len);
    const std::string total = necklace_count_str(len, (int)alpha.size());
    if (verbose) std::cout << "Filtering " << total << " sequences\n";
    uint64_t orderly = 0, count = 0;
    const int bound = order * 2 - paf;
    generate_necklaces(len, alpha, proc_id, proc_num, [&](const Seq& seq) {
        ++orderly;
        int rs = 0;
        for (int v : seq) rs += v;
        for (const auto& d : decomps) {
            if (rs == d.first) {
                auto p = psd(seq);
                if (psd_filter(p, order, paf) && is_canonical(seq, gen_a)) {
                    ++count;
                    out_a.push_back(seq_psd_line(seq, p));
                }
            }
            if (rs == d.second) {
                auto p = psd(seq);
                if (psd_filter(p, order, paf) && is_canonical(seq, gen_b)) {
                    ++count;
                    out_b.push_back(seq_psd_line_invert(seq, p, bound));
                }
            }
        }
    });
    if (verbose) {
        std::cout << "Necklace Count: " << total << "\n";
        std::cout << "Orderly Count: " << orderly << "\n";
        std::cout << "Passed PSD filter: " << count << "\n";
    }
    if (stats) { stats->necklace_count = total; stats->orderly_count = orderly; stats->passed = count; }
    return 0;
}

// ---------------------------------------------------------------------------
// sort

std::vector<std::string> sort_uniq(std::vector<std::string> lines) {  // sort.cpp:17 under LC_ALL=C
    std::sort(lines.begin(), lines.end());  // std::string order == unsigned bytewise
    lines.erase(std::unique(lines.begin(), lines.end()), lines.end());
    return lines;
}

// ---------------------------------------------------------------------------
// match

namespace {
// An istream-like token reader over in-memory text that keeps the
// eof/fail-bit semantics the reference's loops depend on.
struct TokStream {
    const std::string& s;
    size_t pos = 0;
    bool eofbit = false, failbit = false;
    explicit TokStream(const std::string& t) : s(t) {}
    bool good() const { return !eofbit && !failbit; }
    bool peek_eof() const { return pos >= s.size(); }
    static bool ws(char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }
    void read(std::string& tok) {  // operator>>(istream&, string&)
        if (!good()) { failbit = true; return; }
        while (pos < s.size() && ws(s[pos])) ++pos;
        if (pos >= s.size()) { eofbit = failbit = true; return; }
        size_t b = pos;
        while (pos < s.size() && !ws(s[pos])) ++pos;
        if (pos >= s.size()) eofbit = true;
        tok.assign(s, b, pos - b);
    }
    void ignore_line() {  // ignore(numeric_limits<streamsize>::max(), '\n')
        if (!good()) { failbit = true; return; }
        while (pos < s.size() && s[pos] != '\n') ++pos;
        if (pos >= s.size()) { eofbit = true; return; }
        ++pos;
    }
};
}  // namespace

int paf(const Seq& seq, int s) {  // golay.cpp:27-33
    int r = 0;
    for (size_t i = 0; i < seq.size(); i++) r += seq[i] * seq[(i + s) % seq.size()];
    return r;
}

bool check_if_pair(const Seq& a, const Seq& b, int paf_constant, int compress) {  // golay.cpp:35-44
    for (size_t i = 1; i <= a.size() / 2; i++)
        if (paf(a, (int)i) + paf(b, (int)i) != paf_constant * compress) return false;
    return true;
}

int match_pairs(int order, int compress, int paf_c, const std::string& text_a,
                const std::string& text_b, std::string& pairs_out) {  // match_pairs.cpp:15-121
    const int len = order / compress;
    TokStream fa(text_a), fb(text_b);
    std::string a, b, tok_a, tok_b;
    Seq seqa(len), seqb(len);
    std::vector<Seq> match_a, match_b;
    if (fa.peek_eof()) return 1;
    if (fb.peek_eof()) return 1;
    auto read_a = [&] { fa.read(a); for (int i = 0; i < len; i++) { fa.read(tok_a); seqa[i] = std::stoi(tok_a); } };
    auto read_b = [&] { fb.read(b); for (int i = 0; i < len; i++) { fb.read(tok_b); seqb[i] = std::stoi(tok_b); } };
    read_a();
    read_b();
    while (fa.good() && fb.good()) {
        if (a == b) {
            std::string temp = a;
            while (fa.good() && a == temp) { match_a.push_back(seqa); fa.ignore_line(); read_a(); }
            temp = b;
            while (fb.good() && b == temp) { match_b.push_back(seqb); fb.ignore_line(); read_b(); }
            for (const Seq& sa : match_a)
                for (const Seq& sb : match_b)
                    if (check_if_pair(sa, sb, paf_c, compress)) { pairs_out += pair_line(sa, sb); pairs_out += "\n"; }
            match_a.clear();
            match_b.clear();
        } else if (a < b) {
            fa.ignore_line(); read_a();
        } else if (b < a) {
            fb.ignore_line(); read_b();
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// uncompression

std::vector<Seq> get_combinations(int len, const std::vector<int>& alphabet) {  // array.cpp:131-160
    std::vector<Seq> combos;
    if (len == 1) {
        for (int l : alphabet) combos.push_back(Seq{l});
        return combos;
    }
    std::vector<int> sub = alphabet;
    for (int letter : alphabet) {
        std::vector<Seq> part = get_combinations(len - 1, sub);
        for (Seq& c : part) c.insert(c.begin(), letter);
        combos.insert(combos.end(), part.begin(), part.end());
        sub.erase(std::find(sub.begin(), sub.end(), letter));
    }
    return combos;
}

std::set<Seq> generate_uncompress(const Seq& seq) {  // equivalence.cpp:26-44
    std::set<Seq> map{seq};
    size_t size = 0;
    while (map.size() != size) {
        size = map.size();
        {
            std::vector<Seq> snap(map.begin(), map.end());  // shift_equivalence :171-178
            for (Seq s : snap)
                for (size_t i = 0; i < s.size(); i++) { std::rotate(s.begin(), s.begin() + 1, s.end()); map.insert(s); }
        }
        {
            std::vector<Seq> snap(map.begin(), map.end());  // reverse_equivalence :191-197
            for (Seq s : snap) { std::reverse(s.begin(), s.end()); map.insert(s); }
        }
    }
    return map;
}

void uncompress_recursive(Seq& orig, int compress, int new_compress, int paf_c, int seqflag,
                          std::vector<std::string>& lines, uint64_t* leaves, int proc_id,
                          int proc_num) {  // uncompression.cpp:81-210
    const int order = (int)orig.size() * compress;
    const int len = order / compress;
    const int r = compress / new_compress;
    const int newlen = order / new_compress;
    const std::vector<int> alphabet = getalphabet(compress);
    const std::vector<int> newalphabet = getalphabet(new_compress);
    const std::vector<Seq> parts = get_combinations(r, newalphabet);
    std::map<int, std::vector<Seq>> partitions;
    for (int letter : alphabet) {  // :97-111
        std::vector<Seq> partition;
        for (Seq part : parts) {
            int sum = 0;
            for (int v : part) sum += v;
            if (sum == letter) {
                do { partition.push_back(part); } while (std::next_permutation(part.begin(), part.end()));
            }
        }
        partitions.insert({letter, partition});
    }
    // :113-131  best letter and the in-place rotation loop.
    std::set<int> seta(orig.begin(), orig.end());
    size_t max = 0;
    int best = orig[0];
    for (int e : seta)
        if (partitions.at(e).size() > max) { max = partitions.at(e).size(); best = e; }
    for (size_t i = 0; i < orig.size(); i++)
        if (orig[i] == best) std::rotate(orig.begin(), orig.begin() + i, orig.end());
    // :133-142  dihedral-orbit minima of the first position's splits.
    std::set<Seq> perma;
    for (const Seq& perm : partitions.at(orig[0])) perma.insert(*generate_uncompress(perm).begin());
    const std::vector<Seq> newfirsta(perma.begin(), perma.end());

    const int bound = order * 2 - paf_c;
    std::vector<double> p(newlen);
    uint64_t nleaf = 0;
    // :144-165  positions [idx, stop) filled in DFS order, position 0 outermost;
    // `at_leaf` runs on each word with positions [.., stop) filled.
    std::function<void(Seq&, size_t, size_t, const std::function<void(const Seq&)>&)> fill =
        [&](Seq& seq, size_t idx, size_t stop, const std::function<void(const Seq&)>& at_leaf) {
            if (idx >= orig.size() || idx == stop) { at_leaf(seq); return; }
            const std::vector<Seq>& opts = idx == 0 ? newfirsta : partitions.at(orig[idx]);
            for (const Seq& perm : opts) {
                for (int i = 0; i < r; i++) seq[idx + (size_t)len * i] = perm[i];
                fill(seq, idx + 1, stop, at_leaf);
            }
        };
    // :167-191  with proc_num > 1, partial words are expanded one position at a
    // time until there are at least 1000 * proc_num of them (or every position is
    // set); job proc_id keeps the partials whose index is proc_id mod proc_num.
    std::vector<Seq> prev_depth{Seq(newlen, 0)};
    size_t depth = 0;
    if (proc_num > 1) {
        long long count = 0;
        for (; depth < orig.size() && count < 1000LL * proc_num; ++depth) {
            count = 0;
            std::vector<Seq> curr_depth;
            for (Seq& partial : prev_depth)
                fill(partial, depth, depth + 1, [&](const Seq& s) { curr_depth.push_back(s); ++count; });
            prev_depth = std::move(curr_depth);
        }
    }
    std::vector<Seq> partial_list;
    for (size_t i = 0; i < prev_depth.size(); ++i)
        if ((int)i % proc_num == proc_id) partial_list.push_back(prev_depth[i]);
    // :193-207  leaf PSD filter and candidate line.
    for (Seq& partial : partial_list) {
        fill(partial, depth, orig.size(), [&](const Seq& seq) {
            ++nleaf;
            psd(seq.data(), newlen, p.data());
            if (psd_filter(p, order, paf_c)) {
                if (seqflag) lines.push_back(seq_psd_line(seq, p));
                else lines.push_back(seq_psd_line_invert(seq, p, bound));
            }
        });
    }
    if (leaves) *leaves += nleaf;
}

namespace {
std::string join_lines(const std::vector<std::string>& v) {
    std::string out;
    for (const auto& l : v) { out += l; out += "\n"; }
    return out;
}
}  // namespace

int uncompress_pipeline(int order, int compress, int new_compress, int paf_c,
                        const std::string& pairs_in, std::string& pairs_out,
                        uint64_t* leaves, std::vector<std::string>* tmp_files) {  // uncompression.cpp:32-79
    if (compress <= 1) return 0;
    const size_t length = (size_t)(order / compress);  // PairReader length, main.cpp:122
    size_t pos = 0, line_no = 0;
    while (pos < pairs_in.size()) {  // PairReader::operator>> (io.cpp:39-73)
        size_t e = pairs_in.find('\n', pos);
        bool terminated = e != std::string::npos;
        std::string line = pairs_in.substr(pos, terminated ? e - pos : std::string::npos);
        pos = terminated ? e + 1 : pairs_in.size();
        std::vector<int> vals;
        {
            size_t q = 0;
            while (q < line.size()) {
                while (q < line.size() && TokStream::ws(line[q])) ++q;
                if (q >= line.size()) break;
                size_t b = q;
                while (q < line.size() && !TokStream::ws(line[q])) ++q;
                char* endp = nullptr;
                std::string t = line.substr(b, q - b);
                long v = std::strtol(t.c_str(), &endp, 10);
                if (endp == t.c_str()) break;  // istringstream >> int stops at a non-number
                vals.push_back((int)v);
                if (*endp != '\0') break;
            }
        }
        if (vals.size() != length * 2)
            throw std::runtime_error("PairReader: input line " + std::to_string(line_no) + ": length is " +
                                     std::to_string(vals.size()) + ". Expected " + std::to_string(length * 2));
        ++line_no;
        if (!terminated) break;  // getline hit EOF: stream no longer good(), body skipped
        Seq a(vals.begin(), vals.begin() + length), b(vals.begin() + length, vals.end());
        if ((int)a.size() != order / compress || (int)b.size() != order / compress) {
            std::cerr << "Compressed pair has invalid length: " << a.size() << " " << b.size() << "\n";
            return 1;
        }
        std::vector<std::string> la, lb;
        uncompress_recursive(a, compress, new_compress, paf_c, 0, la, leaves);
        uncompress_recursive(b, compress, new_compress, paf_c, 1, lb, leaves);
        std::string ra = join_lines(la), rb = join_lines(lb);
        std::string ta = join_lines(sort_uniq(std::move(la)));
        std::string tb = join_lines(sort_uniq(std::move(lb)));
        match_pairs(order, new_compress, paf_c, ta, tb, pairs_out);
        if (tmp_files) *tmp_files = {ra, rb, ta, tb};
    }
    return 0;
}

// ---------------------------------------------------------------------------
// equivalence filter

namespace {
using PairSet = std::set<PairItem>;

PairSet shift_pair(const PairSet& m) {  // equivalence.cpp:159-169
    PairSet out;
    for (PairItem s : m) {
        int len = (int)s.a.size();
        for (int j = 0; j < len - 1; j++) { std::rotate(s.b.begin(), s.b.begin() + 1, s.b.end()); out.insert(s); }
    }
    return out;
}
PairSet swap_pair(const PairSet& m) {  // equivalence.cpp:146-156
    PairSet out;
    for (PairItem s : m) { std::swap(s.a, s.b); out.insert(s); }
    return out;
}
PairSet reverse_pair(const PairSet& m) {  // equivalence.cpp:180-189
    PairSet out;
    for (PairItem s : m) { std::reverse(s.b.begin(), s.b.end()); out.insert(s); }
    return out;
}
PairSet altnegative_pair(const PairSet& m) {  // equivalence.cpp:219-231
    PairSet out;
    for (PairItem s : m) {
        for (size_t i = 0; i < s.b.size(); i++)
            if (i % 2 == 1) { s.b[i] = -s.b[i]; s.a[i] = -s.a[i]; }
        out.insert(s);
    }
    return out;
}
Seq permute(const Seq& seq, int c) {  // equivalence.cpp:244-252
    int len = (int)seq.size();
    Seq ns(len);
    for (int i = 0; i < len; i++) ns[i] = seq[i * c % len];
    return ns;
}
PairSet decimate_pair(const PairSet& m, const std::vector<int>& coprimes) {  // equivalence.cpp:254-267
    PairSet out;
    for (const PairItem& s : m)
        for (int c : coprimes) out.insert(PairItem{permute(s.a, c), permute(s.b, c)});
    return out;
}
}  // namespace

std::set<PairItem> construct_pair_generators(int len, bool altnegative) {  // equivalence.cpp:46-95
    PairItem seq;
    seq.a.resize(len);
    seq.b.resize(len);
    for (int i = 1; i <= len; i++) { seq.a[i - 1] = i; seq.b[i - 1] = len + i; }
    const std::vector<int> coprimes = getcoprimes(len);
    std::vector<std::function<PairSet(const PairSet&)>> ops = {
        shift_pair, swap_pair, reverse_pair,
        [coprimes](const PairSet& m) { return decimate_pair(m, coprimes); }};
    if (altnegative) ops.push_back(altnegative_pair);  // filter.cpp:80-87
    // generateExhaust :46-82 -- frontier closure.
    PairSet map{seq}, newmap{seq};
    size_t size = 0;
    while (map.size() != size) {
        PairSet iter;
        size = map.size();
        for (const auto& f : ops) {
            PairSet t = f(newmap);
            iter.insert(t.begin(), t.end());
            newmap.insert(iter.begin(), iter.end());
        }
        for (const PairItem& s : map) iter.erase(s);
        map.insert(iter.begin(), iter.end());
        newmap = iter;
    }
    return map;
}

std::set<PairItem> generate_class_pairs(const std::set<PairItem>& gens, const PairItem& seq) {  // equivalence.cpp:97-144
    const int len = (int)seq.a.size();
    std::set<PairItem> equiv;
    for (const PairItem& item : gens) {
        PairItem ns;
        ns.a.resize(len);
        ns.b.resize(len);
        if ((item.a[0] < 0 && -item.a[0] > len) || (item.a[0] > 0 && item.a[0] > len)) {
            // Swapped templates: the reference writes the a-pattern into b and
            // the b-pattern into a (:108-121); kept as is.
            for (int i = 0; i < len; i++) {
                ns.b[i] = item.a[i] < 0 ? -seq.a[(-item.a[i] - 1) % len] : seq.a[(item.a[i] - 1) % len];
                ns.a[i] = item.b[i] < 0 ? -seq.b[(-item.b[i] - 1) % len] : seq.b[(item.b[i] - 1) % len];
            }
        } else {
            for (int i = 0; i < len; i++) {
                ns.a[i] = item.a[i] < 0 ? -seq.a[(-item.a[i] - 1) % len] : seq.a[(item.a[i] - 1) % len];
                ns.b[i] = item.b[i] < 0 ? -seq.b[(-item.b[i] - 1) % len] : seq.b[(item.b[i] - 1) % len];
            }
        }
        equiv.insert(ns);
    }
    return equiv;
}

int cache_filter(int order, int compress, const std::string& pairs_in, std::string& unique_out,
                 FilterStats* stats) {  // filter.cpp:18-163
    const int len = order / compress;
    std::set<PairItem> sequences;
    TokStream in(pairs_in);
    std::string a, b;
    PairItem seq;
    seq.a.resize(len);
    seq.b.resize(len);
    while (in.good()) {  // :46-63
        for (int i = 0; i < len; i++) { in.read(a); seq.a[i] = std::stoi(a); }
        if (!in.good()) break;
        for (int i = 0; i < len; i++) { in.read(b); seq.b[i] = std::stoi(b); }
        sequences.insert(seq);
    }
    const bool alt = !(compress > 1 || len % 2 == 1);  // :83-87
    const std::set<PairItem> gens = construct_pair_generators(len, alt);
    if (stats) { stats->loaded = sequences.size(); stats->group_size = gens.size(); }
    for (auto it = sequences.begin(); it != sequences.end();) {  // :117-144
        std::set<PairItem> classes = generate_class_pairs(gens, *it);
        if (classes.size() > sequences.size()) {
            for (auto iter = std::next(it, 1); iter != sequences.end();) {
                PairItem current = *iter;
                iter++;
                if (classes.find(current) != classes.end()) sequences.erase(current);
            }
            it++;
        } else {
            PairItem base = *it;
            for (const PairItem& s : classes) sequences.erase(s);
            sequences.insert(base);
            it = sequences.find(base);
            it++;
        }
    }
    for (const PairItem& s : sequences) {  // :148-160
        unique_out += pair_line(s.a, s.b);
        unique_out += "\n";
    }
    if (stats) stats->unique = sequences.size();
    return 0;
}

}  // namespace refcpu
