// refcpu -- CPU restatement of the GolayPair search pipeline.
//
// TEST INFRASTRUCTURE ONLY.  This is the oracle: a plain C++ restatement of the
// reference's algorithm, used by tests/, by __graft_entry__.smoke() and by
// bench.py's cpu_baseline leg as the checker and as the reported CPU baseline.
// Nothing in golaypair_amd/ links, loads or executes it.
//
// Every function cites the reference file:line it follows (paths relative to the
// tylerlumsden/GolayPair checkout).  The reference itself cannot be built in this
// image (FFTW3 and Boost.Multiprecision are absent and libstdc++ 11 lacks
// <format>), so the FFT is replaced by a direct FP64 DFT and the multiprecision
// counters by a small decimal bignum.
//
// Pinning (what checks each function beyond retyping it): see the header comment
// of tests/test_oracle.py; functions not pinned by an independent check say
// "parity unpinned" there.
#pragma once

#include <cstdint>
#include <functional>
#include <set>
#include <string>
#include <utility>
#include <vector>

namespace refcpu {

using Seq = std::vector<int>;

// --- small math utilities -------------------------------------------------

// golay.cpp:60-75  even c -> {0,+-2,..,+-c}; odd c -> {+-1,..,+-c}; ascending.
std::vector<int> getalphabet(int compress);
// generate_hybrid.cpp:45-47
int sum_constant(int order, int paf);
// decomps.cpp:6-17  all (x<=y), x,y>=0, x^2+y^2 == k, x-major order.
std::vector<std::pair<int, int>> getdecomps(int k);
// coprimes.cpp:12-33  {1<=i<k : gcd(i,k)==1}; aborts above 400.
std::vector<int> getcoprimes(int k);

// Arbitrary-size unsigned decimal counter used for the printed counts
// (the reference uses boost::multiprecision::cpp_int, lyndon.cpp:38-45).
std::string necklace_count_str(int length, int alphabet_size);

// --- orderly generation ----------------------------------------------------

// orderly_equivalence.cpp:17-51  closure of the identity index vector (1..LEN)
// under rotation, reversal and (flag==0) decimation by units.
std::set<Seq> construct_generators(int flag, int len);
// orderly_equivalence.cpp:53-68  0 if some image seq[item[i]-1] is lex-smaller.
bool is_canonical(const Seq& seq, const std::set<Seq>& generators);

// lyndon.cpp:47-121  FKM prenecklace recursion + round-robin prefix split.
// The callback receives every necklace of the job (LEN % period == 0).
void generate_necklaces(int len, const std::vector<int>& alphabet, int proc_id,
                        int proc_num, const std::function<void(const Seq&)>& cb);
// lyndon.cpp:47-85 (callback on every prenecklace of length `len`, with period).
void generate_prenecklaces(int len, const std::vector<int>& alphabet,
                           const std::function<void(const Seq&, int)>& cb);

// --- PSD -------------------------------------------------------------------

// fourier.cpp:25-43  psd[k] = |sum_j seq[j] e^{-2 pi i jk/n}|^2, k=0..n-1, FP64.
// Direct DFT in place of FFTW's c2c plan (FFTW's last-ulp choices are not
// reproducible; see tie_band_count()).
void psd(const int* seq, int n, double* out);
std::vector<double> psd(const Seq& seq);
// fourier.cpp:45-52  reject iff psd[k] > 2*order - paf + 0.001 for some 1<=k<=n/2.
bool psd_filter(const std::vector<double>& psd, int order, int paf);

// Number of PSD values examined so far that fell within 1e-9 of a rounding
// half-point or of the filter threshold (where FFTW and a direct DFT could
// round differently).  Must stay 0 for parity claims.
uint64_t tie_band_count();
void tie_band_reset();

// io.cpp:14-24 / io.cpp:26-36  candidate line text.
std::string seq_psd_line(const Seq& seq, const std::vector<double>& psd);
std::string seq_psd_line_invert(const Seq& seq, const std::vector<double>& psd, int bound);
// io.cpp:75-95  "a0 a1 .. " + " " + "b0 b1 .. " + "\n"
std::string pair_line(const Seq& a, const Seq& b);

// --- stages ------------------------------------------------------------------

struct GenStats {
    std::string necklace_count;  // necklace_count(LEN, |alphabet|)
    uint64_t orderly_count = 0;  // necklaces visited by this job
    uint64_t passed = 0;         // lines written (A + B)
};

// generate_hybrid.cpp:107-163.  Returns 1 when 2v+(v-1)paf has no
// decomposition into two squares, else 0.  Lines appended in FKM order.
int generate(int order, int compress, int paf, int proc_id, int proc_num,
             std::vector<std::string>& out_a, std::vector<std::string>& out_b,
             GenStats* stats, bool verbose);

// sort.cpp:7-23 run as `LC_ALL=C sort | uniq`: bytewise whole-line order,
// adjacent duplicates removed.
std::vector<std::string> sort_uniq(std::vector<std::string> lines);

// match_pairs.cpp:15-121 over the text of two sorted candidate files.
// Returns 1 if either input is empty (no output), else 0.
int match_pairs(int order, int compress, int paf, const std::string& text_a,
                const std::string& text_b, std::string& pairs_out);

// golay.cpp:27-44
int paf(const Seq& seq, int s);
bool check_if_pair(const Seq& a, const Seq& b, int paf_constant, int compress);

// array.cpp:131-160  nondecreasing `len`-tuples over `alphabet` in lex order.
std::vector<Seq> get_combinations(int len, const std::vector<int>& alphabet);
// equivalence.cpp:26-44  closure of one tuple under rotation and reversal.
std::set<Seq> generate_uncompress(const Seq& seq);

// uncompression.cpp:81-210 (uncompress_pipeline calls it with PROC_ID=0,
// PROC_NUM=1 at :65-66).  `orig` is rotated in place exactly as the reference
// does.  Each leaf of job proc_id of proc_num that passes the PSD filter
// appends one candidate line; `leaves` (if given) counts every leaf examined.
void uncompress_recursive(Seq& orig, int compress, int new_compress, int paf,
                          int seqflag, std::vector<std::string>& lines,
                          uint64_t* leaves, int proc_id = 0, int proc_num = 1);

// uncompression.cpp:32-79  pairs text at `compress` -> appended pairs text at
// `new_compress`.  Returns 1 on a length mismatch, else 0.  Throws on a
// malformed line (io.cpp:52-56).
// `tmp_files` (if given) receives the final text of the four per-pair
// temporary files (uncompress_tmp_a, _b, .sorted) as the reference leaves
// them on disk: the last pair's content, or nothing if no pair was read.
int uncompress_pipeline(int order, int compress, int new_compress, int paf,
                        const std::string& pairs_in, std::string& pairs_out,
                        uint64_t* leaves, std::vector<std::string>* tmp_files = nullptr);

// filter.cpp:18-163 (+ equivalence.cpp:46-267).  Input/output are pair-file text.
// Returns 0; `bad_file` mirrors the "Bad File." early return (:33-37).
struct FilterStats {
    uint64_t loaded = 0;
    uint64_t group_size = 0;
    uint64_t unique = 0;
};
int cache_filter(int order, int compress, const std::string& pairs_in,
                 std::string& unique_out, FilterStats* stats);

// Pair equivalence group as the reference builds it (equivalence.cpp:84-95 with
// the op list of filter.cpp:80-87): each element is (a,b) index template with
// signed 1-based entries.
struct PairItem {
    Seq a, b;
    bool operator<(const PairItem& o) const { return a == o.a ? b < o.b : a < o.a; }
    bool operator==(const PairItem& o) const { return a == o.a && b == o.b; }
};
std::set<PairItem> construct_pair_generators(int len, bool altnegative);
// equivalence.cpp:97-144  class of one pair under the generator templates.
std::set<PairItem> generate_class_pairs(const std::set<PairItem>& gens, const PairItem& seq);

}  // namespace refcpu
