// Host-side helpers of the library (small combinatorics, record sets).
#pragma once

#include <stdint.h>

#include <string>
#include <utility>
#include <vector>

#include "golay_gpu.h"

namespace gp {

// golay.cpp:60-75  even c -> {0,+-2,..,+-c}; odd c -> {+-1,..,+-c}; ascending.
std::vector<int> getalphabet(int compress);
// decomps.cpp:6-17  (x <= y), x^2 + y^2 = k, x-major order.
std::vector<std::pair<int, int>> getdecomps(int k);
// lyndon.cpp:87-106: prefix length of the round-robin job split, or 0 when
// every job enumerates everything (proc_num == 1 or no prefix length < n
// reaches proc_num * 10^4 necklaces).
int shard_prefix_len(int n, int k, int proc_num);
// Digit-index values (base k, first digit most significant) of the
// prenecklaces of length `len` whose 1-based FKM count is = proc_id mod proc_num.
std::vector<uint64_t> shard_prefixes(int len, const std::vector<int>& alphabet, int proc_id, int proc_num);

// Host copy of a candidate set: words (count x n) and key digits (count x nkey).
struct CandSet {
    int n = 0, nkey = 0;
    std::vector<int8_t> words;
    std::vector<uint8_t> keys;
    int64_t size() const { return n ? (int64_t)words.size() / n : 0; }
};
struct GenOutput {
    CandSet a, b;
};
struct PairSet {
    int n = 0;
    std::vector<int8_t> a, b;
    int64_t size() const { return n ? (int64_t)a.size() / n : 0; }
};

// Copies a set into caller buffers (GP_ERR_CAPACITY when too small).
int export_cands(const CandSet& s, gp_cands* out);
int export_pairs(const PairSet& s, gp_pairs* out, bool append);
CandSet import_cands(const gp_cands* in);
PairSet import_pairs(const gp_pairs* in);

}  // namespace gp

struct gp_ctx;
namespace gp {
int generate_level(gp_ctx* ctx, int order, int compress, int paf, int proc_id, int proc_num, GenOutput& out,
                   gp_stats* stats);
}
