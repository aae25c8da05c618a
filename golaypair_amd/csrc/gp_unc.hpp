// Uncompression internals shared by the uncompression translation unit and the C ABI.
#pragma once

#include <stdint.h>

#include <vector>

#include "golay_gpu.h"
#include "gp_host.hpp"

struct gp_ctx;

namespace gp {

// One compressed sequence prepared for expansion (uncompression.cpp:85-142).
struct SideHost {
    int L = 0, r = 0, N = 0, seqflag = 0;
    std::vector<int> orig;                         // after the reference's rotation loop
    std::vector<std::vector<std::vector<int>>> T;  // T[j][d] = r-tuple of position j
    uint64_t leaves = 0;                           // product of |T[j]|
};
SideHost build_side(const int8_t* orig, int L, int compress, int new_compress, int seqflag);

struct UncSide;

// Device-resident result of the filter stage for a batch of sides.
struct UncBatch {
    int N = 0, K = 0, bound = 0;
    int64_t trig_off = 0;
    int nsides_table = 0;
    std::vector<int> side_of;        // table side -> input side
    const UncSide* d_sides = nullptr;
    const int32_t* d_radix = nullptr;
    const int8_t* d_tup = nullptr;
    const uint64_t* d_pm1 = nullptr;
    const double* d_c64 = nullptr;
    const int32_t* d_pos64 = nullptr;
    uint64_t* d_surv = nullptr;      // survivors: (table side << 40) | leaf
    int64_t nsurv = 0;
};

// Emitted pairs of a batch: rows [0, count) are the A words/keys, rows
// [count, 2 count) the B ones; group = input pair index inside the batch.
struct UncEmitted {
    int N = 0, K = 0;
    int64_t count = 0;
    std::vector<int8_t> words;
    std::vector<uint8_t> keys;
    std::vector<int32_t> group;
};

// group[i]: pair index of input side i (a-side seqflag 0, b-side seqflag 1).
UncBatch unc_filter(gp_ctx* ctx, const std::vector<SideHost>& sides, const std::vector<int>& group, int order,
                    int paf, gp_stats* stats);
UncEmitted unc_join(gp_ctx* ctx, const UncBatch& B, int paf_target, gp_stats* stats);
// Survivors of one input side in DFS order with FP64 keys.
CandSet unc_side_cands(gp_ctx* ctx, const UncBatch& B, int input_side);

}  // namespace gp
