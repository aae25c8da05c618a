// Uncompression internals shared by uncompress.hip and the C ABI.
#pragma once

#include <stdint.h>

#include <vector>

#include "golay_gpu.h"

struct gp_ctx;

namespace gp {

// One compressed sequence prepared for expansion (uncompression.cpp:85-142).
struct SideHost {
    int L = 0, r = 0, N = 0, seqflag = 0;
    std::vector<int> orig;                       // after the reference's rotation loop
    std::vector<std::vector<std::vector<int>>> T;  // T[j][d] = r-tuple of position j
    uint64_t leaves = 0;                         // product of |T[j]|
};
SideHost build_side(const int8_t* orig, int L, int compress, int new_compress, int seqflag);

// Survivors of a batch of sides, grouped per side (side_begin) in DFS order.
struct UncResult {
    std::vector<int64_t> side_begin;  // size nsides + 1
    std::vector<uint64_t> leaf;       // leaf index of every survivor
    int8_t* d_words = nullptr;        // device: rows of N
    uint8_t* d_keys = nullptr;        // device: rows of K
    int N = 0, K = 0;
    double filter_ms = 0;
};
UncResult expand_sides(gp_ctx* ctx, const std::vector<SideHost>& sides, int order, int paf, gp_stats* stats);

}  // namespace gp
