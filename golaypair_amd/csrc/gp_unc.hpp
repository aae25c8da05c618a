// Uncompression internals shared by the uncompression translation unit and the C ABI.
#pragma once

#include <stdint.h>

#include <functional>
#include <map>
#include <memory>
#include <span>
#include <vector>

#include "golay_gpu.h"
#include "gp_common.hpp"
#include "gp_host.hpp"

struct gp_ctx;

namespace gp {

// Split lists of one compression step c -> c' (uncompression.cpp:91-111,133-142),
// per letter of the coarse alphabet: every r-tuple over the fine alphabet summing
// to the letter (all permutations of each multiset, lexicographic order), and the
// sorted dihedral-orbit minima of those tuples (used at position 0).
struct SplitTable {
    int r = 0;
    std::map<int, int> letter_id;
    int16_t id_of[256];  // letter + 128 -> letter id, -1 if not in the alphabet
    std::vector<int> nsplit, nfirst;
    std::vector<std::vector<int8_t>> split, first;  // [letter id][tuple * r + i]
};
std::shared_ptr<const SplitTable> make_splits(int compress, int new_compress);

// One compressed sequence prepared for expansion (uncompression.cpp:85-142).
// Plain data (no heap storage), so that many sides live in one reusable buffer;
// positions j < L are valid.
struct SideHost {
    int L, r, N, seqflag;
    int8_t orig[kMaxN];          // after the reference's rotation loop
    int32_t R[kMaxN];            // |T_j|: splits of position j
    const int8_t* tp[kMaxN];     // T_j: R[j] r-tuples (into tab)
    const SplitTable* tab;       // outlives the side (held by the caller)
    uint64_t leaves;             // product of R[j]
    const int8_t* tuple(int j, int d) const { return tp[j] + (size_t)d * r; }
};
SideHost build_side(const SplitTable& tab, const int8_t* orig, int L, int seqflag);

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
    bool all_pm1 = false;            // every side's tuple entries are +-1
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

// FP64 twiddles of length N (the context's PSD table), for unc_prepare.
struct UncTrig {
    int N = 0;
    int64_t trig_off = 0;
    std::vector<double> cs, sn;
};
UncTrig unc_trig(gp_ctx* ctx, int N);

// Host tables of a batch of sides.  unc_prepare touches no device state and may
// run on another host thread while the device works on the previous batch;
// unc_launch uploads the tables and runs the filter (unc_filter = both).
// Every table is written in place into one staging buffer: a pinned buffer
// from `pool` when given, else host memory of the batch (then unc_stage copies
// it into a pinned buffer before the upload).  `before_fill`, when given, runs
// once the layout is known and before the tables are filled on the host
// threads (a caller waits there for an earlier batch's fill).
struct UncPrep;
struct PinnedPool;
std::shared_ptr<UncPrep> unc_prepare(std::span<const SideHost> sides, std::span<const int> group, int order, int paf,
                                     bool fp32, const UncTrig& trig, PinnedPool* pool = nullptr,
                                     const std::function<void()>& before_fill = {});
UncBatch unc_launch(gp_ctx* ctx, UncPrep& prep, gp_stats* stats);
void unc_stage(UncPrep& prep, PinnedPool& pool);

// group[i]: pair index of input side i (a-side seqflag 0, b-side seqflag 1).
UncBatch unc_filter(gp_ctx* ctx, std::span<const SideHost> sides, std::span<const int> group, int order, int paf,
                    gp_stats* stats);
UncEmitted unc_join(gp_ctx* ctx, const UncBatch& B, int paf_target, gp_stats* stats);

// Asynchronous form of unc_launch + unc_join: unc_enqueue queues the upload,
// filter, band recheck and join of a batch on the context's stream, in the
// device buffers of `slot` (0 or 1), with every list length read on the
// device; unc_collect waits for that batch alone and reads back its emitted
// pairs.  Batch i + 1 may be enqueued (in the other slot) before batch i is
// collected; a slot is reused only after its batch is collected.  A batch
// whose lists overflow is run again by unc_collect on the synchronous path.
struct UncInflight;
std::shared_ptr<UncInflight> unc_enqueue(gp_ctx* ctx, std::shared_ptr<UncPrep> prep, int slot, int paf_target);
UncBatch unc_collect(gp_ctx* ctx, UncInflight& F, int paf_target, UncEmitted& E, gp_stats* stats);
// Survivors of one input side in DFS order with FP64 keys.
// Candidates of one input side in DFS order.  With proc_num > 1 only the
// leaves of job proc_id are kept: those whose index divided by `job_suffix`
// (the leaf count below the job-split depth) is proc_id mod proc_num.
CandSet unc_side_cands(gp_ctx* ctx, const UncBatch& B, int input_side, uint64_t job_suffix = 1, int proc_id = 0,
                       int proc_num = 1);

}  // namespace gp
