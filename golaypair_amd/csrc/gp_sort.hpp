// Device radix sort and the candidate join shared by match and uncompress.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

struct gp_ctx;
struct gp_stats;

namespace gp {

// Stable LSD radix sort of (key, value) pairs on bits [begin_bit, end_bit).
void radix_sort_pairs(hipStream_t st, uint64_t* keys, uint32_t* vals, uint64_t* tmp_keys, uint32_t* tmp_vals,
                      uint32_t* hist, int64_t n, int begin_bit, int end_bit);
// Bytes of the `hist` buffer for n elements (histograms and scan scratch).
size_t radix_hist_bytes(int64_t n);
// The same with one 8-bit pass per entry of shifts[0..nshift), least significant first.
void radix_sort_pairs_shifts(hipStream_t st, uint64_t* keys, uint32_t* vals, uint64_t* tmp_keys, uint32_t* tmp_vals,
                             uint32_t* hist, int64_t n, const int* shifts, int nshift);
// In-place exclusive scan of m uint32 values; `sums` is scratch of
// scan_sums_count(m) values.
void scan_exclusive_u32(hipStream_t st, uint32_t* v, int64_t m, uint32_t* sums);
int64_t scan_sums_count(int64_t m);

// Sort + uniq of m candidate records (int8 words of n, uint8 keys of K = n/2)
// held on the device, in the byte order of their text lines; the kept records
// go to out_words / out_keys (not aliasing the inputs).  Returns their number.
int64_t sort_uniq_cands(gp_ctx* ctx, const int8_t* d_words, const uint8_t* d_keys, int64_t m, int n, int K,
                        int8_t* d_out_words, uint8_t* d_out_keys);

// Join of two candidate sets on equal key tuples (nkey bytes per record,
// row stride kstride) followed by the exact integer PAF check
//   PAF_a(s) + PAF_b(s) == paf_target  for s = 1..n/2   (golay.cpp:27-44).
// Words are int8 rows of stride wstride.  Returns the accepted (ia, ib)
// index pairs sorted lexicographically.  All pointers are device memory.
struct JoinResult {
    std::vector<uint32_t> ia, ib;
    uint64_t checks = 0;
};
JoinResult join_pairs(gp_ctx* ctx, const int8_t* d_wa, const uint8_t* d_ka, int64_t na, const int8_t* d_wb,
                      const uint8_t* d_kb, int64_t nb, int n, int nkey, int wstride, int kstride, int paf_target,
                      gp_stats* stats);

}  // namespace gp
