/*
 * golay_gpu.h -- C ABI of the MI355X (gfx950) GolayPair search library.
 *
 * Plain C types only: pointers, sizes, int status codes.  No exceptions and no
 * C++ or torch types cross this boundary.  Every entry point replaces one seam of
 * the reference (tylerlumsden/GolayPair); the citation above each declaration
 * names it (path:line relative to that checkout).
 *
 * Ownership: buffers passed in gp_cands / gp_pairs are host memory owned by the
 * caller.  Device memory, streams and kernels are owned by the gp_ctx.  One
 * gp_ctx is used by one host thread at a time.
 *
 * Status codes: GP_OK (0) on success; GP_NO_INPUT (1) where the reference returns
 * 1 for a benign empty/undecomposable case; negative values are errors whose text
 * gp_last_error() returns.  GP_ERR_CAPACITY means an output buffer was too small:
 * its `count` field then holds the number of records required, nothing else was
 * written, and the call may be repeated with larger buffers.
 */
#ifndef GOLAY_GPU_H
#define GOLAY_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GP_OK 0
#define GP_NO_INPUT 1
#define GP_ERR_ARG (-1)
#define GP_ERR_HIP (-2)
#define GP_ERR_CAPACITY (-3)
#define GP_ERR_UNSUPPORTED (-4)

/* PSD filter arithmetic: split-BF16 MFMA contraction (FP32 accumulate, FP64
 * recheck inside the error band) or the plain FP32 direct DFT on the vector ALUs.
 * Both give identical decisions and keys. */
#define GP_PSD_SPLIT_BF16 0
#define GP_PSD_FP32 1

typedef struct gp_ctx gp_ctx;

/* Candidate records: one length-n word plus its key digits.
 * keys[i*nkey + (k-1)] = rint(PSD(k)) (plain) or bound - rint(PSD(k))
 * (inverted), k = 1..nkey, nkey = floor(n/2); the text key of the reference
 * lists k = 1..n-1, and PSD(n-k) = PSD(k). */
typedef struct gp_cands {
    int32_t n;
    int32_t nkey;
    int64_t count;     /* records held (output: written / required)        */
    int64_t capacity;  /* records the buffers can hold                     */
    int8_t* words;     /* capacity x n                                     */
    uint8_t* keys;     /* capacity x nkey                                  */
} gp_cands;

/* Pairs (a, b) of length n. */
typedef struct gp_pairs {
    int32_t n;
    int64_t count;
    int64_t capacity;
    int8_t* a;  /* capacity x n */
    int8_t* b;  /* capacity x n */
} gp_pairs;

typedef struct gp_stats {
    uint64_t words;          /* words enumerated / leaves expanded              */
    uint64_t psd_decisions;  /* PSD filter decisions made                       */
    uint64_t psd_pass;       /* decisions that accepted                         */
    uint64_t band_rechecks;  /* split-BF16 decisions settled in FP64            */
    uint64_t key_rechecks;   /* keys whose rint was settled in FP64             */
    uint64_t canonical;      /* words kept by the canonical check               */
    uint64_t pair_checks;    /* exact integer PAF checks run                    */
    uint64_t pairs;          /* pairs emitted                                   */
    uint64_t tie_band;       /* FP64 PSD within 1e-9 of a rounding or filter
                                threshold (0 for a parity-grade run)            */
    double kernel_ms;        /* device time of the call's kernels               */
    double psd_kernel_ms;    /* device time of its PSD-filter kernel launches   */
    double band_kernel_ms;   /* device time of its FP64 band-recheck launches   */
    uint64_t psd_launches;   /* PSD-filter kernel launches timed above          */
} gp_stats;

/* Context.  `device` is the HIP ordinal. */
int gp_create(gp_ctx** ctx, int device);
void gp_destroy(gp_ctx* ctx);
const char* gp_last_error(const gp_ctx* ctx);
/* Select GP_PSD_SPLIT_BF16 (default) or GP_PSD_FP32 for later calls. */
int gp_set_psd_mode(gp_ctx* ctx, int mode);

/* Tuning table of the context (no counterpart in the reference): measured launch
 * and batch parameters by key, e.g. "unc_waves/50" (filter waves per SIMD at
 * finer length 50), "unc_leaf_budget/50" (leaves per uncompression batch),
 * "unc_counter_streams/50" and "unc_counter_rounds/50" (how the filter hands out
 * its work items), "gen_ranks_per_lane/10/6" (gen ranks per lane at length 10,
 * 6 letters).  A
 * context starts with the table of the tuning file (GOLAY_TUNING, else
 * golaypair_amd/tuning.tsv, written by python -m golaypair_amd.autotune); every
 * setting gives the same results, only the time changes.  gp_set_tuning with a
 * negative value removes the key (the built-in default applies);
 * gp_get_tuning returns GP_NO_INPUT for a key the table does not hold. */
int gp_set_tuning(gp_ctx* ctx, const char* key, int64_t value);
int gp_get_tuning(const gp_ctx* ctx, const char* key, int64_t* value);

/* Candidate generation for one compression level.
 * Replaces int generate_hybrid(int ORDER, int COMPRESS, int PAF_CONSTANT,
 *   std::ofstream& out_a, std::ofstream& out_b, int PROC_ID, int PROC_NUM)
 *   (src/core/generate/generate_hybrid.h:1, body generate_hybrid.cpp:107-163).
 * out_a/out_b receive the A (plain key) and B (inverted key) candidates of job
 * proc_id of proc_num in the reference's file order (FKM order = lexicographic
 * order of the words).  Returns GP_NO_INPUT when 2v+(v-1)paf is not a sum of
 * two squares (the reference's `return 1`). */
int gp_generate(gp_ctx* ctx, int order, int compress, int paf, int proc_id, int proc_num,
                gp_cands* out_a, gp_cands* out_b, gp_stats* stats);

/* Sort + uniq of candidate records into the byte order of their text lines.
 * Replaces int GNU_sort(const std::vector<std::string>&, const std::string&,
 *   const std::string&) (src/core/sort/sort.h:1, sort.cpp:7-23) run under
 *   LC_ALL=C.  In place. */
int gp_sort_candidates(gp_ctx* ctx, gp_cands* cands);

/* Merge-join of two sorted candidate sets with an exact integer PAF check.
 * Replaces int match_pairs(int ORDER, int COMPRESS, int PAF_CONSTANT,
 *   std::ifstream&, std::ifstream&, IO::PairWriter&)
 *   (src/core/match/match_pairs.h:5, match_pairs.cpp:15-121).
 * Pairs are appended to `out` in the reference's emission order.  Returns
 * GP_NO_INPUT when either input is empty (the reference's `return 1`). */
int gp_match(gp_ctx* ctx, int order, int compress, int paf, const gp_cands* a_sorted,
             const gp_cands* b_sorted, gp_pairs* out, gp_stats* stats);

/* Expansion of one compressed sequence to the finer alphabet + PSD filter.
 * Replaces int uncompress_gpu(std::vector<int>& orig, int COMPRESS,
 *   int NEWCOMPRESS, int PAF_CONSTANT, int PROC_ID, int PROC_NUM,
 *   std::ofstream& outfile, int seqflag)
 *   (src/core/uncompression/gpu/uncompress_gpu.h:8; CPU semantics
 *   uncompression.cpp:81-210).  `orig` (length len) is rotated in place as the
 *   reference does.  Candidates of job proc_id of proc_num are written in the
 *   reference's DFS order (uncompression.cpp:167-191: with proc_num > 1 the
 *   leading positions are expanded until their partial words reach
 *   1000 * proc_num, and the job keeps the partials of index proc_id mod
 *   proc_num; proc_id outside [0, proc_num) keeps nothing); seqflag 0 ->
 *   inverted keys, 1 -> plain keys.  stats counts every leaf of the sequence
 *   (the device filters them all).  proc_num < 1 is GP_ERR_ARG. */
int gp_uncompress_side(gp_ctx* ctx, int8_t* orig, int len, int compress, int new_compress,
                       int paf, int proc_id, int proc_num, int seqflag, gp_cands* out,
                       gp_stats* stats);

/* Uncompression of every pair, re-filter and re-match, batched on the GPU.
 * Replaces int uncompress_pipeline(int order, int compress, int new_compress,
 *   int paf_constant, IO::PairReader&, IO::PairWriter&, const std::string&
 *   work_dir, const std::string& prefix, DeviceType dev)
 *   (src/core/uncompression/uncompression.h:19-29, uncompression.cpp:32-79).
 * Output pairs are appended in the reference's order (input pair order, then
 * match order).  tmp_a / tmp_b (may be NULL) receive the last pair's A and B
 * candidate sets in DFS order, the content the reference leaves in its
 * uncompress_tmp_{a,b} files (gp_sort_candidates gives the .sorted ones). */
int gp_uncompress(gp_ctx* ctx, int order, int compress, int new_compress, int paf,
                  const gp_pairs* in, gp_pairs* out, gp_cands* tmp_a, gp_cands* tmp_b,
                  gp_stats* stats);

/* Equivalence filter: keep the lexicographically smallest input pair of each
 * class, output in (a, b) lexicographic order.
 * Replaces int cache_filter(int ORDER, int COMPRESS, const std::string& PATH_IN,
 *   const std::string& PATH_OUT) (src/core/filter/filter.h:1,
 *   filter.cpp:18-163). */
int gp_filter(gp_ctx* ctx, int order, int compress, const gp_pairs* in, gp_pairs* out,
              gp_stats* stats);

/* Batched PSD filter over caller words: pass[i] = 1 iff PSD_i(k) <= 2*order_i -
 * paf_i + 0.001 for every 1 <= k <= n_i/2.  Row i holds lens[i] entries at
 * words + i*stride (host memory).  max_psd (may be NULL) receives the filter
 * kernel's own (split-BF16 or FP32) value of max_k PSD_i(k), before the FP64
 * recheck, for error measurements.
 * Replaces Fourier::calculate_psd + Fourier::psd_filter
 *   (src/utils/fourier.h:20-21, fourier.cpp:25-52), one call per batch. */
int gp_psd_filter_batch(gp_ctx* ctx, const int8_t* words, int stride, const int32_t* lens,
                        const int32_t* orders, const int32_t* pafs, int64_t count,
                        uint8_t* pass, float* max_psd, gp_stats* stats);

/* The batched PSD filter of one word length without any allocation: words and
 * pass flags are DEVICE memory (d_words: count rows of n int8 entries at
 * `stride` bytes, 8-byte aligned, stride >= 16*ceil(n/16); d_pass: count
 * bytes), pass[i] = 1 iff PSD_i(k) <= 2*order - paf + 0.001 for every
 * 1 <= k <= n/2, computed on the split-BF16 path with the FP64 recheck of the
 * error band (the same decisions as gp_psd_filter_batch).  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code that a
 * caller captures into a graph, tight serving loops, harnesses that forbid
 * hipMalloc).  The call uses no hipMalloc / hipFree; its working memory is a
 * static device array of about 18 KiB per GPU, held until the process ends
 * (word length at most 64, the size of the one PSD table it holds).  All calls on
 * one GPU share that array: calls that overlap on one GPU race.  The first call
 * for a word length copies that length's table into the array (a later call
 * with the same length does not).  The call returns when its work on the
 * context's stream is done; stats (may be NULL) receives the decision counts. */
int gp_psd_filter_batch_static(gp_ctx* ctx, const int8_t* d_words, int stride, int n, int order, int paf,
                               int64_t count, uint8_t* d_pass, gp_stats* stats);

/* FP64 key digits of caller words, the numbers printed after "PSD" by
 * write_seq_psd / write_seq_psd_invert (src/utils/io.cpp:14-36, called at
 * generate_hybrid.cpp:144,152 and uncompression.cpp:200-204):
 * keys[i*(n/2) + k-1] = rint(PSD_i(k)), or bound - rint(PSD_i(k)) when
 * `inverted`, k = 1..n/2.  Row i holds n entries at words + i*stride.  Returns
 * GP_ERR_ARG if a digit falls outside [0, 255]. */
int gp_psd_keys(gp_ctx* ctx, const int8_t* words, int stride, int n, int64_t count, int bound, int inverted,
                uint8_t* keys, gp_stats* stats);

/* Batched PSD microbenchmark on device-generated words: `count` candidates,
 * with mix[j] = {n, c, order, paf, share} rows (share = number of candidates of
 * that kind).  Word i is generated from a splitmix64 counter (seed, i) and its
 * kind from a seeded interleave; nothing is read from the host.  Returns the
 * number of accepted candidates per kind in accepted[j] and a digest of the
 * pass set in *digest.  If pass_out is not NULL the first pass_cap pass flags
 * are copied back.  Same generator as gp_microbench_words() below.
 * No reference counterpart: it drives the same PSD filter as
 *   gp_psd_filter_batch without host input. */
int gp_psd_microbench(gp_ctx* ctx, uint64_t seed, int64_t count, const int32_t* mix, int nmix,
                      uint64_t* accepted, uint64_t* digest, uint8_t* pass_out, int64_t pass_cap,
                      gp_stats* stats);
/* Host copy of the microbenchmark's word generator: for candidates
 * i = first..first+rows-1 writes the kind index to kinds[i-first] and the
 * entries (zero padded) to out + (i-first)*stride. */
int gp_microbench_words(uint64_t seed, int64_t count, const int32_t* mix, int nmix, int64_t first,
                        int64_t rows, int32_t* kinds, int8_t* out, int stride);

#ifdef __cplusplus
}
#endif

#endif /* GOLAY_GPU_H */
