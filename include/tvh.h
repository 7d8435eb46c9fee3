/*
 * tvh.h - C ABI of the MI355X (gfx950) table-commitment library.
 *
 * Drop-in for the reference prover's table-commitment hot path: Goldilocks
 * coset low-degree extension, Tip5 row hashing and Tip5 Merkle trees.  Every
 * entry point is `extern "C"`, takes plain pointers and sizes, returns an int
 * status and never throws.  Field elements are canonical u64 in [0, p),
 * p = 2^64 - 2^32 + 1 (the reference's BFieldElement,
 * include/types/b_field_element.hpp:31-41); `_dev` entry points require
 * canonical inputs and produce canonical outputs.
 *
 * Layouts (the reference GPU layouts):
 *   tables   column-major, element (row r, column c) at c*num_rows + r
 *            (include/gpu/kernels/ntt_kernel.cuh:79-84)
 *   XFE      limb-major columns, limb k of XFE column c at (3c + k)*num_rows + r
 *            (src/gpu/kernels/row_hash_kernel.cu:442-473)
 *   digests  5 contiguous u64 per row
 *   trees    TVH_TREE_LEVEL_ORDER: leaves [0,n), then each level, root at node
 *            2n-2 (src/gpu/kernels/merkle_kernel.cu:60-102);
 *            TVH_TREE_HEAP: node i at i (1-indexed), leaves [n,2n), root at node 1
 *            (src/merkle/merkle_tree.cpp:12-33).  Node j occupies u64 [5j, 5j+5).
 *
 * `_dev` calls take device pointers and a hipStream_t passed as void*
 * (NULL = the default stream); they are asynchronous on that stream, never
 * synchronise it and allocate no device memory, except where stated:
 *  - tvh_ntt_*_batched_dev, tvh_lde_xfe_batch_dev and tvh_merkle_root_dev use
 *    a per-device scratch arena that grows (waiting for the device, then
 *    reallocating) on first use of a larger size; calls that use it must be
 *    ordered on one stream;
 *  - every LDE (tvh_lde_*_dev, tvh_commit_dev, tvh_commit_frugal_dev) reads
 *    per-device coset tables for its (trace_len, target_len, trace_offset,
 *    target_offset); the first call with a new key builds them (one small
 *    allocation and a wait on a private stream).  tvh_lde_prepare builds them
 *    ahead of time, so that later calls allocate nothing and never wait.
 * `_host` calls take host pointers, are synchronous and use device buffers of
 * their own (calls from several threads at once share nothing on the device;
 * their large uploads take turns on the page-locked staging ring of
 * tvh_upload_dev).
 */
#ifndef TVH_H
#define TVH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define TVH_OK 0
#define TVH_ERR_INVALID_ARG 1   /* null pointer, zero/oversized dimension      */
#define TVH_ERR_NOT_POW2 2      /* length must be a power of two               */
#define TVH_ERR_OOM 3           /* device allocation failed                    */
#define TVH_ERR_DEVICE 4        /* HIP runtime or kernel launch error          */
#define TVH_ERR_LAYOUT 5        /* unknown tree layout                         */
#define TVH_ERR_NOT_CANONICAL 6 /* host input value >= p                       */
#define TVH_ERR_UNSUPPORTED 7   /* shape outside the supported range           */
#define TVH_ERR_RANGE 8         /* index out of range                          */

#define TVH_TREE_LEVEL_ORDER 0
#define TVH_TREE_HEAP 1

#define TVH_P 0xFFFFFFFF00000001ull

/* ---- library ------------------------------------------------------------ */

/* Set the current HIP device and build its constant tables (twiddles).
 * Thread-safe; idempotent.  Replaces the implicit ntt_init_constants()
 * (include/gpu/kernels/ntt_kernel.cuh:19-24). */
int tvh_init(int device);

/* Human-readable name of a status code (static string). */
const char* tvh_status_string(int status);

/* Message of the last failing call on this thread (static per-thread buffer). */
const char* tvh_last_error(void);

/* ---- NTT (src/gpu/kernels/ntt_kernel.cu; include/gpu/kernels/ntt_kernel.cuh) */

/* Forward NTT of num_cols columns of length n (power of two, n <= 2^24),
 * column-major, in place, natural order in and out:
 * out[j] = sum_i in[i] * w_n^(i*j), w_n = 7^((p-1)/n).
 * Replaces ntt_forward_batched_gpu (ntt_kernel.cuh:88-93). */
int tvh_ntt_forward_batched_dev(uint64_t* d_data, size_t n, size_t num_cols, void* stream);

/* Inverse NTT including the factor n^-1.
 * Replaces ntt_inverse_batched_gpu (ntt_kernel.cuh:102-107). */
int tvh_ntt_inverse_batched_dev(uint64_t* d_data, size_t n, size_t num_cols, void* stream);

/* tvh_ntt_forward_batched_dev (inverse == 0) or tvh_ntt_inverse_batched_dev
 * (inverse != 0) without any allocation: the step between the two transform
 * halves goes through a static 32 MiB device scratch per GPU (held until the
 * process ends) instead of the growable arena, so n <= 2^22
 * (TVH_ERR_UNSUPPORTED beyond).  tvh_init must have been called for the
 * current device (TVH_ERR_DEVICE otherwise).  Calls that overlap on one GPU
 * (two streams, two threads) race on that scratch; order them on one stream.
 * No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). */
int tvh_ntt_batched_noalloc_dev(uint64_t* d_data, size_t n, size_t num_cols, int inverse,
                                void* stream);

/* Single-column forms. Replace ntt_forward_gpu / ntt_inverse_gpu
 * (ntt_kernel.cuh:37-58). */
int tvh_ntt_forward_dev(uint64_t* d_data, size_t n, void* stream);
int tvh_ntt_inverse_dev(uint64_t* d_data, size_t n, void* stream);

/* ---- low-degree extension (src/gpu/kernels/randomized_lde_kernel.cu) ------ */

/* Randomized coset LDE of num_cols B-field columns:
 *   poly_c = INTT(trace_c) scaled by trace_offset^-i, plus
 *            (x^n - trace_offset^n) * r_c(x), r_c = d_rand[c*rand_len + i]
 *   out_c[j] = poly_c(target_offset * w_m^j), j < target_len
 * trace_len and target_len are powers of two, target_len = E * trace_len
 * with E in {1,2,4,8,16}; rand_len <= trace_len; d_rand may be NULL when
 * rand_len == 0.  d_trace is column-major trace_len x num_cols, d_out is
 * column-major target_len x num_cols; they must not overlap.  d_out is the
 * only working memory (no scratch).
 * Replaces randomized_lde_batch_gpu (randomized_lde_kernel.cuh:62-95) and
 * randomized_lde_batch_gpu_preallocated (:103-117). */
int tvh_lde_batch_dev(const uint64_t* d_trace, size_t num_cols, size_t trace_len,
                      const uint64_t* d_rand, size_t rand_len, uint64_t trace_offset,
                      uint64_t target_offset, size_t target_len, uint64_t* d_out,
                      void* stream);

/* One column.  Replaces randomized_lde_column_gpu (randomized_lde_kernel.cuh:34-60). */
int tvh_lde_column_dev(const uint64_t* d_trace, size_t trace_len, const uint64_t* d_rand,
                       size_t rand_len, uint64_t trace_offset, uint64_t target_offset,
                       size_t target_len, uint64_t* d_out, void* stream);

/* XFE columns in the interleaved layout of randomized_xfe_lde_batch_gpu
 * (randomized_lde_kernel.cuh:184-199): limb k of (col, row) at
 * (col*trace_len + row)*3 + k, output likewise with target_len.  Randomizers
 * are BFE (lifted to limb 0) or, with use_xfe_randomizers, 3 limbs each at
 * (col*rand_len + i)*3 + k.  Evaluation truncates coefficients beyond
 * target_len (src/lde/lde_randomized.cpp:270-275).  Uses d_xfe_output's
 * memory plus a per-device arena for the limb-major staging copy. */
int tvh_lde_xfe_batch_dev(const uint64_t* d_xfe_trace, size_t num_cols, size_t trace_len,
                          const uint64_t* d_rand, size_t rand_len, uint64_t trace_offset,
                          uint64_t target_offset, size_t target_len, uint64_t* d_xfe_out,
                          int use_xfe_randomizers, void* stream);

/* Limb-major XFE columns (the aux-table commit layout, (3c + k)*n + r): the
 * same LDE on 3*num_xfe_cols B-field columns, with the XFE truncation rule and
 * randomizer lifting of tvh_lde_xfe_batch_dev.  No scratch. */
int tvh_lde_xfe_limbs_dev(const uint64_t* d_limbs, size_t num_xfe_cols, size_t trace_len,
                          const uint64_t* d_rand, size_t rand_len, uint64_t trace_offset,
                          uint64_t target_offset, size_t target_len, uint64_t* d_out,
                          int use_xfe_randomizers, void* stream);

/* Builds the per-device coset tables every LDE of this shape and these
 * offsets reads (at most E * (N1 + N2) u64 each, N1 * N2 = trace_len), so
 * that later LDE calls with the same key allocate nothing and never wait.
 * Synchronous; idempotent.  At most 256 keys are kept (TVH_ERR_RANGE beyond;
 * tvh_lde_release_constants frees them all).  No reference counterpart (the
 * reference recomputes its coset powers per call, randomized_lde_kernel.cu:990-1050). */
int tvh_lde_prepare(size_t trace_len, size_t target_len, uint64_t trace_offset,
                    uint64_t target_offset);
/* Waits for the device, then frees every cached coset table. */
int tvh_lde_release_constants(void);

/* ---- LDE kernel forms and autotuning -------------------------------------- */

/* Kernel variants of the LDE, all with identical results.  A variant code is
 * form | (tile << 2).  The form is the kernel shape of the fused row step
 * (inverse row DFT, coset scaling, forward row DFT): TVH_LDE_WIDE runs it as
 * one kernel with all cosets of a row in one 1024-thread LDS tile;
 * TVH_LDE_SPLIT runs the inverse row DFTs as a kernel of their own (one more
 * read and write of the trace-sized data, every thread busy) before the
 * expansion.  Tile 0/1/2 gives the column steps LDS tiles of 2^14/2^13/2^12
 * elements (1024/512/256 threads).  TVH_LDE_ROW_HALF (with TVH_LDE_SPLIT only)
 * gives the two row kernels 2^13-element tiles (512 threads, two blocks per CU)
 * instead of 2^14, with the row length halved where E * row length would not
 * fit; TVH_LDE_INV_HALF (with TVH_LDE_SPLIT only, not with TVH_LDE_ROW_HALF)
 * gives only the inverse row kernel the 2^13 tile; TVH_LDE_ROW_QUARTER (with
 * TVH_LDE_SPLIT only, alone) gives both row kernels 2^12-element tiles (256
 * threads, four blocks per CU).  The variant is chosen per
 * (log2 trace_len, log2 expansion, width band), the width bands being
 * tables of fewer than 64 columns, 64..255 and 256 or more;
 * tvh_lde_set_variant sets every width band, tvh_lde_autotune the band of
 * its num_cols.  TVH_LDE_AUTO restores the default
 * (TVH_LDE_SPLIT with column tile 2 and TVH_LDE_ROW_QUARTER for 2^13..2^16
 * rows, TVH_LDE_ROW_HALF for 2^17..2^18, column tile 1 and TVH_LDE_ROW_HALF
 * from 2^19; TVH_LDE_SPLIT | TVH_LDE_INV_HALF with tile 1 at 2^12;
 * TVH_LDE_WIDE with tile 2 below).  Batch plans use each table's form with
 * column tile 1 and 2^13-element row tiles.
 * No reference counterpart (the reference has one fixed kernel schedule per
 * size, src/gpu/kernels/ntt_kernel.cu). */
enum {
    TVH_LDE_AUTO = 0,
    TVH_LDE_WIDE = 1,
    TVH_LDE_SPLIT = 2,
    TVH_LDE_ROW_HALF = 16,
    TVH_LDE_INV_HALF = 32,
    TVH_LDE_ROW_QUARTER = 64
};
int tvh_lde_set_variant(uint32_t log_n, uint32_t log_e, int variant);
/* The variant used for tables of 256 or more columns. */
int tvh_lde_get_variant(uint32_t log_n, uint32_t log_e, int* variant);
/* The variant used for tables of num_cols columns. */
int tvh_lde_get_variant_for_width(uint32_t log_n, uint32_t log_e, size_t num_cols, int* variant);
/* Number of variants tvh_lde_autotune times (the length of its ms array). */
int tvh_lde_variant_count(void);
/* Times every variant on scratch buffers of this shape (num_cols capped so
 * the codewords stay under 4 GiB), keeps the fastest for the shape and the
 * width band of num_cols, and returns
 * its code in *chosen; ms (may be NULL) receives the average milliseconds per
 * LDE of the variants 1, 2, 5, 6, 9, 10, 18, 22, 26, 34, 38, 42, 66, 70, 74 in
 * that order.  Synchronous; allocates
 * and frees its own buffers; not to be called concurrently with other work. */
int tvh_lde_autotune(size_t num_cols, size_t trace_len, size_t target_len, int* chosen,
                     float* ms);

/* ---- two-phase coset evaluation (the frugal commit's primitives) ---------- */

/* Phase 1: d_coefficients[c*n + i] = INTT_n(trace_c)_i * trace_offset^-i, the
 * trace's interpolant in coefficient form.  d_coefficients may equal d_trace
 * (in place) or must not overlap it.  Uses the per-device arena for the NTT's
 * transpose step.
 * Replaces compute_trace_coefficients_gpu (randomized_lde_kernel.cuh:127-134). */
int tvh_trace_coefficients_dev(const uint64_t* d_trace, size_t num_cols, size_t trace_len,
                               uint64_t trace_offset, uint64_t* d_coefficients, void* stream);

/* Phase 2: d_output[c*n + j] = P_c(coset_offset * w_n^j), j < n = trace_len,
 * with P_c = sum_i coeff_i X^i + (X^n - trace_offset^n) * r_c(X), r_c =
 * d_rand[c*rand_len + i] (rand_len <= trace_len; d_rand may be NULL when 0).
 * Called once per coset g_k = fri_offset * w_M^k of an M-point domain, row j
 * of coset k is row j*(M/n) + k of the M-point codeword.  d_tail_scratch is
 * accepted for the reference's signature and not used (may be NULL).  Uses the
 * per-device arena for the NTT's transpose step and for the powers of
 * coset_offset that its column step applies to the coefficients as it loads
 * them.  d_output may equal d_coefficients (in place, which consumes the
 * coefficients) or must not overlap it.
 * Replaces evaluate_coset_from_coefficients_gpu (randomized_lde_kernel.cuh:152-163). */
int tvh_evaluate_coset_dev(const uint64_t* d_coefficients, size_t num_cols, size_t trace_len,
                           const uint64_t* d_rand, size_t rand_len, uint64_t trace_offset,
                           uint64_t coset_offset, uint64_t* d_output, uint64_t* d_tail_scratch,
                           void* stream);

/* d_leaves[(j*num_cosets + coset_idx)*5 + e] = d_coset_digests[j*5 + e],
 * j < coset_len: a coset's row digests into the full domain's leaf order.
 * Replaces qzc_scatter_digests_strided_kernel (gpu_stark.cu:993-1007). */
int tvh_scatter_digests_strided_dev(const uint64_t* d_coset_digests, uint64_t* d_leaves,
                                    size_t coset_len, size_t coset_idx, size_t num_cosets,
                                    void* stream);

/* ---- layout transforms ------------------------------------------------------ */

/* Row-major n x width (element (r, c) at r*width + c) to column-major (at
 * c*n + r), and back; not in place.  The XFE forms take n x xwidth XFE
 * elements, limbs interleaved row-major ((r*xwidth + c)*3 + k) to limb-major
 * columns ((3c + k)*n + r), and back.  Replace qzc_rowmajor_to_colmajor_{bfe,xfe}
 * and qzc_colmajor_to_rowmajor_{bfe,xfe} (gpu_stark.cu:125-214). */
int tvh_rowmajor_to_colmajor_bfe_dev(const uint64_t* d_rowmajor, uint64_t* d_colmajor, size_t n,
                                     size_t width, void* stream);
int tvh_rowmajor_to_colmajor_xfe_dev(const uint64_t* d_rowmajor, uint64_t* d_colmajor, size_t n,
                                     size_t xwidth, void* stream);
int tvh_colmajor_to_rowmajor_bfe_dev(const uint64_t* d_colmajor, uint64_t* d_rowmajor, size_t n,
                                     size_t width, void* stream);
int tvh_colmajor_to_rowmajor_xfe_dev(const uint64_t* d_colmajor, uint64_t* d_rowmajor, size_t n,
                                     size_t xwidth, void* stream);

/* ---- Tip5 (src/gpu/kernels/row_hash_kernel.cu, tip5_kernel.cu) ------------- */

/* digests[r] = Tip5::hash_varlen(row r) over a column-major table.
 * Replaces hash_bfield_rows_gpu (row_hash_kernel.cuh:27-33). */
int tvh_hash_bfield_rows_dev(const uint64_t* d_table, size_t num_rows, size_t num_cols,
                             uint64_t* d_digests, void* stream);

/* XFE rows in limb-major layout, flattened (c0,c1,c2) per XFE column.
 * Replaces hash_xfield_rows_gpu (row_hash_kernel.cuh:44-50). */
int tvh_hash_xfield_rows_dev(const uint64_t* d_table, size_t num_rows, size_t num_xfe_cols,
                             uint64_t* d_digests, void* stream);

/* In-place Tip5 permutation of num_states 16-element states.
 * Replaces Backend::tip5_permutation_batch (include/backend/backend.hpp:95-101). */
int tvh_tip5_permutation_batch_dev(uint64_t* d_states, size_t num_states, void* stream);

/* Rounds [first_round, first_round + num_rounds) of the permutation on
 * num_states 16-element states, in place (first_round + num_rounds <= 5,
 * TVH_ERR_RANGE otherwise).  d_trace (may be NULL) receives, per state, the
 * input state and the state after each of those rounds:
 * d_trace[(i * (num_rounds + 1) + j) * 16 + k].  Replaces Tip5::round and
 * Tip5::trace (src/hash/tip5.cpp:98-125). */
int tvh_tip5_rounds_dev(uint64_t* d_states, size_t num_states, uint32_t first_round,
                        uint32_t num_rounds, uint64_t* d_trace, void* stream);

/* out[i] = hash_pair(left[i], right[i]).  Replaces Backend::hash_pairs
 * (include/backend/backend.hpp:103-115) and hash_pairs_gpu (tip5_kernel.cu:210-248). */
int tvh_hash_pairs_dev(const uint64_t* d_left, const uint64_t* d_right, uint64_t* d_out,
                       size_t count, void* stream);

/* ---- Merkle (src/gpu/kernels/merkle_kernel.cu; src/merkle/merkle_tree.cpp) -- */

/* Full tree over num_leaves (power of two) digests in the given layout.
 * d_tree holds (2n-1)*5 u64 (level order) or 2n*5 u64 (heap, node 0 unused).
 * d_leaves may alias the tree's leaf level (in-place build).
 * Replaces merkle_tree_gpu (merkle_kernel.cuh:26-38). */
int tvh_merkle_tree_dev(const uint64_t* d_leaves, uint64_t* d_tree, size_t num_leaves,
                        int layout, void* stream);

/* Root only (5 u64).  Replaces merkle_root_gpu (merkle_kernel.cuh:40-52). */
int tvh_merkle_root_dev(const uint64_t* d_leaves, uint64_t* d_root, size_t num_leaves,
                        void* stream);

/* d_out[q] = tree node with heap index d_heap_idx[q] (1 <= idx < 2n), for a
 * tree in either layout: the digest gather behind authentication paths and
 * structures (gpu_stark.cu:1266-1286; gather_kernel.cuh:30-50). */
int tvh_merkle_gather_dev(const uint64_t* d_tree, size_t num_leaves, int layout,
                          const uint64_t* d_heap_idx, size_t count, uint64_t* d_out,
                          void* stream);

/* Host helpers: heap indices of MerkleTree::authentication_path (bottom-up
 * siblings, merkle_tree.cpp:93-113) and of authentication_structure (needed
 * minus computable, sorted descending, merkle_tree.cpp:228-265).
 * *out_count receives the number of indices; out may be NULL to query it. */
int tvh_auth_path_indices(size_t num_leaves, size_t leaf_index, uint64_t* out, size_t cap,
                          size_t* out_count);
int tvh_auth_structure_indices(size_t num_leaves, const size_t* leaf_indices, size_t count,
                               uint64_t* out, size_t cap, size_t* out_count);

/* ---- whole commit (gpu_stark.cu:2132-2610 step_main_table_commitment) ----- */

/* trace -> codewords (LDE) -> row digests (written into the tree's leaf
 * level) -> Merkle tree.  d_codewords: target_len*num_cols (column-major);
 * d_tree: tree buffer of the chosen layout; the root is node 2m-2 (level
 * order) or node 1 (heap).  Inputs as tvh_lde_batch_dev. */
int tvh_commit_dev(const uint64_t* d_trace, size_t num_cols, size_t trace_len,
                   const uint64_t* d_rand, size_t rand_len, uint64_t trace_offset,
                   uint64_t target_offset, size_t target_len, uint64_t* d_codewords,
                   uint64_t* d_tree, int layout, void* stream);

/* ---- FRI codeword commitment and Fiat-Shamir handoff ------------------------ */

/* leaves[i] = (c0, c1, c2, 0, 0) for XFE codeword element i (interleaved
 * limbs xfe[3i + k]): the FRI round leaves, not hashed.
 * Replaces qzc_xfe_to_digest_leaves_kernel (gpu_stark.cu:863-875). */
int tvh_xfe_leaves_dev(const uint64_t* d_xfe, size_t n, uint64_t* d_leaves, void* stream);
/* Leaves as tvh_xfe_leaves_dev written into the tree's leaf level, then the
 * Merkle tree of n (power of two) leaves in the given layout (the FRI round
 * commitment, gpu_stark.cu:5807,5874). */
int tvh_fri_commit_dev(const uint64_t* d_xfe, size_t n, uint64_t* d_tree, int layout,
                       void* stream);

/* Device-resident Fiat-Shamir sponge: 16 canonical u64 in device memory.
 * init = all zeros (variable-length domain, fs_init_sponge_gpu);
 * absorb = pad [data..., 1, 0...] to a multiple of 10 and, per block of 10,
 * overwrite the rate and permute (fs_absorb_device_gpu,
 * fiat_shamir_kernel.cuh:54; Tip5::pad_and_absorb_all); squeeze = out[0..10)
 * = rate, then permute (fs_squeeze_gpu).  The encoded Merkle root item that
 * the prover absorbs after a commitment is [0, root0..root4]
 * (gpu_stark.cu:2576-2587). */
int tvh_sponge_init_dev(uint64_t* d_state, void* stream);
int tvh_sponge_absorb_dev(uint64_t* d_state, const uint64_t* d_data, size_t len, void* stream);
int tvh_sponge_squeeze_dev(uint64_t* d_state, uint64_t* d_out, void* stream);

/* ---- batched commits of many tables ------------------------------------------ */

/* One table commit of a batch: the arguments of tvh_commit_dev. */
typedef struct tvh_table_desc {
    const uint64_t* trace;   /* column-major trace_len x num_cols */
    size_t num_cols;
    size_t trace_len;
    const uint64_t* rand;    /* rand_len coefficients per column, or NULL */
    size_t rand_len;
    uint64_t trace_offset;
    uint64_t target_offset;
    size_t target_len;
    uint64_t* codewords;     /* column-major target_len x num_cols */
    uint64_t* tree;          /* tree in `layout` */
    int layout;
} tvh_table_desc;

typedef struct tvh_batch_plan tvh_batch_plan;

/* Builds a plan that commits every table of the list (buffers as given) with
 * one launch per phase over all tables.  Uploads the per-phase argument
 * arrays (the only allocation).  Same results as tvh_commit_dev per table.
 * No reference counterpart: the reference commits tables one at a time
 * (gpu_stark.cu:2132-2610). */
int tvh_batch_plan_create(const tvh_table_desc* tables, size_t count, tvh_batch_plan** plan);
/* Enqueues the plan on `stream` (no allocation, no memset: capturable). */
int tvh_batch_plan_launch(tvh_batch_plan* plan, void* stream);
/* First call captures the plan's launches on `stream` (non-NULL) into a
 * hipGraph; every call launches that graph. */
int tvh_batch_plan_launch_graph(tvh_batch_plan* plan, void* stream);
int tvh_batch_plan_destroy(tvh_batch_plan* plan);
/* create + launch + wait for the stream + destroy. */
int tvh_commit_batch_dev(const tvh_table_desc* tables, size_t count, void* stream);

/* Copy count u64 from ordinary (pageable) host memory to the device on
 * `stream`.  Large copies go through the library's page-locked staging ring
 * (256 MiB per device, created on first use): worker threads fill its slots
 * in parallel while earlier slots are in flight.  Returns once every piece is
 * queued, so h_src may be reused at once and work queued later on `stream`
 * sees the data.  Calls on one device take turns on the ring.  Not for a
 * stream under graph capture: a captured copy would read a ring slot that
 * later calls refill.  No reference
 * counterpart: it replaces the prover's cudaMemcpy of host tables
 * (main_gpu.cpp:282-322) when the host side keeps its tables in pageable
 * memory. */
int tvh_upload_dev(uint64_t* d_dst, const uint64_t* h_src, size_t count, void* stream);

/* Synthetic uniform field elements for benchmarks and tests (no reference
 * counterpart): d[i] = SplitMix64(seed + (i+1) * golden gamma) reduced below p
 * (values >= p are replaced by v - p). */
int tvh_fill_uniform_dev(uint64_t* d, size_t count, uint64_t seed, void* stream);

/* Synthetic seeded table (no reference counterpart): the column-major
 * rows x cols table whose row-major draw order (element (r, c) is draw
 * r*cols + c) is the SplitMix64 stream with state `seed`, draws >= p rejected
 * -- the same table as the oracle's seeded draw, so a timed table can be
 * reproduced on the host.  Uses 65 u64 of the per-device arena (calls that
 * use the arena must be ordered on one stream). */
int tvh_fill_table_dev(uint64_t* d, size_t rows, size_t cols, uint64_t seed, void* stream);

/* Streaming ("frugal") commit: the same result as tvh_commit_dev without
 * materialising the codewords.  Columns are extended chunk_cols at a time into
 * a chunk buffer and absorbed into a per-row Tip5 sponge state that stays in
 * HBM between chunks (the reference's frugal coset loop bounds memory the same
 * way, gpu_stark.cu:2247-2420).  d_work holds at least
 * tvh_commit_frugal_workspace_elems(...) u64; the tree is written in the chosen
 * layout.  No other memory is used. */
size_t tvh_commit_frugal_workspace_elems(size_t num_cols, size_t trace_len, size_t target_len,
                                         size_t chunk_cols);
int tvh_commit_frugal_dev(const uint64_t* d_trace, size_t num_cols, size_t trace_len,
                          const uint64_t* d_rand, size_t rand_len, uint64_t trace_offset,
                          uint64_t target_offset, size_t target_len, size_t chunk_cols,
                          uint64_t* d_work, uint64_t* d_tree, int layout, void* stream);

/* ---- device-resident commit server --------------------------------------- */

/* No reference counterpart as one API: the reference serves proofs by forking
 * one prover process per request with one job per GPU
 * (src/prover_server/prover.cpp:459-505; rust/prover_server/src/server.rs:22,
 * 70-95) and keeps nothing on the device between requests.  Here one process
 * owns the GPU: jobs queue FIFO and a scheduler thread admits the head job as
 * soon as a stream slot is free and its buffers fit in the HBM budget
 * (sub-allocated from one allocation made at create time); admitted jobs run
 * concurrently on their slot streams; finished jobs keep codewords and tree
 * on the device until released, for opening queries. */
typedef struct tvh_server tvh_server;

typedef struct tvh_server_config {
    size_t hbm_budget_bytes; /* arena for job buffers; 0 = free memory - 16 GiB */
    int max_inflight;        /* stream slots (jobs on the GPU at once); 0 = 4.
                                With jobs from host memory, 8 keeps the uploads
                                running while earlier jobs compute (measured:
                                1.97e10 against 1.71e10 cells/s at 4). */
    size_t max_queue;        /* queued (not yet admitted) jobs; 0 = unbounded */
} tvh_server_config;

/* One LDE + commit job (tvh_commit_dev arguments, level-order tree).  Inputs
 * come from host memory, which must stay valid until the job has finished, or
 * are drawn on the device from `seed` with tvh_fill_uniform_dev when the
 * pointer is NULL (trace from seed, randomizers from seed ^ 0x5bd1e995). */
typedef struct tvh_job_desc {
    const uint64_t* host_trace; /* column-major trace_len x num_cols, or NULL */
    const uint64_t* host_rand;  /* rand_len per column (column-major), or NULL */
    uint64_t seed;
    size_t num_cols;
    size_t trace_len;
    size_t rand_len;
    uint64_t trace_offset;
    uint64_t target_offset;
    size_t target_len;
} tvh_job_desc;

/* steady-clock seconds of a job's submission, admission and completion */
typedef struct tvh_job_times {
    double submit_s;
    double admit_s;
    double done_s;
} tvh_job_times;

enum { TVH_JOB_QUEUED = 0, TVH_JOB_RUNNING = 1, TVH_JOB_DONE = 2, TVH_JOB_FAILED = 3 };

int tvh_server_create(const tvh_server_config* config, tvh_server** server);
/* Queues a job; never blocks on the GPU.  TVH_ERR_RANGE when the queue is
 * full.  A job larger than the whole budget fails with TVH_ERR_OOM. */
int tvh_server_submit(tvh_server* server, const tvh_job_desc* job, uint64_t* job_id);
int tvh_server_poll(tvh_server* server, uint64_t job_id, int* state);
/* Blocks until the job has finished; root (5 u64) and times may be NULL. */
int tvh_server_wait(tvh_server* server, uint64_t job_id, uint64_t* root, tvh_job_times* times);
/* Opening of a finished job at `count` leaf indices (row indices of the
 * codewords): rows_out (count x num_cols, row-major, may be NULL) receives the
 * codeword rows, auth_out the authentication structure digests
 * (MerkleTree::authentication_structure, merkle_tree.cpp:228-265; heap indices
 * from tvh_auth_structure_indices), *auth_count their number (5 u64 each).
 * auth_out may be NULL to query the count. */
int tvh_server_open(tvh_server* server, uint64_t job_id, const size_t* leaf_indices,
                    size_t count, uint64_t* rows_out, uint64_t* auth_out, size_t auth_cap,
                    size_t* auth_count);
/* Authentication paths (bottom-up siblings, merkle_tree.cpp:93-113) of each
 * leaf: paths_out holds count x log2(target_len) digests. */
int tvh_server_paths(tvh_server* server, uint64_t job_id, const size_t* leaf_indices,
                     size_t count, uint64_t* paths_out);
/* Device pointers of a finished job's codewords (column-major) and tree
 * (level order), valid until the job is released. */
int tvh_server_buffers(tvh_server* server, uint64_t job_id, const uint64_t** d_codewords,
                       const uint64_t** d_tree);
/* Frees a finished or failed job's memory. */
int tvh_server_release(tvh_server* server, uint64_t job_id);
int tvh_server_stats(tvh_server* server, size_t* arena_used, size_t* arena_size,
                     size_t* queued, size_t* running);
/* Waits for running jobs, fails queued ones, frees everything; waits for the
 * device so that the arena is back in the free pool when it returns. */
int tvh_server_destroy(tvh_server* server);

/* ---- host-pointer forms (synchronous; CudaBackend-style) ------------------ */

/* Replace CudaBackend::ntt_forward / ntt_inverse (src/gpu/cuda_ntt.cu:13-50). */
int tvh_ntt_forward_host(uint64_t* data, size_t n);
int tvh_ntt_inverse_host(uint64_t* data, size_t n);

/* Replace CudaBackend::lde_batch / RandomizedLDE per column
 * (src/gpu/cuda_lde.cu:26-83; src/lde/lde_randomized.cpp:71-154). */
int tvh_lde_batch_host(const uint64_t* trace, size_t num_cols, size_t trace_len,
                       const uint64_t* rand, size_t rand_len, uint64_t trace_offset,
                       uint64_t target_offset, size_t target_len, uint64_t* out);

/* XFE form of tvh_lde_batch_host in the interleaved layout of
 * tvh_lde_xfe_batch_dev (RandomizedLDE::extend_xfield_column_with_randomizer /
 * _with_xfield_randomizer, src/lde/lde_randomized.cpp:179-395). */
int tvh_lde_xfe_batch_host(const uint64_t* xfe_trace, size_t num_cols, size_t trace_len,
                           const uint64_t* rand, size_t rand_len, uint64_t trace_offset,
                           uint64_t target_offset, size_t target_len, uint64_t* xfe_out,
                           int use_xfe_randomizers);

/* Replace Backend::tip5_permutation_batch / hash_pairs on host buffers. */
int tvh_tip5_permutation_batch_host(uint64_t* states, size_t num_states);
/* tvh_tip5_rounds_dev on host states (trace may be NULL). */
int tvh_tip5_rounds_host(uint64_t* states, size_t num_states, uint32_t first_round,
                         uint32_t num_rounds, uint64_t* trace);
int tvh_hash_pairs_host(const uint64_t* left, const uint64_t* right, uint64_t* out,
                        size_t count);

/* Row digests of a column-major host table. */
int tvh_hash_bfield_rows_host(const uint64_t* table, size_t num_rows, size_t num_cols,
                              uint64_t* digests);

/* Merkle tree of host leaves; tree_out as tvh_merkle_tree_dev (may be NULL),
 * root_out 5 u64.  Replaces Backend::merkle_root / merkle_tree_full. */
int tvh_merkle_tree_host(const uint64_t* leaves, size_t num_leaves, int layout,
                         uint64_t* tree_out, uint64_t* root_out);

/* Whole commit on host buffers (LDETable::compute_lde_and_merkle,
 * src/lde/lde_table.cpp:86-101).  codewords_out (target_len*num_cols,
 * column-major) and digests_out (target_len*5) may be NULL. */
int tvh_commit_host(const uint64_t* trace, size_t num_cols, size_t trace_len,
                    const uint64_t* rand, size_t rand_len, uint64_t trace_offset,
                    uint64_t target_offset, size_t target_len, uint64_t* codewords_out,
                    uint64_t* digests_out, uint64_t* root_out);

#ifdef __cplusplus
}
#endif
#endif /* TVH_H */
