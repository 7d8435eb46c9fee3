/*
 * CPU oracle for the table-commitment path (Goldilocks coset LDE, Tip5 row
 * hashing, Tip5 Merkle tree).
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * prover's CPU algorithm.  It exists to check the HIP implementation and to be
 * timed as a CPU baseline.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The shipped library never calls it.
 *
 * Every function names the reference file:line it restates.  All values are
 * canonical u64 field elements in [0, p), p = 2^64 - 2^32 + 1, exactly like the
 * reference's BFieldElement (include/types/b_field_element.hpp:31-41).
 *
 * Parity pins (see tests/test_oracle.py):
 *   field ops      - u128 arithmetic identities, edge cases around p
 *   roots of unity - order checks, Rust ntt KAT  [1,4,0,0] -> [...]
 *   NTT            - O(n^2) DFT brute force, roundtrip, Rust ntt KAT
 *   coset LDE      - Horner evaluation of the interpolant at 7*w^j, golden roots
 *   Tip5           - Rust hash10 / hash_varlen KATs, S-box closed form, MDS = SHA-256("Tip5")
 *   Merkle         - golden roots from the compiled reference, verify() round trips
 */
#ifndef TVM_ORACLE_H
#define TVM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_P 0xFFFFFFFF00000001ull

/* ---- field (b_field_element.cpp) ---- */
uint64_t orc_add(uint64_t a, uint64_t b);
uint64_t orc_sub(uint64_t a, uint64_t b);
uint64_t orc_mul(uint64_t a, uint64_t b);
uint64_t orc_pow(uint64_t a, uint64_t e);
uint64_t orc_inv(uint64_t a);
uint64_t orc_root_of_unity(uint32_t log2_order);
uint64_t orc_montyred(uint64_t hi, uint64_t lo);

/* ---- NTT (ntt.cpp) ---- */
int orc_ntt_forward(uint64_t* a, size_t n);
int orc_ntt_inverse(uint64_t* a, size_t n);
int orc_evaluate_on_coset(const uint64_t* coeffs, size_t ncoeffs, size_t m,
                          uint64_t offset, uint64_t* out);

/* ---- randomized LDE (lde_randomized.cpp) ---- */
/* BFE column: interpolate over (n, trace_off), add zerofier*randomizer, evaluate
 * on (m, target_off); polynomials longer than m are evaluated in m-chunks. */
int orc_lde_column(const uint64_t* trace, size_t n, uint64_t trace_off,
                   const uint64_t* rand, size_t k, uint64_t target_off,
                   size_t m, uint64_t* out);
/* One limb of an XFE column: as above but coefficients beyond m are truncated
 * (lde_randomized.cpp:270-275).  rand may be NULL (k ignored). */
int orc_lde_limb_truncating(const uint64_t* trace, size_t n, uint64_t trace_off,
                            const uint64_t* rand, size_t k, uint64_t target_off,
                            size_t m, uint64_t* out);
/* Column-major table helpers (cols columns of n rows -> cols columns of m rows). */
int orc_lde_table_colmajor(const uint64_t* trace, size_t n, size_t cols,
                           uint64_t trace_off, const uint64_t* rand, size_t k,
                           uint64_t target_off, size_t m, uint64_t* out);

/* ---- Tip5 (tip5.cpp) ---- */
int orc_lde_sample_colmajor(const uint64_t* trace, size_t n, size_t cols, uint64_t toff,
                            const uint64_t* rand, size_t k, uint64_t goff, size_t m,
                            const uint64_t* rows, size_t nrows, int truncate, uint64_t* out);
void orc_tip5_round(uint64_t st[16], int round_index);
void orc_tip5_permutation(uint64_t st[16]);
void orc_hash_10(const uint64_t in[10], uint64_t out[5]);
void orc_hash_varlen(const uint64_t* in, size_t len, uint64_t out[5]);
void orc_hash_pair(const uint64_t l[5], const uint64_t r[5], uint64_t out[5]);
/* hash_varlen of every row of a column-major table (element (r,c) at c*rows+r). */
void orc_hash_rows_colmajor(const uint64_t* table, size_t rows, size_t cols,
                            uint64_t* digests);
const uint8_t* orc_tip5_lookup_table(void);
const uint64_t* orc_tip5_round_constants(void);
const int64_t* orc_tip5_mds_first_column(void);

/* ---- Merkle (merkle_tree.cpp) ---- */
/* Heap layout: nodes[5*i] for i in [1, 2n); root at node 1, leaves at [n, 2n). */
int orc_merkle_heap(const uint64_t* leaves, size_t n, uint64_t* nodes);
/* Level order (GPU layout, merkle_kernel.cu:60-102): leaves [0,n), then each
 * level, root at node 2n-2. */
int orc_merkle_level_order(const uint64_t* leaves, size_t n, uint64_t* tree);
/* authentication_structure node indices (heap, 1-indexed), sorted descending. */
long orc_auth_structure_indices(size_t n, const size_t* leaf_idx, size_t cnt,
                                size_t* out_idx, size_t out_cap);
int orc_merkle_verify_path(const uint64_t root[5], size_t leaf_index,
                           const uint64_t leaf[5], const uint64_t* path,
                           size_t path_len);

/* ---- whole commit (lde_table.cpp:86-101) ---- */
/* trace column-major n x cols; writes codewords (m x cols, column-major),
 * digests (m x 5) and root (5).  codewords/digests may be NULL. */
int orc_commit(const uint64_t* trace, size_t n, size_t cols, uint64_t trace_off,
               const uint64_t* rand, size_t k, uint64_t target_off, size_t m,
               uint64_t* codewords, uint64_t* digests, uint64_t root[5]);

/* ---- inputs ---- */
/* SplitMix64 uniform Goldilocks stream with rejection of values >= p. */
void orc_splitmix_fill(uint64_t seed, uint64_t* out, size_t count);

#ifdef __cplusplus
}
#endif
#endif
