// Internal declarations shared by the HIP translation units of libtvh.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstddef>
#include <vector>

#include "../../include/tvh.h"
#include "gl64.hpp"
#include "tip5_dev.hpp"

namespace tvh {

using gl::u32;
using gl::u64;

// Twiddle tables of one device.  w = 7^((p-1)/2^24) (the reference's 2^24-th
// root); entries for direction d (0 forward, 1 inverse, using w^-1):
//   r4096[d][e] = w_4096^(+-e)               e < 4096
//   lo[d][e]    = w^(+-e), hi[d][e] = w^(+-4096 e)   e < 4096
// so w_N^x = lo[x & 4095] * hi[x >> 12] for x < 2^24 after scaling x by 2^24/N.
// All entries are stored in Montgomery form (w * 2^64 mod p, canonical) and
// applied with gl::mont_mul, which returns the canonical product for a
// canonical operand.
struct Roots {
    const u64* r4096[2];
    const u64* lo[2];
    const u64* hi[2];
};

constexpr u32 kMaxLogN = 24;

struct DeviceState {
    int device = -1;
    Roots roots{};
    u64* table_mem = nullptr;  // backing store of the twiddle tables
    // growable arena for entry points that need scratch (documented in tvh.h)
    void* arena = nullptr;
    size_t arena_bytes = 0;
    // this device's address of the static scratch of the allocation-free NTT
    u64* noalloc_scratch = nullptr;
};

// Elements of the static scratch behind tvh_ntt_batched_noalloc_dev (32 MiB).
constexpr u64 kNoAllocScratchElems = 1ull << 22;
// Resolves this device's address of that scratch (ntt_kernels.hip); called
// once per device at set-up, with the device current.
int noalloc_scratch_setup(DeviceState* ds);
// The current device if tvh_init (or an earlier call) has set it up; nullptr
// and an error otherwise.  Never allocates.
DeviceState* initialized_device();

// Host-to-device copy of `bytes` from pageable host memory on stream st through
// the device's page-locked staging ring (stager.hip); returns once every piece
// is queued (the source may then be reused).  Calls on one device take turns.
int staged_h2d(void* dst, const void* src, size_t bytes, hipStream_t st);

// State of the current device, initialised on first use (thread-safe).
// Returns nullptr and sets the error on failure.
DeviceState* current_device();

// Arena of at least `bytes` on the current device (grows; not for concurrent
// use of the same device from several streams: calls are serialised by the
// caller's stream order only).
void* arena(DeviceState* ds, size_t bytes, hipStream_t st);

int set_error(int status, const char* fmt, ...);
int check_launch();

inline bool is_pow2(u64 x) { return x && !(x & (x - 1)); }
inline u32 log2u(u64 x) {
    u32 l = 0;
    while ((1ull << l) < x) l++;
    return l;
}

// ---- kernels launchers (ntt_kernels.hip) ----
// Input scaling of a forward batched NTT, applied as its column step loads:
// element i of column c enters as (x + [i < k] rand[c*k + i] * delta) * base^i.
// tables: kNttScaleTableElems u64 of device memory for the launch's powers of
// base; rand is null when k = 0.
struct NttScale {
    u64 base;
    u64* tables;
    const u64* rand;
    u64 k, delta;
};
constexpr u64 kNttScaleTableElems = 8192;  // N1 + N2 <= 2 * 4096
// What the column step reads (NttScale with the launch's tables filled in):
// element i = n1 * N2 + n2 takes base^i = p1[n1] * p2[n2], p1[n1] = base^(N2 n1),
// p2[n2] = base^n2 (Montgomery form).
struct NttScaleDev {
    const u64* p1;
    const u64* p2;
    const u64* rand;
    u64 k, delta;
};
// Batched NTT from src to dst (the same buffer, or buffers that do not
// overlap) through a scratch of scratch_elems u64 (at least n; columns go
// through it in groups of scratch_elems / n).  Launches that share a scratch
// must be ordered (one stream).  scale (forward only) scales the input.
int launch_ntt_batched(const u64* src, u64* dst, u64 n, u64 cols, bool inverse, u64* scratch,
                       u64 scratch_elems, hipStream_t st, const NttScale* scale = nullptr);
// The scratch launch_ntt_batched uses for (n, cols) from the per-device arena
// (groups of at most 2 GiB).
u64 ntt_scratch_elems(u64 n, u64 cols);
// LDE of `cols` B-field columns (column-major in, column-major out).  rand may
// be null (k = 0).  truncate selects the XFE evaluation rule.  lift_stride:
// randomizer of column c is rand[(c / lift_div) * rand_col_stride + ...] with
// only limbs (c % lift_div) == lift_limb ... (see ntt_kernels.hip)
struct RandSpec {
    const u64* ptr = nullptr;  // null: no randomizer
    u64 k = 0;                 // coefficients per column
    // address of coefficient i for B-field column c:
    //   ptr[(c / group) * group_stride + (c % group) * limb_stride + i * elem_stride]
    // columns with (c % group) not in the mask get no randomizer
    u32 group = 1;
    u64 group_stride = 0;
    u64 limb_stride = 0;
    u64 elem_stride = 1;
    u32 limb_mask = 1;  // bit j set: limb j carries a randomizer
};
int launch_lde(const u64* trace, u64 cols, u64 n, const RandSpec& rs, u64 toff, u64 goff, u64 m,
               bool truncate, u64* out, hipStream_t st);

// Kernel arguments of the three LDE steps (one table).  Multi-table launches
// take device arrays of these plus per-segment block offsets.
struct Step1Args {
    const u64* src;
    u64 sstride;
    u64* dst;
    u64 dstride;
    u32 logN, a, E;
    u64 ntr;
    u32 logNT;
};
struct Step3Args {
    u64* out;
    u32 logN, a, logE;
    u64 cols, M;
    u32 logNT;
};
struct LdeArgs {
    u64* out;
    u32 logN, a, E, logE;
    u64 cols, M;
    const u64* src;    // rows of step 2: src[c*src_col_stride + k1*src_row_stride + n2]
    u64 src_col_stride, src_row_stride;
    const u64* pa;     // pa[s*N1 + k1] = N^-1 * beta_s^k1      (Montgomery form)
    const u64* pb;     // pb[s*N2 + k2] = beta_s^(N1*k2)         (Montgomery form)
    u64 delta_n[16];   // (g^N w_E^s - o^N) * N  (or -o^N * N when truncating with E = 1)
    u64 toff;          // o
    RandSpec rs;
    u32 logNT;
};
// LDE kernel variant: bits 0-1 the form of the fused row step (LDE_WIDE: one
// kernel, all E cosets of a row in one 1024-thread LDS tile; LDE_SPLIT: the
// inverse row DFTs as a kernel of their own, then the expansion); bits 2-3 the
// column kernels' tile, 2^(14 - code) elements with 16 per thread; bit 4
// (LDE_SPLIT only) halves the row kernels' tile to 2^13 elements (512 threads,
// two blocks per CU), which also halves the row length N2 when E * N2 would
// not fit (all cosets of a row stay in one block); bit 5 (LDE_SPLIT only, not
// with bit 4) gives only the inverse row kernel the 2^13 tile; bit 6 (LDE_SPLIT
// only, alone) gives both row kernels 2^12-element tiles (256 threads).
enum LdeVariant {
    LDE_AUTO = 0,
    LDE_WIDE = 1,
    LDE_SPLIT = 2,
    LDE_ROW_HALF = 16,
    LDE_INV_HALF = 32,
    LDE_ROW_QUARTER = 64
};
// Tuned choices are kept per (log2 N, log2 E, width class): tables narrower
// than 64 columns, 64..255, and 256 or more (a narrow table launches fewer
// blocks per kernel, which can favour a smaller tile).
constexpr int kWidthClasses = 3;
inline int width_class(u64 cols) { return cols < 64 ? 0 : cols < 256 ? 1 : 2; }
int lde_variant(u32 logN, u32 logE, u64 cols);  // the tuned choice (never LDE_AUTO)
bool lde_variant_valid(int v);
// v = LDE_AUTO restores the default; wc < 0 sets every width class
int set_lde_variant(u32 logN, u32 logE, int v, int wc = -1);

struct LdePlan {
    bool k1, k3;
    int k2_variant;  // full variant code (form | tile << 2 | LDE_ROW_HALF)
    bool split;      // LDE_SPLIT: s2inv (k_lde_rowinv) runs before s2 (k_lde_expand)
    LdeArgs s2inv;
    u32 inv_logNT, inv_blocks;  // s2inv.logNT == inv_logNT
    size_t inv_lds;
    u32 col_lt;      // log2 tile of the column kernels (K1, K3)
    u32 row_lt;      // log2 tile of the row kernels (k_lde_expand; the split's bound)
    u32 inv_lt;      // log2 tile of k_lde_rowinv
    Step1Args s1;
    LdeArgs s2;
    Step3Args s3;
    u32 s1_blocks, s2_blocks, s3_blocks;
    size_t s1_lds, s2_lds, s3_lds;
};
// batch = true plans the column and row kernels with the multi-table tiles
// (kBatchColLt, kBatchRowLt; the wide form keeps its full tile).
int prepare_lde(const u64* trace, u64 cols, u64 n, const RandSpec& rs, u64 toff, u64 goff,
                u64 m, bool truncate, u64* out, LdePlan* plan, bool batch);
// Builds (once) the coset tables every LDE variant of this shape and these
// offsets reads, so that later LDE launches find them cached.
int lde_prepare_consts(u64 n, u64 m, u64 toff, u64 goff);
// Waits for the device and frees every cached coset table.
int lde_release_consts();
// Phases of the multi-table LDE, in launch order.
enum LdePhase {
    LDE_PHASE_STEP1 = 0,   // k_step1_multi<true>
    LDE_PHASE_ROWINV = 1,  // k_lde_rowinv_multi (split-form tables)
    LDE_PHASE_STEP2 = 2,   // k_lde_step2_multi (wide-form tables)
    LDE_PHASE_EXPAND = 3,  // k_lde_expand_multi (split-form tables)
    LDE_PHASE_STEP3 = 4,   // k_lde_step3_multi
    LDE_PHASES = 5
};
constexpr u32 kBatchColLt = 13;  // column-kernel tile of multi-table launches
constexpr u32 kBatchRowLt = 13;  // row-kernel tile of multi-table launches (measured best over 2^12..2^14)
int launch_lde_multi(int phase, const void* d_args, const u32* d_start, u32 nseg, u32 blocks,
                     size_t lds, hipStream_t st);
int launch_fill_uniform(u64* d, u64 count, u64 seed, hipStream_t st);
int launch_xfe_interleave(const u64* src, u64* dst, u64 cols, u64 len, bool to_limb_major,
                          hipStream_t st);

// ---- coset evaluation, digest scatter, transposes (coset_kernels.hip) ----
// y[c*n + i] = (x[c*n + i] + [i < k] rand[c*k + i] * delta) * base^i (x may be y),
// n <= 2^24; tables: scale_table_elems(n) u64 of device scratch for the powers
u64 scale_table_elems(u64 n);
int launch_scale_table(const u64* x, u64* y, u64 n, u64 cols, u64 base, const u64* rand, u64 k,
                       u64 delta, u64* tables, hipStream_t st);
// t[i] = base^i for i < count, Montgomery form
int launch_pow_table(u64* t, u64 base, u64 count, hipStream_t st);
// leaves[(j * num_cosets + coset) * 5 + e] = dig[j * 5 + e], j < len
int launch_scatter_digests(const u64* dig, u64* leaves, u64 len, u64 coset, u64 num_cosets,
                           hipStream_t st);
// dst[c * rows + r] = src[r * cols + c]
int launch_transpose(const u64* src, u64* dst, u64 rows, u64 cols, hipStream_t st);
// d (column-major rows x cols) = the SplitMix64 stream from seed in row-major
// draw order, draws >= p rejected; rej: fill_table_scratch_elems() u64 of scratch
int launch_fill_table(u64* d, u64 rows, u64 cols, u64 seed, u64* rej, hipStream_t st);
u64 fill_table_scratch_elems();

// ---- Tip5 launchers (tip5_kernels.hip) ----
struct HashArgs {
    const u64* tab;
    u64 rows;
    u32 cols;
    u64* dig;
};
struct MerkleArgs {
    const u64* src;  // children level
    u64* dst[9];     // parents of the next `levels` levels
    u32 levels;
    u32 T;           // parents per block at the first level (one thread each, or one
                     // 16-lane group each when group16)
    u32 group16;     // 1: each hash runs on 16 lanes (permutation_group16), 256-thread blocks
};
struct MerkleGroup {
    MerkleArgs args;
    u32 blocks;
};
int merkle_groups(u64* tree, u64 n, int layout, std::vector<MerkleGroup>& groups);
int launch_hash_rows_multi(const HashArgs* d_args, const u32* d_start, u32 nseg, u32 blocks,
                           hipStream_t st);
int launch_merkle_multi(const MerkleArgs* d_args, const u32* d_start, u32 nseg, u32 blocks,
                        hipStream_t st);
int launch_hash_rows(const u64* tab, u64 rows, u64 cols, u64* dig, hipStream_t st);
int launch_permutation(u64* states, u64 num, hipStream_t st);
int launch_permutation_rounds(u64* states, u64 num, u32 first, u32 count, u64* trace,
                              hipStream_t st);
// mode 0 absorb (pad and absorb all), 1 squeeze, 2 init
int launch_sponge(u64* state, const u64* data, u64 len, u64* out, int mode, hipStream_t st);
int launch_xfe_leaves(const u64* xfe, u64 n, u64* leaves, hipStream_t st);
int launch_absorb_chunk(const u64* chunk, u64 rows, u64 ccols, u64* state, u64* pend, u32 npend,
                        bool first, bool final_chunk, u64* dig, hipStream_t st);
int launch_hash_pairs(const u64* l, const u64* r, u64* out, u64 count, hipStream_t st);
u64 merkle_level_base(u64 n, u32 level, int layout);
int launch_merkle_levels(u64* tree, u64 n, int layout, hipStream_t st);
int launch_gather_heap(const u64* tree, u64 n, int layout, const u64* idx, u64 count, u64* out,
                       hipStream_t st);
int ntt_set_attributes();

}  // namespace tvh
