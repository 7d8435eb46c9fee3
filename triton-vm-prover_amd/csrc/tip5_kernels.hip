// Tip5 hashing kernels: row digests, pair hashing, Merkle trees, auth paths.
//
// Row hashing restates Tip5::hash_varlen (src/hash/tip5.cpp:246-282) for one
// row per lane over a column-major table (element (r, c) at c*rows + r, the
// reference GPU layout, include/gpu/kernels/row_hash_kernel.cuh:18-33): zero
// state, overwrite absorb of 10 elements per permutation, last block padded
// [rest, 1, 0...].  Reads of one column by a wavefront are 64 consecutive
// u64, so HBM traffic is coalesced without staging.
//
// Merkle trees follow merkle_tree.cpp:41-61 (parent = hash_pair(left,right)).
// Node (level l, index i) of an n-leaf tree lives at base(l) + i with
//   level order (merkle_kernel.cu:60-102): base(l) = 2n - (2n >> l)
//   heap        (merkle_tree.cpp:12-33):   base(l) = n >> l      (1-indexed)
// Levels of more than 2^16 parents take a launch each, one hash per thread.
// Above them a block hashes up to 16 parents on 16 lanes each, keeps them in
// LDS and continues to their subtree root (up to 5 levels per launch).
#include "tvh_internal.hpp"

namespace tvh {

using gl::u32;
using gl::u64;

// Sponge states are kept in Montgomery form (tip5_dev.hpp); digests leave canonical.
__device__ __forceinline__ void store_digest(u64* dst, const u64 s[tip5::STATE]) {
#pragma unroll
    for (int i = 0; i < tip5::DIGEST; i++) dst[i] = tip5::from_m(s[i]);
}

// Segment lookup for multi-table launches (start[nseg] = total blocks).
__device__ __forceinline__ u32 find_seg(const u32* __restrict__ start, u32 nseg, u32 blk) {
    u32 lo = 0, hi = nseg;
    while (hi - lo > 1) {
        u32 mid = (lo + hi) >> 1;
        if (start[mid] <= blk) lo = mid;
        else hi = mid;
    }
    return lo;
}

// rows x cols column-major table -> rows digests (5 u64 each, row-major).
__device__ __forceinline__ void hash_rows_body(const HashArgs& A, u64 blk, const uint8_t* lut) {
    u64 row = blk * 256 + threadIdx.x;
    if (row >= A.rows) return;
    const u64 rows = A.rows;
    const u32 cols = A.cols;
    u64 s[tip5::STATE];
#pragma unroll
    for (int i = 0; i < tip5::STATE; i++) s[i] = 0;
    const u64* p = A.tab + row;
    u32 used = 0;
    for (; used + tip5::RATE <= cols; used += tip5::RATE) {
#pragma unroll
        for (int i = 0; i < tip5::RATE; i++)
            s[i] = tip5::to_m(__builtin_nontemporal_load(p + (u64)(used + i) * rows));
        tip5::permutation_m(s, lut);
    }
    u32 rest = cols - used;
#pragma unroll
    for (int i = 0; i < tip5::RATE; i++) {
        u32 ui = (u32)i;
        s[i] = ui < rest ? tip5::to_m(__builtin_nontemporal_load(p + (u64)(used + ui) * rows))
                         : (ui == rest ? tip5::kOneM : 0ull);
    }
    tip5::permutation_m(s, lut);
    store_digest(A.dig + row * tip5::DIGEST, s);
}

// At least 4 waves per SIMD (<= 128 VGPRs; 4 to 6 waves time the same).
__global__ void __launch_bounds__(256, 4) k_hash_rows(HashArgs A) {
    __shared__ uint8_t lut[256];
    tip5::load_lut(lut);
    __syncthreads();
    hash_rows_body(A, blockIdx.x, lut);
}

__global__ void __launch_bounds__(256, 4) k_hash_rows_multi(const HashArgs* __restrict__ args,
                                                         const u32* __restrict__ start, u32 nseg) {
    __shared__ uint8_t lut[256];
    tip5::load_lut(lut);
    __syncthreads();
    u32 seg = find_seg(start, nseg, blockIdx.x);
    HashArgs A = args[seg];
    hash_rows_body(A, blockIdx.x - start[seg], lut);
}

// Column-chunk absorption for the streaming ("frugal") commit: each row's
// hash_varlen sponge continues across chunks of columns.  The row's element
// stream is the npend elements left over from the previous chunks (pend,
// column-major by position) followed by this chunk's ccols columns; full rate
// blocks are absorbed by overwrite, the remainder is carried.  All rows have
// the same counts, so lanes never diverge.  State is column-major by word
// (state[i*rows + row]).  The FINAL chunk absorbs the padding block
// [rest, 1, 0...] and writes the canonical digest (tip5.cpp:246-282).  The
// carried state is in Montgomery form, the carried elements are canonical.
template <bool FINAL>
__global__ void __launch_bounds__(256) k_absorb_chunk(const u64* __restrict__ chunk, u64 rows,
                                                      u32 ccols, u64* __restrict__ state,
                                                      u64* __restrict__ pend, u32 npend,
                                                      int first, u64* __restrict__ dig) {
    __shared__ uint8_t lut[256];
    tip5::load_lut(lut);
    __syncthreads();
    u64 row = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= rows) return;
    u64 s[tip5::STATE];
#pragma unroll
    for (int i = 0; i < tip5::STATE; i++) s[i] = first ? 0ull : state[(u64)i * rows + row];
    auto elem = [&](u32 j) -> u64 {
        return j < npend ? pend[(u64)j * rows + row] : chunk[(u64)(j - npend) * rows + row];
    };
    const u32 total = npend + ccols;
    u32 used = 0;
    for (; used + tip5::RATE <= total; used += tip5::RATE) {
#pragma unroll
        for (int i = 0; i < tip5::RATE; i++) s[i] = tip5::to_m(elem(used + i));
        tip5::permutation_m(s, lut);
    }
    const u32 rest = total - used;
    if (FINAL) {
#pragma unroll
        for (int i = 0; i < tip5::RATE; i++) {
            u32 ui = (u32)i;
            s[i] = ui < rest ? tip5::to_m(elem(used + ui)) : (ui == rest ? tip5::kOneM : 0ull);
        }
        tip5::permutation_m(s, lut);
        store_digest(dig + row * tip5::DIGEST, s);
    } else {
        // carried elements first (they may be read from pend itself: read all, then write)
        u64 carry[tip5::RATE];
#pragma unroll
        for (int i = 0; i < tip5::RATE; i++) carry[i] = (u32)i < rest ? elem(used + i) : 0ull;
#pragma unroll
        for (int i = 0; i < tip5::RATE; i++)
            if ((u32)i < rest) pend[(u64)i * rows + row] = carry[i];
#pragma unroll
        for (int i = 0; i < tip5::STATE; i++) state[(u64)i * rows + row] = s[i];
    }
}

// Tip5 permutation of num states (16 u64 each, in place); outputs canonical.
__global__ void __launch_bounds__(256) k_permutation(u64* __restrict__ st, u64 num) {
    __shared__ uint8_t lut[256];
    tip5::load_lut(lut);
    __syncthreads();
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num) return;
    u64 s[tip5::STATE];
#pragma unroll
    for (int k = 0; k < tip5::STATE; k++) s[k] = tip5::to_m(st[i * tip5::STATE + k]);
    tip5::permutation_m(s, lut);
#pragma unroll
    for (int k = 0; k < tip5::STATE; k++) st[i * tip5::STATE + k] = tip5::from_m(s[k]);
}

// Rounds [first, first + count) of the permutation on num states (in place).
// trace (may be null) receives, per state, the input and the state after each
// of those rounds: trace[(i * (count + 1) + j) * 16 + k].  Outputs canonical.
__global__ void __launch_bounds__(256) k_permutation_rounds(u64* __restrict__ st, u64 num,
                                                            u32 first, u32 count,
                                                            u64* __restrict__ trace) {
    __shared__ uint8_t lut[256];
    tip5::load_lut(lut);
    __syncthreads();
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num) return;
    u64 s[tip5::STATE];
#pragma unroll
    for (int k = 0; k < tip5::STATE; k++) s[k] = tip5::to_m(st[i * tip5::STATE + k]);
    u64* tr = trace ? trace + i * (count + 1) * tip5::STATE : nullptr;
    if (tr) {
#pragma unroll
        for (int k = 0; k < tip5::STATE; k++) tr[k] = st[i * tip5::STATE + k];
    }
    for (u32 r = 0; r < count; r++) {
        tip5::round_m(s, (int)(first + r), lut);
        if (tr) {
#pragma unroll
            for (int k = 0; k < tip5::STATE; k++) tr[(r + 1) * tip5::STATE + k] = tip5::from_m(s[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < tip5::STATE; k++) st[i * tip5::STATE + k] = tip5::from_m(s[k]);
}

__device__ __forceinline__ void hash_pair(const u64* l, const u64* r, u64 s[tip5::STATE],
                                          const uint8_t* lut) {
#pragma unroll
    for (int k = 0; k < 5; k++) {
        s[k] = tip5::to_m(l[k]);
        s[5 + k] = tip5::to_m(r[k]);
    }
#pragma unroll
    for (int k = 10; k < tip5::STATE; k++) s[k] = tip5::kOneM;
    tip5::permutation_m(s, lut);
}

// out[i] = hash_pair(left[i], right[i])  (Backend::hash_pairs, backend.hpp:110-115)
__global__ void __launch_bounds__(256) k_hash_pairs(const u64* __restrict__ left,
                                                    const u64* __restrict__ right,
                                                    u64* __restrict__ out, u64 count) {
    __shared__ uint8_t lut[256];
    tip5::load_lut(lut);
    __syncthreads();
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    u64 s[tip5::STATE];
    hash_pair(left + 5 * i, right + 5 * i, s, lut);
    store_digest(out + 5 * i, s);
}

// One launch of the Merkle build: children are the nodes of one level; each
// block hashes 2T children into T parents (T active threads, T <= blockDim)
// and continues for `levels` levels inside LDS.  Level j (1-based) parents go
// to dst[j-1] + index.
__device__ __forceinline__ void merkle_body(const MerkleArgs& A, u64 blk, const uint8_t* lut,
                                            u64* buf) {
    const u32 T = A.T;
    const u32 t = threadIdx.x;
    u64 s[tip5::STATE];
    if (t < T) {
        u64 parent = blk * T + t;
        hash_pair(A.src + 10 * parent, A.src + 10 * parent + 5, s, lut);
        u64 d[5];
#pragma unroll
        for (int k = 0; k < 5; k++) d[k] = tip5::from_m(s[k]);
#pragma unroll
        for (int k = 0; k < 5; k++) {
            A.dst[0][5 * parent + k] = d[k];
            buf[5 * t + k] = d[k];
        }
    }
    u32 width = T;
    for (u32 j = 1; j < A.levels; j++) {
        __syncthreads();
        width >>= 1;
        u64 d[5];
        if (t < width) {
            hash_pair(buf + 10 * t, buf + 10 * t + 5, s, lut);
#pragma unroll
            for (int k = 0; k < 5; k++) d[k] = tip5::from_m(s[k]);
        }
        __syncthreads();
        if (t < width) {
            u64 parent = blk * width + t;
#pragma unroll
            for (int k = 0; k < 5; k++) {
                A.dst[j][5 * parent + k] = d[k];
                buf[5 * t + k] = d[k];
            }
        }
    }
}

// The same levels with each hash on a group of 16 lanes (tip5::permutation_group16):
// a 256-thread block holds 16 groups, hashes T <= 16 parents of its first level
// and continues in LDS to their subtree root.  Group g, lane w: state word w of
// parent g is child word w (the two children are adjacent: words 0..9), then
// capacity ones.
__device__ __forceinline__ void merkle_body_group16(const MerkleArgs& A, u64 blk,
                                                    const uint8_t* lut, u64* buf) {
    const u32 g = threadIdx.x >> 4, w = threadIdx.x & 15;
    u32 width = A.T;
    u64 s = tip5::kOneM;
    if (g < width && w < 10) s = tip5::to_m(A.src[10 * (blk * width + g) + w]);
    s = tip5::permutation_group16(s, w, lut);
    if (g < width && w < tip5::DIGEST) {
        const u64 d = tip5::from_m(s);
        A.dst[0][5 * (blk * width + g) + w] = d;
        buf[5 * g + w] = d;
    }
    for (u32 j = 1; j < A.levels; j++) {
        __syncthreads();  // the previous level's digests are in buf
        width >>= 1;
        s = tip5::kOneM;
        if (g < width && w < 10) s = tip5::to_m(buf[10 * g + w]);
        __syncthreads();  // every read of buf precedes this level's writes
        s = tip5::permutation_group16(s, w, lut);
        if (g < width && w < tip5::DIGEST) {
            const u64 d = tip5::from_m(s);
            A.dst[j][5 * (blk * width + g) + w] = d;
            buf[5 * g + w] = d;
        }
    }
}

__global__ void __launch_bounds__(256) k_merkle_subtrees(MerkleArgs A) {
    __shared__ uint8_t lut[256];
    __shared__ u64 buf[256 * 5];
    tip5::load_lut(lut);
    __syncthreads();
    if (A.group16) merkle_body_group16(A, blockIdx.x, lut, buf);
    else merkle_body(A, blockIdx.x, lut, buf);
}

__global__ void __launch_bounds__(256) k_merkle_subtrees_multi(const MerkleArgs* __restrict__ args,
                                                               const u32* __restrict__ start,
                                                               u32 nseg) {
    __shared__ uint8_t lut[256];
    __shared__ u64 buf[256 * 5];
    tip5::load_lut(lut);
    __syncthreads();
    u32 seg = find_seg(start, nseg, blockIdx.x);
    const MerkleArgs& A = args[seg];  // read in place: a private copy would spill dst[]
    if (A.group16) merkle_body_group16(A, blockIdx.x - start[seg], lut, buf);
    else merkle_body(A, blockIdx.x - start[seg], lut, buf);
}

// Gather: out[q] = node with heap index idx[q] (1 <= idx < 2n) of a tree in
// either layout.  Out-of-range indices produce zeros.
__global__ void k_gather_heap(const u64* __restrict__ tree, u64 n, u32 logn, int layout,
                              const u64* __restrict__ idx, u64 count, u64* __restrict__ out) {
    u64 q = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count * 5) return;
    u64 h = idx[q / 5];
    if (h == 0 || h >= 2 * n) {
        out[q] = 0;
        return;
    }
    u64 pos = h;
    if (layout == TVH_TREE_LEVEL_ORDER) {
        u32 fl = 63 - __builtin_clzll(h);      // floor(log2 h)
        u32 level = logn - fl;                  // 0 = leaves
        u64 base_heap = n >> level;
        pos = (2 * n - ((2 * n) >> level)) + (h - base_heap);
    }
    out[q] = tree[pos * 5 + q % 5];
}

// ------------------------------------------------------------------ host side

int launch_hash_rows(const u64* tab, u64 rows, u64 cols, u64* dig, hipStream_t st) {
    if (rows == 0) return TVH_OK;
    if (cols > 0xFFFFFFFFull) return TVH_ERR_INVALID_ARG;
    u32 blocks = (u32)((rows + 255) / 256);
    hipLaunchKernelGGL(k_hash_rows, dim3(blocks), dim3(256), 0, st, HashArgs{tab, rows, (u32)cols, dig});
    return check_launch();
}

int launch_absorb_chunk(const u64* chunk, u64 rows, u64 ccols, u64* state, u64* pend, u32 npend,
                        bool first, bool final_chunk, u64* dig, hipStream_t st) {
    if (rows == 0) return TVH_OK;
    u32 blocks = (u32)((rows + 255) / 256);
    if (final_chunk)
        hipLaunchKernelGGL(k_absorb_chunk<true>, dim3(blocks), dim3(256), 0, st, chunk, rows,
                           (u32)ccols, state, pend, npend, (int)first, dig);
    else
        hipLaunchKernelGGL(k_absorb_chunk<false>, dim3(blocks), dim3(256), 0, st, chunk, rows,
                           (u32)ccols, state, pend, npend, (int)first, dig);
    return check_launch();
}

// Fiat-Shamir sponge on a device-resident state (16 canonical u64), following
// the reference's device transcript (fiat_shamir_kernel.cu: fs_absorb_device_gpu
// pads [data..., 1, 0...] to a multiple of 10 and overwrites the rate before
// each permutation; squeeze reads the rate, then permutes).  One lane runs the
// whole absorption in one launch; the other 63 lanes of the wave only help
// load the S-box table.
// mode 0 absorb, 1 squeeze, 2 init (zeros; a kernel rather than a memset so
// that the call can be captured in a graph).
__global__ void __launch_bounds__(64) k_sponge(u64* __restrict__ state,
                                               const u64* __restrict__ data, u64 len,
                                               u64* __restrict__ out, int mode) {
    if (mode == 2) {
        if (threadIdx.x < tip5::STATE) state[threadIdx.x] = 0;
        return;
    }
    const bool squeeze = mode == 1;
    __shared__ uint8_t lut[256];
    tip5::load_lut(lut);
    __syncthreads();
    if (threadIdx.x != 0) return;
    u64 s[tip5::STATE];
#pragma unroll
    for (int i = 0; i < tip5::STATE; i++) s[i] = tip5::to_m(state[i]);
    if (squeeze) {
#pragma unroll
        for (int i = 0; i < tip5::RATE; i++) out[i] = tip5::from_m(s[i]);
        tip5::permutation_m(s, lut);
    } else {
        u64 padded = (len + 1 + tip5::RATE - 1) / tip5::RATE * tip5::RATE;
        for (u64 base = 0; base < padded; base += tip5::RATE) {
#pragma unroll
            for (int i = 0; i < tip5::RATE; i++) {
                u64 pos = base + i;
                s[i] = pos < len ? tip5::to_m(data[pos]) : (pos == len ? tip5::kOneM : 0ull);
            }
            tip5::permutation_m(s, lut);
        }
    }
#pragma unroll
    for (int i = 0; i < tip5::STATE; i++) state[i] = tip5::from_m(s[i]);
}

int launch_sponge(u64* state, const u64* data, u64 len, u64* out, int mode, hipStream_t st) {
    hipLaunchKernelGGL(k_sponge, dim3(1), dim3(64), 0, st, state, data, len, out, mode);
    return check_launch();
}

// FRI codeword leaves (reference gpu_stark.cu:863-875): XFE i (3 limbs,
// interleaved) becomes the digest (c0, c1, c2, 0, 0) without hashing.
__global__ void k_xfe_leaves(const u64* __restrict__ xfe, u64 n, u64* __restrict__ leaves) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    u64 a = xfe[3 * i], b = xfe[3 * i + 1], c = xfe[3 * i + 2];
    leaves[5 * i + 0] = a;
    leaves[5 * i + 1] = b;
    leaves[5 * i + 2] = c;
    leaves[5 * i + 3] = 0;
    leaves[5 * i + 4] = 0;
}

int launch_xfe_leaves(const u64* xfe, u64 n, u64* leaves, hipStream_t st) {
    if (n == 0) return TVH_OK;
    hipLaunchKernelGGL(k_xfe_leaves, dim3((u32)((n + 255) / 256)), dim3(256), 0, st, xfe, n,
                       leaves);
    return check_launch();
}

int launch_permutation(u64* states, u64 num, hipStream_t st) {
    if (num == 0) return TVH_OK;
    hipLaunchKernelGGL(k_permutation, dim3((u32)((num + 255) / 256)), dim3(256), 0, st, states,
                       num);
    return check_launch();
}

int launch_permutation_rounds(u64* states, u64 num, u32 first, u32 count, u64* trace,
                              hipStream_t st) {
    if (num == 0) return TVH_OK;
    hipLaunchKernelGGL(k_permutation_rounds, dim3((u32)((num + 255) / 256)), dim3(256), 0, st,
                       states, num, first, count, trace);
    return check_launch();
}

int launch_hash_pairs(const u64* l, const u64* r, u64* out, u64 count, hipStream_t st) {
    if (count == 0) return TVH_OK;
    hipLaunchKernelGGL(k_hash_pairs, dim3((u32)((count + 255) / 256)), dim3(256), 0, st, l, r,
                       out, count);
    return check_launch();
}

u64 merkle_level_base(u64 n, u32 level, int layout) {
    return layout == TVH_TREE_HEAP ? (n >> level) : (2 * n - ((2 * n) >> level));
}

// Launch groups of the Merkle build of an n-leaf tree whose leaves are already
// at level 0: group g covers up to 9 levels; T active threads per block.
int merkle_groups(u64* tree, u64 n, int layout, std::vector<MerkleGroup>& groups) {
    groups.clear();
    u32 logn = log2u(n);
    u32 lvl = 0;
    while (lvl < logn) {
        u64 nchild = n >> lvl;
        u64 parents = nchild / 2;
        // Levels of more than 2^15 parents take a launch each with one hash
        // per thread (continuing inside a block would halve the active threads
        // per level).  Above them one hash on one lane is a latency of ~21 us
        // per level, so each hash runs on 16 lanes (permutation_group16, ~6 us):
        // a 256-thread block hashes up to 16 parents and, once all blocks fit
        // on the GPU at once (<= 2^14 parents), continues to their subtree root,
        // up to 5 levels per launch.  Measured at 2^22 leaves: levels 2^16..1
        // in 387 us with one lane per hash, 152 us in this form.
        const bool group16 = parents <= (1u << 15);
        u32 T = group16 ? (parents >= 16 ? 16 : (u32)parents)
                        : (parents >= 256 ? 256 : (u32)parents);  // parents is a power of two
        u32 maxlev = 0;
        while ((1u << maxlev) <= T) maxlev++;  // T = 16 -> 5 levels
        if (!group16 || parents > (1u << 14)) maxlev = 1;
        u32 levels = logn - lvl < maxlev ? logn - lvl : maxlev;
        MerkleGroup g;
        g.args.group16 = group16 ? 1u : 0u;
        g.args.src = tree + 5 * merkle_level_base(n, lvl, layout);
        for (u32 j = 0; j < 9; j++)
            g.args.dst[j] = j < levels ? tree + 5 * merkle_level_base(n, lvl + j + 1, layout) : nullptr;
        g.args.levels = levels;
        g.args.T = T;
        g.blocks = (u32)(parents / T);
        groups.push_back(g);
        lvl += levels;
    }
    return TVH_OK;
}

// Build levels 1..log2(n) of a tree whose leaves are already at level 0.
int launch_merkle_levels(u64* tree, u64 n, int layout, hipStream_t st) {
    std::vector<MerkleGroup> groups;
    merkle_groups(tree, n, layout, groups);
    for (const MerkleGroup& g : groups) {
        hipLaunchKernelGGL(k_merkle_subtrees, dim3(g.blocks), dim3(g.args.group16 ? 256 : g.args.T),
                           0, st, g.args);
        int rc = check_launch();
        if (rc) return rc;
    }
    return TVH_OK;
}

int launch_hash_rows_multi(const HashArgs* d_args, const u32* d_start, u32 nseg, u32 blocks,
                           hipStream_t st) {
    if (!blocks) return TVH_OK;
    hipLaunchKernelGGL(k_hash_rows_multi, dim3(blocks), dim3(256), 0, st, d_args, d_start, nseg);
    return check_launch();
}

int launch_merkle_multi(const MerkleArgs* d_args, const u32* d_start, u32 nseg, u32 blocks,
                        hipStream_t st) {
    if (!blocks) return TVH_OK;
    hipLaunchKernelGGL(k_merkle_subtrees_multi, dim3(blocks), dim3(256), 0, st, d_args, d_start,
                       nseg);
    return check_launch();
}

int launch_gather_heap(const u64* tree, u64 n, int layout, const u64* idx, u64 count, u64* out,
                       hipStream_t st) {
    if (count == 0) return TVH_OK;
    u64 total = count * 5;
    hipLaunchKernelGGL(k_gather_heap, dim3((u32)((total + 255) / 256)), dim3(256), 0, st, tree, n,
                       log2u(n), layout, idx, count, out);
    return check_launch();
}

}  // namespace tvh
