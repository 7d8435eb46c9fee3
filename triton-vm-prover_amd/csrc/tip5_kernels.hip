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
// One launch builds up to 9 levels: a block loads 2T children, hashes T
// parents, keeps them in LDS and continues to the block's subtree root.
#include "tvh_internal.hpp"

namespace tvh {

using gl::u32;
using gl::u64;

__device__ __forceinline__ void store_digest(u64* dst, const u64 s[tip5::STATE]) {
#pragma unroll
    for (int i = 0; i < tip5::DIGEST; i++) dst[i] = gl::canon(s[i]);
}

// rows x cols column-major table -> rows digests (5 u64 each, row-major).
__global__ void __launch_bounds__(256) k_hash_rows(const u64* __restrict__ tab, u64 rows,
                                                   u32 cols, u64* __restrict__ dig) {
    __shared__ uint8_t lut[256];
    tip5::load_lut(lut);
    __syncthreads();
    u64 row = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= rows) return;
    u64 s[tip5::STATE];
#pragma unroll
    for (int i = 0; i < tip5::STATE; i++) s[i] = 0;
    const u64* p = tab + row;
    u32 used = 0;
    for (; used + tip5::RATE <= cols; used += tip5::RATE) {
#pragma unroll
        for (int i = 0; i < tip5::RATE; i++) s[i] = __builtin_nontemporal_load(p + (u64)(used + i) * rows);
        tip5::permutation(s, lut);
    }
    u32 rest = cols - used;
#pragma unroll
    for (int i = 0; i < tip5::RATE; i++) {
        u32 ui = (u32)i;
        s[i] = ui < rest ? __builtin_nontemporal_load(p + (u64)(used + ui) * rows)
                         : (ui == rest ? 1ull : 0ull);
    }
    tip5::permutation(s, lut);
    store_digest(dig + row * tip5::DIGEST, s);
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
    for (int k = 0; k < tip5::STATE; k++) s[k] = st[i * tip5::STATE + k];
    tip5::permutation(s, lut);
#pragma unroll
    for (int k = 0; k < tip5::STATE; k++) st[i * tip5::STATE + k] = gl::canon(s[k]);
}

__device__ __forceinline__ void hash_pair(const u64* l, const u64* r, u64 s[tip5::STATE],
                                          const uint8_t* lut) {
#pragma unroll
    for (int k = 0; k < 5; k++) {
        s[k] = l[k];
        s[5 + k] = r[k];
    }
#pragma unroll
    for (int k = 10; k < tip5::STATE; k++) s[k] = 1;
    tip5::permutation(s, lut);
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

// One launch of the Merkle build: children are the nodes of level `lvl` (count
// `nchild`, starting at `src`); each block of T threads hashes 2T children into
// T parents and continues for `levels` levels inside LDS.  Level l+j nodes go
// to dst_base[j-1] + index.
struct MerkleLevels {
    u64* dst[9];
};

__global__ void __launch_bounds__(256) k_merkle_subtrees(const u64* __restrict__ src,
                                                         MerkleLevels out, u32 levels) {
    __shared__ uint8_t lut[256];
    __shared__ u64 buf[256 * 5];
    tip5::load_lut(lut);
    __syncthreads();
    const u32 T = blockDim.x;
    const u32 t = threadIdx.x;
    u64 s[tip5::STATE];
    // first level: children from global memory
    {
        u64 parent = (u64)blockIdx.x * T + t;
        hash_pair(src + 10 * parent, src + 10 * parent + 5, s, lut);
        u64 d[5];
#pragma unroll
        for (int k = 0; k < 5; k++) d[k] = gl::canon(s[k]);
#pragma unroll
        for (int k = 0; k < 5; k++) {
            out.dst[0][5 * parent + k] = d[k];
            buf[5 * t + k] = d[k];
        }
    }
    u32 width = T;
    for (u32 j = 1; j < levels; j++) {
        __syncthreads();
        width >>= 1;
        u64 d[5];
        if (t < width) {
            hash_pair(buf + 10 * t, buf + 10 * t + 5, s, lut);
#pragma unroll
            for (int k = 0; k < 5; k++) d[k] = gl::canon(s[k]);
        }
        __syncthreads();
        if (t < width) {
            u64 parent = (u64)blockIdx.x * width + t;
#pragma unroll
            for (int k = 0; k < 5; k++) {
                out.dst[j][5 * parent + k] = d[k];
                buf[5 * t + k] = d[k];
            }
        }
    }
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
    hipLaunchKernelGGL(k_hash_rows, dim3(blocks), dim3(256), 0, st, tab, rows, (u32)cols, dig);
    return check_launch();
}

int launch_permutation(u64* states, u64 num, hipStream_t st) {
    if (num == 0) return TVH_OK;
    hipLaunchKernelGGL(k_permutation, dim3((u32)((num + 255) / 256)), dim3(256), 0, st, states,
                       num);
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

// Build levels 1..log2(n) of a tree whose leaves are already at level 0.
int launch_merkle_levels(u64* tree, u64 n, int layout, hipStream_t st) {
    u32 logn = 0;
    while ((1ull << logn) < n) logn++;
    u32 lvl = 0;
    while (lvl < logn) {
        u64 nchild = n >> lvl;
        u64 parents = nchild / 2;
        u32 T = parents >= 256 ? 256 : (u32)parents;  // parents is a power of two
        u32 maxlev = 0;
        while ((1u << maxlev) <= T) maxlev++;           // T=256 -> 9 levels
        u32 levels = logn - lvl < maxlev ? logn - lvl : maxlev;
        MerkleLevels ml;
        for (u32 j = 0; j < 9; j++)
            ml.dst[j] = j < levels ? tree + 5 * merkle_level_base(n, lvl + j + 1, layout) : nullptr;
        const u64* src = tree + 5 * merkle_level_base(n, lvl, layout);
        hipLaunchKernelGGL(k_merkle_subtrees, dim3((u32)(parents / T)), dim3(T), 0, st, src, ml,
                           levels);
        int rc = check_launch();
        if (rc) return rc;
        lvl += levels;
    }
    return TVH_OK;
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
