// Kernels around the commit path: the two-phase coset evaluation (trace
// coefficients once, then one coset at a time), the coset digest scatter and
// the row-major <-> column-major table transposes.
//
// Two-phase evaluation (src/gpu/kernels/randomized_lde_kernel.cu:1505-1659):
//   phase 1: c_i = INTT_n(trace)_i * o^-i                      (o = trace offset)
//   phase 2: out_j = P(g w_n^j), P(X) = sum c_i X^i + (X^n - o^n) R(X), k <= n.
// On the coset g<w_n> the factor X^n is the constant g^n, so
//   P(g w_n^j) = NTT_n(b)_j,  b_i = (c_i + [i < k] r_i (g^n - o^n)) g^i
// and phase 2 is one scaling pass plus one forward NTT (the reference adds a
// second NTT for the randomizer tail instead).  The scaling pass reads g^i
// from two small per-call tables (g^i = g^(i mod 4096) * (g^4096)^(i / 4096)).
#include "tvh_internal.hpp"

namespace tvh {

namespace {

// t[i] = base^i in Montgomery form (base^i * 2^64 mod p), i < count: one
// square-and-multiply per entry (a table of at most 4096 entries per call).
__global__ void k_pow_table(u64* __restrict__ t, u64 base, u32 count) {
    u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) t[i] = gl::canon(gl::to_mont(gl::pow(base, i)));
}

constexpr int kTabTpb = 256;
constexpr int kTabPer = 4;  // elements per thread, kTabTpb apart (coalesced)

// y[c*n + i] = (x[c*n + i] + [i < k] rand[c*k + i] * delta) * base^i with
// base^i = tlo[i mod 4096] * thi[i / 4096] (Montgomery-form tables): two
// cached loads and two products per element.  x may equal y.
__global__ void __launch_bounds__(kTabTpb) k_scale_table(const u64* x, u64* y, u64 n, u32 log_n,
                                                         u64 total, const u64* __restrict__ tlo,
                                                         const u64* __restrict__ thi,
                                                         const u64* __restrict__ rand, u64 k,
                                                         u64 delta) {
    const u64 e0 = (u64)blockIdx.x * (kTabTpb * kTabPer) + threadIdx.x;
#pragma unroll
    for (int r = 0; r < kTabPer; r++) {
        u64 e = e0 + (u64)r * kTabTpb;
        if (e >= total) break;
        u64 i = e & (n - 1);
        u64 v = x[e];
        if (i < k) v = gl::add(v, gl::mul(rand[(e >> log_n) * k + i], delta));
        u64 f = gl::mont_mul(tlo[i & 4095], thi[i >> 12]);  // base^i in Montgomery form
        y[e] = gl::mont_mul(v, f);
    }
}

// leaves[(j * num_cosets + coset) * 5 + e] = coset_digests[j * 5 + e]
__global__ void k_scatter_digests(const u64* __restrict__ dig, u64* __restrict__ leaves, u64 len,
                                  u64 coset, u64 num_cosets) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len * 5) return;
    u64 j = i / 5, e = i - j * 5;
    leaves[(j * num_cosets + coset) * 5 + e] = dig[i];
}

// dst[c * rows + r] = src[r * cols + c]: 64 x 64 tiles through LDS, both
// sides coalesced (the column-major side writes 64 consecutive rows of one
// column per wave).
constexpr int kTile = 64;
constexpr int kTileRows = 16;  // threads per tile column step: 64 x 16 = 1024 threads
__global__ void __launch_bounds__(kTile * kTileRows) k_transpose(const u64* __restrict__ src,
                                                                 u64* __restrict__ dst, u64 rows,
                                                                 u64 cols, u64 row_tiles) {
    __shared__ u64 tile[kTile][kTile + 1];
    // one-dimensional grid of tiles, row tiles fastest
    const u64 tr = blockIdx.x % row_tiles, tc = blockIdx.x / row_tiles;
    const u64 r0 = tr * kTile, c0 = tc * kTile;
    const int tx = threadIdx.x & (kTile - 1), ty = threadIdx.x / kTile;
#pragma unroll
    for (int k = 0; k < kTile; k += kTileRows) {
        u64 r = r0 + ty + k, c = c0 + tx;
        if (r < rows && c < cols) tile[ty + k][tx] = src[r * cols + c];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kTile; k += kTileRows) {
        u64 c = c0 + ty + k, r = r0 + tx;
        if (r < rows && c < cols) dst[c * rows + r] = tile[tx][ty + k];
    }
}

// ---- seeded field tables in row-major draw order ----
// Draw i + 1 of the SplitMix64 stream with state `seed` (the finaliser of
// seed + (i + 1) * gamma).
__device__ __forceinline__ u64 splitmix_draw(u64 seed, u64 i) {
    u64 z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

constexpr u32 kMaxRejects = 64;

__global__ void k_rejects_clear(u64* rej) {
    if (threadIdx.x == 0) rej[0] = 0;
}

// rej[0] = number of draws >= p among the first `draws`, rej[1..] their
// indices (unordered; the first kMaxRejects kept)
__global__ void k_rejects_find(u64 seed, u64 draws, u64* rej) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= draws) return;
    if (splitmix_draw(seed, i) >= gl::P) {
        u64 slot = atomicAdd((unsigned long long*)rej, 1ull);
        if (slot < kMaxRejects) rej[1 + slot] = i;
    }
}

// d[c*rows + r] = accepted draw number r*cols + c: the draw index k of
// accepted draw e is the fixed point of k = e + #{rejected indices <= k}
__global__ void k_fill_table(u64* __restrict__ d, u64 rows, u64 cols, u64 seed,
                             const u64* __restrict__ rej) {
    u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= rows * cols) return;
    u64 c = t / rows, r = t - c * rows;  // adjacent threads: adjacent rows (coalesced store)
    u64 e = r * cols + c;
    u32 nr = (u32)(rej[0] < kMaxRejects ? rej[0] : kMaxRejects);
    u64 k = e;
    for (u32 it = 0; it <= nr; it++) {
        u64 below = 0;
        for (u32 j = 0; j < nr; j++) below += rej[1 + j] <= k;
        if (e + below == k) break;
        k = e + below;
    }
    d[t] = splitmix_draw(seed, k);
}

}  // namespace

int launch_fill_table(u64* d, u64 rows, u64 cols, u64 seed, u64* rej, hipStream_t st) {
    u64 total = rows * cols;
    if (!total) return TVH_OK;
    u64 draws = total + kMaxRejects;
    hipLaunchKernelGGL(k_rejects_clear, dim3(1), dim3(64), 0, st, rej);
    hipLaunchKernelGGL(k_rejects_find, dim3((u32)((draws + 255) / 256)), dim3(256), 0, st, seed, draws,
                       rej);
    hipLaunchKernelGGL(k_fill_table, dim3((u32)((total + 255) / 256)), dim3(256), 0, st, d, rows, cols,
                       seed, rej);
    return check_launch();
}

u64 fill_table_scratch_elems() { return 1 + kMaxRejects; }

u64 scale_table_elems(u64 n) {
    u64 e = 4096 + (n + 4095) / 4096;  // the two-level table of k_scale_table
    return e > kNttScaleTableElems ? e : kNttScaleTableElems;  // or the NTT's scaling tables
}

int launch_pow_table(u64* t, u64 base, u64 count, hipStream_t st) {
    if (!count) return TVH_OK;
    if (count > (1ull << 24)) return set_error(TVH_ERR_UNSUPPORTED, "power table of %llu",
                                               (unsigned long long)count);
    hipLaunchKernelGGL(k_pow_table, dim3((u32)((count + 255) / 256)), dim3(256), 0, st, t, base,
                       (u32)count);
    return check_launch();
}

int launch_scale_table(const u64* x, u64* y, u64 n, u64 cols, u64 base, const u64* rand, u64 k,
                       u64 delta, u64* tables, hipStream_t st) {
    if (!n || !cols) return TVH_OK;
    if (n > (1ull << 24)) return set_error(TVH_ERR_UNSUPPORTED, "scale table for n > 2^24");
    u64* tlo = tables;
    u64* thi = tables + 4096;
    int rc;
    if ((rc = launch_pow_table(tlo, base, n < 4096 ? n : 4096, st)) ||
        (rc = launch_pow_table(thi, gl::pow(base, 4096), (n + 4095) / 4096, st)))
        return rc;
    const u64 total = n * cols;
    const u64 blocks = (total + kTabTpb * kTabPer - 1) / (kTabTpb * kTabPer);
    if (blocks > 0x7FFFFFFFull) return set_error(TVH_ERR_UNSUPPORTED, "scale of %llu elements",
                                                 (unsigned long long)total);
    hipLaunchKernelGGL(k_scale_table, dim3((u32)blocks), dim3(kTabTpb), 0, st, x, y, n, log2u(n), total,
                       tlo, thi, rand, k, delta);
    return check_launch();
}

int launch_scatter_digests(const u64* dig, u64* leaves, u64 len, u64 coset, u64 num_cosets,
                           hipStream_t st) {
    if (!len) return TVH_OK;
    hipLaunchKernelGGL(k_scatter_digests, dim3((u32)((len * 5 + 255) / 256)), dim3(256), 0, st, dig,
                       leaves, len, coset, num_cosets);
    return check_launch();
}

int launch_transpose(const u64* src, u64* dst, u64 rows, u64 cols, hipStream_t st) {
    if (!rows || !cols) return TVH_OK;
    u64 rt = (rows + kTile - 1) / kTile, ct = (cols + kTile - 1) / kTile;
    if (rt * ct > 0x7FFFFFFFull)
        return set_error(TVH_ERR_UNSUPPORTED, "transpose of %llu x %llu", (unsigned long long)rows,
                         (unsigned long long)cols);
    hipLaunchKernelGGL(k_transpose, dim3((u32)(rt * ct)), dim3(kTile * kTileRows), 0, st, src, dst,
                       rows, cols, rt);
    return check_launch();
}

}  // namespace tvh
