// Host side of the batched Goldilocks NTT and the coset low-degree extension:
// kernel variants and plans, the per-shape coset tables, launches, the
// synthetic fill and the XFE relayout.  The kernels themselves are in
// ntt_dev.hpp (device code) and ntt_step1.hip, lde_rows.hip, lde_expand.hip,
// lde_step3.hip (instantiations and launchers).
#include <atomic>
#include <map>
#include <mutex>
#include <tuple>
#include <utility>

#include "ntt_dev.hpp"

namespace tvh {

// Constant tables pa/pb of one (logN, a, E, o, g).
__global__ void k_lde_consts(u64* pa, u64* pb, u32 a, u32 b, u32 E, const u64* beta, u64 ninv) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 n1 = 1ull << a, n2 = 1ull << b;
    // stored in Montgomery form (x * 2^64 mod p) for gl::mont_mul
    if (i < E * n1) {
        u64 s = i >> a, k1 = i & (n1 - 1);
        pa[i] = gl::canon(gl::to_mont(gl::mul(ninv, gl::pow(beta[s], k1))));
    } else if (i < E * (n1 + n2)) {
        u64 j = i - E * n1;
        u64 s = j >> b, k2 = j & (n2 - 1);
        pb[j] = gl::canon(gl::to_mont(gl::pow(gl::pow(beta[s], n1), k2)));
    }
}

// Synthetic uniform inputs: counter-based SplitMix64 finaliser of seed + (i+1)*gamma.
__global__ void k_fill_uniform(u64* d, u64 count, u64 seed) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    u64 z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    d[i] = z >= gl::P ? z - gl::P : z;
}

int launch_fill_uniform(u64* d, u64 count, u64 seed, hipStream_t st) {
    if (!count) return TVH_OK;
    hipLaunchKernelGGL(k_fill_uniform, dim3((u32)((count + 255) / 256)), dim3(256), 0, st, d,
                       count, seed);
    return check_launch();
}

// XFE interleaved <-> limb-major.  Interleaved: (c*len + r)*3 + k; limb-major: (3c + k)*len + r.
__global__ void k_xfe_relayout(const u64* __restrict__ src, u64* __restrict__ dst, u64 cols,
                               u64 len, int to_limb_major) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cols * len * 3) return;
    // i indexes the interleaved position
    u64 k = i % 3, cr = i / 3, r = cr % len, c = cr / len;
    u64 lm = (3 * c + k) * len + r;
    if (to_limb_major) dst[lm] = src[i];
    else dst[i] = src[lm];
}

// ---------------------------------------------------------------- host side

// Tuned variant per (log2 N, log2 E, width class); 0 = not set.
static std::atomic<int> g_variant[kMaxLogN + 1][5][kWidthClasses];

int lde_variant(u32 logN, u32 logE, u64 cols) {
    int v = (logN <= kMaxLogN && logE <= 4) ? g_variant[logN][logE][width_class(cols)].load() : 0;
    if (v) return v;
    // default (measured on MI355X at widths 156..379, 2^12..2^21 rows): the split
    // form with 2^12-element row tiles for 2^13..2^16 rows, 2^13-element row
    // tiles above (column tiles 2^12, 2^13 from 2^19 rows), 2^13 for the inverse
    // rows only at 2^12 rows, the single fused kernel below
    if (logN >= 19) return LDE_SPLIT | (1 << 2) | LDE_ROW_HALF;
    if (logN >= 17) return LDE_SPLIT | (2 << 2) | LDE_ROW_HALF;
    if (logN >= 13) return LDE_SPLIT | (2 << 2) | LDE_ROW_QUARTER;
    if (logN == 12) return LDE_SPLIT | (1 << 2) | LDE_INV_HALF;
    return LDE_WIDE | (2 << 2);
}

bool lde_variant_valid(int v) {
    int form = v & 3, tile = (v >> 2) & 3, half = (v >> 4) & 1, ihalf = (v >> 5) & 1,
        quarter = (v >> 6) & 1;
    return v == LDE_AUTO || (v < 128 && (form == LDE_WIDE || form == LDE_SPLIT) && tile <= 2 &&
                             half + ihalf + quarter <= (form == LDE_SPLIT));
}

int set_lde_variant(u32 logN, u32 logE, int v, int wc) {
    if (logN > kMaxLogN || logE > 4 || !lde_variant_valid(v) || wc >= kWidthClasses)
        return set_error(TVH_ERR_INVALID_ARG, "lde variant (%u, %u, %d)", logN, logE, v);
    for (int c = 0; c < kWidthClasses; c++)
        if (wc < 0 || c == wc) g_variant[logN][logE][c].store(v);
    return TVH_OK;
}

static u32 choose_split(u32 logN, u32 E, u32 row_lt) {
    // N2 (the contiguous row step) as long as one row tile holds all E cosets
    // of a row (E * N2 <= 2^row_lt, N2 <= 4096: the expansion overwrites its
    // input row in place, so a row's cosets cannot be split over blocks);
    // N1 = N / N2 takes the strided step, whose HBM segments are (column
    // tile) / N1 elements wide.
    u32 logE = log2u(E);
    u32 bmax = row_lt - logE;
    if (bmax > 12) bmax = 12;
    u32 b = logN < bmax ? logN : bmax;
    return logN - b;
}

static u32 tile_log_nt(u32 logL, u32 lt = LOG_TILE) {
    // transforms per block of a 2^lt tile: 2^lt / max(L, 32)
    u32 l = logL < 5 ? 5 : logL;
    return lt - l;
}

static size_t lds_bytes(u32 logL, u32 logNT) {
    return ((size_t)padded_len(1 << logL) << logNT) * sizeof(u64);
}

struct ConstKey {
    u32 logN, a, E;
    u64 toff, goff;
    int dev;
    bool operator<(const ConstKey& o) const {
        return std::tie(logN, a, E, toff, goff, dev) <
               std::tie(o.logN, o.a, o.E, o.toff, o.goff, o.dev);
    }
};
struct ConstVal {
    u64* pa;
    u64* pb;
};
static std::mutex g_const_mu;
static std::map<ConstKey, ConstVal> g_consts;
// At most this many tables (each <= E*(N1+N2)*8 <= 1 MB) are kept; the cache
// never evicts, since kernels in flight may read any entry.
constexpr size_t kMaxConstEntries = 256;

// The coset tables of one (shape, split, offsets) key: found in the cache, or
// computed now (one allocation and a wait on a private stream; tvh_lde_prepare
// does this ahead of time).
static int lde_consts(DeviceState* ds, u32 logN, u32 a, u32 E, u64 toff, u64 goff, ConstVal* out) {
    ConstKey key{logN, a, E, toff, goff, ds->device};
    std::lock_guard<std::mutex> lk(g_const_mu);
    auto it = g_consts.find(key);
    if (it != g_consts.end()) {
        *out = it->second;
        return TVH_OK;
    }
    if (g_consts.size() >= kMaxConstEntries)
        return set_error(TVH_ERR_RANGE,
                         "more than %zu distinct LDE shape/offset tables; release them with "
                         "tvh_lde_release_constants",
                         kMaxConstEntries);
    u32 b = logN - a;
    u64 n1 = 1ull << a, n2 = 1ull << b;
    u64 M = (u64)E << logN;
    u64 wM = gl::root_of_unity(log2u(M));
    u64 oinv = gl::inv(toff);
    u64 beta[16];
    u64 ws = 1;
    for (u32 s = 0; s < E; s++) {
        beta[s] = gl::mul(gl::mul(goff, ws), oinv);
        ws = gl::mul(ws, wM);
    }
    u64 ninv = gl::inv(1ull << logN);
    u64* mem = nullptr;
    size_t bytes = (E * (n1 + n2) + 16) * sizeof(u64);
    if (hipMalloc(&mem, bytes) != hipSuccess) return set_error(TVH_ERR_OOM, "lde constants");
    u64* dbeta = mem + E * (n1 + n2);
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
        return set_error(TVH_ERR_DEVICE, "stream create");
    int rc = TVH_OK;
    if (hipMemcpyAsync(dbeta, beta, sizeof(beta), hipMemcpyHostToDevice, st) != hipSuccess)
        rc = set_error(TVH_ERR_DEVICE, "copy beta");
    if (!rc) {
        u64 total = E * (n1 + n2);
        hipLaunchKernelGGL(k_lde_consts, dim3((u32)((total + 255) / 256)), dim3(256), 0, st, mem,
                           mem + E * n1, a, b, E, dbeta, ninv);
        rc = check_launch();
    }
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = set_error(TVH_ERR_DEVICE, "sync");
    (void)hipStreamDestroy(st);
    if (rc) {
        (void)hipFree(mem);
        return rc;
    }
    ConstVal v{mem, mem + E * n1};
    g_consts[key] = v;
    *out = v;
    return TVH_OK;
}

int lde_prepare_consts(u64 n, u64 m, u64 toff, u64 goff) {
    if (!is_pow2(n) || !is_pow2(m)) return set_error(TVH_ERR_NOT_POW2, "lde lengths");
    if (m < n) return set_error(TVH_ERR_INVALID_ARG, "target_len < trace_len");
    if (m / n > 16) return set_error(TVH_ERR_UNSUPPORTED, "expansion > 16");
    if (toff == 0 || toff >= gl::P || goff >= gl::P)
        return set_error(TVH_ERR_INVALID_ARG, "offsets must be canonical, trace offset nonzero");
    u32 logN = log2u(n);
    if (logN > 23) return set_error(TVH_ERR_UNSUPPORTED, "trace_len > 2^23");
    DeviceState* ds = current_device();
    if (!ds) return TVH_ERR_DEVICE;
    u32 E = (u32)(m / n);
    // every row tile a kernel variant can use (2^12..2^14 elements) gives its own split
    for (u32 lt = LOG_TILE - 2; lt <= LOG_TILE; lt++) {
        u32 a = choose_split(logN, E, lt);
        if (a > 12) continue;
        ConstVal cv;
        int rc = lde_consts(ds, logN, a, E, toff, goff, &cv);
        if (rc) return rc;
    }
    return TVH_OK;
}

int lde_release_consts() {
    std::lock_guard<std::mutex> lk(g_const_mu);
    if (hipDeviceSynchronize() != hipSuccess) return set_error(TVH_ERR_DEVICE, "device sync");
    for (auto& kv : g_consts) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        (void)hipSetDevice(kv.first.dev);
        (void)hipFree(kv.second.pa);
        (void)hipSetDevice(dev);
    }
    g_consts.clear();
    return TVH_OK;
}

int prepare_lde(const u64* trace, u64 cols, u64 n, const RandSpec& rs, u64 toff, u64 goff,
                u64 m, bool truncate, u64* out, LdePlan* P, bool batch) {
    P->k1 = P->k3 = P->split = false;
    P->k2_variant = lde_variant(log2u(n), log2u(m / (n ? n : 1)), cols);
    // multi-table launches run the column and row kernels with one tile each
    // (kBatchColLt, kBatchRowLt) for every table of the launch
    P->col_lt = batch ? kBatchColLt : LOG_TILE - ((P->k2_variant >> 2) & 3);
    P->row_lt = batch                                  ? kBatchRowLt
                : (P->k2_variant & LDE_ROW_QUARTER) ? LOG_TILE - 2
                : (P->k2_variant & LDE_ROW_HALF)    ? LOG_TILE - 1
                                                     : LOG_TILE;
    P->inv_lt = P->row_lt < LOG_TILE                 ? P->row_lt
                : (P->k2_variant & LDE_INV_HALF)    ? LOG_TILE - 1
                                                     : LOG_TILE;
    if ((P->k2_variant & 3) != LDE_SPLIT) P->row_lt = LOG_TILE;  // the wide kernel's tile
    P->s1_blocks = P->s2_blocks = P->s3_blocks = 0;
    if (cols == 0) return TVH_OK;
    if (!is_pow2(n) || !is_pow2(m)) return set_error(TVH_ERR_NOT_POW2, "lde lengths");
    if (m < n) return set_error(TVH_ERR_INVALID_ARG, "target_len < trace_len");
    u64 E64 = m / n;
    if (E64 > 16) return set_error(TVH_ERR_UNSUPPORTED, "expansion > 16");
    if (rs.k > n) return set_error(TVH_ERR_UNSUPPORTED, "rand_len > trace_len");
    if (toff == 0 || toff >= gl::P || goff >= gl::P)
        return set_error(TVH_ERR_INVALID_ARG, "offsets must be canonical, trace offset nonzero");
    u32 logN = log2u(n), E = (u32)E64, logE = log2u(E64);
    if (logN > 23) return set_error(TVH_ERR_UNSUPPORTED, "trace_len > 2^23");
    DeviceState* ds = current_device();
    if (!ds) return TVH_ERR_DEVICE;
    u32 a = choose_split(logN, E, P->row_lt);
    u32 b = logN - a;
    if (a > 12) return set_error(TVH_ERR_UNSUPPORTED, "split");
    ConstVal cv;
    int rc = lde_consts(ds, logN, a, E, toff, goff, &cv);
    if (rc) return rc;

    // K1: INTT step 1 (transforms (c, n2)); a size-1 step is the identity
    // map trace -> Y, so K2 then reads the trace itself.
    P->k1 = a > 0;
    if (P->k1) {
        u32 logNT = tile_log_nt(a, P->col_lt);
        u64 ntr = cols << b;
        P->s1 = Step1Args{trace, n, out, m, logN, a, E, ntr, logNT};
        P->s1_blocks = (u32)((ntr + (1ull << logNT) - 1) >> logNT);
        P->s1_lds = (size_t)fused_pitch(1 << a, 1 << logNT, true) * (1u << logNT) * sizeof(u64);
    }
    // K2: INTT step 2 + coset scaling + forward step 1
    {
        LdeArgs& A = P->s2;
        A.out = out;
        A.logN = logN;
        A.a = a;
        A.E = E;
        A.logE = logE;
        A.cols = cols;
        A.M = m;
        A.pa = cv.pa;
        A.pb = cv.pb;
        A.toff = toff;
        A.rs = rs;
        u64 gN = gl::pow(goff, n), oN = gl::pow(toff, n);
        u64 wE = gl::root_of_unity(logE), ws = 1;
        u64 nmod = n % gl::P;
        for (u32 s = 0; s < 16; s++) {
            u64 d = (truncate && E == 1) ? gl::neg(oN) : gl::sub(gl::mul(gN, ws), oN);
            A.delta_n[s] = gl::mul(d, nmod);
            ws = gl::mul(ws, wE);
        }
        u32 logNT = P->row_lt - (b + logE);  // NT * E * N2 = 2^row_lt
        if (b < 5) logNT = P->row_lt - 5 - logE;
        u64 nrows = cols << a;
        if ((1ull << logNT) > nrows) logNT = log2u(nrows);
        A.logNT = logNT;
        A.src = P->k1 ? out : trace;
        A.src_col_stride = P->k1 ? m : n;
        A.src_row_stride = P->k1 ? ((u64)E << b) : 0;
        P->s2_blocks = (u32)((nrows + (1ull << logNT) - 1) >> logNT);
        P->s2_lds = lds_bytes(b, logNT + logE);
        if ((P->k2_variant & 3) == LDE_SPLIT)  // k_lde_expand: transform-fastest pitch
            P->s2_lds = (size_t)fused_pitch(1 << b, (int)(E << logNT), true) * (E << logNT) * sizeof(u64);
        P->split = (P->k2_variant & 3) == LDE_SPLIT;
        if (P->split) {  // inverse rows first (TILE / N2 rows per block), then expansion from out
            u32 li = tile_log_nt(b, P->inv_lt);
            if ((1ull << li) > nrows) li = log2u(nrows);
            P->inv_logNT = li;
            P->inv_blocks = (u32)((nrows + (1ull << li) - 1) >> li);
            P->inv_lds = lds_bytes(b, li);
            P->s2inv = A;
            P->s2inv.logNT = li;
            A.src = out;
            A.src_col_stride = m;
            A.src_row_stride = (u64)E << b;
        }
    }
    // K3: forward step 2 for all cosets (identity when N1 = 1)
    P->k3 = a > 0;
    if (P->k3) {
        u32 logNT = tile_log_nt(a, P->col_lt);
        u64 ntr = cols << (b + logE);
        P->s3 = Step3Args{out, logN, a, logE, cols, m, logNT};
        P->s3_blocks = (u32)((ntr + (1ull << logNT) - 1) >> logNT);
        P->s3_lds = (size_t)fused_pitch(1 << a, 1 << logNT, true) * (1u << logNT) * sizeof(u64);
    }
    return TVH_OK;
}

int launch_lde(const u64* trace, u64 cols, u64 n, const RandSpec& rs, u64 toff, u64 goff, u64 m,
               bool truncate, u64* out, hipStream_t st) {
    LdePlan P;
    int rc = prepare_lde(trace, cols, n, rs, toff, goff, m, truncate, out, &P, false);
    if (rc || cols == 0) return rc;
    DeviceState* ds = current_device();
    if (P.k1 && (rc = launch_k_step1(true, P.col_lt, P.s1_blocks, P.s1_lds, st, P.s1, ds->roots)))
        return rc;
    if (P.split) {
        if ((rc = launch_k_lde_rowinv(P.inv_lt, P.inv_blocks, P.inv_lds, st, P.s2inv, ds->roots)))
            return rc;
        if ((rc = launch_k_lde_expand(P.row_lt, P.s2_blocks, P.s2_lds, st, P.s2, ds->roots))) return rc;
    } else if ((rc = launch_k_lde_step2(P.s2_blocks, P.s2_lds, st, P.s2, ds->roots))) {
        return rc;
    }
    if (P.k3 && (rc = launch_k_lde_step3(P.col_lt, P.s3_blocks, P.s3_lds, st, P.s3, ds->roots)))
        return rc;
    return TVH_OK;
}

int launch_lde_multi(int phase, const void* d_args, const u32* d_start, u32 nseg, u32 blocks,
                     size_t lds, hipStream_t st) {
    if (blocks == 0) return TVH_OK;
    DeviceState* ds = current_device();
    if (!ds) return TVH_ERR_DEVICE;
    switch (phase) {
        case LDE_PHASE_STEP1:
            return launch_k_step1_multi((const Step1Args*)d_args, d_start, nseg, blocks, lds, st,
                                        ds->roots);
        case LDE_PHASE_ROWINV:
            return launch_k_lde_rowinv_multi((const LdeArgs*)d_args, d_start, nseg, blocks, lds, st,
                                             ds->roots);
        case LDE_PHASE_STEP2:
            return launch_k_lde_step2_multi((const LdeArgs*)d_args, d_start, nseg, blocks, lds, st,
                                            ds->roots);
        case LDE_PHASE_EXPAND:
            return launch_k_lde_expand_multi((const LdeArgs*)d_args, d_start, nseg, blocks, lds, st,
                                             ds->roots);
        default:
            return launch_k_lde_step3_multi((const Step3Args*)d_args, d_start, nseg, blocks, lds, st,
                                            ds->roots);
    }
}

// Scratch of the allocation-free NTT entry point (tvh_ntt_batched_noalloc_dev):
// one file-scope device array per GPU, held until the process ends.  Host
// code never passes `scratch` itself to a kernel (in host code the name is a
// host-side stand-in, not a device address): noalloc_scratch_setup resolves
// this device's address once, at set-up, and launches take that pointer.
static __device__ __align__(256) u64 scratch[kNoAllocScratchElems];

int noalloc_scratch_setup(DeviceState* ds) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return set_error(TVH_ERR_DEVICE, "no-allocation scratch address: %s", hipGetErrorString(e));
    ds->noalloc_scratch = (u64*)p;
    return TVH_OK;
}

u64 ntt_scratch_elems(u64 n, u64 cols) {
    // columns go through the scratch in groups of at most 2 GiB.  Every group
    // is one launch of each step, and each launch pays about one round of
    // blocks to fill and drain the GPU (~35 us at 2^20 rows, a tenth of a
    // 32-column launch), so the groups are made large; the cap bounds the
    // scratch (a 2^20 x 379 table takes two groups).
    u64 group = cols;
    u64 max_group = (2048ull << 20) / (n * sizeof(u64));
    if (max_group < 1) max_group = 1;
    if (group > max_group) group = max_group;
    return group * n;
}

int launch_ntt_batched(const u64* src, u64* dst, u64 n, u64 cols, bool inverse, u64* scratch,
                       u64 scratch_elems, hipStream_t st, const NttScale* scale) {
    if (cols == 0) return TVH_OK;
    if (scale && inverse) return set_error(TVH_ERR_INVALID_ARG, "input scaling of an inverse NTT");
    if (!is_pow2(n)) return set_error(TVH_ERR_NOT_POW2, "NTT size must be a power of 2");
    u32 logN = log2u(n);
    if (logN > kMaxLogN) return set_error(TVH_ERR_UNSUPPORTED, "NTT size > 2^24");
    DeviceState* ds = current_device();
    if (!ds) return TVH_ERR_DEVICE;
    // Split N = N1 * N2 and column tile, measured on MI355X: for 2^18..2^22 rows
    // the column step takes N1 = 2^(logN/2 - 1) points in 2^13-element tiles
    // (two 512-thread blocks per CU), 3-13 % faster than N1 = 2^(logN/2) in
    // 2^14 tiles (2^20 x 379 forward: 5.43 against 5.80 ms).  Above 2^22 the
    // row length is capped at 2^12, and 2^13 tiles would hold only 4 or 2
    // columns' transforms (32 or 16 B segments), so the 2^14 tiles stay.
    const bool big = logN >= 18 && logN <= 22;
    u32 a = logN / 2 - (big ? 1 : 0), b = logN - a;
    if (b > 12) {
        b = 12;
        a = logN - b;
    }
    const u32 col_lt = big ? LOG_TILE - 1 : LOG_TILE;
    NttScaleDev S{};
    if (scale) {  // p1[n1] = base^(N2 n1), p2[n2] = base^n2
        if (!scale->tables) return set_error(TVH_ERR_INVALID_ARG, "no tables for the NTT scaling");
        S = NttScaleDev{scale->tables, scale->tables + (1ull << a), scale->rand, scale->k,
                        scale->delta};
        int rc;
        if ((rc = launch_pow_table(scale->tables, gl::pow(scale->base, 1ull << b), 1ull << a, st)) ||
            (rc = launch_pow_table(scale->tables + (1ull << a), scale->base, 1ull << b, st)))
            return rc;
    }
    // The row step stores each block's rows transposed, in runs as long as its
    // row count: 2^13 tiles (512 threads) win while they still hold 8 rows
    // (N2 <= 2^10: 2^18 x 379 inverse 1.09 against 1.26 ms) and lose when
    // they hold fewer (2^22 x 64, N2 = 2^12: 6.3 against 4.0 ms).
    const u32 row_lt = (big && b <= 10) ? LOG_TILE - 1 : LOG_TILE;
    // step 1 writes each group of columns to the scratch, step 2 writes it back
    u64 group = scratch_elems / n;
    if (!scratch || group == 0)
        return set_error(TVH_ERR_INVALID_ARG, "NTT scratch of %llu elements < n = %llu",
                         (unsigned long long)scratch_elems, (unsigned long long)n);
    if (group > cols) group = cols;
    u64 ninv_m = gl::mul(gl::inv(n % gl::P), gl::EPS);  // N^-1 in Montgomery form
    int rc;
    for (u64 c0 = 0; c0 < cols; c0 += group) {
        u64 g = cols - c0 < group ? cols - c0 : group;
        const u64* sg = src + c0 * n;  // step 1 reads the group from src
        u64* d = dst + c0 * n;         // step 2 writes it to dst
        {
            u32 logNT = tile_log_nt(a, col_lt);
            u64 ntr = g << b;
            u64 blocks = (ntr + (1ull << logNT) - 1) >> logNT;
            Step1Args A{sg, n, scratch, n, logN, a, 1u, ntr, logNT};
            const size_t s1_lds =
                (size_t)fused_pitch(1 << a, 1 << logNT, true) * (1u << logNT) * sizeof(u64);
            if (scale) {
                NttScaleDev Sg = S;
                if (Sg.rand) Sg.rand += c0 * Sg.k;  // the group's first column
                rc = launch_k_step1_scaled(col_lt, (u32)blocks, s1_lds, st, A, Sg, ds->roots);
            } else {
                rc = launch_k_step1(inverse, col_lt, (u32)blocks, s1_lds, st, A, ds->roots);
            }
            if (rc) return rc;
        }
        {
            u32 logNT = tile_log_nt(b, row_lt);
            u64 nrows = g << a;
            if ((1ull << logNT) > nrows) logNT = log2u(nrows);
            u64 blocks = (nrows + (1ull << logNT) - 1) >> logNT;
            const size_t s2_lds =
                (size_t)fused_pitch(1 << b, 1 << logNT, true) * (1u << logNT) * sizeof(u64);
            if ((rc = launch_k_ntt_step2(inverse, row_lt, (u32)blocks, s2_lds, st, scratch, d,
                                         logN, a, nrows, logNT, ninv_m, ds->roots)))
                return rc;
        }
    }
    return TVH_OK;
}

int ntt_set_attributes() {
    const int bytes = 150 * 1024;
    int rc;
    if ((rc = set_attributes_step1(bytes)) || (rc = set_attributes_rows(bytes)) ||
        (rc = set_attributes_expand(bytes)) || (rc = set_attributes_step3(bytes)))
        return rc;
    return TVH_OK;
}

int launch_xfe_interleave(const u64* src, u64* dst, u64 cols, u64 len, bool to_limb_major,
                          hipStream_t st) {
    u64 total = cols * len * 3;
    if (!total) return TVH_OK;
    hipLaunchKernelGGL(k_xfe_relayout, dim3((u32)((total + 255) / 256)), dim3(256), 0, st, src,
                       dst, cols, len, (int)to_limb_major);
    return check_launch();
}

}  // namespace tvh
