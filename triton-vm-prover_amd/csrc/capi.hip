// extern "C" entry points of libtvh (declared in include/tvh.h).
//
// Conventions follow the replacement contract for the reference's GPU API:
// status codes instead of exceptions (cuda_common.cuh:18-36 throws), no
// allocation or synchronisation inside `_dev` calls except the documented
// per-device arena, thread-safe one-time device set-up (the reference's
// twiddle cache is an unsynchronised static, ntt_kernel.cu:34-37,152-174).
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "tvh_internal.hpp"

namespace tvh {

static thread_local char g_err[512] = "";

int set_error(int status, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return status;
}

int check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(TVH_ERR_DEVICE, "launch: %s", hipGetErrorString(e));
    return TVH_OK;
}

static std::mutex g_dev_mu;
static DeviceState g_devs[64];

static int build_tables(DeviceState& ds) {
    // 6 tables of 4096 entries: r4096 fwd/inv, lo fwd/inv, hi fwd/inv
    std::vector<u64> h(6 * 4096);
    u64 w24 = gl::root_of_unity(kMaxLogN);
    u64 w24i = gl::inv(w24);
    u64 w4096 = gl::root_of_unity(12), w4096i = gl::inv(w4096);
    u64 w24_4096 = gl::pow(w24, 4096), w24i_4096 = gl::pow(w24i, 4096);
    u64 a = 1, ai = 1, b = 1, bi = 1, c = 1, ci = 1;
    // stored in Montgomery form (w * 2^64 mod p): the kernels apply them with gl::mont_mul
    auto m = [](u64 x) { return gl::canon(gl::to_mont(x)); };
    for (int e = 0; e < 4096; e++) {
        h[0 * 4096 + e] = m(a);
        h[1 * 4096 + e] = m(ai);
        h[2 * 4096 + e] = m(b);
        h[3 * 4096 + e] = m(bi);
        h[4 * 4096 + e] = m(c);
        h[5 * 4096 + e] = m(ci);
        a = gl::mul(a, w4096);
        ai = gl::mul(ai, w4096i);
        b = gl::mul(b, w24);
        bi = gl::mul(bi, w24i);
        c = gl::mul(c, w24_4096);
        ci = gl::mul(ci, w24i_4096);
    }
    u64* d = nullptr;
    if (hipMalloc(&d, h.size() * sizeof(u64)) != hipSuccess)
        return set_error(TVH_ERR_OOM, "twiddle tables");
    if (hipMemcpy(d, h.data(), h.size() * sizeof(u64), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        return set_error(TVH_ERR_DEVICE, "twiddle upload");
    }
    ds.table_mem = d;
    for (int dir = 0; dir < 2; dir++) {
        ds.roots.r4096[dir] = d + (0 + dir) * 4096;
        ds.roots.lo[dir] = d + (2 + dir) * 4096;
        ds.roots.hi[dir] = d + (4 + dir) * 4096;
    }
    return TVH_OK;
}

DeviceState* current_device() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        set_error(TVH_ERR_DEVICE, "no current HIP device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_dev_mu);
    DeviceState& ds = g_devs[dev];
    if (ds.device < 0) {
        if (build_tables(ds)) return nullptr;
        if (ntt_set_attributes()) return nullptr;
        if (noalloc_scratch_setup(&ds)) return nullptr;
        ds.device = dev;
    }
    return &ds;
}

DeviceState* initialized_device() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        set_error(TVH_ERR_DEVICE, "no current HIP device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_dev_mu);
    if (g_devs[dev].device < 0) {
        set_error(TVH_ERR_DEVICE, "device %d is not set up: call tvh_init(%d) first", dev, dev);
        return nullptr;
    }
    return &g_devs[dev];
}

void* arena(DeviceState* ds, size_t bytes, hipStream_t st) {
    std::lock_guard<std::mutex> lk(g_dev_mu);
    if (ds->arena_bytes >= bytes) return ds->arena;
    if (ds->arena) {
        // work already queued may still use the old arena
        (void)hipDeviceSynchronize();
        (void)hipFree(ds->arena);
        ds->arena = nullptr;
        ds->arena_bytes = 0;
    }
    (void)st;
    if (hipMalloc(&ds->arena, bytes) != hipSuccess) {
        ds->arena = nullptr;
        set_error(TVH_ERR_OOM, "arena of %zu bytes", bytes);
        return nullptr;
    }
    ds->arena_bytes = bytes;
    return ds->arena;
}

}  // namespace tvh

using namespace tvh;

#define STREAM(s) ((hipStream_t)(s))

extern "C" {

int tvh_init(int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return set_error(TVH_ERR_DEVICE, "no HIP device");
    if (device < 0 || device >= n) return set_error(TVH_ERR_INVALID_ARG, "device %d", device);
    if (hipSetDevice(device) != hipSuccess) return set_error(TVH_ERR_DEVICE, "hipSetDevice");
    return current_device() ? TVH_OK : TVH_ERR_DEVICE;
}

const char* tvh_status_string(int s) {
    switch (s) {
        case TVH_OK: return "ok";
        case TVH_ERR_INVALID_ARG: return "invalid argument";
        case TVH_ERR_NOT_POW2: return "length is not a power of two";
        case TVH_ERR_OOM: return "out of device memory";
        case TVH_ERR_DEVICE: return "device error";
        case TVH_ERR_LAYOUT: return "unknown tree layout";
        case TVH_ERR_NOT_CANONICAL: return "non-canonical field element";
        case TVH_ERR_UNSUPPORTED: return "unsupported shape";
        case TVH_ERR_RANGE: return "index out of range";
        default: return "unknown status";
    }
}

const char* tvh_last_error(void) { return g_err; }

// ---------------------------------------------------------------- NTT
// The batched NTT through the per-device arena (grown on first use of a larger
// size), from src (null: d itself, in place) into d.
static int ntt_batched_arena(uint64_t* d, size_t n, size_t cols, bool inv, hipStream_t st,
                             const uint64_t* src = nullptr, const NttScale* scale = nullptr) {
    if (!d && cols) return set_error(TVH_ERR_INVALID_ARG, "null data");
    if (cols == 0) return TVH_OK;
    if (!is_pow2(n)) return set_error(TVH_ERR_NOT_POW2, "NTT size must be a power of 2");
    DeviceState* ds = current_device();
    if (!ds) return TVH_ERR_DEVICE;
    u64 need = ntt_scratch_elems(n, cols);
    u64* scratch = (u64*)arena(ds, need * sizeof(u64), st);
    if (!scratch) return TVH_ERR_OOM;
    return launch_ntt_batched(src ? src : d, d, n, cols, inv, scratch, need, st, scale);
}

int tvh_ntt_forward_batched_dev(uint64_t* d, size_t n, size_t cols, void* s) {
    return ntt_batched_arena(d, n, cols, false, STREAM(s));
}
int tvh_ntt_inverse_batched_dev(uint64_t* d, size_t n, size_t cols, void* s) {
    return ntt_batched_arena(d, n, cols, true, STREAM(s));
}
int tvh_ntt_batched_noalloc_dev(uint64_t* d, size_t n, size_t cols, int inverse, void* s) {
    if (!d && cols) return set_error(TVH_ERR_INVALID_ARG, "null data");
    if (cols == 0) return TVH_OK;
    if (!is_pow2(n)) return set_error(TVH_ERR_NOT_POW2, "NTT size must be a power of 2");
    if (n > kNoAllocScratchElems)
        return set_error(TVH_ERR_UNSUPPORTED,
                         "n = %zu exceeds the no-allocation NTT's scratch of %llu elements "
                         "(use tvh_ntt_*_batched_dev)",
                         n, (unsigned long long)kNoAllocScratchElems);
    // set-up (tables, the scratch's address) must have happened: nothing here allocates
    DeviceState* ds = initialized_device();
    if (!ds) return TVH_ERR_DEVICE;
    return launch_ntt_batched(d, d, n, cols, inverse != 0, ds->noalloc_scratch, kNoAllocScratchElems,
                              STREAM(s));
}

int tvh_ntt_forward_dev(uint64_t* d, size_t n, void* s) {
    return tvh_ntt_forward_batched_dev(d, n, 1, s);
}
int tvh_ntt_inverse_dev(uint64_t* d, size_t n, void* s) {
    return tvh_ntt_inverse_batched_dev(d, n, 1, s);
}

// ---------------------------------------------------------------- LDE
static RandSpec bfe_rand(const uint64_t* r, size_t k) {
    RandSpec rs;
    rs.ptr = k ? r : nullptr;
    rs.k = rs.ptr ? k : 0;
    rs.group = 1;
    rs.group_stride = k;
    rs.limb_stride = 0;
    rs.elem_stride = 1;
    rs.limb_mask = 1;
    return rs;
}

static RandSpec xfe_rand(const uint64_t* r, size_t k, int xfe) {
    RandSpec rs;
    rs.ptr = k ? r : nullptr;
    rs.k = rs.ptr ? k : 0;
    rs.group = 3;
    if (xfe) {
        rs.group_stride = 3 * k;
        rs.limb_stride = 1;
        rs.elem_stride = 3;
        rs.limb_mask = 7;
    } else {
        rs.group_stride = k;
        rs.limb_stride = 0;
        rs.elem_stride = 1;
        rs.limb_mask = 1;
    }
    return rs;
}

int tvh_lde_batch_dev(const uint64_t* trace, size_t cols, size_t n, const uint64_t* rand,
                      size_t k, uint64_t toff, uint64_t goff, size_t m, uint64_t* out, void* s) {
    if (cols && (!trace || !out)) return set_error(TVH_ERR_INVALID_ARG, "null buffer");
    if (k && !rand) return set_error(TVH_ERR_INVALID_ARG, "null randomizers");
    return launch_lde(trace, cols, n, bfe_rand(rand, k), toff, goff, m, false, out, STREAM(s));
}

int tvh_lde_prepare(size_t n, size_t m, uint64_t toff, uint64_t goff) {
    return lde_prepare_consts(n, m, toff, goff);
}

int tvh_lde_release_constants(void) { return lde_release_consts(); }

int tvh_lde_set_variant(uint32_t log_n, uint32_t log_e, int variant) {
    return set_lde_variant(log_n, log_e, variant);
}

// form | column tile << 2 | LDE_ROW_HALF (16), LDE_INV_HALF (32) or LDE_ROW_QUARTER (64)
static const int kLdeVariants[] = {1, 2, 5, 6, 9, 10, 18, 22, 26, 34, 38, 42, 66, 70, 74};
static constexpr int kNumLdeVariants = sizeof(kLdeVariants) / sizeof(kLdeVariants[0]);

int tvh_lde_variant_count(void) { return kNumLdeVariants; }

int tvh_lde_get_variant(uint32_t log_n, uint32_t log_e, int* variant) {
    return tvh_lde_get_variant_for_width(log_n, log_e, 256, variant);
}

int tvh_lde_get_variant_for_width(uint32_t log_n, uint32_t log_e, size_t num_cols, int* variant) {
    if (!variant) return set_error(TVH_ERR_INVALID_ARG, "null variant pointer");
    if (log_n > kMaxLogN || log_e > 4) return set_error(TVH_ERR_INVALID_ARG, "shape");
    *variant = lde_variant(log_n, log_e, num_cols);
    return TVH_OK;
}

int tvh_lde_autotune(size_t cols, size_t n, size_t m, int* chosen, float* ms) {
    if (!is_pow2(n) || !is_pow2(m) || m < n || m / n > 16)
        return set_error(TVH_ERR_NOT_POW2, "autotune shape");
    if (cols == 0) return set_error(TVH_ERR_INVALID_ARG, "autotune needs num_cols > 0");
    DeviceState* ds = current_device();
    if (!ds) return TVH_ERR_DEVICE;
    u32 logN = log2u(n), logE = log2u(m / n);
    // enough columns to fill the GPU, at most 4 GiB of codewords
    size_t tc = cols;
    size_t cap = (4ull << 30) / (8 * m);
    if (tc > cap) tc = cap > 0 ? cap : 1;
    u64 *trace = nullptr, *out = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = TVH_OK;
    if (hipMalloc((void**)&trace, 8 * n * tc) != hipSuccess ||
        hipMalloc((void**)&out, 8 * m * tc) != hipSuccess)
        rc = set_error(TVH_ERR_OOM, "autotune buffers");
    if (!rc && (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = set_error(TVH_ERR_DEVICE, "autotune stream/events");
    if (!rc) rc = launch_fill_uniform(trace, n * tc, 77, st);
    float best = 0;
    int best_v = LDE_WIDE;
    // the variants are timed on tc columns (their width class); the choice is
    // kept for the caller's width class
    const int wc_timed = width_class(tc), wc = width_class(cols);
    int prev_timed = lde_variant(logN, logE, tc);
    int prev = lde_variant(logN, logE, cols);
    if (!rc) rc = lde_prepare_consts(n, m, 1, 7);
    bool any = false;
    for (int i = 0; i < kNumLdeVariants && !rc; i++) {
        set_lde_variant(logN, logE, kLdeVariants[i], wc_timed);
        RandSpec none;
        rc = launch_lde(trace, tc, n, none, 1, 7, m, false, out, st);  // warm-up
        if (rc == TVH_ERR_UNSUPPORTED) {  // this variant cannot run the shape: skip it
            rc = TVH_OK;
            if (ms) ms[i] = -1.0f;
            continue;
        }
        if (!rc && hipEventRecord(e0, st) != hipSuccess) rc = set_error(TVH_ERR_DEVICE, "event");
        for (int r = 0; r < 3 && !rc; r++) rc = launch_lde(trace, tc, n, none, 1, 7, m, false, out, st);
        if (!rc && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
            rc = set_error(TVH_ERR_DEVICE, "autotune timing");
        float t = 0;
        if (!rc && hipEventElapsedTime(&t, e0, e1) != hipSuccess) rc = set_error(TVH_ERR_DEVICE, "elapsed");
        t /= 3;
        if (ms) ms[i] = t;
        if (!rc && (!any || t < best)) {
            best = t;
            best_v = kLdeVariants[i];
            any = true;
        }
    }
    if (!rc && !any) rc = set_error(TVH_ERR_UNSUPPORTED, "no LDE variant runs this shape");
    set_lde_variant(logN, logE, prev_timed, wc_timed);
    set_lde_variant(logN, logE, rc ? prev : best_v, wc);
    if (chosen) *chosen = rc ? prev : best_v;
    if (st) (void)hipStreamSynchronize(st);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (trace) (void)hipFree(trace);
    if (out) (void)hipFree(out);
    return rc;
}

int tvh_lde_column_dev(const uint64_t* trace, size_t n, const uint64_t* rand, size_t k,
                       uint64_t toff, uint64_t goff, size_t m, uint64_t* out, void* s) {
    return tvh_lde_batch_dev(trace, 1, n, rand, k, toff, goff, m, out, s);
}

int tvh_lde_xfe_limbs_dev(const uint64_t* limbs, size_t xcols, size_t n, const uint64_t* rand,
                          size_t k, uint64_t toff, uint64_t goff, size_t m, uint64_t* out,
                          int xfe_r, void* s) {
    if (xcols && (!limbs || !out)) return set_error(TVH_ERR_INVALID_ARG, "null buffer");
    if (k && !rand) return set_error(TVH_ERR_INVALID_ARG, "null randomizers");
    return launch_lde(limbs, 3 * xcols, n, xfe_rand(rand, k, xfe_r), toff, goff, m, true, out,
                      STREAM(s));
}

// Interleaved XFE columns through a limb-major staging copy in `work`
// (3 * xcols * (n + m) u64).
static int lde_xfe_interleaved(const uint64_t* xt, size_t xcols, size_t n, const uint64_t* rand,
                               size_t k, uint64_t toff, uint64_t goff, size_t m, uint64_t* xout,
                               int xfe_r, u64* work, hipStream_t st) {
    u64* in_lm = work;
    u64* out_lm = work + 3 * xcols * n;
    int rc = launch_xfe_interleave(xt, in_lm, xcols, n, true, st);
    if (!rc) rc = tvh_lde_xfe_limbs_dev(in_lm, xcols, n, rand, k, toff, goff, m, out_lm, xfe_r, st);
    if (!rc) rc = launch_xfe_interleave(out_lm, xout, xcols, m, false, st);
    return rc;
}

int tvh_lde_xfe_batch_dev(const uint64_t* xt, size_t xcols, size_t n, const uint64_t* rand,
                          size_t k, uint64_t toff, uint64_t goff, size_t m, uint64_t* xout,
                          int xfe_r, void* s) {
    if (!xcols) return TVH_OK;
    if (!xt || !xout) return set_error(TVH_ERR_INVALID_ARG, "null buffer");
    if (!is_pow2(n) || !is_pow2(m)) return set_error(TVH_ERR_NOT_POW2, "lde lengths");
    DeviceState* ds = current_device();
    if (!ds) return TVH_ERR_DEVICE;
    u64* work = (u64*)arena(ds, 3 * xcols * (n + m) * sizeof(u64), STREAM(s));
    if (!work) return TVH_ERR_OOM;
    return lde_xfe_interleaved(xt, xcols, n, rand, k, toff, goff, m, xout, xfe_r, work, STREAM(s));
}

// ---------------------------------------------------------------- two-phase coset evaluation
// The power tables of the two-phase scaling pass, in the per-device arena
// after the batched NTT's scratch (the arena is sized for both, so the NTT
// that follows on the stream finds it large enough and leaves the tables be).
static u64* coset_tables(size_t n, size_t cols, hipStream_t st) {
    DeviceState* ds = current_device();
    if (!ds) return nullptr;
    if (n > (1ull << kMaxLogN)) {
        set_error(TVH_ERR_UNSUPPORTED, "trace_len > 2^%u", kMaxLogN);
        return nullptr;
    }
    u64 scratch = ntt_scratch_elems(n, cols);
    u64* base = (u64*)arena(ds, (scratch + scale_table_elems(n)) * sizeof(u64), st);
    return base ? base + scratch : nullptr;
}

int tvh_trace_coefficients_dev(const uint64_t* trace, size_t cols, size_t n, uint64_t toff,
                               uint64_t* coeffs, void* s) {
    if (cols && (!trace || !coeffs)) return set_error(TVH_ERR_INVALID_ARG, "null buffer");
    if (!cols) return TVH_OK;
    if (!is_pow2(n)) return set_error(TVH_ERR_NOT_POW2, "trace_len must be a power of 2");
    if (toff == 0 || toff >= TVH_P) return set_error(TVH_ERR_INVALID_ARG, "trace offset must be canonical and nonzero");
    hipStream_t st = STREAM(s);
    u64* tables = nullptr;
    if (toff != 1 && !(tables = coset_tables(n, cols, st))) return TVH_ERR_OOM;
    // the column step reads the trace, the row step writes the coefficients
    int rc = ntt_batched_arena(coeffs, n, cols, true, st, trace);
    if (rc || toff == 1) return rc;
    return launch_scale_table(coeffs, coeffs, n, cols, gl::inv(toff), nullptr, 0, 0, tables, st);
}

int tvh_evaluate_coset_dev(const uint64_t* coeffs, size_t cols, size_t n, const uint64_t* rand,
                           size_t k, uint64_t toff, uint64_t coset_offset, uint64_t* out,
                           uint64_t* tail_scratch, void* s) {
    (void)tail_scratch;  // one NTT covers the randomizer tail (g^n is constant on the coset)
    if (cols && (!coeffs || !out)) return set_error(TVH_ERR_INVALID_ARG, "null buffer");
    if (k && !rand) return set_error(TVH_ERR_INVALID_ARG, "null randomizers");
    if (!cols) return TVH_OK;
    if (!is_pow2(n)) return set_error(TVH_ERR_NOT_POW2, "trace_len must be a power of 2");
    if (k > n) return set_error(TVH_ERR_UNSUPPORTED, "randomizer_len > trace_len");
    if (toff >= TVH_P || coset_offset >= TVH_P) return set_error(TVH_ERR_INVALID_ARG, "offsets must be canonical");
    hipStream_t st = STREAM(s);
    u64* tables = coset_tables(n, cols, st);
    if (!tables) return TVH_ERR_OOM;
    u64 delta = gl::sub(gl::pow(coset_offset, n), gl::pow(toff, n));
    // b_i = (c_i + [i < k] r_i delta) g^i, applied as the NTT's column step
    // loads the coefficients; its row step writes the coset's values to out
    const NttScale S{coset_offset, tables, k ? rand : nullptr, k, delta};
    return ntt_batched_arena(out, n, cols, false, st, coeffs, &S);
}

int tvh_scatter_digests_strided_dev(const uint64_t* coset_digests, uint64_t* leaves,
                                    size_t coset_len, size_t coset_idx, size_t num_cosets, void* s) {
    if (coset_len && (!coset_digests || !leaves)) return set_error(TVH_ERR_INVALID_ARG, "null buffer");
    if (coset_idx >= num_cosets) return set_error(TVH_ERR_RANGE, "coset %zu of %zu", coset_idx, num_cosets);
    return launch_scatter_digests(coset_digests, leaves, coset_len, coset_idx, num_cosets, STREAM(s));
}

// ---------------------------------------------------------------- layout transforms
int tvh_rowmajor_to_colmajor_bfe_dev(const uint64_t* rm, uint64_t* cm, size_t n, size_t width,
                                     void* s) {
    if (n && width && (!rm || !cm)) return set_error(TVH_ERR_INVALID_ARG, "null buffer");
    if (rm == cm && n > 1 && width > 1) return set_error(TVH_ERR_INVALID_ARG, "transpose is not in place");
    return launch_transpose(rm, cm, n, width, STREAM(s));
}
int tvh_rowmajor_to_colmajor_xfe_dev(const uint64_t* rm, uint64_t* cm, size_t n, size_t xwidth,
                                     void* s) {
    return tvh_rowmajor_to_colmajor_bfe_dev(rm, cm, n, 3 * xwidth, s);
}
int tvh_colmajor_to_rowmajor_bfe_dev(const uint64_t* cm, uint64_t* rm, size_t n, size_t width,
                                     void* s) {
    if (n && width && (!rm || !cm)) return set_error(TVH_ERR_INVALID_ARG, "null buffer");
    if (rm == cm && n > 1 && width > 1) return set_error(TVH_ERR_INVALID_ARG, "transpose is not in place");
    return launch_transpose(cm, rm, width, n, STREAM(s));
}
int tvh_colmajor_to_rowmajor_xfe_dev(const uint64_t* cm, uint64_t* rm, size_t n, size_t xwidth,
                                     void* s) {
    return tvh_colmajor_to_rowmajor_bfe_dev(cm, rm, n, 3 * xwidth, s);
}

// ---------------------------------------------------------------- Tip5
int tvh_hash_bfield_rows_dev(const uint64_t* t, size_t rows, size_t cols, uint64_t* dg, void* s) {
    if (rows && (!dg || (cols && !t))) return set_error(TVH_ERR_INVALID_ARG, "null buffer");
    return launch_hash_rows(t, rows, cols, dg, STREAM(s));
}

int tvh_hash_xfield_rows_dev(const uint64_t* t, size_t rows, size_t xcols, uint64_t* dg, void* s) {
    return tvh_hash_bfield_rows_dev(t, rows, 3 * xcols, dg, s);
}

int tvh_tip5_permutation_batch_dev(uint64_t* st, size_t num, void* s) {
    if (num && !st) return set_error(TVH_ERR_INVALID_ARG, "null states");
    return launch_permutation(st, num, STREAM(s));
}

int tvh_tip5_rounds_dev(uint64_t* st, size_t num, uint32_t first, uint32_t count, uint64_t* trace,
                        void* s) {
    if (num && !st) return set_error(TVH_ERR_INVALID_ARG, "null states");
    if (first > (uint32_t)tip5::ROUNDS || count > (uint32_t)tip5::ROUNDS - first)
        return set_error(TVH_ERR_RANGE, "rounds [%u, %u + %u) beyond round %d", first, first, count,
                         tip5::ROUNDS - 1);
    if (!current_device()) return TVH_ERR_DEVICE;
    return launch_permutation_rounds(st, num, first, count, trace, STREAM(s));
}

int tvh_hash_pairs_dev(const uint64_t* l, const uint64_t* r, uint64_t* o, size_t count, void* s) {
    if (count && (!l || !r || !o)) return set_error(TVH_ERR_INVALID_ARG, "null buffer");
    return launch_hash_pairs(l, r, o, count, STREAM(s));
}

// ---------------------------------------------------------------- Merkle
int tvh_merkle_tree_dev(const uint64_t* leaves, uint64_t* tree, size_t n, int layout, void* s) {
    if (layout != TVH_TREE_LEVEL_ORDER && layout != TVH_TREE_HEAP)
        return set_error(TVH_ERR_LAYOUT, "layout %d", layout);
    if (!n) return set_error(TVH_ERR_INVALID_ARG, "Cannot create Merkle tree with no leaves");
    if (!is_pow2(n)) return set_error(TVH_ERR_NOT_POW2, "Number of leaves must be a power of 2");
    if (!leaves || !tree) return set_error(TVH_ERR_INVALID_ARG, "null buffer");
    if (!current_device()) return TVH_ERR_DEVICE;
    u64* leaf_level = tree + 5 * merkle_level_base(n, 0, layout);
    if (leaf_level != leaves) {
        if (hipMemcpyAsync(leaf_level, leaves, n * 5 * sizeof(u64), hipMemcpyDeviceToDevice,
                           STREAM(s)) != hipSuccess)
            return set_error(TVH_ERR_DEVICE, "leaf copy");
    }
    return launch_merkle_levels(tree, n, layout, STREAM(s));
}

int tvh_merkle_root_dev(const uint64_t* leaves, uint64_t* root, size_t n, void* s) {
    if (!n) return set_error(TVH_ERR_INVALID_ARG, "Cannot create Merkle tree with no leaves");
    if (!is_pow2(n)) return set_error(TVH_ERR_NOT_POW2, "Number of leaves must be a power of 2");
    if (!leaves || !root) return set_error(TVH_ERR_INVALID_ARG, "null buffer");
    DeviceState* ds = current_device();
    if (!ds) return TVH_ERR_DEVICE;
    u64* tree = (u64*)arena(ds, (2 * n - 1) * 5 * sizeof(u64), STREAM(s));
    if (!tree) return TVH_ERR_OOM;
    int rc = tvh_merkle_tree_dev(leaves, tree, n, TVH_TREE_LEVEL_ORDER, s);
    if (rc) return rc;
    if (hipMemcpyAsync(root, tree + 5 * (2 * n - 2), 5 * sizeof(u64), hipMemcpyDeviceToDevice,
                       STREAM(s)) != hipSuccess)
        return set_error(TVH_ERR_DEVICE, "root copy");
    return TVH_OK;
}

int tvh_merkle_gather_dev(const uint64_t* tree, size_t n, int layout, const uint64_t* idx,
                          size_t count, uint64_t* out, void* s) {
    if (layout != TVH_TREE_LEVEL_ORDER && layout != TVH_TREE_HEAP)
        return set_error(TVH_ERR_LAYOUT, "layout %d", layout);
    if (!is_pow2(n)) return set_error(TVH_ERR_NOT_POW2, "num_leaves");
    if (count && (!tree || !idx || !out)) return set_error(TVH_ERR_INVALID_ARG, "null buffer");
    return launch_gather_heap(tree, n, layout, idx, count, out, STREAM(s));
}

int tvh_auth_path_indices(size_t n, size_t leaf, uint64_t* out, size_t cap, size_t* cnt) {
    if (!is_pow2(n)) return set_error(TVH_ERR_NOT_POW2, "num_leaves");
    if (leaf >= n) return set_error(TVH_ERR_RANGE, "Leaf index out of range");
    size_t c = 0;
    for (u64 v = n + leaf; v > 1; v >>= 1) {
        if (out && c < cap) out[c] = v ^ 1;
        c++;
    }
    if (cnt) *cnt = c;
    if (out && c > cap) return set_error(TVH_ERR_INVALID_ARG, "capacity %zu < %zu", cap, c);
    return TVH_OK;
}

int tvh_auth_structure_indices(size_t n, const size_t* leaves, size_t count, uint64_t* out,
                               size_t cap, size_t* cnt) {
    if (!n) return set_error(TVH_ERR_INVALID_ARG, "num_leaves must be > 0");
    if (!is_pow2(n)) return set_error(TVH_ERR_NOT_POW2, "num_leaves must be a power of 2");
    if (count && !leaves) return set_error(TVH_ERR_INVALID_ARG, "null leaf indices");
    // needed siblings minus computable nodes over the paths (merkle_tree.cpp:228-265)
    std::vector<uint8_t> mark(2 * n, 0);  // bit0 computable, bit1 needed
    for (size_t j = 0; j < count; j++) {
        if (leaves[j] >= n) return set_error(TVH_ERR_RANGE, "Leaf index out of range");
        for (u64 v = n + leaves[j]; v > 1; v >>= 1) {
            mark[v] |= 1;
            mark[v ^ 1] |= 2;
        }
    }
    size_t c = 0;
    for (u64 v = 2 * n - 1; v >= 2; v--) {
        if (mark[v] == 2) {
            if (out && c < cap) out[c] = v;
            c++;
        }
    }
    if (cnt) *cnt = c;
    if (out && c > cap) return set_error(TVH_ERR_INVALID_ARG, "capacity %zu < %zu", cap, c);
    return TVH_OK;
}

// ---------------------------------------------------------------- commit
int tvh_commit_dev(const uint64_t* trace, size_t cols, size_t n, const uint64_t* rand, size_t k,
                   uint64_t toff, uint64_t goff, size_t m, uint64_t* cw, uint64_t* tree,
                   int layout, void* s) {
    if (layout != TVH_TREE_LEVEL_ORDER && layout != TVH_TREE_HEAP)
        return set_error(TVH_ERR_LAYOUT, "layout %d", layout);
    if (!trace || !cw || !tree) return set_error(TVH_ERR_INVALID_ARG, "null buffer");
    if (!is_pow2(m)) return set_error(TVH_ERR_NOT_POW2, "target_len");
    int rc = tvh_lde_batch_dev(trace, cols, n, rand, k, toff, goff, m, cw, s);
    if (rc) return rc;
    u64* leaf_level = tree + 5 * merkle_level_base(m, 0, layout);
    rc = launch_hash_rows(cw, m, cols, leaf_level, STREAM(s));
    if (rc) return rc;
    return launch_merkle_levels(tree, m, layout, STREAM(s));
}

int tvh_sponge_init_dev(uint64_t* state, void* s) {
    if (!state) return set_error(TVH_ERR_INVALID_ARG, "null state");
    if (!current_device()) return TVH_ERR_DEVICE;
    return launch_sponge(state, nullptr, 0, nullptr, 2, STREAM(s));
}

int tvh_sponge_absorb_dev(uint64_t* state, const uint64_t* data, size_t len, void* s) {
    if (!state || (len && !data)) return set_error(TVH_ERR_INVALID_ARG, "null buffer");
    if (!current_device()) return TVH_ERR_DEVICE;
    return launch_sponge(state, data, len, nullptr, 0, STREAM(s));
}

int tvh_sponge_squeeze_dev(uint64_t* state, uint64_t* out, void* s) {
    if (!state || !out) return set_error(TVH_ERR_INVALID_ARG, "null buffer");
    if (!current_device()) return TVH_ERR_DEVICE;
    return launch_sponge(state, nullptr, 0, out, 1, STREAM(s));
}

int tvh_xfe_leaves_dev(const uint64_t* xfe, size_t n, uint64_t* leaves, void* s) {
    if (n && (!xfe || !leaves)) return set_error(TVH_ERR_INVALID_ARG, "null buffer");
    if (!current_device()) return TVH_ERR_DEVICE;
    return launch_xfe_leaves(xfe, n, leaves, STREAM(s));
}

int tvh_fri_commit_dev(const uint64_t* xfe, size_t n, uint64_t* tree, int layout, void* s) {
    if (layout != TVH_TREE_LEVEL_ORDER && layout != TVH_TREE_HEAP)
        return set_error(TVH_ERR_LAYOUT, "layout %d", layout);
    if (!xfe || !tree) return set_error(TVH_ERR_INVALID_ARG, "null buffer");
    if (!is_pow2(n)) return set_error(TVH_ERR_NOT_POW2, "codeword length must be a power of 2");
    if (!current_device()) return TVH_ERR_DEVICE;
    u64* leaf_level = tree + 5 * merkle_level_base(n, 0, layout);
    int rc = launch_xfe_leaves(xfe, n, leaf_level, STREAM(s));
    if (rc) return rc;
    return launch_merkle_levels(tree, n, layout, STREAM(s));
}

int tvh_fill_uniform_dev(uint64_t* d, size_t count, uint64_t seed, void* s) {
    if (count && !d) return set_error(TVH_ERR_INVALID_ARG, "null buffer");
    if (!current_device()) return TVH_ERR_DEVICE;
    return launch_fill_uniform(d, count, seed, STREAM(s));
}

int tvh_fill_table_dev(uint64_t* d, size_t rows, size_t cols, uint64_t seed, void* s) {
    if (rows && cols && !d) return set_error(TVH_ERR_INVALID_ARG, "null buffer");
    DeviceState* ds = current_device();
    if (!ds) return TVH_ERR_DEVICE;
    if (!rows || !cols) return TVH_OK;
    u64* rej = (u64*)arena(ds, fill_table_scratch_elems() * sizeof(u64), STREAM(s));
    if (!rej) return TVH_ERR_OOM;
    return launch_fill_table(d, rows, cols, seed, rej, STREAM(s));
}

size_t tvh_commit_frugal_workspace_elems(size_t cols, size_t n, size_t m, size_t chunk) {
    (void)n;
    if (chunk == 0) return 0;
    size_t c = chunk < cols ? chunk : cols;
    // chunk codewords + 16 state words + 9 carried elements per row
    return m * (c + 16 + 9);
}

int tvh_commit_frugal_dev(const uint64_t* trace, size_t cols, size_t n, const uint64_t* rand,
                          size_t k, uint64_t toff, uint64_t goff, size_t m, size_t chunk,
                          uint64_t* work, uint64_t* tree, int layout, void* s) {
    if (layout != TVH_TREE_LEVEL_ORDER && layout != TVH_TREE_HEAP)
        return set_error(TVH_ERR_LAYOUT, "layout %d", layout);
    if (!work || !tree || (cols && !trace)) return set_error(TVH_ERR_INVALID_ARG, "null buffer");
    if (k && !rand) return set_error(TVH_ERR_INVALID_ARG, "null randomizers");
    if (chunk == 0) return set_error(TVH_ERR_INVALID_ARG, "chunk_cols must be > 0");
    if (!is_pow2(n) || !is_pow2(m)) return set_error(TVH_ERR_NOT_POW2, "lengths");
    if (!current_device()) return TVH_ERR_DEVICE;
    size_t c = chunk < cols ? chunk : cols;
    u64* cw = work;
    u64* state = work + m * c;
    u64* pend = state + m * 16;
    u64* dig = tree + 5 * merkle_level_base(m, 0, layout);
    hipStream_t st = STREAM(s);
    int rc = TVH_OK;
    if (cols == 0) {
        rc = launch_absorb_chunk(nullptr, m, 0, state, pend, 0, true, true, dig, st);
    }
    for (size_t c0 = 0; c0 < cols && !rc; c0 += c) {
        size_t nc = cols - c0 < c ? cols - c0 : c;
        rc = launch_lde(trace + c0 * n, nc, n, bfe_rand(k ? rand + c0 * k : nullptr, k), toff, goff,
                        m, false, cw, st);
        if (!rc)
            rc = launch_absorb_chunk(cw, m, nc, state, pend, (u32)(c0 % 10), c0 == 0,
                                     c0 + nc == cols, dig, st);
    }
    if (rc) return rc;
    return launch_merkle_levels(tree, m, layout, st);
}

// ---------------------------------------------------------------- host forms
struct HostCall {
    hipStream_t st = nullptr;
    std::vector<void*> bufs;
    int rc = TVH_OK;
    HostCall() {
        if (!current_device()) rc = TVH_ERR_DEVICE;
        else if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
            rc = set_error(TVH_ERR_DEVICE, "stream create");
    }
    ~HostCall() {
        if (st) (void)hipStreamSynchronize(st);
        for (void* p : bufs) (void)hipFree(p);
        if (st) (void)hipStreamDestroy(st);
    }
    u64* alloc(size_t elems) {
        void* p = nullptr;
        if (hipMalloc(&p, (elems ? elems : 1) * sizeof(u64)) != hipSuccess) {
            rc = set_error(TVH_ERR_OOM, "%zu bytes", elems * sizeof(u64));
            return nullptr;
        }
        bufs.push_back(p);
        return (u64*)p;
    }
    int up(u64* d, const u64* h, size_t elems) {
        if (!elems) return TVH_OK;
        int e = staged_h2d(d, h, elems * sizeof(u64), st);
        return e ? (rc = e) : TVH_OK;
    }
    int down(u64* h, const u64* d, size_t elems) {
        if (!elems) return TVH_OK;
        if (hipMemcpyAsync(h, d, elems * sizeof(u64), hipMemcpyDeviceToHost, st) != hipSuccess)
            return rc = set_error(TVH_ERR_DEVICE, "download");
        return TVH_OK;
    }
    int sync() {
        if (hipStreamSynchronize(st) != hipSuccess) return rc = set_error(TVH_ERR_DEVICE, "sync");
        return TVH_OK;
    }
};

static int check_canonical(const u64* p, size_t n) {
    for (size_t i = 0; i < n; i++)
        if (p[i] >= TVH_P) return set_error(TVH_ERR_NOT_CANONICAL, "element %zu >= p", i);
    return TVH_OK;
}

static int ntt_host(uint64_t* data, size_t n, bool inv) {
    if (!data) return set_error(TVH_ERR_INVALID_ARG, "null data");
    if (!is_pow2(n)) return set_error(TVH_ERR_NOT_POW2, "NTT size must be a power of 2");
    int rc = check_canonical(data, n);
    if (rc) return rc;
    HostCall hc;
    if (hc.rc) return hc.rc;
    // a scratch of its own: concurrent host calls share nothing on the device
    u64* d = hc.alloc(n);
    u64* scratch = hc.alloc(n);
    if (hc.rc) return hc.rc;
    if ((rc = hc.up(d, data, n))) return rc;
    if ((rc = launch_ntt_batched(d, d, n, 1, inv, scratch, n, hc.st))) return rc;
    if ((rc = hc.down(data, d, n))) return rc;
    return hc.sync();
}

int tvh_ntt_forward_host(uint64_t* data, size_t n) { return ntt_host(data, n, false); }
int tvh_ntt_inverse_host(uint64_t* data, size_t n) { return ntt_host(data, n, true); }

int tvh_lde_batch_host(const uint64_t* trace, size_t cols, size_t n, const uint64_t* rand,
                       size_t k, uint64_t toff, uint64_t goff, size_t m, uint64_t* out) {
    if (!cols) return TVH_OK;
    if (!trace || !out || (k && !rand)) return set_error(TVH_ERR_INVALID_ARG, "null buffer");
    int rc = check_canonical(trace, cols * n);
    if (!rc && k) rc = check_canonical(rand, cols * k);
    if (rc) return rc;
    HostCall hc;
    if (hc.rc) return hc.rc;
    u64* dt = hc.alloc(cols * n);
    u64* dr = hc.alloc(cols * k);
    u64* dout = hc.alloc(cols * m);
    if (hc.rc) return hc.rc;
    if ((rc = hc.up(dt, trace, cols * n)) || (rc = hc.up(dr, rand, cols * k))) return rc;
    if ((rc = tvh_lde_batch_dev(dt, cols, n, k ? dr : nullptr, k, toff, goff, m, dout, hc.st)))
        return rc;
    if ((rc = hc.down(out, dout, cols * m))) return rc;
    return hc.sync();
}

int tvh_lde_xfe_batch_host(const uint64_t* xt, size_t cols, size_t n, const uint64_t* rand,
                           size_t k, uint64_t toff, uint64_t goff, size_t m, uint64_t* xout,
                           int xfe_r) {
    if (!cols) return TVH_OK;
    if (!xt || !xout || (k && !rand)) return set_error(TVH_ERR_INVALID_ARG, "null buffer");
    size_t rlen = cols * k * (xfe_r ? 3 : 1);
    int rc = check_canonical(xt, 3 * cols * n);
    if (!rc && k) rc = check_canonical(rand, rlen);
    if (rc) return rc;
    HostCall hc;
    if (hc.rc) return hc.rc;
    if (!is_pow2(n) || !is_pow2(m)) return set_error(TVH_ERR_NOT_POW2, "lde lengths");
    u64* dt = hc.alloc(3 * cols * n);
    u64* dr = hc.alloc(rlen);
    u64* dout = hc.alloc(3 * cols * m);
    u64* work = hc.alloc(3 * cols * (n + m));  // private staging: no shared arena
    if (hc.rc) return hc.rc;
    if ((rc = hc.up(dt, xt, 3 * cols * n)) || (rc = hc.up(dr, rand, rlen))) return rc;
    if ((rc = lde_xfe_interleaved(dt, cols, n, k ? dr : nullptr, k, toff, goff, m, dout, xfe_r,
                                  work, hc.st)))
        return rc;
    if ((rc = hc.down(xout, dout, 3 * cols * m))) return rc;
    return hc.sync();
}

int tvh_tip5_permutation_batch_host(uint64_t* states, size_t num) {
    if (!num) return TVH_OK;
    if (!states) return set_error(TVH_ERR_INVALID_ARG, "null states");
    HostCall hc;
    if (hc.rc) return hc.rc;
    u64* d = hc.alloc(num * 16);
    if (!d) return hc.rc;
    int rc;
    if ((rc = hc.up(d, states, num * 16))) return rc;
    if ((rc = launch_permutation(d, num, hc.st))) return rc;
    if ((rc = hc.down(states, d, num * 16))) return rc;
    return hc.sync();
}

int tvh_tip5_rounds_host(uint64_t* states, size_t num, uint32_t first, uint32_t count,
                         uint64_t* trace) {
    if (!num) return TVH_OK;
    if (!states) return set_error(TVH_ERR_INVALID_ARG, "null states");
    if (first > (uint32_t)tip5::ROUNDS || count > (uint32_t)tip5::ROUNDS - first)
        return set_error(TVH_ERR_RANGE, "rounds [%u, %u + %u) beyond round %d", first, first, count,
                         tip5::ROUNDS - 1);
    HostCall hc;
    if (hc.rc) return hc.rc;
    u64* d = hc.alloc(num * 16);
    u64* dt = trace ? hc.alloc(num * 16 * (count + 1)) : nullptr;
    if (hc.rc) return hc.rc;
    int rc;
    if ((rc = hc.up(d, states, num * 16))) return rc;
    if ((rc = launch_permutation_rounds(d, num, first, count, dt, hc.st))) return rc;
    if ((rc = hc.down(states, d, num * 16))) return rc;
    if (trace && (rc = hc.down(trace, dt, num * 16 * (count + 1)))) return rc;
    return hc.sync();
}

int tvh_hash_pairs_host(const uint64_t* l, const uint64_t* r, uint64_t* o, size_t count) {
    if (!count) return TVH_OK;
    if (!l || !r || !o) return set_error(TVH_ERR_INVALID_ARG, "null buffer");
    HostCall hc;
    if (hc.rc) return hc.rc;
    u64* dl = hc.alloc(count * 5);
    u64* dr = hc.alloc(count * 5);
    u64* dout = hc.alloc(count * 5);
    if (hc.rc) return hc.rc;
    int rc;
    if ((rc = hc.up(dl, l, count * 5)) || (rc = hc.up(dr, r, count * 5))) return rc;
    if ((rc = launch_hash_pairs(dl, dr, dout, count, hc.st))) return rc;
    if ((rc = hc.down(o, dout, count * 5))) return rc;
    return hc.sync();
}

int tvh_hash_bfield_rows_host(const uint64_t* t, size_t rows, size_t cols, uint64_t* dg) {
    if (!rows) return TVH_OK;
    if (!dg || (cols && !t)) return set_error(TVH_ERR_INVALID_ARG, "null buffer");
    HostCall hc;
    if (hc.rc) return hc.rc;
    u64* dt = hc.alloc(rows * cols);
    u64* dd = hc.alloc(rows * 5);
    if (hc.rc) return hc.rc;
    int rc;
    if ((rc = hc.up(dt, t, rows * cols))) return rc;
    if ((rc = launch_hash_rows(dt, rows, cols, dd, hc.st))) return rc;
    if ((rc = hc.down(dg, dd, rows * 5))) return rc;
    return hc.sync();
}

int tvh_merkle_tree_host(const uint64_t* leaves, size_t n, int layout, uint64_t* tree_out,
                         uint64_t* root_out) {
    if (layout != TVH_TREE_LEVEL_ORDER && layout != TVH_TREE_HEAP)
        return set_error(TVH_ERR_LAYOUT, "layout %d", layout);
    if (!n) return set_error(TVH_ERR_INVALID_ARG, "Cannot create Merkle tree with no leaves");
    if (!is_pow2(n)) return set_error(TVH_ERR_NOT_POW2, "Number of leaves must be a power of 2");
    if (!leaves) return set_error(TVH_ERR_INVALID_ARG, "null leaves");
    HostCall hc;
    if (hc.rc) return hc.rc;
    size_t nodes = layout == TVH_TREE_HEAP ? 2 * n : 2 * n - 1;
    u64* dtree = hc.alloc(nodes * 5);
    if (!dtree) return hc.rc;
    u64* leaf_level = dtree + 5 * merkle_level_base(n, 0, layout);
    int rc;
    if ((rc = hc.up(leaf_level, leaves, n * 5))) return rc;
    if (layout == TVH_TREE_HEAP && hipMemsetAsync(dtree, 0, 5 * sizeof(u64), hc.st) != hipSuccess)
        return set_error(TVH_ERR_DEVICE, "memset");
    if ((rc = launch_merkle_levels(dtree, n, layout, hc.st))) return rc;
    if (tree_out && (rc = hc.down(tree_out, dtree, nodes * 5))) return rc;
    u64 root_node = layout == TVH_TREE_HEAP ? 1 : 2 * n - 2;
    if (root_out && (rc = hc.down(root_out, dtree + 5 * root_node, 5))) return rc;
    return hc.sync();
}

int tvh_commit_host(const uint64_t* trace, size_t cols, size_t n, const uint64_t* rand, size_t k,
                    uint64_t toff, uint64_t goff, size_t m, uint64_t* cw_out, uint64_t* dg_out,
                    uint64_t* root_out) {
    if (!trace || !root_out || (k && !rand)) return set_error(TVH_ERR_INVALID_ARG, "null buffer");
    if (!is_pow2(m)) return set_error(TVH_ERR_NOT_POW2, "target_len");
    int rc = check_canonical(trace, cols * n);
    if (!rc && k) rc = check_canonical(rand, cols * k);
    if (rc) return rc;
    HostCall hc;
    if (hc.rc) return hc.rc;
    u64* dt = hc.alloc(cols * n);
    u64* dr = hc.alloc(cols * k);
    u64* dcw = hc.alloc(cols * m);
    u64* dtree = hc.alloc((2 * m - 1) * 5);
    if (hc.rc) return hc.rc;
    if ((rc = hc.up(dt, trace, cols * n)) || (rc = hc.up(dr, rand, cols * k))) return rc;
    if ((rc = tvh_commit_dev(dt, cols, n, k ? dr : nullptr, k, toff, goff, m, dcw, dtree,
                             TVH_TREE_LEVEL_ORDER, hc.st)))
        return rc;
    if (cw_out && (rc = hc.down(cw_out, dcw, cols * m))) return rc;
    if (dg_out && (rc = hc.down(dg_out, dtree, m * 5))) return rc;
    if ((rc = hc.down(root_out, dtree + 5 * (2 * m - 2), 5))) return rc;
    return hc.sync();
}

}  // extern "C"
