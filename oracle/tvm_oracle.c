/*
 * CPU oracle for the table-commitment path.  TEST INFRASTRUCTURE ONLY: see
 * tvm_oracle.h.  Plain C11 restatement of the reference prover's CPU
 * algorithm; each function cites the reference file:line it follows.
 */
#include "tvm_oracle.h"

#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------- */
/* Field: canonical Goldilocks (include/types/b_field_element.hpp:28-104,     */
/* src/types/b_field_element.cpp:49-192).  The reference multiplies with      */
/* u128 % p; the reduction below gives the same canonical value.             */
/* ------------------------------------------------------------------------- */

static uint64_t reduce128(u128 x) {
    uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
    uint64_t hh = hi >> 32, hl = hi & 0xffffffffull;
    /* x = lo + hl*2^64 + hh*2^96 ; 2^64 = 2^32-1, 2^96 = -1 (mod p) */
    uint64_t t = lo - hh;
    if (lo < hh) t -= 0xffffffffull;            /* borrow: add p (mod 2^64) */
    uint64_t u = (hl << 32) - hl;               /* hl*(2^32-1) < 2^64 */
    uint64_t r = t + u;
    if (r < u) r += 0xffffffffull;              /* carry: subtract p (mod 2^64) */
    if (r >= ORC_P) r -= ORC_P;
    return r;
}

uint64_t orc_add(uint64_t a, uint64_t b) {       /* b_field_element.cpp:49-59 */
    uint64_t s = a + b;
    if (s < a || s >= ORC_P) s -= ORC_P;
    return s;
}

uint64_t orc_sub(uint64_t a, uint64_t b) {       /* b_field_element.cpp:61-69 */
    uint64_t d = a - b;
    if (a < b) d += ORC_P;
    return d;
}

uint64_t orc_mul(uint64_t a, uint64_t b) {       /* b_field_element.cpp:71-74 */
    return reduce128((u128)a * b);
}

uint64_t orc_pow(uint64_t a, uint64_t e) {       /* b_field_element.cpp:125-138 */
    uint64_t r = 1, b = a;
    while (e) {
        if (e & 1) r = orc_mul(r, b);
        b = orc_mul(b, b);
        e >>= 1;
    }
    return r;
}

uint64_t orc_inv(uint64_t a) {                   /* b_field_element.cpp:140-147 (Fermat) */
    return orc_pow(a, ORC_P - 2);
}

uint64_t orc_root_of_unity(uint32_t k) {         /* b_field_element.cpp:179-192 */
    if (k > 32) return 0;
    return orc_pow(7, (ORC_P - 1) >> k);
}

/* Montgomery reduction (hi:lo) * 2^-64 mod p, bit-for-bit the helper used by
 * the S-box (tip5.cpp:135-152). */
uint64_t orc_montyred(uint64_t hi, uint64_t lo) {
    uint64_t sh = lo << 32;
    uint64_t a = lo + sh;
    int ov = (a < lo) || (a < sh);
    uint64_t b = a - (a >> 32);
    if (ov) b -= 1;
    uint64_t r = hi - b;
    if (hi < b) r -= (1ull + ~ORC_P);
    return r;
}

/* ------------------------------------------------------------------------- */
/* NTT: natural order in and out, radix-2 DIT after a bit-reversal            */
/* (src/ntt/ntt.cpp:8-86).                                                   */
/* ------------------------------------------------------------------------- */

static int log2_exact(size_t n) {
    if (n == 0 || (n & (n - 1))) return -1;
    int l = 0;
    while (((size_t)1 << l) < n) l++;
    return l;
}

static void bit_reverse(uint64_t* a, size_t n, int logn) {   /* ntt.cpp:8-24 */
    for (size_t i = 0; i < n; i++) {
        size_t r = 0;
        for (int j = 0; j < logn; j++)
            if (i & ((size_t)1 << j)) r |= (size_t)1 << (logn - 1 - j);
        if (i < r) { uint64_t t = a[i]; a[i] = a[r]; a[r] = t; }
    }
}

static void ntt_core(uint64_t* a, size_t n, int logn, uint64_t omega, int inverse) {
    /* ntt.cpp:26-57 */
    bit_reverse(a, n, logn);
    for (size_t len = 2; len <= n; len <<= 1) {
        uint64_t wl = orc_pow(omega, n / len);
        size_t h = len / 2;
        for (size_t i = 0; i < n; i += len) {
            uint64_t w = 1;
            for (size_t j = 0; j < h; j++) {
                uint64_t u = a[i + j];
                uint64_t v = orc_mul(a[i + j + h], w);
                a[i + j] = orc_add(u, v);
                a[i + j + h] = orc_sub(u, v);
                w = orc_mul(w, wl);
            }
        }
    }
    if (inverse) {
        uint64_t ninv = orc_inv((uint64_t)n % ORC_P);
        for (size_t i = 0; i < n; i++) a[i] = orc_mul(a[i], ninv);
    }
}

int orc_ntt_forward(uint64_t* a, size_t n) {     /* ntt.cpp:59-71 */
    int l = log2_exact(n);
    if (l < 0 || l > 32) return -1;
    ntt_core(a, n, l, orc_root_of_unity((uint32_t)l), 0);
    return 0;
}

int orc_ntt_inverse(uint64_t* a, size_t n) {     /* ntt.cpp:73-86 */
    int l = log2_exact(n);
    if (l < 0 || l > 32) return -1;
    ntt_core(a, n, l, orc_inv(orc_root_of_unity((uint32_t)l)), 1);
    return 0;
}

int orc_evaluate_on_coset(const uint64_t* coeffs, size_t nc, size_t m,
                          uint64_t offset, uint64_t* out) {
    /* ntt.cpp:94-117: zero-pad to m, c_i *= offset^i, forward NTT(m) */
    if (log2_exact(m) < 0) return -1;
    uint64_t s = 1;
    for (size_t i = 0; i < m; i++) {
        uint64_t c = i < nc ? coeffs[i] : 0;
        out[i] = orc_mul(c, s);
        s = orc_mul(s, offset);
    }
    return orc_ntt_forward(out, m);
}

/* ------------------------------------------------------------------------- */
/* Randomized LDE (src/lde/lde_randomized.cpp:37-154 and :270-300).          */
/* ------------------------------------------------------------------------- */

/* interpolant*offset^-i + (x^n - o^n) * r(x): length max(n, n+k) */
static uint64_t* randomized_interpolant(const uint64_t* trace, size_t n, uint64_t toff,
                                        const uint64_t* rand, size_t k, size_t* out_len) {
    size_t len = n + k;
    uint64_t* p = (uint64_t*)calloc(len ? len : 1, sizeof(uint64_t));
    if (!p) return NULL;
    memcpy(p, trace, n * sizeof(uint64_t));
    orc_ntt_inverse(p, n);                                   /* :78-79 */
    uint64_t oinv = orc_inv(toff), s = 1;                    /* :82-87 */
    for (size_t i = 0; i < n; i++) { p[i] = orc_mul(p[i], s); s = orc_mul(s, oinv); }
    uint64_t on = orc_pow(toff, n);                          /* mul_zerofier_with :37-69 */
    for (size_t i = 0; i < k; i++) {
        p[i] = orc_sub(p[i], orc_mul(rand[i], on));
        p[n + i] = orc_add(p[n + i], rand[i]);
    }
    *out_len = len;
    return p;
}

int orc_lde_column(const uint64_t* trace, size_t n, uint64_t toff,
                   const uint64_t* rand, size_t k, uint64_t goff, size_t m, uint64_t* out) {
    if (log2_exact(n) < 0 || log2_exact(m) < 0) return -1;
    if (!rand) k = 0;
    size_t len;
    uint64_t* p = randomized_interpolant(trace, n, toff, rand, k, &len);
    if (!p) return -2;
    /* :117-152: chunks of m coefficients, chunk c scaled by offset^(c*m) */
    uint64_t* tmp = (uint64_t*)malloc(m * sizeof(uint64_t));
    if (!tmp) { free(p); return -2; }
    for (size_t cs = 0; cs < len || cs == 0; cs += m) {
        size_t cn = len - cs < m ? len - cs : m;
        if (cs == 0) {
            orc_evaluate_on_coset(p, cn, m, goff, out);
        } else {
            orc_evaluate_on_coset(p + cs, cn, m, goff, tmp);
            uint64_t sc = orc_pow(goff, (uint64_t)cs);
            for (size_t i = 0; i < m; i++) out[i] = orc_add(out[i], orc_mul(tmp[i], sc));
        }
        if (len == 0) break;
    }
    free(tmp);
    free(p);
    return 0;
}

int orc_lde_limb_truncating(const uint64_t* trace, size_t n, uint64_t toff,
                            const uint64_t* rand, size_t k, uint64_t goff, size_t m,
                            uint64_t* out) {
    if (log2_exact(n) < 0 || log2_exact(m) < 0) return -1;
    if (!rand) k = 0;
    size_t len;
    uint64_t* p = randomized_interpolant(trace, n, toff, rand, k, &len);
    if (!p) return -2;
    orc_evaluate_on_coset(p, len < m ? len : m, m, goff, out);   /* :270-300 */
    free(p);
    return 0;
}

/* LDE values at sampled rows only: out[c * nrows + q] = p_c(goff * w_m^rows[q]),
 * with p_c the randomized interpolant of column c (the polynomial whose coset
 * evaluation lde_randomized.cpp:71-154 computes, chunking included) or, with
 * truncate, its first m coefficients (the XFE rule, :270-300).  Horner at each
 * point; columns in parallel (OpenMP).  Returns 0, or -1 for bad lengths. */
int orc_lde_sample_colmajor(const uint64_t* trace, size_t n, size_t cols, uint64_t toff,
                            const uint64_t* rand, size_t k, uint64_t goff, size_t m,
                            const uint64_t* rows, size_t nrows, int truncate, uint64_t* out) {
    int lm = log2_exact(m);
    if (log2_exact(n) < 0 || lm < 0) return -1;
    if (!rand) k = 0;
    uint64_t wm = orc_root_of_unity((uint32_t)lm);
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 1)
    for (long c = 0; c < (long)cols; c++) {
        size_t len;
        uint64_t* p = randomized_interpolant(trace + (size_t)c * n, n, toff,
                                             rand ? rand + (size_t)c * k : NULL, k, &len);
        if (!p) { bad = 1; continue; }
        size_t use = truncate && len > m ? m : len;
        for (size_t q = 0; q < nrows; q++) {
            uint64_t x = orc_mul(goff, orc_pow(wm, rows[q] % m));
            uint64_t acc = 0;
            for (size_t i = use; i-- > 0;) acc = orc_add(orc_mul(acc, x), p[i]);
            out[(size_t)c * nrows + q] = acc;
        }
        free(p);
    }
    return bad ? -2 : 0;
}

int orc_lde_table_colmajor(const uint64_t* trace, size_t n, size_t cols, uint64_t toff,
                           const uint64_t* rand, size_t k, uint64_t goff, size_t m,
                           uint64_t* out) {
    /* lde_table.cpp:8-56 (per column; randomizers are rand[c*k + i]) */
    for (size_t c = 0; c < cols; c++) {
        int rc = orc_lde_column(trace + c * n, n, toff, rand ? rand + c * k : NULL, k,
                                goff, m, out + c * m);
        if (rc) return rc;
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Tip5 (src/hash/tip5.cpp).  Constants are the Tip5 specification's tables: */
/* the lookup table equals ((x+1)^3 mod 257) - 1, the MDS first column is     */
/* SHA-256("Tip5") read as 16 little-endian u16, and the 80 round constants   */
/* are the specification's values (tip5.cpp:48-79).                           */
/* ------------------------------------------------------------------------- */

static uint8_t LUT[256];
static int lut_ready = 0;

static void lut_init(void) {
    if (lut_ready) return;
    for (int x = 0; x < 256; x++) {
        int t = x + 1;
        LUT[x] = (uint8_t)(((t * t % 257) * t % 257) - 1);
    }
    lut_ready = 1;
}

static const int64_t MDS[16] = {
    61402, 1108, 28750, 33823, 7454, 43244, 53865, 12034,
    56951, 27521, 41351, 40901, 12021, 59689, 26798, 17845,
};

static const uint64_t RC[80] = {
    /* round 0 */
    0xbd2a3deb61ab60deull, 0xea7df21ad9547ed2ull, 0x900b3677a1de063full, 0x1b46887e876c8677ull,
    0xd364d977889cfb97ull, 0xdc8dfac843699f02ull, 0x375c405d7190db58ull, 0x27924006d2b0d4b1ull,
    0x78dd1172d483cd38ull, 0x3346c66244882a56ull, 0xb0249b279f498aa5ull, 0x94cd51be79338d4dull,
    0xb0e0dc7052c5b218ull, 0xf8dcc4d248adad95ull, 0x68e3c635fec868b7ull, 0xd7d06b3ffb6b0d8cull,
    /* round 1 */
    0xf3500dea20ef032aull, 0x4865bf175bba5803ull, 0xd5f7fe3027287a27ull, 0xa57333f44e193412ull,
    0x8726e153a977eae2ull, 0x3014a98463fc191bull, 0xba145461af39b212ull, 0x03ab70105933202full,
    0x3d90b7eebfcf71e5ull, 0x386322b1cc520bfdull, 0x27c2c8daf774f675ull, 0x4fcb83f50309bc6aull,
    0x5e6d5ce8275f3cb3ull, 0xecc2f6592c8f905cull, 0x837f532461e609b4ull, 0xb2b1f6b95c92c93cull,
    /* round 2 */
    0xc0027af556411dc1ull, 0x16e18c885fc2a26cull, 0x8880ef183d9f2bf3ull, 0xb2930bdb5ca88c45ull,
    0x9c2ec8322e1c1553ull, 0xe5b05eaf3220a674ull, 0xa49cc6ae4b861c4eull, 0x11708e0aeb86ebd7ull,
    0xc09de92bbc3902e0ull, 0x929b3c79516bcbc1ull, 0xe006e5bf738f27d1ull, 0x2d9e1ec0eac8ea38ull,
    0x0984d8d94bf937c5ull, 0x4959273c220e6747ull, 0xfe1d934207e796faull, 0x2b9b9298f2f6dd73ull,
    /* round 3 */
    0x07a1f5a67d6e3a41ull, 0x4407593ee73743d9ull, 0x9f054720ef802e59ull, 0x78d4b711336e6aa6ull,
    0xadc638aef3c8b228ull, 0xa4d6d3e86afb2114ull, 0x9d4808e725531968ull, 0x369804df3866d0efull,
    0xe6dbd9a9d2215024ull, 0x8ed22ca212ee85b2ull, 0x397bb882fcd23eb6ull, 0xeb8f8786d7277531ull,
    0x9999d4cdaff543b5ull, 0xf382a61217f192d6ull, 0x49c37260b026adc1ull, 0x3ff8918ce35c1019ull,
    /* round 4 */
    0x2e7df8b76080bd07ull, 0xf5dbac250b8a28b9ull, 0x853c3727ae9da4ccull, 0xb2f1f5f3d9e5a26dull,
    0x3fce22012d337847ull, 0x6b5a3e6db7eee347ull, 0x171582cd59dde50dull, 0xc0c0b3095ee62a8aull,
    0x665b25c6f6a203d2ull, 0x3099aed93b6ae69full, 0x801df6092be69c38ull, 0x8066ad0cdfff43cdull,
    0x8af9d44a5f4fdc6bull, 0xd80219cd97c0d762ull, 0x10c9ceba14148ebbull, 0x539bd4c3f2f24474ull,
};

const uint8_t* orc_tip5_lookup_table(void) { lut_init(); return LUT; }
const uint64_t* orc_tip5_round_constants(void) { return RC; }
const int64_t* orc_tip5_mds_first_column(void) { return MDS; }

static uint64_t split_and_lookup(uint64_t v) {   /* tip5.cpp:155-179 */
    const uint64_t R2 = 0xFFFFFFFE00000001ull;
    u128 pr = (u128)v * R2;
    uint64_t m = orc_montyred((uint64_t)(pr >> 64), (uint64_t)pr);
    uint64_t o = 0;
    for (int i = 0; i < 8; i++) o |= (uint64_t)LUT[(m >> (8 * i)) & 0xff] << (8 * i);
    return orc_montyred(0, o) % ORC_P;
}

void orc_tip5_round(uint64_t st[16], int r) {   /* tip5.cpp:115-125 */
    lut_init();
    {
        for (int i = 0; i < 4; i++) st[i] = split_and_lookup(st[i]);   /* :121-126 */
        for (int i = 4; i < 16; i++) {                                  /* :128-132 */
            uint64_t sq = orc_mul(st[i], st[i]);
            uint64_t qu = orc_mul(sq, sq);
            st[i] = orc_mul(orc_mul(st[i], sq), qu);
        }
        uint64_t ns[16];                                                /* :181-196 */
        for (int row = 0; row < 16; row++) {
            uint64_t acc = 0;
            for (int col = 0; col < 16; col++)
                acc = orc_add(acc, orc_mul((uint64_t)MDS[(row - col) & 15], st[col]));
            ns[row] = acc;
        }
        for (int i = 0; i < 16; i++) st[i] = orc_add(ns[i], RC[r * 16 + i]);  /* :198-203 */
    }
}

void orc_tip5_permutation(uint64_t st[16]) {      /* tip5.cpp:92-96 */
    for (int r = 0; r < 5; r++) orc_tip5_round(st, r);
}

void orc_hash_10(const uint64_t in[10], uint64_t out[5]) {   /* tip5.cpp:224-244 */
    uint64_t st[16];
    for (int i = 0; i < 10; i++) st[i] = in[i];
    for (int i = 10; i < 16; i++) st[i] = 1;
    orc_tip5_permutation(st);
    memcpy(out, st, 5 * sizeof(uint64_t));
}

void orc_hash_varlen(const uint64_t* in, size_t len, uint64_t out[5]) {   /* tip5.cpp:246-282 */
    uint64_t st[16] = {0};
    size_t used = 0;
    while (used + 10 <= len) {
        for (int i = 0; i < 10; i++) st[i] = in[used + i];
        orc_tip5_permutation(st);
        used += 10;
    }
    size_t rest = len - used;
    for (size_t i = 0; i < 10; i++) st[i] = i < rest ? in[used + i] : (i == rest ? 1 : 0);
    orc_tip5_permutation(st);
    memcpy(out, st, 5 * sizeof(uint64_t));
}

void orc_hash_pair(const uint64_t l[5], const uint64_t r[5], uint64_t out[5]) {  /* tip5.cpp:284-297 */
    uint64_t in[10];
    memcpy(in, l, 5 * sizeof(uint64_t));
    memcpy(in + 5, r, 5 * sizeof(uint64_t));
    orc_hash_10(in, out);
}

void orc_hash_rows_colmajor(const uint64_t* t, size_t rows, size_t cols, uint64_t* dg) {
    /* table_commitment.cpp:9-18,61-63 over a column-major table */
    uint64_t* row = (uint64_t*)malloc((cols ? cols : 1) * sizeof(uint64_t));
    for (size_t r = 0; r < rows; r++) {
        for (size_t c = 0; c < cols; c++) row[c] = t[c * rows + r];
        orc_hash_varlen(row, cols, dg + 5 * r);
    }
    free(row);
}

/* ------------------------------------------------------------------------- */
/* Merkle (src/merkle/merkle_tree.cpp:12-61,93-113,228-265,343-364).         */
/* ------------------------------------------------------------------------- */

int orc_merkle_heap(const uint64_t* leaves, size_t n, uint64_t* nodes) {
    if (log2_exact(n) < 0) return -1;
    memset(nodes, 0, 5 * sizeof(uint64_t));
    memcpy(nodes + 5 * n, leaves, 5 * n * sizeof(uint64_t));
    for (size_t i = n - 1; i >= 1; i--)
        orc_hash_pair(nodes + 5 * (2 * i), nodes + 5 * (2 * i + 1), nodes + 5 * i);
    return 0;
}

int orc_merkle_level_order(const uint64_t* leaves, size_t n, uint64_t* tree) {
    if (log2_exact(n) < 0) return -1;
    memcpy(tree, leaves, 5 * n * sizeof(uint64_t));
    size_t cur = 0, nxt = n, sz = n;
    while (sz > 1) {
        for (size_t i = 0; i < sz / 2; i++)
            orc_hash_pair(tree + 5 * (cur + 2 * i), tree + 5 * (cur + 2 * i + 1),
                          tree + 5 * (nxt + i));
        cur = nxt;
        nxt += sz / 2;
        sz /= 2;
    }
    return 0;
}

static int cmp_desc(const void* a, const void* b) {
    size_t x = *(const size_t*)a, y = *(const size_t*)b;
    return x < y ? 1 : (x > y ? -1 : 0);
}

long orc_auth_structure_indices(size_t n, const size_t* leaf_idx, size_t cnt,
                                size_t* out, size_t cap) {
    /* merkle_tree.cpp:228-265: needed siblings minus computable nodes, descending */
    if (log2_exact(n) < 0) return -1;
    unsigned char* need = (unsigned char*)calloc(2 * n, 1);
    unsigned char* comp = (unsigned char*)calloc(2 * n, 1);
    if (!need || !comp) { free(need); free(comp); return -2; }
    for (size_t j = 0; j < cnt; j++) {
        if (leaf_idx[j] >= n) { free(need); free(comp); return -3; }
        size_t v = leaf_idx[j] + n;
        while (v > 1) { comp[v] = 1; need[v ^ 1] = 1; v >>= 1; }
    }
    long c = 0;
    for (size_t v = 2 * n - 1; v >= 2; v--)
        if (need[v] && !comp[v]) {
            if ((size_t)c < cap) out[c] = v;
            c++;
        }
    free(need);
    free(comp);
    if ((size_t)c <= cap) qsort(out, (size_t)c, sizeof(size_t), cmp_desc);
    return c;
}

int orc_merkle_verify_path(const uint64_t root[5], size_t idx, const uint64_t leaf[5],
                           const uint64_t* path, size_t plen) {
    uint64_t cur[5], nx[5];
    memcpy(cur, leaf, sizeof cur);
    for (size_t i = 0; i < plen; i++) {
        if (idx % 2 == 0) orc_hash_pair(cur, path + 5 * i, nx);
        else orc_hash_pair(path + 5 * i, cur, nx);
        memcpy(cur, nx, sizeof cur);
        idx /= 2;
    }
    return memcmp(cur, root, sizeof cur) == 0;
}

/* ------------------------------------------------------------------------- */
/* Whole commit (src/lde/lde_table.cpp:86-101).                               */
/* ------------------------------------------------------------------------- */

int orc_commit(const uint64_t* trace, size_t n, size_t cols, uint64_t toff,
               const uint64_t* rand, size_t k, uint64_t goff, size_t m,
               uint64_t* codewords, uint64_t* digests, uint64_t root[5]) {
    uint64_t* cw = codewords ? codewords : (uint64_t*)malloc(m * cols * sizeof(uint64_t));
    uint64_t* dg = digests ? digests : (uint64_t*)malloc(m * 5 * sizeof(uint64_t));
    uint64_t* nodes = (uint64_t*)malloc(2 * m * 5 * sizeof(uint64_t));
    int rc = -2;
    if (cw && dg && nodes) {
        rc = orc_lde_table_colmajor(trace, n, cols, toff, rand, k, goff, m, cw);
        if (!rc) {
            orc_hash_rows_colmajor(cw, m, cols, dg);
            rc = orc_merkle_heap(dg, m, nodes);
            if (!rc) memcpy(root, nodes + 5, 5 * sizeof(uint64_t));
        }
    }
    if (!codewords) free(cw);
    if (!digests) free(dg);
    free(nodes);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* Inputs: SplitMix64 (public algorithm) with rejection of values >= p.       */
/* ------------------------------------------------------------------------- */

void orc_splitmix_fill(uint64_t seed, uint64_t* out, size_t count) {
    uint64_t s = seed;
    size_t i = 0;
    while (i < count) {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        if (z < ORC_P) out[i++] = z;
    }
}
