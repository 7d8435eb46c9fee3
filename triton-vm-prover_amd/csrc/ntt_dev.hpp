// Batched Goldilocks NTT and coset low-degree extension for gfx950.
//
// Math restated from src/ntt/ntt.cpp:26-131 and src/lde/lde_randomized.cpp:
// the LDE of a column x of length N onto the coset g*<w_M>, M = E*N, is
//   out[E*q + s] = NTT_N(f_s)[q],  f_s[r] = sum_{i = r mod N} p_i (g w_M^s)^i
// where p = INTT_N(x) * o^-i + (X^N - o^N) * rand(X) (o = trace offset).  The
// E cosets replace the reference's zero-pad + size-M NTT (same exact values,
// a quarter of the butterflies at E = 4).
//
// Each size-N transform is a four-step split N = N1 * N2 (N1 = 2^a):
//   step 1: N1-point DFTs down the columns of the N1 x N2 view, twiddle w_N^(n2 k1)
//   step 2: N2-point DFTs along rows (k1 fixed)
// The LDE fuses INTT step 2, the coset scaling of all E cosets and forward
// step 1 into one kernel (k_lde_step2), so a column is read from HBM three
// times and written three times, with d_out as the only working memory:
//   K1  trace            -> Y   at out[E*N2*k1 + n2]
//   K2  Y (rows k1)      -> Z_s at out[E*(u1 + N2*k1) + s]   (in place per row)
//   K3  Z (u1-tile, all s) -> codewords at out[E*(u1 + N2*u2) + s] (in place)
// Sub-DFTs of length L <= 4096 run in LDS as Stockham radix-16 passes: each
// thread holds R values, applies w_(Ns*R)^(k*r) from a table and an R-point
// DFT whose twiddles are powers of two (w_64 = 2^39), natural order in/out.
//
// This header holds the device code shared by the translation units that
// instantiate the kernels (ntt_step1.hip, lde_rows.hip, lde_expand.hip,
// lde_step3.hip); the host side (plans, constant tables, launches) is
// ntt_kernels.hip.
#pragma once

#include <utility>

#include "tvh_internal.hpp"

namespace tvh {

constexpr int TPB = 1024;    // threads per block of every NTT kernel (16 waves)
constexpr int TILE = 16384;  // u64 elements per block: 16 per thread, 128 KB before padding
constexpr int LOG_TILE = 14;

__device__ __forceinline__ int pidx(int t, int i, int Lp) { return t * Lp + i + (i >> 4); }
__host__ __device__ inline int padded_len(int L) { return L + (L >> 4) + 1; }

template <int LOGR>
__device__ __forceinline__ constexpr int brev(int i) {
    int r = 0;
    for (int b = 0; b < LOGR; b++)
        if (i & (1 << b)) r |= 1 << (LOGR - 1 - b);
    return r;
}

// R-point DFT in registers, natural order in and out: radix-2 DIT on the
// bit-reversed input.  The butterflies of stage LEN use w_LEN^j =
// 2^(39*(64/LEN)*j mod 192) (w_64 = 2^39); a twiddle 2^e with e >= 96 is
// -2^(e-96), folded into the butterfly by swapping add and sub; 2^e itself is
// applied by word shifts (gl::mul_pow2).  Every index and exponent is a
// template constant (index sequences), so the whole DFT is straight-line code.
template <bool INV, int LEN, int J>
__device__ __forceinline__ void dft_bfly(u64& x, u64& y) {
    constexpr int e0 = (39 * (64 / LEN) * J) % 192;
    constexpr int e = INV ? (192 - e0) % 192 : e0;
    constexpr bool negw = e >= 96;
    constexpr int ep = negw ? e - 96 : e;
    u64 t = ep == 0 ? y : gl::mul_pow2<ep>(y);
    u64 u = x;
    if (negw) {
        x = gl::sub(u, t);
        y = gl::add(u, t);
    } else {
        x = gl::add(u, t);
        y = gl::sub(u, t);
    }
}

template <bool INV, int LEN, size_t... K>
__device__ __forceinline__ void dft_stage(u64* a, std::index_sequence<K...>) {
    // butterfly K: group K / (LEN/2), offset j = K % (LEN/2)
    (dft_bfly<INV, LEN, (int)(K % (LEN / 2))>(a[(K / (LEN / 2)) * LEN + K % (LEN / 2)],
                                             a[(K / (LEN / 2)) * LEN + K % (LEN / 2) + LEN / 2]),
     ...);
}

template <int LOGR, bool INV, size_t... S>
__device__ __forceinline__ void dft_stages(u64* a, std::index_sequence<S...>) {
    (dft_stage<INV, (2 << S)>(a, std::make_index_sequence<(1 << LOGR) / 2>{}), ...);
}

template <int LOGR, bool INV>
__device__ __forceinline__ void dft_reg(u64* v) {
    constexpr int R = 1 << LOGR;
    u64 a[R];
#pragma unroll
    for (int i = 0; i < R; i++) a[i] = v[brev<LOGR>(i)];
    dft_stages<LOGR, INV>(a, std::make_index_sequence<LOGR>{});
#pragma unroll
    for (int i = 0; i < R; i++) v[i] = a[i];
}

template <int LOGR, bool INV, int NTPB = TPB, int NTILE = TILE>
__device__ __forceinline__ void stockham_pass(u64* buf, int logL, int Lp, int NT, int Ns,
                                              const u64* __restrict__ r4096) {
    constexpr int R = 1 << LOGR;
    constexpr int MAXIT = NTILE / (R * NTPB);
    const int per = 1 << (logL - LOGR);
    const int total = NT << (logL - LOGR);
    const int step = 4096 / (Ns * R);
    static_assert(MAXIT >= 1, "radix larger than the per-thread element count");
    u64 v[MAXIT][R];
#pragma unroll
    for (int it = 0; it < MAXIT; it++) {
        int w = threadIdx.x + it * NTPB;
        if (w < total) {
            int t = w >> (logL - LOGR);
            int j = w & (per - 1);
            int k = j & (Ns - 1);
#pragma unroll
            for (int r = 0; r < R; r++) v[it][r] = buf[pidx(t, j + r * per, Lp)];
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) v[it][r] = gl::mont_mul(v[it][r], r4096[k * r * step]);
            }
            dft_reg<LOGR, INV>(v[it]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < MAXIT; it++) {
        int w = threadIdx.x + it * NTPB;
        if (w < total) {
            int j = (w) & (per - 1);
            int t = w >> (logL - LOGR);
            int k = j & (Ns - 1);
            int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) buf[pidx(t, base + r * Ns, Lp)] = v[it][r];
        }
    }
    __syncthreads();
}

// NT transforms of length 2^logL (<= 4096) in LDS; NT * L <= NTILE; NTPB
// threads hold NTILE / NTPB values each.
template <bool INV, int NTPB = TPB, int NTILE = TILE>
__device__ __forceinline__ void lds_dft(u64* buf, int logL, int NT, const u64* __restrict__ r4096) {
    const int Lp = padded_len(1 << logL);
    int Ns = 1;
    int rem = logL;
    while (rem >= 4) {
        stockham_pass<4, INV, NTPB, NTILE>(buf, logL, Lp, NT, Ns, r4096);
        Ns <<= 4;
        rem -= 4;
    }
    if (rem == 3) stockham_pass<3, INV, NTPB, NTILE>(buf, logL, Lp, NT, Ns, r4096);
    else if (rem == 2) stockham_pass<2, INV, NTPB, NTILE>(buf, logL, Lp, NT, Ns, r4096);
    else if (rem == 1) stockham_pass<1, INV, NTPB, NTILE>(buf, logL, Lp, NT, Ns, r4096);
}

// ---------------------------------------------------------------- fused-endpoint DFT
// The same Stockham passes as lds_dft, but the first pass takes its inputs
// straight from global memory (a Load functor) and the last pass writes its
// outputs straight to global memory (a Store functor): per transform one LDS
// write and read per middle pass instead of a global->LDS copy before and an
// LDS->global copy after, and 2 barriers per middle pass only.  A one-pass
// transform (L <= 16) never touches LDS.
//
// Work item (t, j) of a radix-R pass: v[r] = x_t[j + r*per] (times the
// Stockham twiddle when Ns > 1), R-point DFT, y_t[(j - k)*R + k + r*Ns] = v[r]
// with k = j mod Ns.  Functors: load(t, j, per, v) fills v[0..R); store(t, j,
// Ns, v) writes it.  TF (transform-fastest): adjacent threads take adjacent
// transforms (column kernels, whose transforms are adjacent in memory);
// otherwise adjacent elements of one transform (row kernels).
struct LdsIO {
    u64* buf;
    int Lp;
    template <int R>
    __device__ __forceinline__ void load(int t, int j, int per, u64 (&v)[R]) const {
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = buf[pidx(t, j + r * per, Lp)];
    }
    template <int R>
    __device__ __forceinline__ void store(int t, int j, int Ns, const u64 (&v)[R]) const {
        int k = j & (Ns - 1);
        int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) buf[pidx(t, base + r * Ns, Lp)] = v[r];
    }
};

// Workgroup barrier for LDS hand-offs only: waits for this wave's LDS
// operations, not for its global loads (the __syncthreads fence also drains
// the vector-memory counter, which would expose the latency of twiddle loads
// issued ahead of the barrier).  No thread of these kernels reads global
// memory that another thread of the block writes before the barrier.
__device__ __forceinline__ void lds_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Stockham twiddles of one pass for the thread's work items, loaded ahead of
// the pass: tw.t[it][r] = r4096[k * r * step] (r >= 1) for work item it.
template <int LOGR, int NTPB, int NTILE>
struct PassTw {
    static constexpr int R = 1 << LOGR;
    static constexpr int MAXIT = NTILE / (R * NTPB);
    u64 t[MAXIT][R];
};

template <int LOGR, int NTPB, int NTILE, bool TF>
__device__ __forceinline__ void prefetch_tw(PassTw<LOGR, NTPB, NTILE>& tw, int logL, int NT,
                                            int logNT, int Ns, const u64* __restrict__ r4096) {
    constexpr int R = 1 << LOGR;
    constexpr int MAXIT = NTILE / (R * NTPB);
    const int lp = logL - LOGR;
    const int per = 1 << lp;
    const int total = NT << lp;
    const int step = 4096 / (Ns * R);
#pragma unroll
    for (int it = 0; it < MAXIT; it++) {
        int w = threadIdx.x + it * NTPB;
        if (w < total) {
            int j = TF ? (w >> logNT) : (w & (per - 1));
            int k = j & (Ns - 1);
#pragma unroll
            for (int r = 1; r < R; r++) tw.t[it][r] = r4096[k * r * step];
        }
    }
}

template <int LOGR, bool INV, int NTPB, int NTILE, bool TF, bool FROM_LDS, bool TO_LDS, class Ld,
          class St>
__device__ __forceinline__ void fused_pass(const LdsIO& io, int logL, int NT, int logNT, int Ns,
                                           const PassTw<LOGR, NTPB, NTILE>& tw, const Ld& ld,
                                           const St& st) {
    constexpr int R = 1 << LOGR;
    constexpr int MAXIT = NTILE / (R * NTPB);
    static_assert(MAXIT >= 1, "radix larger than the per-thread element count");
    const int lp = logL - LOGR;
    const int per = 1 << lp;
    const int total = NT << lp;
    u64 v[MAXIT][R];
#pragma unroll
    for (int it = 0; it < MAXIT; it++) {
        int w = threadIdx.x + it * NTPB;
        if (w < total) {
            int t = TF ? (w & (NT - 1)) : (w >> lp);
            int j = TF ? (w >> logNT) : (w & (per - 1));
            if (FROM_LDS) io.load<R>(t, j, per, v[it]);
            else ld.template load<R>(t, j, per, v[it]);
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) v[it][r] = gl::mont_mul(v[it][r], tw.t[it][r]);
            }
            dft_reg<LOGR, INV>(v[it]);
        }
    }
    if (FROM_LDS && TO_LDS) lds_barrier();  // in place: every read of the pass before any write
#pragma unroll
    for (int it = 0; it < MAXIT; it++) {
        int w = threadIdx.x + it * NTPB;
        if (w < total) {
            int t = TF ? (w & (NT - 1)) : (w >> lp);
            int j = TF ? (w >> logNT) : (w & (per - 1));
            if (TO_LDS) io.store<R>(t, j, Ns, v[it]);
            else st.template store<R>(t, j, Ns, v[it]);
        }
    }
}

// Straight-line pass sequence for L = 2^(4Q + REM): Q radix-16 passes, then a
// radix-2^REM pass (a radix-1 copy when L = 1).  Every radix, stride and
// first/last role is a compile-time constant, which keeps register pressure
// at one pass's worth.  The twiddles of pass P + 1 are loaded before the
// barrier that ends pass P, so their latency overlaps the barrier wait.
// The last pass maps threads by TFL, the others by TF: a row transform whose
// output is stored transposed (adjacent rows adjacent in memory) loads its
// rows element-fastest and stores them transform-fastest, both coalesced.
template <int P, int NP, int Q, int REM, bool INV, int NTPB, int NTILE, bool TF, bool TFL,
          class Ld, class St>
__device__ __forceinline__ void fused_chain(const LdsIO& io, int NT, int logNT,
                                            const u64* __restrict__ r4096, const Ld& ld,
                                            const St& st,
                                            const PassTw<(P < Q ? 4 : REM), NTPB, NTILE>& tw) {
    constexpr int logL = 4 * Q + REM;
    constexpr bool tf = P + 1 == NP ? TFL : TF;
    fused_pass<(P < Q ? 4 : REM), INV, NTPB, NTILE, tf, (P > 0), (P + 1 < NP)>(
        io, logL, NT, logNT, 1 << (4 * P), tw, ld, st);
    if constexpr (P + 1 < NP) {
        constexpr bool tf_next = P + 2 == NP ? TFL : TF;
        PassTw<(P + 1 < Q ? 4 : REM), NTPB, NTILE> next;
        prefetch_tw<(P + 1 < Q ? 4 : REM), NTPB, NTILE, tf_next>(next, logL, NT, logNT,
                                                                1 << (4 * (P + 1)), r4096);
        lds_barrier();
        fused_chain<P + 1, NP, Q, REM, INV, NTPB, NTILE, TF, TFL>(io, NT, logNT, r4096, ld, st,
                                                                  next);
    }
}

template <int Q, int REM, bool INV, int NTPB, int NTILE, bool TF, bool TFL, class Ld, class St>
__device__ __forceinline__ void fused_passes(const LdsIO& io, int NT, int logNT,
                                             const u64* __restrict__ r4096, const Ld& ld,
                                             const St& st) {
    constexpr int NP = Q + ((REM != 0 || Q == 0) ? 1 : 0);
    PassTw<(0 < Q ? 4 : REM), NTPB, NTILE> none;  // the first pass has no twiddles
    fused_chain<0, NP, Q, REM, INV, NTPB, NTILE, TF, TFL>(io, NT, logNT, r4096, ld, st, none);
}

// LDS row pitch of a transform for the fused DFT: L + L/16 + pad, with pad
// chosen so that the lanes of a 16-lane LDS cycle hit distinct banks under the
// transform-fastest mapping (lane = j * NT' + t, NT' = min(NT, 16)):
// t * pitch + j is then distinct mod 16 for t < NT', j < 16 / NT'.
__host__ __device__ inline int fused_pitch(int L, int NT, bool tf) {
    int ntp = NT < 16 ? NT : 16;
    return L + (L >> 4) + (tf ? 16 / ntp : 1);
}

// NT (power of two) transforms of length 2^logL (<= 4096), NT * L <= NTILE.
// The block's LDS (buf, NT * fused_pitch(L, NT, TF || TFL) u64) must not be in
// use by other threads on entry.  TFL: thread mapping of the last pass.
template <bool INV, int NTPB, int NTILE, bool TF, bool TFL = TF, class Ld, class St>
__device__ __forceinline__ void fused_dft(u64* buf, int logL, int NT, int logNT,
                                          const u64* __restrict__ r4096, const Ld& ld,
                                          const St& st) {
    const LdsIO io{buf, fused_pitch(1 << logL, NT, TF || TFL)};
#define TVH_FUSED(q, r) \
    fused_passes<q, r, INV, NTPB, NTILE, TF, TFL>(io, NT, logNT, r4096, ld, st)
    switch (logL) {
        case 0: TVH_FUSED(0, 0); break;
        case 1: TVH_FUSED(0, 1); break;
        case 2: TVH_FUSED(0, 2); break;
        case 3: TVH_FUSED(0, 3); break;
        case 4: TVH_FUSED(1, 0); break;
        case 5: TVH_FUSED(1, 1); break;
        case 6: TVH_FUSED(1, 2); break;
        case 7: TVH_FUSED(1, 3); break;
        case 8: TVH_FUSED(2, 0); break;
        case 9: TVH_FUSED(2, 1); break;
        case 10: TVH_FUSED(2, 2); break;
        case 11: TVH_FUSED(2, 3); break;
        default: TVH_FUSED(3, 0); break;  // 12
    }
#undef TVH_FUSED
}

// w_N^(+-x) for x < N via the two-level 2^24 tables, in Montgomery form (the
// tables hold w * 2^64 mod p): apply with gl::mont_mul.
__device__ __forceinline__ u64 twiddle(const Roots& rt, int dir, u32 x, u32 logN) {
    u32 y = x << (kMaxLogN - logN);
    return gl::mont_mul(rt.lo[dir][y & 4095], rt.hi[dir][y >> 12]);
}

// Per-block segment lookup for multi-table launches: start[i] is the first
// block of segment i (start[nseg] = total blocks); returns the segment of blk.
__device__ __forceinline__ u32 find_segment(const u32* __restrict__ start, u32 nseg, u32 blk) {
    u32 lo = 0, hi = nseg;  // invariant: start[lo] <= blk < start[hi]
    while (hi - lo > 1) {
        u32 mid = (lo + hi) >> 1;
        if (start[mid] <= blk) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Transforms s = (c, n2): element i at src[c*sstride + (i << b) + n2].
// Addressing in the column and row functors: every access is a block-uniform
// base (the first transform of the block) plus a 32-bit element offset, so the
// loads and stores take a scalar base and a vector offset and no 64-bit
// address arithmetic runs per element.  Offsets stay below 2^32 elements: a
// block's transforms span at most its tile's worth of columns, and the
// in-column offset is below M <= 2^27.
struct Step1Load {
    const u64* src;
    u64 sstride, s0, ntr;
    u32 b;
    template <int R>
    __device__ __forceinline__ void load(int t, int j, int per, u64 (&v)[R]) const {
        u64 s = s0 + t;
        if (s < ntr) {
            const u64 mask = (1ull << b) - 1;
            const u64* base = src + (s0 >> b) * sstride + (s0 & mask);
            const u32 o = (u32)(((s >> b) - (s0 >> b)) * sstride + (s & mask) - (s0 & mask));
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = base[o + ((u32)(j + r * per) << b)];
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = 0;
        }
    }
};
// Step1Load with the coset scaling of a forward NTT (NttScale): element
// i = (n1 << b) + n2 of column c enters as (x + [i < k] rand[c*k + i] *
// delta) * base^i, base^i = p1[n1] * p2[n2] (p2[n2] is the same for all of a
// thread's elements).  The p1 entries are loaded with the elements, so all
// loads are in flight together; two products per element replace a separate
// read-and-write pass over the table.
struct Step1ScaleLoad {
    const u64* src;
    u64 sstride, s0, ntr;
    u32 b;
    NttScaleDev S;  // rand already offset to the launch's first column
    template <int R>
    __device__ __forceinline__ void load(int t, int j, int per, u64 (&v)[R]) const {
        u64 s = s0 + t;
        if (s < ntr) {
            const u64 mask = (1ull << b) - 1;
            const u64* base = src + (s0 >> b) * sstride + (s0 & mask);
            const u32 o = (u32)(((s >> b) - (s0 >> b)) * sstride + (s & mask) - (s0 & mask));
            const u32 n2 = (u32)(s & mask);
            const u64 f2 = S.p2[n2];
            u64 f1[R];
#pragma unroll
            for (int r = 0; r < R; r++) {
                v[r] = base[o + ((u32)(j + r * per) << b)];
                f1[r] = S.p1[j + r * per];
            }
            if (n2 < S.k) {  // i >= n2: only such transforms hold randomized elements
                const u64 c = s >> b;
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const u64 i = ((u64)(j + r * per) << b) + n2;
                    if (i < S.k) v[r] = gl::add(v[r], gl::mul(S.rand[c * S.k + i], S.delta));
                }
            }
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = gl::mont_mul(gl::mont_mul(v[r], f1[r]), f2);
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = 0;
        }
    }
};
// Output k1 of transform (c, n2), times w_N^(+-n2*k1), to dst[c*dstride + E*N2*k1 + n2].
template <bool INV>
struct Step1Store {
    u64* dst;
    u64 dstride, s0, ntr;
    u32 b, E, logN;
    Roots rt;
    template <int R>
    __device__ __forceinline__ void store(int t, int j, int Ns, const u64 (&v)[R]) const {
        u64 s = s0 + t;
        if (s >= ntr) return;
        const u64 mask = (1ull << b) - 1;
        u64 n2 = s & mask;
        u64* p0 = dst + (s0 >> b) * dstride + (s0 & mask);
        const u32 o = (u32)(((s >> b) - (s0 >> b)) * dstride + n2 - (s0 & mask));
        int k = j & (Ns - 1), base = (j - k) * R + k;
        // table entries first (index 0 holds 1, so n2 = 0 or k1 = 0 needs no
        // branch), then the products: the loads are in flight together
        const u32 sh = kMaxLogN - logN;
        u64 tl[R], th[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            u32 y = (u32)(n2 * (u32)(base + r * Ns)) << sh;
            tl[r] = rt.lo[INV][y & 4095];
            th[r] = rt.hi[INV][y >> 12];
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            u32 k1 = (u32)(base + r * Ns);
            p0[o + ((E * k1) << b)] = gl::mont_mul(v[r], gl::mont_mul(tl[r], th[r]));
        }
    }
};

// ---------------------------------------------------------------- step 1
// Transforms s = (c, n2), element n1 at src[c*sstride + N2*n1 + n2]; DFT over n1
// (inverse if INV) -> k1; times w_N^(+-n2*k1); to dst[c*dstride + E*N2*k1 + n2].
// Adjacent threads take adjacent transforms (adjacent n2: contiguous in HBM).
template <bool INV, int LT = LOG_TILE>
__device__ __forceinline__ void step1_body(const Step1Args& A, u64 blk, const Roots& rt) {
    extern __shared__ u64 lds[];
    constexpr int NTPB = 1 << (LT - 4);
    const u32 b = A.logN - A.a;
    const u64 s0 = blk << A.logNT;
    const Step1Load ld{A.src, A.sstride, s0, A.ntr, b};
    const Step1Store<INV> st{A.dst, A.dstride, s0, A.ntr, b, A.E, A.logN, rt};
    fused_dft<INV, NTPB, 1 << LT, true>(lds, A.a, 1 << A.logNT, A.logNT, rt.r4096[INV], ld, st);
}

// Forward step 1 with the input scaling (Step1ScaleLoad).
template <int LT>
__device__ __forceinline__ void step1_scaled_body(const Step1Args& A, const NttScaleDev& S,
                                                  u64 blk, const Roots& rt) {
    extern __shared__ u64 lds[];
    constexpr int NTPB = 1 << (LT - 4);
    const u32 b = A.logN - A.a;
    const u64 s0 = blk << A.logNT;
    const Step1ScaleLoad ld{A.src, A.sstride, s0, A.ntr, b, S};
    const Step1Store<false> st{A.dst, A.dstride, s0, A.ntr, b, A.E, A.logN, rt};
    fused_dft<false, NTPB, 1 << LT, true>(lds, A.a, 1 << A.logNT, A.logNT, rt.r4096[0], ld, st);
}



// ---------------------------------------------------------------- plain NTT step 2
// Rows s = (c, k1): n2 at src[c*N + N2*k1 + n2]; DFT over n2 -> k2 (times N^-1
// when INV); to dst[c*N + k1 + N1*k2].

// ---------------------------------------------------------------- LDE step 2
__device__ __forceinline__ u64 rand_term(const LdeArgs& A, u64 c, u32 s, u64 r) {
    const RandSpec& rs = A.rs;
    u32 limb = (u32)(c % rs.group);
    if (!((rs.limb_mask >> limb) & 1)) return 0;
    const u64* p = rs.ptr + (c / rs.group) * rs.group_stride + limb * rs.limb_stride +
                   r * rs.elem_stride;
    // select delta_n[s] without a dynamic index (a dynamic index into the
    // by-value kernel argument would copy it to scratch)
    u64 d = A.delta_n[0];
#pragma unroll
    for (u32 i = 1; i < 16; i++)
        if (s == i) d = A.delta_n[i];
    return gl::mul(gl::mul(*p, d), gl::pow(A.toff, r));
}

// INTT = false: the rows already hold the inverse-transformed coefficients
// (written by k_lde_rowinv), only the coset expansion runs here.
template <bool INTT = true>
__device__ __forceinline__ void lde_step2_body(const LdeArgs& A, u64 blk, const Roots& rt) {
    extern __shared__ u64 lds[];
    u64* out = A.out;
    const u32 a = A.a, b = A.logN - A.a, E = A.E, logNT = A.logNT;
    const int L = 1 << b;
    const int Lp = padded_len(L);
    const int NT = 1 << logNT;           // rows per block
    const u64 nrows = A.cols << a;
    const u64 s0 = blk << logNT;
    // load Y rows into transforms 0..NT-1 (slot s = 0)
    for (int e = threadIdx.x; e < NT * L; e += TPB) {
        int t = e >> b, i = e & (L - 1);
        u64 s = s0 + t;
        if (s < nrows) {
            u64 c = s >> a, k1 = s & ((1ull << a) - 1);
            lds[pidx(t, i, Lp)] = A.src[c * A.src_col_stride + k1 * A.src_row_stride + i];
        }
    }
    __syncthreads();
    if (INTT) lds_dft<true>(lds, b, NT, rt.r4096[1]);
    // expand: transform (sl*NT + t) <- f_sl for row t
    for (int e = threadIdx.x; e < NT * L; e += TPB) {
        int t = e >> b, k2 = e & (L - 1);
        u64 s = s0 + t;
        if (s < nrows) {
            u64 c = s >> a, k1 = s & ((1ull << a) - 1);
            u64 C = lds[pidx(t, k2, Lp)];
            u64 r = k1 + ((u64)k2 << a);
            for (u32 sl = 0; sl < E; sl++) {
                u64 v = C;
                if (r < A.rs.k) v = gl::add(v, rand_term(A, c, sl, r));
                lds[pidx(sl * NT + t, k2, Lp)] = gl::mont_mul(v, A.pb[(sl << b) + k2]);
            }
        }
    }
    __syncthreads();
    lds_dft<false>(lds, b, NT * E, rt.r4096[0]);
    // write Z_s[u1][k1] to out[E*(u1 + N2*k1) + s], contiguous over (u1, s)
    // Output factor pa[s][k1] * w_N^(k1*u1) (N^-1 beta_s^k1 and the four-step
    // twiddle; pb was applied before the row DFT).
    const int logRW = b + A.logE;  // row of all cosets: E * N2 elements
    const int RW = 1 << logRW;
    if (RW == TILE) {
        // one row per block (k1 uniform): thread outputs q = tid + TPB*j have
        // the fixed coset sl = tid mod E and u1 = tid/E + (TPB/E)*j, so the
        // factor advances by w_N^(k1*TPB/E) per j
        u64 s = s0;
        if (s < nrows) {
            u64 c = s >> a, k1 = s & ((1ull << a) - 1);
            const int sl = threadIdx.x & (E - 1);
            const u32 u0 = threadIdx.x >> A.logE;
            u64 f = A.pa[(sl << a) + k1];
            if (k1 && u0) f = gl::mont_mul(f, twiddle(rt, 0, (u32)(k1 * u0), A.logN));
            const u64 step = twiddle(rt, 0, (u32)(k1 * (TPB >> A.logE)), A.logN);
            u64* dst = out + c * A.M + (((u64)E * k1) << b);
#pragma unroll
            for (int j = 0; j < TILE / TPB; j++) {
                int q = threadIdx.x + j * TPB;
                int u1 = q >> A.logE;
                dst[q] = gl::mont_mul(lds[pidx(sl, u1, Lp)], f);
                if (j + 1 < TILE / TPB) f = gl::mont_mul(f, step);
            }
        }
    } else {
        for (int e = threadIdx.x; e < NT * RW; e += TPB) {
            int t = e >> logRW, q = e & (RW - 1);
            int u1 = q >> A.logE, sl = q & (E - 1);
            u64 s = s0 + t;
            if (s < nrows) {
                u64 c = s >> a, k1 = s & ((1ull << a) - 1);
                u64 f = A.pa[(sl << a) + k1];
                if (u1 && k1) f = gl::mont_mul(f, twiddle(rt, 0, (u32)(k1 * u1), A.logN));
                out[c * A.M + (((u64)E * k1) << b) + q] =
                    gl::mont_mul(lds[pidx(sl * NT + t, u1, Lp)], f);
            }
        }
    }
}


// Split form of the fused row step: k_lde_rowinv inverse-transforms TILE / N2
// rows per block with every thread busy and writes the coefficients where the
// row was (out[c*M + E*N2*k1 + n2]); k_lde_expand then runs the coset
// expansion of k_lde_step2 on them.  One extra read and write of N elements
// per column buys full occupancy for the inverse row DFT, which the fused
// kernel runs with a quarter of its threads at E = 4.
// Rows s = (c, k1) of N2 elements at src[c*scs + k1*srs + i].
struct RowLoad {
    const u64* src;
    u64 scs, srs, s0, nrows;
    u32 a;
    template <int R>
    __device__ __forceinline__ void load(int t, int j, int per, u64 (&v)[R]) const {
        u64 s = s0 + t;
        if (s < nrows) {
            const u64 mask = (1ull << a) - 1;
            const u64* p = src + (s0 >> a) * scs + (s0 & mask) * srs;
            const u32 o = (u32)(((s >> a) - (s0 >> a)) * scs + ((s & mask) - (s0 & mask)) * srs);
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = p[o + (u32)(j + r * per)];
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = 0;
        }
    }
};
// Row (c, k1) to out[c*M + E*N2*k1 + i].
struct RowStore {
    u64* out;
    u64 M, s0, nrows;
    u32 a, b, E;
    template <int R>
    __device__ __forceinline__ void store(int t, int j, int Ns, const u64 (&v)[R]) const {
        u64 s = s0 + t;
        if (s >= nrows) return;
        const u64 mask = (1ull << a) - 1;
        u64* p = out + (s0 >> a) * M + (((u64)E * (s0 & mask)) << b);
        const u32 o = (u32)(((s >> a) - (s0 >> a)) * M + (((u64)E * ((s & mask) - (s0 & mask))) << b));
        int k = j & (Ns - 1), base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) p[o + (u32)(base + r * Ns)] = v[r];
    }
};

// Row tiles of 2^LT elements, 16 per thread (LT = 14: 1024 threads, one block
// per CU; LT = 13: 512 threads, two blocks per CU).
template <int LT = LOG_TILE>
__device__ __forceinline__ void lde_rowinv_body(const LdeArgs& A, u64 blk, const Roots& rt) {
    extern __shared__ u64 lds[];
    const u32 b = A.logN - A.a;
    const u64 nrows = A.cols << A.a;
    const u64 s0 = blk << A.logNT;  // rows per block of this kernel: 2^LT / N2
    const RowLoad ld{A.src, A.src_col_stride, A.src_row_stride, s0, nrows, A.a};
    const RowStore st{A.out, A.M, s0, nrows, A.a, b, A.E};
    fused_dft<true, (1 << (LT - 4)), (1 << LT), false>(lds, b, 1 << A.logNT, A.logNT,
                                                        rt.r4096[1], ld, st);
}



// Coset transforms T = t*E + sl of row t (coefficients at src rows, written
// by k_lde_rowinv): element k2 = (C_k2 + randomizer term) * beta_sl^(N1*k2).
// Cosets are the fastest index, so that with the transform-fastest thread
// mapping the E cosets of one output position are written by adjacent lanes
// (out[E*u1 + sl] contiguous).
struct ExpandLoad {
    const LdeArgs& A;
    u64 s0, nrows;
    template <int R>
    __device__ __forceinline__ void load(int T, int j, int per, u64 (&v)[R]) const {
        const u32 a = A.a, b = A.logN - A.a;
        u32 sl = (u32)T & (A.E - 1);
        u64 s = s0 + ((u32)T >> A.logE);
        if (s >= nrows) {
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = 0;
            return;
        }
        const u64 mask = (1ull << a) - 1;
        u64 c = s >> a, k1 = s & mask;
        const u64* p = A.src + (s0 >> a) * A.src_col_stride + (s0 & mask) * A.src_row_stride;
        const u32 o = (u32)((c - (s0 >> a)) * A.src_col_stride + (k1 - (s0 & mask)) * A.src_row_stride);
        const u32 ob = sl << b;
        // every load first, with no branch in between, so that they are in
        // flight together
        u64 f[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            v[r] = p[o + (u32)(j + r * per)];
            f[r] = A.pb[ob + (u32)(j + r * per)];
        }
        // randomizer terms: element k2 of row k1 is trace row k1 + k2 * N1,
        // which carries one when below rand_len (possible only for k1 < rand_len)
        if (k1 < A.rs.k) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                u64 rr = k1 + ((u64)(j + r * per) << a);
                if (rr < A.rs.k) v[r] = gl::add(v[r], rand_term(A, c, sl, rr));
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = gl::mont_mul(v[r], f[r]);
    }
};
// Output u1 of coset sl, row k1, times N^-1 beta_sl^k1 w_N^(k1*u1), to
// out[c*M + E*N2*k1 + E*u1 + sl].  On the last radix-16 pass of a full row the
// thread's outputs are u1 = j + Ns*r, so the factor advances by w_N^(k1*Ns).
struct ExpandStore {
    const LdeArgs& A;
    Roots rt;
    u64 s0, nrows;
    template <int R>
    __device__ __forceinline__ void store(int T, int j, int Ns, const u64 (&v)[R]) const {
        const u32 a = A.a, b = A.logN - A.a, E = A.E;
        u32 sl = (u32)T & (E - 1);
        u64 s = s0 + ((u32)T >> A.logE);
        if (s >= nrows) return;
        u64 k1 = s & ((1ull << a) - 1);
        // row s starts at out + s*E*N2 (c*M + k1*E*N2 with M = N1*E*N2)
        u64* dst = A.out + ((s0 * E) << b);
        const u32 o = (u32)(((s - s0) * E) << b) + sl;
        // (twiddle index 0 holds 1: no branch for k1 = 0 or u1 = 0)
        u64 f = A.pa[((u64)sl << a) + k1];
        int k = j & (Ns - 1), base = (j - k) * R + k;
        if (R * Ns == (1 << b)) {  // base = j: u1 = j + Ns*r
            f = gl::mont_mul(f, twiddle(rt, 0, (u32)(k1 * j), A.logN));
            const u64 step = twiddle(rt, 0, (u32)(k1 * Ns), A.logN);
            // factors f * step^r as C interleaved chains step^C apart: a
            // dependency depth of about R / C products instead of R - 1
            constexpr int C = R >= 4 ? 4 : R;
            u64 fr[C];
            fr[0] = f;
            u64 stride = step;
            if constexpr (C == 4) {
                const u64 step2 = gl::mont_sqr(step);
                fr[1] = gl::mont_mul(f, step);
                fr[2] = gl::mont_mul(f, step2);
                fr[3] = gl::mont_mul(fr[1], step2);
                stride = gl::mont_sqr(step2);
            } else if constexpr (C == 2) {
                fr[1] = gl::mont_mul(f, step);
                stride = gl::mont_sqr(step);
            }
#pragma unroll
            for (int r = 0; r < R; r++) {
                dst[o + E * (u32)(j + r * Ns)] = gl::mont_mul(v[r], fr[r % C]);
                if (r + C < R) fr[r % C] = gl::mont_mul(fr[r % C], stride);
            }
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) {
                u32 u1 = (u32)(base + r * Ns);
                u64 g = gl::mont_mul(f, twiddle(rt, 0, (u32)(k1 * u1), A.logN));
                dst[o + E * u1] = gl::mont_mul(v[r], g);
            }
        }
    }
};

// Split form, second kernel: for NT rows per block (NT * E * N2 = 2^LT), the E
// coset transforms straight from the coefficient rows to the codewords.  The
// outputs overwrite the coefficient row in place, so every coset of a row
// must be computed by the block that reads the row (its reads all precede its
// first barrier, its writes follow its last one).
template <int LT = LOG_TILE>
__device__ __forceinline__ void lde_expand_body(const LdeArgs& A, u64 blk, const Roots& rt) {
    extern __shared__ u64 lds[];
    const u32 b = A.logN - A.a;
    const u64 nrows = A.cols << A.a;
    const u64 s0 = blk << A.logNT;
    const ExpandLoad ld{A, s0, nrows};
    const ExpandStore st{A, rt, s0, nrows};
    fused_dft<false, (1 << (LT - 4)), (1 << LT), true>(
        lds, b, (int)(A.E << A.logNT), (int)(A.logE + A.logNT), rt.r4096[0], ld, st);
}




// ---------------------------------------------------------------- LDE step 3
// Transforms (c, q = u1*E + s), element k1 at out[c*M + E*N2*k1 + q]; forward
// DFT over k1 -> u2, written back to out[c*M + E*N2*u2 + q].
// Transforms (c, q) of one column: element i at out[c*M + (i << logQ) + q], in place.
struct Step3IO {
    u64* out;
    u64 M, s0, ntr;
    u32 logQ;
    __device__ __forceinline__ u64* base() const {
        return out + (s0 >> logQ) * M + (s0 & ((1ull << logQ) - 1));
    }
    __device__ __forceinline__ u32 off(u64 s) const {
        const u64 mask = (1ull << logQ) - 1;
        return (u32)(((s >> logQ) - (s0 >> logQ)) * M + (s & mask) - (s0 & mask));
    }
    template <int R>
    __device__ __forceinline__ void load(int t, int j, int per, u64 (&v)[R]) const {
        u64 s = s0 + t;
        if (s < ntr) {
            const u64* p = base();
            const u32 o = off(s);
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = p[o + ((u32)(j + r * per) << logQ)];
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = 0;
        }
    }
    template <int R>
    __device__ __forceinline__ void store(int t, int j, int Ns, const u64 (&v)[R]) const {
        u64 s = s0 + t;
        if (s >= ntr) return;
        u64* p = base();
        const u32 o = off(s);
        int k = j & (Ns - 1), base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) p[o + ((u32)(base + r * Ns) << logQ)] = v[r];
    }
};

template <int LT = LOG_TILE>
__device__ __forceinline__ void lde_step3_body(const Step3Args& A, u64 blk, const Roots& rt) {
    extern __shared__ u64 lds[];
    constexpr int NTPB = 1 << (LT - 4);
    const u32 b = A.logN - A.a;
    const u32 logQ = b + A.logE;  // transforms per column = E*N2
    const Step3IO io{A.out, A.M, blk << A.logNT, A.cols << logQ, logQ};
    // every input of the block is read in the first pass, before any output is
    // written: in place is safe
    fused_dft<false, NTPB, 1 << LT, true>(lds, A.a, 1 << A.logNT, A.logNT, rt.r4096[0], io, io);
}

// ---- host launchers of the kernels (one translation unit each) ----
// ntt_step1.hip
int launch_k_step1(bool inv, u32 lt, u32 blocks, size_t lds, hipStream_t st, const Step1Args& A,
                   const Roots& rt);
int launch_k_step1_multi(const Step1Args* args, const u32* start, u32 nseg, u32 blocks, size_t lds,
                         hipStream_t st, const Roots& rt);
int launch_k_step1_scaled(u32 lt, u32 blocks, size_t lds, hipStream_t st, const Step1Args& A,
                          const NttScaleDev& S, const Roots& rt);
int launch_k_ntt_step2(bool inv, u32 lt, u32 blocks, size_t lds, hipStream_t st, const u64* src,
                       u64* dst, u32 logN, u32 a, u64 nrows, u32 logNT, u64 ninv_m,
                       const Roots& rt);
int set_attributes_step1(int lds_bytes);
// lde_rows.hip
int launch_k_lde_rowinv(u32 lt, u32 blocks, size_t lds, hipStream_t st, const LdeArgs& A,
                        const Roots& rt);
int launch_k_lde_rowinv_multi(const LdeArgs* args, const u32* start, u32 nseg, u32 blocks,
                              size_t lds, hipStream_t st, const Roots& rt);
int launch_k_lde_step2(u32 blocks, size_t lds, hipStream_t st, const LdeArgs& A, const Roots& rt);
int launch_k_lde_step2_multi(const LdeArgs* args, const u32* start, u32 nseg, u32 blocks,
                             size_t lds, hipStream_t st, const Roots& rt);
int set_attributes_rows(int lds_bytes);
// lde_expand.hip
int launch_k_lde_expand(u32 lt, u32 blocks, size_t lds, hipStream_t st, const LdeArgs& A,
                        const Roots& rt);
int launch_k_lde_expand_multi(const LdeArgs* args, const u32* start, u32 nseg, u32 blocks,
                              size_t lds, hipStream_t st, const Roots& rt);
int set_attributes_expand(int lds_bytes);
// lde_step3.hip
int launch_k_lde_step3(u32 lt, u32 blocks, size_t lds, hipStream_t st, const Step3Args& A,
                       const Roots& rt);
int launch_k_lde_step3_multi(const Step3Args* args, const u32* start, u32 nseg, u32 blocks,
                             size_t lds, hipStream_t st, const Roots& rt);
int set_attributes_step3(int lds_bytes);

}  // namespace tvh
